// k_point.hip -- pointwise kernels for gfx950 (FP64).
//
//   state_kernel  update_flow_states (src/solver/StateUpdater.cpp:15-168) and
//                 calculate_time_step (src/solver/TimeIntegrator.cpp:20-178)
//                 fused per point: c and T on the interior plus the ghost
//                 "cross" layers (|rho| there), Sutherland mu on every point,
//                 dq=0, then the spectral radii and dt/vol on interior points
//                 (the only points whose radii the solver reads) and the
//                 minimum physical time step.
//   residual      ResidualMonitor::compute (src/solver/ResidualMonitor.cpp:11-138):
//                 deterministic two-pass reduction of sum r^2 and of the
//                 first maximum in (block,k,j,i,var) order.
#include "orion_dev.h"
#include "orion_kernels.h"

namespace orion_hip {

__global__ void __launch_bounds__(256) state_kernel(const BlockDev* __restrict__ blocks,
                                                    Span SP, int mode, unsigned long long* dtmin_base,
                                                    const int* step, int nreq) {
  __shared__ double smin[256];
  double mydt = 1000.0;
  int bi;
  long long n;
  if (span_item(SP, &bi, &n)) {
    const BlockDev& b = blocks[bi];
    const Params& p = b.par;  // the inputs of the block's request
    unsigned q, ui, q2, uj;
    udivmod((unsigned)n, (unsigned)b.nx, &q, &ui);
    udivmod(q, (unsigned)b.ny, &q2, &uj);
    const int i = (int)ui, j = (int)uj, k = (int)q2;
    const int oi = (i < NG || i >= NG + b.idim), oj = (j < NG || j >= NG + b.jdim), ok = (k < NG || k >= NG + b.kdim);
    const int nout = oi + oj + ok;
    // mode 8: ghost points only (state_pack_kernel forms the interior)
    if (!((mode & 8) && nout == 0)) {
      const long long P = b.P;
      const long long id = n;
      double c = 0.0, mu = 0.0;
      if (mode & 1) {
        if (nout <= 1) {
          double rho = b.prim[id];
          const double pp = b.prim[4 * P + id];
          if (nout == 1) rho = fabs(rho);
          const double a2 = p.gama * pp / rho;
          c = sqrt(fabs(a2));
          b.c[id] = c;
          if (p.nvis == 1) b.prim[5 * P + id] = (p.moo * p.moo) * a2;
        }
        if (p.nvis == 1) {
          double tm = b.prim[5 * P + id];
          if (tm < 1e-6) tm = 1e-6;
          mu = tm * sqrt(tm) * (1.0 + p.visc) / (tm + p.visc);
          b.mu[id] = mu;
        }
        if (b.q && nout <= 1) {
          const double rho = b.prim[id], u = b.prim[P + id], v = b.prim[2 * P + id], w = b.prim[3 * P + id];
          const double pp = b.prim[4 * P + id];
          const double rc = (nout == 1 || rho < 0.0) ? fabs(rho) : rho;
          b.q[id] = rc;
          b.q[P + id] = rc * u;
          b.q[2 * P + id] = rc * v;
          b.q[3 * P + id] = rc * w;
          const double ke = 0.5 * rc * (u * u + v * v + w * w);
          b.q[4 * P + id] = pp * (1.0 / (p.gama - 1.0)) + ke;
        }
        if (nout == 0 && !(mode & 4)) {  // mode 4: the first right-hand-side kernel writes dq whole
  #pragma unroll
          for (int m = 0; m < 5; ++m) b.dq[m * P + id] = 0.0;
        }
      }
      if ((mode & 2) && nout == 0) {
        if (!(mode & 1)) {
          c = b.c[id];
          mu = b.mu[id];
        }
        const double u = b.prim[P + id], v = b.prim[2 * P + id], w = b.prim[3 * P + id];
        const double rho = b.prim[id], vol = b.vol[id];
        const double csrv = (double)p.csrv;
        double sum = 0.0;
  #pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double mx = b.met[(3 * d) * P + id], my = b.met[(3 * d + 1) * P + id], mz = b.met[(3 * d + 2) * P + id];
          const double U = fabs(u * mx + v * my + w * mz);
          const double gl = sqrt(mx * mx + my * my + mz * mz);
          const double spec = U + c * gl;
          double sv = 0.0;
          if (p.nvis == 1) {
            const double g2 = mx * mx + my * my + mz * mz;
            const double denom = p.reynolds * rho * vol + 1.0e-30;
            sv = (2.0 * mu / denom) * g2 * csrv;
          }
          sum += spec + sv;
        }
        const double dtp = p.cfl / (sum + 1.0e-30);
        const double dtv = dtp / vol;
        b.dt[id] = dtv;
        const double back = dtv * vol;
        mydt = back < mydt ? back : mydt;
      }
    }
  }
  if (!dtmin_base) return;
  // slot (step, request): a workgroup holds items of one block
  const int req = blocks[SP.wgb[xcd_remap(blockIdx.x, gridDim.x)]].req;
  unsigned long long* dtmin_bits = dtmin_base + (long long)(step ? *step : 0) * nreq + req;
  smin[threadIdx.x] = mydt;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const double o = smin[threadIdx.x + s];
      if (o < smin[threadIdx.x]) smin[threadIdx.x] = o;
    }
    __syncthreads();
  }
  // positive doubles order like their bit patterns
  if (threadIdx.x == 0 && smin[0] < 1000.0) atomicMin(dtmin_bits, (unsigned long long)__double_as_longlong(smin[0]));
}

// ---------------------------------------------------------------------------
// residual
// ---------------------------------------------------------------------------
struct MaxLoc {
  double v;
  long long key;
};
__device__ inline bool better(const MaxLoc& a, const MaxLoc& b) {
  return a.v > b.v || (a.v == b.v && a.key < b.key);
}

constexpr int RES_TB = 256;
constexpr int RES_MAXPARTS = 1024;

int residual_parts(long long n) {
  long long w = (n + RES_TB - 1) / RES_TB;
  if (w < 1) w = 1;
  return (int)(w < RES_MAXPARTS ? w : RES_MAXPARTS);
}

__global__ void __launch_bounds__(RES_TB) residual_kernel(const BlockDev* __restrict__ blocks,
                                                          const long long* __restrict__ off, int nb, double* partial,
                                                          double* pmax, long long* pidx_) {
  __shared__ double ssum[RES_TB];
  __shared__ double smax[RES_TB];
  __shared__ long long skey[RES_TB];
  const long long total = off[nb];
  // one contiguous chunk of the interior nodes per workgroup: the block
  // changes only at block boundaries inside a chunk
  const long long chunk = ((total + gridDim.x - 1) / gridDim.x + RES_TB - 1) / RES_TB * RES_TB;
  const long long g0 = (long long)blockIdx.x * chunk;
  const long long g1 = g0 + chunk < total ? g0 + chunk : total;
  double sum = 0.0;
  MaxLoc best{-1.0, 0x7fffffffffffffffLL};
  int bi = find_block(off, nb, g0 < total ? g0 : 0);
  for (long long g = g0 + threadIdx.x; g < g1; g += RES_TB) {
    while (g >= off[bi + 1]) ++bi;
    const BlockDev& b = blocks[bi];
    int i, j, k;
    decode_interior(b, g - off[bi], &i, &j, &k);
    const long long id = pidx(b, i, j, k);
    double vol = b.vol[id], dt = b.dt[id];
    if (vol < 1e-30) vol = 1e-30;
    if (dt < 1e-30) dt = 1e-30;
    const double inv = 1.0 / (dt * vol);
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      const double r = fabs(b.dq[m * b.P + id]) * inv;
      sum += r * r;
      MaxLoc c{r, g * 5 + m};
      if (better(c, best)) best = c;
    }
  }
  ssum[threadIdx.x] = sum;
  smax[threadIdx.x] = best.v;
  skey[threadIdx.x] = best.key;
  __syncthreads();
  for (int s = RES_TB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      ssum[threadIdx.x] = ssum[threadIdx.x] + ssum[threadIdx.x + s];
      MaxLoc a{smax[threadIdx.x], skey[threadIdx.x]}, c{smax[threadIdx.x + s], skey[threadIdx.x + s]};
      if (better(c, a)) {
        smax[threadIdx.x] = c.v;
        skey[threadIdx.x] = c.key;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = ssum[0];
    pmax[blockIdx.x] = smax[0];
    pidx_[blockIdx.x] = skey[0];
  }
}

// Workgroup r reduces request r's partials, seg[r] .. seg[r + 1] (all
// nparts when seg is null), over tot5[r] (total5) terms into row
// (*step) * nreq + r of out.  With step: the last workgroup to finish (done
// counts them, and is left at 0) advances *step, after every workgroup has
// read it.
__global__ void __launch_bounds__(RES_MAXPARTS) residual_final_kernel(const double* partial, const double* pmax,
                                                                      const long long* pidx_, const int* seg,
                                                                      int nparts, const long long* tot5,
                                                                      long long total5, double* out_base,
                                                                      int* step, int* done, int nreq) {
  __shared__ double ssum[RES_MAXPARTS];
  __shared__ double smax[RES_MAXPARTS];
  __shared__ long long skey[RES_MAXPARTS];
  const int t = threadIdx.x, r = blockIdx.x;
  const int q0 = seg ? seg[r] : 0, q1 = seg ? seg[r + 1] : nparts;
  // partials q0 + t, q0 + t + 1024, ... in order, then a tree over the threads
  double sum = 0.0;
  MaxLoc best{-1.0, 0x7fffffffffffffffLL};
  for (int q = q0 + t; q < q1; q += RES_MAXPARTS) {
    sum = q == q0 + t ? partial[q] : sum + partial[q];
    const MaxLoc c{pmax[q], pidx_[q]};
    if (better(c, best)) best = c;
  }
  ssum[t] = sum;
  smax[t] = best.v;
  skey[t] = best.key;
  __syncthreads();
  for (int s = RES_MAXPARTS / 2; s > 0; s >>= 1) {
    if (t < s) {
      ssum[t] = ssum[t] + ssum[t + s];
      MaxLoc a{smax[t], skey[t]}, c{smax[t + s], skey[t + s]};
      if (better(c, a)) {
        smax[t] = c.v;
        skey[t] = c.key;
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    const int sidx = step ? *step : 0;
    const long long t5 = tot5 ? tot5[r] : total5;
    double* out = out_base + 4 * ((long long)sidx * nreq + r);
    out[0] = t5 > 0 ? sqrt(ssum[0] / (double)t5) : 0.0;
    out[1] = smax[0];
    out[2] = (double)skey[0];
    if (step && atomicAdd(done, 1) == nreq - 1) {
      // every workgroup has read *step (its value addressed the stores above)
      *done = 0;
      *step = sidx + 1;
    }
  }
}

__global__ void fill_kernel(double* p, long long n, double v) {
  const long long g = xcd_thread();
  if (g < n) p[g] = v;
}

// update_temperature, src/preprocess/InitialCondition.cpp:91-147
__global__ void __launch_bounds__(256) temperature_kernel(const BlockDev* __restrict__ blocks,
                                                          Span SP) {
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;
  const BlockDev& b = blocks[bi];
  const double gama = b.par.gama, moo = b.par.moo;
  const long long id = gi;
  double rho = b.prim[id];
  const double p = b.prim[4 * b.P + id];
  if (rho < 1.0e-30) rho = 1.0e-30;
  const double a2 = gama * p / rho;
  b.prim[5 * b.P + id] = (moo * moo) * a2;
}

// The point fields that PostProcess::write_block_vts_binary writes
// (src/postprocess/PostProcess.cpp:183-240), on the interior nodes of block
// bi, i fastest, as float: Density, Pressure, Temperature and Mach (n each),
// then Velocity as (u, v, w) per node.  Mach is sqrt(v2 / c2) with
// c2 = gamma p / rho, 0 where c2 <= 1e-30, in double, then rounded.
__global__ void __launch_bounds__(256) output_kernel(const BlockDev* __restrict__ blocks, int bi, float* out,
                                                     double gamma) {
  const BlockDev& b = blocks[bi];
  const long long n = (long long)b.idim * b.jdim * b.kdim;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  unsigned q, ui, q2, uj;
  udivmod((unsigned)g, (unsigned)b.idim, &q, &ui);
  udivmod(q, (unsigned)b.jdim, &q2, &uj);
  const long long P = b.P;
  const long long id = pidx(b, (int)ui + NG, (int)uj + NG, (int)q2 + NG);
  const double r = b.prim[id], u = b.prim[P + id], v = b.prim[2 * P + id], w = b.prim[3 * P + id];
  const double p = b.prim[4 * P + id], t = b.prim[5 * P + id];
  out[g] = (float)r;
  out[n + g] = (float)p;
  out[2 * n + g] = (float)t;
  const double v2 = u * u + v * v + w * w;
  const double c2 = gamma * p / r;
  out[3 * n + g] = (float)((c2 > 1e-30) ? sqrt(v2 / c2) : 0.0);
  out[4 * n + 3 * g] = (float)u;
  out[4 * n + 3 * g + 1] = (float)v;
  out[4 * n + 3 * g + 2] = (float)w;
}

static inline unsigned grid_of(long long n, int tb) { return (unsigned)((n + tb - 1) / tb); }

void launch_state(const Launch& L, long long n, int mode, unsigned long long* dtmin_base, const int* step,
                  int nreq) {
  if (n <= 0) return;
  hipLaunchKernelGGL(state_kernel, dim3((unsigned)L.nwg), dim3(SPAN_TB), 0, L.stream, L.blocks, L.span, mode,
                     dtmin_base, step, nreq);
}
void launch_residual(const Launch& L, long long n, long long total5, double* partial, long long* pidx_, int nparts,
                     double* out) {
  hipLaunchKernelGGL(residual_kernel, dim3(nparts), dim3(RES_TB), 0, L.stream, L.blocks, L.off, L.nb, partial,
                     partial + RES_MAXPARTS, pidx_);
  hipLaunchKernelGGL(residual_final_kernel, dim3(1), dim3(RES_MAXPARTS), 0, L.stream, partial, partial + RES_MAXPARTS,
                     pidx_, nullptr, nparts, nullptr, total5, out, nullptr, nullptr, 1);
}
void launch_residual_final(const double* psum, const double* pmax, const long long* pkey, const int* seg,
                           const long long* tot5, int nreq, double* out, int* step, int* done, hipStream_t s) {
  hipLaunchKernelGGL(residual_final_kernel, dim3(nreq), dim3(RES_MAXPARTS), 0, s, psum, pmax, pkey, seg, 0, tot5, 0LL,
                     out, step, done, nreq);
}
void launch_output(const BlockDev* blocks, int bi, long long n, float* out, double gamma, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(output_kernel, dim3(grid_of(n, 256)), dim3(256), 0, s, blocks, bi, out, gamma);
}
void launch_fill(double* p, long long n, double v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(grid_of(n, 256)), dim3(256), 0, s, p, n, v);
}
void launch_temperature(const Launch& L, long long n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(temperature_kernel, dim3((unsigned)L.nwg), dim3(SPAN_TB), 0, L.stream, L.blocks, L.span);
}

}  // namespace orion_hip
