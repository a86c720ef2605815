// k_lusgs.hip -- LU-SGS update for gfx950 (FP64).
//
// StateUpdater::update_solution (src/solver/StateUpdater.cpp:297-580) sweeps
// each block lexicographically (k, then j, then i; reversed for the backward
// sweeps) and every cell reads only its six face neighbours: a cell sees the
// new values of the neighbours before it in sweep order and the old values
// of those after it.  Any schedule with that property gives bit-identical
// results; here it is the hyperplane order s = i + j + k (neighbours before a
// cell lie on s - 1, those after it on s + 1), run as a pipeline:
//
//  * Each block's interior is viewed along (march, col, row) axes -- march
//    the longest axis, col = i whenever i is not the march axis -- and the
//    (col, row) cross-section is cut into 16 x 16 tiles.  One workgroup of
//    256 lanes owns a tile; lane (row, col) marches along its pencil, and at
//    pipeline step s processes cell m = s - row - col, so all lanes of a tile
//    are on hyperplane s together.
//  * Inside a tile, the neighbours before a cell along col and row are lanes
//    of the same workgroup and come through LDS (buffers alternating by step
//    parity, one barrier per step); along the march axis it is the lane's
//    own previous cell.
//  * Across tiles, a tile publishes how far it has come in a progress flag
//    and its downstream tiles wait on it (bounded spin) before reading edge
//    values.  dq is stored with the cache policy of agent-scope atomic
//    stores (16-byte pair stores), plane s is
//    published after every wave has waited for its stores of that plane,
//    and edge dq is read with agent-scope atomic loads.  No cache
//    maintenance is needed (fence-based release/acquire around plain stores
//    flushes and invalidates L2 and measured 2-8x slower).
//    Tiles are launched in dependency order and a wait only ever targets a
//    tile launched earlier, so the persistent grid always drains; a spin
//    that exceeds its bound records an error instead of hanging.
//
// Everything in a cell's update that does not depend on dq is either packed
// once per update by lu_pack's state part (the diagonal 1/(1/dt + sum of
// radii) and each cell's u, v, w, c^2, H, a_f, c and viscous coefficient;
// its rhs part copies the right-hand side) or recomputed per
// neighbour slot in the sweep from this cell's metrics and the neighbour's
// packed values, with the reference's expressions (calc_spectral_radius
// :175-191, calc_visc_spectral_radius :194-207, compute_flux_splitting
// :209-264, get_diag_coeff :267-295, sweep bodies :318-365, :392-441).
// The Gauss-Seidel sweeps take their after-neighbour terms from the sweep
// before them (see lusgs_tile_kernel).
#include "orion_dev.h"
#include "orion_kernels.h"
#include <type_traits>

namespace orion_hip {

constexpr int TW = SWEEP_TW;   // tile width along col
constexpr int TR = SWEEP_TR;   // tile height along row
constexpr int TL = TW * TR;    // lanes per tile workgroup (of each kind)
constexpr int KEY_BITS = 16;   // progress flag: sweep serial << KEY_BITS | plane key
constexpr int CPOL_SC1 = 16;   // cache-policy operand of a memory builtin: the sc1 bit (gfx950)

__device__ inline double spec_rad(const double* pr, double g, double kx, double ky, double kz) {
  double rho = pr[0], p = pr[4];
  const double u = pr[1], v = pr[2], w = pr[3];
  if (rho < 1e-30) rho = 1e-30;
  if (p < 1e-30) p = 1e-30;
  const double c2 = g * p / rho;
  const double c = sqrt(c2);
  const double U = u * kx + v * ky + w * kz + 0.0;
  const double gm = sqrt(kx * kx + ky * ky + kz * kz);
  return fabs(U) + c * gm;
}

__device__ inline double visc_rad(double rho, double mu, double vol, double re, double kx, double ky, double kz) {
  if (rho < 1e-30) rho = 1e-30;
  const double g2 = kx * kx + ky * ky + kz * kz;
  const double rmv = rho * vol;
  const double coef = 2.0 * mu / (fabs(re) * rmv + 1.0e-30);
  return coef * g2;
}

__device__ inline void natural_to_sweep(const BlockDev& b, int i, int j, int k, int* m, int* r, int* c) {
  const int x[3] = {i - NG, j - NG, k - NG};
  *m = x[b.sax[0]];
  *c = x[b.sax[1]];
  *r = x[b.sax[2]];
}

// ---------------------------------------------------------------------------
// packing and unpacking between the block arrays and the sweep layout
//
// A hyperplane row of the sweep layout holds cells of consecutive columns
// and decreasing march positions, so neither natural-order nor sweep-order
// threads reach both sides contiguously.  Each workgroup takes a tile of
// 16 x 16 (march, col) cells of one row and stages it through LDS: the
// block arrays are accessed with lane = (march, col) (contiguous along the
// col axis whenever col = i), the sweep arrays with the tile's cells in
// sweep order (contiguous runs of up to 16 columns per hyperplane row).
// ---------------------------------------------------------------------------
constexpr int PT = 16;

// the tile of launch item `item` (256 items per tile)
__device__ inline void pack_tile(const BlockDev& b, long long item, int* m0, int* c0, int* r) {
  const unsigned tc = (b.sn[1] + PT - 1) / PT, tm = (b.sn[0] + PT - 1) / PT;
  unsigned q, rc, q2, rm;
  udivmod((unsigned)(item / (PT * PT)), tc, &q, &rc);
  udivmod(q, tm, &q2, &rm);
  *c0 = (int)rc * PT;
  *m0 = (int)rm * PT;
  *r = (int)q2;
}

// tile position (mi * PT + ci) of the tile's u-th cell in sweep order:
// by diagonal mi + ci (the hyperplane), then by column
__device__ inline int sweep_slot(int u) {
  int sd = 0, cnt = 0;
  for (;;) {
    const int len = sd < PT ? sd + 1 : 2 * PT - 1 - sd;
    if (u < cnt + len) break;
    cnt += len;
    ++sd;
  }
  const int ci = (sd < PT ? 0 : sd - PT + 1) + (u - cnt);
  return (sd - ci) * PT + ci;
}

// padded block index of sweep cell (m, r, c)
__device__ inline long long natural_id(const BlockDev& b, int m, int r, int c) {
  return pidx(b, NG, NG, NG) + (long long)m * dstride(b, b.sax[0]) + (long long)c * dstride(b, b.sax[1]) +
         (long long)r * dstride(b, b.sax[2]);
}

// Once, after the metrics: the static own pairs 0..3 (met0..met7).
__global__ void __launch_bounds__(256) lu_geom_kernel(const BlockDev* __restrict__ blocks, Span SP, SweepArrays S) {
  __shared__ double sh[8][PT * PT];
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;  // whole tiles: every lane of a workgroup has an item
  const BlockDev& b = blocks[bi];
  int m0, c0, r;
  pack_tile(b, gi, &m0, &c0, &r);
  const int lane = threadIdx.x;
  {
    const int m = m0 + lane / PT, c = c0 + lane % PT;
    if (m < b.sn[0] && c < b.sn[1]) {
      const long long id = natural_id(b, m, r, c);
#pragma unroll
      for (int f = 0; f < 8; ++f) sh[f][lane] = b.met[f * b.P + id];
    }
  }
  __syncthreads();
  const int src = sweep_slot(lane);
  const int m = m0 + src / PT, c = c0 + src % PT;
  if (m < b.sn[0] && c < b.sn[1]) {
    const long long rb = sweep_index(b, m, r, c) - c;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      S.own[rb * SW_OWN + (long long)k * b.sncp + c] = make_double2(sh[2 * k][src], sh[2 * k + 1][src]);
  }
}

// Per update, in two parts.  PART 1 (the state part, which depends only on
// the flow state, the time step and the grid, so it can run beside the
// right-hand side): each cell's met8 and diagonal coed of get_diag_coeff
// (own pair 4) and, for its neighbours, u, v, w, c^2, H and the viscous
// coefficient (D; the sweep recomputes a_f and c from them with the same
// expressions, so each value is the one packed before).  PART 2 (after the right-hand side): the rhs
// copy (own pairs 5..7).
template <int PART>
__global__ void __launch_bounds__(256) lu_pack_kernel(const BlockDev* __restrict__ blocks,
                                                      Span SP, SweepArrays S) {
  constexpr int NPK = PART == 1 ? 8 : 5;  // staged words per cell
  __shared__ double sh[NPK][PT * PT];
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;  // whole tiles: every lane of a workgroup has an item
  const BlockDev& b = blocks[bi];
  const Params& p = b.par;  // the inputs of the block's request
  int m0, c0, r;
  pack_tile(b, gi, &m0, &c0, &r);
  const int lane = threadIdx.x;
  {
    const int m = m0 + lane / PT, c = c0 + lane % PT;
    if (m < b.sn[0] && c < b.sn[1]) {
      const long long P = b.P;
      const long long id = natural_id(b, m, r, c);
      if constexpr (PART == 2) {
        // rhs copy with the physical-BC windows zeroed (StateUpdater.cpp:557-570)
        const bool fz = b.frozen[id] != 0;
#pragma unroll
        for (int f = 0; f < 5; ++f) sh[f][lane] = fz ? 0.0 : b.dq[f * P + id];
      } else {
        const double gamma = p.gama, re = p.reynolds;
        double pr[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) pr[f] = b.prim[f * P + id];
        double met[9];
#pragma unroll
        for (int f = 0; f < 9; ++f) met[f] = b.met[f * P + id];
        sh[0][lane] = met[8];
        const double mu = b.mu[id];
        double vol = b.vol[id];
        if (vol < 1e-30) vol = 1e-30;
        {  // get_diag_coeff
          const double ra = spec_rad(pr, gamma, met[0], met[1], met[2]);
          const double rb = spec_rad(pr, gamma, met[3], met[4], met[5]);
          const double rc = spec_rad(pr, gamma, met[6], met[7], met[8]);
          double rva = 0, rvb = 0, rvc = 0;
          if (p.nvis == 1) {
            rva = visc_rad(pr[0], mu, vol, re, met[0], met[1], met[2]);
            rvb = visc_rad(pr[0], mu, vol, re, met[3], met[4], met[5]);
            rvc = visc_rad(pr[0], mu, vol, re, met[6], met[7], met[8]);
          }
          const double rs = ra + rb + rc + rva + rvb + rvc;
          const double term = (1.0 / b.dt[id]) + 1.0 * 1.0 * rs;
          sh[1][lane] = 1.0 / term;
        }
        // this cell seen as a neighbour: compute_flux_splitting's per-cell
        // terms, the sound speed of calc_spectral_radius and the viscous
        // coefficient of calc_visc_spectral_radius (same clamps)
        double rm = pr[0], pm = pr[4];
        const double v2 = pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3];
        if (rm < 1e-30) rm = 1e-30;
        if (pm < 1e-30) pm = 1e-30;
        const double c2 = gamma * pm / rm;
        sh[2][lane] = pr[1];
        sh[3][lane] = pr[2];
        sh[4][lane] = pr[3];
        sh[5][lane] = c2;
        sh[6][lane] = (gamma / (gamma - 1.0)) * (pm / rm) + 0.5 * v2;
        sh[7][lane] = 2.0 * mu / (fabs(re) * (rm * vol) + 1.0e-30);
      }
    }
  }
  __syncthreads();
  const int src = sweep_slot(lane);
  const int m = m0 + src / PT, c = c0 + src % PT;
  if (m < b.sn[0] && c < b.sn[1]) {
    const int ncp = b.sncp;
    const long long rb = sweep_index(b, m, r, c) - c;
    if constexpr (PART == 2) {
#pragma unroll
      for (int k = 0; k < 3; ++k)
        S.own[rb * SW_OWN + (long long)(5 + k) * ncp + c] = make_double2(sh[2 * k][src], k < 2 ? sh[2 * k + 1][src] : 0.0);
    } else {
      S.own[rb * SW_OWN + 4LL * ncp + c] = make_double2(sh[0][src], sh[1][src]);
#pragma unroll
      for (int k = 0; k < SW_D; ++k)
        S.D[rb * SW_D + (long long)k * ncp + c] = make_double2(sh[2 + 2 * k][src], sh[3 + 2 * k][src]);
    }
  }
}

// The step's flow states and time step at the interior points
// (state_kernel, modes 1 | 2 | 4, whose interior branch this restates
// expression for expression: c, T, mu, q, dt and the time step's minimum)
// fused with the LU-SGS pack's state part (lu_pack_kernel<1>, from the same
// values in registers instead of reading back mu and dt), over the pack's
// tiles: a workgroup takes 16 x 16 (march, col) cells of one row, forms the
// state with lane = (march, col) and stages the pack's words through LDS
// into the sweep layout.  The ghost points keep their own launch
// (state_kernel, mode 8).  Each stored value is the one the two kernels
// store; the minimum of dt vol is order-free.
__global__ void __launch_bounds__(256) state_pack_kernel(const BlockDev* __restrict__ blocks, Span SP,
                                                         SweepArrays S, unsigned long long* dtmin_base,
                                                         const int* step, int nreq) {
  __shared__ double sh[8][PT * PT];
  __shared__ double smin[PT * PT];
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;  // whole tiles: every lane of a workgroup has an item
  const BlockDev& b = blocks[bi];
  const Params& p = b.par;  // the inputs of the block's request
  int m0, c0, r;
  pack_tile(b, gi, &m0, &c0, &r);
  const int lane = threadIdx.x;
  double mydt = 1000.0;
  {
    const int m = m0 + lane / PT, c = c0 + lane % PT;
    if (m < b.sn[0] && c < b.sn[1]) {
      const long long P = b.P;
      const long long id = natural_id(b, m, r, c);
      // update_flow_states at an interior point (state_kernel, mode 1)
      double cs = 0.0, mu = 0.0;
      {
        double rho = b.prim[id];
        const double pp = b.prim[4 * P + id];
        const double a2 = p.gama * pp / rho;
        cs = sqrt(fabs(a2));
        b.c[id] = cs;
        if (p.nvis == 1) b.prim[5 * P + id] = (p.moo * p.moo) * a2;
      }
      if (p.nvis == 1) {
        double tm = b.prim[5 * P + id];
        if (tm < 1e-6) tm = 1e-6;
        mu = tm * sqrt(tm) * (1.0 + p.visc) / (tm + p.visc);
        b.mu[id] = mu;
      }
      if (b.q) {
        const double rho = b.prim[id], u = b.prim[P + id], v = b.prim[2 * P + id], w = b.prim[3 * P + id];
        const double pp = b.prim[4 * P + id];
        const double rc = (rho < 0.0) ? fabs(rho) : rho;
        b.q[id] = rc;
        b.q[P + id] = rc * u;
        b.q[2 * P + id] = rc * v;
        b.q[3 * P + id] = rc * w;
        const double ke = 0.5 * rc * (u * u + v * v + w * w);
        b.q[4 * P + id] = pp * (1.0 / (p.gama - 1.0)) + ke;
      }
      // calculate_time_step (state_kernel, mode 2)
      double pr[5];
#pragma unroll
      for (int f = 0; f < 5; ++f) pr[f] = b.prim[f * P + id];
      double met[9];
#pragma unroll
      for (int f = 0; f < 9; ++f) met[f] = b.met[f * P + id];
      const double vol0 = b.vol[id];
      double dtv;
      {
        const double u = pr[1], v = pr[2], w = pr[3];
        const double rho = pr[0], vol = vol0;
        const double csrv = (double)p.csrv;
        double sum = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double mx = met[3 * d], my = met[3 * d + 1], mz = met[3 * d + 2];
          const double U = fabs(u * mx + v * my + w * mz);
          const double gl = sqrt(mx * mx + my * my + mz * mz);
          const double spec = U + cs * gl;
          double sv = 0.0;
          if (p.nvis == 1) {
            const double g2 = mx * mx + my * my + mz * mz;
            const double denom = p.reynolds * rho * vol + 1.0e-30;
            sv = (2.0 * mu / denom) * g2 * csrv;
          }
          sum += spec + sv;
        }
        const double dtp = p.cfl / (sum + 1.0e-30);
        dtv = dtp / vol;
        b.dt[id] = dtv;
        const double back = dtv * vol;
        mydt = back < mydt ? back : mydt;
      }
      // the pack's state part (lu_pack_kernel<1>) from the same values;
      // without viscosity mu is the array's value, as the pack reads it
      if (p.nvis != 1) mu = b.mu[id];
      const double gamma = p.gama, re = p.reynolds;
      sh[0][lane] = met[8];
      double vol = vol0;
      if (vol < 1e-30) vol = 1e-30;
      {  // get_diag_coeff
        const double ra = spec_rad(pr, gamma, met[0], met[1], met[2]);
        const double rb = spec_rad(pr, gamma, met[3], met[4], met[5]);
        const double rcd = spec_rad(pr, gamma, met[6], met[7], met[8]);
        double rva = 0, rvb = 0, rvc = 0;
        if (p.nvis == 1) {
          rva = visc_rad(pr[0], mu, vol, re, met[0], met[1], met[2]);
          rvb = visc_rad(pr[0], mu, vol, re, met[3], met[4], met[5]);
          rvc = visc_rad(pr[0], mu, vol, re, met[6], met[7], met[8]);
        }
        const double rs = ra + rb + rcd + rva + rvb + rvc;
        const double term = (1.0 / dtv) + 1.0 * 1.0 * rs;
        sh[1][lane] = 1.0 / term;
      }
      double rm = pr[0], pm = pr[4];
      const double v2 = pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3];
      if (rm < 1e-30) rm = 1e-30;
      if (pm < 1e-30) pm = 1e-30;
      const double c2 = gamma * pm / rm;
      sh[2][lane] = pr[1];
      sh[3][lane] = pr[2];
      sh[4][lane] = pr[3];
      sh[5][lane] = c2;
      sh[6][lane] = (gamma / (gamma - 1.0)) * (pm / rm) + 0.5 * v2;
      sh[7][lane] = 2.0 * mu / (fabs(re) * (rm * vol) + 1.0e-30);
    }
  }
  smin[lane] = mydt;
  __syncthreads();
  {
    const int src = sweep_slot(lane);
    const int m = m0 + src / PT, c = c0 + src % PT;
    if (m < b.sn[0] && c < b.sn[1]) {
      const int ncp = b.sncp;
      const long long rb = sweep_index(b, m, r, c) - c;
      S.own[rb * SW_OWN + 4LL * ncp + c] = make_double2(sh[0][src], sh[1][src]);
#pragma unroll
      for (int k = 0; k < SW_D; ++k)
        S.D[rb * SW_D + (long long)k * ncp + c] = make_double2(sh[2 + 2 * k][src], sh[3 + 2 * k][src]);
    }
  }
  // the time step's minimum over the workgroup, into the slot of (step,
  // request) as state_kernel leaves it
  for (int h = PT * PT / 2; h > 0; h >>= 1) {
    if (lane < h) {
      const double o = smin[lane + h];
      if (o < smin[lane]) smin[lane] = o;
    }
    __syncthreads();
  }
  if (lane == 0 && smin[0] < 1000.0) {
    unsigned long long* dtmin_bits = dtmin_base + (long long)(step ? *step : 0) * nreq + b.req;
    atomicMin(dtmin_bits, (unsigned long long)__double_as_longlong(smin[0]));
  }
}

// dq back to the block arrays (zero on physical-BC windows,
// StateUpdater.cpp:577) and update_limited (:470-542).
// The update also forms the residual's terms (ResidualMonitor.cpp, as in
// residual_kernel) of its cells from the dq it stores: each workgroup leaves
// the sum of squares and the largest term with its key (interior index x 5
// + component) in part[blockIdx.x], for residual_final_kernel.
__global__ void __launch_bounds__(256) update_kernel(const BlockDev* __restrict__ blocks,
                                                     Span SP, SweepArrays S,
                                                     double gama, const long long* __restrict__ ioff,
                                                     double* psum, double* pmax, long long* pkey,
                                                     unsigned long long* epoch) {
  __shared__ double sh[5][PT * PT];
  __shared__ double ssum[PT * PT];
  __shared__ double smax[PT * PT];
  __shared__ long long skey[PT * PT];
  // the step's four sweeps are done (they precede this kernel on the
  // stream): the next step's take the next four serials
  if (blockIdx.x == 0 && threadIdx.x == 0) *epoch += 1;
  int bi;
  long long gi;
  // partials in tile order (the span's workgroup index): a block's, and a
  // request's, are contiguous and ordered alike in any context
  const unsigned wq = xcd_remap(blockIdx.x, gridDim.x);
  if (!span_item(SP, &bi, &gi)) {  // whole tiles: the workgroup has no tile
    if (threadIdx.x == 0) {
      psum[wq] = 0.0;
      pmax[wq] = -1.0;
      pkey[wq] = 0x7fffffffffffffffLL;
    }
    return;
  }
  const BlockDev& b = blocks[bi];
  int m0, c0, r;
  pack_tile(b, gi, &m0, &c0, &r);
  const int lane = threadIdx.x;
  {
    const int src = sweep_slot(lane);
    const int m = m0 + src / PT, c = c0 + src % PT;
    if (m < b.sn[0] && c < b.sn[1]) {
      const int ncp = b.sncp;
      const long long rb = sweep_index(b, m, r, c) - c;
      const double2 a = S.dq[rb * SW_DQ + c], e = S.dq[rb * SW_DQ + ncp + c];
      const double g = reinterpret_cast<const double*>(S.dq)[2 * (rb * SW_DQ + 2 * (long long)ncp + c)];
      sh[0][src] = a.x;
      sh[1][src] = a.y;
      sh[2][src] = e.x;
      sh[3][src] = e.y;
      sh[4][src] = g;
    }
  }
  __syncthreads();
  const int m = m0 + lane / PT, c = c0 + lane % PT;
  double rsum = 0.0, rmax = -1.0;
  long long rkey = 0x7fffffffffffffffLL;
  if (m < b.sn[0] && c < b.sn[1]) {
    const long long id = natural_id(b, m, r, c);
    const long long P = b.P;
    double d[5];
    const bool fz = b.frozen[id] != 0;
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      d[f] = fz ? 0.0 : sh[f][lane];
      b.dq[f * P + id] = d[f];
    }
    {  // the residual's terms |dq| / (dt vol), same clamps as residual_kernel
      double vol = b.vol[id], dt = b.dt[id];
      if (vol < 1e-30) vol = 1e-30;
      if (dt < 1e-30) dt = 1e-30;
      const double inv = 1.0 / (dt * vol);
      auto along = [&](int ax) { return b.sax[0] == ax ? m : (b.sax[1] == ax ? c : r); };
      const long long g = ioff[bi] + ((long long)along(2) * b.jdim + along(1)) * b.idim + along(0);
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        const double x = fabs(d[f]) * inv;
        rsum += x * x;
        if (x > rmax) {  // components in increasing key order: the first of equal terms stays
          rmax = x;
          rkey = g * 5 + f;
        }
      }
    }
    const double gm1 = gama - 1.0;
    const double rr = b.prim[id], u = b.prim[P + id], v = b.prim[2 * P + id], w = b.prim[3 * P + id];
    const double pp = b.prim[4 * P + id];
    const double E = pp / gm1 + 0.5 * rr * (u * u + v * v + w * w);
    const double q0 = rr + d[0], q1 = rr * u + d[1], q2 = rr * v + d[2], q3 = rr * w + d[3], q4 = E + d[4];
    double rn = q0;
    double ir = 1.0 / rn;
    const double a = q1 * ir, bb = q2 * ir, cc = q3 * ir;
    const double pn = gm1 * (q4 - 0.5 * rn * (a * a + bb * bb + cc * cc));
    const double dpp = fabs((pn - pp) / pp);
    const double x = -0.3 + dpp;
    const double term = (0.0 < x) ? x : 0.0;  // std::max(0.0, x)
    const double Fa = 1.0 / (1.0 + 2.0 * term);
    const double f0 = rr + d[0] * Fa, f1 = rr * u + d[1] * Fa, f2 = rr * v + d[2] * Fa, f3 = rr * w + d[3] * Fa;
    const double f4 = E + d[4] * Fa;
    rn = f0;
    ir = 1.0 / rn;
    const double un = f1 * ir, vn = f2 * ir, wn = f3 * ir;
    const double pf = gm1 * (f4 - 0.5 * rn * (un * un + vn * vn + wn * wn));
    if (!(rn <= 1e-6 || pf <= 1e-6)) {
      b.prim[id] = rn;
      b.prim[P + id] = un;
      b.prim[2 * P + id] = vn;
      b.prim[3 * P + id] = wn;
      b.prim[4 * P + id] = pf;
    }
  }
  ssum[lane] = rsum;
  smax[lane] = rmax;
  skey[lane] = rkey;
  __syncthreads();
  for (int h = PT * PT / 2; h > 0; h >>= 1) {
    if (lane < h) {
      ssum[lane] = ssum[lane] + ssum[lane + h];
      const double ov = smax[lane + h];
      const long long ok = skey[lane + h];
      if (ov > smax[lane] || (ov == smax[lane] && ok < skey[lane])) {
        smax[lane] = ov;
        skey[lane] = ok;
      }
    }
    __syncthreads();
  }
  if (lane == 0) {
    psum[wq] = ssum[0];
    pmax[wq] = smax[0];
    pkey[wq] = skey[0];
  }
}

// ---------------------------------------------------------------------------
// one neighbour slot: the inviscid part of compute_flux_splitting (:209-264,
// with calc_spectral_radius) from this cell's metrics and the neighbour's
// packed values and dq
// ---------------------------------------------------------------------------
struct AxisGeo {
  double g2, sq, gm, im;  // |n|^2, sqrt(|n|^2), clamped, 1/clamped
};

__device__ __forceinline__ void axis_geo(const double* n, AxisGeo& a) {
  a.g2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  a.sq = sqrt(a.g2);
  a.gm = (a.sq < 1.0e-30) ? 1.0e-30 : a.sq;
  a.im = 1.0 / a.gm;
}

// n: this cell's metrics of the slot's axis; Dn, d: the neighbour's packed
// values and dq.  a_f = 0.5 (gamma - 1) |v|^2 and c = sqrt(c^2) are formed
// here as compute_flux_splitting and calc_spectral_radius form them (and as
// the pack formed them before: the same operations on the same values).
template <int SL>
__device__ __forceinline__ void slot_inv(const double* n, const AxisGeo& a, const double* Dn, const double* d,
                                         double gm1, double* res) {
  constexpr int sgn = SL < 3 ? 1 : -1;
  const double nx = n[0], ny = n[1], nz = n[2];
  const double um = Dn[0], vm = Dn[1], wm = Dn[2], c2 = Dn[3], H = Dn[4];
  const double v2 = um * um + vm * vm + wm * wm;
  const double af = 0.5 * gm1 * v2;
  const double cm = sqrt(c2);
  const double U = um * nx + vm * ny + wm * nz + 0.0;
  const double r = fabs(U) + cm * a.sq;
  const double ct = nx * um + ny * vm + nz * wm + 0.0;
  const double ssr = (double)sgn * r;
  const double l1 = 0.5 * (ct + ssr);
  const double l4 = 0.5 * ((ct + cm * a.gm) + ssr);
  const double l5 = 0.5 * ((ct - cm * a.gm) + ssr);
  const double x1 = (2.0 * l1 - l4 - l5) / (2.0 * c2);
  const double x2 = (l4 - l5) / (2.0 * cm);
  const double ctn = (ct - 0.0) * a.im;
  const double nxn = nx * a.im, nyn = ny * a.im, nzn = nz * a.im;
  const double drm = d[0], dum = d[1], dvm = d[2], dwm = d[3], dEm = d[4];
  const double dc = ctn * drm - (nxn * dum + nyn * dvm + nzn * dwm);
  const double dh = af * drm - gm1 * (um * dum + vm * dvm + wm * dwm - dEm);
  const double c2dc = c2 * dc;
  res[0] = l1 * drm - dh * x1 - dc * x2;
  res[1] = l1 * dum + (nxn * c2dc - um * dh) * x1 + (nxn * dh - um * dc) * x2;
  res[2] = l1 * dvm + (nyn * c2dc - vm * dh) * x1 + (nyn * dh - vm * dc) * x2;
  res[3] = l1 * dwm + (nzn * c2dc - wm * dh) * x1 + (nzn * dh - wm * dc) * x2;
  res[4] = l1 * dEm + (ctn * c2dc - H * dh) * x1 + (ctn * dh - H * dc) * x2;
}

// ---------------------------------------------------------------------------
// tile sweep: progress keys, waits and loads
// ---------------------------------------------------------------------------
constexpr int NW = 11;  // exchanged words per cell: D (6) + dq (5)

__device__ inline unsigned long long plane_key(bool fwd, int s) {
  return fwd ? (unsigned long long)(s + 1) : (unsigned long long)((1 << KEY_BITS) - s);
}

// progress a tile must have published before its cells on plane s may be
// read (0: the tile has none on or before s in sweep order)
__device__ inline unsigned long long need_key(bool fwd, const SweepTile& u, int s) {
  if (fwd) {
    if (s < u.s_first) return 0;
    return plane_key(true, s < u.s_last ? s : u.s_last);
  }
  if (s > u.s_last) return 0;
  return plane_key(false, s > u.s_first ? s : u.s_first);
}

// s_waitcnt immediate that waits until at most n vector memory operations
// of the wave are outstanding (GFX9 encoding; expcnt and lgkmcnt at their
// maxima, i.e. not waited on)
constexpr unsigned waitcnt_vm(unsigned n) { return (n & 15u) | ((n >> 4) << 14) | (7u << 4) | (15u << 8); }

// Sweep layout pair loads: 16-byte loads of two fields of a cell.  (rb, c):
// the cell's row base q - c and column.
__device__ __forceinline__ double2 ld_pair(const double2* g, long long rb, int c, int ncp, int K, int k) {
  return g[rb * K + (long long)k * ncp + c];
}

// a pair whose second half is padding is loaded as one double (the
// register of a loaded but unused half could be reused while the load is in
// flight, which costs a wait)
__device__ __forceinline__ double ld_half(const double2* g, long long rb, int c, int ncp, int K, int k) {
  return reinterpret_cast<const double*>(g)[2 * (rb * K + (long long)k * ncp + c)];
}

// Direct global-to-LDS load of one sweep-layout pair per lane: lane l of
// the wave writes its 16 bytes to wbase[l], where wbase (the wave's first
// slot) is the same for every lane of the wave.  Lanes that are switched off
// write nothing.
__device__ __forceinline__ void dma_pair(const double2* g, long long rb, int c, int ncp, int K, int k,
                                         double2* wbase) {
  __builtin_amdgcn_global_load_lds((const void*)(g + rb * K + (long long)k * ncp + c),
                                   (__attribute__((address_space(3))) void*)wbase, 16, 0, 0);
}

// A block's share of a sweep array (dq, and in the chained launch ga and
// pre) as a buffer resource, for 16-byte stores and loads with the cache
// policy of agent-scope atomics (sc1: written through to, and read from,
// where every XCD sees it).  Base: the block's first cell; the host checks
// that a block's share stays below 4 GiB, so byte offsets fit 32 bits.
// Word 3 is the gfx9 raw-buffer format (32-bit data, no swizzle).
typedef unsigned v4u32 __attribute__((ext_vector_type(4)));
typedef double v2f64 __attribute__((ext_vector_type(2)));
// a block's share of a sweep array of K pairs per cell
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pair_resource(const double2* arr, const BlockDev& b, int K) {
  const long long bytes = (long long)(b.sn[0] + b.sn[1] + b.sn[2] - 2) * b.sn[2] * b.sncp * K * 16;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double2*>(arr) + b.sbase * K, 0, (int)(unsigned)bytes,
                                           0x00020000);
}
// byte offset of pair k of cell (q, c) in the block's share of an array of K pairs
__device__ __forceinline__ unsigned pair_off(const BlockDev& b, long long q, int c, int ncp, int K, int k) {
  return (unsigned)(((q - c - b.sbase) * K + (long long)k * ncp + c) * 16);
}
// one cell's pair (x, y) at byte offset off of the resource, coherent for the device
__device__ __forceinline__ void store_pair_coherent(__amdgpu_buffer_rsrc_t r, unsigned off, double x, double y) {
  const v2f64 v = {x, y};
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u32, v), r, off, 0, CPOL_SC1);
}
// and the coherent load of one
__device__ __forceinline__ double2 load_pair_coherent(__amdgpu_buffer_rsrc_t r, unsigned off) {
  const v2f64 v = __builtin_bit_cast(v2f64, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, CPOL_SC1));
  return make_double2(v.x, v.y);
}

// the same with the cache policy of an agent-scope atomic load (sc1: the
// load is served coherently for the whole device), for values that other
// workgroups store while this one runs
__device__ __forceinline__ void dma_pair_coherent(const double2* g, long long rb, int c, int ncp, int K, int k,
                                                  double2* wbase) {
  __builtin_amdgcn_global_load_lds((const void*)(g + rb * K + (long long)k * ncp + c),
                                   (__attribute__((address_space(3))) void*)wbase, 16, 0, CPOL_SC1);
}

// ---------------------------------------------------------------------------
// tile sweep
// ---------------------------------------------------------------------------
// Sweep types: 0 LU forward, 1 LU backward, 2 Gauss-Seidel forward,
// 3 Gauss-Seidel backward.
//
// In every sweep a cell's update sums six neighbour terms in the order of
// slots 0..5 (the -1 neighbours along i, j, k, then the +1 ones).  The
// terms of the neighbours before the cell in sweep order use their new dq
// and are computed here; for the Gauss-Seidel sweeps the other three use
// values that an earlier sweep already finished, and they come in ready:
//  * Gauss-Seidel forward: its after-neighbours (slots 3..5) carry the LU
//    backward sweep's final dq, and the LU backward sweep evaluates exactly
//    those terms for the same cell pairs (its own before-neighbours).  It
//    stores them (with the Gauss-Seidel sign of the viscous part) in ga,
//    zero for a missing neighbour, so that r - 0 leaves r unchanged.
//  * Gauss-Seidel backward: its sum starts with slots 0..2, whose
//    neighbours carry the forward sweep's final dq, which is what the
//    forward sweep used for them: the forward sweep's partial sum after its
//    first three slots is stored in pre and is the backward sweep's start.
//
// Pipeline of one tile, with the memory traffic kept off the update waves:
// a tile's workgroup has 256 update lanes and 256 loader lanes, lane l of
// either kind working on pencil l.  Step t works on hyperplane
// s(t) = s0 + dir*t:
//  * update lanes: the cell's inputs, the neighbours' (D, dq) and the
//    upstream edge values come from LDS; the new dq goes to LDS;
//  * loader lanes: store the dq of step t-1 (agent scope; for the LU
//    backward sweep first the Gauss-Seidel terms the update lanes left in
//    the own-input slot of step t-1), then load the own inputs of step t+1
//    and the D of step t straight into LDS (global_load_lds, no registers);
//    the first loader wave also puts the edge values of step t+1 (fetched a
//    step earlier) into LDS, checks the upstream progress, fetches the edge
//    values of step t+2 and reads the upstream flags for the next check,
//    and its lane 0 publishes;
//  * each loader wave waits for its direct loads; one barrier ends the step.
// Plane s(t) is published at the start of step t+2: each loader wave's
// stores of it precede the direct loads it waited for at the end of step
// t+1 (the counter is in order), and a barrier followed.  LDS: own inputs
// by step parity, D by step parity (a step's update reads the D of the step
// before while the loaders fill the step's own), dq by parity, edge values
// by parity; own inputs and D are kept as the global pairs.
// LDS of one tile by sweep type: own inputs (three step buffers, two in the
// Gauss-Seidel forward sweep), D by step parity, new dq (and the Gauss-Seidel
// forward partial sums) by parity, edge D and edge dq by step mod 3.
__host__ __device__ constexpr int sweep_nob(int type) { return type != 2 ? 3 : 2; }
__host__ __device__ constexpr int sweep_nq(int type) { return type == 2 ? 10 : 5; }
__host__ __device__ constexpr unsigned sweep_lds_bytes(int type) {
  return (unsigned)(16 * sweep_nob(type) * SW_OWN * TL + 16 * 2 * SW_D * TL + 8 * 2 * sweep_nq(type) * TL +
                    16 * 2 * 3 * 128);
}
constexpr unsigned SWEEP_LDS_MAX = sweep_lds_bytes(0) > sweep_lds_bytes(2) ? sweep_lds_bytes(0) : sweep_lds_bytes(2);

// One tile's sweep of type TYPE.  e4: four times the step epoch (the sweep
// serial base).  CHAIN: the tile runs in the chained launch of all four
// sweeps, where its sweep TYPE waits for its own sweep TYPE - 1 to have
// published its last plane, and the values one sweep leaves for the next
// (the forward dq, ga, pre) are stored and loaded with the device-coherent
// policy, as the two sweeps of a tile may run on different XCDs.
template <int TYPE, bool VIS, bool CHAIN>
__device__ __forceinline__ void sweep_tile(const SweepArrays& S, const BlockDev* __restrict__ blocks,
                                           const SweepTile* __restrict__ tiles, int tix, unsigned long long* flags,
                                           unsigned long long e4, int* err, double gm1, unsigned char* lds) {
  constexpr bool FWD = (TYPE == 0 || TYPE == 2);
  constexpr int dir = FWD ? 1 : -1;
  constexpr int SL0 = FWD ? 0 : 3;
  constexpr int NO = 15;  // own input words
  // own inputs loaded two steps ahead (three buffers), except in the
  // Gauss-Seidel forward sweep, whose LDS holds one buffer less
  constexpr bool OWN2 = TYPE != 2;
  constexpr int NOB = sweep_nob(TYPE);
  constexpr int NQ = sweep_nq(TYPE);  // new dq, then (Gauss-Seidel forward) the partial sums
  constexpr unsigned OFF_D = 16u * NOB * SW_OWN * TL;
  constexpr unsigned OFF_Q = OFF_D + 16u * 2 * SW_D * TL;
  constexpr unsigned OFF_XD = OFF_Q + 8u * 2 * NQ * TL;
  constexpr unsigned OFF_XQ = OFF_XD + 16u * 3 * 128;
  double2(*ownP)[SW_OWN][TL] = reinterpret_cast<double2(*)[SW_OWN][TL]>(lds);
  double2(*Dp)[SW_D][TL] = reinterpret_cast<double2(*)[SW_D][TL]>(lds + OFF_D);
  double(*dqB)[NQ][TL] = reinterpret_cast<double(*)[NQ][TL]>(lds + OFF_Q);
  double2(*shXD)[128] = reinterpret_cast<double2(*)[128]>(lds + OFF_XD);  // edge D: [step mod 3][item] (96 used)
  double2(*shXQ)[128] = reinterpret_cast<double2(*)[128]>(lds + OFF_XQ);  // edge dq: [step mod 3][item] (96 used)

  auto ob = [](int u) { return OWN2 ? (u + 3) % 3 : u & 1; };  // own buffer of step u >= -1
  const SweepTile T = tiles[tix];
  const BlockDev& b = blocks[T.block];
  const int nm = b.sn[0], nc = b.sn[1], nr = b.sn[2];
  const int am = b.sax[0], ac = b.sax[1];
  const bool loader = threadIdx.x >= TL;
  const bool lw0 = loader && threadIdx.x < TL + 64;  // first loader wave
  const int lane = threadIdx.x & (TL - 1), lc = lane % TW, lr = lane / TW;
  const int xl = threadIdx.x & 63;
  const int wb = lane & ~63;  // the wave's first pencil: LDS base of its direct loads
  const int c = T.c0 + lc, r = T.r0 + lr;
  const bool lane_ok = c < nc && r < nr;
  const int up_c = FWD ? T.nbr[0] : T.nbr[2], up_r = FWD ? T.nbr[1] : T.nbr[3];
  // the sweep's serial: four sweeps per step epoch (the update kernel after
  // them advances the epoch), so serials rise from launch to launch and a
  // flag left by an earlier sweep is always short of this one's keys
  const unsigned long long ep = (e4 + TYPE + 1) << KEY_BITS;
  SweepTile ut[2];
  ut[0] = up_c >= 0 ? tiles[up_c] : T;
  ut[1] = up_r >= 0 ? tiles[up_r] : T;
  const int nsteps = T.s_last - T.s_first + 1;
  const int s0 = FWD ? T.s_first : T.s_last;
  const int ncp = b.sncp;
  const __amdgpu_buffer_rsrc_t dq_rsrc = pair_resource(S.dq, b, SW_DQ);
  // the chained launch's cross-sweep arrays (unused otherwise)
  const __amdgpu_buffer_rsrc_t ga_rsrc = pair_resource(S.ga, b, SW_GA);
  const __amdgpu_buffer_rsrc_t pre_rsrc = pair_resource(S.pre, b, SW_PRE);
  const long long qstep = dir * (long long)nr * ncp;
  const int m0 = s0 - r - c;
  const long long rb0 = b.sbase + ((long long)s0 * nr + r) * ncp;
  auto cell_ok = [&](int t) {
    const int m = m0 + dir * t;
    return lane_ok && t >= 0 && t < nsteps && m >= 0 && m < nm;
  };
  auto cell_rb = [&](int t) { return cell_ok(t) ? rb0 + (long long)t * qstep : b.sbase; };
  auto cell_c = [&](int t) { return cell_ok(t) ? c : 0; };

  // Edge values, loaded by the first loader wave straight into LDS: the D
  // pairs and the dq pairs (three pairs each per cell; items x and x + 64 of
  // each, those below 96: role i / 48, index (i % 48) / 3, pair i % 3) of
  // the cells across the upstream col edge (role 0) and row edge (role 1).
  // shXD[b][item] and shXQ[b][item] put lane x of each load at 16 x bytes
  // from the load's base; item h of the lane is x + 64 h.
  int er[2], ec[2];
  bool ein[2];
  int ek[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int i = xl + 64 * h;
    bool item = i < 96;
    const int role = i < 48 ? 0 : 1;
    const int idx = (i % 48) / 3;
    ek[h] = i % 3;
    if (role == 0) {  // the col edge has TR cells
      item = item && idx < TR;
      ec[h] = FWD ? T.c0 - 1 : T.c0 + TW;
      er[h] = T.r0 + idx;
    } else {
      er[h] = FWD ? T.r0 - 1 : T.r0 + TR;
      ec[h] = T.c0 + idx;
    }
    ein[h] = lw0 && item && er[h] >= 0 && er[h] < nr && ec[h] >= 0 && ec[h] < nc;
  }
  // the edge cell of item h for step u lies on plane s(u-1); out-of-range
  // items read the block's first cell, which no lane reads as a neighbour
  auto edge_rb = [&](int h, int u, int* cc) -> long long {
    const int sp = s0 + dir * (u - 1);
    const int mm = sp - er[h] - ec[h];
    const bool ok = ein[h] && u < nsteps && mm >= 0 && mm < nm;
    *cc = ok ? ec[h] : 0;
    return ok ? b.sbase + ((long long)sp * nr + er[h]) * ncp : b.sbase;
  };
  // first loader wave: the edge values of step u into LDS, four loads on
  // every lane (items 96..127 read the block's first cell into slots no lane
  // reads); dq with the coherent cache policy, as it was stored by other
  // tiles during this sweep
  auto dma_edges = [&](int u) {
    const int e = u % 3;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      int cc;
      const long long rb = edge_rb(h, u, &cc);
      dma_pair(S.D, rb, cc, ncp, SW_D, ek[h], &shXD[e][64 * h]);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      int cc;
      const long long rb = edge_rb(h, u, &cc);
      dma_pair_coherent(S.dq, rb, cc, ncp, SW_DQ, ek[h], &shXQ[e][64 * h]);
    }
  };
  // the own inputs of the cell of step u into ownP[ob(u)], as pairs:
  // met0..8, coed, r5 (rhs; for the LU backward sweep the cell's dq from the
  // forward sweep).  Eight loads on every lane: a lane without a cell reads
  // the block's first cell into its own slot, which nothing reads.
  auto dma_own = [&](int u) {
    const long long rb = cell_rb(u);
    const int cc = cell_c(u);
    double2(*dst)[TL] = ownP[ob(u)];
    if constexpr (TYPE == 1) {
#pragma unroll
      for (int k = 0; k < 5; ++k) dma_pair(S.own, rb, cc, ncp, SW_OWN, k, &dst[k][wb]);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if constexpr (CHAIN)
          dma_pair_coherent(S.dq, rb, cc, ncp, SW_DQ, k, &dst[5 + k][wb]);
        else
          dma_pair(S.dq, rb, cc, ncp, SW_DQ, k, &dst[5 + k][wb]);
      }
    } else {
#pragma unroll
      for (int k = 0; k < SW_OWN; ++k) dma_pair(S.own, rb, cc, ncp, SW_OWN, k, &dst[k][wb]);
    }
  };
  auto dma_D = [&](int u) {
    const long long rb = rb0 + (long long)u * qstep;
#pragma unroll
    for (int k = 0; k < SW_D; ++k) dma_pair(S.D, rb, c, ncp, SW_D, k, &Dp[u & 1][k][wb]);
  };
  // first loader wave: the flags for the next check, two loads on every
  // lane (a missing upstream reads this tile's own flag and is never
  // checked)
  auto read_up = [&](unsigned long long* pf) {
    pf[0] = __hip_atomic_load(&flags[(long long)(up_c >= 0 ? up_c : tix) * FLAG_STRIDE], __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
    pf[1] = __hip_atomic_load(&flags[(long long)(up_r >= 0 ? up_r : tix) * FLAG_STRIDE], __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
  };
  // first loader wave: the upstream tiles must have finished plane s; pf
  // holds flag values read a step earlier, re-read (bounded spin) only when
  // short.  Every lane of the wave runs the check (all read the same flag
  // words), so the flag registers are consumed on every path.
  auto ensure_up = [&](int s, unsigned long long* pf) {
    for (int k = 0; k < 2; ++k) {
      const int u = k ? up_r : up_c;
      if (u < 0) continue;
      const unsigned long long key = need_key(FWD, ut[k], s);
      if (key == 0) continue;
      const unsigned long long want = ep + key;
      unsigned long long v = pf[k];
      long long spins = 0;
      while (v < want) {
        // far beyond any legal wait (a whole sweep takes milliseconds), or
        // another tile already gave up: report and carry on, never hang
        if (++spins > (1LL << 20) ||
            ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
          if (xl == 0) atomicOr(err, 1);
          break;
        }
        if (spins > 1) __builtin_amdgcn_s_sleep(2);
        v = __hip_atomic_load(&flags[(long long)u * FLAG_STRIDE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      pf[k] = v;
    }
  };

  // before-neighbour of the lane's cell along real axis ax at step t: its
  // (D, dq) and whether it exists
  auto nbr = [&](int ax, int m, int t, double* v) -> bool {
    constexpr int d = -dir;
    const int p = t & 1;
    const bool march = (ax == am);
    const bool col = (ax == ac);
    const int pos = col ? lc : lr, lim = col ? nc : nr, coord = col ? c : r;
    const bool inside = march || (pos + d >= 0 && pos + d < (col ? TW : TR));
    const int ln = march ? lane : lane + (col ? d : d * TW);
    const int role = col ? 0 : 1, idx = col ? lr : lc;
    const int e = t % 3;
    const double2* bd = inside ? &Dp[p ^ 1][0][ln] : &shXD[e][(role * TW + idx) * SW_D];
    const double* bq = inside ? &dqB[p ^ 1][0][ln] : reinterpret_cast<const double*>(&shXQ[e][(role * TW + idx) * 3]);
    const int st = inside ? TL : 1;
#pragma unroll
    for (int k = 0; k < SW_D; ++k) {
      const double2 x = bd[k * st];
      v[2 * k] = x.x;
      v[2 * k + 1] = x.y;
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) v[2 * SW_D + f] = bq[f * st];
    return march ? (m + d >= 0 && m + d < nm) : (coord + d >= 0 && coord + d < lim);
  };

  // update lanes, a step ahead into registers: the Gauss-Seidel forward
  // sweep's ga terms, the Gauss-Seidel backward sweep's starting sums
  auto load_ga = [&](int u, double* g) {
    if constexpr (TYPE == 3) {
      const long long rb = cell_rb(u);
      const int cc = cell_c(u);
      if constexpr (CHAIN) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double2 v = load_pair_coherent(pre_rsrc, pair_off(b, rb + cc, cc, ncp, SW_PRE, k));
          g[2 * k] = v.x;
          if (k < 2) g[2 * k + 1] = v.y;
        }
      } else {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          const double2 v = ld_pair(S.pre, rb, cc, ncp, SW_PRE, k);
          g[2 * k] = v.x;
          g[2 * k + 1] = v.y;
        }
        g[4] = ld_half(S.pre, rb, cc, ncp, SW_PRE, 2);
      }
    }
    if constexpr (TYPE == 2) {
      const long long rb = cell_rb(u);
      const int cc = cell_c(u);
      if constexpr (CHAIN) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const double2 v = load_pair_coherent(ga_rsrc, pair_off(b, rb + cc, cc, ncp, SW_GA, k));
          g[2 * k] = v.x;
          if (k < 7) g[2 * k + 1] = v.y;
        }
      } else {
#pragma unroll
        for (int k = 0; k < 7; ++k) {
          const double2 v = ld_pair(S.ga, rb, cc, ncp, SW_GA, k);
          g[2 * k] = v.x;
          g[2 * k + 1] = v.y;
        }
        g[14] = ld_half(S.ga, rb, cc, ncp, SW_GA, 7);
      }
    }
  };

  // one step of a loader lane; W0: the first loader wave, which runs its
  // own loop and also keeps the tile's progress: it publishes, checks the
  // upstream flags it read in the step before, and loads the edge values of
  // step t+2 last, so that they stay in flight across the barrier.
  auto lstep = [&](auto w0, int t, unsigned long long* pf) __attribute__((always_inline)) {
    constexpr bool W0 = decltype(w0)::value;
    const int s = s0 + dir * t;
    if constexpr (W0) {
      // plane s(t-2): every loader wave's stores of it preceded the direct
      // loads that the wave waited for before the barrier ending step t-1
      if (xl == 0 && t >= 2 && t - 2 < nsteps)
        __hip_atomic_store(&flags[(long long)tix * FLAG_STRIDE], ep + plane_key(FWD, s - 2 * dir), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      if (t + 2 < nsteps) ensure_up(s + dir, pf);  // the edge values of step t+2 lie on plane s(t+1)
    }
    if (TYPE == 1 && cell_ok(t - 1)) {  // the terms the update lanes left in the own-input slot of step t-1
      const int q = ob(t - 1);
      if constexpr (CHAIN) {
        const long long qc = rb0 + (long long)(t - 1) * qstep + c;
#pragma unroll
        for (int k = 0; k < SW_GA; ++k) {
          const double2 v = ownP[q][k][lane];
          store_pair_coherent(ga_rsrc, pair_off(b, qc, c, ncp, SW_GA, k), v.x, v.y);
        }
      } else {
        double2* gp = S.ga + (rb0 + (long long)(t - 1) * qstep) * SW_GA + c;
#pragma unroll
        for (int k = 0; k < SW_GA; ++k) gp[(long long)k * ncp] = ownP[q][k][lane];
      }
    }
    if (cell_ok(t - 1)) {
      const long long q = rb0 + (long long)(t - 1) * qstep + c;
      // the cell's three dq pairs (the last one's second half is padding)
      // as whole 16-byte stores: half-pair stores written through cost
      // twice the write traffic
      const unsigned off = pair_off(b, q, c, ncp, SW_DQ, 0);
      const unsigned kst = (unsigned)ncp * 16u;
      const double* w = &dqB[(t - 1) & 1][0][lane];
      store_pair_coherent(dq_rsrc, off, w[0], w[TL]);
      store_pair_coherent(dq_rsrc, off + kst, w[2 * TL], w[3 * TL]);
      store_pair_coherent(dq_rsrc, off + 2u * kst, w[4 * TL], 0.0);
      if constexpr (TYPE == 2) {  // the partial sums the update lanes left beside the new dq
        const double* w = &dqB[(t - 1) & 1][5][lane];
        if constexpr (CHAIN) {
          store_pair_coherent(pre_rsrc, pair_off(b, q, c, ncp, SW_PRE, 0), w[0], w[TL]);
          store_pair_coherent(pre_rsrc, pair_off(b, q, c, ncp, SW_PRE, 1), w[2 * TL], w[3 * TL]);
          store_pair_coherent(pre_rsrc, pair_off(b, q, c, ncp, SW_PRE, 2), w[4 * TL], 0.0);
        } else {
          double2* pp = S.pre + (q - c) * SW_PRE + c;
          pp[0] = make_double2(w[0], w[TL]);
          pp[ncp] = make_double2(w[2 * TL], w[3 * TL]);
          pp[2 * ncp] = make_double2(w[4 * TL], 0.0);
        }
      }
    }
    // the stores above took their values from LDS before they issued, so
    // the direct loads may now refill the own buffer of step t-1
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (W0) read_up(pf);  // checked at the next step
    if (cell_ok(t)) dma_D(t);
    if constexpr (OWN2) {
      dma_own(t + 2);
    } else {
      if (cell_ok(t + 1)) dma_own(t + 1);
    }
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (W0) dma_edges(t + 2);
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    // Loads that may stay in flight across the barrier: the own inputs of
    // step t+2 (eight loads) and the first wave's edge values of step t+2
    // (four), the last issued.  Everything before them has completed (the
    // counter is in order): the stores, the flags, the D of step t and the
    // own inputs of step t+1 (issued a step earlier); LDS accesses too.  The
    // barrier carries no fence, which would wait for every direct load.
    constexpr unsigned NFLY = (OWN2 ? SW_OWN : 0) + (W0 ? 4 : 0);
    if constexpr (NFLY > 0) {
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(NFLY) & ~(15u << 8));
      __builtin_amdgcn_sched_barrier(0);
      __builtin_amdgcn_s_barrier();
      asm volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
    } else {
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
      __builtin_amdgcn_sched_barrier(0);
      __syncthreads();
    }
  };

  // one step of an update lane
  auto ustep = [&](int t, double* gcur, double* gnext) __attribute__((always_inline)) {
    if (t < nsteps) {
      const int m = m0 + dir * t;
      const bool act = cell_ok(t);
      double o[NO];
#pragma unroll
      for (int k = 0; k < SW_OWN; ++k) {
        const double2 x = ownP[ob(t)][k][lane];
        o[2 * k] = x.x;
        if (2 * k + 1 < 15) o[2 * k + 1] = x.y;
      }
      double r0[5];
#pragma unroll
      for (int f = 0; f < 5; ++f) r0[f] = TYPE == 3 ? gcur[f] : 0.0;
      double gout[16];
      gout[15] = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        AxisGeo ag;
        axis_geo(o + 3 * k, ag);
        double v[NW];
        const bool ok = nbr(k, m, t, v);
        double inv[5], tv[5];
        slot_inv<SL0>(o + 3 * k, ag, v, v + 2 * SW_D, gm1, inv);
        if (VIS) {
          const double hv = 0.5 * (v[2 * SW_D - 1] * ag.g2);
#pragma unroll
          for (int f = 0; f < 5; ++f) tv[f] = hv * v[2 * SW_D + f];
        }
#pragma unroll
        for (int f = 0; f < 5; ++f) {
          if (TYPE == 3) {  // Gauss-Seidel term of a +1 neighbour: r - (inviscid - viscous)
            const double res = VIS ? inv[f] + tv[f] * -1.0 : inv[f];
            const double acc = r0[f] - res;
            r0[f] = ok ? acc : r0[f];
          } else {  // LU terms, and Gauss-Seidel terms of -1 neighbours: r + (inviscid + viscous)
            const double res = VIS ? inv[f] + tv[f] : inv[f];
            const double acc = r0[f] + res;
            r0[f] = ok ? acc : r0[f];
          }
          if (TYPE == 1) {  // the Gauss-Seidel forward sweep's term for the same pair
            const double g = VIS ? inv[f] + tv[f] * -1.0 : inv[f];
            gout[5 * k + f] = ok ? g : 0.0;
          }
        }
      }
      double pre5[5];
      if (TYPE == 2) {
#pragma unroll
        for (int f = 0; f < 5; ++f) pre5[f] = r0[f];
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int f = 0; f < 5; ++f) r0[f] = r0[f] - gcur[5 * k + f];
      }
      const double wmig = 1.0;
      const double* r5 = o + 10;
      const double coed = o[9];
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        double x;
        if (TYPE == 0) x = (-r5[f] + wmig * r0[f]) * coed;
        else if (TYPE == 1) x = r5[f] - (wmig * r0[f]) * coed;
        else x = (wmig * r0[f] + -r5[f]) * coed;
        dqB[t & 1][f][lane] = act ? x : 0.0;
      }
      if (TYPE == 1) {  // into the lane's own-input slot (read above); the loaders store it
#pragma unroll
        for (int k = 0; k < SW_GA; ++k) ownP[ob(t)][k][lane] = make_double2(gout[2 * k], gout[2 * k + 1]);
      }
      if constexpr (TYPE == 2) {  // the loaders store them with the dq
#pragma unroll
        for (int f = 0; f < 5; ++f) dqB[t & 1][5 + f][lane] = pre5[f];
      }
      load_ga(t + 1, gnext);
    }
    __syncthreads();
  };

  // Chained launch: the tile's previous sweep must have published its last
  // plane (its last stores had completed before that publication, as for
  // every plane) before anything of it is read; the first loader wave waits
  // (bounded spin, as for the upstream tiles) and a barrier releases all.
  if constexpr (CHAIN && TYPE > 0) {
    if (lw0) {
      constexpr bool PFWD = (TYPE - 1) == 0 || (TYPE - 1) == 2;
      const unsigned long long want =
          ((e4 + TYPE) << KEY_BITS) + (PFWD ? plane_key(true, T.s_last) : plane_key(false, T.s_first));
      unsigned long long v =
          __hip_atomic_load(&flags[(long long)tix * FLAG_STRIDE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      long long spins = 0;
      while (v < want) {
        if (++spins > (1LL << 20) ||
            ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
          if (xl == 0) atomicOr(err, 2);
          break;
        }
        if (spins > 1) __builtin_amdgcn_s_sleep(2);
        v = __hip_atomic_load(&flags[(long long)tix * FLAG_STRIDE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    __syncthreads();
  }

  // Both kinds of lanes run their own loop with one barrier per step (and
  // one after the set-up), so every barrier is reached by all waves.
  const int TT = (nsteps + 3) & ~1;  // whole pairs of steps, at least nsteps + 2
  if (loader) {
    unsigned long long pf[2] = {0, 0};
    if (lw0) {
      ensure_up(s0 - dir, pf);  // the edge values of step 0 lie on plane s(-1)
      dma_edges(0);
      ensure_up(s0, pf);  // those of step 1 on plane s(0)
      dma_edges(1);
      read_up(pf);
    }
    dma_own(0);
    if constexpr (OWN2) dma_own(1);
    __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    __syncthreads();
    if (lw0) {
      for (int t = 0; t < TT; ++t) lstep(std::true_type{}, t, pf);
    } else {
      for (int t = 0; t < TT; ++t) lstep(std::false_type{}, t, pf);
    }
    __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
  } else {
    double gaR[2][TYPE == 2 ? 15 : TYPE == 3 ? 5 : 1];  // ga (GS forward) or starting sums (GS backward) by step parity
#pragma unroll
    for (int f = 0; f < 5; ++f) dqB[1][f][lane] = 0.0;
#pragma unroll
    for (int k = 0; k < SW_D; ++k) Dp[1][k][lane] = make_double2(0.0, 0.0);
    load_ga(0, gaR[0]);
    __syncthreads();
    for (int t = 0; t < TT; t += 2) {
      ustep(t, gaR[0], gaR[1]);
      ustep(t + 1, gaR[1], gaR[0]);
    }
  }
}

// One launch per sweep: tile order[w] on workgroup w.
template <int TYPE, bool VIS>
__global__ void __launch_bounds__(2 * TL, 1) lusgs_tile_kernel(SweepArrays S, const BlockDev* __restrict__ blocks,
                                                            const SweepTile* __restrict__ tiles,
                                                            const int* __restrict__ order, unsigned long long* flags,
                                                            const unsigned long long* __restrict__ epoch, int* err,
                                                            double gm1) {
  __shared__ __align__(16) unsigned char lds[sweep_lds_bytes(TYPE)];
  sweep_tile<TYPE, VIS, false>(S, blocks, tiles, order[blockIdx.x], flags, (*epoch) * 4, err, gm1, lds);
}

// The four sweeps as one chained launch: workgroup w runs task tasks[w] =
// tile << 2 | sweep type.  The task order is topological (a tile's sweep
// after its previous sweep, and after the tiles it waits for in its sweep),
// so every wait targets a task launched earlier and the grid drains; a tile
// whose blocks' smaller neighbours finished early starts its next sweep
// while larger blocks are still in an earlier one.
template <bool VIS>
__global__ void __launch_bounds__(2 * TL, 1) lusgs_chain_kernel(SweepArrays S, const BlockDev* __restrict__ blocks,
                                                             const SweepTile* __restrict__ tiles,
                                                             const int* __restrict__ tasks, unsigned long long* flags,
                                                             const unsigned long long* __restrict__ epoch, int* err,
                                                             double gm1) {
  __shared__ __align__(16) unsigned char lds[SWEEP_LDS_MAX];
  const int task = tasks[blockIdx.x];
  const int tix = task >> 2;
  const unsigned long long e4 = (*epoch) * 4;
  switch (task & 3) {
    case 0: sweep_tile<0, VIS, true>(S, blocks, tiles, tix, flags, e4, err, gm1, lds); break;
    case 1: sweep_tile<1, VIS, true>(S, blocks, tiles, tix, flags, e4, err, gm1, lds); break;
    case 2: sweep_tile<2, VIS, true>(S, blocks, tiles, tix, flags, e4, err, gm1, lds); break;
    default: sweep_tile<3, VIS, true>(S, blocks, tiles, tix, flags, e4, err, gm1, lds); break;
  }
}

void launch_lusgs_chain(const SweepArrays& S, const BlockDev* blocks, const SweepTile* tiles, const int* tasks,
                        int ntasks, unsigned long long* flags, unsigned long long* epoch, int* err, const Params& p,
                        hipStream_t s) {
  if (ntasks <= 0) return;
  const double gm1 = p.gama - 1.0;
  if (p.nvis == 1)
    hipLaunchKernelGGL((lusgs_chain_kernel<true>), dim3(ntasks), dim3(2 * TL), 0, s, S, blocks, tiles, tasks, flags,
                       epoch, err, gm1);
  else
    hipLaunchKernelGGL((lusgs_chain_kernel<false>), dim3(ntasks), dim3(2 * TL), 0, s, S, blocks, tiles, tasks, flags,
                       epoch, err, gm1);
}

void launch_lu_geom(const Launch& La, long long n, const SweepArrays& S) {
  if (n <= 0) return;
  hipLaunchKernelGGL(lu_geom_kernel, dim3((unsigned)La.nwg), dim3(SPAN_TB), 0, La.stream, La.blocks, La.span, S);
}

void launch_lu_pack(const Launch& La, long long n, const SweepArrays& S, int part) {
  if (n <= 0) return;
  if (part == 1)
    hipLaunchKernelGGL(lu_pack_kernel<1>, dim3((unsigned)La.nwg), dim3(SPAN_TB), 0, La.stream, La.blocks, La.span, S);
  else
    hipLaunchKernelGGL(lu_pack_kernel<2>, dim3((unsigned)La.nwg), dim3(SPAN_TB), 0, La.stream, La.blocks, La.span, S);
}

void launch_state_pack(const Launch& La, long long n, const SweepArrays& S, unsigned long long* dtmin_base,
                       const int* step, int nreq) {
  if (n <= 0) return;
  hipLaunchKernelGGL(state_pack_kernel, dim3((unsigned)La.nwg), dim3(SPAN_TB), 0, La.stream, La.blocks, La.span, S,
                     dtmin_base, step, nreq);
}

void launch_update(const Launch& La, long long n, const SweepArrays& S, double gamma, const long long* ioff,
                   double* psum, double* pmax, long long* pkey, unsigned long long* epoch) {
  if (n <= 0) return;
  hipLaunchKernelGGL(update_kernel, dim3((unsigned)La.nwg), dim3(SPAN_TB), 0, La.stream, La.blocks, La.span, S,
                     gamma, ioff, psum, pmax, pkey, epoch);
}

void launch_lusgs_tiles(const SweepArrays& S, const BlockDev* blocks, const SweepTile* tiles, const int* order,
                        int ntiles, unsigned long long* flags, unsigned long long* epoch, int* err, int type,
                        const Params& p, hipStream_t s) {
  if (ntiles <= 0) return;
  const double gm1 = p.gama - 1.0;
  const dim3 gr(ntiles);
  const bool v = p.nvis == 1;
  const dim3 tb2(2 * TL);
  switch (type * 2 + (v ? 1 : 0)) {
    case 0: hipLaunchKernelGGL((lusgs_tile_kernel<0, false>), gr, tb2, 0, s, S, blocks, tiles, order, flags, epoch, err, gm1); break;
    case 1: hipLaunchKernelGGL((lusgs_tile_kernel<0, true>), gr, tb2, 0, s, S, blocks, tiles, order, flags, epoch, err, gm1); break;
    case 2: hipLaunchKernelGGL((lusgs_tile_kernel<1, false>), gr, tb2, 0, s, S, blocks, tiles, order, flags, epoch, err, gm1); break;
    case 3: hipLaunchKernelGGL((lusgs_tile_kernel<1, true>), gr, tb2, 0, s, S, blocks, tiles, order, flags, epoch, err, gm1); break;
    case 4: hipLaunchKernelGGL((lusgs_tile_kernel<2, false>), gr, tb2, 0, s, S, blocks, tiles, order, flags, epoch, err, gm1); break;
    case 5: hipLaunchKernelGGL((lusgs_tile_kernel<2, true>), gr, tb2, 0, s, S, blocks, tiles, order, flags, epoch, err, gm1); break;
    case 6: hipLaunchKernelGGL((lusgs_tile_kernel<3, false>), gr, tb2, 0, s, S, blocks, tiles, order, flags, epoch, err, gm1); break;
    default: hipLaunchKernelGGL((lusgs_tile_kernel<3, true>), gr, tb2, 0, s, S, blocks, tiles, order, flags, epoch, err, gm1); break;
  }
}

}  // namespace orion_hip
