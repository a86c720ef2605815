// k_bc.hip -- block-interface and boundary kernels for gfx950 (FP64).
//
//   halo        exchange_bc (src/solver/HaloExchanger.cpp:17-179) as one
//               gather over a precomputed (destination, source) list of
//               arena offsets: the sender window (face node plus ng-1 inward,
//               tangentially widened by one except at block edges) unpacked
//               in the receiver's own order, last writer kept, built once
//               from the BC topology.  Sources are interior points and
//               destinations ghosts, so the gather has no ordering hazards.
//               Blocks whose windows pair equal point counts move whole
//               points (five fields per entry, halo_point_kernel).
//   bc          apply_physical_bc (src/bc/PhysicalBC.cpp:12-613): one launch
//               per region rank (regions keep file order within a block),
//               one thread per face node.
//   avg_stage   average_interface_residuals phase 1 (:200-282): dq/vol of the
//               partner face node, per staging entry.
//   avg_apply   phase 3 (:330-397): dq <- 0.5*(stage*vol + dq), entries of a
//               face node applied in region order.
#include "dd_pow.h"
#include "orion_dev.h"
#include "orion_kernels.h"

namespace orion_hip {

__global__ void __launch_bounds__(256) halo_kernel(const long long* __restrict__ pairs, long long n,
                                                   double* __restrict__ arena) {
  const long long g = xcd_thread();
  if (g >= n) return;
  arena[pairs[2 * g]] = arena[pairs[2 * g + 1]];
}

// The same gather by whole points: entry g moves the five halo fields of
// one source point to one ghost point (halo_point packs each side's field 0
// offset and field stride), all five loaded before any is stored.
__global__ void __launch_bounds__(256) halo_point_kernel(const unsigned long long* __restrict__ pts, long long n,
                                                         double* __restrict__ arena) {
  const long long g = xcd_thread();
  if (g >= n) return;
  const unsigned long long d = pts[2 * g], s = pts[2 * g + 1];
  const unsigned long long d0 = d & HALO_OFF_MASK, dP = d >> HALO_OFF_BITS;
  const unsigned long long s0 = s & HALO_OFF_MASK, sP = s >> HALO_OFF_BITS;
  double v[5];
#pragma unroll
  for (int m = 0; m < 5; ++m) v[m] = arena[s0 + (unsigned long long)m * sP];
#pragma unroll
  for (int m = 0; m < 5; ++m) arena[d0 + (unsigned long long)m * dP] = v[m];
}

__device__ inline double& PR(const BlockDev& b, int i, int j, int k, int m) {
  return b.prim[(long long)m * b.P + pidx(b, i, j, k)];
}
__device__ inline double MT(const BlockDev& b, int i, int j, int k, int m) {
  return b.met[(long long)m * b.P + pidx(b, i, j, k)];
}

// get_normal (PhysicalBC.cpp:12-44)
__device__ inline void get_normal(const BlockDev& b, int i, int j, int k, int s_nd, double* nx, double* ny,
                                  double* nz) {
  const int o = (s_nd - 1) * 3;
  const double mx = MT(b, i, j, k, o), my = MT(b, i, j, k, o + 1), mz = MT(b, i, j, k, o + 2);
  double len = sqrt(mx * mx + my * my + mz * mz);
  if (len < 1.0e-30) len = 1.0e-30;
  *nx = mx / len;
  *ny = my / len;
  *nz = mz / len;
}

__device__ inline void dif_average(const BlockDev& b, const int* o, int dir, int s_lr) {
  int g[3] = {o[0], o[1], o[2]}, n[3] = {o[0], o[1], o[2]};
  if (s_lr == 1) {
    g[dir] += 1;
    n[dir] -= 1;
  } else {
    g[dir] -= 1;
    n[dir] += 1;
  }
  for (int m = 0; m < 6; ++m) {
    const double vg = PR(b, g[0], g[1], g[2], m), vi = PR(b, n[0], n[1], n[2], m);
    PR(b, o[0], o[1], o[2], m) = 0.5 * (vg + vi);
  }
}

__global__ void __launch_bounds__(256) bc_kernel(const BlockDev* __restrict__ blocks, const BCTask* __restrict__ tasks,
                                                 int ntasks, long long total) {
  const long long g = xcd_thread();
  if (g >= total) return;
  int t = 0;
  {
    int lo = 0, hi = ntasks - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tasks[mid].start <= g) lo = mid;
      else hi = mid - 1;
    }
    t = lo;
  }
  const BCTask T = tasks[t];
  const BlockDev& b = blocks[T.block];
  const Params& p = b.par;  // the inputs of the block's request
  const long long e = g - T.start;
  const int ni = T.hi[0] - T.lo[0] + 1, nj = T.hi[1] - T.lo[1] + 1;
  const int i = T.lo[0] + (int)(e % ni);
  const int j = T.lo[1] + (int)((e / ni) % nj);
  const int k = T.lo[2] + (int)(e / ((long long)ni * nj));
  const int dir = T.s_nd - 1;
  const int lr = T.s_lr;
  const int NGL = NG;
  auto ghost = [&](int dist, int* gg) {
    gg[0] = i;
    gg[1] = j;
    gg[2] = k;
    gg[dir] += (lr == 1) ? dist : -dist;
  };
  auto inner = [&](int dist, int* nn) {
    nn[0] = i;
    nn[1] = j;
    nn[2] = k;
    nn[dir] += (lr == 1) ? -dist : dist;
  };
  switch (T.bctype) {
    case 2:
      if (p.nvis == 1) {
        // apply_viscid_wall (:254-363)
        const double moocp = p.moo * p.moo * p.gama;
        const bool iso = (p.twall > 1.0e-6);
        const double twall_dim = (iso && fabs(p.tref) > 1e-16) ? p.twall / p.tref : p.twall;
        int a1[3], a2[3];
        inner(1, a1);
        inner(2, a2);
        const double p1 = PR(b, a1[0], a1[1], a1[2], 4), p2 = PR(b, a2[0], a2[1], a2[2], 4);
        double p_wall = (4.0 * p1 - p2) / 3.0;
        if (p_wall < 1.0e-6) p_wall = p1;
        double t_wall;
        if (iso) {
          t_wall = twall_dim;
        } else {
          const double t1 = PR(b, a1[0], a1[1], a1[2], 5), t2 = PR(b, a2[0], a2[1], a2[2], 5);
          t_wall = (4.0 * t1 - t2) / 3.0;
          if (t_wall < 1e-6) t_wall = t1;
        }
        const double rho_wall = moocp * p_wall / t_wall;
        for (int l = 1; l <= NGL; ++l) {
          int gg[3], nn[3];
          ghost(l, gg);
          inner(l, nn);
          PR(b, gg[0], gg[1], gg[2], 1) = -PR(b, nn[0], nn[1], nn[2], 1);
          PR(b, gg[0], gg[1], gg[2], 2) = -PR(b, nn[0], nn[1], nn[2], 2);
          PR(b, gg[0], gg[1], gg[2], 3) = -PR(b, nn[0], nn[1], nn[2], 3);
          PR(b, gg[0], gg[1], gg[2], 4) = PR(b, nn[0], nn[1], nn[2], 4);
          if (iso) PR(b, gg[0], gg[1], gg[2], 5) = 2.0 * t_wall - PR(b, nn[0], nn[1], nn[2], 5);
          else PR(b, gg[0], gg[1], gg[2], 5) = PR(b, nn[0], nn[1], nn[2], 5);
          const double pg = PR(b, gg[0], gg[1], gg[2], 4);
          double tg = PR(b, gg[0], gg[1], gg[2], 5);
          if (tg < 1e-6) tg = 1.0;
          const double rc = moocp * pg / tg;
          PR(b, gg[0], gg[1], gg[2], 0) = (l == 3) ? -rc : rc;
        }
        PR(b, i, j, k, 0) = rho_wall;
        PR(b, i, j, k, 1) = 0.0;
        PR(b, i, j, k, 2) = 0.0;
        PR(b, i, j, k, 3) = 0.0;
        PR(b, i, j, k, 4) = p_wall;
        PR(b, i, j, k, 5) = t_wall;
      } else {
        // apply_inviscid_wall (:191-248)
        double nx, ny, nz;
        get_normal(b, i, j, k, T.s_nd, &nx, &ny, &nz);
        for (int l = 1; l <= NGL; ++l) {
          int gg[3], nn[3];
          ghost(l, gg);
          inner(l, nn);
          const double ri = PR(b, nn[0], nn[1], nn[2], 0), ui = PR(b, nn[0], nn[1], nn[2], 1);
          const double vi = PR(b, nn[0], nn[1], nn[2], 2), wi = PR(b, nn[0], nn[1], nn[2], 3);
          const double pi = PR(b, nn[0], nn[1], nn[2], 4), ti = PR(b, nn[0], nn[1], nn[2], 5);
          const double vdn = ui * nx + vi * ny + wi * nz;
          PR(b, gg[0], gg[1], gg[2], 0) = (l == 3) ? -ri : ri;
          PR(b, gg[0], gg[1], gg[2], 1) = ui - 2.0 * vdn * nx;
          PR(b, gg[0], gg[1], gg[2], 2) = vi - 2.0 * vdn * ny;
          PR(b, gg[0], gg[1], gg[2], 3) = wi - 2.0 * vdn * nz;
          PR(b, gg[0], gg[1], gg[2], 4) = pi;
          PR(b, gg[0], gg[1], gg[2], 5) = ti;
        }
        if (p.method == 1) {
          const int o[3] = {i, j, k};
          dif_average(b, o, dir, lr);
        }
      }
      break;
    case 3: {
      // apply_symmetry_bc (:97-183)
      const int kphys = b.nz - 2 * NG;
      const bool is2d = (dir != 2 && kphys <= 2);
      if (is2d) {
        for (int l = 1; l <= NGL; ++l) {
          int gg[3], nn[3];
          ghost(l, gg);
          inner(l, nn);
          PR(b, gg[0], gg[1], gg[2], 0) = PR(b, nn[0], nn[1], nn[2], 0);
          PR(b, gg[0], gg[1], gg[2], 1) = PR(b, nn[0], nn[1], nn[2], 1);
          PR(b, gg[0], gg[1], gg[2], 2) = PR(b, nn[0], nn[1], nn[2], 2);
          PR(b, gg[0], gg[1], gg[2], 4) = PR(b, nn[0], nn[1], nn[2], 4);
          PR(b, gg[0], gg[1], gg[2], 5) = PR(b, nn[0], nn[1], nn[2], 5);
          PR(b, gg[0], gg[1], gg[2], 3) = -PR(b, nn[0], nn[1], nn[2], 3);
        }
      } else {
        double nx, ny, nz;
        get_normal(b, T.mid[0], T.mid[1], T.mid[2], T.s_nd, &nx, &ny, &nz);
        for (int l = 1; l <= NGL; ++l) {
          int gg[3], nn[3];
          ghost(l, gg);
          inner(l, nn);
          const double rho = PR(b, nn[0], nn[1], nn[2], 0), u = PR(b, nn[0], nn[1], nn[2], 1);
          const double v = PR(b, nn[0], nn[1], nn[2], 2), w = PR(b, nn[0], nn[1], nn[2], 3);
          const double pp = PR(b, nn[0], nn[1], nn[2], 4);
          const double vdn = u * nx + v * ny + w * nz;
          PR(b, gg[0], gg[1], gg[2], 0) = rho;
          PR(b, gg[0], gg[1], gg[2], 1) = u - 2.0 * vdn * nx;
          PR(b, gg[0], gg[1], gg[2], 2) = v - 2.0 * vdn * ny;
          PR(b, gg[0], gg[1], gg[2], 3) = w - 2.0 * vdn * nz;
          PR(b, gg[0], gg[1], gg[2], 4) = pp;
          PR(b, gg[0], gg[1], gg[2], 5) = PR(b, nn[0], nn[1], nn[2], 5);
        }
      }
      if (p.method == 1) {
        const int o[3] = {i, j, k};
        dif_average(b, o, dir, lr);
      }
    } break;
    case 4: {
      // apply_farfield_bc (:369-472)
      const double gama = p.gama, gm1 = gama - 1.0;
      const double rinf = p.roo, uinf = p.uoo, vinf = p.voo, winf = p.woo, pinf = p.poo;
      const double cinf = sqrt(gama * pinf / rinf);
      double nx, ny, nz;
      get_normal(b, i, j, k, T.s_nd, &nx, &ny, &nz);
      if (lr == -1) {
        nx = -nx;
        ny = -ny;
        nz = -nz;
      }
      int nn[3];
      inner(1, nn);
      const double rin = PR(b, nn[0], nn[1], nn[2], 0), uin = PR(b, nn[0], nn[1], nn[2], 1);
      const double vin = PR(b, nn[0], nn[1], nn[2], 2), win = PR(b, nn[0], nn[1], nn[2], 3);
      const double pin = PR(b, nn[0], nn[1], nn[2], 4);
      const double cin = sqrt(gama * pin / rin);
      // s_in = p_in / rho_in^gamma is formed only where it is used (subsonic
      // outflow): the same value, one double-double power fewer elsewhere
      const double sinf = pinf / orion_dd::pow_cr(rinf, gama);
      const double vnin = uin * nx + vin * ny + win * nz;
      const double vninf = uinf * nx + vinf * ny + winf * nz;
      const double vel = sqrt(uin * uin + vin * vin + win * win);
      const double mach = vel / cin;
      double rb, ub, vb, wb, pb;
      if (mach < 1.0) {
        const double Rp = vnin + 2.0 * cin / gm1;
        const double Rm = vninf - 2.0 * cinf / gm1;
        const double vnb = 0.5 * (Rp + Rm);
        const double cb = 0.25 * gm1 * (Rp - Rm);
        double sb, ur, vr, wr, vnr;
        if (vnb > 0.0) {
          sb = pin / orion_dd::pow_cr(rin, gama); ur = uin; vr = vin; wr = win; vnr = vnin;
        } else {
          sb = sinf; ur = uinf; vr = vinf; wr = winf; vnr = vninf;
        }
        rb = orion_dd::pow_cr(cb * cb / (gama * sb), 1.0 / gm1);
        pb = sb * orion_dd::pow_cr(rb, gama);
        ub = ur + (vnb - vnr) * nx;
        vb = vr + (vnb - vnr) * ny;
        wb = wr + (vnb - vnr) * nz;
      } else if (vnin > 0.0) {
        rb = rin; ub = uin; vb = vin; wb = win; pb = pin;
      } else {
        rb = rinf; ub = uinf; vb = vinf; wb = winf; pb = pinf;
      }
      const double tb = (mach * mach * gama) * pb / rb;
      for (int l = 1; l <= NGL; ++l) {
        int gg[3];
        ghost(l, gg);
        PR(b, gg[0], gg[1], gg[2], 0) = (l == 3) ? -rb : rb;
        PR(b, gg[0], gg[1], gg[2], 1) = ub;
        PR(b, gg[0], gg[1], gg[2], 2) = vb;
        PR(b, gg[0], gg[1], gg[2], 3) = wb;
        PR(b, gg[0], gg[1], gg[2], 4) = pb;
        PR(b, gg[0], gg[1], gg[2], 5) = tb;
      }
    } break;
    case 5: {
      // apply_freestream_bc (:477-521)
      const double v6[6] = {p.roo, p.uoo, p.voo, p.woo, p.poo, p.too};
      for (int l = 1; l <= NGL; ++l) {
        int gg[3];
        ghost(l, gg);
        for (int m = 0; m < 6; ++m) PR(b, gg[0], gg[1], gg[2], m) = v6[m];
      }
    } break;
    case 6: {
      // apply_extrapolation_bc (:526-571)
      int nn[3];
      inner(1, nn);
      double v6[6];
      for (int m = 0; m < 6; ++m) v6[m] = PR(b, nn[0], nn[1], nn[2], m);
      for (int l = 1; l <= NGL; ++l) {
        int gg[3];
        ghost(l, gg);
        for (int m = 0; m < 6; ++m) PR(b, gg[0], gg[1], gg[2], m) = v6[m];
      }
    } break;
    default:
      break;
  }
}

__global__ void __launch_bounds__(256) avg_stage_kernel(const BlockDev* __restrict__ blocks,
                                                        const AvgSrc* __restrict__ src, long long n,
                                                        double* __restrict__ stage) {
  const long long g = xcd_thread();
  if (g >= n) return;
  const AvgSrc s = src[g];
  const BlockDev& b = blocks[s.block];
  const double vol = b.vol[s.idx];
#pragma unroll
  for (int m = 0; m < 5; ++m) stage[g * 5 + m] = b.dq[(long long)m * b.P + s.idx] / vol;
}

__global__ void __launch_bounds__(256) avg_apply_kernel(const BlockDev* __restrict__ blocks,
                                                        const AvgDst* __restrict__ dst, long long n,
                                                        const double* __restrict__ stage) {
  const long long g = xcd_thread();
  if (g >= n) return;
  const AvgDst D = dst[g];
  const BlockDev& b = blocks[D.block];
  const double vs = b.vol[D.idx];
  double v[5];
#pragma unroll
  for (int m = 0; m < 5; ++m) v[m] = b.dq[(long long)m * b.P + D.idx];
  for (int e = 0; e < D.count; ++e) {
    const double* st = stage + (D.start + e) * 5;
#pragma unroll
    for (int m = 0; m < 5; ++m) v[m] = 0.5 * (st[m] * vs + v[m]);
  }
#pragma unroll
  for (int m = 0; m < 5; ++m) b.dq[(long long)m * b.P + D.idx] = v[m];
}

static inline unsigned grid_of(long long n, int tb) { return (unsigned)((n + tb - 1) / tb); }

void launch_halo(const long long* pairs, long long n, double* arena, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(halo_kernel, dim3(grid_of(n, 256)), dim3(256), 0, s, pairs, n, arena);
}
void launch_halo_points(const unsigned long long* pts, long long n, double* arena, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(halo_point_kernel, dim3(grid_of(n, 256)), dim3(256), 0, s, pts, n, arena);
}
void launch_bc(const BlockDev* blocks, const BCTask* tasks, int ntasks, long long n, hipStream_t s) {
  if (n <= 0 || ntasks <= 0) return;
  hipLaunchKernelGGL(bc_kernel, dim3(grid_of(n, 256)), dim3(256), 0, s, blocks, tasks, ntasks, n);
}
void launch_avg_stage(const BlockDev* blocks, const AvgSrc* src, long long n, double* stage, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(avg_stage_kernel, dim3(grid_of(n, 256)), dim3(256), 0, s, blocks, src, n, stage);
}
void launch_avg_apply(const BlockDev* blocks, const AvgDst* dst, long long n, const double* stage, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(avg_apply_kernel, dim3(grid_of(n, 256)), dim3(256), 0, s, blocks, dst, n, stage);
}

}  // namespace orion_hip
