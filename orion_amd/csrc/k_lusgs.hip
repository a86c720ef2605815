// k_lusgs.hip -- LU-SGS update for gfx950 (FP64).
//
// StateUpdater::update_solution (src/solver/StateUpdater.cpp:297-580) sweeps
// each block lexicographically (k, then j, then i; reversed for the backward
// sweeps) and every cell reads only its six face neighbours.  On the
// hyperplane s = i+j+k, lower neighbours lie on s-1 and upper ones on s+1,
// so processing planes in increasing (decreasing) s with all cells of a
// plane in parallel reads exactly the same neighbour values as the
// lexicographic sweep: the results are identical, not an approximation.
//
// Layout: the sweeps work on a plane-major copy of their data.  Cells are
// numbered plane by plane over all blocks (plane s of block 0, then of block
// 1, ...; within a block by k then j), so the cells of one plane are
// contiguous and a plane's neighbours on s-1 and s+1 are contiguous too.
//
// Split form: everything in a cell's update that does not depend on the
// evolving dq -- the diagonal 1/(1/dt + sum of radii), the normalised own
// metrics, and for each of the six neighbour slots the split eigenvalue
// coefficients l1, x1, x2, ct_n and the viscous 0.5*r_v, plus each cell's
// own c^2, H and a_f used when it is a neighbour -- is computed once per
// update by lu_coef (fully parallel, natural order) with the reference's
// expressions (calc_spectral_radius :175-191, calc_visc_spectral_radius
// :194-207, compute_flux_splitting :209-254, get_diag_coeff :267-295).
// The sweeps evaluate only the dq-linear remainder of
// compute_flux_splitting (:255-264) and the sweep bodies (:318-365,
// :392-441) with the same expressions, so results are bit-identical to
// recomputing everything in the sweep.
#include "orion_dev.h"
#include "orion_kernels.h"

namespace orion_hip {

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

// packed coefficient fields ([field][ncell])
//   0 coed | 1..9 nx_n,ny_n,nz_n per axis | 10+5*slot+{0..4} l1,x1,x2,ct_n,hv
// source fields: 0..2 u,v,w | 3 c2 | 4 H | 5 af
// slots: 0 i-1, 1 j-1, 2 k-1, 3 i+1, 4 j+1, 5 k+1
constexpr int LC_COED = 0, LC_NRM = 1, LC_SLOT = 10;

// One thread per interior point in natural order: reads are coalesced and
// the plane-major writes of neighbouring rows merge in L2 (iterating in
// plane order instead, with scattered reads, measured 2.5x slower).
__global__ void __launch_bounds__(256) lu_coef_kernel(const BlockDev* __restrict__ blocks,
                                                      const long long* __restrict__ off, int nb, LuPacked L,
                                                      Params p) {
  const long long g = xcd_thread();
  if (g >= off[nb]) return;
  const int bi = find_block(off, nb, g);
  const BlockDev& b = blocks[bi];
  int i, j, k;
  decode_interior(b, g - off[bi], &i, &j, &k);
  const long long P = b.P;
  const long long id = pidx(b, i, j, k);
  const long long N = L.ncell;
  const long long q = b.lpos[id];
  const double gamma = p.gama, re = p.reynolds;
  // rhs copy with the physical-BC windows zeroed (StateUpdater.cpp:557-570)
  const bool fz = b.frozen[id] != 0;
#pragma unroll
  for (int m = 0; m < 5; ++m) L.rhs[m * N + q] = fz ? 0.0 : b.dq[m * P + id];
  double pr[5];
#pragma unroll
  for (int m = 0; m < 5; ++m) pr[m] = b.prim[m * P + id];
  double met[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) met[m] = b.met[m * P + id];
  {  // get_diag_coeff
    const double ra = spec_rad(pr, gamma, met[0], met[1], met[2]);
    const double rb = spec_rad(pr, gamma, met[3], met[4], met[5]);
    const double rc = spec_rad(pr, gamma, met[6], met[7], met[8]);
    double rva = 0, rvb = 0, rvc = 0;
    if (p.nvis == 1) {
      const double mu = b.mu[id];
      double vol = b.vol[id];
      if (vol < 1e-30) vol = 1e-30;
      rva = visc_rad(pr[0], mu, vol, re, met[0], met[1], met[2]);
      rvb = visc_rad(pr[0], mu, vol, re, met[3], met[4], met[5]);
      rvc = visc_rad(pr[0], mu, vol, re, met[6], met[7], met[8]);
    }
    const double rs = ra + rb + rc + rva + rvb + rvc;
    const double term = (1.0 / b.dt[id]) + 1.0 * 1.0 * rs;
    L.coef[LC_COED * N + q] = 1.0 / term;
  }
  {  // this cell as a neighbour
    double rm = pr[0], pm = pr[4];
    const double v2 = pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3];
    if (rm < 1e-30) rm = 1e-30;
    if (pm < 1e-30) pm = 1e-30;
    L.src[0 * N + q] = pr[1];
    L.src[1 * N + q] = pr[2];
    L.src[2 * N + q] = pr[3];
    L.src[3 * N + q] = gamma * pm / rm;
    L.src[4 * N + q] = (gamma / (gamma - 1.0)) * (pm / rm) + 0.5 * v2;
    L.src[5 * N + q] = 0.5 * (gamma - 1.0) * v2;
  }
  double gmv[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double nx = met[3 * a], ny = met[3 * a + 1], nz = met[3 * a + 2];
    double gm = sqrt(nx * nx + ny * ny + nz * nz);
    if (gm < 1.0e-30) gm = 1.0e-30;
    gmv[a] = gm;
    const double im = 1.0 / gm;
    L.coef[(LC_NRM + 3 * a) * N + q] = nx * im;
    L.coef[(LC_NRM + 3 * a + 1) * N + q] = ny * im;
    L.coef[(LC_NRM + 3 * a + 2) * N + q] = nz * im;
  }
#pragma unroll
  for (int sl = 0; sl < 6; ++sl) {
    const int ax = sl % 3;
    const int step = sl < 3 ? -1 : 1;
    const int sgn = sl < 3 ? 1 : -1;
    int in = i, jn = j, kn = k;
    if (ax == 0) in += step;
    else if (ax == 1) jn += step;
    else kn += step;
    double* S = L.coef + (long long)(LC_SLOT + 5 * sl) * N + q;
    if (in < NG || in >= b.nx - NG || jn < NG || jn >= b.ny - NG || kn < NG || kn >= b.nz - NG) {
#pragma unroll
      for (int t = 0; t < 5; ++t) S[t * N] = 0.0;
      continue;
    }
    const long long idn = pidx(b, in, jn, kn);
    double pn[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) pn[m] = b.prim[m * P + idn];
    const double nx = met[3 * ax], ny = met[3 * ax + 1], nz = met[3 * ax + 2];
    const double r = spec_rad(pn, gamma, nx, ny, nz);
    double rm = pn[0], pm = pn[4];
    const double um = pn[1], vm = pn[2], wm = pn[3];
    if (rm < 1e-30) rm = 1e-30;
    if (pm < 1e-30) pm = 1e-30;
    const double c2 = gamma * pm / rm;
    const double cm = sqrt(c2);
    const double nt = 0.0;
    const double ct = nx * um + ny * vm + nz * wm + nt;
    const double gm = gmv[ax];
    const double ssr = (double)sgn * r;
    const double l1 = 0.5 * (ct + ssr);
    const double l4 = 0.5 * ((ct + cm * gm) + ssr);
    const double l5 = 0.5 * ((ct - cm * gm) + ssr);
    S[0] = l1;
    S[N] = (2.0 * l1 - l4 - l5) / (2.0 * c2);
    S[2 * N] = (l4 - l5) / (2.0 * cm);
    S[3 * N] = (ct - nt) * (1.0 / gm);
    double hv = 0.0;
    if (p.nvis == 1) {
      const double mun = b.mu[idn];
      double voln = b.vol[idn];
      if (voln < 1e-30) voln = 1e-30;
      hv = 0.5 * visc_rad(pn[0], mun, voln, re, nx, ny, nz);
    }
    S[4 * N] = hv;
  }
}

// dq-linear part of one neighbour's contribution (compute_flux_splitting
// :255-264 plus the viscous off-diagonal :341-343 / :419-421).  Loads are
// issued for every slot up front (an absent neighbour reads the cell
// itself) and an absent neighbour's result is dropped with a select, which
// leaves r0 exactly as skipping it does.
struct SlotIn {
  double l1, x1, x2, ctn, hv, nxn, nyn, nzn;
  double um, vm, wm, c2, H, af;
  double d[5];
};

template <int SL>
__device__ __forceinline__ void slot_load(const LuPacked& L, long long q, long long qn, SlotIn& s) {
  const long long N = L.ncell;
  constexpr int ax = SL % 3;
  const double* C = L.coef;
  s.l1 = C[(LC_SLOT + 5 * SL) * N + q];
  s.x1 = C[(LC_SLOT + 5 * SL + 1) * N + q];
  s.x2 = C[(LC_SLOT + 5 * SL + 2) * N + q];
  s.ctn = C[(LC_SLOT + 5 * SL + 3) * N + q];
  s.hv = C[(LC_SLOT + 5 * SL + 4) * N + q];
  s.nxn = C[(LC_NRM + 3 * ax) * N + q];
  s.nyn = C[(LC_NRM + 3 * ax + 1) * N + q];
  s.nzn = C[(LC_NRM + 3 * ax + 2) * N + q];
  s.um = L.src[qn];
  s.vm = L.src[N + qn];
  s.wm = L.src[2 * N + qn];
  s.c2 = L.src[3 * N + qn];
  s.H = L.src[4 * N + qn];
  s.af = L.src[5 * N + qn];
#pragma unroll
  for (int m = 0; m < 5; ++m) s.d[m] = L.dq[m * N + qn];
}

template <bool GS, int SL>
__device__ __forceinline__ void slot_apply(const SlotIn& s, bool valid, double gm1, int nvis, double* r0) {
  constexpr int sgn = SL < 3 ? 1 : -1;
  const double drm = s.d[0], dum = s.d[1], dvm = s.d[2], dwm = s.d[3], dEm = s.d[4];
  const double dc = s.ctn * drm - (s.nxn * dum + s.nyn * dvm + s.nzn * dwm);
  const double dh = s.af * drm - gm1 * (s.um * dum + s.vm * dvm + s.wm * dwm - dEm);
  const double c2dc = s.c2 * dc;
  double res[5];
  res[0] = s.l1 * drm - dh * s.x1 - dc * s.x2;
  res[1] = s.l1 * dum + (s.nxn * c2dc - s.um * dh) * s.x1 + (s.nxn * dh - s.um * dc) * s.x2;
  res[2] = s.l1 * dvm + (s.nyn * c2dc - s.vm * dh) * s.x1 + (s.nyn * dh - s.vm * dc) * s.x2;
  res[3] = s.l1 * dwm + (s.nzn * c2dc - s.wm * dh) * s.x1 + (s.nzn * dh - s.wm * dc) * s.x2;
  res[4] = s.l1 * dEm + (s.ctn * c2dc - s.H * dh) * s.x1 + (s.ctn * dh - s.H * dc) * s.x2;
  if (nvis == 1) {
    if (GS) {
#pragma unroll
      for (int m = 0; m < 5; ++m) res[m] += s.hv * s.d[m] * (double)sgn;
    } else {
#pragma unroll
      for (int m = 0; m < 5; ++m) res[m] += s.hv * s.d[m];
    }
  }
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    const double acc = (GS && sgn == -1) ? r0[m] - res[m] : r0[m] + res[m];
    r0[m] = valid ? acc : r0[m];
  }
}

// Sweep types: 0 LU forward, 1 LU backward, 2 GS forward, 3 GS backward.
// One launch per plane: the plane's cells are L[start, start+count).
template <int TYPE>
__global__ void __launch_bounds__(256) lusgs_sweep_kernel(LuPacked L, long long start, long long count, double gm1,
                                                          int nvis) {
  const long long t = xcd_thread();
  if (t >= count) return;
  const long long q = start + t;
  const long long N = L.ncell;
  double r0[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  constexpr bool GS = TYPE >= 2;
  constexpr int S0 = (TYPE == 1) ? 3 : 0;        // first slot used
  constexpr int NS = GS ? 6 : 3;                 // slots used
  int nb[NS];
#pragma unroll
  for (int u = 0; u < NS; ++u) nb[u] = L.nbr[(long long)(S0 + u) * N + q];
  SlotIn in[NS];
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    const long long qn = nb[u] >= 0 ? (long long)nb[u] : q;
    switch (S0 + u) {
      case 0: slot_load<0>(L, q, qn, in[u]); break;
      case 1: slot_load<1>(L, q, qn, in[u]); break;
      case 2: slot_load<2>(L, q, qn, in[u]); break;
      case 3: slot_load<3>(L, q, qn, in[u]); break;
      case 4: slot_load<4>(L, q, qn, in[u]); break;
      default: slot_load<5>(L, q, qn, in[u]); break;
    }
  }
  const double coed = L.coef[LC_COED * N + q];
  double rh[5], dqo[5];
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    rh[m] = (TYPE == 1) ? 0.0 : L.rhs[m * N + q];
    dqo[m] = (TYPE == 1) ? L.dq[m * N + q] : 0.0;
  }
  if (TYPE == 1) {
    slot_apply<false, 3>(in[0], nb[0] >= 0, gm1, nvis, r0);
    slot_apply<false, 4>(in[1], nb[1] >= 0, gm1, nvis, r0);
    slot_apply<false, 5>(in[2], nb[2] >= 0, gm1, nvis, r0);
  } else {
    slot_apply<GS, 0>(in[0], nb[0] >= 0, gm1, nvis, r0);
    slot_apply<GS, 1>(in[1], nb[1] >= 0, gm1, nvis, r0);
    slot_apply<GS, 2>(in[2], nb[2] >= 0, gm1, nvis, r0);
    if (GS) {
      slot_apply<true, 3>(in[NS > 3 ? 3 : 0], nb[NS > 3 ? 3 : 0] >= 0, gm1, nvis, r0);
      slot_apply<true, 4>(in[NS > 4 ? 4 : 0], nb[NS > 4 ? 4 : 0] >= 0, gm1, nvis, r0);
      slot_apply<true, 5>(in[NS > 5 ? 5 : 0], nb[NS > 5 ? 5 : 0] >= 0, gm1, nvis, r0);
    }
  }
  const double wmig = 1.0;
  if (TYPE == 0) {
#pragma unroll
    for (int m = 0; m < 5; ++m) L.dq[m * N + q] = (-rh[m] + wmig * r0[m]) * coed;
  } else if (TYPE == 1) {
#pragma unroll
    for (int m = 0; m < 5; ++m) L.dq[m * N + q] = dqo[m] - (wmig * r0[m]) * coed;
  } else {
#pragma unroll
    for (int m = 0; m < 5; ++m) L.dq[m * N + q] = (wmig * r0[m] + -rh[m]) * coed;
  }
}

static inline unsigned grid_of(long long n, int tb) { return (unsigned)((n + tb - 1) / tb); }

void launch_lu_coef(const Launch& La, long long n, const LuPacked& L, const Params& p) {
  if (n <= 0) return;
  hipLaunchKernelGGL(lu_coef_kernel, dim3(grid_of(n, 256)), dim3(256), 0, La.stream, La.blocks, La.off, La.nb, L, p);
}

void launch_lusgs_sweep(const LuPacked& L, long long start, long long count, int type, const Params& p,
                        hipStream_t s) {
  if (count <= 0) return;
  const double gm1 = p.gama - 1.0;
  const dim3 gr(grid_of(count, 256)), tb(256);
  switch (type) {
    case 0: hipLaunchKernelGGL(lusgs_sweep_kernel<0>, gr, tb, 0, s, L, start, count, gm1, p.nvis); break;
    case 1: hipLaunchKernelGGL(lusgs_sweep_kernel<1>, gr, tb, 0, s, L, start, count, gm1, p.nvis); break;
    case 2: hipLaunchKernelGGL(lusgs_sweep_kernel<2>, gr, tb, 0, s, L, start, count, gm1, p.nvis); break;
    default: hipLaunchKernelGGL(lusgs_sweep_kernel<3>, gr, tb, 0, s, L, start, count, gm1, p.nvis); break;
  }
}

}  // namespace orion_hip
