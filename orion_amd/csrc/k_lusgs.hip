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
//  * Inside a tile, neighbours along col and row are lanes of the same
//    workgroup and are exchanged through LDS (two barriers per step); along
//    the march axis they are the lane's own previous / next cell (registers).
//  * Across tiles, a tile publishes how far it has come in a progress flag
//    and its downstream tiles wait on it (bounded spin) before reading edge
//    values.  dq is stored with agent-scope atomic stores and each wave
//    waits for their completion before the barrier that precedes the flag
//    store; edge dq is read with agent-scope atomic loads.  No cache
//    maintenance is needed (fence-based release/acquire around plain stores
//    flushes and invalidates L2 and measured 2-8x slower).
//    Tiles are launched in dependency order and a wait only ever targets a
//    tile launched earlier, so the persistent grid always drains; a spin
//    that exceeds its bound records an error instead of hanging.
//
// Everything in a cell's update that does not depend on dq is either packed
// once per update by lu_pack (the diagonal 1/(1/dt + sum of radii) and each
// cell's u, v, w, c^2, H, a_f, c and viscous coefficient) or recomputed per
// neighbour slot in the sweep from this cell's metrics and the neighbour's
// packed values, with the reference's expressions (calc_spectral_radius
// :175-191, calc_visc_spectral_radius :194-207, compute_flux_splitting
// :209-264, get_diag_coeff :267-295, sweep bodies :318-365, :392-441).
#include "orion_dev.h"
#include "orion_kernels.h"

namespace orion_hip {

constexpr int TW = 16;         // tile width along col and row
constexpr int TL = TW * TW;    // lanes per tile workgroup
constexpr int KEY_BITS = 16;   // progress flag: epoch << KEY_BITS | plane key

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
// packing (one thread per interior point, natural order)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) lu_geom_kernel(const BlockDev* __restrict__ blocks,
                                                      Span SP, SweepArrays S) {
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;
  const BlockDev& b = blocks[bi];
  int i, j, k, m, r, c;
  decode_interior(b, gi, &i, &j, &k);
  natural_to_sweep(b, i, j, k, &m, &r, &c);
  const long long q = sweep_index(b, m, r, c);
  const long long id = pidx(b, i, j, k);
  double* own = reinterpret_cast<double*>(S.own);
#pragma unroll
  for (int f = 0; f < 9; ++f) own[sweep_field(q, c, b.sncp, SW_OWN, f)] = b.met[f * b.P + id];
}

__global__ void __launch_bounds__(256) lu_pack_kernel(const BlockDev* __restrict__ blocks,
                                                      Span SP, SweepArrays S,
                                                      Params p) {
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;
  const BlockDev& b = blocks[bi];
  int i, j, k, m, r, c;
  decode_interior(b, gi, &i, &j, &k);
  natural_to_sweep(b, i, j, k, &m, &r, &c);
  const long long q = sweep_index(b, m, r, c);
  const int ncp = b.sncp;
  double* own = reinterpret_cast<double*>(S.own);
  double* Dp = reinterpret_cast<double*>(S.D);
  const long long P = b.P;
  const long long id = pidx(b, i, j, k);
  const double gamma = p.gama, re = p.reynolds;
  // rhs copy with the physical-BC windows zeroed (StateUpdater.cpp:557-570)
  const bool fz = b.frozen[id] != 0;
#pragma unroll
  for (int f = 0; f < 5; ++f) own[sweep_field(q, c, ncp, SW_OWN, 10 + f)] = fz ? 0.0 : b.dq[f * P + id];
  double pr[5];
#pragma unroll
  for (int f = 0; f < 5; ++f) pr[f] = b.prim[f * P + id];
  double met[9];
#pragma unroll
  for (int f = 0; f < 9; ++f) met[f] = b.met[f * P + id];
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
    own[sweep_field(q, c, ncp, SW_OWN, 9)] = 1.0 / term;
  }
  // this cell seen as a neighbour: compute_flux_splitting's per-cell terms,
  // the sound speed of calc_spectral_radius and the viscous coefficient of
  // calc_visc_spectral_radius (same clamps)
  double rm = pr[0], pm = pr[4];
  const double v2 = pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3];
  if (rm < 1e-30) rm = 1e-30;
  if (pm < 1e-30) pm = 1e-30;
  const double c2 = gamma * pm / rm;
  Dp[sweep_field(q, c, ncp, SW_D, 0)] = pr[1];
  Dp[sweep_field(q, c, ncp, SW_D, 1)] = pr[2];
  Dp[sweep_field(q, c, ncp, SW_D, 2)] = pr[3];
  Dp[sweep_field(q, c, ncp, SW_D, 3)] = c2;
  Dp[sweep_field(q, c, ncp, SW_D, 4)] = (gamma / (gamma - 1.0)) * (pm / rm) + 0.5 * v2;
  Dp[sweep_field(q, c, ncp, SW_D, 5)] = 0.5 * (gamma - 1.0) * v2;
  Dp[sweep_field(q, c, ncp, SW_D, 6)] = sqrt(c2);
  Dp[sweep_field(q, c, ncp, SW_D, 7)] = 2.0 * mu / (fabs(re) * (rm * vol) + 1.0e-30);
}

// dq back to the block arrays (zero on physical-BC windows,
// StateUpdater.cpp:577) and update_limited (:470-542).
__global__ void __launch_bounds__(256) update_kernel(const BlockDev* __restrict__ blocks,
                                                     Span SP, SweepArrays S,
                                                     double gama) {
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;
  const BlockDev& b = blocks[bi];
  int i, j, k, m, r, c;
  decode_interior(b, gi, &i, &j, &k);
  const long long id = pidx(b, i, j, k);
  const long long P = b.P;
  double d[5];
  if (b.frozen[id]) {
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      d[f] = 0.0;
      b.dq[f * P + id] = 0.0;
    }
  } else {
    natural_to_sweep(b, i, j, k, &m, &r, &c);
    const long long q = sweep_index(b, m, r, c);
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      d[f] = reinterpret_cast<const double*>(S.dq)[sweep_field(q, c, b.sncp, SW_DQ, f)];
      b.dq[f * P + id] = d[f];
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

// ---------------------------------------------------------------------------
// one neighbour slot: coefficients from this cell's metrics and the
// neighbour's packed values (compute_flux_splitting :209-254 with
// calc_spectral_radius and calc_visc_spectral_radius), then the dq-linear
// remainder (:255-264) and the viscous off-diagonal (:341-343 / :419-421)
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
// values and dq
template <bool GS, int SL>
__device__ __forceinline__ void slot(const double* n, const AxisGeo& a, const double* Dn, const double* d, bool valid,
                                     double gm1, int nvis, double* r0) {
  constexpr int sgn = SL < 3 ? 1 : -1;
  const double nx = n[0], ny = n[1], nz = n[2];
  const double um = Dn[0], vm = Dn[1], wm = Dn[2], c2 = Dn[3], H = Dn[4], af = Dn[5], cm = Dn[6];
  // calc_spectral_radius with this cell's metrics and the neighbour's state
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
  double res[5];
  res[0] = l1 * drm - dh * x1 - dc * x2;
  res[1] = l1 * dum + (nxn * c2dc - um * dh) * x1 + (nxn * dh - um * dc) * x2;
  res[2] = l1 * dvm + (nyn * c2dc - vm * dh) * x1 + (nyn * dh - vm * dc) * x2;
  res[3] = l1 * dwm + (nzn * c2dc - wm * dh) * x1 + (nzn * dh - wm * dc) * x2;
  res[4] = l1 * dEm + (ctn * c2dc - H * dh) * x1 + (ctn * dh - H * dc) * x2;
  if (nvis == 1) {
    const double hv = 0.5 * (Dn[7] * a.g2);
    if (GS) {
#pragma unroll
      for (int f = 0; f < 5; ++f) res[f] += hv * d[f] * (double)sgn;
    } else {
#pragma unroll
      for (int f = 0; f < 5; ++f) res[f] += hv * d[f];
    }
  }
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    const double acc = (GS && sgn == -1) ? r0[f] - res[f] : r0[f] + res[f];
    r0[f] = valid ? acc : r0[f];
  }
}

// ---------------------------------------------------------------------------
// tile sweep
// ---------------------------------------------------------------------------
constexpr int NW = 13;  // exchanged words per cell: D (8) + dq (5)

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

struct Own {  // a cell's own inputs to its update
  double met[9];
  double coed;
  double r5[5];  // rhs, or (backward LU) the forward sweep's dq
};

// Loads are unconditional: a lane whose cell is absent reads the block's
// first cell instead (always a valid address) and never uses the values.
// (rb, c): the cell's row base q - c and column; pair k of a K-pair group
// is one 16-byte load
__device__ __forceinline__ double2 ld_pair(const double2* g, long long rb, int c, int ncp, int K, int k) {
  return g[rb * K + (long long)k * ncp + c];
}

template <int TYPE>
__device__ __forceinline__ void load_own(const SweepArrays& S, long long rb, int c, int ncp, Own& o) {
  double2 v[SW_OWN];
#pragma unroll
  for (int k = 0; k < (TYPE == 1 ? 5 : SW_OWN); ++k) v[k] = ld_pair(S.own, rb, c, ncp, SW_OWN, k);
  if (TYPE == 1) {
#pragma unroll
    for (int k = 0; k < SW_DQ; ++k) v[5 + k] = ld_pair(S.dq, rb, c, ncp, SW_DQ, k);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    o.met[2 * k] = v[k].x;
    o.met[2 * k + 1] = v[k].y;
  }
  o.met[8] = v[4].x;
  o.coed = v[4].y;
  o.r5[0] = v[5].x;
  o.r5[1] = v[5].y;
  o.r5[2] = v[6].x;
  o.r5[3] = v[6].y;
  o.r5[4] = v[7].x;
}

// D (8) and, for the Gauss-Seidel sweeps, the cell's old dq (5); x holds
// 13 words for the Gauss-Seidel sweeps and 8 for the LU ones
template <bool GS>
__device__ __forceinline__ void load_nbr(const SweepArrays& S, long long rb, int c, int ncp, double* x) {
#pragma unroll
  for (int k = 0; k < SW_D; ++k) {
    const double2 v = ld_pair(S.D, rb, c, ncp, SW_D, k);
    x[2 * k] = v.x;
    x[2 * k + 1] = v.y;
  }
  if constexpr (GS) {
    const double2 a = ld_pair(S.dq, rb, c, ncp, SW_DQ, 0), b2 = ld_pair(S.dq, rb, c, ncp, SW_DQ, 1);
    const double2 e = ld_pair(S.dq, rb, c, ncp, SW_DQ, 2);
    x[8] = a.x;
    x[9] = a.y;
    x[10] = b2.x;
    x[11] = b2.y;
    x[12] = e.x;
  }
}

// Sweep types: 0 LU forward, 1 LU backward, 2 GS forward, 3 GS backward.
//
// Pipeline of one tile.  Step t works on hyperplane s(t) = s0 + dir*t:
//   phase 1  edge values (loaded two steps ago) and, for Gauss-Seidel, the
//            lanes' next cells go to LDS; barrier;
//   phase 2  the update of each lane's cell from LDS and registers; dq is
//            stored with agent-scope stores; when publishing, each wave waits
//            for its stores; barrier;
//   phase 3  thread 0 publishes plane s(t); the lanes' (D, new dq) go to
//            LDS; thread 0 makes sure the upstream tiles have finished plane
//            s(t+1), from flag values it read one step earlier (it spins
//            only when they have not); then the loads for later steps are
//            issued: own inputs of t+2, neighbour views of t+3, edge values
//            of t+2.
// Issuing every prefetch after the store wait keeps that wait from draining
// them, and reading the flags a step early keeps the flag round trip off
// the critical path, at the price of one step of lag per tile edge.
template <int TYPE>
__global__ void __launch_bounds__(TL, 1) lusgs_tile_kernel(SweepArrays S, const BlockDev* __restrict__ blocks,
                                                          const SweepTile* __restrict__ tiles,
                                                          const int* __restrict__ order, unsigned long long* flags,
                                                          const unsigned long long* __restrict__ epoch, int* err,
                                                          int pub, double gm1, int nvis) {
  constexpr bool FWD = (TYPE == 0 || TYPE == 2);
  constexpr bool GS = TYPE >= 2;
  constexpr int dir = FWD ? 1 : -1;
  constexpr int NROLE = GS ? 4 : 2;
  constexpr int NB = GS ? NW : 8;              // words of a neighbour view
  __shared__ double shA[NW][TL];               // (D, new dq) of each lane's cell of the last step
  __shared__ double shB[GS ? NW : 1][TL];      // (D, old dq) of each lane's cell of the next step
  __shared__ double shX[NROLE][NW][TW];        // edge values from neighbouring tiles

  const int tix = order[blockIdx.x];
  const SweepTile T = tiles[tix];
  const BlockDev& b = blocks[T.block];
  const int nm = b.sn[0], nc = b.sn[1], nr = b.sn[2];
  const int am = b.sax[0], ac = b.sax[1];
  const int lane = threadIdx.x, lc = lane % TW, lr = lane / TW;
  const int c = T.c0 + lc, r = T.r0 + lr;
  const bool lane_ok = c < nc && r < nr;
  const int up_c = FWD ? T.nbr[0] : T.nbr[2], up_r = FWD ? T.nbr[1] : T.nbr[3];
  const unsigned long long ep = (*epoch) << KEY_BITS;
  SweepTile ut[2];  // plane ranges of the upstream tiles (clamped wait targets)
  ut[0] = up_c >= 0 ? tiles[up_c] : T;
  ut[1] = up_r >= 0 ? tiles[up_r] : T;
  const int nsteps = T.s_last - T.s_first + 1;
  const int s0 = FWD ? T.s_first : T.s_last;
  const long long plane_stride = (long long)nr * b.sncp;  // one hyperplane in the sweep layout
  const long long qstep = dir * plane_stride;

  // this lane's cell at step t: valid flag and sweep index (block's first
  // cell when absent)
  const int ncp = b.sncp;
  const int m0 = s0 - r - c;
  const long long rb0 = b.sbase + ((long long)s0 * nr + r) * ncp;  // row base q - c at step 0
  auto cell_ok = [&](int t) {
    const int m = m0 + dir * t;
    return lane_ok && t < nsteps && m >= 0 && m < nm;
  };
  // row base and column of the lane's cell at step t (the block's first
  // cell when absent)
  auto cell_rb = [&](int t) { return cell_ok(t) ? rb0 + (long long)t * qstep : b.sbase; };
  auto cell_c = [&](int t) { return cell_ok(t) ? c : 0; };

  // edge loaders: lanes 0..16*NROLE-1 of the first wave, one (role, index)
  // each.  Roles: 0 col-upstream, 1 row-upstream, 2 col-downstream, 3
  // row-downstream; the upstream ones read new values (agent-scope loads),
  // the others old ones.  A neighbour of a cell on plane s lies on s -/+ dir.
  const int xrole = lane / TW, xi = lane % TW;
  const bool xloader = lane < NROLE * TW;
  const bool xup = xrole < 2;
  int xr, xc;
  {
    const int off = xup ? -dir : dir;
    if ((xrole & 1) == 0) {  // across the col edge
      xc = (xup == FWD ? T.c0 : T.c0 + TW - 1) + off;
      xr = T.r0 + xi;
    } else {  // across the row edge
      xr = (xup == FWD ? T.r0 : T.r0 + TW - 1) + off;
      xc = T.c0 + xi;
    }
  }
  const bool xin = xloader && xr >= 0 && xr < nr && xc >= 0 && xc < nc;
  auto xload = [&](int t, double* xv) {  // edge values for step t
    const int sp = s0 + dir * t + (xup ? -dir : dir);
    const int mm = sp - xr - xc;
    const bool ok = xin && t < nsteps && mm >= 0 && mm < nm;
    const long long rb = ok ? b.sbase + ((long long)sp * nr + xr) * ncp : b.sbase;
    const int cc = ok ? xc : 0;
    if (xup) {
      load_nbr<false>(S, rb, cc, ncp, xv);
      const double* dqp = reinterpret_cast<const double*>(S.dq);
#pragma unroll
      for (int f = 0; f < 5; ++f)
        xv[8 + f] = __hip_atomic_load(&dqp[sweep_field(rb + cc, cc, ncp, SW_DQ, f)], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
    } else {
      load_nbr<GS>(S, rb, cc, ncp, xv);
    }
  };

  // thread 0: the upstream tiles must have finished plane s; pre[] holds the
  // flag values read earlier, re-read (bounded spin) only when short
  unsigned long long pre[2] = {0, 0};
  auto ensure_up = [&](int s) {
    if (lane != 0) return;
    for (int k = 0; k < 2; ++k) {
      const int u = k ? up_r : up_c;
      if (u < 0) continue;
      const unsigned long long key = need_key(FWD, ut[k], s);
      if (key == 0) continue;
      const unsigned long long want = ep + key;
      unsigned long long v = pre[k];
      long long spins = 0;
      while (v < want) {
        // far beyond any legal wait (a whole sweep takes milliseconds), or
        // another tile already gave up: report and carry on, never hang
        if (++spins > (1LL << 20) ||
            ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
          atomicOr(err, 1);
          v = want;
          break;
        }
        if (spins > 1) __builtin_amdgcn_s_sleep(2);
        v = __hip_atomic_load(&flags[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      pre[k] = v;
    }
  };
  auto read_up = [&]() {  // thread 0: early flag reads
    if (lane != 0) return;
    if (up_c >= 0) pre[0] = __hip_atomic_load(&flags[up_c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (up_r >= 0) pre[1] = __hip_atomic_load(&flags[up_r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };

  Own own0, own1;
  double Dcur[8], nb1[NB], nb2[NB], xv0[NW], xv1[NW];
  {
    double t0[NB];
    load_own<TYPE>(S, cell_rb(0), cell_c(0), ncp, own0);
    load_own<TYPE>(S, cell_rb(1), cell_c(1), ncp, own1);
    load_nbr<GS>(S, cell_rb(0), cell_c(0), ncp, t0);
#pragma unroll
    for (int f = 0; f < 8; ++f) Dcur[f] = t0[f];
    load_nbr<GS>(S, cell_rb(1), cell_c(1), ncp, nb1);
    load_nbr<GS>(S, cell_rb(2), cell_c(2), ncp, nb2);
  }
#pragma unroll
  for (int f = 0; f < NW; ++f) shA[f][lane] = 0.0;
  ensure_up(s0);  // covers the edge values of steps 0 and 1
  if (xloader) {
    xload(0, xv0);
    xload(1, xv1);
  }
  read_up();

  for (int t = 0; t < nsteps; ++t) {
    const int s = s0 + dir * t;
    const int m = m0 + dir * t;
    const bool act = cell_ok(t);
    // phase 1
    if (xloader) {
#pragma unroll
      for (int f = 0; f < NW; ++f) shX[xrole][f][xi] = xv0[f];
    }
    if (GS) {
#pragma unroll
      for (int f = 0; f < NW; ++f) shB[f][lane] = nb1[f];
    }
    __syncthreads();
    // phase 2
    double dqn[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (act) {
      double r0[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      AxisGeo ag[3];
      axis_geo(&own0.met[0], ag[0]);
      axis_geo(&own0.met[3], ag[1]);
      axis_geo(&own0.met[6], ag[2]);
      // neighbour of real slot sl: its (D, dq) and validity.  The lane's
      // previous cell is its own entry of shA; for the Gauss-Seidel sweeps
      // its next cell is its own entry of shB.
      auto nb = [&](int sl, double* v) -> bool {
        const int ax = sl % 3, d = sl < 3 ? -1 : 1;
        const bool before = (d == -dir);  // precedes this cell in sweep order
        if (ax == am) {
          const bool ok = m + d >= 0 && m + d < nm;
          if (before) {
#pragma unroll
            for (int f = 0; f < NW; ++f) v[f] = shA[f][lane];
          } else {
#pragma unroll
            for (int f = 0; f < NW; ++f) v[f] = GS ? shB[GS ? f : 0][lane] : 0.0;
          }
          return ok;
        }
        const bool col = (ax == ac);
        const int pos = col ? lc : lr, lim = col ? nc : nr, coord = col ? c : r;
        const bool ok = coord + d >= 0 && coord + d < lim;
        if (pos + d >= 0 && pos + d < TW) {
          const int ln = lane + (col ? d : d * TW);
          if (before) {
#pragma unroll
            for (int f = 0; f < NW; ++f) v[f] = shA[f][ln];
          } else {
#pragma unroll
            for (int f = 0; f < NW; ++f) v[f] = GS ? shB[GS ? f : 0][ln] : 0.0;
          }
        } else {
          const int role = (before ? 0 : 2) + (col ? 0 : 1);
          const int idx = col ? lr : lc;
#pragma unroll
          for (int f = 0; f < NW; ++f) v[f] = shX[role < NROLE ? role : 0][f][idx];
        }
        return ok;
      };
      double v[NW];
      if (TYPE == 1) {
        bool ok;
        ok = nb(3, v); slot<false, 3>(&own0.met[0], ag[0], v, v + 8, ok, gm1, nvis, r0);
        ok = nb(4, v); slot<false, 4>(&own0.met[3], ag[1], v, v + 8, ok, gm1, nvis, r0);
        ok = nb(5, v); slot<false, 5>(&own0.met[6], ag[2], v, v + 8, ok, gm1, nvis, r0);
      } else {
        bool ok;
        ok = nb(0, v); slot<GS, 0>(&own0.met[0], ag[0], v, v + 8, ok, gm1, nvis, r0);
        ok = nb(1, v); slot<GS, 1>(&own0.met[3], ag[1], v, v + 8, ok, gm1, nvis, r0);
        ok = nb(2, v); slot<GS, 2>(&own0.met[6], ag[2], v, v + 8, ok, gm1, nvis, r0);
        if (GS) {
          ok = nb(3, v); slot<true, 3>(&own0.met[0], ag[0], v, v + 8, ok, gm1, nvis, r0);
          ok = nb(4, v); slot<true, 4>(&own0.met[3], ag[1], v, v + 8, ok, gm1, nvis, r0);
          ok = nb(5, v); slot<true, 5>(&own0.met[6], ag[2], v, v + 8, ok, gm1, nvis, r0);
        }
      }
      const double wmig = 1.0;
      if (TYPE == 0) {
#pragma unroll
        for (int f = 0; f < 5; ++f) dqn[f] = (-own0.r5[f] + wmig * r0[f]) * own0.coed;
      } else if (TYPE == 1) {
#pragma unroll
        for (int f = 0; f < 5; ++f) dqn[f] = own0.r5[f] - (wmig * r0[f]) * own0.coed;
      } else {
#pragma unroll
        for (int f = 0; f < 5; ++f) dqn[f] = (wmig * r0[f] + -own0.r5[f]) * own0.coed;
      }
      const long long q = rb0 + (long long)t * qstep + c;
      double* dqp = reinterpret_cast<double*>(S.dq);
#pragma unroll
      for (int f = 0; f < 5; ++f)
        __hip_atomic_store(&dqp[sweep_field(q, c, ncp, SW_DQ, f)], dqn[f], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const bool publish = (t % pub == pub - 1) || t == nsteps - 1;
    if (publish) __builtin_amdgcn_s_waitcnt(0);  // this wave's agent-coherent stores have completed
    __syncthreads();
    // phase 3
    if (publish && lane == 0)
      __hip_atomic_store(&flags[tix], ep + plane_key(FWD, s), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int f = 0; f < 8; ++f) shA[f][lane] = Dcur[f];
#pragma unroll
    for (int f = 0; f < 5; ++f) shA[8 + f][lane] = dqn[f];
    if (t + 2 < nsteps) ensure_up(s + dir);  // edge values of step t+2 lie on plane s(t+1)
    double xv2[NW];
    if (xloader) xload(t + 2, xv2);
    Own own2;
    double nb3[NB];
    load_own<TYPE>(S, cell_rb(t + 2), cell_c(t + 2), ncp, own2);
    load_nbr<GS>(S, cell_rb(t + 3), cell_c(t + 3), ncp, nb3);
    read_up();
#pragma unroll
    for (int f = 0; f < 8; ++f) Dcur[f] = nb1[f];
    own0 = own1;
    own1 = own2;
#pragma unroll
    for (int f = 0; f < NB; ++f) {
      nb1[f] = nb2[f];
      nb2[f] = nb3[f];
    }
#pragma unroll
    for (int f = 0; f < NW; ++f) {
      xv0[f] = xv1[f];
      xv1[f] = xloader ? xv2[f] : 0.0;
    }
  }
}

__global__ void epoch_kernel(unsigned long long* epoch) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *epoch += 1;
}

static inline unsigned grid_of(long long n, int tb) { return (unsigned)((n + tb - 1) / tb); }

void launch_lu_geom(const Launch& La, long long n, const SweepArrays& S) {
  if (n <= 0) return;
  hipLaunchKernelGGL(lu_geom_kernel, dim3((unsigned)La.nwg), dim3(SPAN_TB), 0, La.stream, La.blocks, La.span, S);
}

void launch_lu_pack(const Launch& La, long long n, const SweepArrays& S, const Params& p) {
  if (n <= 0) return;
  hipLaunchKernelGGL(lu_pack_kernel, dim3((unsigned)La.nwg), dim3(SPAN_TB), 0, La.stream, La.blocks, La.span, S, p);
}

void launch_update(const Launch& La, long long n, const SweepArrays& S, double gamma) {
  if (n <= 0) return;
  hipLaunchKernelGGL(update_kernel, dim3((unsigned)La.nwg), dim3(SPAN_TB), 0, La.stream, La.blocks, La.span, S,
                     gamma);
}

void launch_lusgs_tiles(const SweepArrays& S, const BlockDev* blocks, const SweepTile* tiles, const int* order,
                        int ntiles, unsigned long long* flags, unsigned long long* epoch, int* err, int pub, int type,
                        const Params& p, hipStream_t s) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(epoch_kernel, dim3(1), dim3(64), 0, s, epoch);
  const double gm1 = p.gama - 1.0;
  const dim3 gr(ntiles), tb(TL);
  switch (type) {
    case 0: hipLaunchKernelGGL(lusgs_tile_kernel<0>, gr, tb, 0, s, S, blocks, tiles, order, flags, epoch, err, pub, gm1, p.nvis); break;
    case 1: hipLaunchKernelGGL(lusgs_tile_kernel<1>, gr, tb, 0, s, S, blocks, tiles, order, flags, epoch, err, pub, gm1, p.nvis); break;
    case 2: hipLaunchKernelGGL(lusgs_tile_kernel<2>, gr, tb, 0, s, S, blocks, tiles, order, flags, epoch, err, pub, gm1, p.nvis); break;
    default: hipLaunchKernelGGL(lusgs_tile_kernel<3>, gr, tb, 0, s, S, blocks, tiles, order, flags, epoch, err, pub, gm1, p.nvis); break;
  }
}

}  // namespace orion_hip
