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
//    values.  Visibility (argument `sync`): 2 (default) stores dq with
//    agent-scope atomic stores, waits for their completion before the flag
//    store, and reads edge dq with agent-scope atomic loads, so no cache
//    maintenance is needed; 1 uses agent-scope release/acquire fences and
//    0 sequentially consistent fences around plain stores (both flush and
//    invalidate L2, several times slower).
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
#pragma unroll
  for (int f = 0; f < 9; ++f) S.met[f * S.n + q] = b.met[f * b.P + id];
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
  const long long N = S.n;
  const long long P = b.P;
  const long long id = pidx(b, i, j, k);
  const double gamma = p.gama, re = p.reynolds;
  // rhs copy with the physical-BC windows zeroed (StateUpdater.cpp:557-570)
  const bool fz = b.frozen[id] != 0;
#pragma unroll
  for (int f = 0; f < 5; ++f) S.rhs[f * N + q] = fz ? 0.0 : b.dq[f * P + id];
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
    S.coed[q] = 1.0 / term;
  }
  // this cell seen as a neighbour: compute_flux_splitting's per-cell terms,
  // the sound speed of calc_spectral_radius and the viscous coefficient of
  // calc_visc_spectral_radius (same clamps)
  double rm = pr[0], pm = pr[4];
  const double v2 = pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3];
  if (rm < 1e-30) rm = 1e-30;
  if (pm < 1e-30) pm = 1e-30;
  const double c2 = gamma * pm / rm;
  S.D[0 * N + q] = pr[1];
  S.D[1 * N + q] = pr[2];
  S.D[2 * N + q] = pr[3];
  S.D[3 * N + q] = c2;
  S.D[4 * N + q] = (gamma / (gamma - 1.0)) * (pm / rm) + 0.5 * v2;
  S.D[5 * N + q] = 0.5 * (gamma - 1.0) * v2;
  S.D[6 * N + q] = sqrt(c2);
  S.D[7 * N + q] = 2.0 * mu / (fabs(re) * (rm * vol) + 1.0e-30);
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
      d[f] = S.dq[f * S.n + q];
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

template <int TYPE>
__device__ __forceinline__ void load_own(const SweepArrays& S, long long q, bool ok, Own& o) {
  const long long N = S.n;
  if (!ok) q = 0;
#pragma unroll
  for (int f = 0; f < 9; ++f) o.met[f] = ok ? S.met[f * N + q] : 0.0;
  o.coed = ok ? S.coed[q] : 0.0;
#pragma unroll
  for (int f = 0; f < 5; ++f) o.r5[f] = ok ? (TYPE == 1 ? S.dq[f * N + q] : S.rhs[f * N + q]) : 0.0;
}

template <bool GS>
__device__ __forceinline__ void load_nbr(const SweepArrays& S, long long q, bool ok, double* x) {
  // D (8) and, for the Gauss-Seidel sweeps, the cell's old dq (5)
  const long long N = S.n;
  if (!ok) q = 0;
#pragma unroll
  for (int f = 0; f < 8; ++f) x[f] = ok ? S.D[f * N + q] : 0.0;
#pragma unroll
  for (int f = 0; f < 5; ++f) x[8 + f] = (GS && ok) ? S.dq[f * N + q] : 0.0;
}

// Sweep types: 0 LU forward, 1 LU backward, 2 GS forward, 3 GS backward.
// WAVES: the minimum resident waves per SIMD the register allocation targets
// (2: two tiles per CU; 1: one tile per CU and no spills).
template <int TYPE, int WAVES>
__global__ void __launch_bounds__(TL, WAVES) lusgs_tile_kernel(SweepArrays S, const BlockDev* __restrict__ blocks,
                                                       const SweepTile* __restrict__ tiles,
                                                       const int* __restrict__ order, unsigned long long* flags,
                                                       const unsigned long long* __restrict__ epoch, int* err,
                                                       int pub, int sync, double gm1, int nvis) {
  constexpr bool FWD = (TYPE == 0 || TYPE == 2);
  constexpr bool GS = TYPE >= 2;
  constexpr int dir = FWD ? 1 : -1;
  constexpr int NROLE = GS ? 4 : 2;
  __shared__ double shA[NW][TL];               // (D, new dq) of each lane's cell of the last step
  __shared__ double shB[GS ? NW : 1][TL];      // (D, old dq) of each lane's cell of the next step
  __shared__ double shX[NROLE][NW][TW];        // edge values from neighbouring tiles

  const int tix = order[blockIdx.x];
  const SweepTile T = tiles[tix];
  const BlockDev& b = blocks[T.block];
  const int nm = b.sn[0], nc = b.sn[1], nr = b.sn[2];
  const int am = b.sax[0], ac = b.sax[1];
  const long long N = S.n;
  const int lane = threadIdx.x, lc = lane % TW, lr = lane / TW;
  const int c = T.c0 + lc, r = T.r0 + lr;
  const bool lane_ok = c < nc && r < nr;
  // upstream tiles (their cells precede ours in sweep order) and downstream
  const int up_c = FWD ? T.nbr[0] : T.nbr[2], up_r = FWD ? T.nbr[1] : T.nbr[3];
  const unsigned long long ep = (*epoch) << KEY_BITS;
  // plane ranges of the upstream tiles (for the clamped wait targets)
  SweepTile ut[2];
  ut[0] = up_c >= 0 ? tiles[up_c] : T;
  ut[1] = up_r >= 0 ? tiles[up_r] : T;
  const int nsteps = T.s_last - T.s_first + 1;
  const int s0 = FWD ? T.s_first : T.s_last;

  // edge loaders: lanes 0..16*NROLE-1 of the first wave, one (role, index)
  // each.  Roles: 0 col-upstream, 1 row-upstream, 2 col-downstream, 3
  // row-downstream; the upstream ones read new values, the others old ones.
  const int xrole = lane / TW, xi = lane % TW;
  const bool xloader = lane < NROLE * TW;
  auto xload = [&](int sig, double* xv) {
    bool ok = false;
    long long q = 0;
    if (xloader) {
      const bool upr = xrole < 2;
      const int sp = upr ? sig - dir : sig + dir;       // plane of the neighbour cells
      const int off = upr ? -dir : dir;
      int rr, cc;
      if ((xrole & 1) == 0) {  // across the col edge
        cc = (upr == FWD ? T.c0 : T.c0 + TW - 1) + off;
        rr = T.r0 + xi;
      } else {  // across the row edge
        rr = (upr == FWD ? T.r0 : T.r0 + TW - 1) + off;
        cc = T.c0 + xi;
      }
      const int mm = sp - rr - cc;
      ok = rr >= 0 && rr < nr && cc >= 0 && cc < nc && mm >= 0 && mm < nm;
      if (ok) q = sweep_index(b, mm, rr, cc);
      if (ok) {
#pragma unroll
        for (int f = 0; f < 8; ++f) xv[f] = S.D[f * N + q];
        if (upr) {
#pragma unroll
          for (int f = 0; f < 5; ++f)
            xv[8 + f] = __hip_atomic_load(&S.dq[f * N + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
#pragma unroll
          for (int f = 0; f < 5; ++f) xv[8 + f] = S.dq[f * N + q];
        }
      }
    }
    if (!ok) {
#pragma unroll
      for (int f = 0; f < NW; ++f) xv[f] = 0.0;
    }
  };
  // wait (thread 0) until the upstream tiles have published plane s
  unsigned long long seen_c = 0, seen_r = 0;
  auto wait_up = [&](int s) {
    if (lane != 0) return;
    bool got = false;
    for (int q = 0; q < 2; ++q) {
      const int u = q ? up_r : up_c;
      if (u < 0) continue;
      const unsigned long long k = need_key(FWD, ut[q], s);
      if (k == 0) continue;
      const unsigned long long want = ep + k;
      unsigned long long& seen = q ? seen_r : seen_c;
      if (seen >= want) continue;
      unsigned long long v = __hip_atomic_load(&flags[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
        __builtin_amdgcn_s_sleep(2);
        v = __hip_atomic_load(&flags[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      seen = v;
      got = true;
    }
    if (got) {
      if (sync == 0) __threadfence();
      else if (sync == 1) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
  };

  // own inputs of step t (cur), the neighbour view (D, old dq) of this
  // lane's cell of step t+1 (nxt); each step loads the own inputs of t+1 and
  // the neighbour view of t+2
  auto cell_of = [&](int t, int* m) {
    const int s = s0 + dir * t;
    *m = s - r - c;
    return lane_ok && t >= 0 && t < nsteps && *m >= 0 && *m < nm;
  };
  Own cur;
  double nxt[NW];
  double Dcur[8];
  {
    int m0, m1;
    const bool ok0 = cell_of(0, &m0), ok1 = cell_of(1, &m1);
    load_own<TYPE>(S, ok0 ? sweep_index(b, m0, r, c) : 0, ok0, cur);
    double t0[NW];
    load_nbr<false>(S, ok0 ? sweep_index(b, m0, r, c) : 0, ok0, t0);
#pragma unroll
    for (int f = 0; f < 8; ++f) Dcur[f] = t0[f];
    load_nbr<GS>(S, ok1 ? sweep_index(b, m1, r, c) : 0, ok1, nxt);
  }
#pragma unroll
  for (int f = 0; f < NW; ++f) shA[f][lane] = 0.0;
  double xv[NW];
  wait_up(s0 - dir);
  __syncthreads();
  xload(s0, xv);

  for (int t = 0; t < nsteps; ++t) {
    const int s = s0 + dir * t;
    int m;
    const bool act = cell_of(t, &m);
    // phase 1: publish edge values and (GS) next-cell values, wait upstream
    if (xloader) {
#pragma unroll
      for (int f = 0; f < NW; ++f) shX[xrole][f][xi] = xv[f];
    }
    if (GS) {
#pragma unroll
      for (int f = 0; f < NW; ++f) shB[f][lane] = nxt[f];
    }
    wait_up(s);
    __syncthreads();
    // phase 2: loads for later steps, then this cell
    xload(s + dir, xv);
    Own own_n;
    double nxt2[NW];
    {
      int m1, m2;
      const bool ok1 = cell_of(t + 1, &m1), ok2 = cell_of(t + 2, &m2);
      load_own<TYPE>(S, ok1 ? sweep_index(b, m1, r, c) : 0, ok1, own_n);
      load_nbr<GS>(S, ok2 ? sweep_index(b, m2, r, c) : 0, ok2, nxt2);
    }
    double dqn[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (act) {
      double r0[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      AxisGeo ag[3];
      axis_geo(&cur.met[0], ag[0]);
      axis_geo(&cur.met[3], ag[1]);
      axis_geo(&cur.met[6], ag[2]);
      // the lane's previous cell is its own entry of shA; for the
      // Gauss-Seidel sweeps its next cell is its own entry of shB
      // neighbour of real slot sl: returns its (D, dq) and validity
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
        ok = nb(3, v); slot<false, 3>(&cur.met[0], ag[0], v, v + 8, ok, gm1, nvis, r0);
        ok = nb(4, v); slot<false, 4>(&cur.met[3], ag[1], v, v + 8, ok, gm1, nvis, r0);
        ok = nb(5, v); slot<false, 5>(&cur.met[6], ag[2], v, v + 8, ok, gm1, nvis, r0);
      } else {
        bool ok;
        ok = nb(0, v); slot<GS, 0>(&cur.met[0], ag[0], v, v + 8, ok, gm1, nvis, r0);
        ok = nb(1, v); slot<GS, 1>(&cur.met[3], ag[1], v, v + 8, ok, gm1, nvis, r0);
        ok = nb(2, v); slot<GS, 2>(&cur.met[6], ag[2], v, v + 8, ok, gm1, nvis, r0);
        if (GS) {
          ok = nb(3, v); slot<true, 3>(&cur.met[0], ag[0], v, v + 8, ok, gm1, nvis, r0);
          ok = nb(4, v); slot<true, 4>(&cur.met[3], ag[1], v, v + 8, ok, gm1, nvis, r0);
          ok = nb(5, v); slot<true, 5>(&cur.met[6], ag[2], v, v + 8, ok, gm1, nvis, r0);
        }
      }
      const double wmig = 1.0;
      if (TYPE == 0) {
#pragma unroll
        for (int f = 0; f < 5; ++f) dqn[f] = (-cur.r5[f] + wmig * r0[f]) * cur.coed;
      } else if (TYPE == 1) {
#pragma unroll
        for (int f = 0; f < 5; ++f) dqn[f] = cur.r5[f] - (wmig * r0[f]) * cur.coed;
      } else {
#pragma unroll
        for (int f = 0; f < 5; ++f) dqn[f] = (wmig * r0[f] + -cur.r5[f]) * cur.coed;
      }
      const long long q = sweep_index(b, m, r, c);
      if (sync == 2) {
#pragma unroll
        for (int f = 0; f < 5; ++f) __hip_atomic_store(&S.dq[f * N + q], dqn[f], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
#pragma unroll
        for (int f = 0; f < 5; ++f) S.dq[f * N + q] = dqn[f];
      }
    }
    const bool publish = (t % pub == pub - 1) || t == nsteps - 1;
    if (publish) {
      if (sync == 0) __threadfence();
      else if (sync == 1) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      else __builtin_amdgcn_s_waitcnt(0);  // this wave's agent-coherent stores have completed
    }
    __syncthreads();
    // phase 3: progress flag, this step's values for the next step
    if (publish && lane == 0)
      __hip_atomic_store(&flags[tix], ep + plane_key(FWD, s), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int f = 0; f < 8; ++f) shA[f][lane] = Dcur[f];
#pragma unroll
    for (int f = 0; f < 5; ++f) shA[8 + f][lane] = dqn[f];
#pragma unroll
    for (int f = 0; f < 8; ++f) Dcur[f] = nxt[f];
    cur = own_n;
#pragma unroll
    for (int f = 0; f < NW; ++f) nxt[f] = nxt2[f];
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
                        int ntiles, unsigned long long* flags, unsigned long long* epoch, int* err, int pub, int sync,
                        int waves, int type, const Params& p, hipStream_t s) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(epoch_kernel, dim3(1), dim3(64), 0, s, epoch);
  const double gm1 = p.gama - 1.0;
  const dim3 gr(ntiles), tb(TL);
#define ORION_TILE_LAUNCH(T, W) \
  hipLaunchKernelGGL((lusgs_tile_kernel<T, W>), gr, tb, 0, s, S, blocks, tiles, order, flags, epoch, err, pub, sync, gm1, p.nvis)
  const bool one = waves == 1;
  switch (type) {
    case 0: if (one) ORION_TILE_LAUNCH(0, 1); else ORION_TILE_LAUNCH(0, 2); break;
    case 1: if (one) ORION_TILE_LAUNCH(1, 1); else ORION_TILE_LAUNCH(1, 2); break;
    case 2: if (one) ORION_TILE_LAUNCH(2, 1); else ORION_TILE_LAUNCH(2, 2); break;
    default: if (one) ORION_TILE_LAUNCH(3, 1); else ORION_TILE_LAUNCH(3, 2); break;
#undef ORION_TILE_LAUNCH
  }
}

}  // namespace orion_hip
