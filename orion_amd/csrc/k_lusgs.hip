// k_lusgs.hip -- LU-SGS sweeps for gfx950 (FP64).
//
// StateUpdater::update_solution (src/solver/StateUpdater.cpp:297-445) sweeps
// each block lexicographically (k, then j, then i; reversed for the backward
// sweeps) and every cell reads only its six face neighbours.  On the
// hyperplane s = i+j+k, lower neighbours lie on s-1 and upper ones on s+1,
// so processing planes in increasing (decreasing) s with all cells of a
// plane in parallel reads exactly the same neighbour values as the
// lexicographic sweep: the results are identical, not an approximation.
//
// Sweep types: 0 LU forward, 1 LU backward, 2 GS forward, 3 GS backward.
// Per-cell arithmetic follows calc_spectral_radius (:175-191),
// calc_visc_spectral_radius (:194-207), compute_flux_splitting (:209-265),
// get_diag_coeff (:267-295) and the sweep bodies (:318-365, :392-441).
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

__device__ inline void flux_split(const double* pr, const double* dq, double nx, double ny, double nz, double sr,
                                  int sign, double gamma, double* res) {
  double rm = pr[0], pm = pr[4];
  const double um = pr[1], vm = pr[2], wm = pr[3];
  const double drm = dq[0], dum = dq[1], dvm = dq[2], dwm = dq[3], dEm = dq[4];
  const double nt = 0.0;
  const double v2 = um * um + vm * vm + wm * wm;
  if (rm < 1e-30) rm = 1e-30;
  if (pm < 1e-30) pm = 1e-30;
  const double c2 = gamma * pm / rm;
  const double cm = sqrt(c2);
  const double H = (gamma / (gamma - 1.0)) * (pm / rm) + 0.5 * v2;
  const double ct = nx * um + ny * vm + nz * wm + nt;
  double gm = sqrt(nx * nx + ny * ny + nz * nz);
  if (gm < 1.0e-30) gm = 1.0e-30;
  const double ssr = (double)sign * sr;
  const double l1 = 0.5 * (ct + ssr);
  const double l4 = 0.5 * ((ct + cm * gm) + ssr);
  const double l5 = 0.5 * ((ct - cm * gm) + ssr);
  const double x1 = (2.0 * l1 - l4 - l5) / (2.0 * c2);
  const double x2 = (l4 - l5) / (2.0 * cm);
  const double im = 1.0 / gm;
  const double nxn = nx * im, nyn = ny * im, nzn = nz * im, ctn = (ct - nt) * im;
  const double gm1 = gamma - 1.0;
  const double af = 0.5 * gm1 * v2;
  const double dc = ctn * drm - (nxn * dum + nyn * dvm + nzn * dwm);
  const double dh = af * drm - gm1 * (um * dum + vm * dvm + wm * dwm - dEm);
  const double c2dc = c2 * dc;
  res[0] = l1 * drm - dh * x1 - dc * x2;
  res[1] = l1 * dum + (nxn * c2dc - um * dh) * x1 + (nxn * dh - um * dc) * x2;
  res[2] = l1 * dvm + (nyn * c2dc - vm * dh) * x1 + (nyn * dh - vm * dc) * x2;
  res[3] = l1 * dwm + (nzn * c2dc - wm * dh) * x1 + (nzn * dh - wm * dc) * x2;
  res[4] = l1 * dEm + (ctn * c2dc - H * dh) * x1 + (ctn * dh - H * dc) * x2;
}

__global__ void __launch_bounds__(256) lusgs_plane_kernel(const BlockDev* __restrict__ blocks,
                                                          const long long* __restrict__ off, int nb, int plane,
                                                          int type, Params p) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= off[nb]) return;
  const int bi = find_block(off, nb, g);
  const BlockDev& b = blocks[bi];
  const long long n = g - off[bi];
  const int jj = (int)(n % b.jdim), kk = (int)(n / b.jdim);
  const int ii = plane - jj - kk;
  if (ii < 0 || ii >= b.idim) return;
  const int i = ii + NG, j = jj + NG, k = kk + NG;
  const long long P = b.P;
  const long long id = pidx(b, i, j, k);
  const double gamma = p.gama, re = p.reynolds;
  const int nvis = p.nvis;
  // get_diag_coeff
  double pr[5];
#pragma unroll
  for (int m = 0; m < 5; ++m) pr[m] = b.prim[m * P + id];
  double met[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) met[m] = b.met[m * P + id];
  const double ra = spec_rad(pr, gamma, met[0], met[1], met[2]);
  const double rb = spec_rad(pr, gamma, met[3], met[4], met[5]);
  const double rc = spec_rad(pr, gamma, met[6], met[7], met[8]);
  double rva = 0, rvb = 0, rvc = 0;
  if (nvis == 1) {
    const double mu = b.mu[id];
    double vol = b.vol[id];
    if (vol < 1e-30) vol = 1e-30;
    rva = visc_rad(pr[0], mu, vol, re, met[0], met[1], met[2]);
    rvb = visc_rad(pr[0], mu, vol, re, met[3], met[4], met[5]);
    rvc = visc_rad(pr[0], mu, vol, re, met[6], met[7], met[8]);
  }
  const double rs = ra + rb + rc + rva + rvb + rvc;
  const double term = (1.0 / b.dt[id]) + 1.0 * 1.0 * rs;
  const double coed = 1.0 / term;

  double r0[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const bool gs = type >= 2;
  const int dir = (type == 0 || type == 2) ? 1 : -1;
  // neighbour list: LU uses the three on the sweep's trailing side with
  // sign=dir; GS uses lower (sign +1) then upper (sign -1).
  const int nnb = gs ? 6 : 3;
  for (int t = 0; t < nnb; ++t) {
    int ax, sgn, step;
    if (gs) {
      ax = t % 3;
      sgn = t < 3 ? 1 : -1;
      step = t < 3 ? -1 : 1;
    } else {
      ax = t;
      sgn = dir;
      step = -dir;
    }
    int in = i, jn = j, kn = k;
    if (ax == 0) in += step;
    else if (ax == 1) jn += step;
    else kn += step;
    if (in < NG || in >= b.nx - NG || jn < NG || jn >= b.ny - NG || kn < NG || kn >= b.nz - NG) continue;
    const long long idn = pidx(b, in, jn, kn);
    double pn[5], dn[5], fs[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      pn[m] = b.prim[m * P + idn];
      dn[m] = b.dq[m * P + idn];
    }
    const double nx = met[3 * ax], ny = met[3 * ax + 1], nz = met[3 * ax + 2];
    const double r = spec_rad(pn, gamma, nx, ny, nz);
    flux_split(pn, dn, nx, ny, nz, r, sgn, gamma, fs);
    if (nvis == 1) {
      const double mun = b.mu[idn];
      double voln = b.vol[idn];
      if (voln < 1e-30) voln = 1e-30;
      const double rv = visc_rad(pn[0], mun, voln, re, nx, ny, nz);
      if (gs) {
#pragma unroll
        for (int m = 0; m < 5; ++m) fs[m] += 0.5 * rv * dn[m] * (double)sgn;
      } else {
#pragma unroll
        for (int m = 0; m < 5; ++m) fs[m] += 0.5 * rv * dn[m];
      }
    }
    if (gs && sgn == -1) {
#pragma unroll
      for (int m = 0; m < 5; ++m) r0[m] -= fs[m];
    } else {
#pragma unroll
      for (int m = 0; m < 5; ++m) r0[m] += fs[m];
    }
  }
  const double wmig = 1.0;
  if (type == 0) {
#pragma unroll
    for (int m = 0; m < 5; ++m) b.dq[m * P + id] = (-b.rhs[m * P + id] + wmig * r0[m]) * coed;
  } else if (type == 1) {
#pragma unroll
    for (int m = 0; m < 5; ++m) b.dq[m * P + id] = b.dq[m * P + id] - (wmig * r0[m]) * coed;
  } else {
#pragma unroll
    for (int m = 0; m < 5; ++m) b.dq[m * P + id] = (wmig * r0[m] + -b.rhs[m * P + id]) * coed;
  }
}

static inline unsigned grid_of(long long n, int tb) { return (unsigned)((n + tb - 1) / tb); }

void launch_lusgs_plane(const Launch& L, long long n, int plane, int type, const Params& p) {
  if (n <= 0) return;
  hipLaunchKernelGGL(lusgs_plane_kernel, dim3(grid_of(n, 256)), dim3(256), 0, L.stream, L.blocks, L.off, L.nb, plane,
                     type, p);
}

}  // namespace orion_hip
