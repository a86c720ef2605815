// k_flux.hip -- right-hand-side kernels for gfx950 (FP64).
//
// Inviscid RHS (src/solver/InviscidFluxComputer.cpp:14-647):
//   inv_face    one thread per face along direction d: WCNS-E-5 left/right
//               states (fixed boundary stencils where the third ghost rho is
//               below 1e-20), low-order fallback, Roe flux with Harten fix.
//               Face f of a line is stored at the padded index of its left
//               node (position f+2 along d).
//   inv_closure one thread per line whose end carries the boundary signal:
//               replays the reference's boundary recalculation (:377-521) on
//               a private window and writes the three end-node differences.
//   inv_diff    one thread per interior node: dq += fc_xi, fc_eta, fc_zeta in
//               that order (6th-order difference, 3rd/4th-order ends, or the
//               closure value).
// Viscous RHS (src/solver/FluxComputer.cpp:13-732):
//   visc_deriv  one thread per interior node: 4th-order node derivatives of
//               (u,v,w,T) along xi, eta, zeta (one-sided next to rho<0 ghosts).
//   visc_face   one thread per face: face interpolation of (u,v,w,T) (flagged),
//               of mu, vol, metrics and cross derivatives (always central,
//               so the zero ghost metrics and volumes enter the end faces),
//               face derivative, chain rule, stress and heat flux.
//   visc_diff   one thread per interior node: dq -= Re^-1 dFv per direction.
//
// Expression forms follow the cited reference lines operand for operand.
#include "orion_dev.h"
#include "orion_kernels.h"

namespace orion_hip {

// ---------------------------------------------------------------------------
// Roe flux, InviscidFluxComputer.cpp:14-131
// ---------------------------------------------------------------------------
__device__ inline void phys_flux(const double* q, double g, double nx, double ny, double nz, double nt, double* f) {
  const double rho = q[0], u = q[1], v = q[2], w = q[3], p = q[4];
  const double V = u * nx + v * ny + w * nz + nt;
  const double E = p / (g - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
  const double H = (E + p) / rho;
  f[0] = rho * V;
  f[1] = rho * u * V + p * nx;
  f[2] = rho * v * V + p * ny;
  f[3] = rho * w * V + p * nz;
  f[4] = rho * H * V + p * nt;
}

__device__ inline void roe_flux(const double* ql, const double* qr, double nx, double ny, double nz, double nt,
                                double g, double* f) {
  const double efix = 0.1;
  double fl[5], fr[5];
  phys_flux(ql, g, nx, ny, nz, nt, fl);
  phys_flux(qr, g, nx, ny, nz, nt, fr);
  const double drho = qr[0] - ql[0], du = qr[1] - ql[1], dv = qr[2] - ql[2], dw = qr[3] - ql[3];
  const double dp = qr[4] - ql[4];
  const double g1 = g - 1.0;
  const double g2 = g / g1;
  const double hl = g2 * ql[4] / ql[0] + 0.5 * (ql[1] * ql[1] + ql[2] * ql[2] + ql[3] * ql[3]);
  const double hr = g2 * qr[4] / qr[0] + 0.5 * (qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
  const double sl = sqrt(ql[0]), sr = sqrt(qr[0]);
  const double iden = 1.0 / (sl + sr);
  const double rroe = sl * sr;
  const double uroe = (sl * ql[1] + sr * qr[1]) * iden;
  const double vroe = (sl * ql[2] + sr * qr[2]) * iden;
  const double wroe = (sl * ql[3] + sr * qr[3]) * iden;
  const double hroe = (sl * hl + sr * hr) * iden;
  const double v2 = uroe * uroe + vroe * vroe + wroe * wroe;
  double c2 = (hroe - 0.5 * v2) * g1;
  if (c2 <= 0.0) c2 = 1.0e-6;
  const double c = sqrt(c2);
  const double U = uroe * nx + vroe * ny + wroe * nz + nt;
  double gm = sqrt(nx * nx + ny * ny + nz * nz);
  gm = (gm < 1.0e-30) ? 1.0e-30 : gm;
  const double im = 1.0 / gm;
  const double nxn = nx * im, nyn = ny * im, nzn = nz * im;
  const double cm = c * gm;
  double l1 = fabs(U - cm), l4 = fabs(U), l5 = fabs(U + cm);
  const double de = efix * cm;
  const double de2 = de * de;
  if (l1 < de) l1 = sqrt(l1 * l1 + de2);
  if (l4 < de) l4 = sqrt(l4 * l4 + de2);
  if (l5 < de) l5 = sqrt(l5 * l5 + de2);
  const double rodcta = rroe * (du * nxn + dv * nyn + dw * nzn);
  const double dpc2 = dp / c2;
  const double rodctac = rodcta / c;
  const double a1 = l4 * (drho - dpc2);
  const double a2 = l5 * (dpc2 + rodctac) * 0.5;
  const double a3 = l1 * (dpc2 - rodctac) * 0.5;
  const double a4 = a1 + a2 + a3;
  const double a5 = c * (a2 - a3);
  const double a6 = l4 * (rroe * du - nxn * rodcta);
  const double a7 = l4 * (rroe * dv - nyn * rodcta);
  const double a8 = l4 * (rroe * dw - nzn * rodcta);
  double df[5];
  df[0] = a4;
  df[1] = uroe * a4 + nxn * a5 + a6;
  df[2] = vroe * a4 + nyn * a5 + a7;
  df[3] = wroe * a4 + nzn * a5 + a8;
  df[4] = hroe * a4 + U * im * a5 + uroe * a6 + vroe * a7 + wroe * a8 - c2 * a1 / g1;
#pragma unroll
  for (int m = 0; m < 5; ++m) f[m] = 0.5 * (fl[m] + fr[m] - df[m]);
}

// Correctly rounded quotients a / d over one denominator d.  The reciprocal
// is refined once from the hardware estimate by two Newton-Raphson steps,
// r <- r + r (1 - d r), and each quotient is corrected with its exact
// residual (Markstein): q = a r, e = fma(-d, q, a), q' = fma(e, r, q).
// These are the steps of the correctly rounded division sequence without its
// exponent pre-scaling and special-value fix-up, which change nothing while
// d, a and a / d stay well inside the normal range (the callers check).
// So each quotient is the IEEE a / d, and quotients sharing d share the
// reciprocal.
struct SharedDiv {
  double d, r;
  __device__ __forceinline__ explicit SharedDiv(double den) : d(den) {
    double x = __builtin_amdgcn_rcp(den);
    double e = __builtin_fma(-den, x, 1.0);
    x = __builtin_fma(x, e, x);
    e = __builtin_fma(-den, x, 1.0);
    r = __builtin_fma(x, e, x);
  }
  __device__ __forceinline__ double operator()(double a) const {
    const double q = a * r;
    const double e = __builtin_fma(-d, q, a);
    return __builtin_fma(e, r, q);
  }
};

// WCNS-E-5 left/right states of one variable at the face between line
// positions p+2 and p+3, from q[p..p+5] (InviscidFluxComputer.cpp:268-314).
// SHARED: with smoothness indicators below 1e100 every divisor lies in
// [1e-12, 1e200] and every weight in [6e-202, 1], so the weights are formed
// with SharedDiv (one reciprocal for the three normalised weights of a
// side); otherwise (never for a physical state; NaN included), and without
// SHARED (kernels whose register budget is tight), with '/'.
template <bool SHARED>
__device__ inline void weno5(double qm2, double qm1, double q0, double qp1, double qp2, double qp3, double* ql,
                             double* qr) {
  const double CL1 = 1.0 / 16.0, CL2 = 10.0 / 16.0, CL3 = 5.0 / 16.0, EPS = 1.0e-6;
  const double s1 = qm2 - 2.0 * qm1 + q0;
  const double s2 = qm1 - 2.0 * q0 + qp1;
  const double s3 = q0 - 2.0 * qp1 + qp2;
  const double g1 = 0.5 * (qm2 - 4.0 * qm1 + 3.0 * q0);
  const double g2 = 0.5 * (qp1 - qm1);
  const double g3 = 0.5 * (-3.0 * q0 + 4.0 * qp1 - qp2);
  const double is1 = g1 * g1 + s1 * s1, is2 = g2 * g2 + s2 * s2, is3 = g3 * g3 + s3 * s3;
  const double r1s = s2, r2s = s3, r3s = qp1 - 2.0 * qp2 + qp3;
  const double h1 = 0.5 * (qm1 - 4.0 * q0 + 3.0 * qp1);
  const double h2 = 0.5 * (qp2 - q0);
  const double h3 = 0.5 * (-3.0 * qp1 + 4.0 * qp2 - qp3);
  const double js1 = h1 * h1 + r1s * r1s, js2 = h2 * h2 + r2s * r2s, js3 = h3 * h3 + r3s * r3s;
  const double ismax = fmax(fmax(fmax(is1, is2), fmax(is3, js1)), fmax(js2, js3));
  if (SHARED && ismax < 1.0e100) {
    const double w1 = SharedDiv((EPS + is1) * (EPS + is1))(CL1);
    const double w2 = SharedDiv((EPS + is2) * (EPS + is2))(CL2);
    const double w3 = SharedDiv((EPS + is3) * (EPS + is3))(CL3);
    const SharedDiv sw(w1 + w2 + w3);
    *ql = q0 + 0.125 * (sw(w1) * (s1 + 4.0 * g1) + sw(w2) * (s2 + 4.0 * g2) + sw(w3) * (s3 + 4.0 * g3));
    const double v1 = SharedDiv((EPS + js1) * (EPS + js1))(CL3);
    const double v2 = SharedDiv((EPS + js2) * (EPS + js2))(CL2);
    const double v3 = SharedDiv((EPS + js3) * (EPS + js3))(CL1);
    const SharedDiv sv(v1 + v2 + v3);
    *qr = qp1 + 0.125 * (sv(v1) * (r1s - 4.0 * h1) + sv(v2) * (r2s - 4.0 * h2) + sv(v3) * (r3s - 4.0 * h3));
  } else {
    const double w1 = CL1 / ((EPS + is1) * (EPS + is1));
    const double w2 = CL2 / ((EPS + is2) * (EPS + is2));
    const double w3 = CL3 / ((EPS + is3) * (EPS + is3));
    const double sw = w1 + w2 + w3;
    *ql = q0 + 0.125 * ((w1 / sw) * (s1 + 4.0 * g1) + (w2 / sw) * (s2 + 4.0 * g2) + (w3 / sw) * (s3 + 4.0 * g3));
    const double v1 = CL3 / ((EPS + js1) * (EPS + js1));
    const double v2 = CL2 / ((EPS + js2) * (EPS + js2));
    const double v3 = CL1 / ((EPS + js3) * (EPS + js3));
    const double sv = v1 + v2 + v3;
    *qr = qp1 + 0.125 * ((v1 / sv) * (r1s - 4.0 * h1) + (v2 / sv) * (r2s - 4.0 * h2) + (v3 / sv) * (r3s - 4.0 * h3));
  }
}

// The prim values of one line: field m at position p (0 .. n+5).  Offsets
// are 32-bit (a block's prim holds 6 fields of P < 2^25 points).
struct LineRef {
  const double* prim;
  unsigned P;
  unsigned base;  // padded index of position 0
  unsigned s;     // stride along the line
  __device__ inline double at(int m, int p) const { return ldo(prim, boff((unsigned)m, P, base + (unsigned)p * s)); }
};

// Initial face states of face f (after the WCNS pass and the fixed boundary
// stencils, InviscidFluxComputer.cpp:262-337).  The WCNS stencil of every
// face lies inside the padded line, so it is evaluated for all faces -- its
// loads then need not wait for the boundary flags -- and the faces outside
// [ist, ied] (f < 2 or f > n - 2) have both states replaced below, exactly
// as the reference leaves them to the fixed stencils.
// The fixed boundary stencils' states of face f where they replace the WCNS
// states (face_states after its WCNS pass).
__device__ inline void boundary_states(const LineRef& L, int n, int ist, int ied, int f, double* qwl, double* qwr) {
  if (ist > 1 && f <= 2) {
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      const double q1 = L.at(m, 3), q2 = L.at(m, 4), q3 = L.at(m, 5), q4 = L.at(m, 6);
      if (f == 0) qwl[m] = (35. * q1 - 35. * q2 + 21. * q3 - 5. * q4) / 16.;
      else if (f == 1) qwl[m] = (5. * q1 + 15. * q2 - 5. * q3 + q4) / 16.;
      else qwl[m] = (-q1 + 9. * q2 + 9. * q3 - q4) / 16.;
      if (f <= 1) qwr[m] = qwl[m];
    }
  }
  if (ied < n && f >= n - 2) {
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      const double qn = L.at(m, n + 2), qn1 = L.at(m, n + 1), qn2 = L.at(m, n), qn3 = L.at(m, n - 1);
      if (f == n) qwr[m] = (35. * qn - 35. * qn1 + 21. * qn2 - 5. * qn3) / 16.;
      else if (f == n - 1) qwr[m] = (5. * qn + 15. * qn1 - 5. * qn2 + qn3) / 16.;
      else qwr[m] = (-qn + 9. * qn1 + 9. * qn2 - qn3) / 16.;
      if (f >= n - 1) qwl[m] = qwr[m];
    }
  }
}

template <bool SHARED>
__device__ inline void face_states(const LineRef& L, int n, int ist, int ied, int f, double* qwl, double* qwr) {
#pragma unroll
  for (int m = 0; m < 5; ++m)
    weno5<SHARED>(L.at(m, f), L.at(m, f + 1), L.at(m, f + 2), L.at(m, f + 3), L.at(m, f + 4), L.at(m, f + 5), &qwl[m],
          &qwr[m]);
  boundary_states(L, n, ist, ied, f, qwl, qwr);
}

// weno5 of two neighbouring faces of a line, face 0 from v[0..5] and face 1
// from v[1..6].  Face 1's left state is built from the three stencils of
// face 0's right state: its s, g and smoothness indicators are face 0's
// r1s, r2s, r3s, h1, h2, h3 and js1..3, formed by the same expressions from
// the same values, so they are computed once; its weight of the middle
// stencil (CL2 / (EPS + js2)^2) is face 0's right weight v2, and its other
// two weights divide by face 0's right denominators, whose reciprocals
// (SharedDiv) serve both quotients.  When either face has an indicator of
// 1e100 or more, each face is formed by weno5 alone.  Either way every
// state is the one weno5 forms for its face, bit for bit.
__device__ inline void weno5_pair(const double* v, double* ql0, double* qr0, double* ql1, double* qr1) {
  const double CL1 = 1.0 / 16.0, CL2 = 10.0 / 16.0, CL3 = 5.0 / 16.0, EPS = 1.0e-6;
  const double s1 = v[0] - 2.0 * v[1] + v[2];
  const double s2 = v[1] - 2.0 * v[2] + v[3];
  const double s3 = v[2] - 2.0 * v[3] + v[4];
  const double g1 = 0.5 * (v[0] - 4.0 * v[1] + 3.0 * v[2]);
  const double g2 = 0.5 * (v[3] - v[1]);
  const double g3 = 0.5 * (-3.0 * v[2] + 4.0 * v[3] - v[4]);
  const double is1 = g1 * g1 + s1 * s1, is2 = g2 * g2 + s2 * s2, is3 = g3 * g3 + s3 * s3;
  const double r3s = v[3] - 2.0 * v[4] + v[5];
  const double h1 = 0.5 * (v[1] - 4.0 * v[2] + 3.0 * v[3]);
  const double h2 = 0.5 * (v[4] - v[2]);
  const double h3 = 0.5 * (-3.0 * v[3] + 4.0 * v[4] - v[5]);
  const double js1 = h1 * h1 + s2 * s2, js2 = h2 * h2 + s3 * s3, js3 = h3 * h3 + r3s * r3s;
  // face 1's right state
  const double r3t = v[4] - 2.0 * v[5] + v[6];
  const double k1 = 0.5 * (v[2] - 4.0 * v[3] + 3.0 * v[4]);
  const double k2 = 0.5 * (v[5] - v[3]);
  const double k3 = 0.5 * (-3.0 * v[4] + 4.0 * v[5] - v[6]);
  const double ks1 = k1 * k1 + s3 * s3, ks2 = k2 * k2 + r3s * r3s, ks3 = k3 * k3 + r3t * r3t;
  const double m0 = fmax(fmax(fmax(is1, is2), fmax(is3, js1)), fmax(js2, js3));
  const double m1 = fmax(fmax(fmax(js1, js2), fmax(js3, ks1)), fmax(ks2, ks3));
  if (m0 < 1.0e100 && m1 < 1.0e100) {
    const double w1 = SharedDiv((EPS + is1) * (EPS + is1))(CL1);
    const double w2 = SharedDiv((EPS + is2) * (EPS + is2))(CL2);
    const double w3 = SharedDiv((EPS + is3) * (EPS + is3))(CL3);
    const SharedDiv sw(w1 + w2 + w3);
    *ql0 = v[2] + 0.125 * (sw(w1) * (s1 + 4.0 * g1) + sw(w2) * (s2 + 4.0 * g2) + sw(w3) * (s3 + 4.0 * g3));
    const SharedDiv dj1((EPS + js1) * (EPS + js1)), dj3((EPS + js3) * (EPS + js3));
    const double v1 = dj1(CL3);
    const double v2 = SharedDiv((EPS + js2) * (EPS + js2))(CL2);
    const double v3 = dj3(CL1);
    const SharedDiv sv(v1 + v2 + v3);
    *qr0 = v[3] + 0.125 * (sv(v1) * (s2 - 4.0 * h1) + sv(v2) * (s3 - 4.0 * h2) + sv(v3) * (r3s - 4.0 * h3));
    const double u1 = dj1(CL1), u3 = dj3(CL3);
    const SharedDiv su(u1 + v2 + u3);
    *ql1 = v[3] + 0.125 * (su(u1) * (s2 + 4.0 * h1) + su(v2) * (s3 + 4.0 * h2) + su(u3) * (r3s + 4.0 * h3));
    const double x1 = SharedDiv((EPS + ks1) * (EPS + ks1))(CL3);
    const double x2 = SharedDiv((EPS + ks2) * (EPS + ks2))(CL2);
    const double x3 = SharedDiv((EPS + ks3) * (EPS + ks3))(CL1);
    const SharedDiv sx(x1 + x2 + x3);
    *qr1 = v[4] + 0.125 * (sx(x1) * (s3 - 4.0 * k1) + sx(x2) * (r3s - 4.0 * k2) + sx(x3) * (r3t - 4.0 * k3));
  } else {
    weno5<true>(v[0], v[1], v[2], v[3], v[4], v[5], ql0, qr0);
    weno5<true>(v[1], v[2], v[3], v[4], v[5], v[6], ql1, qr1);
  }
}

// Half-node metric (kx,ky,kz) at face f, interp_metrics_half (:136-192).
__device__ inline void metric_face(const BlockDev& b, int d, long long base, long long s, int n, int f, double* k3) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double* M = b.met + (long long)(3 * d + c) * b.P + base;
    double v;
    if (f == 0) v = (35.0 * M[3 * s] + -35.0 * M[4 * s] + 21.0 * M[5 * s] + -5.0 * M[6 * s]) / 16.0;
    else if (f == 1) v = (5.0 * M[3 * s] + 15.0 * M[4 * s] + -5.0 * M[5 * s] + 1.0 * M[6 * s]) / 16.0;
    else if (f == n)
      v = (35.0 * M[(n + 2) * s] + -35.0 * M[(n + 1) * s] + 21.0 * M[n * s] + -5.0 * M[(n - 1) * s]) / 16.0;
    else if (f == n - 1)
      v = (5.0 * M[(n + 2) * s] + 15.0 * M[(n + 1) * s] + -5.0 * M[n * s] + 1.0 * M[(n - 1) * s]) / 16.0;
    else
      v = (9.0 * (M[(f + 2) * s] + M[(f + 3) * s]) + -1.0 * (M[(f + 4) * s] + M[(f + 1) * s])) / 16.0;
    k3[c] = v;
  }
}

// Roe flux from face states with the low-order fallback (:342-364).
__device__ inline void face_flux(const LineRef& L, int n, int f, const double* qwl, const double* qwr,
                                 const double* k3, double g, double* out) {
  double a[5], c[5];
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    a[m] = qwl[m];
    c[m] = qwr[m];
  }
  if (a[0] <= 1e-6 || a[4] <= 1e-6) {
    const int ix = (f - 1 > 0 ? f - 1 : 0) + 2;
#pragma unroll
    for (int m = 0; m < 5; ++m) a[m] = L.at(m, ix);
  }
  if (c[0] <= 1e-6 || c[4] <= 1e-6) {
    const int ix = (n - 1 < f ? n - 1 : f) + 2 + 1;
#pragma unroll
    for (int m = 0; m < 5; ++m) c[m] = L.at(m, ix);
  }
  roe_flux(a, c, k3[0], k3[1], k3[2], 0.0, g, out);
}

__global__ void __launch_bounds__(256) inv_face_kernel(const BlockDev* __restrict__ blocks,
                                                       Span SP, int d,
                                                       double gamma) {
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;
  if (d < 0) {  // the three directions in one launch: virtual block d * nb + b
    const int nbr = SP.nb / 3;
    d = bi / nbr;
    bi -= d * nbr;
  }
  const BlockDev& b = blocks[bi];
  long long base;
  int f;
  decode_face(b, d, gi, &base, &f);
  const int n = dcount(b, d);
  const unsigned s = (unsigned)dstride(b, d);
  const unsigned P = (unsigned)b.P;
  const LineRef L{b.prim, P, (unsigned)base, s};
  const int ist = (L.at(0, 0) < 1.0e-20) ? 2 : 0;
  const int ied = (L.at(0, n + 5) < 1.0e-20) ? n - 2 : n;
  double qwl[5], qwr[5], k3[3], fo[5];
  face_states<true>(L, n, ist, ied, f, qwl, qwr);
  const unsigned g = (unsigned)base + (unsigned)(f + 2) * s;  // the face's slot
  {
    const double* G = b.fgeo + (long long)(d * FG_N + 10) * b.P;  // this direction's inviscid metrics
    k3[0] = ldo(G, boff(0, P, g));
    k3[1] = ldo(G, boff(1, P, g));
    k3[2] = ldo(G, boff(2, P, g));
  }
  face_flux(L, n, f, qwl, qwr, k3, gamma, fo);
  double* F = b.fxi + (long long)d * 5 * b.P;
#pragma unroll
  for (int m = 0; m < 5; ++m) sto(F, boff((unsigned)m, P, g), fo[m]);
}

// The inviscid face fluxes of inv_face_kernel, faces f0 = 2g and f0 + 1 of
// a line per thread (the face-pair space), the two faces' WCNS states formed
// together (weno5_pair).  The pair's stencils cover positions f0 .. f0+6;
// for a last pair with one face (f0 = n) position f0+6 is read at n+5
// instead, so face 1's states are not those of any face and are neither
// differenced into a flux nor stored.
__global__ void __launch_bounds__(256) inv_face_pair_kernel(const BlockDev* __restrict__ blocks, Span SP,
                                                            double gamma) {
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;
  int d;
  {  // the three directions in one launch: virtual block d * nb + b
    const int nbr = SP.nb / 3;
    d = bi / nbr;
    bi -= d * nbr;
  }
  const BlockDev& b = blocks[bi];
  long long base;
  int f0;
  decode_face_pair(b, d, gi, &base, &f0);
  const int n = dcount(b, d);
  const unsigned s = (unsigned)dstride(b, d);
  const unsigned P = (unsigned)b.P;
  const LineRef L{b.prim, P, (unsigned)base, s};
  const int ist = (L.at(0, 0) < 1.0e-20) ? 2 : 0;
  const int ied = (L.at(0, n + 5) < 1.0e-20) ? n - 2 : n;
  const bool two = f0 + 1 <= n;
  const int p6 = two ? f0 + 6 : n + 5;
  double ql0[5], qr0[5], ql1[5], qr1[5];
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    double v[7];
#pragma unroll
    for (int j = 0; j < 6; ++j) v[j] = L.at(m, f0 + j);
    v[6] = L.at(m, p6);
    weno5_pair(v, &ql0[m], &qr0[m], &ql1[m], &qr1[m]);
  }
  const double* G = b.fgeo + (long long)(d * FG_N + 10) * b.P;  // this direction's inviscid metrics
  double* F = b.fxi + (long long)d * 5 * b.P;
  {
    boundary_states(L, n, ist, ied, f0, ql0, qr0);
    const unsigned g = (unsigned)base + (unsigned)(f0 + 2) * s;
    const double k3[3] = {ldo(G, boff(0, P, g)), ldo(G, boff(1, P, g)), ldo(G, boff(2, P, g))};
    double fo[5];
    face_flux(L, n, f0, ql0, qr0, k3, gamma, fo);
#pragma unroll
    for (int m = 0; m < 5; ++m) sto(F, boff((unsigned)m, P, g), fo[m]);
  }
  if (two) {
    const int f1 = f0 + 1;
    boundary_states(L, n, ist, ied, f1, ql1, qr1);
    const unsigned g = (unsigned)base + (unsigned)(f1 + 2) * s;
    const double k3[3] = {ldo(G, boff(0, P, g)), ldo(G, boff(1, P, g)), ldo(G, boff(2, P, g))};
    double fo[5];
    face_flux(L, n, f1, ql1, qr1, k3, gamma, fo);
#pragma unroll
    for (int m = 0; m < 5; ++m) sto(F, boff((unsigned)m, P, g), fo[m]);
  }
}

// Boundary recalculation (InviscidFluxComputer.cpp:377-521) at a line end
// whose boundary signal is set.  Each end has three stages; each sets the
// states of a few faces from fixed boundary stencils (the third also takes
// the WCNS states of the end faces), recomputes their fluxes and combines
// them into one closure value.  The stages only read line values and face
// states, never each other's fluxes, so the six (end, stage) tasks of a
// line run on six threads.  The right end reads face states as the left
// end's stages left them where the two ends' faces overlap (lines of at
// most 9 cells), as the reference, which treats the left end first, does.
__global__ void __launch_bounds__(256) inv_closure_kernel(const BlockDev* __restrict__ blocks,
                                                         Span SP, int d,
                                                         double gamma) {
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;
  if (d < 0) {  // the three directions in one launch: virtual block d * nb + b
    const int nbr = SP.nb / 3;
    d = bi / nbr;
    bi -= d * nbr;
  }
  const BlockDev& b = blocks[bi];
  // task slowest: a wavefront's lanes are lines of one task (no divergence)
  unsigned ln, task;
  udivmod((unsigned)gi, (unsigned)nlines(b, d), &task, &ln);
  const long long base = decode_line(b, d, ln);
  const int n = dcount(b, d);
  const long long s = dstride(b, d);
  const LineRef L{b.prim, (unsigned)b.P, (unsigned)base, (unsigned)s};
  const int ist = (L.at(0, 0) < 1.0e-20) ? 2 : 0;
  const int ied = (L.at(0, n + 5) < 1.0e-20) ? n - 2 : n;
  const bool right = task >= 3;
  if (right ? !(ied < n) : !(ist > 1)) return;
  const int stage = (int)task % 3;
  auto flux_at = [&](int f, const double* ql, const double* qr, double* fo) {
    const double* G = b.fgeo + (long long)(d * FG_N + 10) * b.P + base + (long long)(f + 2) * s;
    const double k3[3] = {G[0], G[b.P], G[2 * b.P]};
    face_flux(L, n, f, ql, qr, k3, gamma, fo);
  };
  // the four line values next to this end, outermost first
  double A[5], B[5], C[5], D[5];
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    A[m] = L.at(m, right ? n + 2 : 3);
    B[m] = L.at(m, right ? n + 1 : 4);
    C[m] = L.at(m, right ? n : 5);
    D[m] = L.at(m, right ? n - 1 : 6);
  }
  // face k of the stage (k = 0 the outermost face it touches)
  auto face = [&](int k, int first) { return right ? n - first - k : first + k; };
  double o[5];
  if (stage == 0) {  // faces 0..3
    double q[4][5], F[4][5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      q[0][m] = (3. * A[m] - 1. * B[m]) / 2.0;
      q[1][m] = (19. * A[m] + 25. * B[m] - 2. * C[m]) / 42.0;
      q[2][m] = (4. * A[m] + 3. * B[m] + 9. * C[m] + 2. * D[m]) / 18.0;
      q[3][m] = (-2. * A[m] + 3. * B[m] + 3. * C[m] + 2. * D[m]) / 6.0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) flux_at(face(k, 0), q[k], q[k], F[k]);
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      if (right) o[m] = -(-23. * F[0][m] + 21. * F[1][m] + 3. * F[2][m] - F[3][m]) / 24.0;
      else o[m] = (-23. * F[0][m] + 21. * F[1][m] + 3. * F[2][m] - F[3][m]) / 24.0;
    }
  } else if (stage == 1) {  // faces 1..5
    double q[5][5], q2[5], F[5][5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      q[0][m] = (3. * A[m] + 6. * B[m] - C[m]) / 8.0;
      q[1][m] = (-30. * A[m] + 145. * B[m] + 29. * C[m] - 8. * D[m]) / 136.0;  // the end-side state of face 2
      q2[m] = (2. * A[m] + 49. * B[m] + 125. * C[m] - 40. * D[m]) / 136.0;     // its other state
      q[2][m] = (-B[m] + 6. * C[m] + 3. * D[m]) / 8.0;
      q[3][m] = (B[m] + 2. * C[m] + 5. * D[m]) / 8.0;
      q[4][m] = (B[m] + C[m] + 6. * D[m]) / 8.0;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      if (k == 1) {
        if (right) flux_at(face(k, 1), q2, q[1], F[k]);
        else flux_at(face(k, 1), q[1], q2, F[k]);
      } else {
        flux_at(face(k, 1), q[k], q[k], F[k]);
      }
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      if (right) o[m] = -(-22. * F[0][m] + 17. * F[1][m] + 9. * F[2][m] - 5. * F[3][m] + F[4][m]) / 24.0;
      else o[m] = (-22. * F[0][m] + 17. * F[1][m] + 9. * F[2][m] - 5. * F[3][m] + F[4][m]) / 24.0;
    }
  } else {  // faces 1..4: end-side states from the stencils, the others as before the end's first stage
    double qe[4][5], qo[4][5], F[4][5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      qe[0][m] = (3. * A[m] + 6. * B[m] - C[m]) / 8.0;
      qe[1][m] = (-29. * A[m] + 170. * B[m] + 63. * C[m] + 12. * D[m]) / 216.0;
      qe[2][m] = (-B[m] + 6. * C[m] + 3. * D[m]) / 8.0;
      qe[3][m] = (B[m] + 2. * C[m] + 5. * D[m]) / 8.0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int f = face(k, 1);
      double ql[5], qr[5];
      face_states<false>(L, n, ist, ied, f, ql, qr);
#pragma unroll
      for (int m = 0; m < 5; ++m) qo[k][m] = right ? ql[m] : qr[m];
      if (right && ist > 1 && f <= 5) {  // a face the left end's stages set last
#pragma unroll
        for (int m = 0; m < 5; ++m) {
          const double a = L.at(m, 3), bb = L.at(m, 4), cc = L.at(m, 5), dd = L.at(m, 6);
          double v;
          if (f == 0) v = (3. * a - 1. * bb) / 2.0;
          else if (f == 1) v = (3. * a + 6. * bb - cc) / 8.0;
          else if (f == 2) v = (-29. * a + 170. * bb + 63. * cc + 12. * dd) / 216.0;
          else if (f == 3) v = (-bb + 6. * cc + 3. * dd) / 8.0;
          else if (f == 4) v = (bb + 2. * cc + 5. * dd) / 8.0;
          else v = (bb + cc + 6. * dd) / 8.0;
          qo[k][m] = v;
        }
      }
      if (right) flux_at(f, qo[k], qe[k], F[k]);
      else flux_at(f, qe[k], qo[k], F[k]);
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      if (right) o[m] = -(F[0][m] - 27. * F[1][m] + 27. * F[2][m] - F[3][m]) / 24.0;
      else o[m] = (F[0][m] - 27. * F[1][m] + 27. * F[2][m] - F[3][m]) / 24.0;
    }
  }
  double* out = b.fcov + b.fcov_off[d] + (long long)ln * 30 + (right ? 5 - stage : stage) * 5;
#pragma unroll
  for (int m = 0; m < 5; ++m) out[m] = o[m];
}


// 6th-order difference with 3rd/4th-order ends
// (compute_flux_diff_high_order :197-233 == compute_flux_diff_1d :257-296)
// for interior node t (0-based along the line); F(f) is face f.


// ---------------------------------------------------------------------------
// Viscous RHS
// ---------------------------------------------------------------------------
__device__ inline double uvwt_at(const BlockDev& b, long long id, int v) {
  // (u,v,w,T) = prim 1,2,3,5
  return b.prim[(long long)(v < 3 ? v + 1 : 5) * b.P + id];
}

// compute_node_deriv_line (FluxComputer.cpp:13-67) at 1-based position i1.
__device__ inline double node_deriv(const double* q, long long s, int n, int i1, bool fl, bool fr) {
  // q points at line position 0; Q(p) = value at position p
  auto Q = [&](int p) { return q[(long long)p * s]; };
  if (fl && i1 == 1) return div_c<6>((-11.0 * Q(3) + 18.0 * Q(4) - 9.0 * Q(5) + 2.0 * Q(6)));
  if (fl && i1 == 2) return div_c<12>((-3.0 * Q(3) + -10.0 * Q(4) + 18.0 * Q(5) + -6.0 * Q(6) + 1.0 * Q(7)));
  if (fr && i1 == n) return div_c<6>(-(-11.0 * Q(n + 2) + 18.0 * Q(n + 1) - 9.0 * Q(n) + 2.0 * Q(n - 1)));
  if (fr && i1 == n - 1)
    return div_c<12>(-(-3.0 * Q(n + 2) + -10.0 * Q(n + 1) + 18.0 * Q(n) + -6.0 * Q(n - 1) + 1.0 * Q(n - 2)));
  return div_c<12>((8.0 * (Q(i1 + 3) - Q(i1 + 1)) + -1.0 * (Q(i1 + 4) - Q(i1))));
}

__global__ void __launch_bounds__(256) visc_deriv_kernel(const BlockDev* __restrict__ blocks,
                                                         Span SP, int* queue_head) {
  // the face-flux work queue, which runs next, starts from its first item
  if (queue_head && blockIdx.x == 0 && threadIdx.x == 0) *queue_head = 0;
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;
  const BlockDev& b = blocks[bi];
  int i, j, k;
  decode_interior(b, gi, &i, &j, &k);
  const long long id = pidx(b, i, j, k);
  const int pos[3] = {i, j, k};
  for (int d = 0; d < 3; ++d) {
    const int n = dcount(b, d);
    const long long s = dstride(b, d);
    const long long lbase = id - (long long)pos[d] * s;
    const bool fl = b.prim[lbase] < 0.0;
    const bool fr = b.prim[lbase + (long long)(n + 5) * s] < 0.0;
    const int i1 = pos[d] - NG + 1;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double* q = b.prim + (long long)(v < 3 ? v + 1 : 5) * b.P + lbase;
      b.duvwt[(long long)(4 * d + v) * b.P + id] = node_deriv(q, s, n, i1, fl, fr);
    }
  }
}

// interp_line_to_half central form at face f (positions f+1..f+4)
__device__ inline double half_central(const double* q, long long s, int f) {
  return (9.0 * (q[(long long)(f + 2) * s] + q[(long long)(f + 3) * s]) +
          -1.0 * (q[(long long)(f + 4) * s] + q[(long long)(f + 1) * s])) /
         16.0;
}
// interp_line_to_half with flags (FluxComputer.cpp:129-178); the central
// stencil is loaded first so its loads do not wait for the flags
__device__ inline double half_flagged(const double* q, long long s, int n, int f, bool fl, bool fr) {
  auto Q = [&](int p) { return q[(long long)p * s]; };
  const double c1 = Q(f + 1), c2 = Q(f + 2), c3 = Q(f + 3), c4 = Q(f + 4);
  if (fl && f == 0) return (35.0 * Q(3) + -35.0 * Q(4) + 21.0 * Q(5) + -5.0 * Q(6)) / 16.0;
  if (fl && f == 1) return (5.0 * Q(3) + 15.0 * Q(4) + -5.0 * Q(5) + 1.0 * Q(6)) / 16.0;
  if (fr && f == n) return (35.0 * Q(n + 2) + -35.0 * Q(n + 1) + 21.0 * Q(n) + -5.0 * Q(n - 1)) / 16.0;
  if (fr && f == n - 1) return (5.0 * Q(n + 2) + 15.0 * Q(n + 1) + -5.0 * Q(n) + 1.0 * Q(n - 1)) / 16.0;
  return (9.0 * (c2 + c3) + -1.0 * (c4 + c1)) / 16.0;
}
// compute_half_deriv_line (FluxComputer.cpp:72-124)
__device__ inline double half_deriv(const double* q, long long s, int n, int f, bool fl, bool fr) {
  auto Q = [&](int p) { return q[(long long)p * s]; };
  const double c1 = Q(f + 1), c2 = Q(f + 2), c3 = Q(f + 3), c4 = Q(f + 4);
  if (fl && f == 0) return div_c<24>((-71.0 * Q(3) + 141.0 * Q(4) + -93.0 * Q(5) + 23.0 * Q(6)));
  if (fl && f == 1) return div_c<24>((-22.0 * Q(3) + 17.0 * Q(4) + 9.0 * Q(5) + -5.0 * Q(6) + 1.0 * Q(7)));
  if (fr && f == n) return div_c<24>(-(-71.0 * Q(n + 2) + 141.0 * Q(n + 1) + -93.0 * Q(n) + 23.0 * Q(n - 1)));
  if (fr && f == n - 1)
    return div_c<24>(-(-22.0 * Q(n + 2) + 17.0 * Q(n + 1) + 9.0 * Q(n) + -5.0 * Q(n - 1) + 1.0 * Q(n - 2)));
  return div_c<24>((27.0 * (c3 - c2) + -1.0 * (c4 - c1)));
}

// Face geometry (once, after the metrics): the central interpolation of
// the metrics and the volume used by the viscous flux (:183-205 via
// interp_line_to_half) and the inviscid metrics with their end formulas.
__global__ void __launch_bounds__(256) face_geom_kernel(const BlockDev* __restrict__ blocks, Span SP, int d) {
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;
  const BlockDev& b = blocks[bi];
  long long base;
  int f;
  decode_face(b, d, gi, &base, &f);
  const int n = dcount(b, d);
  const long long s = dstride(b, d);
  const long long P = b.P;
  double* G = b.fgeo + (long long)(d * FG_N) * P + base + (long long)(f + 2) * s;
#pragma unroll
  for (int m = 0; m < 9; ++m) G[(long long)m * P] = half_central(b.met + (long long)m * P + base, s, f);
  G[9 * P] = half_central(b.vol + base, s, f);
  double k3[3];
  metric_face(b, d, base, s, n, f, k3);
#pragma unroll
  for (int c = 0; c < 3; ++c) G[(long long)(10 + c) * P] = k3[c];
}

// interp_line_to_half / compute_half_deriv_line at face f with the central
// stencil values c1..c4 (positions f+1..f+4) already loaded; Q(p) loads the
// line's value at position p (the boundary forms only)
template <class QF>
__device__ __forceinline__ double half_flagged_c(const QF& Q, int n, int f, bool fl, bool fr, double c1, double c2,
                                                 double c3, double c4) {
  if (fl && f == 0) return (35.0 * Q(3) + -35.0 * Q(4) + 21.0 * Q(5) + -5.0 * Q(6)) / 16.0;
  if (fl && f == 1) return (5.0 * Q(3) + 15.0 * Q(4) + -5.0 * Q(5) + 1.0 * Q(6)) / 16.0;
  if (fr && f == n) return (35.0 * Q(n + 2) + -35.0 * Q(n + 1) + 21.0 * Q(n) + -5.0 * Q(n - 1)) / 16.0;
  if (fr && f == n - 1) return (5.0 * Q(n + 2) + 15.0 * Q(n + 1) + -5.0 * Q(n) + 1.0 * Q(n - 1)) / 16.0;
  return (9.0 * (c2 + c3) + -1.0 * (c4 + c1)) / 16.0;
}
template <class QF>
__device__ __forceinline__ double half_deriv_c(const QF& Q, int n, int f, bool fl, bool fr, double c1, double c2,
                                               double c3, double c4) {
  if (fl && f == 0) return div_c<24>((-71.0 * Q(3) + 141.0 * Q(4) + -93.0 * Q(5) + 23.0 * Q(6)));
  if (fl && f == 1) return div_c<24>((-22.0 * Q(3) + 17.0 * Q(4) + 9.0 * Q(5) + -5.0 * Q(6) + 1.0 * Q(7)));
  if (fr && f == n) return div_c<24>(-(-71.0 * Q(n + 2) + 141.0 * Q(n + 1) + -93.0 * Q(n) + 23.0 * Q(n - 1)));
  if (fr && f == n - 1)
    return div_c<24>(-(-22.0 * Q(n + 2) + 17.0 * Q(n + 1) + 9.0 * Q(n) + -5.0 * Q(n - 1) + 1.0 * Q(n - 2)));
  return div_c<24>((27.0 * (c3 - c2) + -1.0 * (c4 - c1)));
}

// prim field of viscous variable v (u, v, w, T = prim 1, 2, 3, 5)
__device__ __forceinline__ constexpr unsigned uvwt_field(int v) { return v < 3 ? (unsigned)v + 1u : 5u; }

// One viscous face flux (compute_viscous_flux_1d, FluxComputer.cpp:183-252)
// at face f of a line along D, from the windows at positions f+1..f+4:
// Wv(v, j) of u, v, w, T, Mv(j) of mu and Xv(c, j) of the cross
// derivatives (c < 4: along o0, c >= 4: along o1, at positions clamped to
// the interior), j = 0..3; the face geometry is the static one
// interpolated after the metrics.  Writes the five face fluxes of face f.
// base is the padded index of the line's position 0 and s the stride
// along D (32-bit: field offsets are formed with boff).  The direction is a
// template parameter: with it the derivative and metric arrays are indexed
// by constants and stay in registers (a run-time index puts them in
// scratch memory).
template <int D, class WF, class MF, class XF>
__device__ __forceinline__ void visc_face_val(const BlockDev& b, unsigned base, unsigned s, int n, int f,
                                              bool fl, bool fr, const WF& Wv, const MF& Mv, const XF& Xv,
                                              double* out) {
  constexpr int d = D;
  const unsigned P = (unsigned)b.P;
  constexpr int o0 = (d == 0) ? 1 : 0, o1 = (d == 2) ? 1 : 2;
  const double* prim = b.prim;
  double uvwt[4], dmid[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    auto Q = [&](int p) { return ldo(prim, boff(uvwt_field(v), P, base + (unsigned)p * s)); };
    uvwt[v] = half_flagged_c(Q, n, f, fl, fr, Wv(v, 0), Wv(v, 1), Wv(v, 2), Wv(v, 3));
    dmid[v] = half_deriv_c(Q, n, f, fl, fr, Wv(v, 0), Wv(v, 1), Wv(v, 2), Wv(v, 3));
  }
  const double vs1 = (9.0 * (Mv(1) + Mv(2)) + -1.0 * (Mv(3) + Mv(0))) / 16.0;
  const double vs2 = vs1 * b.cp_prl;
  const double* G = b.fgeo + (long long)(d * FG_N) * b.P;  // this direction's face geometry
  const unsigned g = base + (unsigned)(f + 2) * s;
  const double volh = ldo(G, boff(9, P, g));
  double kx[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) kx[m] = ldo(G, boff(m, P, g));
  double duv[12];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    duv[4 * d + v] = dmid[v];
    duv[4 * o0 + v] = (9.0 * (Xv(v, 1) + Xv(v, 2)) + -1.0 * (Xv(v, 3) + Xv(v, 0))) / 16.0;
    duv[4 * o1 + v] = (9.0 * (Xv(4 + v, 1) + Xv(4 + v, 2)) + -1.0 * (Xv(4 + v, 3) + Xv(4 + v, 0))) / 16.0;
  }
  // transform_derivs_to_phys (:183-205)
  const double vinv = 1.0 / volh;
  double gx[12];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const double k1 = kx[m], k2 = kx[m + 3], k3 = kx[m + 6];
#pragma unroll
    for (int v = 0; v < 4; ++v) gx[v * 3 + m] = (k1 * duv[v] + k2 * duv[v + 4] + k3 * duv[v + 8]) * vinv;
  }
  // compute_viscous_flux_1d (:210-252)
  const double vscc = vs1 * (2.0 / 3.0);
  const double dudx = gx[0], dudy = gx[1], dudz = gx[2];
  const double dvdx = gx[3], dvdy = gx[4], dvdz = gx[5];
  const double dwdx = gx[6], dwdy = gx[7], dwdz = gx[8];
  const double txx = vscc * (2.0 * dudx - dvdy - dwdz);
  const double tyy = vscc * (2.0 * dvdy - dwdz - dudx);
  const double tzz = vscc * (2.0 * dwdz - dudx - dvdy);
  const double txy = vs1 * (dudy + dvdx);
  const double txz = vs1 * (dudz + dwdx);
  const double tyz = vs1 * (dvdz + dwdy);
  const double nx = kx[3 * d], ny = kx[3 * d + 1], nz = kx[3 * d + 2];
  const double f1 = txx * nx + txy * ny + txz * nz;
  const double f2 = txy * nx + tyy * ny + tyz * nz;
  const double f3 = txz * nx + tyz * ny + tzz * nz;
  const double work = uvwt[0] * f1 + uvwt[1] * f2 + uvwt[2] * f3;
  const double heat = vs2 * (gx[9] * nx + gx[10] * ny + gx[11] * nz);
  // components 1..4; the mass component is zero, and the viscous
  // difference skips it (dq0 - Re^-1 0 == dq0)
  out[0] = f1;
  out[1] = f2;
  out[2] = f3;
  out[3] = work + heat;
}

// visc_face_val stored at the face's slot of this direction's flux arrays
// (components 1..4; the zero mass component is never stored)
template <int D, class WF, class MF, class XF>
__device__ __forceinline__ void visc_face_at(const BlockDev& b, unsigned base, unsigned s, int n, int f,
                                             bool fl, bool fr, const WF& Wv, const MF& Mv, const XF& Xv) {
  double o[4];
  visc_face_val<D>(b, base, s, n, f, fl, fr, Wv, Mv, Xv, o);
  const unsigned P = (unsigned)b.P;
  const unsigned g = base + (unsigned)(f + 2) * s;
  double* F = b.fx + (long long)D * 5 * b.P;
#pragma unroll
  for (int m = 0; m < 4; ++m) sto(F, boff(1 + m, P, g), o[m]);
}

__device__ __forceinline__ int clamp_interior(int p, int n) { return p < NG ? NG : (p > n + NG - 1 ? n + NG - 1 : p); }

// the stencil values of line position p: u, v, w, T and mu at p, the cross
// derivatives at p clamped to the interior
template <int D>
__device__ __forceinline__ void visc_window_load(const BlockDev& b, unsigned base, unsigned s, int n, int p,
                                                 double* W4, double& M, double* X8) {
  constexpr int o0 = (D == 0) ? 1 : 0, o1 = (D == 2) ? 1 : 2;
  const unsigned P = (unsigned)b.P;
  const unsigned e = base + (unsigned)p * s;
#pragma unroll
  for (int v = 0; v < 4; ++v) W4[v] = ldo(b.prim, boff(uvwt_field(v), P, e));
  M = ldo(b.mu, e << 3);
  const unsigned ec = base + (unsigned)clamp_interior(p, n) * s;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    X8[v] = ldo(b.duvwt, boff(4 * o0 + v, P, ec));
    X8[4 + v] = ldo(b.duvwt, boff(4 * o1 + v, P, ec));
  }
}

// Faces f0 = 2g and f0 + 1 of one line along D: the two faces share their
// stencil windows (positions f0+1..f0+5).
template <int D>
__device__ __forceinline__ void visc_face_pair(const BlockDev& b, long long gi) {
  constexpr int d = D;
  long long base64;
  int f0;
  decode_face_pair(b, d, gi, &base64, &f0);
  const unsigned base = (unsigned)base64;
  const int n = dcount(b, d);
  const unsigned s = (unsigned)dstride(b, d);
  const bool fl = ldo(b.prim, base << 3) < 0.0;
  const bool fr = ldo(b.prim, (base + (unsigned)(n + 5) * s) << 3) < 0.0;
  double W[5][4], M[5], X[5][8];
#pragma unroll
  for (int j = 0; j < 5; ++j) visc_window_load<D>(b, base, s, n, f0 + 1 + j, W[j], M[j], X[j]);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int f = f0 + k;
    if (f > n) break;
    visc_face_at<D>(b, base, s, n, f, fl, fr, [&](int v, int j) { return W[k + j][v]; },
                    [&](int j) { return M[k + j]; }, [&](int c, int j) { return X[k + j][c]; });
  }
}

// Viscous face fluxes over a span of face pairs (one launch per direction,
// or, d < 0, the three directions in one launch: virtual block d * nb + b).
__global__ void __launch_bounds__(256) visc_face_kernel(const BlockDev* __restrict__ blocks,
                                                        Span SP, int d) {
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;
  if (d < 0) {
    const int nbr = SP.nb / 3;
    d = bi / nbr;
    bi -= d * nbr;
  }
  if (d == 0)
    visc_face_pair<0>(blocks[bi], gi);
  else if (d == 1)
    visc_face_pair<1>(blocks[bi], gi);
  else
    visc_face_pair<2>(blocks[bi], gi);
}

// Faces f0 .. f1-1 of one line along D (j or k) in order: the windows of
// the thirteen stencil arrays slide one position per face, so every value
// is loaded once, and the 64 lanes of a wavefront hold neighbouring lines
// (loads of consecutive i).
template <int D>
__device__ __forceinline__ void visc_face_march(const BlockDev& b, long long line, int f0, int f1) {
  constexpr int d = D;
  unsigned q, r;
  udivmod((unsigned)line, (unsigned)b.idim, &q, &r);
  const unsigned base =
      (unsigned)((d == 1) ? pidx(b, (int)r + NG, 0, (int)q + NG) : pidx(b, (int)r + NG, (int)q + NG, 0));
  const int n = dcount(b, d);
  const unsigned s = (unsigned)dstride(b, d);
  const bool fl = ldo(b.prim, base << 3) < 0.0;
  const bool fr = ldo(b.prim, (base + (unsigned)(n + 5) * s) << 3) < 0.0;
  double W[4][4], M[4], X[4][8];
#pragma unroll
  for (int j = 0; j < 4; ++j) visc_window_load<D>(b, base, s, n, f0 + 1 + j, W[j], M[j], X[j]);
  for (int f = f0; f < f1; ++f) {
    visc_face_at<D>(b, base, s, n, f, fl, fr, [&](int v, int j) { return W[j][v]; }, [&](int j) { return M[j]; },
                    [&](int c, int j) { return X[j][c]; });
    if (f + 1 < f1) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
#pragma unroll
        for (int v = 0; v < 4; ++v) W[j][v] = W[j + 1][v];
        M[j] = M[j + 1];
#pragma unroll
        for (int c = 0; c < 8; ++c) X[j][c] = X[j + 1][c];
      }
      visc_window_load<D>(b, base, s, n, f + 5, W[3], M[3], X[3]);
    }
  }
}

// The viscous face fluxes of every block and direction as one persistent
// launch fed from a block work queue: each workgroup takes the next item
// (an atomic ticket on *head) until the queue is empty.  An item is
//   d = 0: 256 consecutive face pairs of block b's lines along i (one pair
//          per lane, as visc_face_kernel);
//   d = 1, 2: 256 consecutive lines along j or k (lane = line, i fastest)
//          over faces f0 .. f1-1, each lane marching along its line.
// *head is zeroed by the derivative kernel that runs before this one.
// Held to two waves per SIMD (256 VGPRs): beside it, on the side stream,
// the inviscid face kernel then gets more waves (measured faster per step
// than the unbounded allocation, which leaves one wave of this kernel and
// one of the other per SIMD).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) visc_face_queue_kernel(const BlockDev* __restrict__ blocks,
                                                              const ViscItem* __restrict__ items, int nitems,
                                                              int* head) {
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(head, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);  // the same for the whole workgroup
    __syncthreads();
    if (it >= nitems) return;
    const ViscItem Q = items[it];
    const BlockDev& b = blocks[Q.block];
    const long long g = Q.a0 + threadIdx.x;
    if (g < Q.a1) {
      if (Q.d == 0)
        visc_face_pair<0>(b, g);
      else if (Q.d == 1)
        visc_face_march<1>(b, g, Q.f0, Q.f1);
      else
        visc_face_march<2>(b, g, Q.f0, Q.f1);
    }
  }
}


template <class FF>
__device__ __forceinline__ double diff_node_f(const FF& F, int n, int t) {
  if (t >= 2 && t < n - 2)
    return div_c<1920>((2250.0 * (F(t + 1) - F(t)) + -125.0 * (F(t + 2) - F(t - 1)) + 9.0 * (F(t + 3) - F(t - 2))));
  if (t == 1) return div_c<24>((F(0) - 27.0 * F(1) + 27.0 * F(2) - F(3)));
  if (t == n - 2) return div_c<24>(-(F(n) - 27.0 * F(n - 1) + 27.0 * F(n - 2) - F(n - 3)));
  if (t == 0) return div_c<24>((-23.0 * F(0) + 21.0 * F(1) + 3.0 * F(2) - F(3)));
  return div_c<24>(-(-23.0 * F(n) + 21.0 * F(n - 1) + 3.0 * F(n - 2) - F(n - 3)));
}

// The difference of diff_node_f at node t over a window W(w) of the six
// faces t-2 .. t+3 (w = 0..5), with the same branches in the same order and
// the same expressions; each branch reads its faces at constant window
// slots (face f at slot f - t + 2), so a window held in registers stays
// there.  Slots of faces outside 0..n are never read.
template <class WF>
__device__ __forceinline__ double diff_node_w(const WF& W, int n, int t) {
  if (t >= 2 && t < n - 2)
    return div_c<1920>((2250.0 * (W(3) - W(2)) + -125.0 * (W(4) - W(1)) + 9.0 * (W(5) - W(0))));
  if (t == 1) return div_c<24>((W(1) - 27.0 * W(2) + 27.0 * W(3) - W(4)));
  if (t == n - 2) return div_c<24>(-(W(4) - 27.0 * W(3) + 27.0 * W(2) - W(1)));
  if (t == 0) return div_c<24>((-23.0 * W(2) + 21.0 * W(3) + 3.0 * W(4) - W(5)));
  return div_c<24>(-(-23.0 * W(3) + 21.0 * W(2) + 3.0 * W(1) - W(0)));
}

// first and last face that diff_node_f reads for node t of a line of n >= 4
// nodes; both are non-decreasing in t, so the faces of nodes ta..tb lie in
// [face_lo(ta), face_hi(tb)], at most tb - ta + 6 of them
__device__ __forceinline__ int face_lo(int t, int n) { return t < 2 ? 0 : (t >= n - 2 ? n - 3 : t - 2); }
__device__ __forceinline__ int face_hi(int t, int n) { return t < 2 ? 3 : (t >= n - 2 ? n : t + 3); }

// One viscous face flux (visc_face_val) of the line along D whose position
// 0 is at padded index base, every stencil value loaded from the arrays.
template <int D>
__device__ __forceinline__ void visc_face_direct(const BlockDev& b, unsigned base, unsigned s, int n, int f,
                                                 double* o) {
  constexpr int o0 = (D == 0) ? 1 : 0, o1 = (D == 2) ? 1 : 2;
  const unsigned P = (unsigned)b.P;
  const bool fl = ldo(b.prim, base << 3) < 0.0;
  const bool fr = ldo(b.prim, (base + (unsigned)(n + 5) * s) << 3) < 0.0;
  visc_face_val<D>(
      b, base, s, n, f, fl, fr,
      [&](int v, int j) { return ldo(b.prim, boff(uvwt_field(v), P, base + (unsigned)(f + 1 + j) * s)); },
      [&](int j) { return ldo(b.mu, (base + (unsigned)(f + 1 + j) * s) << 3); },
      [&](int c, int j) {
        const unsigned e = base + (unsigned)clamp_interior(f + 1 + j, n) * s;
        return ldo(b.duvwt, boff(c < 4 ? 4 * o0 + c : 4 * o1 + c - 4, P, e));
      },
      o);
}

// The viscous right-hand side after the node derivatives, as one
// persistent launch fed from a block work queue (the derivative kernel
// zeroes *head).  An item is a tile of VT x VT (i, j) interior nodes of one
// block over the planes [k0, k1); its workgroup marches along k, and at
// each plane
//  * forms the i faces and j faces that the tile's nodes difference (at most
//    VT + 5 per row or column) into LDS, one face per thread in turn, with
//    loads along i (i faces) or across columns (j faces) coalesced;
//  * forms, per column, the k face three planes ahead, into a window of the
//    six k faces the node differences, held in registers (the item starts
//    by forming the five faces before its first plane's lead face);
//  * differences i, j and k in that order into dq: dq -= Re^-1 dF per
//    direction (compute_flux_diff_1d / compute_flux_diff_high_order).
// Each face flux is the value the face kernels store for it and each
// difference the expression diff_kernel forms over the same values, so dq
// is the same bit for bit; no face flux goes through HBM.  Face arrays in
// LDS alternate by plane, so one barrier per plane separates a plane's
// faces from its difference and the difference from the faces of the plane
// after next.  FIRST: the step's first right-hand-side kernel (dq starts
// from zero; component 0 is written as +0).  Lines of fewer than 4 nodes
// are not handled (the host keeps such contexts on the face-flux queue).
constexpr int VT = 16;
template <bool FIRST>
__global__ void __launch_bounds__(256) visc_fused_kernel(const BlockDev* __restrict__ blocks,
                                                         const ViscTile* __restrict__ items, int nitems,
                                                         int* head) {
  __shared__ double Fi[2][4][VT][VT + 5];  // [plane parity][component][row][face - first i face]
  __shared__ double Fj[2][4][VT + 5][VT];  // [plane parity][component][face - first j face][column]
  __shared__ int s_item;
  const int lane = threadIdx.x, li = lane % VT, lj = lane / VT;
  for (;;) {
    if (lane == 0) s_item = atomicAdd(head, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);  // the same for the whole workgroup
    __syncthreads();
    if (it >= nitems) return;
    const ViscTile Q = items[it];
    const BlockDev& b = blocks[Q.block];
    const unsigned P = (unsigned)b.P;
    const int ni = b.idim, nj = b.jdim, nk = b.kdim;
    const int ti1 = min(Q.i0 + VT, ni), tj1 = min(Q.j0 + VT, nj);
    const int ncol = ti1 - Q.i0, nrow = tj1 - Q.j0;
    const int fi0 = face_lo(Q.i0, ni), nfi = face_hi(ti1 - 1, ni) - fi0 + 1;
    const int fj0 = face_lo(Q.j0, nj), nfj = face_hi(tj1 - 1, nj) - fj0 + 1;
    const int nface_i = nfi * nrow, nface = nface_i + nfj * ncol;
    const bool mine = li < ncol && lj < nrow;
    const int ti = Q.i0 + li, tj = Q.j0 + lj;  // this lane's node (0-based interior)
    const unsigned sk = (unsigned)b.nx * (unsigned)b.ny;
    const unsigned kbase = mine ? (unsigned)pidx(b, ti + NG, tj + NG, 0) : 0u;  // its column's k line
    const double re_inv = b.re_inv;
    double KW[6][4];
#pragma unroll
    for (int w = 0; w < 6; ++w)
#pragma unroll
      for (int m = 0; m < 4; ++m) KW[w][m] = 0.0;
#pragma unroll
    for (int w = 0; w < 5; ++w) {
      const int f = Q.k0 - 2 + w;
      if (mine && f >= 0 && f <= nk) visc_face_direct<2>(b, kbase, sk, nk, f, KW[w]);
    }
    int par = 0;
    for (int tk = Q.k0; tk < Q.k1; ++tk) {
      if (mine && tk + 3 <= nk) visc_face_direct<2>(b, kbase, sk, nk, tk + 3, KW[5]);
      const int k = tk + NG;
      for (int e = lane; e < nface; e += 256) {
        double o[4];
        if (e < nface_i) {
          unsigned row, sl;
          udivmod((unsigned)e, (unsigned)nfi, &row, &sl);
          visc_face_direct<0>(b, (unsigned)pidx(b, 0, Q.j0 + (int)row + NG, k), 1u, ni, fi0 + (int)sl, o);
#pragma unroll
          for (int m = 0; m < 4; ++m) Fi[par][m][row][sl] = o[m];
        } else {
          unsigned sl, col;
          udivmod((unsigned)(e - nface_i), (unsigned)ncol, &sl, &col);
          visc_face_direct<1>(b, (unsigned)pidx(b, Q.i0 + (int)col + NG, 0, k), (unsigned)b.nx, nj, fj0 + (int)sl,
                              o);
#pragma unroll
          for (int m = 0; m < 4; ++m) Fj[par][m][sl][col] = o[m];
        }
      }
      __syncthreads();
      if (mine) {
        const unsigned id = kbase + (unsigned)k * sk;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          double acc = FIRST ? 0.0 : ldo(b.dq, boff(1 + m, P, id));
          acc -= re_inv * diff_node_f([&](int f) { return Fi[par][m][lj][f - fi0]; }, ni, ti);
          acc -= re_inv * diff_node_f([&](int f) { return Fj[par][m][f - fj0][li]; }, nj, tj);
          acc -= re_inv * diff_node_w([&](int w) { return KW[w][m]; }, nk, tk);
          sto(b.dq, boff(1 + m, P, id), acc);
        }
        if (FIRST) sto(b.dq, boff(0, P, id), 0.0);
      }
#pragma unroll
      for (int w = 0; w < 5; ++w)
#pragma unroll
        for (int m = 0; m < 4; ++m) KW[w][m] = KW[w + 1][m];
      par ^= 1;
    }
  }
}

// Flux differences of the inviscid (INV, closure overrides included) or the
// viscous face fluxes into dq, per direction i, j, k in that order
// (compute_flux_diff_1d / compute_flux_diff_high_order; viscous: dq -=
// diff / Re, inviscid: dq += diff).  A workgroup takes DIFF_TI x DIFF_TJ
// (i, j) nodes of DIFF_TK consecutive k planes.  Each lane first forms the
// k differences of its column's DIFF_TK nodes from one register window of
// the k faces they read (faces t0-2 .. t0+DIFF_TK+2 of the line, loaded
// once for the DIFF_TK nodes); then, plane by plane, the i and j stencils
// of the tile (5 face slots more than nodes per row or column) are staged
// through LDS once.  Node t of a line reads faces t-2..t+3, which sit at the
// padded slots t..t+5.  FIRST: the step's first right-hand-side kernel,
// which starts from dq = 0 (what update_flow_states leaves) instead of
// reading it.
// The viscous mass flux is identically zero (compute_viscous_flux_1d sets
// it to 0): its difference is +-0, and dq0 - Re^-1 (+-0) is dq0 bit for bit
// for every dq0 the step produces (dq starts from +0), so the viscous
// difference leaves component 0 as it is and the face kernel never stores
// it.
template <bool INV, bool FIRST>
__global__ void __launch_bounds__(256) diff_kernel(const BlockDev* __restrict__ blocks, Span SP) {
  constexpr int TI = DIFF_TI, TJ = DIFF_TJ, TK = DIFF_TK;  // the tile's nodes along i, j and k
  constexpr int M0 = INV ? 0 : 1, NM = 5 - M0;               // flux components differenced
  __shared__ double X[NM][TJ][TI + 5];  // i slots i0-3 .. i0+TI+1 of each row of the tile
  __shared__ double Y[NM][TJ + 5][TI];  // j slots j0-3 .. j0+TJ+1 of each column
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;  // whole tiles: every lane of a workgroup has an item
  const BlockDev& b = blocks[bi];
  const double re_inv = b.re_inv;
  // 32-bit offsets: the flux arrays hold 15 fields of P < 2^25 points
  const unsigned P = (unsigned)b.P, nx = (unsigned)b.nx, nxy = (unsigned)b.nx * (unsigned)b.ny;
  const unsigned tci = (unsigned)(b.idim + TI - 1) / TI, tcj = (unsigned)(b.jdim + TJ - 1) / TJ;
  unsigned q, ti, q2, tj;
  udivmod((unsigned)(gi / 256), tci, &q, &ti);
  udivmod(q, tcj, &q2, &tj);
  const int k0 = TK * (int)q2 + NG;
  const int i0 = NG + TI * (int)ti, j0 = NG + TJ * (int)tj;
  const int lane = threadIdx.x, li = lane % TI, lj = lane / TI;
  const int i = i0 + li, j = j0 + lj;
  const bool in_ij = i < NG + b.idim && j < NG + b.jdim;
  const double* FX = INV ? b.fxi : b.fx;
  // the k differences of nodes (i, j, k0 .. k0+TK-1), closure values
  // applied.  Window faces outside 0..n are never read; their slots are
  // clamped to the padded line (slot n+5 at most: a last group of fewer
  // than TK planes would otherwise reach past it)
  const int nk = b.kdim, t0 = k0 - NG;
  const unsigned lbk = (unsigned)i + (unsigned)j * nx;  // the k line's position 0
  double fk[TK][5];
  if (in_ij) {
    const double* cov[TK];
#pragma unroll
    for (int kk = 0; kk < TK; ++kk) cov[kk] = nullptr;
    if (INV) {
      const bool lflag = ldo(b.prim, lbk << 3) < 1.0e-20;
      const bool rflag = ldo(b.prim, (lbk + (unsigned)(nk + 5) * nxy) << 3) < 1.0e-20;
#pragma unroll
      for (int kk = 0; kk < TK; ++kk) {
        const int t = t0 + kk;
        if (t >= nk) continue;  // no node (a last group of fewer than TK planes)
        if (lflag && t <= 2) cov[kk] = b.fcov + b.fcov_off[2] + line_id(b, 2, i, j, k0 + kk) * 30 + t * 5;
        else if (rflag && t >= nk - 3)
          cov[kk] = b.fcov + b.fcov_off[2] + line_id(b, 2, i, j, k0 + kk) * 30 + (3 + t - (nk - 3)) * 5;
      }
    }
#pragma unroll
    for (int m = M0; m < 5; ++m) {
      double W[TK + 5];
#pragma unroll
      for (int w = 0; w < TK + 5; ++w)
        W[w] = ldo(FX, boff((unsigned)(10 + m), P, lbk + (unsigned)min(t0 + w, nk + 5) * nxy));
#pragma unroll
      for (int kk = 0; kk < TK; ++kk) {
        fk[kk][m] = diff_node_w([&](int w) { return W[kk + w]; }, nk, t0 + kk);
        if (INV && cov[kk]) fk[kk][m] = cov[kk][m];
      }
    }
  }
  // each lane stages up to two (row, slot) positions of each direction, all
  // components of a position from one offset
  static_assert(TI * TJ == 256 && TJ * (TI + 5) <= 512 && (TJ + 5) * TI <= 512, "tile of 256 nodes");
#pragma unroll
  for (int kk = 0; kk < TK; ++kk) {
    const int k = k0 + kk;
    if (k >= NG + nk) break;  // the same for the whole workgroup
    if (kk > 0) __syncthreads();  // every lane has read the plane before
    const unsigned kplane = (unsigned)k * nxy;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int p = lane + 256 * h;
      if (p < TJ * (TI + 5)) {
        const int row = p / (TI + 5), col = p % (TI + 5);
        const int jj = j0 + row, ii = i0 - 3 + col;
        const bool ok = jj < NG + b.jdim && ii < b.nx;
        const unsigned e = (unsigned)ii + (unsigned)jj * nx + kplane;
#pragma unroll
        for (int m = 0; m < NM; ++m) X[m][row][col] = ok ? ldo(FX, boff((unsigned)(M0 + m), P, e)) : 0.0;
      }
      if (p < (TJ + 5) * TI) {
        const int row = p / TI, col = p % TI;
        const int ii = i0 + col, jj = j0 - 3 + row;
        const bool ok = ii < NG + b.idim && jj < b.ny;
        const unsigned e = (unsigned)ii + (unsigned)jj * nx + kplane;
#pragma unroll
        for (int m = 0; m < NM; ++m) Y[m][row][col] = ok ? ldo(FX, boff((unsigned)(5 + M0 + m), P, e)) : 0.0;
      }
    }
    __syncthreads();
    if (in_ij) {
      const unsigned id = lbk + kplane;
      double acc[5];
#pragma unroll
      for (int m = 0; m < 5; ++m) acc[m] = FIRST ? 0.0 : (m < M0 ? 0.0 : ldo(b.dq, boff((unsigned)m, P, id)));
      const int pos[2] = {i, j};
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        const int n = dcount(b, d);
        const unsigned s = d == 0 ? 1u : nx;
        const int t = pos[d] - NG;
        const unsigned lbase = id - (unsigned)pos[d] * s;
        const double* cov = nullptr;
        if (INV) {
          const bool lflag = ldo(b.prim, lbase << 3) < 1.0e-20;
          const bool rflag = ldo(b.prim, (lbase + (unsigned)(n + 5) * s) << 3) < 1.0e-20;
          if (lflag && t <= 2) cov = b.fcov + b.fcov_off[d] + line_id(b, d, i, j, k) * 30 + t * 5;
          else if (rflag && t >= n - 3) cov = b.fcov + b.fcov_off[d] + line_id(b, d, i, j, k) * 30 + (3 + t - (n - 3)) * 5;
        }
#pragma unroll
        for (int m = M0; m < 5; ++m) {
          double fc;
          if (d == 0) fc = diff_node_f([&](int f) { return X[m - M0][lj][f + 2 - (i0 - 3)]; }, n, t);
          else fc = diff_node_f([&](int f) { return Y[m - M0][f + 2 - (j0 - 3)][li]; }, n, t);
          if (INV) {
            // the difference is formed for every node and replaced by the
            // closure value where the line has one
            if (cov) fc = cov[m];
            acc[m] += fc;
          } else {
            acc[m] -= re_inv * fc;
          }
        }
      }
#pragma unroll
      for (int m = M0; m < 5; ++m) {
        if (INV) acc[m] += fk[kk][m];
        else acc[m] -= re_inv * fk[kk][m];
      }
      // component 0 of the viscous difference: written (as +0) only by the
      // step's first kernel, otherwise left as it is
#pragma unroll
      for (int m = 0; m < 5; ++m)
        if (m >= M0 || FIRST) sto(b.dq, boff((unsigned)m, P, id), acc[m]);
    }
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned grid_of(long long n, int tb) { return (unsigned)((n + tb - 1) / tb); }

void launch_inv_face(const Launch& L, int d, long long n, double gamma) {
  if (n <= 0) return;
  hipLaunchKernelGGL(inv_face_kernel, dim3((unsigned)L.nwg), dim3(SPAN_TB), 0, L.stream, L.blocks, L.span, d, gamma);
}
void launch_inv_face_pairs(const Launch& L, long long n, double gamma) {
  if (n <= 0) return;
  hipLaunchKernelGGL(inv_face_pair_kernel, dim3((unsigned)L.nwg), dim3(SPAN_TB), 0, L.stream, L.blocks, L.span,
                     gamma);
}
void launch_inv_closure(const Launch& L, int d, long long n, double gamma) {
  if (n <= 0) return;
  hipLaunchKernelGGL(inv_closure_kernel, dim3((unsigned)L.nwg), dim3(SPAN_TB), 0, L.stream, L.blocks, L.span, d, gamma);
}
void launch_inv_diff(const Launch& L, long long n, bool first) {
  if (n <= 0) return;
  if (first)
    hipLaunchKernelGGL((diff_kernel<true, true>), dim3((unsigned)L.nwg), dim3(SPAN_TB), 0, L.stream, L.blocks, L.span);
  else
    hipLaunchKernelGGL((diff_kernel<true, false>), dim3((unsigned)L.nwg), dim3(SPAN_TB), 0, L.stream, L.blocks, L.span);
}
void launch_visc_deriv(const Launch& L, long long n, int* queue_head) {
  if (n <= 0) return;
  hipLaunchKernelGGL(visc_deriv_kernel, dim3((unsigned)L.nwg), dim3(SPAN_TB), 0, L.stream, L.blocks, L.span,
                     queue_head);
}
void launch_visc_face_queue(const BlockDev* blocks, const ViscItem* items, int nitems, int* head, int grid,
                            hipStream_t s) {
  if (nitems <= 0) return;
  const int g = grid < nitems ? grid : nitems;
  hipLaunchKernelGGL(visc_face_queue_kernel, dim3((unsigned)g), dim3(256), 0, s, blocks, items, nitems, head);
}
void launch_visc_fused(const BlockDev* blocks, const ViscTile* items, int nitems, int* head, int grid, bool first,
                       hipStream_t s) {
  if (nitems <= 0) return;
  const int g = grid < nitems ? grid : nitems;
  if (first)
    hipLaunchKernelGGL(visc_fused_kernel<true>, dim3((unsigned)g), dim3(256), 0, s, blocks, items, nitems, head);
  else
    hipLaunchKernelGGL(visc_fused_kernel<false>, dim3((unsigned)g), dim3(256), 0, s, blocks, items, nitems, head);
}
void launch_face_geom(const Launch& L, int d, long long n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(face_geom_kernel, dim3((unsigned)L.nwg), dim3(SPAN_TB), 0, L.stream, L.blocks, L.span, d);
}
void launch_visc_face(const Launch& L, int d, long long n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(visc_face_kernel, dim3((unsigned)L.nwg), dim3(SPAN_TB), 0, L.stream, L.blocks, L.span, d);
}
void launch_visc_diff(const Launch& L, long long n, bool first) {
  if (n <= 0) return;
  if (first)
    hipLaunchKernelGGL((diff_kernel<false, true>), dim3((unsigned)L.nwg), dim3(SPAN_TB), 0, L.stream, L.blocks,
                       L.span);
  else
    hipLaunchKernelGGL((diff_kernel<false, false>), dim3((unsigned)L.nwg), dim3(SPAN_TB), 0, L.stream, L.blocks,
                       L.span);
}

}  // namespace orion_hip
