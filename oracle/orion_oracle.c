/*
 * orion_oracle.c -- CPU restatement of ORION's per-step hot path (plain C).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never part of it.
 *
 * Every floating-point expression keeps the operand order of the reference
 * line it cites, and the file is built with -ffp-contract=off, so on x86-64
 * (SSE2 doubles, no FMA) it reproduces the reference bit for bit.
 *
 * How each part is pinned against something other than itself:
 *   - the whole step (every operator, 1..N steps, per-stage fields) against
 *     the reference compiled from its own sources (oracle/_ref, built by
 *     oracle/Makefile; tests/test_oracle_vs_reference.py) and against the
 *     committed fixtures that build wrote (tests/golden/);
 *   - the reference's own known-answer checks (src/benchmark.cpp:39-677):
 *     unit-lattice volume, BC layer-3 sign flips, state closed forms,
 *     dt closed form, viscous u=y^2 exactness, halo and interface-average
 *     fingerprints (tests/test_oracle_kat.py);
 *   - mathematics (tests/test_oracle_kat.py): the inviscid RHS is exact
 *     when every flux component is quadratic (WCNS-E-5 exact on quadratics,
 *     Roe dissipation zero for equal states, exact 6th-order difference),
 *     metrics exact under quadratic stretching, freestream preservation
 *     with the GCL metrics, LU-SGS gives dq = 0 for a zero right-hand side.
 *     The LU-SGS sweep order and the stage order are pinned only by the
 *     stage-by-stage comparison with the reference build.
 */
#include "orion_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NG OB_NG

typedef struct {
  int idim, jdim, kdim, nx, ny, nz;
  long P;
  double *prim, *metrics, *vol, *q, *dq, *dt, *c, *mu, *spec, *spec_visc;
  double *x, *y, *z;
} ob_block;

typedef struct {
  ob_region_desc d;
  int s_nd, s_lr;
  int t_nd, t_lr;
  int map_dims[3];
  int *image, *jmage, *kmage;
  double* qpvpack;
  long qpv_len;
} ob_region;

struct ob_case {
  int nb;
  ob_block* blocks;
  int* nreg;
  ob_region** reg;
  ob_params p;
};

static __thread char g_err[512];
const char* ob_last_error(void) { return g_err; }

/* How often the reference's guard branches were taken (test
 * instrumentation; no effect on results): 0 inviscid left-state fallback,
 * 1 right-state fallback (InviscidFluxComputer.cpp:346-355), 2 update
 * damped by the pressure-change limiter, 3 update refused and the old
 * state kept (StateUpdater.cpp:502-538). */
static long g_guard[4];
static void guard_hit(int which) { __atomic_fetch_add(&g_guard[which], 1, __ATOMIC_RELAXED); }
void ob_guard_counts(long* out) {
  for (int i = 0; i < 4; ++i) out[i] = __atomic_load_n(&g_guard[i], __ATOMIC_RELAXED);
}
void ob_guard_reset(void) {
  for (int i = 0; i < 4; ++i) __atomic_store_n(&g_guard[i], 0, __ATOMIC_RELAXED);
}
static int fail(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return -1;
}

static int g_threads = 0;
void ob_set_threads(int n) { g_threads = n; }
static int nthreads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

#define IX(b, i, j, k) ((long)(i) + (long)(b)->nx * ((long)(j) + (long)(b)->ny * (long)(k)))
#define FLD(arr, b, m, i, j, k) (arr)[(long)(m) * (b)->P + IX(b, i, j, k)]
#define PRIM(b, i, j, k, m) FLD((b)->prim, b, m, i, j, k)
#define MET(b, i, j, k, m) FLD((b)->metrics, b, m, i, j, k)
#define DQ(b, i, j, k, m) FLD((b)->dq, b, m, i, j, k)
#define S1(arr, b, i, j, k) (arr)[IX(b, i, j, k)]

static double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */

/* ========================================================================
 * Setup: metrics (src/mesh/MetricComputer.cpp:21-369)
 * ======================================================================== */

/* value_half_node_1d, MetricComputer.cpp:21-45 (requires ni > 4) */
static void m_half(int nv, int ni, const double* q, double* qh) {
#define Q(v, i) q[(v) + nv * (i)]
#define QH(v, i) qh[(v) + nv * (i)]
  for (int i = 2; i < ni - 1; ++i)
    for (int m = 0; m < nv; ++m)
      QH(m, i) = (9.0 * (Q(m, i - 1) + Q(m, i)) + -1.0 * (Q(m, i + 1) + Q(m, i - 2))) / 16.0;
  for (int m = 0; m < nv; ++m) {
    QH(m, 1) = (5.0 * Q(m, 0) + 15.0 * Q(m, 1) + -5.0 * Q(m, 2) + 1.0 * Q(m, 3)) / 16.0;
    QH(m, ni - 1) = (5.0 * Q(m, ni - 1) + 15.0 * Q(m, ni - 2) + -5.0 * Q(m, ni - 3) + 1.0 * Q(m, ni - 4)) / 16.0;
    QH(m, 0) = (35.0 * Q(m, 0) - 35.0 * Q(m, 1) + 21.0 * Q(m, 2) - 5.0 * Q(m, 3)) / 16.0;
    QH(m, ni) = (35.0 * Q(m, ni - 1) - 35.0 * Q(m, ni - 2) + 21.0 * Q(m, ni - 3) - 5.0 * Q(m, ni - 4)) / 16.0;
  }
#undef Q
#undef QH
}

/* flux_dxyz_1d, MetricComputer.cpp:47-99 */
static void m_diff(int nv, int ni, const double* f, double* df) {
#define FF(v, i) f[(v) + nv * (i)]
#define DF(v, i) df[(v) + nv * (i)]
  for (int i = 2; i <= ni - 3; ++i)
    for (int m = 0; m < nv; ++m) {
      double t1 = FF(m, i + 1) - FF(m, i);
      double t2 = FF(m, i + 2) - FF(m, i - 1);
      double t3 = FF(m, i + 3) - FF(m, i - 2);
      DF(m, i) = (2250.0 * t1 - 125.0 * t2 + 9.0 * t3) / 1920.0;
    }
  for (int m = 0; m < nv; ++m) DF(m, 0) = (-23.0 * FF(m, 0) + 21.0 * FF(m, 1) + 3.0 * FF(m, 2) - FF(m, 3)) / 24.0;
  for (int m = 0; m < nv; ++m) DF(m, 1) = (FF(m, 0) - 27.0 * FF(m, 1) + 27.0 * FF(m, 2) - FF(m, 3)) / 24.0;
  for (int m = 0; m < nv; ++m)
    DF(m, ni - 2) = -(FF(m, ni) - 27.0 * FF(m, ni - 1) + 27.0 * FF(m, ni - 2) - FF(m, ni - 3)) / 24.0;
  for (int m = 0; m < nv; ++m)
    DF(m, ni - 1) = -(-23.0 * FF(m, ni) + 21.0 * FF(m, ni - 1) + 3.0 * FF(m, ni - 2) - FF(m, ni - 3)) / 24.0;
#undef FF
#undef DF
}

/* Differentiate nv interleaved quantities along axis ax of an
 * (ni,nj,nk) node box: in(v, i,j,k) -> out written through put(). */
static void metric_block(ob_block* b, int gcl) {
  const int ni = b->idim, nj = b->jdim, nk = b->kdim;
  const long n = (long)ni * nj * nk;
  double* D = calloc((size_t)18 * n, sizeof(double));
  int nmax = ni > nj ? ni : nj;
  if (nk > nmax) nmax = nk;
  double* lin = malloc(sizeof(double) * 6 * (nmax + 1));
  double* lhf = malloc(sizeof(double) * 6 * (nmax + 2));
  double* lout = malloc(sizeof(double) * 6 * (nmax + 1));
#define NODE(i, j, k) ((long)(i) + (long)ni * ((long)(j) + (long)nj * (long)(k)))
#define DX(c, i, j, k) D[(c) + 18 * NODE(i, j, k)]
  const double *X = b->x, *Y = b->y, *Z = b->z;
  /* pass 1: xi (:176-191), eta (:193-208), zeta (:210-225) */
  for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j) {
      for (int i = 0; i < ni; ++i) {
        lin[3 * i] = X[NODE(i, j, k)];
        lin[3 * i + 1] = Y[NODE(i, j, k)];
        lin[3 * i + 2] = Z[NODE(i, j, k)];
      }
      m_half(3, ni, lin, lhf);
      m_diff(3, ni, lhf, lout);
      for (int i = 0; i < ni; ++i) {
        DX(0, i, j, k) = lout[3 * i];
        DX(3, i, j, k) = lout[3 * i + 1];
        DX(6, i, j, k) = lout[3 * i + 2];
      }
    }
  for (int k = 0; k < nk; ++k)
    for (int i = 0; i < ni; ++i) {
      for (int j = 0; j < nj; ++j) {
        lin[3 * j] = X[NODE(i, j, k)];
        lin[3 * j + 1] = Y[NODE(i, j, k)];
        lin[3 * j + 2] = Z[NODE(i, j, k)];
      }
      m_half(3, nj, lin, lhf);
      m_diff(3, nj, lhf, lout);
      for (int j = 0; j < nj; ++j) {
        DX(1, i, j, k) = lout[3 * j];
        DX(4, i, j, k) = lout[3 * j + 1];
        DX(7, i, j, k) = lout[3 * j + 2];
      }
    }
  for (int j = 0; j < nj; ++j)
    for (int i = 0; i < ni; ++i) {
      for (int k = 0; k < nk; ++k) {
        lin[3 * k] = X[NODE(i, j, k)];
        lin[3 * k + 1] = Y[NODE(i, j, k)];
        lin[3 * k + 2] = Z[NODE(i, j, k)];
      }
      m_half(3, nk, lin, lhf);
      m_diff(3, nk, lhf, lout);
      for (int k = 0; k < nk; ++k) {
        DX(2, i, j, k) = lout[3 * k];
        DX(5, i, j, k) = lout[3 * k + 1];
        DX(8, i, j, k) = lout[3 * k + 2];
      }
    }
  /* standard cross products (:228-252) */
  for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j)
      for (int i = 0; i < ni; ++i) {
        double x_xi = DX(0, i, j, k), x_et = DX(1, i, j, k), x_zt = DX(2, i, j, k);
        double y_xi = DX(3, i, j, k), y_et = DX(4, i, j, k), y_zt = DX(5, i, j, k);
        double z_xi = DX(6, i, j, k), z_et = DX(7, i, j, k), z_zt = DX(8, i, j, k);
        double kcx = y_et * z_zt - z_et * y_zt;
        double kcy = z_et * x_zt - x_et * z_zt;
        double kcz = x_et * y_zt - y_et * x_zt;
        double etx = y_zt * z_xi - z_zt * y_xi;
        double ety = z_zt * x_xi - x_zt * z_xi;
        double etz = x_zt * y_xi - y_zt * x_xi;
        double ctx = y_xi * z_et - z_xi * y_et;
        double cty = z_xi * x_et - x_xi * z_et;
        double ctz = x_xi * y_et - y_xi * x_et;
        double vol = x_xi * kcx + x_et * etx + x_zt * ctx;
        int I = i + NG, J = j + NG, K = k + NG;
        MET(b, I, J, K, 0) = kcx; MET(b, I, J, K, 1) = kcy; MET(b, I, J, K, 2) = kcz;
        MET(b, I, J, K, 3) = etx; MET(b, I, J, K, 4) = ety; MET(b, I, J, K, 5) = etz;
        MET(b, I, J, K, 6) = ctx; MET(b, I, J, K, 7) = cty; MET(b, I, J, K, 8) = ctz;
        S1(b->vol, b, I, J, K) = vol;
      }
  if (gcl == 1) { /* Visbal-Gaitonde conservative form (:257-367) */
    double* XK = malloc(sizeof(double) * 3 * n);
    for (int k = 0; k < nk; ++k)
      for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i) {
          double xm = X[NODE(i, j, k)], ym = Y[NODE(i, j, k)], zm = Z[NODE(i, j, k)];
          long nd = NODE(i, j, k);
          XK[3 * nd] = DX(0, i, j, k);
          XK[3 * nd + 1] = DX(1, i, j, k);
          XK[3 * nd + 2] = DX(2, i, j, k);
          int I = i + NG, J = j + NG, K = k + NG;
          MET(b, I, J, K, 0) = DX(0, i, j, k) * ym;
          MET(b, I, J, K, 1) = DX(1, i, j, k) * ym;
          MET(b, I, J, K, 2) = DX(2, i, j, k) * ym;
          MET(b, I, J, K, 3) = DX(3, i, j, k) * zm;
          MET(b, I, J, K, 4) = DX(4, i, j, k) * zm;
          MET(b, I, J, K, 5) = DX(5, i, j, k) * zm;
          MET(b, I, J, K, 6) = DX(6, i, j, k) * xm;
          MET(b, I, J, K, 7) = DX(7, i, j, k) * xm;
          MET(b, I, J, K, 8) = DX(8, i, j, k) * xm;
        }
    static const int sel[3][6] = {{7, 8, 1, 2, 4, 5}, {6, 8, 0, 2, 3, 5}, {6, 7, 0, 1, 3, 4}};
    for (int k = 0; k < nk; ++k)
      for (int j = 0; j < nj; ++j) {
        for (int i = 0; i < ni; ++i)
          for (int c = 0; c < 6; ++c) lin[c + 6 * i] = MET(b, i + NG, j + NG, k + NG, sel[0][c]);
        m_half(6, ni, lin, lhf);
        m_diff(6, ni, lhf, lout);
        for (int i = 0; i < ni; ++i)
          for (int c = 0; c < 6; ++c) DX(c, i, j, k) = lout[c + 6 * i];
      }
    for (int k = 0; k < nk; ++k)
      for (int i = 0; i < ni; ++i) {
        for (int j = 0; j < nj; ++j)
          for (int c = 0; c < 6; ++c) lin[c + 6 * j] = MET(b, i + NG, j + NG, k + NG, sel[1][c]);
        m_half(6, nj, lin, lhf);
        m_diff(6, nj, lhf, lout);
        for (int j = 0; j < nj; ++j)
          for (int c = 0; c < 6; ++c) DX(c + 6, i, j, k) = lout[c + 6 * j];
      }
    for (int j = 0; j < nj; ++j)
      for (int i = 0; i < ni; ++i) {
        for (int k = 0; k < nk; ++k)
          for (int c = 0; c < 6; ++c) lin[c + 6 * k] = MET(b, i + NG, j + NG, k + NG, sel[2][c]);
        m_half(6, nk, lin, lhf);
        m_diff(6, nk, lhf, lout);
        for (int k = 0; k < nk; ++k)
          for (int c = 0; c < 6; ++c) DX(c + 12, i, j, k) = lout[c + 6 * k];
      }
    for (int k = 0; k < nk; ++k)
      for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i) {
          double kcx = DX(17, i, j, k) - DX(11, i, j, k);
          double kcy = DX(13, i, j, k) - DX(7, i, j, k);
          double kcz = DX(15, i, j, k) - DX(9, i, j, k);
          double etx = DX(5, i, j, k) - DX(16, i, j, k);
          double ety = DX(1, i, j, k) - DX(12, i, j, k);
          double etz = DX(3, i, j, k) - DX(14, i, j, k);
          double ctx = DX(10, i, j, k) - DX(4, i, j, k);
          double cty = DX(6, i, j, k) - DX(0, i, j, k);
          double ctz = DX(8, i, j, k) - DX(2, i, j, k);
          int I = i + NG, J = j + NG, K = k + NG;
          MET(b, I, J, K, 0) = kcx; MET(b, I, J, K, 1) = kcy; MET(b, I, J, K, 2) = kcz;
          MET(b, I, J, K, 3) = etx; MET(b, I, J, K, 4) = ety; MET(b, I, J, K, 5) = etz;
          MET(b, I, J, K, 6) = ctx; MET(b, I, J, K, 7) = cty; MET(b, I, J, K, 8) = ctz;
          long nd = NODE(i, j, k);
          S1(b->vol, b, I, J, K) = XK[3 * nd] * kcx + XK[3 * nd + 1] * etx + XK[3 * nd + 2] * ctx;
        }
    free(XK);
  }
#undef DX
#undef NODE
  free(D);
  free(lin);
  free(lhf);
  free(lout);
}

/* ========================================================================
 * Setup: interface maps (src/bc/BCPreprocess.cpp:16-192) and region
 * geometry (src/bc/BCReader.cpp:125-143)
 * ======================================================================== */

static int analyze_connect(ob_region* r) {
  const int* s_st = r->d.s_st;
  const int* s_ed = r->d.s_ed;
  const int* t_st = r->d.t_st;
  const int* t_ed = r->d.t_ed;
  const int s_nd = r->s_nd;
  r->t_nd = 0;
  for (int m = 0; m < 3; ++m)
    if (t_st[m] == t_ed[m]) {
      r->t_nd = m + 1;
      r->t_lr = (t_st[m] == 1) ? -1 : 1;
    }
  const int t_nd = r->t_nd;
  if (t_nd == 0 || s_nd == 0) return fail("interface region has no constant axis");
  int dirm[3][3] = {{0}};
  for (int m = 0; m < 3; ++m)
    for (int n = 0; n < 3; ++n)
      if ((m + 1) != s_nd && (n + 1) != t_nd) {
        int js1 = s_st[m];
        if (abs(js1) < abs(s_ed[m])) js1 = s_ed[m];
        int js2 = t_st[n];
        if (abs(js2) < abs(t_ed[n])) js2 = t_ed[n];
        if (js1 * js2 > 0) {
          dirm[n][m] = 1;
          dirm[t_nd - 1][s_nd - 1] = 1;
          int rem_t = 6 - (n + 1) - t_nd;
          int rem_s = 6 - (m + 1) - s_nd;
          dirm[rem_t - 1][rem_s - 1] = 1;
          goto found;
        }
      }
found:;
  int s_sign[3], t_sign[3], st_sign[3];
  for (int m = 0; m < 3; ++m) {
    s_sign[m] = 1;
    if ((m + 1) != s_nd && abs(s_st[m]) > abs(s_ed[m])) s_sign[m] = -1;
    t_sign[m] = 1;
    if ((m + 1) != t_nd && abs(t_st[m]) > abs(t_ed[m])) t_sign[m] = -1;
  }
  for (int m = 0; m < 3; ++m) {
    int co = 0;
    for (int n = 0; n < 3; ++n)
      if (dirm[n][m] == 1) co = n;
    st_sign[m] = t_sign[co] * s_sign[m];
  }
  int lo[3], hi[3];
  for (int m = 0; m < 3; ++m) {
    int a = abs(s_st[m]), bb = abs(s_ed[m]);
    lo[m] = a < bb ? a : bb;
    hi[m] = a < bb ? bb : a;
    r->map_dims[m] = hi[m] - lo[m] + 1;
  }
  long tot = (long)r->map_dims[0] * r->map_dims[1] * r->map_dims[2];
  r->image = calloc(tot, sizeof(int));
  r->jmage = calloc(tot, sizeof(int));
  r->kmage = calloc(tot, sizeof(int));
  for (int k = abs(s_st[2]); k != abs(s_ed[2]) + s_sign[2]; k += s_sign[2])
    for (int j = abs(s_st[1]); j != abs(s_ed[1]) + s_sign[1]; j += s_sign[1])
      for (int i = abs(s_st[0]); i != abs(s_ed[0]) + s_sign[0]; i += s_sign[0]) {
        int idelt = (i - abs(s_st[0])) * st_sign[0];
        int jdelt = (j - abs(s_st[1])) * st_sign[1];
        int kdelt = (k - abs(s_st[2])) * st_sign[2];
        int vi = abs(t_st[0]) + dirm[0][0] * idelt + dirm[0][1] * jdelt + dirm[0][2] * kdelt;
        int vj = abs(t_st[1]) + dirm[1][0] * idelt + dirm[1][1] * jdelt + dirm[1][2] * kdelt;
        int vk = abs(t_st[2]) + dirm[2][0] * idelt + dirm[2][1] * jdelt + dirm[2][2] * kdelt;
        long f = ((long)(k - lo[2]) * r->map_dims[1] + (j - lo[1])) * r->map_dims[0] + (i - lo[0]);
        r->image[f] = vi;
        r->jmage[f] = vj;
        r->kmage[f] = vk;
      }
  return 0;
}

/* ========================================================================
 * Create / destroy
 * ======================================================================== */

static void free_case(ob_case* c) {
  if (!c) return;
  for (int b = 0; b < c->nb; ++b) {
    ob_block* k = &c->blocks[b];
    free(k->prim); free(k->metrics); free(k->vol); free(k->q); free(k->dq);
    free(k->dt); free(k->c); free(k->mu); free(k->spec); free(k->spec_visc);
    free(k->x); free(k->y); free(k->z);
    if (c->reg && c->reg[b]) {
      for (int r = 0; r < c->nreg[b]; ++r) {
        free(c->reg[b][r].image); free(c->reg[b][r].jmage); free(c->reg[b][r].kmage);
        free(c->reg[b][r].qpvpack);
      }
      free(c->reg[b]);
    }
  }
  free(c->blocks);
  free(c->nreg);
  free(c->reg);
  free(c);
}
void ob_destroy(ob_case* c) { free_case(c); }

/* init_inflow (src/preprocess/Inflow.cpp:300-459) restricted to the custom-
 * atmosphere path (height<0) with tref given: the quantities the hot path
 * reads.  setup_perfect_gas_air (:15-46) fixes gama=1.4, prl=0.72. */
static void init_params(ob_params* p, const ob_inputs* in) {
  const double PAI = 3.14159265358979323846;
  memset(p, 0, sizeof *p);
  p->gama = 1.40;
  p->prl = 0.72;
  double attack = in->attack * PAI / 180.0;
  double sideslip = in->sideslip * PAI / 180.0;
  p->uoo = cos(attack) * cos(sideslip);
  p->voo = sin(attack) * cos(sideslip);
  p->woo = sin(sideslip);
  p->roo = 1.0;
  p->too = 1.0;
  p->reynolds = (in->reynolds / 1.0) * 1.0; /* (reynolds/lfref)*lref, both 1 */
  p->moo = in->moo;
  p->poo = 1.0 / (p->gama * p->moo * p->moo);
  p->tref = in->tref;
  p->twall = in->twall;
  p->visc = (in->tref != 0.0) ? 110.4 / in->tref : 0.0;
  p->cfl = in->cfl;
  p->csrv = in->csrv;
  p->nvis = in->nvis;
  p->method = in->method;
  p->ntmst = in->ntmst;
  p->gcl = in->gcl;
  p->phydtime = 0.0;
}

ob_case* ob_create(int nblocks, const int* dims, const double* xyz, int nregions,
                   const ob_region_desc* regions, const ob_inputs* in) {
  if (nblocks <= 0) { fail("nblocks <= 0"); return NULL; }
  ob_case* c = calloc(1, sizeof *c);
  c->nb = nblocks;
  c->blocks = calloc(nblocks, sizeof(ob_block));
  c->nreg = calloc(nblocks, sizeof(int));
  c->reg = calloc(nblocks, sizeof(ob_region*));
  init_params(&c->p, in);
  long off = 0;
  for (int b = 0; b < nblocks; ++b) {
    ob_block* k = &c->blocks[b];
    k->idim = dims[3 * b];
    k->jdim = dims[3 * b + 1];
    k->kdim = dims[3 * b + 2];
    if (k->idim <= 2 * NG || k->jdim <= 2 * NG || k->kdim <= 2 * NG) {
      /* MetricComputer.cpp:154-157 leaves such blocks with zero volume, which
       * check_grid_metrics then rejects. */
      fail("block %d too small (%d,%d,%d): every dimension must exceed %d", b, k->idim, k->jdim,
           k->kdim, 2 * NG);
      free_case(c);
      return NULL;
    }
    k->nx = k->idim + 2 * NG;
    k->ny = k->jdim + 2 * NG;
    k->nz = k->kdim + 2 * NG;
    k->P = (long)k->nx * k->ny * k->nz;
    k->prim = calloc(6 * k->P, sizeof(double));
    k->metrics = calloc(9 * k->P, sizeof(double));
    k->vol = calloc(k->P, sizeof(double));
    k->q = calloc(5 * k->P, sizeof(double));
    k->dq = calloc(5 * k->P, sizeof(double));
    k->dt = calloc(k->P, sizeof(double));
    k->c = calloc(k->P, sizeof(double));
    k->mu = calloc(k->P, sizeof(double));
    k->spec = calloc(3 * k->P, sizeof(double));
    k->spec_visc = calloc(3 * k->P, sizeof(double));
    long n = (long)k->idim * k->jdim * k->kdim;
    k->x = malloc(n * sizeof(double));
    k->y = malloc(n * sizeof(double));
    k->z = malloc(n * sizeof(double));
    memcpy(k->x, xyz + off, n * sizeof(double));
    memcpy(k->y, xyz + off + n, n * sizeof(double));
    memcpy(k->z, xyz + off + 2 * n, n * sizeof(double));
    off += 3 * n;
  }
  /* regions, file order per block */
  for (int r = 0; r < nregions; ++r) {
    int b = regions[r].block;
    if (b < 0 || b >= nblocks) { fail("region %d: bad block", r); free_case(c); return NULL; }
    c->nreg[b]++;
  }
  for (int b = 0; b < nblocks; ++b) c->reg[b] = calloc(c->nreg[b] ? c->nreg[b] : 1, sizeof(ob_region));
  int* fill = calloc(nblocks, sizeof(int));
  for (int r = 0; r < nregions; ++r) {
    const ob_region_desc* d = &regions[r];
    int b = d->block;
    ob_region* R = &c->reg[b][fill[b]++];
    R->d = *d;
    if (d->bctype >= 0) { R->d.nbt = -1; R->d.ibcwin = 0; }
    int bt = d->bctype;
    if (!(bt == -1 || bt == 2 || bt == 3 || bt == 4 || bt == 5 || bt == 6)) {
      fail("block %d: bctype %d not among {-1,4,5,6,2,3}", b + 1, bt);
      free(fill);
      free_case(c);
      return NULL;
    }
    const ob_block* k = &c->blocks[b];
    int dm[3] = {k->idim, k->jdim, k->kdim};
    R->s_nd = 0;
    R->s_lr = 0;
    for (int m = 0; m < 3; ++m)
      if (d->s_st[m] == d->s_ed[m]) {
        R->s_nd = m + 1;
        if (d->s_st[m] == 1) R->s_lr = -1;
        else if (d->s_st[m] == dm[m]) R->s_lr = 1;
        else R->s_lr = 0;
        break;
      }
    if (bt < 0) {
      if (R->s_nd == 0 || d->ibcwin <= 0 || d->nbt < 1 || d->nbt > nblocks) {
        fail("block %d region %d: invalid interface", b + 1, fill[b]);
        free(fill);
        free_case(c);
        return NULL;
      }
      if (analyze_connect(R) != 0) { free(fill); free_case(c); return NULL; }
    }
  }
  free(fill);
  /* interface windows must agree in size with their partner (the reference
   * indexes the partner buffer without a bound check) */
  for (int b = 0; b < nblocks; ++b)
    for (int r = 0; r < c->nreg[b]; ++r) {
      ob_region* R = &c->reg[b][r];
      if (R->d.bctype >= 0) continue;
      int tb = R->d.nbt - 1, tw = R->d.ibcwin - 1;
      if (tw >= c->nreg[tb]) { fail("block %d region %d: target window out of range", b + 1, r + 1); free_case(c); return NULL; }
      ob_region* T = &c->reg[tb][tw];
      long ns = 1, nt = 1;
      for (int m = 0; m < 3; ++m) {
        ns *= abs(abs(R->d.s_st[m]) - abs(R->d.s_ed[m])) + 1;
        nt *= abs(abs(T->d.s_st[m]) - abs(T->d.s_ed[m])) + 1;
      }
      if (ns != nt) { fail("block %d region %d: partner window size mismatch", b + 1, r + 1); free_case(c); return NULL; }
    }
  /* metrics + check (check_grid_metrics, MetricComputer.cpp:371-500) */
  for (int b = 0; b < nblocks; ++b) {
    ob_block* k = &c->blocks[b];
    metric_block(k, c->p.gcl);
    for (int kk = 0; kk < k->kdim; ++kk)
      for (int j = 0; j < k->jdim; ++j)
        for (int i = 0; i < k->idim; ++i)
          if (S1(k->vol, k, i + NG, j + NG, kk + NG) < 1.0e-15) {
            fail("negative Jacobian in block %d at (%d,%d,%d)", b + 1, i + 1, j + 1, kk + 1);
            free_case(c);
            return NULL;
          }
  }
  /* initialize_flow (InitialCondition.cpp:49-80) then update_temperature */
  const ob_params* p = &c->p;
  double v2 = p->uoo * p->uoo + p->voo * p->voo + p->woo * p->woo;
  double eoo = p->poo / (p->gama - 1.0) + 0.5 * p->roo * v2;
  double qoo[5] = {p->roo, p->roo * p->uoo, p->roo * p->voo, p->roo * p->woo, eoo};
  double poo6[6] = {p->roo, p->uoo, p->voo, p->woo, p->poo, p->too};
  for (int b = 0; b < nblocks; ++b) {
    ob_block* k = &c->blocks[b];
    for (int m = 0; m < 5; ++m)
      for (long i = 0; i < k->P; ++i) k->q[m * k->P + i] = qoo[m];
    for (int m = 0; m < 6; ++m)
      for (long i = 0; i < k->P; ++i) k->prim[m * k->P + i] = poo6[m];
  }
  ob_update_temperature(c);
  return c;
}

int ob_nblocks(const ob_case* c) { return c->nb; }
void ob_block_dims(const ob_case* c, int b, int* d) {
  d[0] = c->blocks[b].nx;
  d[1] = c->blocks[b].ny;
  d[2] = c->blocks[b].nz;
}
const ob_params* ob_get_params(const ob_case* c) { return &c->p; }

double* ob_field(ob_case* c, int b, int f) {
  ob_block* k = &c->blocks[b];
  switch (f) {
    case OB_PRIM: return k->prim;
    case OB_METRICS: return k->metrics;
    case OB_VOL: return k->vol;
    case OB_Q: return k->q;
    case OB_DQ: return k->dq;
    case OB_DT: return k->dt;
    case OB_C: return k->c;
    case OB_MU: return k->mu;
    case OB_SPEC: return k->spec;
    case OB_SPEC_VISC: return k->spec_visc;
  }
  return NULL;
}

/* update_temperature, InitialCondition.cpp:91-147 */
void ob_update_temperature(ob_case* c) {
  if (c->p.nvis == 0) return;
  const double gama = c->p.gama, mach = c->p.moo, mach2 = mach * mach;
  for (int b = 0; b < c->nb; ++b) {
    ob_block* k = &c->blocks[b];
    for (long i = 0; i < k->P; ++i) {
      double rho = k->prim[i];
      double p = k->prim[4 * k->P + i];
      if (rho < 1.0e-30) rho = 1.0e-30;
      double a2 = gama * p / rho;
      k->prim[5 * k->P + i] = mach2 * a2;
    }
  }
}

/* ========================================================================
 * Halo exchange (src/solver/HaloExchanger.cpp:17-179)
 * ======================================================================== */

static void memory_window(const ob_region* r, const int* dphy, int mode, int* rb, int* re) {
  int cc[3];
  for (int m = 0; m < 3; ++m) cc[m] = abs(r->d.s_st[m]) - 1 + NG;
  int dir = r->s_nd - 1;
  if (mode == 0) {
    if (r->s_lr == -1) { rb[dir] = cc[dir]; re[dir] = cc[dir] + NG - 1; }
    else { rb[dir] = cc[dir] - NG + 1; re[dir] = cc[dir]; }
  } else {
    if (r->s_lr == -1) { rb[dir] = cc[dir] - NG; re[dir] = cc[dir] - 1; }
    else { rb[dir] = cc[dir] + 1; re[dir] = cc[dir] + NG; }
  }
  for (int m = 0; m < 3; ++m) {
    if (m == dir) continue;
    int a = abs(r->d.s_st[m]), bb = abs(r->d.s_ed[m]);
    int ps = a < bb ? a : bb, pe = a < bb ? bb : a;
    if (ps > 1) ps -= 1;
    if (pe < dphy[m]) pe += 1;
    rb[m] = ps - 1 + NG;
    re[m] = pe - 1 + NG;
  }
}

typedef struct {
  int tag;
  int used;
  long n;
  double* buf;
} ob_msg;

int ob_exchange_bc(ob_case* c) {
  int nmsg = 0;
  for (int b = 0; b < c->nb; ++b)
    for (int r = 0; r < c->nreg[b]; ++r)
      if (c->reg[b][r].d.nbt > 0) nmsg++;
  ob_msg* msgs = calloc(nmsg ? nmsg : 1, sizeof(ob_msg));
  int im = 0;
  for (int b = 0; b < c->nb; ++b) {
    ob_block* k = &c->blocks[b];
    int dphy[3] = {k->idim, k->jdim, k->kdim};
    for (int r = 0; r < c->nreg[b]; ++r) {
      const ob_region* R = &c->reg[b][r];
      if (R->d.nbt <= 0) continue;
      int rb[3], re[3];
      memory_window(R, dphy, 0, rb, re);
      long n = (long)(re[0] - rb[0] + 1) * (re[1] - rb[1] + 1) * (re[2] - rb[2] + 1);
      ob_msg* M = &msgs[im++];
      M->buf = malloc(sizeof(double) * 5 * n);
      M->n = 5 * n;
      M->tag = R->d.nbt * 1000 + R->d.ibcwin;
      long p = 0;
      for (int m = 0; m < 5; ++m)
        for (int kk = rb[2]; kk <= re[2]; ++kk)
          for (int j = rb[1]; j <= re[1]; ++j)
            for (int i = rb[0]; i <= re[0]; ++i) M->buf[p++] = PRIM(k, i, j, kk, m);
    }
  }
  int rc = 0;
  for (int b = 0; b < c->nb && rc == 0; ++b) {
    ob_block* k = &c->blocks[b];
    int dphy[3] = {k->idim, k->jdim, k->kdim};
    for (int r = 0; r < c->nreg[b]; ++r) {
      const ob_region* R = &c->reg[b][r];
      if (R->d.nbt <= 0) continue;
      int tag = (b + 1) * 1000 + (r + 1);
      ob_msg* M = NULL;
      for (int q = 0; q < nmsg; ++q)
        if (!msgs[q].used && msgs[q].tag == tag) { M = &msgs[q]; break; }
      if (!M) { rc = fail("halo: no message for block %d window %d", b + 1, r + 1); break; }
      M->used = 1;
      int rb[3], re[3];
      memory_window(R, dphy, 1, rb, re);
      long p = 0;
      for (int m = 0; m < 5; ++m)
        for (int kk = rb[2]; kk <= re[2]; ++kk)
          for (int j = rb[1]; j <= re[1]; ++j)
            for (int i = rb[0]; i <= re[0]; ++i)
              if (p < M->n) PRIM(k, i, j, kk, m) = M->buf[p++];
    }
  }
  for (int q = 0; q < nmsg; ++q) free(msgs[q].buf);
  free(msgs);
  return rc;
}

/* ========================================================================
 * Physical boundary conditions (src/bc/PhysicalBC.cpp)
 * ======================================================================== */

static void get_normal(const ob_block* k, int i, int j, int kk, int s_nd, double* nx, double* ny, double* nz) {
  int off = (s_nd - 1) * 3;
  double mx = MET(k, i, j, kk, off), my = MET(k, i, j, kk, off + 1), mz = MET(k, i, j, kk, off + 2);
  double len = sqrt(mx * mx + my * my + mz * mz);
  if (len < 1.0e-30) len = 1.0e-30;
  *nx = mx / len;
  *ny = my / len;
  *nz = mz / len;
}

static void dif_average(ob_block* k, int i, int j, int kk, int dir, int s_lr) {
  int o[3] = {i, j, kk}, g[3] = {i, j, kk}, n[3] = {i, j, kk};
  if (s_lr == 1) { g[dir] += 1; n[dir] -= 1; }
  else { g[dir] -= 1; n[dir] += 1; }
  for (int m = 0; m < 6; ++m) {
    double vg = PRIM(k, g[0], g[1], g[2], m);
    double vi = PRIM(k, n[0], n[1], n[2], m);
    PRIM(k, o[0], o[1], o[2], m) = 0.5 * (vg + vi);
  }
}

#define WIN                                   \
  int i_st = R->d.s_st[0] - 1 + NG;           \
  int i_ed = R->d.s_ed[0] - 1 + NG;           \
  int j_st = R->d.s_st[1] - 1 + NG;           \
  int j_ed = R->d.s_ed[1] - 1 + NG;           \
  int k_st = R->d.s_st[2] - 1 + NG;           \
  int k_ed = R->d.s_ed[2] - 1 + NG;           \
  int dir = R->s_nd - 1

static void bc_symmetry(const ob_region* R, ob_block* k, const ob_params* p) {
  WIN;
  int kphys = k->nz - 2 * NG;
  int is2d = (dir != 2 && kphys <= 2);
  if (is2d) {
    for (int kk = k_st; kk <= k_ed; ++kk)
      for (int j = j_st; j <= j_ed; ++j)
        for (int i = i_st; i <= i_ed; ++i) {
          for (int l = 0; l < NG; ++l) {
            int d = l + 1, g[3] = {i, j, kk}, n[3] = {i, j, kk};
            if (R->s_lr == 1) { g[dir] += d; n[dir] -= d; }
            else { g[dir] -= d; n[dir] += d; }
            PRIM(k, g[0], g[1], g[2], 0) = PRIM(k, n[0], n[1], n[2], 0);
            PRIM(k, g[0], g[1], g[2], 1) = PRIM(k, n[0], n[1], n[2], 1);
            PRIM(k, g[0], g[1], g[2], 2) = PRIM(k, n[0], n[1], n[2], 2);
            PRIM(k, g[0], g[1], g[2], 4) = PRIM(k, n[0], n[1], n[2], 4);
            PRIM(k, g[0], g[1], g[2], 5) = PRIM(k, n[0], n[1], n[2], 5);
            PRIM(k, g[0], g[1], g[2], 3) = -PRIM(k, n[0], n[1], n[2], 3);
          }
          if (p->method == 1) dif_average(k, i, j, kk, dir, R->s_lr);
        }
  } else {
    double nx, ny, nz;
    get_normal(k, (i_st + i_ed) / 2, (j_st + j_ed) / 2, (k_st + k_ed) / 2, R->s_nd, &nx, &ny, &nz);
    for (int kk = k_st; kk <= k_ed; ++kk)
      for (int j = j_st; j <= j_ed; ++j)
        for (int i = i_st; i <= i_ed; ++i) {
          for (int l = 0; l < NG; ++l) {
            int d = l + 1, g[3] = {i, j, kk}, n[3] = {i, j, kk};
            if (R->s_lr == 1) { g[dir] += d; n[dir] -= d; }
            else { g[dir] -= d; n[dir] += d; }
            double rho = PRIM(k, n[0], n[1], n[2], 0);
            double u = PRIM(k, n[0], n[1], n[2], 1);
            double v = PRIM(k, n[0], n[1], n[2], 2);
            double w = PRIM(k, n[0], n[1], n[2], 3);
            double pp = PRIM(k, n[0], n[1], n[2], 4);
            double vdn = u * nx + v * ny + w * nz;
            PRIM(k, g[0], g[1], g[2], 0) = rho;
            PRIM(k, g[0], g[1], g[2], 1) = u - 2.0 * vdn * nx;
            PRIM(k, g[0], g[1], g[2], 2) = v - 2.0 * vdn * ny;
            PRIM(k, g[0], g[1], g[2], 3) = w - 2.0 * vdn * nz;
            PRIM(k, g[0], g[1], g[2], 4) = pp;
            PRIM(k, g[0], g[1], g[2], 5) = PRIM(k, n[0], n[1], n[2], 5);
          }
          if (p->method == 1) dif_average(k, i, j, kk, dir, R->s_lr);
        }
  }
}

static void bc_inviscid_wall(const ob_region* R, ob_block* k, const ob_params* p) {
  WIN;
  for (int kk = k_st; kk <= k_ed; ++kk)
    for (int j = j_st; j <= j_ed; ++j)
      for (int i = i_st; i <= i_ed; ++i) {
        double nx, ny, nz;
        get_normal(k, i, j, kk, R->s_nd, &nx, &ny, &nz);
        for (int l = 0; l < NG; ++l) {
          int d = l + 1, g[3] = {i, j, kk}, n[3] = {i, j, kk};
          if (R->s_lr == 1) { g[dir] += d; n[dir] -= d; }
          else { g[dir] -= d; n[dir] += d; }
          double rho_i = PRIM(k, n[0], n[1], n[2], 0);
          double u_i = PRIM(k, n[0], n[1], n[2], 1);
          double v_i = PRIM(k, n[0], n[1], n[2], 2);
          double w_i = PRIM(k, n[0], n[1], n[2], 3);
          double p_i = PRIM(k, n[0], n[1], n[2], 4);
          double t_i = PRIM(k, n[0], n[1], n[2], 5);
          double vdn = u_i * nx + v_i * ny + w_i * nz;
          PRIM(k, g[0], g[1], g[2], 0) = (d == 3) ? -rho_i : rho_i;
          PRIM(k, g[0], g[1], g[2], 1) = u_i - 2.0 * vdn * nx;
          PRIM(k, g[0], g[1], g[2], 2) = v_i - 2.0 * vdn * ny;
          PRIM(k, g[0], g[1], g[2], 3) = w_i - 2.0 * vdn * nz;
          PRIM(k, g[0], g[1], g[2], 4) = p_i;
          PRIM(k, g[0], g[1], g[2], 5) = t_i;
        }
        if (p->method == 1) dif_average(k, i, j, kk, dir, R->s_lr);
      }
}

static void bc_viscid_wall(const ob_region* R, ob_block* k, const ob_params* p) {
  WIN;
  double gama = p->gama, mach = p->moo, twall = p->twall, tref = p->tref;
  double moocp = mach * mach * gama;
  int iso = (twall > 1.0e-6);
  double twall_dim = (iso && fabs(tref) > 1e-16) ? twall / tref : twall;
  for (int kk = k_st; kk <= k_ed; ++kk)
    for (int j = j_st; j <= j_ed; ++j)
      for (int i = i_st; i <= i_ed; ++i) {
        int a1[3] = {i, j, kk}, a2[3] = {i, j, kk};
        if (R->s_lr == 1) { a1[dir] -= 1; a2[dir] -= 2; }
        else { a1[dir] += 1; a2[dir] += 2; }
        double p1 = PRIM(k, a1[0], a1[1], a1[2], 4);
        double p2 = PRIM(k, a2[0], a2[1], a2[2], 4);
        double p_wall = (4.0 * p1 - p2) / 3.0;
        if (p_wall < 1.0e-6) p_wall = p1;
        double t_wall;
        if (iso) {
          t_wall = twall_dim;
        } else {
          double t1 = PRIM(k, a1[0], a1[1], a1[2], 5);
          double t2 = PRIM(k, a2[0], a2[1], a2[2], 5);
          t_wall = (4.0 * t1 - t2) / 3.0;
          if (t_wall < 1e-6) t_wall = t1;
        }
        double rho_wall = moocp * p_wall / t_wall;
        for (int l = 0; l < NG; ++l) {
          int d = l + 1, g[3] = {i, j, kk}, n[3] = {i, j, kk};
          if (R->s_lr == 1) { g[dir] += d; n[dir] -= d; }
          else { g[dir] -= d; n[dir] += d; }
          double u_i = PRIM(k, n[0], n[1], n[2], 1);
          double v_i = PRIM(k, n[0], n[1], n[2], 2);
          double w_i = PRIM(k, n[0], n[1], n[2], 3);
          PRIM(k, g[0], g[1], g[2], 1) = -u_i;
          PRIM(k, g[0], g[1], g[2], 2) = -v_i;
          PRIM(k, g[0], g[1], g[2], 3) = -w_i;
          PRIM(k, g[0], g[1], g[2], 4) = PRIM(k, n[0], n[1], n[2], 4);
          if (iso) {
            double t_in = PRIM(k, n[0], n[1], n[2], 5);
            PRIM(k, g[0], g[1], g[2], 5) = 2.0 * t_wall - t_in;
          } else {
            PRIM(k, g[0], g[1], g[2], 5) = PRIM(k, n[0], n[1], n[2], 5);
          }
          double pg = PRIM(k, g[0], g[1], g[2], 4);
          double tg = PRIM(k, g[0], g[1], g[2], 5);
          if (tg < 1e-6) tg = 1.0;
          double rho_calc = moocp * pg / tg;
          PRIM(k, g[0], g[1], g[2], 0) = (d == 3) ? -rho_calc : rho_calc;
        }
        PRIM(k, i, j, kk, 0) = rho_wall;
        PRIM(k, i, j, kk, 1) = 0.0;
        PRIM(k, i, j, kk, 2) = 0.0;
        PRIM(k, i, j, kk, 3) = 0.0;
        PRIM(k, i, j, kk, 4) = p_wall;
        PRIM(k, i, j, kk, 5) = t_wall;
      }
}

static void bc_farfield(const ob_region* R, ob_block* k, const ob_params* p) {
  WIN;
  double gama = p->gama, gm1 = gama - 1.0;
  double rho_inf = p->roo, u_inf = p->uoo, v_inf = p->voo, w_inf = p->woo, p_inf = p->poo;
  double c_inf = sqrt(gama * p_inf / rho_inf);
  for (int kk = k_st; kk <= k_ed; ++kk)
    for (int j = j_st; j <= j_ed; ++j)
      for (int i = i_st; i <= i_ed; ++i) {
        double nx, ny, nz;
        get_normal(k, i, j, kk, R->s_nd, &nx, &ny, &nz);
        if (R->s_lr == -1) { nx = -nx; ny = -ny; nz = -nz; }
        int n[3] = {i, j, kk};
        if (R->s_lr == 1) n[dir] -= 1;
        else n[dir] += 1;
        double rho_in = PRIM(k, n[0], n[1], n[2], 0);
        double u_in = PRIM(k, n[0], n[1], n[2], 1);
        double v_in = PRIM(k, n[0], n[1], n[2], 2);
        double w_in = PRIM(k, n[0], n[1], n[2], 3);
        double p_in = PRIM(k, n[0], n[1], n[2], 4);
        double c_in = sqrt(gama * p_in / rho_in);
        double s_in = p_in / pow(rho_in, gama);
        double s_inf = p_inf / pow(rho_inf, gama);
        double vn_in = u_in * nx + v_in * ny + w_in * nz;
        double vn_inf = u_inf * nx + v_inf * ny + w_inf * nz;
        double vel = sqrt(u_in * u_in + v_in * v_in + w_in * w_in);
        double mach = vel / c_in;
        double rb, ub, vb, wb, pb;
        if (mach < 1.0) {
          double Rp = vn_in + 2.0 * c_in / gm1;
          double Rm = vn_inf - 2.0 * c_inf / gm1;
          double vn_b = 0.5 * (Rp + Rm);
          double c_b = 0.25 * gm1 * (Rp - Rm);
          double s_b, ur, vr, wr, vnr;
          if (vn_b > 0.0) { s_b = s_in; ur = u_in; vr = v_in; wr = w_in; vnr = vn_in; }
          else { s_b = s_inf; ur = u_inf; vr = v_inf; wr = w_inf; vnr = vn_inf; }
          rb = pow(c_b * c_b / (gama * s_b), 1.0 / gm1);
          pb = s_b * pow(rb, gama);
          ub = ur + (vn_b - vnr) * nx;
          vb = vr + (vn_b - vnr) * ny;
          wb = wr + (vn_b - vnr) * nz;
        } else {
          if (vn_in > 0.0) { rb = rho_in; ub = u_in; vb = v_in; wb = w_in; pb = p_in; }
          else { rb = rho_inf; ub = u_inf; vb = v_inf; wb = w_inf; pb = p_inf; }
        }
        double tb = (mach * mach * gama) * pb / rb;
        for (int l = 0; l < NG; ++l) {
          int d = l + 1, g[3] = {i, j, kk};
          if (R->s_lr == 1) g[dir] += d;
          else g[dir] -= d;
          PRIM(k, g[0], g[1], g[2], 0) = (d == 3) ? -rb : rb;
          PRIM(k, g[0], g[1], g[2], 1) = ub;
          PRIM(k, g[0], g[1], g[2], 2) = vb;
          PRIM(k, g[0], g[1], g[2], 3) = wb;
          PRIM(k, g[0], g[1], g[2], 4) = pb;
          PRIM(k, g[0], g[1], g[2], 5) = tb;
        }
      }
}

static void bc_freestream(const ob_region* R, ob_block* k, const ob_params* p) {
  WIN;
  double v[6] = {p->roo, p->uoo, p->voo, p->woo, p->poo, p->too};
  for (int kk = k_st; kk <= k_ed; ++kk)
    for (int j = j_st; j <= j_ed; ++j)
      for (int i = i_st; i <= i_ed; ++i)
        for (int l = 0; l < NG; ++l) {
          int d = l + 1, g[3] = {i, j, kk};
          if (R->s_lr == 1) g[dir] += d;
          else g[dir] -= d;
          for (int m = 0; m < 6; ++m) PRIM(k, g[0], g[1], g[2], m) = v[m];
        }
}

static void bc_extrapolation(const ob_region* R, ob_block* k) {
  WIN;
  for (int kk = k_st; kk <= k_ed; ++kk)
    for (int j = j_st; j <= j_ed; ++j)
      for (int i = i_st; i <= i_ed; ++i) {
        int s[3] = {i, j, kk};
        if (R->s_lr == 1) s[dir] -= 1;
        else s[dir] += 1;
        double v[6];
        for (int m = 0; m < 6; ++m) v[m] = PRIM(k, s[0], s[1], s[2], m);
        for (int l = 0; l < NG; ++l) {
          int d = l + 1, g[3] = {i, j, kk};
          if (R->s_lr == 1) g[dir] += d;
          else g[dir] -= d;
          for (int m = 0; m < 6; ++m) PRIM(k, g[0], g[1], g[2], m) = v[m];
        }
      }
}

int ob_apply_physical_bc(ob_case* c) {
  for (int b = 0; b < c->nb; ++b) {
    ob_block* k = &c->blocks[b];
    for (int r = 0; r < c->nreg[b]; ++r) {
      const ob_region* R = &c->reg[b][r];
      if (R->d.bctype < 0) continue;
      switch (R->d.bctype) {
        case 2:
          if (c->p.nvis == 1) bc_viscid_wall(R, k, &c->p);
          else bc_inviscid_wall(R, k, &c->p);
          break;
        case 3: bc_symmetry(R, k, &c->p); break;
        case 4: bc_farfield(R, k, &c->p); break;
        case 5: bc_freestream(R, k, &c->p); break;
        case 6: bc_extrapolation(R, k); break;
        default: break;
      }
    }
  }
  return 0;
}

/* ========================================================================
 * Flow states (src/solver/StateUpdater.cpp:15-168)
 * ======================================================================== */

static void thermo_point(ob_block* k, long id, double gamma, double moo2, int nvis, int ghost) {
  double rho = k->prim[id];
  double p = k->prim[4 * k->P + id];
  if (ghost) rho = fabs(rho);
  double a2 = gamma * p / rho;
  k->c[id] = sqrt(fabs(a2));
  if (nvis == 1) k->prim[5 * k->P + id] = moo2 * a2;
}

static void cons_point(ob_block* k, long id, double gm1_inv, int use_abs) {
  double rho = k->prim[id], u = k->prim[k->P + id], v = k->prim[2 * k->P + id];
  double w = k->prim[3 * k->P + id], p = k->prim[4 * k->P + id];
  double rc = (use_abs || rho < 0.0) ? fabs(rho) : rho;
  k->q[id] = rc;
  k->q[k->P + id] = rc * u;
  k->q[2 * k->P + id] = rc * v;
  k->q[3 * k->P + id] = rc * w;
  double ke = 0.5 * rc * (u * u + v * v + w * w);
  k->q[4 * k->P + id] = p * gm1_inv + ke;
}

/* Interior plus the ng ghost layers of each face, tangentially interior
 * (StateUpdater.cpp:29-72, 126-148).  kind 0: thermodynamics, 1: q. */
static void cross_layers(ob_block* k, int kind, double a, double b, int nvis) {
  int dims[3] = {k->idim, k->jdim, k->kdim};
  int is = NG, ie = NG + k->idim - 1, js = NG, je = NG + k->jdim - 1, ks = NG, ke = NG + k->kdim - 1;
  for (int kk = ks; kk <= ke; ++kk)
    for (int j = js; j <= je; ++j)
      for (int i = is; i <= ie; ++i) {
        if (kind == 0) thermo_point(k, IX(k, i, j, kk), a, b, nvis, 0);
        else cons_point(k, IX(k, i, j, kk), a, 0);
      }
  for (int dir = 0; dir < 3; ++dir)
    for (int l = 1; l <= NG; ++l) {
      int faces[2] = {NG - l, NG + dims[dir] - 1 + l};
      for (int f = 0; f < 2; ++f) {
        int fi = faces[f];
        int k0 = dir == 2 ? fi : ks, k1 = dir == 2 ? fi : ke;
        int j0 = dir == 1 ? fi : js, j1 = dir == 1 ? fi : je;
        int i0 = dir == 0 ? fi : is, i1 = dir == 0 ? fi : ie;
        for (int kk = k0; kk <= k1; ++kk)
          for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) {
              if (kind == 0) thermo_point(k, IX(k, i, j, kk), a, b, nvis, 1);
              else cons_point(k, IX(k, i, j, kk), a, 1);
            }
      }
    }
}

void ob_update_flow_states(ob_case* c) {
  const ob_params* p = &c->p;
  double moo2 = p->moo * p->moo;
  double gm1_inv = 1.0 / (p->gama - 1.0);
#pragma omp parallel for schedule(dynamic) num_threads(nthreads())
  for (int b = 0; b < c->nb; ++b) {
    ob_block* k = &c->blocks[b];
    cross_layers(k, 0, p->gama, moo2, p->nvis);
    if (p->nvis == 1)
      for (long id = 0; id < k->P; ++id) {
        double tm = k->prim[5 * k->P + id];
        if (tm < 1e-6) tm = 1e-6;
        k->mu[id] = tm * sqrt(tm) * (1.0 + p->visc) / (tm + p->visc);
      }
    cross_layers(k, 1, gm1_inv, 0.0, 0);
    memset(k->dq, 0, sizeof(double) * 5 * k->P);
  }
}

/* ========================================================================
 * Time step (src/solver/TimeIntegrator.cpp:20-178)
 * ======================================================================== */

void ob_calculate_time_step(ob_case* c) {
  const ob_params* p = &c->p;
  double reynolds = p->reynolds, csrv = (double)p->csrv, cfl = p->cfl;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads())
  for (int b = 0; b < c->nb; ++b) {
    ob_block* k = &c->blocks[b];
    long P = k->P;
    for (int dir = 0; dir < 3; ++dir)
      for (long id = 0; id < P; ++id) {
        double u = k->prim[P + id], v = k->prim[2 * P + id], w = k->prim[3 * P + id];
        double cc = k->c[id];
        double mx = k->metrics[(3 * dir) * P + id], my = k->metrics[(3 * dir + 1) * P + id];
        double mz = k->metrics[(3 * dir + 2) * P + id];
        double U = fabs(u * mx + v * my + w * mz);
        double gl = sqrt(mx * mx + my * my + mz * mz);
        k->spec[dir * P + id] = U + cc * gl;
      }
    if (p->nvis == 1) {
      for (int dir = 0; dir < 3; ++dir)
        for (long id = 0; id < P; ++id) {
          double mx = k->metrics[(3 * dir) * P + id], my = k->metrics[(3 * dir + 1) * P + id];
          double mz = k->metrics[(3 * dir + 2) * P + id];
          double g2 = mx * mx + my * my + mz * mz;
          double rho = k->prim[id], vol = k->vol[id], mu = k->mu[id];
          double denom = reynolds * rho * vol + 1.0e-30;
          k->spec_visc[dir * P + id] = (2.0 * mu / denom) * g2 * csrv;
        }
    } else {
      memset(k->spec_visc, 0, sizeof(double) * 3 * P);
    }
    for (int kk = NG; kk < k->nz - NG; ++kk)
      for (int j = NG; j < k->ny - NG; ++j)
        for (int i = NG; i < k->nx - NG; ++i) {
          long id = IX(k, i, j, kk);
          double sum = 0.0;
          for (int dir = 0; dir < 3; ++dir) sum += k->spec[dir * P + id] + k->spec_visc[dir * P + id];
          double dt_phys = cfl / (sum + 1.0e-30);
          k->dt[id] = dt_phys / k->vol[id];
        }
  }
  if (p->ntmst != 1) {
    double dtmin = 1000.0;
    for (int b = 0; b < c->nb; ++b) {
      ob_block* k = &c->blocks[b];
      for (int kk = NG; kk < k->nz - NG; ++kk)
        for (int j = NG; j < k->ny - NG; ++j)
          for (int i = NG; i < k->nx - NG; ++i) {
            long id = IX(k, i, j, kk);
            double d = k->dt[id] * k->vol[id];
            if (d < dtmin) dtmin = d;
          }
    }
    c->p.phydtime = dtmin;
  }
}

/* ========================================================================
 * Shared line stencils (FluxComputer.cpp:13-296, InviscidFluxComputer.cpp:136-233)
 * Line arrays use the reference's convention: position p <-> padded index p.
 * ======================================================================== */

/* compute_node_deriv_line, FluxComputer.cpp:13-67 */
static void node_deriv(int n, int ni, const double* q, double* dq, int fl, int fr) {
  int st = 1, ed = ni;
  if (fl < 0) {
    st = 3;
    for (int m = 0; m < n; ++m) {
      double q1 = q[3 * n + m], q2 = q[4 * n + m], q3 = q[5 * n + m], q4 = q[6 * n + m], q5 = q[7 * n + m];
      dq[0 * n + m] = (-11.0 * q1 + 18.0 * q2 - 9.0 * q3 + 2.0 * q4) / 6.0;
      dq[1 * n + m] = (-3.0 * q1 + -10.0 * q2 + 18.0 * q3 + -6.0 * q4 + 1.0 * q5) / 12.0;
    }
  }
  if (fr < 0) {
    ed = ni - 2;
    for (int m = 0; m < n; ++m) {
      double qn = q[(ni + 2) * n + m], qn1 = q[(ni + 1) * n + m], qn2 = q[ni * n + m];
      double qn3 = q[(ni - 1) * n + m], qn4 = q[(ni - 2) * n + m];
      dq[(ni - 1) * n + m] = -(-11.0 * qn + 18.0 * qn1 - 9.0 * qn2 + 2.0 * qn3) / 6.0;
      dq[(ni - 2) * n + m] = -(-3.0 * qn + -10.0 * qn1 + 18.0 * qn2 + -6.0 * qn3 + 1.0 * qn4) / 12.0;
    }
  }
  for (int i = st; i <= ed; ++i)
    for (int m = 0; m < n; ++m) {
      double qp1 = q[(i + 3) * n + m], qm1 = q[(i + 1) * n + m];
      double qp2 = q[(i + 4) * n + m], qm2 = q[(i + 0) * n + m];
      dq[(i - 1) * n + m] = (8.0 * (qp1 - qm1) + -1.0 * (qp2 - qm2)) / 12.0;
    }
}

/* compute_half_deriv_line, FluxComputer.cpp:72-124 */
static void half_deriv(int n, int ni, const double* q, double* dq, int fl, int fr) {
  int st = 0, ed = ni;
  if (fl < 0) {
    st = 2;
    for (int m = 0; m < n; ++m) {
      double q1 = q[3 * n + m], q2 = q[4 * n + m], q3 = q[5 * n + m], q4 = q[6 * n + m], q5 = q[7 * n + m];
      dq[0 * n + m] = (-71.0 * q1 + 141.0 * q2 + -93.0 * q3 + 23.0 * q4) / 24.0;
      dq[1 * n + m] = (-22.0 * q1 + 17.0 * q2 + 9.0 * q3 + -5.0 * q4 + 1.0 * q5) / 24.0;
    }
  }
  if (fr < 0) {
    ed = ni - 2;
    for (int m = 0; m < n; ++m) {
      double qn = q[(ni + 2) * n + m], qn1 = q[(ni + 1) * n + m], qn2 = q[ni * n + m];
      double qn3 = q[(ni - 1) * n + m], qn4 = q[(ni - 2) * n + m];
      dq[ni * n + m] = -(-71.0 * qn + 141.0 * qn1 + -93.0 * qn2 + 23.0 * qn3) / 24.0;
      dq[(ni - 1) * n + m] = -(-22.0 * qn + 17.0 * qn1 + 9.0 * qn2 + -5.0 * qn3 + 1.0 * qn4) / 24.0;
    }
  }
  for (int i = st; i <= ed; ++i)
    for (int m = 0; m < n; ++m) {
      double qi = q[(i + 2) * n + m], qip1 = q[(i + 3) * n + m];
      double qip2 = q[(i + 4) * n + m], qim1 = q[(i + 1) * n + m];
      dq[i * n + m] = (27.0 * (qip1 - qi) + -1.0 * (qip2 - qim1)) / 24.0;
    }
}

/* interp_line_to_half, FluxComputer.cpp:129-178 (flags 0,0 equal
 * interp_metrics_half, InviscidFluxComputer.cpp:136-192, for ni>2) */
static void half_interp(int n, int ni, const double* q, double* val, int fl, int fr) {
  int st = 0, ed = ni;
  if (fl < 0) {
    st = 2;
    for (int m = 0; m < n; ++m) {
      double q1 = q[3 * n + m], q2 = q[4 * n + m], q3 = q[5 * n + m], q4 = q[6 * n + m];
      val[0 * n + m] = (35.0 * q1 + -35.0 * q2 + 21.0 * q3 + -5.0 * q4) / 16.0;
      val[1 * n + m] = (5.0 * q1 + 15.0 * q2 + -5.0 * q3 + 1.0 * q4) / 16.0;
    }
  }
  if (fr < 0) {
    ed = ni - 2;
    for (int m = 0; m < n; ++m) {
      double qn = q[(ni + 2) * n + m], qn1 = q[(ni + 1) * n + m], qn2 = q[ni * n + m];
      double qn3 = q[(ni - 1) * n + m];
      val[ni * n + m] = (35.0 * qn + -35.0 * qn1 + 21.0 * qn2 + -5.0 * qn3) / 16.0;
      val[(ni - 1) * n + m] = (5.0 * qn + 15.0 * qn1 + -5.0 * qn2 + 1.0 * qn3) / 16.0;
    }
  }
  for (int i = st; i <= ed; ++i)
    for (int m = 0; m < n; ++m) {
      double qi = q[(i + 2) * n + m], qip1 = q[(i + 3) * n + m];
      double qip2 = q[(i + 4) * n + m], qim1 = q[(i + 1) * n + m];
      val[i * n + m] = (9.0 * (qi + qip1) + -1.0 * (qip2 + qim1)) / 16.0;
    }
}

/* interp_metrics_half, InviscidFluxComputer.cpp:136-192 */
static void metric_half(int n, int ni, const double* q, double* val) {
  for (int m = 0; m < n; ++m) {
    double q1 = q[3 * n + m], q2 = q[4 * n + m], q3 = q[5 * n + m], q4 = q[6 * n + m];
    val[0 * n + m] = (35.0 * q1 + -35.0 * q2 + 21.0 * q3 + -5.0 * q4) / 16.0;
    val[1 * n + m] = (5.0 * q1 + 15.0 * q2 + -5.0 * q3 + 1.0 * q4) / 16.0;
  }
  for (int m = 0; m < n; ++m) {
    double qn = q[(ni + 2) * n + m], qn1 = q[(ni + 1) * n + m], qn2 = q[ni * n + m], qn3 = q[(ni - 1) * n + m];
    val[ni * n + m] = (35.0 * qn + -35.0 * qn1 + 21.0 * qn2 + -5.0 * qn3) / 16.0;
    val[(ni - 1) * n + m] = (5.0 * qn + 15.0 * qn1 + -5.0 * qn2 + 1.0 * qn3) / 16.0;
  }
  for (int i = 2; i <= ni - 2; ++i)
    for (int m = 0; m < n; ++m) {
      double qi = q[(i + 2) * n + m], qip1 = q[(i + 3) * n + m];
      double qip2 = q[(i + 4) * n + m], qim1 = q[(i + 1) * n + m];
      val[i * n + m] = (9.0 * (qi + qip1) + -1.0 * (qip2 + qim1)) / 16.0;
    }
}

/* compute_flux_diff_high_order / compute_flux_diff_1d (identical), nl=5 */
static void flux_diff(int ni, const double* f, double* df) {
  const int nl = 5;
  for (int i = 2; i < ni - 2; ++i)
    for (int m = 0; m < nl; ++m) {
      double t1 = 2250.0 * (f[(i + 1) * nl + m] - f[i * nl + m]);
      double t2 = -125.0 * (f[(i + 2) * nl + m] - f[(i - 1) * nl + m]);
      double t3 = 9.0 * (f[(i + 3) * nl + m] - f[(i - 2) * nl + m]);
      df[i * nl + m] = (t1 + t2 + t3) / 1920.0;
    }
  for (int m = 0; m < nl; ++m) {
    df[1 * nl + m] = (f[0 * nl + m] - 27.0 * f[1 * nl + m] + 27.0 * f[2 * nl + m] - f[3 * nl + m]) / 24.0;
    df[(ni - 2) * nl + m] =
        -(f[ni * nl + m] - 27.0 * f[(ni - 1) * nl + m] + 27.0 * f[(ni - 2) * nl + m] - f[(ni - 3) * nl + m]) / 24.0;
    df[0 * nl + m] = (-23.0 * f[0 * nl + m] + 21.0 * f[1 * nl + m] + 3.0 * f[2 * nl + m] - f[3 * nl + m]) / 24.0;
    df[(ni - 1) * nl + m] =
        -(-23.0 * f[ni * nl + m] + 21.0 * f[(ni - 1) * nl + m] + 3.0 * f[(ni - 2) * nl + m] - f[(ni - 3) * nl + m]) /
        24.0;
  }
}

/* Line addressing: direction d, the two other padded coords (a,b). */
typedef struct {
  long base, stride;
  int n; /* interior count along d */
} ob_line;

static ob_line line_of(const ob_block* k, int d, int a, int b) {
  ob_line L;
  if (d == 0) { L.base = IX(k, 0, a, b); L.stride = 1; L.n = k->idim; }
  else if (d == 1) { L.base = IX(k, a, 0, b); L.stride = k->nx; L.n = k->jdim; }
  else { L.base = IX(k, a, b, 0); L.stride = (long)k->nx * k->ny; L.n = k->kdim; }
  return L;
}

/* ========================================================================
 * Viscous RHS (src/solver/FluxComputer.cpp:301-732)
 * ======================================================================== */

static void viscous_block(ob_block* k, const ob_params* p) {
  const long P = k->P;
  const int nmax = (k->idim > k->jdim ? (k->idim > k->kdim ? k->idim : k->kdim) : (k->jdim > k->kdim ? k->jdim : k->kdim));
  double* duvwt = calloc(12 * P, sizeof(double));
  double* dmid = calloc(12 * P, sizeof(double));
  double* ql = calloc((nmax + 6) * 9, sizeof(double));
  double* dl = calloc((nmax + 1) * 4, sizeof(double));
  double* uvwt_h = calloc(4 * (nmax + 1), sizeof(double));
  double* duv_h = calloc(12 * (nmax + 1), sizeof(double));
  double* kx_h = calloc(9 * (nmax + 1), sizeof(double));
  double* vol_h = calloc(nmax + 1, sizeof(double));
  double* vs1 = calloc(nmax + 1, sizeof(double));
  double* vs2 = calloc(nmax + 1, sizeof(double));
  double* fv = calloc(5 * (nmax + 1), sizeof(double));
  double* dfv = calloc(5 * nmax, sizeof(double));
  double* cc8 = calloc(8 * (nmax + 1), sizeof(double));
  double* dxyz = calloc(12 * (nmax + 1), sizeof(double));
  const double re = p->reynolds;
  const double re_inv = (fabs(re) > 1.0e-30) ? 1.0 / re : 0.0;
  const double cp = 1.0 / ((p->gama - 1.0) * p->moo * p->moo);
  const double cp_prl = cp / p->prl;
  const int na[3] = {k->ny, k->nx, k->nx}, nbn[3] = {k->nz, k->nz, k->ny};
  /* stage 1: cell and face derivatives per direction */
  for (int d = 0; d < 3; ++d)
    for (int bb = NG; bb < nbn[d] - NG; ++bb)
      for (int aa = NG; aa < na[d] - NG; ++aa) {
        ob_line L = line_of(k, d, aa, bb);
        int n = L.n;
        int fl = (k->prim[L.base] < 0) ? -1 : 0;
        int fr = (k->prim[L.base + (long)(n + 5) * L.stride] < 0) ? -1 : 0;
        for (int ps = 1; ps <= n + 5; ++ps) {
          long id = L.base + (long)ps * L.stride;
          ql[ps * 4 + 0] = k->prim[P + id];
          ql[ps * 4 + 1] = k->prim[2 * P + id];
          ql[ps * 4 + 2] = k->prim[3 * P + id];
          ql[ps * 4 + 3] = k->prim[5 * P + id];
        }
        node_deriv(4, n, ql, dl, fl, fr);
        for (int i = 0; i < n; ++i)
          for (int m = 0; m < 4; ++m) duvwt[(4 * d + m) * P + L.base + (long)(NG + i) * L.stride] = dl[i * 4 + m];
        half_deriv(4, n, ql, dl, fl, fr);
        for (int i = 0; i <= n; ++i)
          for (int m = 0; m < 4; ++m) dmid[(4 * d + m) * P + L.base + (long)(NG + i - 1) * L.stride] = dl[i * 4 + m];
      }
  /* stage 2: face fluxes and the 6th-order difference per direction */
  for (int d = 0; d < 3; ++d) {
    /* cross-derivative components carried to the face: the two other dirs */
    int oth[2] = {d == 0 ? 1 : 0, d == 2 ? 1 : 2};
    for (int bb = NG; bb < nbn[d] - NG; ++bb)
      for (int aa = NG; aa < na[d] - NG; ++aa) {
        ob_line L = line_of(k, d, aa, bb);
        int n = L.n;
        int fl = (k->prim[L.base] < 0) ? -1 : 0;
        int fr = (k->prim[L.base + (long)(n + 5) * L.stride] < 0) ? -1 : 0;
        for (int ps = 1; ps <= n + 5; ++ps) {
          long id = L.base + (long)ps * L.stride;
          ql[ps * 4 + 0] = k->prim[P + id];
          ql[ps * 4 + 1] = k->prim[2 * P + id];
          ql[ps * 4 + 2] = k->prim[3 * P + id];
          ql[ps * 4 + 3] = k->prim[5 * P + id];
        }
        half_interp(4, n, ql, uvwt_h, fl, fr);
        for (int ps = 1; ps <= n + 5; ++ps) ql[ps] = k->mu[L.base + (long)ps * L.stride];
        half_interp(1, n, ql, vs1, 0, 0);
        for (int i = 0; i <= n; ++i) vs2[i] = vs1[i] * cp_prl;
        for (int ps = 1; ps <= n + 5; ++ps) ql[ps] = k->vol[L.base + (long)ps * L.stride];
        half_interp(1, n, ql, vol_h, 0, 0);
        for (int ps = 1; ps <= n + 5; ++ps)
          for (int m = 0; m < 9; ++m) ql[ps * 9 + m] = k->metrics[m * P + L.base + (long)ps * L.stride];
        half_interp(9, n, ql, kx_h, 0, 0);
        for (int ps = 1; ps <= n + 5; ++ps) {
          int src = ps < NG ? NG : (ps > n + NG - 1 ? n + NG - 1 : ps);
          long id = L.base + (long)src * L.stride;
          for (int m = 0; m < 4; ++m) ql[ps * 8 + m] = duvwt[(4 * oth[0] + m) * P + id];
          for (int m = 0; m < 4; ++m) ql[ps * 8 + 4 + m] = duvwt[(4 * oth[1] + m) * P + id];
        }
        half_interp(8, n, ql, cc8, 0, 0);
        for (int i = 0; i <= n; ++i) {
          long mid = L.base + (long)(NG + i - 1) * L.stride;
          for (int m = 0; m < 4; ++m) duv_h[i * 12 + 4 * d + m] = dmid[(4 * d + m) * P + mid];
          for (int m = 0; m < 4; ++m) duv_h[i * 12 + 4 * oth[0] + m] = cc8[i * 8 + m];
          for (int m = 0; m < 4; ++m) duv_h[i * 12 + 4 * oth[1] + m] = cc8[i * 8 + 4 + m];
        }
        /* transform_derivs_to_phys (:183-205) */
        for (int i = 0; i <= n; ++i) {
          double vinv = 1.0 / vol_h[i];
          for (int m = 0; m < 3; ++m) {
            double k1 = kx_h[i * 9 + m], k2 = kx_h[i * 9 + m + 3], k3 = kx_h[i * 9 + m + 6];
            for (int v = 0; v < 4; ++v) {
              double dx = duv_h[i * 12 + v], de = duv_h[i * 12 + v + 4], dz = duv_h[i * 12 + v + 8];
              dxyz[i * 12 + v * 3 + m] = (k1 * dx + k2 * de + k3 * dz) * vinv;
            }
          }
        }
        /* compute_viscous_flux_1d (:210-252) */
        for (int i = 0; i <= n; ++i) {
          double vs = vs1[i], kcp = vs2[i], vscc = vs * (2.0 / 3.0);
          const double* g = &dxyz[i * 12];
          double dudx = g[0], dudy = g[1], dudz = g[2];
          double dvdx = g[3], dvdy = g[4], dvdz = g[5];
          double dwdx = g[6], dwdy = g[7], dwdz = g[8];
          double txx = vscc * (2.0 * dudx - dvdy - dwdz);
          double tyy = vscc * (2.0 * dvdy - dwdz - dudx);
          double tzz = vscc * (2.0 * dwdz - dudx - dvdy);
          double txy = vs * (dudy + dvdx);
          double txz = vs * (dudz + dwdx);
          double tyz = vs * (dvdz + dwdy);
          double nx = kx_h[i * 9 + 3 * d], ny = kx_h[i * 9 + 3 * d + 1], nz = kx_h[i * 9 + 3 * d + 2];
          fv[i * 5 + 0] = 0.0;
          fv[i * 5 + 1] = txx * nx + txy * ny + txz * nz;
          fv[i * 5 + 2] = txy * nx + tyy * ny + tyz * nz;
          fv[i * 5 + 3] = txz * nx + tyz * ny + tzz * nz;
          double u = uvwt_h[i * 4], v = uvwt_h[i * 4 + 1], w = uvwt_h[i * 4 + 2];
          double work = u * fv[i * 5 + 1] + v * fv[i * 5 + 2] + w * fv[i * 5 + 3];
          double heat = kcp * (g[9] * nx + g[10] * ny + g[11] * nz);
          fv[i * 5 + 4] = work + heat;
        }
        flux_diff(n, fv, dfv);
        for (int i = 0; i < n; ++i) {
          long id = L.base + (long)(NG + i) * L.stride;
          for (int m = 0; m < 5; ++m) k->dq[m * P + id] -= re_inv * dfv[i * 5 + m];
        }
      }
  }
  free(duvwt); free(dmid); free(ql); free(dl); free(uvwt_h); free(duv_h); free(kx_h);
  free(vol_h); free(vs1); free(vs2); free(fv); free(dfv); free(cc8); free(dxyz);
}

void ob_compute_viscous_rhs(ob_case* c) {
  if (c->p.nvis != 1) return;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads())
  for (int b = 0; b < c->nb; ++b) viscous_block(&c->blocks[b], &c->p);
}

/* ========================================================================
 * Interface residual averaging (src/solver/HaloExchanger.cpp:186-398)
 * ======================================================================== */

int ob_average_interface_residuals(ob_case* c) {
  /* phase 1: every connect region packs dq/vol of its face window into the
   * partner region's buffer */
  for (int b = 0; b < c->nb; ++b) {
    ob_block* k = &c->blocks[b];
    for (int r = 0; r < c->nreg[b]; ++r) {
      ob_region* R = &c->reg[b][r];
      if (R->d.bctype >= 0) continue;
      int lo[3], hi[3];
      for (int d = 0; d < 3; ++d) {
        int a = abs(R->d.s_st[d]), bb = abs(R->d.s_ed[d]);
        lo[d] = a < bb ? a : bb;
        hi[d] = a < bb ? bb : a;
      }
      int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
      double* buf = malloc(sizeof(double) * 5 * (long)nx * ny * nz);
      long p = 0;
      for (int kk = 0; kk < nz; ++kk)
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i) {
            int I = lo[0] - 1 + NG + i, J = lo[1] - 1 + NG + j, K = lo[2] - 1 + NG + kk;
            double vol = S1(k->vol, k, I, J, K);
            for (int m = 0; m < 5; ++m) buf[p++] = DQ(k, I, J, K, m) / vol;
          }
      ob_region* T = &c->reg[R->d.nbt - 1][R->d.ibcwin - 1];
      free(T->qpvpack);
      T->qpvpack = buf;
      T->qpv_len = p;
    }
  }
  /* phase 3: average through the image/jmage/kmage maps */
  for (int b = 0; b < c->nb; ++b) {
    ob_block* k = &c->blocks[b];
    for (int r = 0; r < c->nreg[b]; ++r) {
      ob_region* R = &c->reg[b][r];
      if (R->d.bctype >= 0) continue;
      if (!R->qpvpack) return fail("interface buffer missing for block %d window %d", b + 1, r + 1);
      int lo[3], hi[3], tlo[3], thi[3];
      for (int d = 0; d < 3; ++d) {
        int a = abs(R->d.s_st[d]), bb = abs(R->d.s_ed[d]);
        lo[d] = a < bb ? a : bb;
        hi[d] = a < bb ? bb : a;
        a = abs(R->d.t_st[d]);
        bb = abs(R->d.t_ed[d]);
        tlo[d] = a < bb ? a : bb;
        thi[d] = a < bb ? bb : a;
      }
      int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
      int nxt = thi[0] - tlo[0] + 1, nyt = thi[1] - tlo[1] + 1;
      for (int kk = 0; kk < nz; ++kk)
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i) {
            int I = lo[0] - 1 + NG + i, J = lo[1] - 1 + NG + j, K = lo[2] - 1 + NG + kk;
            long f = (long)kk * (nx * ny) + (long)j * nx + i;
            int ir = R->image[f] - tlo[0], jr = R->jmage[f] - tlo[1], kr = R->kmage[f] - tlo[2];
            long rem = (long)kr * (nxt * nyt) + (long)jr * nxt + ir;
            if (rem < 0 || 5 * rem + 4 >= R->qpv_len)
              return fail("interface map out of range (block %d window %d)", b + 1, r + 1);
            double vs = S1(k->vol, k, I, J, K);
            for (int m = 0; m < 5; ++m) {
              double dt_ = R->qpvpack[rem * 5 + m];
              double ds = DQ(k, I, J, K, m);
              DQ(k, I, J, K, m) = 0.5 * (dt_ * vs + ds);
            }
          }
    }
  }
  return 0;
}

/* ========================================================================
 * Inviscid RHS (src/solver/InviscidFluxComputer.cpp:14-647)
 * ======================================================================== */

static void phys_flux(const double* q, double g, double nx, double ny, double nz, double nt, double* f) {
  double rho = q[0], u = q[1], v = q[2], w = q[3], p = q[4];
  double V = u * nx + v * ny + w * nz + nt;
  double E = p / (g - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
  double H = (E + p) / rho;
  f[0] = rho * V;
  f[1] = rho * u * V + p * nx;
  f[2] = rho * v * V + p * ny;
  f[3] = rho * w * V + p * nz;
  f[4] = rho * H * V + p * nt;
}

static void roe(const double* ql, const double* qr, double nx, double ny, double nz, double nt, double efix,
                double gamma, double* f) {
  double fl[5], fr[5], dq[5], df[5];
  phys_flux(ql, gamma, nx, ny, nz, nt, fl);
  phys_flux(qr, gamma, nx, ny, nz, nt, fr);
  for (int m = 0; m < 5; ++m) dq[m] = qr[m] - ql[m];
  double g1 = gamma - 1.0, g2 = gamma / g1;
  double hl = g2 * ql[4] / ql[0] + 0.5 * (ql[1] * ql[1] + ql[2] * ql[2] + ql[3] * ql[3]);
  double hr = g2 * qr[4] / qr[0] + 0.5 * (qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
  double sL = sqrt(ql[0]), sR = sqrt(qr[0]);
  double inv_den = 1.0 / (sL + sR);
  double rr = sL * sR;
  double ur = (sL * ql[1] + sR * qr[1]) * inv_den;
  double vr = (sL * ql[2] + sR * qr[2]) * inv_den;
  double wr = (sL * ql[3] + sR * qr[3]) * inv_den;
  double hh = (sL * hl + sR * hr) * inv_den;
  double v2 = ur * ur + vr * vr + wr * wr;
  double c2 = (hh - 0.5 * v2) * g1;
  if (c2 <= 0.0) c2 = 1.0e-6;
  double c = sqrt(c2);
  double U = ur * nx + vr * ny + wr * nz + nt;
  double gm = sqrt(nx * nx + ny * ny + nz * nz);
  gm = dmax(gm, 1.0e-30);
  double im = 1.0 / gm;
  double nxn = nx * im, nyn = ny * im, nzn = nz * im;
  double cm = c * gm;
  double l1 = fabs(U - cm), l4 = fabs(U), l5 = fabs(U + cm);
  double de = efix * cm, de2 = de * de;
  if (l1 < de) l1 = sqrt(l1 * l1 + de2);
  if (l4 < de) l4 = sqrt(l4 * l4 + de2);
  if (l5 < de) l5 = sqrt(l5 * l5 + de2);
  double du = dq[1], dv = dq[2], dw = dq[3], dp = dq[4], drho = dq[0];
  double rodcta = rr * (du * nxn + dv * nyn + dw * nzn);
  double dp_c2 = dp / c2;
  double rodctac = rodcta / c;
  double a1 = l4 * (drho - dp_c2);
  double a2 = l5 * (dp_c2 + rodctac) * 0.5;
  double a3 = l1 * (dp_c2 - rodctac) * 0.5;
  double a4 = a1 + a2 + a3;
  double a5 = c * (a2 - a3);
  double a6 = l4 * (rr * du - nxn * rodcta);
  double a7 = l4 * (rr * dv - nyn * rodcta);
  double a8 = l4 * (rr * dw - nzn * rodcta);
  df[0] = a4;
  df[1] = ur * a4 + nxn * a5 + a6;
  df[2] = vr * a4 + nyn * a5 + a7;
  df[3] = wr * a4 + nzn * a5 + a8;
  df[4] = hh * a4 + U * im * a5 + ur * a6 + vr * a7 + wr * a8 - c2 * a1 / g1;
  for (int m = 0; m < 5; ++m) f[m] = 0.5 * (fl[m] + fr[m] - df[m]);
}

/* compute_flux_line, InviscidFluxComputer.cpp:238-522.  q: (ni+6)*6 line
 * of prim, mh: (ni+1)*5 half-node metrics. */
static void flux_line(int ni, const double* q, const double* mh, double* fc, double gamma, double* qwl, double* qwr,
                      double* f) {
  const int nl = 5;
  const double efix = 0.1, small = 1.0e-20;
  const double CL1 = 1.0 / 16.0, CL2 = 10.0 / 16.0, CL3 = 5.0 / 16.0, EPS = 1.0e-6;
  double ulo[9 * 5], uro[9 * 5];
  int ist = 0, ied = ni;
  if (q[0 * 6 + 0] < small) ist = 2;
  if (q[(ni + 5) * 6 + 0] < small) ied = ni - 2;
#define W(g, s, C) ((C) / ((EPS + ((g) * (g) + (s) * (s))) * (EPS + ((g) * (g) + (s) * (s)))))
  for (int m = 0; m < nl; ++m)
    for (int i = ist; i <= ied; ++i) {
      double qm2 = q[i * 6 + m], qm1 = q[(i + 1) * 6 + m], q0 = q[(i + 2) * 6 + m];
      double qp1 = q[(i + 3) * 6 + m], qp2 = q[(i + 4) * 6 + m], qp3 = q[(i + 5) * 6 + m];
      double s1 = qm2 - 2.0 * qm1 + q0;
      double s2 = qm1 - 2.0 * q0 + qp1;
      double s3 = q0 - 2.0 * qp1 + qp2;
      double g1 = 0.5 * (qm2 - 4.0 * qm1 + 3.0 * q0);
      double g2 = 0.5 * (qp1 - qm1);
      double g3 = 0.5 * (-3.0 * q0 + 4.0 * qp1 - qp2);
      double w1 = W(g1, s1, CL1), w2 = W(g2, s2, CL2), w3 = W(g3, s3, CL3);
      double sw = w1 + w2 + w3;
      qwl[i * nl + m] = q0 + 0.125 * ((w1 / sw) * (s1 + 4.0 * g1) + (w2 / sw) * (s2 + 4.0 * g2) + (w3 / sw) * (s3 + 4.0 * g3));
      double sr1 = s2, sr2 = s3, sr3 = qp1 - 2.0 * qp2 + qp3;
      double gr1 = 0.5 * (qm1 - 4.0 * q0 + 3.0 * qp1);
      double gr2 = 0.5 * (qp2 - q0);
      double gr3 = 0.5 * (-3.0 * qp1 + 4.0 * qp2 - qp3);
      double r1 = W(gr1, sr1, CL3), r2 = W(gr2, sr2, CL2), r3 = W(gr3, sr3, CL1);
      double sr = r1 + r2 + r3;
      qwr[i * nl + m] =
          qp1 + 0.125 * ((r1 / sr) * (sr1 - 4.0 * gr1) + (r2 / sr) * (sr2 - 4.0 * gr2) + (r3 / sr) * (sr3 - 4.0 * gr3));
    }
#undef W
  if (ist > 1)
    for (int m = 0; m < nl; ++m) {
      double q1 = q[3 * 6 + m], q2 = q[4 * 6 + m], q3 = q[5 * 6 + m], q4 = q[6 * 6 + m];
      qwl[0 * nl + m] = (35. * q1 - 35. * q2 + 21. * q3 - 5. * q4) / 16.;
      qwl[1 * nl + m] = (5. * q1 + 15. * q2 - 5. * q3 + q4) / 16.;
      qwl[2 * nl + m] = (-q1 + 9. * q2 + 9. * q3 - q4) / 16.;
      qwr[0 * nl + m] = qwl[0 * nl + m];
      qwr[1 * nl + m] = qwl[1 * nl + m];
    }
  if (ied < ni)
    for (int m = 0; m < nl; ++m) {
      double qn = q[(ni + 2) * 6 + m], qn1 = q[(ni + 1) * 6 + m], qn2 = q[ni * 6 + m], qn3 = q[(ni - 1) * 6 + m];
      qwr[ni * nl + m] = (35. * qn - 35. * qn1 + 21. * qn2 - 5. * qn3) / 16.;
      qwr[(ni - 1) * nl + m] = (5. * qn + 15. * qn1 - 5. * qn2 + qn3) / 16.;
      qwr[(ni - 2) * nl + m] = (-qn + 9. * qn1 + 9. * qn2 - qn3) / 16.;
      qwl[ni * nl + m] = qwr[ni * nl + m];
      qwl[(ni - 1) * nl + m] = qwr[(ni - 1) * nl + m];
    }
  /* calc_flux_loop (:342-364) */
#define FLUXES(st, en)                                                        \
  for (int i = (st); i <= (en); ++i) {                                        \
    double a_[5], b_[5], o_[5];                                               \
    for (int m = 0; m < nl; ++m) { a_[m] = qwl[i * nl + m]; b_[m] = qwr[i * nl + m]; } \
    if (a_[0] <= 1e-6 || a_[4] <= 1e-6) {                                     \
      int ix = (i - 1 > 0 ? i - 1 : 0) + 2;                                   \
      for (int m = 0; m < nl; ++m) a_[m] = q[ix * 6 + m];                     \
      guard_hit(0);                                                           \
    }                                                                         \
    if (b_[0] <= 1e-6 || b_[4] <= 1e-6) {                                     \
      int ix = (ni - 1 < i ? ni - 1 : i) + 2 + 1;                             \
      for (int m = 0; m < nl; ++m) b_[m] = q[ix * 6 + m];                     \
      guard_hit(1);                                                           \
    }                                                                         \
    roe(a_, b_, mh[i * 5 + 0], mh[i * 5 + 1], mh[i * 5 + 2], mh[i * 5 + 3], efix, gamma, o_); \
    for (int m = 0; m < nl; ++m) f[i * nl + m] = o_[m];                       \
  }
  FLUXES(0, ni);
  flux_diff(ni, f, fc);
  if (ist > 1) {
    for (int i = 1; i <= 4; ++i)
      for (int m = 0; m < nl; ++m) { ulo[i * nl + m] = qwl[i * nl + m]; uro[i * nl + m] = qwr[i * nl + m]; }
    for (int m = 0; m < nl; ++m) {
      double a = q[3 * 6 + m], b = q[4 * 6 + m], cc = q[5 * 6 + m], d = q[6 * 6 + m];
      qwl[0 * nl + m] = (3. * a - 1. * b) / 2.0;
      qwr[0 * nl + m] = qwl[0 * nl + m];
      qwl[1 * nl + m] = (19. * a + 25. * b - 2. * cc) / 42.0;
      qwr[1 * nl + m] = qwl[1 * nl + m];
      qwl[2 * nl + m] = (4. * a + 3. * b + 9. * cc + 2. * d) / 18.0;
      qwr[2 * nl + m] = qwl[2 * nl + m];
      qwl[3 * nl + m] = (-2. * a + 3. * b + 3. * cc + 2. * d) / 6.0;
      qwr[3 * nl + m] = qwl[3 * nl + m];
    }
    FLUXES(0, 3);
    for (int m = 0; m < nl; ++m)
      fc[0 * nl + m] = (-23. * f[0 * nl + m] + 21. * f[1 * nl + m] + 3. * f[2 * nl + m] - f[3 * nl + m]) / 24.0;
    for (int m = 0; m < nl; ++m) {
      double a = q[3 * 6 + m], b = q[4 * 6 + m], cc = q[5 * 6 + m], d = q[6 * 6 + m];
      qwl[1 * nl + m] = (3. * a + 6. * b - cc) / 8.0;
      qwr[1 * nl + m] = qwl[1 * nl + m];
      qwl[2 * nl + m] = (-30. * a + 145. * b + 29. * cc - 8. * d) / 136.0;
      qwr[2 * nl + m] = (2. * a + 49. * b + 125. * cc - 40. * d) / 136.0;
      qwl[3 * nl + m] = (-b + 6. * cc + 3. * d) / 8.0;
      qwr[3 * nl + m] = qwl[3 * nl + m];
      qwl[4 * nl + m] = (b + 2. * cc + 5. * d) / 8.0;
      qwr[4 * nl + m] = qwl[4 * nl + m];
      qwl[5 * nl + m] = (b + cc + 6. * d) / 8.0;
      qwr[5 * nl + m] = qwl[5 * nl + m];
    }
    FLUXES(1, 5);
    for (int m = 0; m < nl; ++m)
      fc[1 * nl + m] =
          (-22. * f[1 * nl + m] + 17. * f[2 * nl + m] + 9. * f[3 * nl + m] - 5. * f[4 * nl + m] + f[5 * nl + m]) / 24.0;
    for (int m = 0; m < nl; ++m) {
      double a = q[3 * 6 + m], b = q[4 * 6 + m], cc = q[5 * 6 + m], d = q[6 * 6 + m];
      qwr[1 * nl + m] = uro[1 * nl + m];
      qwl[2 * nl + m] = (-29. * a + 170. * b + 63. * cc + 12. * d) / 216.0;
      qwr[2 * nl + m] = uro[2 * nl + m];
      qwr[3 * nl + m] = uro[3 * nl + m];
      qwr[4 * nl + m] = uro[4 * nl + m];
    }
    FLUXES(1, 4);
    for (int m = 0; m < nl; ++m)
      fc[2 * nl + m] = (f[1 * nl + m] - 27. * f[2 * nl + m] + 27. * f[3 * nl + m] - f[4 * nl + m]) / 24.0;
  }
  if (ied < ni) {
    for (int i = 1; i <= 4; ++i) {
      int i1 = ni - i;
      for (int m = 0; m < nl; ++m) { ulo[(9 - i) * nl + m] = qwl[i1 * nl + m]; uro[(9 - i) * nl + m] = qwr[i1 * nl + m]; }
    }
    const int N = ni;
    for (int m = 0; m < nl; ++m) {
      double a = q[(N + 2) * 6 + m], b = q[(N + 1) * 6 + m], cc = q[N * 6 + m], d = q[(N - 1) * 6 + m];
      qwr[N * nl + m] = (3. * a - 1. * b) / 2.0;
      qwl[N * nl + m] = qwr[N * nl + m];
      qwr[(N - 1) * nl + m] = (19. * a + 25. * b - 2. * cc) / 42.0;
      qwl[(N - 1) * nl + m] = qwr[(N - 1) * nl + m];
      qwr[(N - 2) * nl + m] = (4. * a + 3. * b + 9. * cc + 2. * d) / 18.0;
      qwl[(N - 2) * nl + m] = qwr[(N - 2) * nl + m];
      qwr[(N - 3) * nl + m] = (-2. * a + 3. * b + 3. * cc + 2. * d) / 6.0;
      qwl[(N - 3) * nl + m] = qwr[(N - 3) * nl + m];
    }
    FLUXES(N - 3, N);
    for (int m = 0; m < nl; ++m)
      fc[(N - 1) * nl + m] =
          -(-23. * f[N * nl + m] + 21. * f[(N - 1) * nl + m] + 3. * f[(N - 2) * nl + m] - f[(N - 3) * nl + m]) / 24.0;
    for (int m = 0; m < nl; ++m) {
      double a = q[(N + 2) * 6 + m], b = q[(N + 1) * 6 + m], cc = q[N * 6 + m], d = q[(N - 1) * 6 + m];
      qwr[(N - 1) * nl + m] = (3. * a + 6. * b - cc) / 8.0;
      qwl[(N - 1) * nl + m] = qwr[(N - 1) * nl + m];
      qwr[(N - 2) * nl + m] = (-30. * a + 145. * b + 29. * cc - 8. * d) / 136.0;
      qwl[(N - 2) * nl + m] = (2. * a + 49. * b + 125. * cc - 40. * d) / 136.0;
      qwr[(N - 3) * nl + m] = (-b + 6. * cc + 3. * d) / 8.0;
      qwl[(N - 3) * nl + m] = qwr[(N - 3) * nl + m];
      qwr[(N - 4) * nl + m] = (b + 2. * cc + 5. * d) / 8.0;
      qwl[(N - 4) * nl + m] = qwr[(N - 4) * nl + m];
      qwr[(N - 5) * nl + m] = (b + cc + 6. * d) / 8.0;
      qwl[(N - 5) * nl + m] = qwr[(N - 5) * nl + m];
    }
    FLUXES(N - 5, N - 1);
    for (int m = 0; m < nl; ++m)
      fc[(N - 2) * nl + m] = -(-22. * f[(N - 1) * nl + m] + 17. * f[(N - 2) * nl + m] + 9. * f[(N - 3) * nl + m] -
                               5. * f[(N - 4) * nl + m] + f[(N - 5) * nl + m]) /
                             24.0;
    for (int m = 0; m < nl; ++m) {
      double a = q[(N + 2) * 6 + m], b = q[(N + 1) * 6 + m], cc = q[N * 6 + m], d = q[(N - 1) * 6 + m];
      qwl[(N - 1) * nl + m] = ulo[8 * nl + m];
      qwr[(N - 2) * nl + m] = (-29. * a + 170. * b + 63. * cc + 12. * d) / 216.0;
      qwl[(N - 2) * nl + m] = ulo[7 * nl + m];
      qwl[(N - 3) * nl + m] = ulo[6 * nl + m];
      qwl[(N - 4) * nl + m] = ulo[5 * nl + m];
    }
    FLUXES(N - 4, N - 1);
    for (int m = 0; m < nl; ++m)
      fc[(N - 3) * nl + m] =
          -(f[(N - 1) * nl + m] - 27. * f[(N - 2) * nl + m] + 27. * f[(N - 3) * nl + m] - f[(N - 4) * nl + m]) / 24.0;
  }
#undef FLUXES
}

static void inviscid_block(ob_block* k, const ob_params* p) {
  const long P = k->P;
  const int nmax = (k->idim > k->jdim ? (k->idim > k->kdim ? k->idim : k->kdim) : (k->jdim > k->kdim ? k->jdim : k->kdim));
  double* ql = calloc((nmax + 6) * 6, sizeof(double));
  double* ml = calloc((nmax + 6) * 5, sizeof(double));
  double* mh = calloc((nmax + 1) * 5, sizeof(double));
  double* fc = calloc((nmax + 6) * 5, sizeof(double));
  double* qwl = calloc((nmax + 1) * 5, sizeof(double));
  double* qwr = calloc((nmax + 1) * 5, sizeof(double));
  double* f = calloc((nmax + 1) * 5, sizeof(double));
  const int na[3] = {k->ny, k->nx, k->nx}, nbn[3] = {k->nz, k->nz, k->ny};
  for (int d = 0; d < 3; ++d)
    for (int bb = NG; bb < nbn[d] - NG; ++bb)
      for (int aa = NG; aa < na[d] - NG; ++aa) {
        ob_line L = line_of(k, d, aa, bb);
        int n = L.n;
        for (int ps = 0; ps <= n + 5; ++ps) {
          long id = L.base + (long)ps * L.stride;
          for (int m = 0; m < 6; ++m) ql[ps * 6 + m] = k->prim[m * P + id];
          ml[ps * 5 + 0] = k->metrics[(3 * d) * P + id];
          ml[ps * 5 + 1] = k->metrics[(3 * d + 1) * P + id];
          ml[ps * 5 + 2] = k->metrics[(3 * d + 2) * P + id];
          ml[ps * 5 + 3] = 0.0;
          ml[ps * 5 + 4] = k->vol[id];
        }
        metric_half(5, n, ml, mh);
        flux_line(n, ql, mh, fc, p->gama, qwl, qwr, f);
        for (int i = 0; i < n; ++i) {
          long id = L.base + (long)(NG + i) * L.stride;
          for (int m = 0; m < 5; ++m) k->dq[m * P + id] += fc[i * 5 + m];
        }
      }
  free(ql); free(ml); free(mh); free(fc); free(qwl); free(qwr); free(f);
}

void ob_compute_inviscid_rhs(ob_case* c) {
#pragma omp parallel for schedule(dynamic) num_threads(nthreads())
  for (int b = 0; b < c->nb; ++b) inviscid_block(&c->blocks[b], &c->p);
}

/* ========================================================================
 * LU-SGS update (src/solver/StateUpdater.cpp:175-580)
 * ======================================================================== */

static double spec_rad(const double* pr, double g, double kx, double ky, double kz, double kt) {
  double rho = pr[0], u = pr[1], v = pr[2], w = pr[3], p = pr[4];
  if (rho < 1e-30) rho = 1e-30;
  if (p < 1e-30) p = 1e-30;
  double c2 = g * p / rho;
  double c = sqrt(c2);
  double U = u * kx + v * ky + w * kz + kt;
  double gm = sqrt(kx * kx + ky * ky + kz * kz);
  return fabs(U) + c * gm;
}

static double visc_rad(const double* pr, double mu, double vol, double re, double kx, double ky, double kz) {
  double rho = pr[0];
  if (rho < 1e-30) rho = 1e-30;
  double g2 = kx * kx + ky * ky + kz * kz;
  double rmv = rho * vol;
  double coef = 2.0 * mu / (fabs(re) * rmv + 1.0e-30);
  return coef * g2;
}

static void flux_split(const double* pr, const double* dq, const double* met, double sr, int sign, double gamma,
                       double* res) {
  double rm = pr[0], um = pr[1], vm = pr[2], wm = pr[3], pm = pr[4];
  double drm = dq[0], dum = dq[1], dvm = dq[2], dwm = dq[3], dEm = dq[4];
  double nx = met[0], ny = met[1], nz = met[2], nt = met[3];
  double v2 = um * um + vm * vm + wm * wm;
  if (rm < 1e-30) rm = 1e-30;
  if (pm < 1e-30) pm = 1e-30;
  double c2 = gamma * pm / rm;
  double cm = sqrt(c2);
  double H = (gamma / (gamma - 1.0)) * (pm / rm) + 0.5 * v2;
  double ct = nx * um + ny * vm + nz * wm + nt;
  double gm = sqrt(nx * nx + ny * ny + nz * nz);
  if (gm < 1.0e-30) gm = 1.0e-30;
  double l1 = ct, l4 = ct + cm * gm, l5 = ct - cm * gm;
  l1 = 0.5 * (l1 + sign * sr);
  l4 = 0.5 * (l4 + sign * sr);
  l5 = 0.5 * (l5 + sign * sr);
  double x1 = (2.0 * l1 - l4 - l5) / (2.0 * c2);
  double x2 = (l4 - l5) / (2.0 * cm);
  double im = 1.0 / gm;
  double nxn = nx * im, nyn = ny * im, nzn = nz * im, ctn = (ct - nt) * im;
  double gm1 = gamma - 1.0;
  double af = 0.5 * gm1 * v2;
  double dc = ctn * drm - (nxn * dum + nyn * dvm + nzn * dwm);
  double dh = af * drm - gm1 * (um * dum + vm * dvm + wm * dwm - dEm);
  double c2dc = c2 * dc;
  res[0] = l1 * drm - dh * x1 - dc * x2;
  res[1] = l1 * dum + (nxn * c2dc - um * dh) * x1 + (nxn * dh - um * dc) * x2;
  res[2] = l1 * dvm + (nyn * c2dc - vm * dh) * x1 + (nyn * dh - vm * dc) * x2;
  res[3] = l1 * dwm + (nzn * c2dc - wm * dh) * x1 + (nzn * dh - wm * dc) * x2;
  res[4] = l1 * dEm + (ctn * c2dc - H * dh) * x1 + (ctn * dh - H * dc) * x2;
}

static void load5(const ob_block* k, const double* arr, long id, double* o) {
  for (int m = 0; m < 5; ++m) o[m] = arr[m * k->P + id];
}

static double diag_coeff(const ob_block* k, long id, double g, double dtl, int nvis, double re) {
  double pr[5];
  load5(k, k->prim, id, pr);
  const long P = k->P;
  const double* M = k->metrics;
  double ra = spec_rad(pr, g, M[id], M[P + id], M[2 * P + id], 0.0);
  double rb = spec_rad(pr, g, M[3 * P + id], M[4 * P + id], M[5 * P + id], 0.0);
  double rc = spec_rad(pr, g, M[6 * P + id], M[7 * P + id], M[8 * P + id], 0.0);
  double rva = 0, rvb = 0, rvc = 0;
  if (nvis == 1) {
    double mu = k->mu[id], vol = k->vol[id];
    if (vol < 1e-30) vol = 1e-30;
    rva = visc_rad(pr, mu, vol, re, M[id], M[P + id], M[2 * P + id]);
    rvb = visc_rad(pr, mu, vol, re, M[3 * P + id], M[4 * P + id], M[5 * P + id]);
    rvc = visc_rad(pr, mu, vol, re, M[6 * P + id], M[7 * P + id], M[8 * P + id]);
  }
  double rs = ra + rb + rc + rva + rvb + rvc;
  double beta = 1.0, wmig = 1.0;
  double term = (1.0 / dtl) + beta * wmig * rs;
  return 1.0 / term;
}

/* One neighbour's contribution (the add_nb lambdas, :326-347 and :401-426).
 * Returns 0 if the neighbour is outside the block interior. */
static int nb_contrib(const ob_block* k, int i, int j, int kk, int in, int jn, int kn, int midx, int sgn, int gs,
                      double g, int nvis, double re, double* out) {
  if (in < NG || in >= k->nx - NG || jn < NG || jn >= k->ny - NG || kn < NG || kn >= k->nz - NG) return 0;
  long idn = IX(k, in, jn, kn), id = IX(k, i, j, kk);
  double pn[5], dn[5];
  load5(k, k->prim, idn, pn);
  load5(k, k->dq, idn, dn);
  const long P = k->P;
  double met[4] = {k->metrics[midx * P + id], k->metrics[(midx + 1) * P + id], k->metrics[(midx + 2) * P + id], 0.0};
  double r = spec_rad(pn, g, met[0], met[1], met[2], 0);
  flux_split(pn, dn, met, r, sgn, g, out);
  if (nvis == 1) {
    double mun = k->mu[idn], voln = k->vol[idn];
    if (voln < 1e-30) voln = 1e-30;
    double rv = visc_rad(pn, mun, voln, re, met[0], met[1], met[2]);
    if (gs)
      for (int m = 0; m < 5; ++m) out[m] += 0.5 * rv * dn[m] * (double)sgn;
    else
      for (int m = 0; m < 5; ++m) out[m] += 0.5 * rv * dn[m];
  }
  return 1;
}

static void lusgs_sweep(ob_block* k, const double* rhs, double g, int nvis, double re, int dir) {
  int ist = dir == 1 ? NG : k->nx - NG - 1, ied = dir == 1 ? k->nx - NG : NG - 1;
  int jst = dir == 1 ? NG : k->ny - NG - 1, jed = dir == 1 ? k->ny - NG : NG - 1;
  int kst = dir == 1 ? NG : k->nz - NG - 1, ked = dir == 1 ? k->nz - NG : NG - 1;
  const long P = k->P;
  for (int kk = kst; dir == 1 ? kk < ked : kk > ked; kk += dir)
    for (int j = jst; dir == 1 ? j < jed : j > jed; j += dir)
      for (int i = ist; dir == 1 ? i < ied : i > ied; i += dir) {
        long id = IX(k, i, j, kk);
        double coed = diag_coeff(k, id, g, k->dt[id], nvis, re);
        double r0[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, fs[5];
        if (nb_contrib(k, i, j, kk, i - dir, j, kk, 0, dir, 0, g, nvis, re, fs))
          for (int m = 0; m < 5; ++m) r0[m] += fs[m];
        if (nb_contrib(k, i, j, kk, i, j - dir, kk, 3, dir, 0, g, nvis, re, fs))
          for (int m = 0; m < 5; ++m) r0[m] += fs[m];
        if (nb_contrib(k, i, j, kk, i, j, kk - dir, 6, dir, 0, g, nvis, re, fs))
          for (int m = 0; m < 5; ++m) r0[m] += fs[m];
        double wmig = 1.0;
        if (dir == 1) {
          for (int m = 0; m < 5; ++m) {
            double ex = -rhs[m * P + id];
            k->dq[m * P + id] = (ex + wmig * r0[m]) * coed;
          }
        } else {
          for (int m = 0; m < 5; ++m) k->dq[m * P + id] = k->dq[m * P + id] - (wmig * r0[m]) * coed;
        }
      }
}

static void gs_sweep(ob_block* k, const double* rhs, double g, int nvis, double re, int dir) {
  int ist = dir == 1 ? NG : k->nx - NG - 1, ied = dir == 1 ? k->nx - NG : NG - 1;
  int jst = dir == 1 ? NG : k->ny - NG - 1, jed = dir == 1 ? k->ny - NG : NG - 1;
  int kst = dir == 1 ? NG : k->nz - NG - 1, ked = dir == 1 ? k->nz - NG : NG - 1;
  const long P = k->P;
  static const int off[6][4] = {{-1, 0, 0, 0}, {0, -1, 0, 3}, {0, 0, -1, 6}, {1, 0, 0, 0}, {0, 1, 0, 3}, {0, 0, 1, 6}};
  for (int kk = kst; dir == 1 ? kk < ked : kk > ked; kk += dir)
    for (int j = jst; dir == 1 ? j < jed : j > jed; j += dir)
      for (int i = ist; dir == 1 ? i < ied : i > ied; i += dir) {
        long id = IX(k, i, j, kk);
        double coed = diag_coeff(k, id, g, k->dt[id], nvis, re);
        double r0[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, fs[5];
        for (int n = 0; n < 6; ++n) {
          int sgn = n < 3 ? 1 : -1;
          if (nb_contrib(k, i, j, kk, i + off[n][0], j + off[n][1], kk + off[n][2], off[n][3], sgn, 1, g, nvis, re, fs)) {
            if (sgn == 1)
              for (int m = 0; m < 5; ++m) r0[m] += fs[m];
            else
              for (int m = 0; m < 5; ++m) r0[m] -= fs[m];
          }
        }
        double wmig = 1.0;
        for (int m = 0; m < 5; ++m) {
          double resid = -rhs[m * P + id];
          k->dq[m * P + id] = (wmig * r0[m] + resid) * coed;
        }
      }
}

static void zero_bc_dq(ob_block* k, ob_region* regs, int nreg) {
  for (int r = 0; r < nreg; ++r) {
    const ob_region* R = &regs[r];
    if (R->d.bctype <= 0) continue;
    for (int kk = R->d.s_st[2] - 1 + NG; kk <= R->d.s_ed[2] - 1 + NG; ++kk)
      for (int j = R->d.s_st[1] - 1 + NG; j <= R->d.s_ed[1] - 1 + NG; ++j)
        for (int i = R->d.s_st[0] - 1 + NG; i <= R->d.s_ed[0] - 1 + NG; ++i)
          for (int m = 0; m < 5; ++m) DQ(k, i, j, kk, m) = 0.0;
  }
}

static void update_limited(ob_block* k, const ob_params* p) {
  const double g = p->gama, gm1 = g - 1.0, p_min = 1e-6, r_min = 1e-6;
  const long P = k->P;
  for (int kk = NG; kk < k->nz - NG; ++kk)
    for (int j = NG; j < k->ny - NG; ++j)
      for (int i = NG; i < k->nx - NG; ++i) {
        long id = IX(k, i, j, kk);
        double r = k->prim[id], u = k->prim[P + id], v = k->prim[2 * P + id], w = k->prim[3 * P + id];
        double pp = k->prim[4 * P + id];
        double E = pp / gm1 + 0.5 * r * (u * u + v * v + w * w);
        double d0 = k->dq[id], d1 = k->dq[P + id], d2 = k->dq[2 * P + id], d3 = k->dq[3 * P + id];
        double d4 = k->dq[4 * P + id];
        double qn[5] = {r + d0, r * u + d1, r * v + d2, r * w + d3, E + d4};
        double rn = qn[0];
        double ir = 1.0 / rn;
        double a = qn[1] * ir, b = qn[2] * ir, cc = qn[3] * ir;
        double pn = gm1 * (qn[4] - 0.5 * rn * (a * a + b * b + cc * cc));
        double dpp = fabs((pn - pp) / pp);
        double term = dmax(0.0, -0.3 + dpp);
        double Fa = 1.0 / (1.0 + 2.0 * term);
        if (term > 0.0) guard_hit(2);
        double qf[5] = {r + d0 * Fa, r * u + d1 * Fa, r * v + d2 * Fa, r * w + d3 * Fa, E + d4 * Fa};
        rn = qf[0];
        ir = 1.0 / rn;
        double un = qf[1] * ir, vn = qf[2] * ir, wn = qf[3] * ir;
        double pf = gm1 * (qf[4] - 0.5 * rn * (un * un + vn * vn + wn * wn));
        if (rn <= r_min || pf <= p_min) {
          /* keep the old state */
          guard_hit(3);
        } else {
          k->prim[id] = rn;
          k->prim[P + id] = un;
          k->prim[2 * P + id] = vn;
          k->prim[3 * P + id] = wn;
          k->prim[4 * P + id] = pf;
        }
      }
}

void ob_lusgs_sweeps(ob_case* c, int b, const double* rhs) {
  ob_block* k = &c->blocks[b];
  const ob_params* p = &c->p;
  lusgs_sweep(k, rhs, p->gama, p->nvis, p->reynolds, 1);
  lusgs_sweep(k, rhs, p->gama, p->nvis, p->reynolds, -1);
  gs_sweep(k, rhs, p->gama, p->nvis, p->reynolds, 1);
  gs_sweep(k, rhs, p->gama, p->nvis, p->reynolds, -1);
}

void ob_update_solution(ob_case* c) {
#pragma omp parallel for schedule(dynamic) num_threads(nthreads())
  for (int b = 0; b < c->nb; ++b) {
    ob_block* k = &c->blocks[b];
    zero_bc_dq(k, c->reg[b], c->nreg[b]);
    double* rhs = malloc(sizeof(double) * 5 * k->P);
    memcpy(rhs, k->dq, sizeof(double) * 5 * k->P);
    ob_lusgs_sweeps(c, b, rhs);
    free(rhs);
    zero_bc_dq(k, c->reg[b], c->nreg[b]);
    update_limited(k, &c->p);
  }
}

/* ========================================================================
 * Residual (src/solver/ResidualMonitor.cpp:11-138)
 * ======================================================================== */

/* Test option (not the reference's arithmetic): form the RMS residual's
 * sum of squares with Neumaier's compensated summation instead of the
 * reference's loop-order sum.  At millions of nodes the loop-order sum
 * carries a rounding error of order 1e-12 relative, larger than the parity
 * bar; the compensated sum is the exact sum to about 1e-16, which a
 * tree-ordered device sum approaches too.  The maximum and its location
 * are unaffected. */
static int g_accurate_sum = 0;
void ob_set_accurate_sum(int on) { g_accurate_sum = on; }

ob_residual ob_residual_compute(const ob_case* c) {
  ob_residual s;
  memset(&s, 0, sizeof s);
  double sum = 0.0, comp = 0.0;
  long long tot = 0;
  double mx = -1.0;
  int loc[5] = {-1, -1, -1, -1, -1};
  for (int b = 0; b < c->nb; ++b) {
    const ob_block* k = &c->blocks[b];
    for (int kk = NG; kk < k->nz - NG; ++kk)
      for (int j = NG; j < k->ny - NG; ++j)
        for (int i = NG; i < k->nx - NG; ++i) {
          long id = IX(k, i, j, kk);
          double vol = k->vol[id], dt = k->dt[id];
          if (vol < 1e-30) vol = 1e-30;
          if (dt < 1e-30) dt = 1e-30;
          double inv = 1.0 / (dt * vol);
          for (int m = 0; m < 5; ++m) {
            double r = fabs(k->dq[m * k->P + id]) * inv;
            if (g_accurate_sum) {
              const double x = r * r, t = sum + x;
              comp += fabs(sum) >= fabs(x) ? (sum - t) + x : (x - t) + sum;
              sum = t;
            } else {
              sum += r * r;
            }
            if (r > mx) {
              mx = r;
              loc[0] = b;
              loc[1] = i - NG + 1;
              loc[2] = j - NG + 1;
              loc[3] = kk - NG + 1;
              loc[4] = m + 1;
            }
          }
          tot += 5;
        }
  }
  s.max = mx;
  memcpy(s.max_loc, loc, sizeof loc);
  if (g_accurate_sum) sum += comp;
  if (tot > 0) s.rms = sqrt(sum / tot);
  return s;
}

/* ========================================================================
 * Step driver (src/core/OrionSolver.cpp:139-209, debug block removed)
 * ======================================================================== */

int ob_step(ob_case* c, ob_residual* out) {
  if (ob_exchange_bc(c)) return -1;
  ob_apply_physical_bc(c);
  ob_update_flow_states(c);
  ob_calculate_time_step(c);
  if (c->p.nvis == 1) {
    ob_compute_viscous_rhs(c);
    if (ob_average_interface_residuals(c)) return -1;
  }
  ob_compute_inviscid_rhs(c);
  if (ob_average_interface_residuals(c)) return -1;
  ob_update_solution(c);
  ob_residual r = ob_residual_compute(c);
  if (out) *out = r;
  return 0;
}

int ob_run(ob_case* c, int nsteps, double* hist) {
  for (int s = 0; s < nsteps; ++s) {
    ob_residual r;
    if (ob_step(c, &r)) return -1;
    if (hist) { hist[2 * s] = r.rms; hist[2 * s + 1] = r.max; }
  }
  return 0;
}

/* PostProcess::write_block_vts_binary's extraction loops
 * (src/postprocess/PostProcess.cpp:174-240): k, j, i over the interior,
 * each double value cast to float; Mach from prim with gama
 * (params.inflow.gama), sqrt(v2 / c2) where c2 > 1e-30, else 0. */
void ob_output_fields(const ob_case* c, int b, float* out) {
  const ob_block* k = &c->blocks[b];
  const long n = (long)k->idim * k->jdim * k->kdim;
  const double gamma = c->p.gama;
  long g = 0;
  for (int kk = 0; kk < k->kdim; ++kk)
    for (int j = 0; j < k->jdim; ++j)
      for (int i = 0; i < k->idim; ++i, ++g) {
        const long id = IX(k, i + NG, j + NG, kk + NG);
        const double r = k->prim[id], u = k->prim[k->P + id], v = k->prim[2 * k->P + id];
        const double w = k->prim[3 * k->P + id], p = k->prim[4 * k->P + id];
        out[g] = (float)r;
        out[n + g] = (float)p;
        out[2 * n + g] = (float)k->prim[5 * k->P + id];
        const double v2 = u * u + v * v + w * w;
        const double c2 = gamma * p / r;
        out[3 * n + g] = (float)((c2 > 1e-30) ? sqrt(v2 / c2) : 0.0);
        out[4 * n + 3 * g] = (float)u;
        out[4 * n + 3 * g + 1] = (float)v;
        out[4 * n + 3 * g + 2] = (float)w;
      }
}
