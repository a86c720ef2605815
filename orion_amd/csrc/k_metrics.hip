// k_metrics.hip -- grid metrics on the device (setup, gfx950, FP64).
//
// compute_grid_metrics (src/mesh/MetricComputer.cpp:103-369): 4th-order
// half-node interpolation (value_half_node_1d :21-45) followed by the
// 6th-order difference (flux_dxyz_1d :47-99) of x,y,z along each index
// direction, then either the plain cross products (:228-252) or the
// Visbal-Gaitonde conservative form (gcl=1, :257-367).  One thread per node;
// the half-node values a node needs are recomputed locally.  Metrics and
// volume are written on interior points only; ghost entries stay zero, as in
// the reference.  Scratch layout during setup (the face-flux scratch):
//   s[0..2]  x,y,z     s[3..11] GCL products     s[12..14] x_xi,x_et,x_zt
#include "orion_dev.h"
#include "orion_kernels.h"

namespace orion_hip {

// node value p (0..n-1) of a line whose node 0 sits at q
struct NodeLine {
  const double* q;
  long long s;
  __device__ inline double operator()(int p) const { return q[(long long)p * s]; }
};

__device__ inline double mhalf(const NodeLine& Q, int n, int h) {
  if (h >= 2 && h <= n - 2) return (9.0 * (Q(h - 1) + Q(h)) + -1.0 * (Q(h + 1) + Q(h - 2))) / 16.0;
  if (h == 1) return (5.0 * Q(0) + 15.0 * Q(1) + -5.0 * Q(2) + 1.0 * Q(3)) / 16.0;
  if (h == n - 1) return (5.0 * Q(n - 1) + 15.0 * Q(n - 2) + -5.0 * Q(n - 3) + 1.0 * Q(n - 4)) / 16.0;
  if (h == 0) return (35.0 * Q(0) - 35.0 * Q(1) + 21.0 * Q(2) - 5.0 * Q(3)) / 16.0;
  return (35.0 * Q(n - 1) - 35.0 * Q(n - 2) + 21.0 * Q(n - 3) - 5.0 * Q(n - 4)) / 16.0;
}

__device__ inline double mderiv(const NodeLine& Q, int n, int t) {
  auto F = [&](int h) { return mhalf(Q, n, h); };
  if (t >= 2 && t <= n - 3) {
    const double t1 = F(t + 1) - F(t);
    const double t2 = F(t + 2) - F(t - 1);
    const double t3 = F(t + 3) - F(t - 2);
    return (2250.0 * t1 - 125.0 * t2 + 9.0 * t3) / 1920.0;
  }
  if (t == 0) return (-23.0 * F(0) + 21.0 * F(1) + 3.0 * F(2) - F(3)) / 24.0;
  if (t == 1) return (F(0) - 27.0 * F(1) + 27.0 * F(2) - F(3)) / 24.0;
  if (t == n - 2) return -(F(n) - 27.0 * F(n - 1) + 27.0 * F(n - 2) - F(n - 3)) / 24.0;
  return -(-23.0 * F(n) + 21.0 * F(n - 1) + 3.0 * F(n - 2) - F(n - 3)) / 24.0;
}

__device__ inline double dnode(const BlockDev& b, const double* fld, long long id, const int* pos, int d) {
  const long long s = dstride(b, d);
  const int t = pos[d] - NG;
  const NodeLine Q{fld + id - (long long)t * s, s};
  return mderiv(Q, dcount(b, d), t);
}

__global__ void __launch_bounds__(256) metric_a_kernel(const BlockDev* __restrict__ blocks,
                                                       Span SP, int gcl) {
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;
  const BlockDev& b = blocks[bi];
  int i, j, k;
  decode_interior(b, gi, &i, &j, &k);
  const long long id = pidx(b, i, j, k);
  const int pos[3] = {i, j, k};
  const long long P = b.P;
  const double* S = b.fx;
  double D[9];
  for (int c = 0; c < 3; ++c)
    for (int d = 0; d < 3; ++d) D[3 * c + d] = dnode(b, S + (long long)c * P, id, pos, d);
  if (gcl != 1) {
    const double x_xi = D[0], x_et = D[1], x_zt = D[2];
    const double y_xi = D[3], y_et = D[4], y_zt = D[5];
    const double z_xi = D[6], z_et = D[7], z_zt = D[8];
    const double kcx = y_et * z_zt - z_et * y_zt;
    const double kcy = z_et * x_zt - x_et * z_zt;
    const double kcz = x_et * y_zt - y_et * x_zt;
    const double etx = y_zt * z_xi - z_zt * y_xi;
    const double ety = z_zt * x_xi - x_zt * z_xi;
    const double etz = x_zt * y_xi - y_zt * x_xi;
    const double ctx = y_xi * z_et - z_xi * y_et;
    const double cty = z_xi * x_et - x_xi * z_et;
    const double ctz = x_xi * y_et - y_xi * x_et;
    const double mm[9] = {kcx, kcy, kcz, etx, ety, etz, ctx, cty, ctz};
    for (int m = 0; m < 9; ++m) b.met[(long long)m * P + id] = mm[m];
    b.vol[id] = x_xi * kcx + x_et * etx + x_zt * ctx;
  } else {
    const double xm = S[id], ym = S[P + id], zm = S[2 * P + id];
    double* T = b.fx + 3 * P;
    for (int m = 0; m < 3; ++m) {
      T[(long long)m * P + id] = D[m] * ym;
      T[(long long)(m + 3) * P + id] = D[m + 3] * zm;
      T[(long long)(m + 6) * P + id] = D[m + 6] * xm;
    }
    double* XK = b.fx + 12 * P;
    for (int m = 0; m < 3; ++m) XK[(long long)m * P + id] = D[m];
  }
}

__global__ void __launch_bounds__(256) metric_gcl_kernel(const BlockDev* __restrict__ blocks,
                                                         Span SP) {
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;
  const BlockDev& b = blocks[bi];
  int i, j, k;
  decode_interior(b, gi, &i, &j, &k);
  const long long id = pidx(b, i, j, k);
  const int pos[3] = {i, j, k};
  const long long P = b.P;
  const double* T = b.fx + 3 * P;
  const int sel[3][6] = {{7, 8, 1, 2, 4, 5}, {6, 8, 0, 2, 3, 5}, {6, 7, 0, 1, 3, 4}};
  double D[18];
  for (int d = 0; d < 3; ++d)
    for (int c = 0; c < 6; ++c) D[6 * d + c] = dnode(b, T + (long long)sel[d][c] * P, id, pos, d);
  const double kcx = D[17] - D[11];
  const double kcy = D[13] - D[7];
  const double kcz = D[15] - D[9];
  const double etx = D[5] - D[16];
  const double ety = D[1] - D[12];
  const double etz = D[3] - D[14];
  const double ctx = D[10] - D[4];
  const double cty = D[6] - D[0];
  const double ctz = D[8] - D[2];
  const double mm[9] = {kcx, kcy, kcz, etx, ety, etz, ctx, cty, ctz};
  for (int m = 0; m < 9; ++m) b.met[(long long)m * P + id] = mm[m];
  const double* XK = b.fx + 12 * P;
  b.vol[id] = XK[id] * kcx + XK[P + id] * etx + XK[2 * P + id] * ctx;
}

__global__ void __launch_bounds__(256) small_vol_kernel(const BlockDev* __restrict__ blocks,
                                                        Span SP,
                                                        unsigned long long* count) {
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;
  const BlockDev& b = blocks[bi];
  int i, j, k;
  decode_interior(b, gi, &i, &j, &k);
  // check_grid_metrics (MetricComputer.cpp:433-470): vol < 1e-15 is fatal
  if (!(b.vol[pidx(b, i, j, k)] >= 1.0e-15)) atomicAdd(count, 1ULL);
}

// The node coordinates of each block, copied compactly (x, y, z, each
// idim*jdim*kdim values, i fastest) to fx + 3P, spread into the padded
// layout at fx[0 .. 3P) that the metric kernels read.  The arena starts
// zeroed, so the ghost points stay 0 as the reference's grid arrays have
// them.  Source and destination do not overlap.
__global__ void __launch_bounds__(256) grid_scatter_kernel(const BlockDev* __restrict__ blocks, Span SP) {
  int bi;
  long long gi;
  if (!span_item(SP, &bi, &gi)) return;
  const BlockDev& b = blocks[bi];
  int i, j, k;
  decode_interior(b, gi, &i, &j, &k);
  const long long n = (long long)b.idim * b.jdim * b.kdim;
  const double* stage = b.fx + 3 * b.P;
  const long long id = pidx(b, i, j, k);
  for (int m = 0; m < 3; ++m) b.fx[m * b.P + id] = stage[m * n + gi];
}

static inline unsigned grid_of(long long n, int tb) { return (unsigned)((n + tb - 1) / tb); }

void launch_grid_scatter(const Launch& L, long long n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(grid_scatter_kernel, dim3((unsigned)L.nwg), dim3(SPAN_TB), 0, L.stream, L.blocks, L.span);
}

void launch_metrics(const Launch& L, long long n, int gcl, double*, hipStream_t) {
  if (n <= 0) return;
  hipLaunchKernelGGL(metric_a_kernel, dim3((unsigned)L.nwg), dim3(SPAN_TB), 0, L.stream, L.blocks, L.span, gcl);
  if (gcl == 1)
    hipLaunchKernelGGL(metric_gcl_kernel, dim3((unsigned)L.nwg), dim3(SPAN_TB), 0, L.stream, L.blocks, L.span);
}
void launch_count_small_vol(const Launch& L, long long n, unsigned long long* count) {
  if (n <= 0) return;
  hipLaunchKernelGGL(small_vol_kernel, dim3((unsigned)L.nwg), dim3(SPAN_TB), 0, L.stream, L.blocks, L.span, count);
}

}  // namespace orion_hip
