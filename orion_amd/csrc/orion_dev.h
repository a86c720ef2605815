// orion_dev.h -- device data model shared by the gfx950 kernels.
//
// Every block lives in one HBM arena as ghost-padded SoA fields with the
// reference's layout (src/core/OrionArray.hpp:241-261): box
// (idim+6, jdim+6, kdim+6), i fastest, variable slowest.  Kernels cover all
// blocks in one launch: a thread's flat index is mapped to (block, local
// index) through a prefix-offset table of the iteration space in use.
//
// All arithmetic is FP64 and every translation unit is compiled with
// -ffp-contract=off, so a*b+c stays two rounded operations exactly as in
// the reference's x86-64 build; f64 division and sqrt on gfx950 are the
// correctly rounded sequences.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define ORION_NG 3

namespace orion_hip {

constexpr int NG = ORION_NG;

// A device pointer held in a descriptor table: on the device it is typed as
// a global-address-space pointer, so that the loads and stores made through
// it (and through pointers derived from it) are global instructions with a
// scalar base, not flat ones (which also count against the LDS/scalar
// counter and need 64-bit addresses in vector registers).  It converts to
// and from a plain pointer.
#if defined(__HIP_DEVICE_COMPILE__)
#define ORION_GLOBAL_AS __attribute__((address_space(1)))
#else
#define ORION_GLOBAL_AS
#endif
template <class T>
struct GPtr {
  ORION_GLOBAL_AS T* p;
  GPtr() = default;
  __host__ __device__ GPtr(T* x) : p((ORION_GLOBAL_AS T*)x) {}
  __host__ __device__ operator T*() const { return (T*)p; }
};

// Accesses at a 32-bit byte offset from a wave-uniform global base.  A
// block's field arrays hold at most 15 fields of P doubles and blocks are
// held to P < 2^25 padded points (checked when a context is built), so a
// word's byte offset from its array's base fits in 32 bits.  Written as
// (uniform base) + (32-bit offset), an access is one global instruction
// with a scalar base and a 32-bit vector offset; a 64-bit index instead
// needs a 64-bit address (two registers and a 64-bit add) per access.
constexpr long long MAX_BLOCK_POINTS = 1LL << 25;
__device__ __forceinline__ double ldo(const double* base, unsigned byte_off) {
  return *(const ORION_GLOBAL_AS double*)((const ORION_GLOBAL_AS char*)base + byte_off);
}
__device__ __forceinline__ void sto(double* base, unsigned byte_off, double v) {
  *(ORION_GLOBAL_AS double*)((ORION_GLOBAL_AS char*)base + byte_off) = v;
}
// byte offset of element e of field k in an array of fields of P words
__device__ __forceinline__ unsigned boff(unsigned k, unsigned P, unsigned e) { return (k * P + e) << 3; }

// x / C for a constant C, correctly rounded: q = x * RN(1/C), then one
// correction with the exact residual, e = fma(-C, q, x), q' = fma(e, rc, q)
// (Markstein: with the correctly rounded reciprocal and a faithful q the
// corrected quotient is the correctly rounded one, short of underflow).
// Where the residual is 0 (q exact, x = +-0 included: q keeps x's sign) or
// not a number (x infinite or NaN: q is then x / C) q itself is the
// quotient; a quotient that may be subnormal (|x| < 2^-960) takes the
// division.  Three FP64 operations instead of the division sequence's ten.
template <int C>
__device__ __forceinline__ double div_c(double x) {
  constexpr double rc = 1.0 / (double)C;
  const double q = x * rc;
  const double e = __builtin_fma(-(double)C, q, x);
  double r = (e == 0.0 || e != e) ? q : __builtin_fma(e, rc, q);
  if (__builtin_expect(__builtin_fabs(x) < 0x1p-960 && x != 0.0, 0)) r = x / (double)C;
  return r;
}

struct Params {
  double gama, prl, moo, reynolds, visc, tref, twall, cfl;
  double roo, uoo, voo, woo, poo, too;
  int csrv, nvis, method, ntmst, gcl;
};

// Fields of one block (device pointers into the arena).
struct BlockDev {
  int idim, jdim, kdim;  // interior sizes
  int nx, ny, nz;        // padded sizes
  long long P;           // nx*ny*nz
  GPtr<double> prim;          // [6][P] rho,u,v,w,p,T
  GPtr<double> met;           // [9][P]
  GPtr<double> vol;           // [P]
  GPtr<double> dq;            // [5][P]
  GPtr<double> rhs;           // [5][P] LU-SGS right-hand side copy
  GPtr<double> dt;            // [P]
  GPtr<double> c;             // [P]
  GPtr<double> mu;            // [P]
  GPtr<double> q;             // [5][P] conservative (write-only; may be null)
  GPtr<double> duvwt;         // [12][P] viscous cell derivatives (scratch)
  GPtr<double> fx;            // [3][5][P] viscous face fluxes per direction (scratch)
  GPtr<double> fxi;           // [3][5][P] inviscid face fluxes (scratch; the same array as fx without viscosity)
  GPtr<double> fcov;          // closure overrides, [3][lines][6][5] (scratch)
  GPtr<double> fgeo;          // [3][FG_N][P] face geometry per direction (static, see FG_N)
  GPtr<unsigned char> frozen; // [P] 1 inside a physical-BC window
  long long fcov_off[3]; // offset (doubles) of each direction's overrides
  // LU-SGS sweep layout of this block: axes (march, col, row) as real axes
  // 0=i,1=j,2=k, their interior sizes, the padded column count and the
  // block's first cell in the sweep arrays.
  int sax[3];
  int sn[3];
  int sncp;
  long long sbase;
  // The block's request: its index in a context that batches several
  // independent solves (0 when the context holds one), its inputs, and the
  // viscous coefficients the host derives from them.  Kernels take every
  // per-request value from here, so blocks of different requests share a
  // launch.
  int req;
  Params par;
  double cp_prl;  // cp / Pr, cp = 1 / ((gama - 1) moo^2)
  double re_inv;  // 1 / Re, 0 for |Re| <= 1e-30
};

// Face geometry, interpolated once after the metrics (the grid is static)
// and stored per direction d at the face's slot (f + 2) along d, like the
// face fluxes: 0..8 the nine metrics and 9 the volume in the central form
// of interp_line_to_half (viscous flux), 10..12 the direction's metrics with
// the one-sided end formulas of the inviscid flux.
constexpr int FG_N = 13;

// LU-SGS working set in sweep layout (all blocks, n cells incl. padding).
// A block's cell (m, r, c) along (march, row, col) lies on hyperplane
// s = m + r + c; its cell index q = sbase + (s * nrow + r) * ncol_padded + c
// puts one hyperplane row in contiguous cells.  Fields are stored in pairs,
// so that one 16-byte load brings two fields of a cell: pair k of a group
// of K pairs lies at (q - c) * K + k * ncol_padded + c, i.e. every row holds
// K planes of pairs and consecutive columns are consecutive pairs.
constexpr int SW_OWN = 8;  // met0..met8, coed, rhs0..rhs4 (+1 unused)
constexpr int SW_D = 3;    // u, v, w, c2, H, 2mu/(|Re| rho vol + 1e-30) (a_f and c are recomputed)
constexpr int SW_DQ = 3;   // dq0..dq4 (+1 unused)
constexpr int SW_GA = 8;   // Gauss-Seidel forward: terms of the +1 neighbours along i, j, k (3 x 5, +1 unused)
constexpr int SW_PRE = 3;  // Gauss-Seidel backward: sum of the -1 neighbours' terms (5, +1 unused)
struct SweepArrays {
  double2* own;
  double2* D;
  double2* dq;
  double2* ga;
  double2* pre;
  long long n;
};

// field f of the cell (q, c) in a pair group of K pairs
__host__ __device__ inline long long sweep_field(long long q, int c, int ncp, int K, int f) {
  return 2 * ((q - c) * K + (long long)(f >> 1) * ncp + c) + (f & 1);
}

// A halo point entry: the arena offset of field 0 of a point (low
// HALO_OFF_BITS bits) and the field stride P of its block (the bits above),
// one entry for the destination and one for the source.
constexpr int HALO_OFF_BITS = 39;
constexpr unsigned long long HALO_OFF_MASK = (1ull << HALO_OFF_BITS) - 1;
__host__ __device__ inline unsigned long long halo_point(long long off, long long P) {
  return (unsigned long long)off | ((unsigned long long)P << HALO_OFF_BITS);
}

// Progress flags of the LU-SGS tiles, one per 256 bytes: a flag is polled by
// the tile's downstream neighbours while the tile stores it, and flags of
// unrelated tiles in one line would make those accesses contend.
constexpr int FLAG_STRIDE = 32;

// The flux differences' tiles: DIFF_TI x DIFF_TJ (i, j) nodes of DIFF_TK
// consecutive k planes per workgroup (256 lanes, a column of DIFF_TK nodes
// each).
constexpr int DIFF_TI = 16;
constexpr int DIFF_TJ = 16;
constexpr int DIFF_TK = 2;

// LU-SGS tiles: SWEEP_TW columns by SWEEP_TR rows of pencils (the block
// layout pads columns to whole multiples of SWEEP_TW).  16 x 8 tiles were
// measured slower than 16 x 16 on the 64-block mesh: a step's time is set
// by the update lanes' per-cell chain, not by the tile's bytes, and the
// extra tile diagonals add lag.
constexpr int SWEEP_TW = 16;
constexpr int SWEEP_TR = 16;

// One tile of the LU-SGS pipeline: SWEEP_TR x SWEEP_TW (row, col) pencils of a block.
struct SweepTile {
  int block;
  int r0, c0;           // first row and column
  int s_first, s_last;  // hyperplanes holding cells of the tile
  int nbr[4];           // tiles across the col-1, row-1, col+1, row+1 edges (-1: none)
};

__host__ __device__ inline long long sweep_index(const BlockDev& b, int m, int r, int c) {
  return b.sbase + ((long long)(m + r + c) * b.sn[2] + r) * b.sncp + c;
}

__host__ __device__ inline long long pidx(const BlockDev& b, int i, int j, int k) {
  return (long long)i + (long long)b.nx * ((long long)j + (long long)b.ny * (long long)k);
}

// Workgroups of a launch are dispatched to the device's 8 XCDs round-robin
// (workgroup w runs on XCD w % 8) and every XCD has its own L2.  xcd_block()
// renumbers them so that each XCD walks one contiguous range of the
// iteration space: neighbouring tiles, which share stencil data, then meet in
// the same L2.  It is a bijection of [0, gridDim.x).
__host__ __device__ inline unsigned xcd_remap(unsigned b, unsigned n) {
  const unsigned per = n >> 3, rem = n & 7u, x = b & 7u, q = b >> 3;
  return x < rem ? x * (per + 1) + q : rem * (per + 1) + (x - rem) * per + q;
}
__device__ inline long long xcd_thread() {
  return (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
}

// A multi-block iteration space as launched: each block's items are padded
// to whole 256-thread workgroups, so a workgroup never spans two blocks and
// its block comes from one uniform table load (and the block's fields from
// scalar loads), instead of a per-lane search.
constexpr int SPAN_TB = 256;
struct Span {
  const long long* off;  // [nb+1] prefix of item counts
  const int* wgb;        // [nwg] block of each workgroup
  const int* wg0;        // [nb] first workgroup of each block
  int nb;
};

// This thread's (block, item) of span S, false for the padding threads.
// Workgroups are renumbered XCD-contiguously first (xcd_remap).
__device__ inline bool span_item(const Span& S, int* bi, long long* item) {
  const unsigned wg = xcd_remap(blockIdx.x, gridDim.x);
  const int b = S.wgb[wg];
  const long long n = (long long)(wg - (unsigned)S.wg0[b]) * SPAN_TB + threadIdx.x;
  if (n >= S.off[b + 1] - S.off[b]) return false;
  *bi = b;
  *item = n;
  return true;
}

// block lookup: largest b with off[b] <= g (off has nb+1 entries)
__device__ inline int find_block(const long long* __restrict__ off, int nb, long long g) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Per-direction geometry helpers: stride of padded index along d and the
// interior count along d.
__host__ __device__ inline long long dstride(const BlockDev& b, int d) {
  return d == 0 ? 1LL : (d == 1 ? (long long)b.nx : (long long)b.nx * b.ny);
}
__host__ __device__ inline int dcount(const BlockDev& b, int d) {
  return d == 0 ? b.idim : (d == 1 ? b.jdim : b.kdim);
}

// Index decodes work in 32-bit unsigned arithmetic (a block's item counts
// are checked to fit when the iteration spaces are built): a 64-bit
// division costs several times the instructions of a 32-bit one.
__device__ inline void udivmod(unsigned n, unsigned d, unsigned* q, unsigned* r) {
  *q = n / d;
  *r = n - *q * d;
}

// Decode an index of the "faces along d" space (faces 0..n_d on every
// interior line, i fastest in memory order) into the padded index of the
// face's line start (position 0 along d) and the face number f.
__device__ inline void decode_face(const BlockDev& b, int d, long long n64, long long* base, int* f) {
  const unsigned n = (unsigned)n64;
  unsigned q, r, q2, r2;
  if (d == 0) {
    udivmod(n, (unsigned)b.idim + 1u, &q, &r);
    udivmod(q, (unsigned)b.jdim, &q2, &r2);
    *f = (int)r;
    *base = pidx(b, 0, (int)r2 + NG, (int)q2 + NG);
  } else if (d == 1) {
    udivmod(n, (unsigned)b.idim, &q, &r);
    udivmod(q, (unsigned)b.jdim + 1u, &q2, &r2);
    *f = (int)r2;
    *base = pidx(b, (int)r + NG, 0, (int)q2 + NG);
  } else {
    udivmod(n, (unsigned)b.idim, &q, &r);
    udivmod(q, (unsigned)b.jdim, &q2, &r2);
    *f = (int)q2;
    *base = pidx(b, (int)r + NG, (int)r2 + NG, 0);
  }
}

// Decode an index of the "face pairs along d" space (faces 2g and 2g+1,
// g = 0..(n_d+1)/2 on every interior line, same order as the face space).
__device__ inline void decode_face_pair(const BlockDev& b, int d, long long n64, long long* base, int* f0) {
  const unsigned n = (unsigned)n64;
  unsigned q, r, q2, r2;
  if (d == 0) {
    udivmod(n, (unsigned)(b.idim + 2) / 2u, &q, &r);
    udivmod(q, (unsigned)b.jdim, &q2, &r2);
    *f0 = 2 * (int)r;
    *base = pidx(b, 0, (int)r2 + NG, (int)q2 + NG);
  } else if (d == 1) {
    udivmod(n, (unsigned)b.idim, &q, &r);
    udivmod(q, (unsigned)(b.jdim + 2) / 2u, &q2, &r2);
    *f0 = 2 * (int)r2;
    *base = pidx(b, (int)r + NG, 0, (int)q2 + NG);
  } else {
    udivmod(n, (unsigned)b.idim, &q, &r);
    udivmod(q, (unsigned)b.jdim, &q2, &r2);
    *f0 = 2 * (int)q2;
    *base = pidx(b, (int)r + NG, (int)r2 + NG, 0);
  }
}

// Decode an index of the "lines along d" space (interior lines, i fastest).
__device__ inline long long decode_line(const BlockDev& b, int d, long long n64) {
  const unsigned n = (unsigned)n64;
  unsigned q, r;
  if (d == 0) {
    udivmod(n, (unsigned)b.jdim, &q, &r);
    return pidx(b, 0, (int)r + NG, (int)q + NG);
  } else if (d == 1) {
    udivmod(n, (unsigned)b.idim, &q, &r);
    return pidx(b, (int)r + NG, 0, (int)q + NG);
  }
  udivmod(n, (unsigned)b.idim, &q, &r);
  return pidx(b, (int)r + NG, (int)q + NG, 0);
}
__host__ __device__ inline long long nlines(const BlockDev& b, int d) {
  return d == 0 ? (long long)b.jdim * b.kdim : (d == 1 ? (long long)b.idim * b.kdim : (long long)b.idim * b.jdim);
}
// line number (within direction d) of an interior point
__device__ inline long long line_id(const BlockDev& b, int d, int i, int j, int k) {
  if (d == 0) return (long long)(j - NG) + (long long)b.jdim * (k - NG);
  if (d == 1) return (long long)(i - NG) + (long long)b.idim * (k - NG);
  return (long long)(i - NG) + (long long)b.idim * (j - NG);
}

// decode an interior point index into padded (i,j,k)
__device__ inline void decode_interior(const BlockDev& b, long long n64, int* i, int* j, int* k) {
  unsigned q, r, q2, r2;
  udivmod((unsigned)n64, (unsigned)b.idim, &q, &r);
  udivmod(q, (unsigned)b.jdim, &q2, &r2);
  *i = (int)r + NG;
  *j = (int)r2 + NG;
  *k = (int)q2 + NG;
}

}  // namespace orion_hip
