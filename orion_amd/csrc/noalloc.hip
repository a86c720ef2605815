// noalloc.hip -- the residual monitor over caller-held device fields,
// allocating nothing (orion_hip_residual_noalloc, include/orion_hip.h).
//
// It runs the kernels of orion_hip_residual (ResidualMonitor::compute,
// src/solver/ResidualMonitor.cpp:11-138): one pass of per-workgroup partial
// sums and first maxima, then one workgroup that reduces them.  Their
// scratch -- the block table, the interior-node prefix, the partials and
// the four result words -- lives in `scratch`, a file-scope device array of
// NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until the
// process ends.  In host code `scratch` names a host-side stand-in, not
// device memory, so the host only ever uses the address that
// hipGetSymbolAddress returns for it, once per device, on the first call
// made with that device current.
//
// Calls on one GPU share its array and keep the last call's contents;
// every word a call reads was written earlier by the same call (the table
// and prefix by its copies, the partials by its first kernel), so nothing
// needs clearing, but two calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/orion_hip.h"
#include "orion_dev.h"
#include "orion_kernels.h"

namespace orion_hip {
int fail(const char* fmt, ...);  // context.hip: records the orion_hip_last_error message, returns -1
}

using namespace orion_hip;

namespace {

constexpr int MAX_BLOCKS = ORION_NOALLOC_MAX_BLOCKS;
constexpr int MAXPARTS = 1024;  // the residual kernels' partial count (RES_MAXPARTS)
constexpr size_t al256(size_t x) { return (x + 255) / 256 * 256; }
constexpr size_t OFF_TABLE = 0;
constexpr size_t OFF_PREFIX = al256(OFF_TABLE + sizeof(BlockDev) * MAX_BLOCKS);
constexpr size_t OFF_PART = al256(OFF_PREFIX + sizeof(long long) * (MAX_BLOCKS + 1));
constexpr size_t OFF_KEY = al256(OFF_PART + sizeof(double) * 2 * MAXPARTS);
constexpr size_t OFF_OUT = al256(OFF_KEY + sizeof(long long) * MAXPARTS);
constexpr size_t NOALLOC_SCRATCH_BYTES = al256(OFF_OUT + sizeof(double) * 4);

}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mutex;
unsigned char* g_scratch[MAX_DEVICES] = {};  // device address of `scratch`, per device

int fail_msg(const char* what, hipError_t e) {
  return fail("residual_noalloc: %s failed: %s", what, hipGetErrorString(e));
}

// this device's copy of `scratch`
int scratch_of(int device, unsigned char** out) {
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail_msg("hipGetSymbolAddress", e);
    g_scratch[device] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[device];
  return 0;
}

}  // namespace

extern "C" int orion_hip_residual_noalloc(int nblocks, const int* dims, const double* const* dq,
                                          const double* const* dt, const double* const* vol, void* stream,
                                          orion_residual_stats* out) {
  if (nblocks <= 0 || nblocks > MAX_BLOCKS)
    return fail("residual_noalloc: %d blocks (1 to %d are accepted)", nblocks, MAX_BLOCKS);
  if (!dims || !dq || !dt || !vol || !out) return fail("residual_noalloc: null argument");
  int device = 0;
  hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess) return fail_msg("hipGetDevice", e);
  if (device < 0 || device >= MAX_DEVICES) return fail("residual_noalloc: device %d out of range", device);
  unsigned char* base = nullptr;
  if (scratch_of(device, &base)) return -1;
  // the block table (only the fields the residual kernels read) and the
  // prefix of interior node counts, built on the host
  std::vector<BlockDev> hb(nblocks);
  std::vector<long long> prefix(nblocks + 1, 0);
  for (int b = 0; b < nblocks; ++b) {
    const int i = dims[3 * b], j = dims[3 * b + 1], k = dims[3 * b + 2];
    if (i <= 0 || j <= 0 || k <= 0) return fail("residual_noalloc: block %d has empty sizes", b);
    if (!dq[b] || !dt[b] || !vol[b]) return fail("residual_noalloc: block %d has a null field", b);
    BlockDev B;
    std::memset(&B, 0, sizeof B);
    B.idim = i;
    B.jdim = j;
    B.kdim = k;
    B.nx = i + 2 * NG;
    B.ny = j + 2 * NG;
    B.nz = k + 2 * NG;
    B.P = (long long)B.nx * B.ny * B.nz;
    B.dq = const_cast<double*>(dq[b]);
    B.dt = const_cast<double*>(dt[b]);
    B.vol = const_cast<double*>(vol[b]);
    const long long n = (long long)i * j * k;
    if (n >= (1LL << 31)) return fail("residual_noalloc: block %d has %lld nodes (below 2^31 accepted)", b, n);
    hb[b] = B;
    prefix[b + 1] = prefix[b] + n;
  }
  const hipStream_t s = static_cast<hipStream_t>(stream);
  BlockDev* d_table = reinterpret_cast<BlockDev*>(base + OFF_TABLE);
  long long* d_prefix = reinterpret_cast<long long*>(base + OFF_PREFIX);
  double* d_part = reinterpret_cast<double*>(base + OFF_PART);
  long long* d_key = reinterpret_cast<long long*>(base + OFF_KEY);
  double* d_out = reinterpret_cast<double*>(base + OFF_OUT);
  if ((e = hipMemcpyAsync(d_table, hb.data(), sizeof(BlockDev) * nblocks, hipMemcpyHostToDevice, s)) != hipSuccess)
    return fail_msg("hipMemcpyAsync (block table)", e);
  if ((e = hipMemcpyAsync(d_prefix, prefix.data(), sizeof(long long) * (nblocks + 1), hipMemcpyHostToDevice, s)) !=
      hipSuccess)
    return fail_msg("hipMemcpyAsync (prefix)", e);
  const long long total = prefix[nblocks];
  const int nparts = residual_parts(total);
  Launch L{s, d_table, d_prefix, nblocks, Span{d_prefix, nullptr, nullptr, nblocks}, 0};
  launch_residual(L, total, 5 * total, d_part, d_key, nparts, d_out);
  if ((e = hipGetLastError()) != hipSuccess) return fail_msg("residual kernels", e);
  double r[4];
  if ((e = hipMemcpyAsync(r, d_out, sizeof r, hipMemcpyDeviceToHost, s)) != hipSuccess)
    return fail_msg("hipMemcpyAsync (result)", e);
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail_msg("hipStreamSynchronize", e);
  // the key of the first maximum: 5 x (global interior node) + variable
  out->rms_residual = r[0];
  out->max_residual = r[1];
  const long long key = (long long)r[2];
  const long long g = key / 5;
  int b = 0;
  while (b + 1 < nblocks && prefix[b + 1] <= g) ++b;
  const long long n = g - prefix[b];
  const BlockDev& B = hb[b];
  out->max_loc[0] = b;
  out->max_loc[1] = (int)(n % B.idim) + 1;
  out->max_loc[2] = (int)((n / B.idim) % B.jdim) + 1;
  out->max_loc[3] = (int)(n / ((long long)B.idim * B.jdim)) + 1;
  out->max_loc[4] = (int)(key % 5) + 1;
  out->phydtime = 0.0;  // formed by the time step, not by the residual monitor
  return 0;
}
