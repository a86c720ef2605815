// context.hip -- host side of the C ABI (include/orion_hip.h).
//
// A context owns one HBM arena with every block's fields, the multi-block
// iteration tables, and the boundary topology compiled once into device
// maps: a halo gather list, per-rank physical-BC task lists, interface-
// averaging staging maps and the frozen-node masks.  Operators enqueue
// kernels on the context's stream; nothing synchronises inside a step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <functional>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <filesystem>
#include <iomanip>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/orion_hip.h"
#include "orion_dev.h"
#include "orion_kernels.h"

using namespace orion_hip;

static thread_local std::string g_err;
static thread_local int g_device = 0;

// Contexts may be used from several host threads, one context per thread
// (the serving workers).  Creating or destroying a context makes
// synchronous allocations and copies on the legacy stream, which HIP
// refuses while another thread captures a graph; these and the capture of
// a context's step are therefore serialised.  (Creation destroys on its
// error paths, hence recursive.)
static std::recursive_mutex g_capture_mutex;

// records the message returned by orion_hip_last_error; returns -1
namespace orion_hip {
int fail(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return -1;
}
}  // namespace orion_hip

#define HIPCHK(x)                                                                              \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess) return fail("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

struct Space {
  std::vector<long long> h;  // prefix offsets (nb+1)
  std::vector<int> hwg0;     // first workgroup of each block (host copy)
  long long* d = nullptr;
  int* wgb = nullptr;        // block of each workgroup of the padded space
  int* wg0 = nullptr;        // first workgroup of each block
  long long nwg = 0;
  long long total() const { return h.back(); }
};

struct orion_hip_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  hipStream_t side_stream = nullptr;  // inviscid face fluxes beside the viscous chain (captured step)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int nb = 0;
  std::vector<BlockDev> hb;
  BlockDev* db = nullptr;
  Params p{};                          // request 0's inputs (the shared ones for a batch)
  // requests batched in the context: blocks req_first[r] .. req_first[r+1]
  // are request r's; each has its own residual and time step per step
  int nreq = 1;
  std::vector<int> req_first;
  int* d_rseg = nullptr;               // [nreq+1] each request's update partials
  long long* d_rtot = nullptr;         // [nreq] 5 x each request's interior nodes
  int* d_rdone = nullptr;              // residual workgroups finished (left at 0)
  char* arena = nullptr;
  size_t arena_bytes = 0;
  size_t map_bytes = 0;
  Space pad, intr, face[3], line[3];
  Space fpair[3];                      // face pairs (2g, 2g+1) per line along d
  Space ltask[3];                      // six boundary-closure tasks per line along d
  // the same three spaces each joined over the directions (virtual blocks
  // d * nb + b): one launch per step for all directions
  Space face3, fpair3, ltask3;
  Space dtile;                         // flux differences: DIFF_TI x DIFF_TJ (i, j) nodes of DIFF_TK k planes
  Space ptile;                         // LU-SGS pack tiles: 16 x 16 (march, col) cells of one row
  SweepArrays sw{};                    // LU-SGS working set in sweep layout
  std::vector<SweepTile> tiles;
  SweepTile* d_tiles = nullptr;
  int* d_order[2] = {nullptr, nullptr};  // tile launch order: forward, backward sweeps
  std::vector<int> sched_order[ORION_SCHEDULES][2];  // candidate orders (host; empty for the chained schedule)
  int schedule = 0;                      // the one in d_order, or ORION_SCHEDULE_CHAINED
  int* d_chain = nullptr;                // chained launch: tasks tile << 2 | sweep, topological order
  int n_chain = 0;                       // its task count (0: the chained schedule is unavailable)
  unsigned long long* d_flags = nullptr; // per-tile progress
  unsigned long long* d_epoch = nullptr; // sweep counter
  int* d_err = nullptr;                  // set by a sweep whose wait ran out
  long long* d_halo = nullptr;          // halo: (destination, source) arena offsets, one pair per field value
  long long n_halo = 0;
  unsigned long long* d_halo_pt = nullptr;  // halo: whole points, packed (halo_point_kernel)
  long long n_halo_pt = 0;
  BCTask* d_bc = nullptr;
  std::vector<int> bc_first, bc_count;  // per rank: task offset and count
  std::vector<long long> bc_total;      // per rank: face nodes
  AvgSrc* d_avg_src = nullptr;
  AvgDst* d_avg_dst = nullptr;
  long long n_avg_src = 0, n_avg_dst = 0;
  double* d_stage = nullptr;
  double* d_part = nullptr;
  long long* d_pidx = nullptr;
  double* d_upart = nullptr;      // residual partials of the update kernel: sums, then maxima
  long long* d_ukey = nullptr;    // and the maxima's keys (one per update workgroup)
  int nparts = 1;
  double* d_res = nullptr;  // 4 doubles per residual slot
  int res_cap = 0;
  unsigned long long* d_dtmin = nullptr;
  int dtmin_cap = 0;
  unsigned long long* d_count = nullptr;
  std::vector<void*> allocs;
  char* pool = nullptr;  // current chunk for small device tables
  size_t pool_used = 0, pool_size = 0;
  int keep_q = 0;
  int* d_step = nullptr;             // step index for per-step outputs
  hipGraph_t graph = nullptr;        // one captured pseudo-time step
  hipGraphExec_t gexec = nullptr;
  hipStream_t graph_stream = nullptr;
  int use_graph = 1;
  bool overlap = true;  // inviscid face fluxes on side_stream (viscous steps)
  // plain per-block launches: every kernel once per block, in block order,
  // no captured step (orion_hip_set_per_block)
  bool per_block = false;
  int* d_zeros = nullptr;               // zero workgroup tables of one-block spans
  std::vector<int> tile_first, tile_count;  // each block's tiles in `tiles`
  int* d_border[2] = {nullptr, nullptr};   // per-block tile orders at the tiles' positions
  // the viscous face-flux work queue (one persistent launch per step):
  // items, their count, the ticket counter, faces per marching item and
  // the launch's workgroups
  ViscItem* d_vitems = nullptr;
  long long vitems_cap = 0;
  int n_vitems = 0;
  int* d_vhead = nullptr;
  int visc_seg = 32;
  int visc_grid = 2048;
  // the fused viscous form (visc_fused_kernel: face fluxes and their
  // difference in one pass, items of 16 x 16 columns by visc_seg planes),
  // the default where every block line has at least 4 nodes
  bool visc_fused = false;
  bool visc_fused_ok = false;
  ViscTile* d_vtiles = nullptr;
  long long vtiles_cap = 0;
  int n_vtiles = 0;
};

// Device allocations of a context, freed with it.  Tables up to 64 KB share
// 1 MB chunks (256-byte aligned), so a context makes a few dozen fewer
// hipMalloc and hipFree calls: hipFree waits for the device, which a
// serving batch that is rebuilt while the device works would feel.
static int dalloc(orion_hip_ctx* c, void** p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  constexpr size_t SMALL = 64 << 10, CHUNK = 1 << 20, ALIGN = 256;
  if (bytes <= SMALL) {
    const size_t need = (bytes + ALIGN - 1) / ALIGN * ALIGN;
    if (!c->pool || c->pool_used + need > c->pool_size) {
      void* q = nullptr;
      hipError_t e = hipMalloc(&q, CHUNK);
      if (e != hipSuccess) return fail("hipMalloc(%zu) failed: %s", CHUNK, hipGetErrorString(e));
      c->allocs.push_back(q);
      c->map_bytes += CHUNK;
      c->pool = (char*)q;
      c->pool_used = 0;
      c->pool_size = CHUNK;
    }
    *p = c->pool + c->pool_used;
    c->pool_used += need;
    return 0;
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  c->allocs.push_back(*p);
  c->map_bytes += bytes;
  return 0;
}

// Every set-up copy and fill goes to the context's stream, in order after
// what was queued there before (the arena's zeroing among it): the stream
// does not block on the legacy stream, so a legacy-stream copy could land
// before an earlier fill of the same memory that is still queued behind
// other contexts' work.  The wait keeps the host vector alive until read.
template <class T>
static int upload_vec(orion_hip_ctx* c, T** dst, const std::vector<T>& v) {
  if (dalloc(c, (void**)dst, sizeof(T) * std::max<size_t>(v.size(), 1))) return -1;
  if (!v.empty()) {
    HIPCHK(hipMemcpyAsync(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

static int make_space(orion_hip_ctx* c, Space& s, const std::vector<long long>& counts) {
  s.h.assign(counts.size() + 1, 0);
  for (size_t b = 0; b < counts.size(); ++b) {
    // kernels decode a block's item index in 32-bit arithmetic
    if (counts[b] >= (1LL << 31)) return fail("block %d: %lld items exceed the 32-bit index range", (int)b + 1, counts[b]);
    s.h[b + 1] = s.h[b] + counts[b];
  }
  std::vector<int> wgb, wg0(counts.size());
  for (size_t b = 0; b < counts.size(); ++b) {
    wg0[b] = (int)wgb.size();
    const long long n = (counts[b] + SPAN_TB - 1) / SPAN_TB;
    for (long long w = 0; w < n; ++w) wgb.push_back((int)b);
  }
  if (wgb.size() >= (size_t)INT32_MAX) return fail("iteration space too large");
  s.nwg = (long long)wgb.size();
  s.hwg0 = wg0;
  return upload_vec(c, &s.d, s.h) || upload_vec(c, &s.wgb, wgb) || upload_vec(c, &s.wg0, wg0);
}

// workgroups and items of block b in space s
static long long nwg_of(const Space& s, int b) {
  const long long hi = b + 1 < (int)s.hwg0.size() ? s.hwg0[b + 1] : s.nwg;
  return hi - s.hwg0[b];
}
static long long items_of(const Space& s, int b) { return s.h[b + 1] - s.h[b]; }
// one block of a space as a launch of its own: the block table and the
// prefix start at block b, and every workgroup maps to block 0 of that view
static Launch LAB(orion_hip_ctx* c, const Space& s, int b) {
  return Launch{c->stream, c->db + b, s.d + b, 1, Span{s.d + b, c->d_zeros, c->d_zeros, 1}, nwg_of(s, b)};
}

static Launch LA(orion_hip_ctx* c, const Space& s, hipStream_t st = nullptr) {
  return Launch{st ? st : c->stream, c->db, s.d, c->nb, Span{s.d, s.wgb, s.wg0, c->nb}, s.nwg};
}
// a space joined over the three directions (3 nb virtual blocks)
static Launch LA3(orion_hip_ctx* c, const Space& s, hipStream_t st = nullptr) {
  return Launch{st ? st : c->stream, c->db, s.d, 3 * c->nb, Span{s.d, s.wgb, s.wg0, 3 * c->nb}, s.nwg};
}

// ---------------------------------------------------------------------------
// region preprocessing (BCReader.cpp:125-143, BCPreprocess.cpp:16-192)
// ---------------------------------------------------------------------------
struct Reg {
  orion_region_desc d;
  int s_nd = 0, s_lr = 0, t_nd = 0;
  int map_dims[3] = {0, 0, 0};
  std::vector<int> image, jmage, kmage;
};

static int analyze_connect(Reg& r) {
  const int* s_st = r.d.s_st;
  const int* s_ed = r.d.s_ed;
  const int* t_st = r.d.t_st;
  const int* t_ed = r.d.t_ed;
  const int s_nd = r.s_nd;
  r.t_nd = 0;
  for (int m = 0; m < 3; ++m)
    if (t_st[m] == t_ed[m]) r.t_nd = m + 1;
  const int t_nd = r.t_nd;
  if (t_nd == 0 || s_nd == 0) return fail("interface region without a constant axis");
  int dirm[3][3] = {{0}};
  bool done = false;
  for (int m = 0; m < 3 && !done; ++m)
    for (int n = 0; n < 3 && !done; ++n)
      if ((m + 1) != s_nd && (n + 1) != t_nd) {
        int js1 = s_st[m];
        if (std::abs(js1) < std::abs(s_ed[m])) js1 = s_ed[m];
        int js2 = t_st[n];
        if (std::abs(js2) < std::abs(t_ed[n])) js2 = t_ed[n];
        if (js1 * js2 > 0) {
          dirm[n][m] = 1;
          dirm[t_nd - 1][s_nd - 1] = 1;
          dirm[6 - (n + 1) - t_nd - 1][6 - (m + 1) - s_nd - 1] = 1;
          done = true;
        }
      }
  int s_sign[3], t_sign[3], st_sign[3];
  for (int m = 0; m < 3; ++m) {
    s_sign[m] = ((m + 1) != s_nd && std::abs(s_st[m]) > std::abs(s_ed[m])) ? -1 : 1;
    t_sign[m] = ((m + 1) != t_nd && std::abs(t_st[m]) > std::abs(t_ed[m])) ? -1 : 1;
  }
  for (int m = 0; m < 3; ++m) {
    int co = 0;
    for (int n = 0; n < 3; ++n)
      if (dirm[n][m] == 1) co = n;
    st_sign[m] = t_sign[co] * s_sign[m];
  }
  int lo[3];
  for (int m = 0; m < 3; ++m) {
    const int a = std::abs(s_st[m]), b = std::abs(s_ed[m]);
    lo[m] = std::min(a, b);
    r.map_dims[m] = std::max(a, b) - lo[m] + 1;
  }
  const long long tot = (long long)r.map_dims[0] * r.map_dims[1] * r.map_dims[2];
  r.image.assign(tot, 0);
  r.jmage.assign(tot, 0);
  r.kmage.assign(tot, 0);
  for (int k = std::abs(s_st[2]); k != std::abs(s_ed[2]) + s_sign[2]; k += s_sign[2])
    for (int j = std::abs(s_st[1]); j != std::abs(s_ed[1]) + s_sign[1]; j += s_sign[1])
      for (int i = std::abs(s_st[0]); i != std::abs(s_ed[0]) + s_sign[0]; i += s_sign[0]) {
        const int di = (i - std::abs(s_st[0])) * st_sign[0];
        const int dj = (j - std::abs(s_st[1])) * st_sign[1];
        const int dk = (k - std::abs(s_st[2])) * st_sign[2];
        const long long f = ((long long)(k - lo[2]) * r.map_dims[1] + (j - lo[1])) * r.map_dims[0] + (i - lo[0]);
        r.image[f] = std::abs(t_st[0]) + dirm[0][0] * di + dirm[0][1] * dj + dirm[0][2] * dk;
        r.jmage[f] = std::abs(t_st[1]) + dirm[1][0] * di + dirm[1][1] * dj + dirm[1][2] * dk;
        r.kmage[f] = std::abs(t_st[2]) + dirm[2][0] * di + dirm[2][1] * dj + dirm[2][2] * dk;
      }
  return 0;
}

// get_memory_window (HaloExchanger.cpp:17-64): padded inclusive ranges
static void memory_window(const Reg& r, const BlockDev& b, int mode, int* rb, int* re) {
  const int dphy[3] = {b.idim, b.jdim, b.kdim};
  int cc[3];
  for (int m = 0; m < 3; ++m) cc[m] = std::abs(r.d.s_st[m]) - 1 + NG;
  const int dir = r.s_nd - 1;
  if (mode == 0) {
    if (r.s_lr == -1) { rb[dir] = cc[dir]; re[dir] = cc[dir] + NG - 1; }
    else { rb[dir] = cc[dir] - NG + 1; re[dir] = cc[dir]; }
  } else {
    if (r.s_lr == -1) { rb[dir] = cc[dir] - NG; re[dir] = cc[dir] - 1; }
    else { rb[dir] = cc[dir] + 1; re[dir] = cc[dir] + NG; }
  }
  for (int m = 0; m < 3; ++m) {
    if (m == dir) continue;
    const int a = std::abs(r.d.s_st[m]), bb = std::abs(r.d.s_ed[m]);
    int ps = std::min(a, bb), pe = std::max(a, bb);
    if (ps > 1) ps -= 1;
    if (pe < dphy[m]) pe += 1;
    rb[m] = ps - 1 + NG;
    re[m] = pe - 1 + NG;
  }
}

static void init_params(Params& p, const orion_inputs* in) {
  const double PAI = 3.14159265358979323846;
  std::memset(&p, 0, sizeof p);
  p.gama = 1.40;
  p.prl = 0.72;
  const double attack = in->attack * PAI / 180.0;
  const double sideslip = in->sideslip * PAI / 180.0;
  p.uoo = std::cos(attack) * std::cos(sideslip);
  p.voo = std::sin(attack) * std::cos(sideslip);
  p.woo = std::sin(sideslip);
  p.roo = 1.0;
  p.too = 1.0;
  p.reynolds = (in->reynolds / 1.0) * 1.0;
  p.moo = in->moo;
  p.poo = 1.0 / (p.gama * p.moo * p.moo);
  p.tref = in->tref;
  p.twall = in->twall;
  p.visc = (in->tref != 0.0) ? 110.4 / in->tref : 0.0;
  p.cfl = in->cfl;
  p.csrv = in->csrv;
  p.nvis = in->nvis;
  p.method = in->method;
  p.ntmst = in->ntmst;
  p.gcl = in->gcl;
}

static long long arena_fields(const BlockDev& b, int keep_q, int viscous, long long lines_total) {
  const long long P = b.P;
  long long n = 0;
  auto al = [](long long x) { return (x + 31) / 32 * 32; };  // 256-byte alignment
  n += al(6 * P) + al(9 * P) + al(P) + al(5 * P) + al(5 * P) + al(P) + al(P) + al(P);
  if (keep_q) n += al(5 * P);
  n += al(12 * P) + al(15 * P) + al(30 * lines_total) + al((P + 7) / 8) + al(3 * FG_N * P);
  if (viscous) n += al(15 * P);
  return n;
}

// Runs fn(b, scratch) for b = 0 .. n-1 on up to 16 host threads (blocks
// taken in turn from a shared counter); scratch is a per-thread vector that
// fn may size and reuse (32-bit: it indexes at most 5 P < 2^28 values).  Returns the error message fn gave for the lowest
// block that gave one, or an empty string.
template <class F>
static std::string parallel_blocks(int n, const F& fn) {
  const int hw = (int)std::thread::hardware_concurrency();
  const int nt = std::max(1, std::min(std::min(n, 16), hw > 0 ? hw : 1));
  std::atomic<int> next(0);
  std::vector<std::string> err((size_t)std::max(n, 0));
  auto work = [&]() {
    std::vector<int> scratch;
    for (;;) {
      const int b = next.fetch_add(1);
      if (b >= n) break;
      err[(size_t)b] = fn(b, scratch);
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(work);
  work();
  for (auto& t : th) t.join();
  for (auto& e : err)
    if (!e.empty()) return e;
  return std::string();
}

static int build_topology(orion_hip_ctx* c, const std::vector<std::vector<Reg>>& R,
                          const std::function<void(const char*)>& timing) {
  const int nb = c->nb;
  // ---------------- halo gather list ----------------
  struct Msg {
    int tag, block, reg;
    bool used;
  };
  std::vector<Msg> msgs;
  std::unordered_map<long long, std::vector<int>> by_tag;  // tag -> messages in send order
  for (int b = 0; b < nb; ++b)
    for (int r = 0; r < (int)R[b].size(); ++r)
      if (R[b][r].d.nbt > 0) {
        by_tag[(long long)R[b][r].d.nbt * 1000 + R[b][r].d.ibcwin].push_back((int)msgs.size());
        msgs.push_back({R[b][r].d.nbt * 1000 + R[b][r].d.ibcwin, b, r, false});
      }
  // Every ghost write of every receiving region, in the order the
  // reference performs them (receiving blocks, their regions, then m, k, j,
  // i of the window); a later write to the same ghost replaces an earlier
  // one.  Destinations only repeat within one receiving block (where its
  // windows overlap), so each block's writes are reduced through an index of
  // its own destinations, last writer kept, in linear time.
  auto arena_off = [&](int b, int m, long long id) {
    return (long long)((c->hb[b].prim - (double*)c->arena) + (long long)m * c->hb[b].P + id);
  };
  // the partner message of every receiving window, matched in the order
  // the reference receives (receiving blocks, then their windows)
  struct Link {
    int reg, sblock, sreg;
  };
  std::vector<std::vector<Link>> links(nb);
  for (int b = 0; b < nb; ++b)
    for (int r = 0; r < (int)R[b].size(); ++r) {
      const Reg& rg = R[b][r];
      if (rg.d.nbt <= 0) continue;
      const int tag = (b + 1) * 1000 + (r + 1);
      Msg* M = nullptr;
      auto it = by_tag.find(tag);
      if (it != by_tag.end())
        for (int q : it->second)
          if (!msgs[q].used) { M = &msgs[q]; break; }
      if (!M) return fail("halo: no partner message for block %d window %d", b + 1, r + 1);
      M->used = true;
      links[b].push_back({r, M->block, M->reg});
    }
  // each receiving block's writes, reduced to its last writers, on host
  // threads (a block's destinations are its own, so blocks are independent).
  // A block whose windows each unpack as many points as their senders pack
  // receives whole points: field m of a ghost comes from field m of one
  // source point, so its writes reduce per point to one entry of five
  // fields (halo_point).  A window of another length shifts the fields
  // across points; its block keeps one pair per field value.
  // ORION_HIP_HALO_FIELDS=1 keeps every block on the per-field pairs (the
  // two forms' results are compared in the tests)
  const char* hf = std::getenv("ORION_HIP_HALO_FIELDS");
  const bool fields_only = hf && std::atoi(hf) != 0;
  std::vector<std::vector<long long>> bpairs(nb);
  std::vector<std::vector<unsigned long long>> bpts(nb);
  parallel_blocks(nb, [&](int b, std::vector<int>& where) {
    const BlockDev& B = c->hb[b];
    std::vector<long long> touched;
    bool whole = !fields_only && B.P < (1ll << (64 - HALO_OFF_BITS));
    for (const Link& L : links[b]) {
      const BlockDev& A = c->hb[L.sblock];
      int sb[3], se[3], rb[3], re[3];
      memory_window(R[L.sblock][L.sreg], A, 0, sb, se);
      memory_window(R[b][L.reg], B, 1, rb, re);
      const long long ns = (long long)(se[0] - sb[0] + 1) * (se[1] - sb[1] + 1) * (se[2] - sb[2] + 1);
      const long long nr = (long long)(re[0] - rb[0] + 1) * (re[1] - rb[1] + 1) * (re[2] - rb[2] + 1);
      const long long end = arena_off(L.sblock, 4, A.P);
      if (ns != nr || A.P >= (1ll << (64 - HALO_OFF_BITS)) || end > (long long)HALO_OFF_MASK ||
          arena_off(b, 4, B.P) > (long long)HALO_OFF_MASK)
        whole = false;
    }
    // the index of each destination's entry (-1: none yet), over the block's
    // points (whole points) or its five fields' values; every slot is -1
    // again when the block is done
    const long long nslot = whole ? B.P : 5 * B.P;
    if ((long long)where.size() < nslot) where.assign((size_t)nslot, -1);
    if (whole) {
      std::vector<unsigned long long>& pts = bpts[b];
      for (const Link& L : links[b]) {
        const BlockDev& A = c->hb[L.sblock];
        int sb[3], se[3], rb[3], re[3];
        memory_window(R[L.sblock][L.sreg], A, 0, sb, se);
        memory_window(R[b][L.reg], B, 1, rb, re);
        // the sender's packing order (k, j, i over its window) walked
        // alongside the receiver's unpacking order, the same for each field
        int si = sb[0], sj = sb[1], sk = sb[2];
        for (int k = rb[2]; k <= re[2]; ++k)
          for (int j = rb[1]; j <= re[1]; ++j)
            for (int i = rb[0]; i <= re[0]; ++i) {
              const long long local = pidx(B, i, j, k);
              const unsigned long long src = halo_point(arena_off(L.sblock, 0, pidx(A, si, sj, sk)), A.P);
              if (++si > se[0]) {
                si = sb[0];
                if (++sj > se[1]) {
                  sj = sb[1];
                  ++sk;
                }
              }
              int& w = where[(size_t)local];
              if (w >= 0) {
                pts[2 * (size_t)w + 1] = src;
              } else {
                w = (int)(pts.size() / 2);
                pts.push_back(halo_point(arena_off(b, 0, local), B.P));
                pts.push_back(src);
                touched.push_back(local);
              }
            }
      }
      for (long long t : touched) where[(size_t)t] = -1;
      return std::string();
    }
    std::vector<long long>& pairs = bpairs[b];
    for (const Link& L : links[b]) {
      const Reg& rg = R[b][L.reg];
      const BlockDev& A = c->hb[L.sblock];
      int sb[3], se[3], rb[3], re[3];
      memory_window(R[L.sblock][L.sreg], A, 0, sb, se);
      memory_window(rg, B, 1, rb, re);
      const long long npts = (long long)(se[0] - sb[0] + 1) * (se[1] - sb[1] + 1) * (se[2] - sb[2] + 1);
      const long long count = 5 * npts;
      long long ptr = 0;
      // the sender's packing order (m, k, j, i over its window) walked
      // alongside the receiver's unpacking order
      int sm = 0, si = sb[0], sj = sb[1], sk = sb[2];
      for (int m = 0; m < 5; ++m)
        for (int k = rb[2]; k <= re[2]; ++k)
          for (int j = rb[1]; j <= re[1]; ++j)
            for (int i = rb[0]; i <= re[0]; ++i) {
              if (ptr >= count) continue;
              ++ptr;
              const long long local = (long long)m * B.P + pidx(B, i, j, k);
              const long long src = arena_off(L.sblock, sm, pidx(A, si, sj, sk));
              if (++si > se[0]) {
                si = sb[0];
                if (++sj > se[1]) {
                  sj = sb[1];
                  if (++sk > se[2]) {
                    sk = sb[2];
                    ++sm;
                  }
                }
              }
              int& w = where[(size_t)local];
              if (w >= 0) {
                pairs[2 * (size_t)w + 1] = src;
              } else {
                w = (int)(pairs.size() / 2);
                pairs.push_back(arena_off(b, m, pidx(B, i, j, k)));
                pairs.push_back(src);
                touched.push_back(local);
              }
            }
    }
    for (long long t : touched) where[(size_t)t] = -1;
    return std::string();
  });
  {  // the blocks' lists one after another, copied from where they were built
    size_t tot = 0;
    for (auto& v : bpairs) tot += v.size();
    c->n_halo = (long long)(tot / 2);
    if (dalloc(c, (void**)&c->d_halo, sizeof(long long) * std::max<size_t>(tot, 1))) return -1;
    size_t off = 0;
    for (auto& v : bpairs) {
      if (!v.empty())
        HIPCHK(hipMemcpyAsync(c->d_halo + off, v.data(), sizeof(long long) * v.size(), hipMemcpyHostToDevice,
                              c->stream));
      off += v.size();
    }
    size_t ptot = 0;
    for (auto& v : bpts) ptot += v.size();
    c->n_halo_pt = (long long)(ptot / 2);
    if (dalloc(c, (void**)&c->d_halo_pt, sizeof(unsigned long long) * std::max<size_t>(ptot, 1))) return -1;
    off = 0;
    for (auto& v : bpts) {
      if (!v.empty())
        HIPCHK(hipMemcpyAsync(c->d_halo_pt + off, v.data(), sizeof(unsigned long long) * v.size(),
                              hipMemcpyHostToDevice, c->stream));
      off += v.size();
    }
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  timing("  halo gather list");

  // ---------------- physical BC tasks per rank ----------------
  int maxrank = 0;
  std::vector<std::vector<const Reg*>> phys(nb);
  for (int b = 0; b < nb; ++b) {
    for (auto& rg : R[b])
      if (rg.d.bctype > 0) phys[b].push_back(&rg);
    maxrank = std::max(maxrank, (int)phys[b].size());
  }
  std::vector<BCTask> tasks;
  for (int rk = 0; rk < maxrank; ++rk) {
    c->bc_first.push_back((int)tasks.size());
    long long tot = 0;
    for (int b = 0; b < nb; ++b) {
      if (rk >= (int)phys[b].size()) continue;
      const Reg& rg = *phys[b][rk];
      BCTask t{};
      t.block = b;
      t.bctype = rg.d.bctype;
      for (int m = 0; m < 3; ++m) {
        t.lo[m] = rg.d.s_st[m] - 1 + NG;
        t.hi[m] = rg.d.s_ed[m] - 1 + NG;
        t.mid[m] = (t.lo[m] + t.hi[m]) / 2;
      }
      t.s_nd = rg.s_nd;
      t.s_lr = rg.s_lr;
      if (t.s_nd == 0) return fail("block %d: physical BC region is not a face", b + 1);
      long long cnt = 1;
      for (int m = 0; m < 3; ++m) cnt *= std::max(0, t.hi[m] - t.lo[m] + 1);
      if (cnt == 0) continue;
      t.start = tot;
      tot += cnt;
      tasks.push_back(t);
    }
    c->bc_count.push_back((int)tasks.size() - c->bc_first.back());
    c->bc_total.push_back(tot);
  }
  if (upload_vec(c, &c->d_bc, tasks)) return -1;
  timing("  BC tasks");

  // ---------------- interface averaging maps ----------------
  // Every face node of every interface window W (block b) takes the dq/vol
  // that the region X packed for it (X of block a targets W; the last such
  // packer wins), at the partner node W's map names.  A node on several
  // windows is averaged once per window, in region order: entries sorted by
  // (block, node, region order).
  std::map<std::pair<int, int>, std::pair<int, int>> fill_of;
  for (int a = 0; a < nb; ++a)
    for (int x = 0; x < (int)R[a].size(); ++x) {
      const Reg& X = R[a][x];
      if (X.d.bctype >= 0) continue;
      fill_of[{X.d.nbt - 1, X.d.ibcwin - 1}] = {a, x};  // last packer wins
    }
  struct AE {
    int b;
    long long didx, seq;
    int a;
    long long sidx;
  };
  // each block's entries (in its windows' order, so seq orders them as one
  // global count would), sorted by (node, region order) on host threads
  std::vector<std::vector<AE>> bae(nb);
  std::vector<long long> bruns(nb, 0);  // each block's runs of equal nodes
  const std::string aerr = parallel_blocks(nb, [&](int b, std::vector<int>&) -> std::string {
    std::vector<AE>& ae = bae[b];
    long long seq = 0;
    for (int w = 0; w < (int)R[b].size(); ++w) {
      const Reg& W = R[b][w];
      if (W.d.bctype >= 0) continue;
      auto it = fill_of.find({b, w});
      if (it == fill_of.end())
        return "interface buffer never filled for block " + std::to_string(b + 1) + " window " + std::to_string(w + 1);
      const int a = it->second.first;
      const Reg& X = R[a][it->second.second];
      int xlo[3], xn[3], lo[3], n[3], tlo[3], tn[3];
      for (int m = 0; m < 3; ++m) {
        int u = std::abs(X.d.s_st[m]), v = std::abs(X.d.s_ed[m]);
        xlo[m] = std::min(u, v);
        xn[m] = std::max(u, v) - xlo[m] + 1;
        u = std::abs(W.d.s_st[m]);
        v = std::abs(W.d.s_ed[m]);
        lo[m] = std::min(u, v);
        n[m] = std::max(u, v) - lo[m] + 1;
        u = std::abs(W.d.t_st[m]);
        v = std::abs(W.d.t_ed[m]);
        tlo[m] = std::min(u, v);
        tn[m] = std::max(u, v) - tlo[m] + 1;
      }
      const long long xcount = (long long)xn[0] * xn[1] * xn[2];
      for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
          for (int i = 0; i < n[0]; ++i) {
            const long long f = (long long)k * (n[0] * n[1]) + (long long)j * n[0] + i;
            const int ir = W.image[f] - tlo[0], jr = W.jmage[f] - tlo[1], kr = W.kmage[f] - tlo[2];
            const long long rem = (long long)kr * (tn[0] * tn[1]) + (long long)jr * tn[0] + ir;
            if (rem < 0 || rem >= xcount)
              return "interface map of block " + std::to_string(b + 1) + " window " + std::to_string(w + 1) +
                     " leaves its partner window";
            const int xi = (int)(rem % xn[0]), xj = (int)((rem / xn[0]) % xn[1]), xk = (int)(rem / ((long long)xn[0] * xn[1]));
            const long long sidx = pidx(c->hb[a], xlo[0] - 1 + NG + xi, xlo[1] - 1 + NG + xj, xlo[2] - 1 + NG + xk);
            const long long didx = pidx(c->hb[b], lo[0] - 1 + NG + i, lo[1] - 1 + NG + j, lo[2] - 1 + NG + k);
            ae.push_back({b, didx, seq++, a, sidx});
          }
    }
    std::sort(ae.begin(), ae.end(), [](const AE& x, const AE& y) {
      return x.didx < y.didx || (x.didx == y.didx && x.seq < y.seq);
    });
    for (size_t q = 0; q < ae.size(); ++q) bruns[b] += (q == 0 || ae[q].didx != ae[q - 1].didx) ? 1 : 0;
    return std::string();
  });
  if (!aerr.empty()) return fail("%s", aerr.c_str());
  // the staging entries (blocks in turn, each in its sorted order) and one
  // destination per run of equal nodes, each block's written at its offsets
  // on host threads
  std::vector<long long> s_off(nb + 1, 0), d_off(nb + 1, 0);
  for (int b = 0; b < nb; ++b) {
    s_off[b + 1] = s_off[b] + (long long)bae[b].size();
    d_off[b + 1] = d_off[b] + bruns[b];
  }
  std::vector<AvgSrc> src((size_t)s_off[nb]);
  std::vector<AvgDst> dst((size_t)d_off[nb]);
  parallel_blocks(nb, [&](int b, std::vector<int>&) -> std::string {
    const std::vector<AE>& v = bae[b];
    long long dn = d_off[b];
    for (size_t q = 0; q < v.size(); ++q) {
      const long long sq = s_off[b] + (long long)q;
      if (q == 0 || v[q].didx != v[q - 1].didx) {
        AvgDst& D = dst[(size_t)dn++];
        D.block = b;
        D.idx = v[q].didx;
        D.start = sq;
        D.count = 0;
      }
      dst[(size_t)(dn - 1)].count++;
      src[(size_t)sq] = AvgSrc{v[q].a, v[q].sidx};
    }
    std::vector<AE>().swap(bae[b]);
    return std::string();
  });
  c->n_avg_src = (long long)src.size();
  c->n_avg_dst = (long long)dst.size();
  if (upload_vec(c, &c->d_avg_src, src)) return -1;
  if (upload_vec(c, &c->d_avg_dst, dst)) return -1;
  if (dalloc(c, (void**)&c->d_stage, sizeof(double) * 5 * std::max<long long>(1, c->n_avg_src))) return -1;
  timing("  interface averaging lists");

  // ---------------- frozen masks (set_boundary_dq_zero windows) ----------------
  for (int b = 0; b < nb; ++b) {
    const BlockDev& B = c->hb[b];
    std::vector<unsigned char> mask(B.P, 0);
    for (auto& rg : R[b]) {
      if (rg.d.bctype <= 0) continue;
      for (int k = rg.d.s_st[2] - 1 + NG; k <= rg.d.s_ed[2] - 1 + NG; ++k)
        for (int j = rg.d.s_st[1] - 1 + NG; j <= rg.d.s_ed[1] - 1 + NG; ++j)
          for (int i = rg.d.s_st[0] - 1 + NG; i <= rg.d.s_ed[0] - 1 + NG; ++i)
            if (i >= 0 && j >= 0 && k >= 0 && i < B.nx && j < B.ny && k < B.nz) mask[pidx(B, i, j, k)] = 1;
    }
    // the mask lies in the arena, whose zeroing is queued on the stream
    HIPCHK(hipMemcpyAsync(B.frozen, mask.data(), B.P, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

// The viscous face-flux work queue with `seg` faces per marching item:
// along j and k, 256 lines (i fastest) by seg faces; along i, 256 face pairs.
// The marching items come first (they are the longer ones).
static std::vector<ViscItem> visc_items(const orion_hip_ctx* c, int seg) {
  std::vector<ViscItem> v;
  for (int d = 1; d < 3; ++d)
    for (int b = 0; b < c->nb; ++b) {
      const BlockDev& B = c->hb[b];
      const long long lines = (long long)B.idim * (d == 1 ? B.kdim : B.jdim);
      const int faces = dcount(B, d) + 1;
      for (long long a = 0; a < lines; a += 256)
        for (int f = 0; f < faces; f += seg)
          v.push_back(ViscItem{b, d, a, std::min(lines, a + 256), f, std::min(faces, f + seg)});
    }
  for (int b = 0; b < c->nb; ++b) {
    const BlockDev& B = c->hb[b];
    const long long pairs = nlines(B, 0) * ((B.idim + 2) / 2);
    for (long long a = 0; a < pairs; a += 256) v.push_back(ViscItem{b, 0, a, std::min(pairs, a + 256), 0, 0});
  }
  return v;
}

// The fused viscous form's items: per block, tiles of 16 x 16 (i, j)
// columns by `seg` planes along k, the items with the most nodes first.
static std::vector<ViscTile> column_tiles(const orion_hip_ctx* c, int seg) {
  std::vector<ViscTile> v;
  for (int b = 0; b < c->nb; ++b) {
    const BlockDev& B = c->hb[b];
    for (int k0 = 0; k0 < B.kdim; k0 += seg)
      for (int j0 = 0; j0 < B.jdim; j0 += 16)
        for (int i0 = 0; i0 < B.idim; i0 += 16) v.push_back(ViscTile{b, i0, j0, k0, std::min(B.kdim, k0 + seg)});
  }
  auto nodes = [&](const ViscTile& t) {
    const BlockDev& B = c->hb[t.block];
    return (long long)std::min(16, B.idim - t.i0) * std::min(16, B.jdim - t.j0) * (t.k1 - t.k0);
  };
  std::stable_sort(v.begin(), v.end(), [&](const ViscTile& a, const ViscTile& b) { return nodes(a) > nodes(b); });
  return v;
}

static int set_visc_queue(orion_hip_ctx* c, int seg, int grid) {
  if (seg < 1 || grid < 1) return fail("viscous queue: segment and grid must be positive");
  if (c->visc_fused) {
    std::vector<ViscTile> v = column_tiles(c, seg);
    if (v.size() >= (size_t)INT32_MAX) return fail("viscous queue: too many items");
    if ((long long)v.size() > c->vtiles_cap) {
      const long long cap = std::max<long long>((long long)v.size(), (long long)column_tiles(c, 8).size());
      if (dalloc(c, (void**)&c->d_vtiles, sizeof(ViscTile) * cap)) return -1;
      c->vtiles_cap = cap;
    }
    HIPCHK(hipMemcpyAsync(c->d_vtiles, v.data(), sizeof(ViscTile) * v.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->n_vtiles = (int)v.size();
    c->visc_seg = seg;
    c->visc_grid = grid;
    c->graph_stream = nullptr;  // the captured step holds the old count: capture again
    return 0;
  }
  std::vector<ViscItem> v = visc_items(c, seg);
  if (v.size() >= (size_t)INT32_MAX) return fail("viscous queue: too many items");
  if ((long long)v.size() > c->vitems_cap) {
    // sized once for the shortest segment the tuner tries, so later changes reuse it
    const long long cap = std::max<long long>((long long)v.size(), (long long)visc_items(c, 8).size());
    if (dalloc(c, (void**)&c->d_vitems, sizeof(ViscItem) * cap)) return -1;
    c->vitems_cap = cap;
  }
  HIPCHK(hipMemcpyAsync(c->d_vitems, v.data(), sizeof(ViscItem) * v.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->n_vitems = (int)v.size();
  c->visc_seg = seg;
  c->visc_grid = grid;
  c->graph_stream = nullptr;  // the captured step holds the old count: capture again
  return 0;
}

// The chained launch of the four sweeps (lusgs_chain_kernel): one task per
// tile and sweep, dispatched in the order a greedy list schedule on the
// device's CUs (one tile workgroup per CU) starts them.  Units: pipeline
// steps.  A task runs its tile's planes plus two steps of publication lag;
// it may start once its tile's previous sweep has finished and the tiles it
// waits for in its own sweep have published the plane before its first
// (the upstream tile's start + the planes up to it + 3).  Among the tasks
// that may start, the one with the longest remaining path to the end of the
// fourth sweep goes first.  Each task starts strictly after the tasks it
// waits for, so the start order is topological: every wait in the launch
// targets a workgroup dispatched earlier.  The chained schedule needs each
// block's share of the Gauss-Seidel term array (its largest) addressable
// with 32-bit offsets; otherwise it is unavailable.
static int build_chain(orion_hip_ctx* c) {
  c->n_chain = 0;
  const int nt = (int)c->tiles.size();
  for (int b = 0; b < c->nb; ++b) {
    const BlockDev& B = c->hb[b];
    const long long cells = (long long)(B.sn[0] + B.sn[1] + B.sn[2] - 2) * B.sn[2] * B.sncp;
    if (cells * SW_GA * 16 >= (1LL << 32)) return 0;  // unavailable, not an error
  }
  if (nt == 0) return 0;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || ncu <= 0)
    ncu = 256;
  const int NTK = 4 * nt;
  struct Edge { int to; long long off; };
  std::vector<std::vector<Edge>> succ(NTK);
  std::vector<std::vector<std::pair<int, long long>>> pred(NTK);
  auto dur = [&](int t) { const SweepTile& T = c->tiles[t]; return (long long)(T.s_last - T.s_first + 1) + 2; };
  auto add = [&](int a, int b, long long off) {
    succ[a].push_back({b, off});
    pred[b].push_back({a, off});
  };
  for (int x = 0; x < nt; ++x) {
    const SweepTile& X = c->tiles[x];
    for (int k = 0; k < 4; ++k) {
      const int id = 4 * x + k;
      if (k > 0) add(4 * x + k - 1, id, dur(x));
      const bool fwd = (k == 0 || k == 2);
      for (int e = 0; e < 2; ++e) {
        const int u = fwd ? X.nbr[e] : X.nbr[2 + e];
        if (u < 0) continue;
        const SweepTile& U = c->tiles[u];
        if (fwd) {
          const int p = X.s_first - 1;  // the plane X's first edge values lie on
          if (p < U.s_first) continue;
          add(4 * u + k, id, (long long)(std::min(p, U.s_last) - U.s_first) + 3);
        } else {
          const int p = X.s_last + 1;
          if (p > U.s_last) continue;
          add(4 * u + k, id, (long long)(U.s_last - std::max(p, U.s_first)) + 3);
        }
      }
    }
  }
  // remaining path to the end of the last sweep, in reverse topological
  // order (sweep 3 first; within a sweep, tiles in reverse dependency order)
  std::vector<long long> bl(NTK, 0);
  std::vector<int> topo;  // a topological order of the tasks: by sweep, then by tile diagonal
  {
    std::vector<int> ord(nt);
    for (int i = 0; i < nt; ++i) ord[i] = i;
    auto diag = [&](int i) { return c->tiles[i].r0 / SWEEP_TR + c->tiles[i].c0 / SWEEP_TW; };
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return diag(a) < diag(b); });
    for (int k = 0; k < 4; ++k) {
      const bool fwd = (k == 0 || k == 2);
      for (int i = 0; i < nt; ++i) topo.push_back(4 * ord[fwd ? i : nt - 1 - i] + k);
    }
  }
  for (int i = NTK - 1; i >= 0; --i) {
    const int a = topo[i];
    long long v = dur(a / 4);
    for (const Edge& e : succ[a]) v = std::max(v, e.off + bl[e.to]);
    bl[a] = v;
  }
  // greedy list schedule
  std::vector<long long> start(NTK, -1), avail(NTK, 0);
  std::vector<int> npred(NTK), left(NTK);
  for (int i = 0; i < NTK; ++i) npred[i] = left[i] = (int)pred[i].size();
  std::vector<int> ready;  // all predecessors started; may start at avail
  for (int i = 0; i < NTK; ++i)
    if (npred[i] == 0) ready.push_back(i);
  std::vector<long long> running;  // end times of running tasks
  std::vector<int> order;
  order.reserve(NTK);
  long long now = 0;
  while ((int)order.size() < NTK) {
    // free the slots of tasks that have ended
    running.erase(std::remove_if(running.begin(), running.end(), [&](long long e) { return e <= now; }),
                  running.end());
    // start what may start, longest remaining path first
    std::vector<int> cand;
    for (int t : ready)
      if (avail[t] <= now) cand.push_back(t);
    std::stable_sort(cand.begin(), cand.end(), [&](int a, int b) { return bl[a] > bl[b]; });
    bool started = false;
    for (int t : cand) {
      if ((int)running.size() >= ncu) break;
      start[t] = now;
      order.push_back(t);
      running.push_back(now + dur(t / 4));
      ready.erase(std::find(ready.begin(), ready.end(), t));
      for (const Edge& e : succ[t]) {
        avail[e.to] = std::max(avail[e.to], now + e.off);
        if (--left[e.to] == 0) ready.push_back(e.to);
      }
      started = true;
    }
    if (started) continue;
    // advance to the next event: a slot freeing or a ready task becoming available
    long long next = -1;
    for (long long e : running)
      if (next < 0 || e < next) next = e;
    for (int t : ready)
      if (avail[t] > now && (next < 0 || avail[t] < next)) next = avail[t];
    if (next < 0 || next <= now) return fail("chained sweep schedule: no progress at step %lld", now);
    now = next;
  }
  std::vector<int> tasks(NTK);
  for (int i = 0; i < NTK; ++i) tasks[i] = (order[i] / 4) << 2 | (order[i] % 4);
  if (upload_vec(c, &c->d_chain, tasks)) return -1;
  c->n_chain = NTK;
  return 0;
}

// LU-SGS sweep layout (k_lusgs.hip): per block the march axis is the longest
// one and col is i whenever i is not the march axis; the (col, row)
// cross-section is cut into 16 x 16 tiles.  Tiles are launched so that every
// tile's upstream neighbours come first: by increasing (row tile + col tile)
// for the forward sweeps, by decreasing for the backward ones.
static int build_sweep(orion_hip_ctx* c) {
  const int TWD = SWEEP_TW, TRD = SWEEP_TR;
  long long n = 0;
  c->tiles.clear();
  for (int b = 0; b < c->nb; ++b) {
    BlockDev& B = c->hb[b];
    const int d[3] = {B.idim, B.jdim, B.kdim};
    int am = 0;
    for (int a = 1; a < 3; ++a)
      if (d[a] > d[am]) am = a;
    int ac, ar;
    if (am != 0) {
      ac = 0;
      ar = 3 - am;
    } else {
      ac = d[1] >= d[2] ? 1 : 2;
      ar = 3 - ac;
    }
    B.sax[0] = am; B.sax[1] = ac; B.sax[2] = ar;
    B.sn[0] = d[am]; B.sn[1] = d[ac]; B.sn[2] = d[ar];
    B.sncp = (d[ac] + TWD - 1) / TWD * TWD;
    B.sbase = n;
    const long long planes = (long long)d[0] + d[1] + d[2] - 2;
    if (planes >= 60000) return fail("block %d: %lld hyperplanes exceed the sweep limit", b + 1, planes);
    // the sweep stores a block's dq through a buffer resource with 32-bit byte offsets
    if (planes * B.sn[2] * ((B.sn[1] + TWD - 1) / TWD * TWD) * SW_DQ * 16 >= (1LL << 32))
      return fail("block %d: its LU-SGS sweep arrays exceed 4 GiB", b + 1);
    n += planes * B.sn[2] * B.sncp;
    const int tc = (B.sn[1] + TWD - 1) / TWD, tr = (B.sn[2] + TRD - 1) / TRD;
    const int first = (int)c->tiles.size();
    for (int R = 0; R < tr; ++R)
      for (int C = 0; C < tc; ++C) {
        SweepTile t{};
        t.block = b;
        t.r0 = R * TRD;
        t.c0 = C * TWD;
        const int rhi = std::min(t.r0 + TRD - 1, B.sn[2] - 1), chi = std::min(t.c0 + TWD - 1, B.sn[1] - 1);
        t.s_first = t.r0 + t.c0;
        t.s_last = rhi + chi + B.sn[0] - 1;
        t.nbr[0] = C > 0 ? first + R * tc + C - 1 : -1;
        t.nbr[1] = R > 0 ? first + (R - 1) * tc + C : -1;
        t.nbr[2] = C + 1 < tc ? first + R * tc + C + 1 : -1;
        t.nbr[3] = R + 1 < tr ? first + (R + 1) * tc + C : -1;
        c->tiles.push_back(t);
      }
  }
  n = (n + 31) / 32 * 32;
  const int nt = (int)c->tiles.size();
  // Launch orders of the tiles (forward sweeps, backward sweeps).  Every
  // order launches a tile after the upstream tiles it waits for, so the
  // persistent grid drains whatever the dispatch.
  //   0 diagonal: by tile diagonal R + C over all blocks, reversed for the
  //     backward sweeps;
  //   1 critical path first: by the planes left from the tile's start to
  //     the end of its block's sweep (forward: last plane of the block minus
  //     the tile's first; backward: the tile's last plane), longest first,
  //     so that the tiles of the longest chains take the CUs first when the
  //     tiles outnumber them.  Upstream tiles have more planes left.
  {
    std::vector<int> ord(nt);
    for (int i = 0; i < nt; ++i) ord[i] = i;
    auto diag = [&](int i) { return c->tiles[i].r0 / TRD + c->tiles[i].c0 / TWD; };
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return diag(a) < diag(b); });
    c->sched_order[0][0] = ord;
    std::reverse(ord.begin(), ord.end());
    c->sched_order[0][1] = ord;
    auto last_plane = [&](int i) {
      const BlockDev& B = c->hb[c->tiles[i].block];
      return B.sn[0] + B.sn[1] + B.sn[2] - 3;
    };
    for (int i = 0; i < nt; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
      return last_plane(a) - c->tiles[a].s_first > last_plane(b) - c->tiles[b].s_first;
    });
    c->sched_order[1][0] = ord;
    for (int i = 0; i < nt; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return c->tiles[a].s_last > c->tiles[b].s_last; });
    c->sched_order[1][1] = ord;
  }
  // per-block launches: each block's tiles by diagonal, at the positions of
  // its tiles (forward sweeps; reversed for the backward ones)
  {
    c->tile_first.assign(c->nb, 0);
    c->tile_count.assign(c->nb, 0);
    std::vector<int> f(nt), bk(nt);
    for (int i = nt - 1; i >= 0; --i) c->tile_first[c->tiles[i].block] = i;
    for (int i = 0; i < nt; ++i) c->tile_count[c->tiles[i].block]++;
    for (int b = 0; b < c->nb; ++b) {
      const int f0 = c->tile_first[b], n = c->tile_count[b];
      std::vector<int> o(n);
      for (int i = 0; i < n; ++i) o[i] = f0 + i;
      auto diag = [&](int i) { return c->tiles[i].r0 / TRD + c->tiles[i].c0 / TWD; };
      std::stable_sort(o.begin(), o.end(), [&](int a, int b2) { return diag(a) < diag(b2); });
      for (int i = 0; i < n; ++i) {
        f[f0 + i] = o[i];
        bk[f0 + i] = o[n - 1 - i];
      }
    }
    if (upload_vec(c, &c->d_border[0], f) || upload_vec(c, &c->d_border[1], bk)) return -1;
  }
  if (build_chain(c)) return -1;
  // the four sweeps chained where they can be, else critical path first,
  // until a tuner says otherwise (d_order holds the critical-path orders
  // either way)
  c->schedule = c->n_chain > 0 ? ORION_SCHEDULE_CHAINED : 1;
  if (upload_vec(c, &c->d_order[0], c->sched_order[1][0])) return -1;
  if (upload_vec(c, &c->d_order[1], c->sched_order[1][1])) return -1;
  if (upload_vec(c, &c->d_tiles, c->tiles)) return -1;
  c->sw.n = n;
  if (dalloc(c, (void**)&c->sw.own, sizeof(double2) * SW_OWN * n) ||
      dalloc(c, (void**)&c->sw.D, sizeof(double2) * SW_D * n) ||
      dalloc(c, (void**)&c->sw.dq, sizeof(double2) * SW_DQ * n) ||
      dalloc(c, (void**)&c->sw.ga, sizeof(double2) * SW_GA * n) ||
      dalloc(c, (void**)&c->sw.pre, sizeof(double2) * SW_PRE * n) ||
      dalloc(c, (void**)&c->d_flags, sizeof(unsigned long long) * FLAG_STRIDE * std::max(nt, 1)) ||
      dalloc(c, (void**)&c->d_epoch, sizeof(unsigned long long)) || dalloc(c, (void**)&c->d_err, sizeof(int)))
    return -1;
  HIPCHK(hipMemsetAsync(c->sw.dq, 0, sizeof(double2) * SW_DQ * n, c->stream));
  HIPCHK(hipMemsetAsync(c->sw.own, 0, sizeof(double2) * SW_OWN * n, c->stream));
  HIPCHK(hipMemsetAsync(c->sw.D, 0, sizeof(double2) * SW_D * n, c->stream));
  HIPCHK(hipMemsetAsync(c->sw.ga, 0, sizeof(double2) * SW_GA * n, c->stream));
  HIPCHK(hipMemsetAsync(c->sw.pre, 0, sizeof(double2) * SW_PRE * n, c->stream));
  HIPCHK(hipMemsetAsync(c->d_flags, 0, sizeof(unsigned long long) * FLAG_STRIDE * std::max(nt, 1), c->stream));
  HIPCHK(hipMemsetAsync(c->d_epoch, 0, sizeof(unsigned long long), c->stream));
  HIPCHK(hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* orion_hip_last_error(void) { return g_err.c_str(); }

int orion_hip_backend_init(int device) {
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail("device %d out of range (%d devices)", device, n);
  HIPCHK(hipSetDevice(device));
  g_device = device;
  return 0;
}

// Streams of destroyed contexts, per device, handed to the next contexts
// (creating a stream takes milliseconds, which a serving batch rebuilt at
// every chunk boundary would pay twice); used under g_capture_mutex.
static std::map<int, std::vector<hipStream_t>> g_stream_pool;

static int take_stream(int device, hipStream_t* s) {
  std::vector<hipStream_t>& v = g_stream_pool[device];
  if (!v.empty()) {
    *s = v.back();
    v.pop_back();
    return 0;
  }
  return hipStreamCreateWithFlags(s, hipStreamNonBlocking) == hipSuccess ? 0 : -1;
}

static void give_stream(int device, hipStream_t s) {
  if (!s) return;
  if (hipStreamSynchronize(s) != hipSuccess) {
    hipStreamDestroy(s);
    return;
  }
  g_stream_pool[device].push_back(s);
}

void orion_hip_ctx_destroy(orion_hip_ctx* c) {
  if (!c) return;
  std::lock_guard<std::recursive_mutex> guard(g_capture_mutex);
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  for (void* p : c->allocs) hipFree(p);
  if (c->arena) hipFree(c->arena);
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->graph) hipGraphDestroy(c->graph);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  give_stream(c->device, c->side_stream);
  give_stream(c->device, c->own_stream);
  delete c;
}

// The requests of a context, checked on the host before any device work:
// block counts that add up, the inputs that select kernels or are
// constants of the discretisation shared, and no interface between two
// requests' blocks.
static int check_requests(int nreq, const int* req_nblocks, int nblocks, int nregions,
                          const orion_region_desc* regions, const orion_inputs* in) {
  if (nreq <= 0 || !req_nblocks || !in) return fail("a context needs at least one request");
  std::vector<int> owner;
  for (int r = 0; r < nreq; ++r) {
    if (req_nblocks[r] <= 0) return fail("request %d has no blocks", r);
    owner.insert(owner.end(), req_nblocks[r], r);
    const orion_inputs& a = in[r];
    const orion_inputs& z = in[0];
    if (a.nvis != z.nvis || a.method != z.method || a.csrv != z.csrv || a.ntmst != z.ntmst || a.gcl != z.gcl)
      return fail("request %d: a batch's requests must share nvis, method, csrv, ntmst and gcl", r);
  }
  if ((int)owner.size() != nblocks) return fail("the requests' block counts do not add up to nblocks");
  for (int q = 0; q < nregions; ++q) {
    const orion_region_desc& d = regions[q];
    if (d.bctype >= 0 || d.block < 0 || d.block >= nblocks || d.nbt < 1 || d.nbt > nblocks) continue;
    if (owner[d.nbt - 1] != owner[d.block])
      return fail("block %d region %d: an interface joins blocks of requests %d and %d", d.block + 1, q + 1,
                  owner[d.block], owner[d.nbt - 1]);
  }
  return 0;
}

// xyz: every block's coordinates one after another; or, with req_xyz,
// request r's blocks' coordinates one after another at req_xyz[r]
static int create_ctx(int nreq, const int* req_nblocks, int nblocks, const int* dims, const double* xyz,
                      const double* const* req_xyz, int nregions, const orion_region_desc* regions,
                      const orion_inputs* in, int keep_q, orion_hip_ctx** out) {
  // ORION_HIP_TIMING=1: the set-up's phases on stderr (each waits for the
  // device work queued before it)
  static const bool timed = std::getenv("ORION_HIP_TIMING") && std::atoi(std::getenv("ORION_HIP_TIMING")) != 0;
  auto tlast = std::chrono::steady_clock::now();
  hipStream_t tstream = nullptr;
  auto timing = [&](const char* what) {
    if (!timed) return;
    if (tstream) hipStreamSynchronize(tstream);
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[orion_hip] set-up %-40s %8.1f ms\n", what,
                 std::chrono::duration<double, std::milli>(now - tlast).count());
    tlast = now;
  };
  *out = nullptr;
  if (nblocks <= 0) return fail("nblocks must be positive");
  if (nregions > 0 && !regions) return fail("null regions");
  if (check_requests(nreq, req_nblocks, nblocks, nregions, regions, in)) return -1;
  std::lock_guard<std::recursive_mutex> guard(g_capture_mutex);
  orion_hip_ctx* c = new orion_hip_ctx();
  auto bail = [&](int rc) {
    orion_hip_ctx_destroy(c);
    return rc;
  };
  c->device = g_device;
  if (hipSetDevice(c->device) != hipSuccess) return bail(fail("hipSetDevice(%d) failed", c->device));
  // default-priority streams
  if (take_stream(c->device, &c->own_stream) || take_stream(c->device, &c->side_stream))
    return bail(fail("hipStreamCreate failed"));
  if (hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess)
    return bail(fail("hipEventCreate failed"));
  c->stream = c->own_stream;
  tstream = c->stream;
  timing("streams");
  c->nb = nblocks;
  c->keep_q = keep_q;
  init_params(c->p, in);
  // requests: their blocks and inputs; what selects kernels or is a
  // constant of the discretisation must agree across a batch
  c->nreq = nreq;
  c->req_first.assign(nreq + 1, 0);
  std::vector<Params> preq(nreq);
  for (int r = 0; r < nreq; ++r) {
    c->req_first[r + 1] = c->req_first[r] + req_nblocks[r];
    init_params(preq[r], &in[r]);
  }
  // blocks
  c->hb.resize(nblocks);
  std::vector<long long> lines_tot(nblocks);
  for (int b = 0; b < nblocks; ++b) {
    BlockDev& B = c->hb[b];
    B = BlockDev{};
    B.idim = dims[3 * b];
    B.jdim = dims[3 * b + 1];
    B.kdim = dims[3 * b + 2];
    if (B.idim <= 2 * NG || B.jdim <= 2 * NG || B.kdim <= 2 * NG)
      return bail(fail("block %d too small (%d,%d,%d): every dimension must exceed %d (MetricComputer.cpp:154)", b,
                       B.idim, B.jdim, B.kdim, 2 * NG));
    B.nx = B.idim + 2 * NG;
    B.ny = B.jdim + 2 * NG;
    B.nz = B.kdim + 2 * NG;
    B.P = (long long)B.nx * B.ny * B.nz;
    if (B.P >= MAX_BLOCK_POINTS)
      return bail(fail("block %d: %lld padded points (kernels address a block's fields with 32-bit offsets: "
                       "fewer than %lld accepted)", b, B.P, MAX_BLOCK_POINTS));
    lines_tot[b] = nlines(B, 0) + nlines(B, 1) + nlines(B, 2);
    int r = 0;
    while (c->req_first[r + 1] <= b) ++r;
    B.req = r;
    B.par = preq[r];
    const double cp = 1.0 / ((B.par.gama - 1.0) * B.par.moo * B.par.moo);
    B.cp_prl = cp / B.par.prl;
    B.re_inv = (std::fabs(B.par.reynolds) > 1.0e-30) ? 1.0 / B.par.reynolds : 0.0;
  }
  // arena
  long long total = 0;
  for (int b = 0; b < nblocks; ++b) total += arena_fields(c->hb[b], keep_q, c->p.nvis == 1, lines_tot[b]);
  c->arena_bytes = (size_t)total * sizeof(double);
  if (hipMalloc((void**)&c->arena, c->arena_bytes) != hipSuccess)
    return bail(fail("arena hipMalloc(%zu bytes) failed", c->arena_bytes));
  if (hipMemsetAsync(c->arena, 0, c->arena_bytes, c->stream) != hipSuccess) return bail(fail("arena memset failed"));
  timing("arena allocation and zeroing");
  {
    double* a = (double*)c->arena;
    auto al = [](long long x) { return (x + 31) / 32 * 32; };
    for (int b = 0; b < nblocks; ++b) {
      BlockDev& B = c->hb[b];
      const long long P = B.P;
      B.prim = a; a += al(6 * P);
      B.met = a; a += al(9 * P);
      B.vol = a; a += al(P);
      B.dq = a; a += al(5 * P);
      B.rhs = a; a += al(5 * P);
      B.dt = a; a += al(P);
      B.c = a; a += al(P);
      B.mu = a; a += al(P);
      B.q = nullptr;
      if (keep_q) { B.q = a; a += al(5 * P); }
      B.duvwt = a; a += al(12 * P);
      B.fx = a; a += al(15 * P);
      B.fcov = a; a += al(30 * lines_tot[b]);
      B.fcov_off[0] = 0;
      B.fcov_off[1] = 30 * nlines(B, 0);
      B.fcov_off[2] = 30 * (nlines(B, 0) + nlines(B, 1));
      B.frozen = (unsigned char*)a; a += al((P + 7) / 8);
      B.fgeo = a; a += al(3 * FG_N * P);
      // the inviscid fluxes get their own array when a viscous flux runs
      // beside them; otherwise they reuse fx
      B.fxi = B.fx;
      if (c->p.nvis == 1) { B.fxi = a; a += al(15 * P); }
    }
  }
  // regions
  std::vector<std::vector<Reg>> R(nblocks);
  for (int r = 0; r < nregions; ++r) {
    const orion_region_desc& d = regions[r];
    if (d.block < 0 || d.block >= nblocks) return bail(fail("region %d: bad block %d", r, d.block));
    Reg rg;
    rg.d = d;
    if (d.bctype >= 0) {
      rg.d.nbt = -1;
      rg.d.ibcwin = 0;
    }
    const int bt = d.bctype;
    if (!(bt == -1 || bt == 2 || bt == 3 || bt == 4 || bt == 5 || bt == 6))
      return bail(fail("block %d: bctype %d not among {-1,4,5,6,2,3} (BCPreprocess.cpp:196)", d.block + 1, bt));
    const BlockDev& B = c->hb[d.block];
    const int dm[3] = {B.idim, B.jdim, B.kdim};
    // windows must lie inside their blocks (the reference does not check;
    // here an out-of-range window would address outside the arena)
    for (int m = 0; m < 3; ++m) {
      const int a = std::abs(d.s_st[m]), e = std::abs(d.s_ed[m]);
      if (a < 1 || e < 1 || a > dm[m] || e > dm[m])
        return bail(fail("block %d region %d: window outside the block", d.block + 1, r + 1));
      if (bt < 0) {
        if (d.nbt < 1 || d.nbt > nblocks) return bail(fail("block %d region %d: bad target block", d.block + 1, r + 1));
        const BlockDev& T = c->hb[d.nbt - 1];
        const int tm[3] = {T.idim, T.jdim, T.kdim};
        const int ta = std::abs(d.t_st[m]), te = std::abs(d.t_ed[m]);
        if (ta < 1 || te < 1 || ta > tm[m] || te > tm[m])
          return bail(fail("block %d region %d: target window outside block %d", d.block + 1, r + 1, d.nbt));
      }
    }
    for (int m = 0; m < 3; ++m)
      if (d.s_st[m] == d.s_ed[m]) {
        rg.s_nd = m + 1;
        rg.s_lr = (d.s_st[m] == 1) ? -1 : (d.s_st[m] == dm[m] ? 1 : 0);
        break;
      }
    if (bt < 0) {
      if (rg.s_nd == 0 || d.ibcwin <= 0 || d.nbt < 1 || d.nbt > nblocks)
        return bail(fail("block %d: invalid interface region", d.block + 1));
      if (analyze_connect(rg)) return bail(-1);
    }
    R[d.block].push_back(std::move(rg));
  }
  for (int b = 0; b < nblocks; ++b)
    for (auto& rg : R[b])
      if (rg.d.bctype < 0) {
        const int tb = rg.d.nbt - 1, tw = rg.d.ibcwin - 1;
        if (tw >= (int)R[tb].size()) return bail(fail("block %d: target window out of range", b + 1));
      }
  // tables
  std::vector<long long> cp(nblocks), ci(nblocks), cf[3], cl[3];
  for (int d = 0; d < 3; ++d) {
    cf[d].resize(nblocks);
    cl[d].resize(nblocks);
  }
  for (int b = 0; b < nblocks; ++b) {
    const BlockDev& B = c->hb[b];
    cp[b] = B.P;
    ci[b] = (long long)B.idim * B.jdim * B.kdim;
    for (int d = 0; d < 3; ++d) {
      cl[d][b] = nlines(B, d);
      cf[d][b] = nlines(B, d) * (dcount(B, d) + 1);
    }
  }
  if (make_space(c, c->pad, cp) || make_space(c, c->intr, ci)) return bail(-1);
  for (int d = 0; d < 3; ++d)
    if (make_space(c, c->face[d], cf[d]) || make_space(c, c->line[d], cl[d])) return bail(-1);
  for (int d = 0; d < 3; ++d) {
    std::vector<long long> cp2(nblocks);
    for (int b = 0; b < nblocks; ++b) cp2[b] = nlines(c->hb[b], d) * ((dcount(c->hb[b], d) + 2) / 2);
    if (make_space(c, c->fpair[d], cp2)) return bail(-1);
    for (int b = 0; b < nblocks; ++b) cp2[b] = nlines(c->hb[b], d) * 6;
    if (make_space(c, c->ltask[d], cp2)) return bail(-1);
  }
  {
    std::vector<long long> f3, p3, l3;
    for (int d = 0; d < 3; ++d)
      for (int b = 0; b < nblocks; ++b) {
        f3.push_back(cf[d][b]);
        p3.push_back(c->fpair[d].h[b + 1] - c->fpair[d].h[b]);
        l3.push_back(c->ltask[d].h[b + 1] - c->ltask[d].h[b]);
      }
    if (make_space(c, c->face3, f3) || make_space(c, c->fpair3, p3) || make_space(c, c->ltask3, l3)) return bail(-1);
  }
  {
    std::vector<long long> ct(nblocks);
    for (int b = 0; b < nblocks; ++b) {
      const BlockDev& B = c->hb[b];
      ct[b] = (long long)((B.idim + DIFF_TI - 1) / DIFF_TI) * ((B.jdim + DIFF_TJ - 1) / DIFF_TJ) *
              ((B.kdim + DIFF_TK - 1) / DIFF_TK) * 256;
    }
    if (make_space(c, c->dtile, ct)) return bail(-1);
  }
  timing("regions and iteration spaces");
  if (build_sweep(c)) return bail(-1);
  timing("sweep layout and tile orders");
  {
    std::vector<long long> ct(nblocks);
    for (int b = 0; b < nblocks; ++b) {
      const BlockDev& B = c->hb[b];
      ct[b] = (long long)((B.sn[0] + 15) / 16) * ((B.sn[1] + 15) / 16) * B.sn[2] * 256;
    }
    if (make_space(c, c->ptile, ct)) return bail(-1);
  }
  {
    // the zero workgroup tables of one-block launches: as long as the most
    // workgroups any block has in any space
    long long most = 1;
    std::vector<const Space*> all = {&c->pad, &c->intr, &c->dtile, &c->ptile};
    for (int d = 0; d < 3; ++d) {
      all.push_back(&c->face[d]);
      all.push_back(&c->fpair[d]);
      all.push_back(&c->ltask[d]);
    }
    for (const Space* sp : all)
      for (int b = 0; b < nblocks; ++b) most = std::max(most, nwg_of(*sp, b));
    if (upload_vec(c, &c->d_zeros, std::vector<int>((size_t)most, 0))) return bail(-1);
  }
  if (upload_vec(c, &c->db, c->hb)) return bail(-1);
  if (c->p.nvis == 1) {
    c->visc_fused_ok = true;
    for (const BlockDev& B : c->hb) c->visc_fused_ok = c->visc_fused_ok && B.idim >= 4 && B.jdim >= 4 && B.kdim >= 4;
    c->visc_fused = false;  // the face-flux queue measured faster; the tuner tries both
    if (dalloc(c, (void**)&c->d_vhead, sizeof(int)) || set_visc_queue(c, c->visc_seg, c->visc_grid)) return bail(-1);
  }
  timing("block table, viscous work queue");
  if (build_topology(c, R, timing)) return bail(-1);
  timing("boundary topology (halo, BC, averaging, masks)");
  c->nparts = residual_parts(c->intr.total());
  if (c->ptile.nwg >= (1LL << 31)) return bail(fail("too many update workgroups"));
  if (dalloc(c, (void**)&c->d_part, sizeof(double) * 2 * 1024) || dalloc(c, (void**)&c->d_pidx, sizeof(long long) * 1024) ||
      dalloc(c, (void**)&c->d_upart, sizeof(double) * 2 * c->ptile.nwg) ||
      dalloc(c, (void**)&c->d_ukey, sizeof(long long) * c->ptile.nwg) ||
      dalloc(c, (void**)&c->d_count, sizeof(unsigned long long) * 2) || dalloc(c, (void**)&c->d_step, sizeof(int)))
    return bail(-1);
  {
    // each request's update partials (whole workgroups per block, in tile
    // order) and residual term count
    std::vector<int> seg(nreq + 1, 0);
    std::vector<long long> tot(nreq, 0);
    for (int r = 0; r < nreq; ++r) {
      long long w = 0;
      for (int b = c->req_first[r]; b < c->req_first[r + 1]; ++b)
        w += (c->ptile.h[b + 1] - c->ptile.h[b] + SPAN_TB - 1) / SPAN_TB;
      seg[r + 1] = seg[r] + (int)w;
      tot[r] = 5 * (c->intr.h[c->req_first[r + 1]] - c->intr.h[c->req_first[r]]);
    }
    if (seg[nreq] != c->ptile.nwg) return bail(fail("internal: request partials do not cover the update"));
    if (upload_vec(c, &c->d_rseg, seg) || upload_vec(c, &c->d_rtot, tot) || dalloc(c, (void**)&c->d_rdone, sizeof(int)))
      return bail(-1);
    if (hipMemsetAsync(c->d_rdone, 0, sizeof(int), c->stream) != hipSuccess) return bail(fail("memset"));
  }
  if (const char* e = std::getenv("ORION_HIP_GRAPH")) c->use_graph = std::atoi(e) != 0;
  // initial state (InitialCondition.cpp:49-80): freestream prim everywhere,
  // each block with its request's freestream
  const Params& p = c->p;
  for (int b = 0; b < nblocks; ++b) {
    const BlockDev& B = c->hb[b];
    const Params& f = B.par;
    const double prim0[6] = {f.roo, f.uoo, f.voo, f.woo, f.poo, f.too};
    const double v2 = f.uoo * f.uoo + f.voo * f.voo + f.woo * f.woo;
    const double q0[5] = {f.roo, f.roo * f.uoo, f.roo * f.voo, f.roo * f.woo, f.poo / (f.gama - 1.0) + 0.5 * f.roo * v2};
    for (int m = 0; m < 6; ++m) launch_fill(B.prim + m * B.P, B.P, prim0[m], c->stream);
    if (B.q)
      for (int m = 0; m < 5; ++m) launch_fill(B.q + m * B.P, B.P, q0[m], c->stream);
  }
  // grid -> metrics on the device
  {
    // each block's coordinates as given (no padding) to a staging part of
    // its face-flux scratch, then into the padded layout on the device
    long long off = 0;
    int r = 0, rb = 0;  // the block's request and its place in it
    for (int b = 0; b < nblocks; ++b) {
      const BlockDev& B = c->hb[b];
      const long long n = (long long)B.idim * B.jdim * B.kdim;
      if (req_xyz) {
        while (rb == req_nblocks[r]) {
          ++r;
          rb = 0;
          off = 0;
        }
        ++rb;
      }
      const double* src = req_xyz ? req_xyz[r] + off : xyz + off;
      if (!src) return bail(fail("null grid for request %d", r));
      if (hipMemcpyAsync(B.fx + 3 * B.P, src, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream) !=
          hipSuccess)
        return bail(fail("grid upload failed"));
      off += 3 * n;
    }
    launch_grid_scatter(LA(c, c->intr), c->intr.total());
    timing("grid upload");
    launch_metrics(LA(c, c->intr), c->intr.total(), p.gcl, nullptr, c->stream);
    if (hipMemsetAsync(c->d_count, 0, sizeof(unsigned long long), c->stream) != hipSuccess) return bail(fail("memset"));
    launch_count_small_vol(LA(c, c->intr), c->intr.total(), c->d_count);
    unsigned long long bad = 0;
    if (hipMemcpyAsync(&bad, c->d_count, sizeof bad, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      return bail(fail("metric check failed: %s", hipGetErrorString(hipGetLastError())));
    if (bad) return bail(fail("ERROR(Jacobi<0): %llu points with volume below 1e-15 (MetricComputer.cpp:488-491)", bad));
    for (int d = 0; d < 3; ++d) launch_face_geom(LA(c, c->face[d]), d, c->face[d].total());
    launch_lu_geom(LA(c, c->ptile), c->ptile.total(), c->sw);
  }
  if (p.nvis != 0) launch_temperature(LA(c, c->pad), c->pad.total());
  if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(fail("setup failed: %s", hipGetErrorString(hipGetLastError())));
  timing("metrics, face geometry, initial state");
  *out = c;
  return 0;
}

int orion_hip_ctx_create(int nblocks, const int* dims, const double* xyz, int nregions,
                         const orion_region_desc* regions, const orion_inputs* in, int keep_q,
                         orion_hip_ctx** out) {
  return create_ctx(1, &nblocks, nblocks, dims, xyz, nullptr, nregions, regions, in, keep_q, out);
}

int orion_hip_batch_create(int nreq, const int* req_nblocks, int nblocks, const int* dims, const double* xyz,
                           int nregions, const orion_region_desc* regions, const orion_inputs* inputs, int keep_q,
                           orion_hip_ctx** out) {
  if (!out) return fail("batch_create: null output");
  return create_ctx(nreq, req_nblocks, nblocks, dims, xyz, nullptr, nregions, regions, inputs, keep_q, out);
}

int orion_hip_batch_create_grids(int nreq, const int* req_nblocks, int nblocks, const int* dims,
                                 const double* const* req_xyz, int nregions, const orion_region_desc* regions,
                                 const orion_inputs* inputs, int keep_q, orion_hip_ctx** out) {
  if (!out || !req_xyz) return fail("batch_create_grids: null output or grids");
  return create_ctx(nreq, req_nblocks, nblocks, dims, nullptr, req_xyz, nregions, regions, inputs, keep_q, out);
}

int orion_hip_requests(const orion_hip_ctx* c) { return c ? c->nreq : 0; }

static int field_info(orion_hip_ctx* c, int b, int f, double** ptr, long long* n) {
  if (!c || b < 0 || b >= c->nb) return fail("bad context or block");
  const BlockDev& B = c->hb[b];
  switch (f) {
    case ORION_FIELD_PRIM: *ptr = B.prim; *n = 6 * B.P; break;
    case ORION_FIELD_METRICS: *ptr = B.met; *n = 9 * B.P; break;
    case ORION_FIELD_VOL: *ptr = B.vol; *n = B.P; break;
    case ORION_FIELD_Q:
      if (!B.q) return fail("context created without keep_q");
      *ptr = B.q; *n = 5 * B.P; break;
    case ORION_FIELD_DQ: *ptr = B.dq; *n = 5 * B.P; break;
    case ORION_FIELD_DT: *ptr = B.dt; *n = B.P; break;
    case ORION_FIELD_C: *ptr = B.c; *n = B.P; break;
    case ORION_FIELD_MU: *ptr = B.mu; *n = B.P; break;
    default: return fail("unknown field %d", f);
  }
  return 0;
}

int orion_hip_field(orion_hip_ctx* c, int b, int f, double** ptr, long long* n) { return field_info(c, b, f, ptr, n); }

int orion_hip_upload(orion_hip_ctx* c, int b, int f, const double* host, long long count) {
  double* p;
  long long n;
  if (field_info(c, b, f, &p, &n)) return -1;
  if (count != n) return fail("upload: field %d of block %d holds %lld values, got %lld", f, b, n, count);
  HIPCHK(hipMemcpyAsync(p, host, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int orion_hip_copy_block_state(orion_hip_ctx* dst, int dst_block, orion_hip_ctx* src, int src_block) {
  if (!dst || !src) return fail("copy_block_state: null context");
  if (dst->device != src->device) return fail("copy_block_state: the contexts are on different devices");
  double *pd, *ps;
  long long nd, ns;
  if (field_info(dst, dst_block, ORION_FIELD_PRIM, &pd, &nd) || field_info(src, src_block, ORION_FIELD_PRIM, &ps, &ns))
    return -1;
  const BlockDev &D = dst->hb[dst_block], &S = src->hb[src_block];
  if (D.idim != S.idim || D.jdim != S.jdim || D.kdim != S.kdim)
    return fail("copy_block_state: block sizes differ (%d,%d,%d) vs (%d,%d,%d)", D.idim, D.jdim, D.kdim, S.idim,
                S.jdim, S.kdim);
  // the source's queued steps first, then the copy in the destination's order
  HIPCHK(hipStreamSynchronize(src->stream));
  HIPCHK(hipMemcpyAsync(pd, ps, sizeof(double) * nd, hipMemcpyDeviceToDevice, dst->stream));
  HIPCHK(hipStreamSynchronize(dst->stream));
  return 0;
}

int orion_hip_download(orion_hip_ctx* c, int b, int f, double* host, long long count) {
  double* p;
  long long n;
  if (field_info(c, b, f, &p, &n)) return -1;
  if (count != n) return fail("download: field %d of block %d holds %lld values, got %lld", f, b, n, count);
  HIPCHK(hipMemcpyAsync(host, p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

static int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail("%s: kernel launch failed: %s", what, hipGetErrorString(e));
  return 0;
}

int orion_hip_update_temperature(orion_hip_ctx* c) {
  if (c->p.nvis == 0) return 0;
  launch_temperature(LA(c, c->pad), c->pad.total());
  return check_launch("update_temperature");
}

// the state kernel over the padded points (per block when per_block)
static void op_state(orion_hip_ctx* c, int mode, unsigned long long* dtmin, const int* step) {
  if (c->per_block) {
    for (int b = 0; b < c->nb; ++b) launch_state(LAB(c, c->pad, b), items_of(c->pad, b), mode, dtmin, step, c->nreq);
    return;
  }
  launch_state(LA(c, c->pad), c->pad.total(), mode, dtmin, step, c->nreq);
}

// the blocks that receive whole points, then those that receive field
// values (their destinations are disjoint)
static void op_halo(orion_hip_ctx* c) {
  launch_halo_points(c->d_halo_pt, c->n_halo_pt, (double*)c->arena, c->stream);
  launch_halo(c->d_halo, c->n_halo, (double*)c->arena, c->stream);
}

static void op_bc(orion_hip_ctx* c) {
  for (size_t rk = 0; rk < c->bc_first.size(); ++rk)
    launch_bc(c->db, c->d_bc + c->bc_first[rk], c->bc_count[rk], c->bc_total[rk], c->stream);
}

// first: the step's first right-hand-side kernel, which writes dq whole
// (the step's state kernel skipped the zeroing); the per-operator calls
// accumulate into dq as the reference does
static void op_visc(orion_hip_ctx* c, bool first = false) {
  if (c->p.nvis != 1) return;
  if (c->per_block) {
    for (int b = 0; b < c->nb; ++b) {
      launch_visc_deriv(LAB(c, c->intr, b), items_of(c->intr, b));
      for (int d = 0; d < 3; ++d) launch_visc_face(LAB(c, c->fpair[d], b), d, items_of(c->fpair[d], b));
      launch_visc_diff(LAB(c, c->dtile, b), items_of(c->dtile, b), first);
    }
    return;
  }
  launch_visc_deriv(LA(c, c->intr), c->intr.total(), c->d_vhead);
  if (c->visc_fused) {
    launch_visc_fused(c->db, c->d_vtiles, c->n_vtiles, c->d_vhead, c->visc_grid, first, c->stream);
    return;
  }
  launch_visc_face_queue(c->db, c->d_vitems, c->n_vitems, c->d_vhead, c->visc_grid, c->stream);
  launch_visc_diff(LA(c, c->dtile), c->dtile.total(), first);
}

static void op_avg(orion_hip_ctx* c) {
  launch_avg_stage(c->db, c->d_avg_src, c->n_avg_src, c->d_stage, c->stream);
  launch_avg_apply(c->db, c->d_avg_dst, c->n_avg_dst, c->d_stage, c->stream);
}

// Inviscid face fluxes and boundary closures: they read the flow state
// and the static face geometry only, and write fxi and fcov.
static void op_inv_faces(orion_hip_ctx* c, hipStream_t st) {
  launch_inv_face_pairs(LA3(c, c->fpair3, st), c->fpair3.total(), c->p.gama);
  launch_inv_closure(LA3(c, c->ltask3, st), -1, c->ltask3.total(), c->p.gama);
}

static void op_inv(orion_hip_ctx* c, bool first = false) {
  if (c->per_block) {
    for (int b = 0; b < c->nb; ++b) {
      for (int d = 0; d < 3; ++d) {
        launch_inv_face(LAB(c, c->face[d], b), d, items_of(c->face[d], b), c->p.gama);
        launch_inv_closure(LAB(c, c->ltask[d], b), d, items_of(c->ltask[d], b), c->p.gama);
      }
      launch_inv_diff(LAB(c, c->dtile, b), items_of(c->dtile, b), first);
    }
    return;
  }
  op_inv_faces(c, c->stream);
  launch_inv_diff(LA(c, c->dtile), c->dtile.total(), first);
}

static void op_sweep(orion_hip_ctx* c, int type) {
  const int bwd = (type == 1 || type == 3) ? 1 : 0;
  launch_lusgs_tiles(c->sw, c->db, c->d_tiles, c->d_order[bwd], (int)c->tiles.size(), c->d_flags, c->d_epoch,
                     c->d_err, type, c->p, c->stream);
}

// the four sweeps: one chained launch, or a launch per sweep
static void op_sweeps(orion_hip_ctx* c) {
  if (c->schedule == ORION_SCHEDULE_CHAINED) {
    launch_lusgs_chain(c->sw, c->db, c->d_tiles, c->d_chain, c->n_chain, c->d_flags, c->d_epoch, c->d_err, c->p,
                       c->stream);
    return;
  }
  for (int t = 0; t < 4; ++t) op_sweep(c, t);
}

// state_packed: the step already formed the pack's state part (with the
// interior flow states, state_pack_kernel)
static void op_update(orion_hip_ctx* c, bool state_packed = false) {
  if (c->per_block) {
    // the reference's update_solution loop: pack, four sweeps and the
    // limited update of one block, then the next (StateUpdater.cpp:544-580);
    // the partials go to the block's own slots of the step's partials
    for (int b = 0; b < c->nb; ++b) {
      launch_lu_pack(LAB(c, c->ptile, b), items_of(c->ptile, b), c->sw, 1);
      launch_lu_pack(LAB(c, c->ptile, b), items_of(c->ptile, b), c->sw, 2);
      for (int t = 0; t < 4; ++t) {
        const int bwd = (t == 1 || t == 3) ? 1 : 0;
        launch_lusgs_tiles(c->sw, c->db, c->d_tiles, c->d_border[bwd] + c->tile_first[b], c->tile_count[b],
                           c->d_flags, c->d_epoch, c->d_err, t, c->p, c->stream);
      }
      const long long w0 = c->ptile.hwg0[b];
      launch_update(LAB(c, c->ptile, b), items_of(c->ptile, b), c->sw, c->p.gama, c->intr.d + b, c->d_upart + w0,
                    c->d_upart + c->ptile.nwg + w0, c->d_ukey + w0, c->d_epoch);
    }
    return;
  }
  if (!state_packed) launch_lu_pack(LA(c, c->ptile), c->ptile.total(), c->sw, 1);
  launch_lu_pack(LA(c, c->ptile), c->ptile.total(), c->sw, 2);
  op_sweeps(c);
  launch_update(LA(c, c->ptile), c->ptile.total(), c->sw, c->p.gama, c->intr.d, c->d_upart,
                c->d_upart + c->ptile.nwg, c->d_ukey, c->d_epoch);
}

// the residual of the update op_update just made, from its partials
// (one row per request: out + 4 (step nreq + r))
static void op_update_residual(orion_hip_ctx* c, double* out, int* step) {
  launch_residual_final(c->d_upart, c->d_upart + c->ptile.nwg, c->d_ukey, c->d_rseg, c->d_rtot, c->nreq, out, step,
                        c->d_rdone, c->stream);
}

// a sweep whose wait for an upstream tile ran out has left wrong values
static int check_sweeps(orion_hip_ctx* c) {
  int e = 0;
  HIPCHK(hipMemcpyAsync(&e, c->d_err, sizeof e, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (e) return fail("LU-SGS sweep: a tile waited past its bound for an upstream tile; results are invalid");
  return 0;
}

// per-step outputs of n steps: one row per (step, request)
static int ensure_res(orion_hip_ctx* c, int n) {
  if (n <= c->res_cap) return 0;
  // grow: old buffers stay owned by the context until destroy
  const size_t rows = (size_t)n * c->nreq;
  if (dalloc(c, (void**)&c->d_res, sizeof(double) * 4 * rows)) return -1;
  if (dalloc(c, (void**)&c->d_dtmin, sizeof(unsigned long long) * rows)) return -1;
  c->res_cap = n;
  if (c->gexec) {  // the captured step points at the old buffers
    hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  return 0;
}

// row of request req: the location's block is numbered within the request
static void decode_stats(const orion_hip_ctx* c, const double* r, unsigned long long dtbits, orion_residual_stats* o,
                         int req = 0) {
  o->rms_residual = r[0];
  o->max_residual = r[1];
  const long long key = (long long)r[2];
  const long long g = key / 5;
  const int m = (int)(key % 5);
  int b = c->req_first[req];
  while (b + 1 < c->req_first[req + 1] && c->intr.h[b + 1] <= g) ++b;
  const BlockDev& B = c->hb[b];
  const long long n = g - c->intr.h[b];
  o->max_loc[0] = b - c->req_first[req];
  o->max_loc[1] = (int)(n % B.idim) + 1;
  o->max_loc[2] = (int)((n / B.idim) % B.jdim) + 1;
  o->max_loc[3] = (int)(n / ((long long)B.idim * B.jdim)) + 1;
  o->max_loc[4] = m + 1;
  double d;
  std::memcpy(&d, &dtbits, sizeof d);
  o->phydtime = d;
}

int orion_hip_exchange_bc(orion_hip_ctx* c) {
  op_halo(c);
  return check_launch("exchange_bc");
}
int orion_hip_apply_physical_bc(orion_hip_ctx* c) {
  op_bc(c);
  return check_launch("apply_physical_bc");
}
int orion_hip_update_flow_states(orion_hip_ctx* c) {
  op_state(c, 1, nullptr, nullptr);
  return check_launch("update_flow_states");
}
int orion_hip_calculate_time_step(orion_hip_ctx* c, double* phydtime) {
  if (c->nreq != 1) return fail("calculate_time_step: a batch reports each request's time step through orion_hip_run");
  if (ensure_res(c, 1)) return -1;
  const double init = 1000.0;
  unsigned long long bits;
  std::memcpy(&bits, &init, sizeof bits);
  HIPCHK(hipMemcpyAsync(c->d_dtmin, &bits, sizeof bits, hipMemcpyHostToDevice, c->stream));
  op_state(c, 2, c->d_dtmin, nullptr);
  if (check_launch("calculate_time_step")) return -1;
  if (phydtime) {
    HIPCHK(hipMemcpyAsync(&bits, c->d_dtmin, sizeof bits, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::memcpy(phydtime, &bits, sizeof bits);
  }
  return 0;
}
int orion_hip_compute_viscous_rhs(orion_hip_ctx* c) {
  op_visc(c);
  return check_launch("compute_viscous_rhs");
}
int orion_hip_average_interface_residuals(orion_hip_ctx* c) {
  op_avg(c);
  return check_launch("average_interface_residuals");
}
int orion_hip_compute_inviscid_rhs(orion_hip_ctx* c) {
  op_inv(c);
  return check_launch("compute_inviscid_rhs");
}
int orion_hip_update_solution(orion_hip_ctx* c) {
  op_update(c);
  return check_launch("update_solution");
}
int orion_hip_residual(orion_hip_ctx* c, orion_residual_stats* out) {
  if (c->nreq != 1) return fail("residual: a batch reports each request's residual through orion_hip_run");
  if (ensure_res(c, 1)) return -1;
  launch_residual(LA(c, c->intr), c->intr.total(), 5 * c->intr.total(), c->d_part, c->d_pidx, c->nparts, c->d_res);
  if (check_launch("residual")) return -1;
  double r[4];
  unsigned long long bits = 0;
  HIPCHK(hipMemcpyAsync(r, c->d_res, sizeof r, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&bits, c->d_dtmin, sizeof bits, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (check_sweeps(c)) return -1;
  if (out) decode_stats(c, r, bits, out);
  return 0;
}

// One pseudo-time step; per-step outputs go to slot *d_step, which the last
// kernel advances, so the same sequence can be captured once and replayed.
static void step_body(orion_hip_ctx* c) {
  op_halo(c);
  op_bc(c);
  if (c->per_block) {
    // one stream, the operators in the reference's order, each per block
    op_state(c, 3 | 4, c->d_dtmin, c->d_step);
    if (c->p.nvis == 1) {
      op_visc(c, true);
      op_avg(c);
      op_inv(c, false);
    } else {
      op_inv(c, true);
    }
    op_avg(c);
    op_update(c);
    op_update_residual(c, c->d_res, c->d_step);
    return;
  }
  // The step's flow states: the ghost points (state kernel, mode 8) and
  // the interior with the LU-SGS pack's state part (state_pack_kernel);
  // mode 4: the first right-hand-side kernel after them writes dq whole.
  auto states = [&]() {
    launch_state(LA(c, c->pad), c->pad.total(), 3 | 4 | 8, c->d_dtmin, c->d_step, c->nreq);
    launch_state_pack(LA(c, c->ptile), c->ptile.total(), c->sw, c->d_dtmin, c->d_step, c->nreq);
  };
  if (c->p.nvis == 1 && !c->overlap) {
    states();
    op_visc(c, true);
    op_avg(c);
    op_inv(c, false);
  } else if (c->p.nvis == 1) {
    // The inviscid face fluxes (FP64-bound) run on the side stream beside
    // the flow states and the viscous chain (load-bound).  They read prim's
    // first five fields and the static face geometry, which neither writes
    // (the state kernels write T, c, mu, q, dt and the sweep arrays' state
    // part), write their own flux arrays, and join before the inviscid
    // difference, the first kernel after them that touches dq.  dq is thus
    // updated in the operators' order.
    hipEventRecord(c->ev_fork, c->stream);
    hipStreamWaitEvent(c->side_stream, c->ev_fork, 0);
    op_inv_faces(c, c->side_stream);
    hipEventRecord(c->ev_join, c->side_stream);
    states();
    op_visc(c, true);
    op_avg(c);
    hipStreamWaitEvent(c->stream, c->ev_join, 0);
    launch_inv_diff(LA(c, c->dtile), c->dtile.total(), false);
  } else {
    states();
    op_inv(c, true);
  }
  op_avg(c);
  op_update(c, true);
  op_update_residual(c, c->d_res, c->d_step);  // and advances *d_step
}

static int ensure_graph(orion_hip_ctx* c) {
  if (c->gexec && c->graph_stream == c->stream) return 0;
  if (c->gexec) {
    hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  if (c->graph) {
    hipGraphDestroy(c->graph);
    c->graph = nullptr;
  }
  std::lock_guard<std::recursive_mutex> guard(g_capture_mutex);
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  step_body(c);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (e != hipSuccess) return fail("step capture failed: %s", hipGetErrorString(e));
  c->graph = g;
  HIPCHK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
  c->graph_stream = c->stream;
  return 0;
}

int orion_hip_run(orion_hip_ctx* c, int nsteps, orion_residual_stats* hist) {
  if (nsteps <= 0) return 0;
  if (ensure_res(c, nsteps)) return -1;
  const size_t rows = (size_t)nsteps * c->nreq;
  {
    const double init = 1000.0;
    unsigned long long bits;
    std::memcpy(&bits, &init, sizeof bits);
    std::vector<unsigned long long> v(rows, bits);
    const int zero = 0;
    HIPCHK(hipMemcpyAsync(c->d_dtmin, v.data(), sizeof(unsigned long long) * rows, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_step, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (c->use_graph && !c->per_block) {
    if (ensure_graph(c)) return -1;
    for (int s = 0; s < nsteps; ++s) HIPCHK(hipGraphLaunch(c->gexec, c->stream));
  } else {
    for (int s = 0; s < nsteps; ++s) step_body(c);
  }
  if (check_launch("run") || check_sweeps(c)) return -1;
  if (hist) {
    std::vector<double> r(4 * rows);
    std::vector<unsigned long long> bits(rows);
    HIPCHK(hipMemcpyAsync(r.data(), c->d_res, sizeof(double) * 4 * rows, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(bits.data(), c->d_dtmin, sizeof(unsigned long long) * rows, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t w = 0; w < rows; ++w) decode_stats(c, &r[4 * w], bits[w], &hist[w], (int)(w % c->nreq));
  }
  return 0;
}

int orion_hip_profile_run(orion_hip_ctx* c, int nsteps, double* phase_ms, long long* phase_launches) {
  if (nsteps <= 0) return 0;
  if (c->per_block) return fail("profile_run times the all-blocks launches; switch per-block launches off first");
  if (ensure_res(c, nsteps)) return -1;
  const int NP = ORION_PHASES;
  hipEvent_t ev[ORION_PHASES + 1];
  for (int i = 0; i <= NP; ++i) HIPCHK(hipEventCreate(&ev[i]));
  std::vector<double> acc(NP, 0.0);
  std::vector<long long> nl(NP, 0);
  for (int s = 0; s < nsteps; ++s) {
    int k = 0;
    auto mark = [&](long long launches) {
      hipEventRecord(ev[k + 1], c->stream);
      nl[k] += launches;
      ++k;
    };
    hipEventRecord(ev[0], c->stream);
    op_halo(c); mark(1);
    op_bc(c); mark((long long)c->bc_first.size());
    launch_state(LA(c, c->pad), c->pad.total(), 3 | 4 | 8, c->d_dtmin + (long long)s * c->nreq, nullptr, c->nreq);
    launch_state_pack(LA(c, c->ptile), c->ptile.total(), c->sw, c->d_dtmin + (long long)s * c->nreq, nullptr,
                      c->nreq);
    mark(2);
    if (c->p.nvis == 1) {
      launch_visc_deriv(LA(c, c->intr), c->intr.total(), c->d_vhead); mark(1);
      if (c->visc_fused) {  // faces and difference in one launch, timed as the face phase
        launch_visc_fused(c->db, c->d_vtiles, c->n_vtiles, c->d_vhead, c->visc_grid, true, c->stream);
        mark(1);
        mark(0);
      } else {
        launch_visc_face_queue(c->db, c->d_vitems, c->n_vitems, c->d_vhead, c->visc_grid, c->stream);
        mark(1);
        launch_visc_diff(LA(c, c->dtile), c->dtile.total(), true); mark(1);
      }
      op_avg(c); mark(2);
    } else {
      mark(0); mark(0); mark(0); mark(0);
    }
    launch_inv_face_pairs(LA3(c, c->fpair3), c->fpair3.total(), c->p.gama);
    mark(1);
    launch_inv_closure(LA3(c, c->ltask3), -1, c->ltask3.total(), c->p.gama);
    mark(1);
    mark(0);  // the pack's state part is in the state phase
    launch_inv_diff(LA(c, c->dtile), c->dtile.total(), c->p.nvis != 1); mark(1);
    op_avg(c); mark(2);
    launch_lu_pack(LA(c, c->ptile), c->ptile.total(), c->sw, 2); mark(1);
    if (c->schedule == ORION_SCHEDULE_CHAINED) {
      mark(0); mark(0); mark(0); mark(0);
      op_sweeps(c); mark(1);
    } else {
      for (int t = 0; t < 4; ++t) { op_sweep(c, t); mark(1); }
      mark(0);
    }
    launch_update(LA(c, c->ptile), c->ptile.total(), c->sw, c->p.gama, c->intr.d, c->d_upart,
                  c->d_upart + c->ptile.nwg, c->d_ukey, c->d_epoch);
    mark(1);
    op_update_residual(c, c->d_res + 4 * (long long)s * c->nreq, nullptr);
    mark(1);
    if (check_launch("profile_run")) return -1;
    HIPCHK(hipEventSynchronize(ev[k]));
    for (int i = 0; i < k; ++i) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
      acc[i] += ms;
    }
  }
  for (int i = 0; i <= NP; ++i) hipEventDestroy(ev[i]);
  if (check_sweeps(c)) return -1;
  for (int i = 0; i < NP; ++i) {
    if (phase_ms) phase_ms[i] = acc[i];
    if (phase_launches) phase_launches[i] = nl[i];
  }
  return 0;
}

int orion_hip_output_fields(orion_hip_ctx* c, int b, float* host, long long count) {
  if (!c || b < 0 || b >= c->nb) return fail("output_fields: bad block %d", b);
  const BlockDev& B = c->hb[b];
  const long long n = (long long)B.idim * B.jdim * B.kdim;
  if (count != 7 * n) return fail("output_fields: block %d has %lld values, got %lld", b, 7 * n, count);
  // staged in the block's face-flux scratch (15 P doubles >= 7 n floats),
  // which every step rewrites before reading
  float* dev = reinterpret_cast<float*>((double*)B.fx);
  launch_output(c->db, b, n, dev, c->p.gama, c->stream);
  if (check_launch("output_fields")) return -1;
  HIPCHK(hipMemcpyAsync(host, dev, sizeof(float) * 7 * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int orion_hip_write_solution(orion_hip_ctx* c, const char* flowname, int step, const double* xyz) {
  if (!c || !xyz) return fail("write_solution: null context or grid");
  // output directory and file stem (PostProcess.cpp:53-71)
  std::string prefix = flowname ? flowname : "";
  if (prefix.empty()) prefix = "output/flow";
  const std::filesystem::path path(prefix);
  std::string dir = path.parent_path().string();
  const std::string stem = path.filename().string();
  if (dir.empty()) dir = ".";
  {
    std::error_code ec;
    if (!std::filesystem::exists(dir, ec)) std::filesystem::create_directories(dir, ec);
  }
  const double* g = xyz;
  std::vector<float> f;
  for (int b = 0; b < c->nb; ++b) {
    const BlockDev& B = c->hb[b];
    const long long n = (long long)B.idim * B.jdim * B.kdim;
    f.resize(7 * n);
    if (orion_hip_output_fields(c, b, f.data(), 7 * n)) return -1;
    std::ostringstream name;
    name << dir << "/" << stem << "_" << std::setfill('0') << std::setw(6) << step << "_block_" << b << ".vts";
    if (orion_hip_write_vts(name.str().c_str(), B.idim, B.jdim, B.kdim, g, g + n, g + 2 * n, f.data())) return -1;
    g += 3 * n;
  }
  return orion_hip_write_vtm(dir.c_str(), stem.c_str(), c->nb, step);
}

// state file: magic, block count, the interior sizes of every block, the
// step, then every block's padded prim
static const char kStateMagic[8] = {'O', 'R', 'I', 'O', 'N', 'S', 'T', '1'};

int orion_hip_save_state(orion_hip_ctx* c, const char* path, long long step) {
  if (!c || !path) return fail("save_state: null context or path");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail("save_state: cannot open %s", path);
  bool ok = std::fwrite(kStateMagic, 1, 8, f) == 8;
  const long long nb = c->nb;
  ok = ok && std::fwrite(&nb, sizeof nb, 1, f) == 1;
  for (int b = 0; ok && b < c->nb; ++b) {
    const int d[3] = {c->hb[b].idim, c->hb[b].jdim, c->hb[b].kdim};
    ok = std::fwrite(d, sizeof(int), 3, f) == 3;
  }
  ok = ok && std::fwrite(&step, sizeof step, 1, f) == 1;
  std::vector<double> h;
  for (int b = 0; ok && b < c->nb; ++b) {
    h.resize(6 * c->hb[b].P);
    if (orion_hip_download(c, b, ORION_FIELD_PRIM, h.data(), (long long)h.size())) {
      std::fclose(f);
      return -1;
    }
    ok = std::fwrite(h.data(), sizeof(double), h.size(), f) == h.size();
  }
  if (std::fclose(f) != 0) ok = false;
  return ok ? 0 : fail("save_state: write to %s failed", path);
}

int orion_hip_load_state(orion_hip_ctx* c, const char* path, long long* step) {
  if (!c || !path) return fail("load_state: null context or path");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail("load_state: cannot open %s", path);
  auto bad = [&](const char* why) {
    std::fclose(f);
    return fail("load_state: %s: %s", path, why);
  };
  char magic[8];
  long long nb = 0, st = 0;
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kStateMagic, 8) != 0) return bad("not a state file");
  if (std::fread(&nb, sizeof nb, 1, f) != 1 || nb != c->nb) return bad("block count differs from the context's");
  for (int b = 0; b < c->nb; ++b) {
    int d[3];
    if (std::fread(d, sizeof(int), 3, f) != 3) return bad("truncated header");
    if (d[0] != c->hb[b].idim || d[1] != c->hb[b].jdim || d[2] != c->hb[b].kdim)
      return bad("block sizes differ from the context's");
  }
  if (std::fread(&st, sizeof st, 1, f) != 1) return bad("truncated header");
  std::vector<double> h;
  for (int b = 0; b < c->nb; ++b) {
    h.resize(6 * c->hb[b].P);
    if (std::fread(h.data(), sizeof(double), h.size(), f) != h.size()) return bad("truncated fields");
    if (orion_hip_upload(c, b, ORION_FIELD_PRIM, h.data(), (long long)h.size())) {
      std::fclose(f);
      return -1;
    }
  }
  std::fclose(f);
  if (step) *step = st;
  return 0;
}

int orion_hip_synchronize(orion_hip_ctx* c) {
  HIPCHK(hipStreamSynchronize(c->stream));
  return check_sweeps(c);
}

int orion_hip_set_stream(orion_hip_ctx* c, void* s) {
  c->stream = s ? (hipStream_t)s : c->own_stream;
  return 0;
}

int orion_hip_set_schedule(orion_hip_ctx* c, int which) {
  if (!c || which < 0 || which >= ORION_SCHEDULES) return fail("set_schedule: no schedule %d", which);
  if (which == c->schedule) return 0;
  if (which == ORION_SCHEDULE_CHAINED) {
    if (c->n_chain == 0)
      return fail("set_schedule: the chained schedule needs every block's LU-SGS arrays below 4 GiB");
  } else {
    // the captured step reads the order arrays in place: only their contents change
    for (int k = 0; k < 2; ++k) {
      const std::vector<int>& v = c->sched_order[which][k];
      if (!v.empty())
        HIPCHK(hipMemcpyAsync(c->d_order[k], v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  // the chained schedule is another launch: capture the step again
  if (which == ORION_SCHEDULE_CHAINED || c->schedule == ORION_SCHEDULE_CHAINED) c->graph_stream = nullptr;
  c->schedule = which;
  return 0;
}

int orion_hip_get_schedule(const orion_hip_ctx* c) { return c ? c->schedule : -1; }

// Trials of the autotuners: each runs one untimed step (capture and warm-up)
// and then, twice, `steps` timed steps from the state saved at the start,
// keeping the faster, and restores that state (prim alone fixes the next
// step).
struct TrialState {
  orion_hip_ctx* c;
  std::vector<long long> at;
  double* saved = nullptr;
  explicit TrialState(orion_hip_ctx* ctx) : c(ctx), at(ctx->nb + 1, 0) {
    for (int b = 0; b < c->nb; ++b) at[b + 1] = at[b] + 6 * c->hb[b].P;
  }
  ~TrialState() {
    if (saved) {
      hipStreamSynchronize(c->stream);
      hipFree(saved);
    }
  }
  int copy(bool save) {
    for (int b = 0; b < c->nb; ++b) {
      double* prim = c->hb[b].prim;
      double* dst = save ? saved + at[b] : prim;
      const double* src = save ? prim : saved + at[b];
      if (hipMemcpyAsync(dst, src, sizeof(double) * 6 * c->hb[b].P, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
        return fail("tuning: state copy failed");
    }
    return hipStreamSynchronize(c->stream) == hipSuccess ? 0 : fail("tuning: state copy failed");
  }
  int begin() {
    if (hipMalloc((void**)&saved, sizeof(double) * at[c->nb]) != hipSuccess)
      return fail("tuning: no memory for the state copy");
    return copy(true);
  }
  // the faster of two timed runs from the saved state (shapes often differ
  // by less than one run's noise)
  int trial(int steps, double* ms) {
    int rc = orion_hip_run(c, 1, nullptr);
    for (int rep = 0; rc == 0 && rep < 2; ++rep) {
      rc = copy(false);
      if (rc != 0) break;
      const auto t0 = std::chrono::steady_clock::now();
      rc = orion_hip_run(c, steps, nullptr);  // synchronises at the end
      const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps;
      if (rep == 0 || t < *ms) *ms = t;
    }
    return rc == 0 ? copy(false) : rc;
  }
};

int orion_hip_tune_schedule(orion_hip_ctx* c, int steps, int* chosen, double* ms_per_step) {
  if (!c || steps <= 0) return fail("tune_schedule: bad context or step count");
  TrialState T(c);
  int rc = T.begin();
  int best = c->schedule;
  double best_ms = 0.0;
  for (int k = 0; rc == 0 && k < ORION_SCHEDULES; ++k) {
    double ms = 0.0;
    if (k == ORION_SCHEDULE_CHAINED && c->n_chain == 0) {
      if (ms_per_step) ms_per_step[k] = 0.0;
      continue;
    }
    rc = orion_hip_set_schedule(c, k);
    if (rc == 0) rc = T.trial(steps, &ms);
    if (rc == 0) {
      if (ms_per_step) ms_per_step[k] = ms;
      if (k == 0 || ms < best_ms) {
        best = k;
        best_ms = ms;
      }
    }
  }
  if (rc == 0) rc = orion_hip_set_schedule(c, best);
  if (rc == 0 && chosen) *chosen = best;
  return rc;
}

int orion_hip_set_visc_tiles(orion_hip_ctx* c, int segment, int grid) {
  if (!c) return fail("set_visc_tiles: null context");
  if (c->p.nvis != 1) return 0;  // no viscous flux, nothing to tile
  if (segment == c->visc_seg && grid == c->visc_grid) return 0;
  return set_visc_queue(c, segment, grid);
}

int orion_hip_set_visc_form(orion_hip_ctx* c, int fused) {
  if (!c) return fail("set_visc_form: null context");
  if (c->p.nvis != 1) return 0;  // no viscous flux
  if (fused && !c->visc_fused_ok) return fail("set_visc_form: the fused form needs at least 4 nodes on every block line");
  if ((fused != 0) == c->visc_fused) return 0;
  c->visc_fused = fused != 0;
  return set_visc_queue(c, c->visc_seg, c->visc_grid);
}

int orion_hip_get_visc_form(const orion_hip_ctx* c) { return c && c->visc_fused ? 1 : 0; }

int orion_hip_get_visc_tiles(const orion_hip_ctx* c, int* segment, int* grid) {
  if (!c) return fail("get_visc_tiles: null context");
  if (segment) *segment = c->visc_seg;
  if (grid) *grid = c->visc_grid;
  return 0;
}

int orion_hip_tune(orion_hip_ctx* c, int steps, int* schedule, int* segment, int* grid, double* best_ms,
                   int* trials) {
  if (!c || steps <= 0) return fail("tune: bad context or step count");
  TrialState T(c);
  int rc = T.begin();
  double best = -1.0;
  int n = 0;
  auto measure = [&](double* ms) {
    const int r = T.trial(steps, ms);
    ++n;
    return r;
  };
  // the sweep schedule
  int bs = c->schedule;
  for (int k = 0; rc == 0 && k < ORION_SCHEDULES; ++k) {
    double ms = 0.0;
    if (k == ORION_SCHEDULE_CHAINED && c->n_chain == 0) continue;
    rc = orion_hip_set_schedule(c, k);
    if (rc == 0) rc = measure(&ms);
    if (rc == 0 && (best < 0.0 || ms < best)) {
      best = ms;
      bs = k;
    }
  }
  if (rc == 0) rc = orion_hip_set_schedule(c, bs);
  // the viscous form, then its planes or faces per item, then the
  // persistent grid
  if (c->p.nvis == 1) {
    if (c->visc_fused_ok && rc == 0) {
      const bool was = c->visc_fused;
      double ms = 0.0;
      rc = orion_hip_set_visc_form(c, was ? 0 : 1);
      if (rc == 0) rc = measure(&ms);
      if (rc == 0 && !(ms < best)) rc = orion_hip_set_visc_form(c, was ? 1 : 0);
      else if (rc == 0) best = ms;
    }
    int bseg = c->visc_seg, bgrid = c->visc_grid;
    static const int SEGS[] = {8, 16, 32, 64, 128};
    static const int GRIDS[] = {256, 512, 1024, 2048, 4096};
    for (int seg : SEGS) {
      if (rc != 0 || seg == bseg) continue;
      double ms = 0.0;
      rc = orion_hip_set_visc_tiles(c, seg, bgrid);
      if (rc == 0) rc = measure(&ms);
      if (rc == 0 && ms < best) {
        best = ms;
        bseg = seg;
      }
    }
    for (int g : GRIDS) {
      if (rc != 0 || g == bgrid) continue;
      double ms = 0.0;
      rc = orion_hip_set_visc_tiles(c, bseg, g);
      if (rc == 0) rc = measure(&ms);
      if (rc == 0 && ms < best) {
        best = ms;
        bgrid = g;
      }
    }
    if (rc == 0) rc = orion_hip_set_visc_tiles(c, bseg, bgrid);
  }
  if (rc == 0) {
    if (schedule) *schedule = c->schedule;
    orion_hip_get_visc_tiles(c, segment, grid);
    if (best_ms) *best_ms = best;
    if (trials) *trials = n;
  }
  return rc;
}

int orion_hip_set_overlap(orion_hip_ctx* c, int on) {
  if (!c) return fail("set_overlap: null context");
  if ((on != 0) == c->overlap) return 0;
  c->overlap = on != 0;
  c->graph_stream = nullptr;  // the captured step has the other form: capture again
  return 0;
}

int orion_hip_set_per_block(orion_hip_ctx* c, int on) {
  if (!c) return fail("set_per_block: null context");
  c->per_block = on != 0;
  return 0;
}

size_t orion_hip_device_bytes(const orion_hip_ctx* c) { return c->arena_bytes + c->map_bytes; }
long long orion_hip_interior_nodes(const orion_hip_ctx* c) { return c->intr.total(); }

}  // extern "C"
