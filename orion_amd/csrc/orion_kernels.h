// orion_kernels.h -- host-side launch wrappers for the gfx950 kernels.
#pragma once
#include "orion_dev.h"

namespace orion_hip {

// One multi-block launch: every kernel maps its flat thread index onto
// blocks through `off` (device prefix table of the iteration space, nb+1).
struct Launch {
  hipStream_t stream;
  const BlockDev* blocks;  // device
  const long long* off;    // device prefix of item counts (nb+1)
  int nb;
  Span span;               // the same space padded per block to whole workgroups
  long long nwg;           // workgroups of the padded space
};

// physical-BC work item: one region of one block (memory coordinates)
struct BCTask {
  int block, bctype;
  int lo[3], hi[3];  // face window, padded 0-based inclusive
  int s_nd, s_lr;    // 1-based normal axis, -1/0/+1 face side
  int mid[3];        // window mid point (symmetry normal)
  long long start;   // first flat index of this task in the launch
};

// interface averaging maps
struct AvgSrc {
  int block;
  long long idx;
};
struct AvgDst {
  int block;
  int count;
  long long idx;
  long long start;  // first staging entry (entries are in region order)
};

// k_flux.hip
// face kernels: d = 0, 1, 2 for one direction's space; d = -1 for a space of
// 3 nb virtual blocks d * nb + b holding the three directions' spaces
void launch_inv_face(const Launch& L, int d, long long n, double gamma);
// the same fluxes, two neighbouring faces of a line per thread, over the
// three directions' face-pair spaces as 3 nb virtual blocks
void launch_inv_face_pairs(const Launch& L, long long n, double gamma);
void launch_inv_closure(const Launch& L, int d, long long n, double gamma);
// first: the step's first right-hand-side kernel, starting from dq = 0
// without reading it (the step's state kernel then skips the zeroing)
void launch_inv_diff(const Launch& L, long long n, bool first);
// queue_head: when not null, the viscous face work queue's ticket counter,
// zeroed by this launch
void launch_visc_deriv(const Launch& L, long long n, int* queue_head = nullptr);
// one item of the viscous face-flux work queue (visc_face_queue_kernel):
// d = 0: face pairs [a0, a1) of block's lines along i; d = 1, 2: lines
// [a0, a1) along j or k (i fastest), faces [f0, f1)
struct ViscItem {
  int block, d;
  long long a0, a1;
  int f0, f1;
};
void launch_visc_face_queue(const BlockDev* blocks, const ViscItem* items, int nitems, int* head, int grid,
                            hipStream_t s);
// one item of the fused viscous right-hand side (visc_fused_kernel): the
// tile of 16 x 16 (i, j) interior nodes from (i0, j0) (0-based) of block
// over the planes [k0, k1) along k
struct ViscTile {
  int block, i0, j0, k0, k1;
};
// the face fluxes and their difference in one persistent launch over the
// item queue (*head zeroed by the derivative kernel before it); first: the
// step's first right-hand-side kernel (dq from zero)
void launch_visc_fused(const BlockDev* blocks, const ViscTile* items, int nitems, int* head, int grid, bool first,
                       hipStream_t s);
void launch_face_geom(const Launch& L, int d, long long n);
void launch_visc_face(const Launch& L, int d, long long n);
void launch_visc_diff(const Launch& L, long long n, bool first);

// k_point.hip
// dtmin slot of (step, request): dtmin_base[*step * nreq + request]
void launch_state(const Launch& L, long long n, int mode, unsigned long long* dtmin_base, const int* step,
                  int nreq);
// into row 0 of out
void launch_residual(const Launch& L, long long n, long long total5, double* partial, long long* pidx,
                     int nparts, double* out);
int residual_parts(long long n);
// the VTS point fields of block bi (n interior nodes) as float, 7 n values
void launch_output(const BlockDev* blocks, int bi, long long n, float* out, double gamma, hipStream_t s);
// the final reduction alone, over partials left by another kernel (any
// count): one workgroup per request r over partials seg[r] .. seg[r + 1],
// row (*step) * nreq + r of out; then *step advances (done: a zeroed counter)
void launch_residual_final(const double* psum, const double* pmax, const long long* pkey, const int* seg,
                           const long long* tot5, int nreq, double* out, int* step, int* done, hipStream_t s);
void launch_fill(double* p, long long n, double v, hipStream_t s);
void launch_temperature(const Launch& L, long long n);

// k_bc.hip
void launch_halo(const long long* pairs, long long n, double* arena, hipStream_t s);
void launch_halo_points(const unsigned long long* pts, long long n, double* arena, hipStream_t s);
void launch_bc(const BlockDev* blocks, const BCTask* tasks, int ntasks, long long n, hipStream_t s);
void launch_avg_stage(const BlockDev* blocks, const AvgSrc* src, long long n, double* stage, hipStream_t s);
void launch_avg_apply(const BlockDev* blocks, const AvgDst* dst, long long n, const double* stage, hipStream_t s);

// k_lusgs.hip
void launch_lu_geom(const Launch& La, long long n, const SweepArrays& S);
// part 1: met8, coed and D (state only); part 2: the rhs copy
void launch_lu_pack(const Launch& La, long long n, const SweepArrays& S, int part);
// the step's interior flow states and time step (state modes 1 | 2 | 4)
// with the pack's state part, over the pack's tile space; the ghost points
// take launch_state with mode 8
void launch_state_pack(const Launch& La, long long n, const SweepArrays& S, unsigned long long* dtmin_base,
                       const int* step, int nreq);
void launch_lusgs_tiles(const SweepArrays& S, const BlockDev* blocks, const SweepTile* tiles, const int* order,
                        int ntiles, unsigned long long* flags, unsigned long long* epoch, int* err, int type,
                        const Params& p, hipStream_t s);
// the four sweeps as one launch of tasks (tile << 2 | sweep type) in a
// topological order
void launch_lusgs_chain(const SweepArrays& S, const BlockDev* blocks, const SweepTile* tiles, const int* tasks,
                        int ntasks, unsigned long long* flags, unsigned long long* epoch, int* err, const Params& p,
                        hipStream_t s);
// ioff: prefix of the interior node counts; psum/pmax/pkey: one residual
// partial per workgroup of the launch
void launch_update(const Launch& La, long long n, const SweepArrays& S, double gamma, const long long* ioff,
                   double* psum, double* pmax, long long* pkey, unsigned long long* epoch);

// k_metrics.hip
void launch_metrics(const Launch& L, long long n, int gcl, double* tmp, hipStream_t s);
void launch_count_small_vol(const Launch& L, long long n, unsigned long long* count);
void launch_grid_scatter(const Launch& L, long long n);

}  // namespace orion_hip
