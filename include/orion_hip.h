/*
 * orion_hip.h -- C ABI of the MI355X (gfx950) implementation of ORION's
 * per-pseudo-time-step hot path.
 *
 * The entry points replace the reference's operator surface one for one;
 * each comment names the reference interface it stands in for.  All state
 * lives on the device in one HBM arena per context, laid out exactly like
 * the reference's OrionArray fields (ghost-padded box (idim+6, jdim+6,
 * kdim+6), i fastest, variable slowest; src/core/OrionArray.hpp:241-261), so
 * host buffers in the reference's layout copy in and out unchanged.
 *
 * Conventions: every function returns 0 on success and a negative value on
 * failure, with the message available from orion_hip_last_error() (the
 * reference throws std::runtime_error, src/core/OrionArray.hpp:25-27).
 * Operators are asynchronous on the context's stream and run in the order
 * they are called; only functions that return host values synchronise.
 * One context must not be used from two threads at once.
 */
#ifndef ORION_HIP_H
#define ORION_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orion_hip_ctx orion_hip_ctx;

/* One BC region in file order (bc::BCRegion, src/bc/BCData.hpp:8-50):
 * 1-based windows as read by BCReader (src/bc/BCReader.cpp:107-175); block
 * is the 0-based owner; nbt/ibcwin are the 1-based target block and window
 * of an interface (bctype < 0). */
typedef struct {
  int block;
  int bctype;
  int s_st[3], s_ed[3];
  int t_st[3], t_ed[3];
  int nbt;
  int ibcwin;
} orion_region_desc;

/* The namelist inputs the hot path depends on (core::Params,
 * src/core/Params.hpp:11-106).  The freestream constants are derived as in
 * preprocess::init_inflow (src/preprocess/Inflow.cpp:300-459) for the
 * custom-atmosphere path with tref given. */
typedef struct {
  double moo, reynolds, tref, twall, cfl;
  double attack, sideslip; /* degrees */
  int csrv;                /* integer, as the reference parses it */
  int nvis, method, ntmst, gcl;
} orion_inputs;

/* solver::ResidualStats (src/solver/ResidualMonitor.hpp:8-12) plus the
 * step's params.step.phydtime. */
typedef struct {
  double rms_residual;
  double max_residual;
  int max_loc[5]; /* 0-based global block, 1-based i, j, k, var */
  double phydtime;
} orion_residual_stats;

/* Field ids (preprocess::BlockField, src/preprocess/FlowField.hpp:19-71). */
enum {
  ORION_FIELD_PRIM = 0,    /* 6 vars: rho,u,v,w,p,T */
  ORION_FIELD_METRICS = 1, /* 9 vars */
  ORION_FIELD_VOL = 2,
  ORION_FIELD_Q = 3,       /* 5 vars; only when created with keep_q */
  ORION_FIELD_DQ = 4,      /* 5 vars */
  ORION_FIELD_DT = 5,      /* dt/vol */
  ORION_FIELD_C = 6,
  ORION_FIELD_MU = 7
};

/* Backend init hook; replaces backend_cpu_init (src/backend/cpu/BackendCPU.cpp:3).
 * Selects the device for contexts created afterwards on this thread. */
int orion_hip_backend_init(int device);
const char* orion_hip_last_error(void);

/* Setup: replaces the per-case part of OrionSolver::preprocess
 * (src/core/OrionSolver.cpp:61-93): bc::set_bc_index / prepare_bc_topology
 * (src/bc/BCPreprocess.cpp:194-270), allocate_other_variable
 * (src/preprocess/FlowField.cpp:25-107), mesh::compute_grid_metrics and
 * check_grid_metrics (src/mesh/MetricComputer.cpp:103-500, on the device),
 * init_inflow, initialize_flow and update_temperature.  dims holds
 * 3*nblocks interior sizes (each > 6); xyz holds per block x, y then z
 * (i fastest).  keep_q: also maintain the write-only conservative field. */
int orion_hip_ctx_create(int nblocks, const int* dims, const double* xyz, int nregions,
                         const orion_region_desc* regions, const orion_inputs* in, int keep_q,
                         orion_hip_ctx** out);
/* No counterpart in the reference (one solver per process there): a context
 * that batches nreq independent solves, for serving many small requests.
 * Request r owns req_nblocks[r] consecutive blocks (their sum is nblocks)
 * and has its own inputs, inputs[r]; dims, xyz and regions are laid out as
 * for orion_hip_ctx_create over all blocks (region block numbers over all
 * blocks), and an interface may not join blocks of two requests.  The
 * requests must share nvis, method, csrv, ntmst and gcl.  Every kernel of a
 * step covers all requests; each request's results are those it has in a
 * context of its own, bit for bit.  orion_hip_run then reports nsteps x
 * nreq residuals; the per-operator residual and time step calls refuse a
 * batch.  orion_hip_ctx_create is the batch of one request. */
int orion_hip_batch_create(int nreq, const int* req_nblocks, int nblocks, const int* dims, const double* xyz,
                           int nregions, const orion_region_desc* regions, const orion_inputs* inputs, int keep_q,
                           orion_hip_ctx** out);
/* No counterpart in the reference: orion_hip_batch_create with each
 * request's grid where the caller holds it (req_xyz[r]: request r's blocks'
 * coordinates one after another), so that a serving layer rebuilding a
 * batch of resident requests does not gather their grids into one array. */
int orion_hip_batch_create_grids(int nreq, const int* req_nblocks, int nblocks, const int* dims,
                                 const double* const* req_xyz, int nregions, const orion_region_desc* regions,
                                 const orion_inputs* inputs, int keep_q, orion_hip_ctx** out);
/* Requests in the context (1 unless batched). */
int orion_hip_requests(const orion_hip_ctx* ctx);
void orion_hip_ctx_destroy(orion_hip_ctx* ctx);

/* Device pointer and element count of one block field (OrionArray::devicePtr,
 * src/core/OrionArray.hpp:139-141). */
int orion_hip_field(orion_hip_ctx* ctx, int block, int field, double** dev_ptr, long long* count);
/* Host <-> device copies of a whole block field (OrionArray::toDevice /
 * fromDevice, src/core/OrionArray.hpp:146-155); synchronous. */
int orion_hip_upload(orion_hip_ctx* ctx, int block, int field, const double* host, long long count);
int orion_hip_download(orion_hip_ctx* ctx, int block, int field, double* host, long long count);
/* No counterpart in the reference: one block's solution state (prim, which
 * alone fixes the next step, see orion_hip_save_state) copied from a block
 * of another context on the same device, without a host round trip; for
 * rebuilding a serving batch around requests that stay.  The blocks must
 * have the same sizes.  Synchronous. */
int orion_hip_copy_block_state(orion_hip_ctx* dst, int dst_block, orion_hip_ctx* src, int src_block);

/* preprocess::update_temperature (src/preprocess/InitialCondition.cpp:91-147). */
int orion_hip_update_temperature(orion_hip_ctx* ctx);

/* solver::HaloExchanger::exchange_bc (src/solver/HaloExchanger.hpp:20). */
int orion_hip_exchange_bc(orion_hip_ctx* ctx);
/* bc::apply_physical_bc (src/bc/PhysicalBC.hpp:17). */
int orion_hip_apply_physical_bc(orion_hip_ctx* ctx);
/* solver::StateUpdater::update_flow_states (src/solver/StateUpdater.hpp:14). */
int orion_hip_update_flow_states(orion_hip_ctx* ctx);
/* solver::TimeIntegrator::calculate_time_step (src/solver/TimeIntegrator.hpp:10);
 * phydtime may be NULL (then nothing synchronises). */
int orion_hip_calculate_time_step(orion_hip_ctx* ctx, double* phydtime);
/* solver::FluxComputer::compute_viscous_rhs (src/solver/FluxComputer.hpp:15). */
int orion_hip_compute_viscous_rhs(orion_hip_ctx* ctx);
/* solver::HaloExchanger::average_interface_residuals (src/solver/HaloExchanger.hpp:25). */
int orion_hip_average_interface_residuals(orion_hip_ctx* ctx);
/* solver::InviscidFluxComputer::compute_inviscid_rhs (src/solver/InviscidFluxComputer.hpp:17). */
int orion_hip_compute_inviscid_rhs(orion_hip_ctx* ctx);
/* solver::StateUpdater::update_solution (src/solver/StateUpdater.hpp:21). */
int orion_hip_update_solution(orion_hip_ctx* ctx);
/* solver::ResidualMonitor::compute (src/solver/ResidualMonitor.hpp:20); synchronises. */
int orion_hip_residual(orion_hip_ctx* ctx, orion_residual_stats* out);

/* nsteps pseudo-time steps in the order of OrionSolver::solve
 * (src/core/OrionSolver.cpp:139-209, without the debug block at :156-201).
 * hist (nsteps x requests entries, step-major, may be NULL) receives each
 * step's residual per request, max_loc's block numbered within the
 * request; it is read back once, after the last step. */
int orion_hip_run(orion_hip_ctx* ctx, int nsteps, orion_residual_stats* hist);

/* Measurement helper with no counterpart in the reference: runs nsteps
 * like orion_hip_run with HIP events between the kernels of each phase and
 * returns, per phase, the summed device time (ms) and the number of kernel
 * launches.  Phases, in order: halo, bc, state+dt, visc_deriv, visc_face,
 * visc_diff, avg1, inv_face, inv_closure, lu_pack_state, inv_diff, avg2,
 * lu_pack, lu_fwd, lu_bwd, gs_fwd, gs_bwd, lu_chain (the four sweeps as
 * one launch under the chained schedule; the four before it are then
 * empty), update, residual. */
#define ORION_PHASES 20
int orion_hip_profile_run(orion_hip_ctx* ctx, int nsteps, double* phase_ms, long long* phase_launches);

/* No counterpart in the reference: the launch order of the LU-SGS tiles.
 * 0 diagonal (by tile diagonal over all blocks), 1 critical path first
 * (by the planes left to the end of the tile's block sweep, longest first;
 * the default), both one launch per sweep; 2 chained: the four sweeps as
 * one launch of (tile, sweep) tasks, a tile's sweep waiting for its own
 * previous sweep, in the order a greedy list schedule on the device's CUs
 * starts them (unavailable, with an error, for blocks whose LU-SGS arrays
 * exceed 4 GiB).  Every order launches a task after the tasks it waits
 * for, and the results are the same under every order; they differ in how
 * soon the longest chains get CUs when the tiles outnumber them. */
#define ORION_SCHEDULES 3
#define ORION_SCHEDULE_CHAINED 2
int orion_hip_set_schedule(orion_hip_ctx* ctx, int schedule);
int orion_hip_get_schedule(const orion_hip_ctx* ctx);
/* The autotuner of the sweep schedule: from the current state, times
 * `steps` pseudo-steps (after one untimed step; the faster of two runs from
 * the same state) under each schedule, restoring the state after each
 * trial, and keeps the fastest.  chosen
 * (may be NULL) gets its index, ms_per_step (ORION_SCHEDULES entries, may
 * be NULL) each schedule's time.  The solution state afterwards is the one
 * before the call, bit for bit (prim alone fixes the next step).
 * Synchronous; holds a device copy of prim while it runs. */
int orion_hip_tune_schedule(orion_hip_ctx* ctx, int steps, int* chosen, double* ms_per_step);

/* No counterpart in the reference: the tiling of the viscous face-flux
 * stage, one persistent launch of `grid` workgroups that take items from a
 * block work queue.  In the fused form (orion_hip_set_visc_form) an item is
 * 16 x 16 columns by `segment` planes; in the face-flux queue form, along j
 * and k, 256 lines by `segment` faces, each lane marching along its line,
 * and along i, 256 face pairs.  Results do not depend on the tiling.
 * Defaults: segment 32, grid 2048. */
int orion_hip_set_visc_tiles(orion_hip_ctx* ctx, int segment, int grid);
int orion_hip_get_visc_tiles(const orion_hip_ctx* ctx, int* segment, int* grid);
/* No counterpart in the reference: the form of the viscous right-hand side
 * after the node derivatives.  1 (where every block has at least 4 nodes
 * along each axis): fused -- one persistent launch whose
 * items are tiles of 16 x 16 (i, j) columns by `segment` planes along k;
 * each forms its i, j and k face fluxes on the chip and differences them
 * into dq, so no face flux goes through device memory.  0 (the default):
 * the face-flux queue above, then a difference kernel.  Results are the same bit for
 * bit.  get returns the form in use; orion_hip_tune tries both. */
int orion_hip_set_visc_form(orion_hip_ctx* ctx, int fused);
int orion_hip_get_visc_form(const orion_hip_ctx* ctx);
/* The autotuner of the step's launch shapes: from the current state, times
 * `steps` pseudo-steps (after one untimed step; the faster of two runs from
 * the same state) per candidate, restoring the
 * state after each, and keeps the fastest by coordinate search -- the LU-SGS
 * tile schedule (ORION_SCHEDULES), then the viscous form (fused or face
 * queue, where both apply), then the viscous segment (8, 16, 32, 64, 128),
 * then its grid (256 to 4096 workgroups).  Outputs (each may be NULL):
 * the chosen schedule, segment and grid, the best ms per step, the trials
 * run.  The solution state afterwards is the one before the call, bit for
 * bit.  Synchronous; holds a device copy of prim while it runs. */
int orion_hip_tune(orion_hip_ctx* ctx, int steps, int* schedule, int* segment, int* grid, double* best_ms,
                   int* trials);

/* Output (postprocess::PostProcess, src/postprocess/PostProcess.hpp:10-33).
 *
 * The point fields that PostProcess::write_block_vts_binary writes
 * (src/postprocess/PostProcess.cpp:183-240) for one block, formed on the
 * device from the current state and copied back as float: Density,
 * Pressure, Temperature and Mach over the interior nodes (i fastest), then
 * Velocity as (u, v, w) per node.  count = 7 * idim * jdim * kdim.
 * Synchronous. */
int orion_hip_output_fields(orion_hip_ctx* ctx, int block, float* host, long long count);
/* PostProcess::write_solution (src/postprocess/PostProcess.cpp:44-123) for
 * every block of the context at step `step`: the directory and file stem
 * come from flowname (NULL or "" means "output/flow"), the directory is
 * created when missing, then <dir>/<stem>_<step:06>_block_<b>.vts per block
 * and <dir>/<stem>_<step:06>.vtm.  xyz: the grid as passed to
 * orion_hip_ctx_create (x, y, z of block 0, then block 1, ...). */
int orion_hip_write_solution(orion_hip_ctx* ctx, const char* flowname, int step, const double* xyz);
/* Host only (no device use): one block's VTS file as
 * write_block_vts_binary writes it (src/postprocess/PostProcess.cpp:126-252)
 * from its grid (x, y, z: idim * jdim * kdim each, i fastest) and the fields
 * laid out as orion_hip_output_fields returns them. */
int orion_hip_write_vts(const char* path, int idim, int jdim, int kdim, const double* x, const double* y,
                        const double* z, const float* fields);
/* Host only: the multiblock index of a step (write_global_vtm,
 * src/postprocess/PostProcess.cpp:254-283). */
int orion_hip_write_vtm(const char* output_dir, const char* stem, int nblocks, int step);

/* Restart, which the reference does not have (it always starts from the
 * freestream, src/preprocess/InitialCondition.cpp:10-18, and
 * OrionSolver::solve starts at step 0, src/core/OrionSolver.cpp:132).
 * The state file holds a header (magic "ORIONST1", block count, the
 * interior sizes of every block, the step) and then every block's padded
 * prim (6 variables, the reference's layout, doubles).  prim alone fixes
 * the next pseudo-time step, so a context of the same grid and inputs that
 * loads the file continues bit for bit as the one that saved it.  Loading
 * checks the block count and sizes.  Both synchronise. */
int orion_hip_save_state(orion_hip_ctx* ctx, const char* path, long long step);
int orion_hip_load_state(orion_hip_ctx* ctx, const char* path, long long* step);

/* No counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  solver::ResidualMonitor::compute
 * (src/solver/ResidualMonitor.hpp:20) over caller-held device fields, with
 * the kernels of orion_hip_residual, allocating nothing: its scratch (a
 * block table, partial sums and maxima, the result) is a fixed device array
 * of about 130 KB that each GPU holds until the process ends.  nblocks
 * blocks (1 to ORION_NOALLOC_MAX_BLOCKS, each under 2^31 interior nodes);
 * dims holds 3 * nblocks interior sizes; dq[b], dt[b] and vol[b] are device
 * pointers to block b's fields in the OrionArray layout of this header
 * (dq: 5 variables, dt holding dt/vol as calculate_time_step leaves it).
 * stream: a hipStream_t (NULL: the null stream).  out->phydtime is 0 (the
 * time step forms it, not the monitor).  Synchronises the stream.  Calls
 * on one GPU share its scratch: calls that overlap on one GPU race. */
#define ORION_NOALLOC_MAX_BLOCKS 256
int orion_hip_residual_noalloc(int nblocks, const int* dims, const double* const* dq, const double* const* dt,
                               const double* const* vol, void* stream, orion_residual_stats* out);

int orion_hip_synchronize(orion_hip_ctx* ctx);
/* Use a caller-owned hipStream_t (NULL restores the context's own stream). */
int orion_hip_set_stream(orion_hip_ctx* ctx, void* stream);
/* No counterpart in the reference.  With viscosity the step runs the
 * inviscid face fluxes on a second stream beside the viscous chain (on, the
 * default): that fills a GPU that one context has to itself.  Off runs the
 * step on one stream, for callers that keep many contexts busy at once
 * (their steps then overlap each other instead, without a second stream per
 * context sharing the hardware queues).  The kernels and the order of the
 * updates of dq are the same either way, so are the results. */
int orion_hip_set_overlap(orion_hip_ctx* ctx, int on);
/* No counterpart in the reference.  On: the plain per-block form -- every
 * kernel of a step launched once per block, in block order, as the
 * reference's operators loop over their blocks (e.g.
 * StateUpdater::update_solution, src/solver/StateUpdater.cpp:544-580), on
 * one stream and without the captured step; the boundary, halo and
 * averaging gathers stay one launch each.  Off (the default): the batched
 * form -- each kernel covers every block, the LU-SGS tiles of all blocks
 * share one launch per sweep, and orion_hip_run replays the step captured
 * as a graph.  The two forms give the same results bit for bit;
 * orion_hip_profile_run measures the batched form only. */
int orion_hip_set_per_block(orion_hip_ctx* ctx, int on);
/* Device bytes held by the context (arena plus maps). */
size_t orion_hip_device_bytes(const orion_hip_ctx* ctx);
/* Interior node count (the unit of cell-steps). */
long long orion_hip_interior_nodes(const orion_hip_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* ORION_HIP_H */
