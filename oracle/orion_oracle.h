/*
 * orion_oracle.h -- CPU restatement of ORION's per-pseudo-time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in orion_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path never links,
 * imports or falls back to it.
 *
 * It restates, in plain C with the reference's operation order, the
 * operators called by OrionSolver::solve (src/core/OrionSolver.cpp:139-209)
 * plus the setup they depend on:
 *   metrics + GCL           src/mesh/MetricComputer.cpp:21-369
 *   interface maps          src/bc/BCPreprocess.cpp:16-192
 *   freestream constants    src/preprocess/Inflow.cpp:300-459
 *   initial field           src/preprocess/InitialCondition.cpp:7-147
 *   halo exchange           src/solver/HaloExchanger.cpp:17-179
 *   physical BCs            src/bc/PhysicalBC.cpp:12-613
 *   flow states             src/solver/StateUpdater.cpp:15-168
 *   time step               src/solver/TimeIntegrator.cpp:20-178
 *   viscous RHS             src/solver/FluxComputer.cpp:13-732
 *   interface averaging     src/solver/HaloExchanger.cpp:186-398
 *   inviscid RHS            src/solver/InviscidFluxComputer.cpp:14-647
 *   LU-SGS update           src/solver/StateUpdater.cpp:175-580
 *   residual monitor        src/solver/ResidualMonitor.cpp:11-138
 *
 * Storage follows OrionArray (src/core/OrionArray.hpp:241-261): per block a
 * ghost-padded box (idim+6, jdim+6, kdim+6), i fastest, variable slowest.
 *
 * How each function is pinned (compared against an independent source):
 * see the header comment of orion_oracle.c.
 */
#ifndef ORION_ORACLE_H
#define ORION_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OB_NG 3

/* Field ids for ob_field(). */
enum {
  OB_PRIM = 0,     /* 6 vars: rho,u,v,w,p,T */
  OB_METRICS = 1,  /* 9 vars: kcx,kcy,kcz, etx,ety,etz, ctx,cty,ctz */
  OB_VOL = 2,
  OB_Q = 3,        /* 5 vars */
  OB_DQ = 4,       /* 5 vars */
  OB_DT = 5,
  OB_C = 6,
  OB_MU = 7,
  OB_SPEC = 8,     /* 3 vars */
  OB_SPEC_VISC = 9 /* 3 vars */
};

/* One BC region as read from the 1976-format file (src/bc/BCData.hpp:8-50),
 * windows 1-based and possibly signed.  block is 0-based. */
typedef struct {
  int block;
  int bctype;
  int s_st[3], s_ed[3];
  int t_st[3], t_ed[3];
  int nbt;    /* 1-based target block (bctype<0), else -1 */
  int ibcwin; /* 1-based target window (bctype<0), else 0 */
} ob_region_desc;

/* Raw namelist inputs that the hot path depends on (src/core/Params.hpp). */
typedef struct {
  double moo, reynolds, tref, twall, cfl;
  double attack, sideslip; /* degrees */
  int csrv;                /* parsed as int by the reference */
  int nvis, method, ntmst, gcl;
} ob_inputs;

/* Derived freestream constants (Inflow.cpp init_inflow). */
typedef struct {
  double gama, prl, moo, reynolds, visc, tref, twall, cfl;
  double roo, uoo, voo, woo, poo, too;
  int csrv, nvis, method, ntmst, gcl;
  double phydtime;
} ob_params;

typedef struct {
  double rms, max;
  int max_loc[5]; /* 0-based block id, 1-based i,j,k, 1-based var */
} ob_residual;

typedef struct ob_case ob_case;

/* Build a case: dims is 3*nblocks interior sizes; xyz holds, per block in
 * order, x then y then z (i fastest).  Runs metrics (+GCL), the interface
 * map, freestream constants and the uniform initial field.  Returns NULL on
 * error (message via ob_last_error()). */
ob_case* ob_create(int nblocks, const int* dims, const double* xyz,
                   int nregions, const ob_region_desc* regions,
                   const ob_inputs* in);
void ob_destroy(ob_case* c);
const char* ob_last_error(void);

int ob_nblocks(const ob_case* c);
void ob_block_dims(const ob_case* c, int b, int* padded3);
double* ob_field(ob_case* c, int b, int field);
const ob_params* ob_get_params(const ob_case* c);
void ob_set_threads(int n);

/* Recompute T from p/rho over all points (InitialCondition.cpp:91-147). */
void ob_update_temperature(ob_case* c);

/* Operators, in OrionSolver::solve order. */
int ob_exchange_bc(ob_case* c);
int ob_apply_physical_bc(ob_case* c);
void ob_update_flow_states(ob_case* c);
void ob_calculate_time_step(ob_case* c);
void ob_compute_viscous_rhs(ob_case* c);
int ob_average_interface_residuals(ob_case* c);
void ob_compute_inviscid_rhs(ob_case* c);
void ob_update_solution(ob_case* c);
ob_residual ob_residual_compute(const ob_case* c);

/* One full pseudo-time step (OrionSolver.cpp:141-209 without the debug
 * block at :156-201).  Returns 0 on success. */
int ob_step(ob_case* c, ob_residual* out);
/* n steps; writes rms,max per step into hist[2*n] (may be NULL). */
int ob_run(ob_case* c, int nsteps, double* hist);

/* LU-SGS building blocks exposed for the lexicographic-vs-wavefront test. */
void ob_lusgs_sweeps(ob_case* c, int b, const double* rhs_soa);

/* The point fields PostProcess::write_block_vts_binary writes for block b
 * (src/postprocess/PostProcess.cpp:183-240), as float: Density, Pressure,
 * Temperature, Mach (idim*jdim*kdim each, interior nodes, i fastest), then
 * Velocity as (u, v, w) per node -- 7 * idim*jdim*kdim values in out. */
void ob_output_fields(const ob_case* c, int b, float* out);

/* Test instrumentation: how often the reference's guard branches were taken
 * since the last reset -- out[0] inviscid left-state and out[1] right-state
 * low-order fallbacks (src/solver/InviscidFluxComputer.cpp:346-355),
 * out[2] updates damped by the pressure-change limiter and out[3] updates
 * refused (old state kept) (src/solver/StateUpdater.cpp:502-538). */
void ob_guard_counts(long* out);
/* Test option: the RMS residual's sum of squares by compensated summation
 * (about the exact sum) instead of the reference's loop order, whose
 * rounding error reaches 1e-12 relative at millions of nodes. */
void ob_set_accurate_sum(int on);
void ob_guard_reset(void);

#ifdef __cplusplus
}
#endif
#endif
