// orion_host.hpp -- C++ host mirror of ORION's operator surface over the C
// ABI of the gfx950 library (include/orion_hip.h).
//
// The namespaces, class names and operator signatures are the reference's
// (orion::core / mesh / bc / preprocess / solver), so code written against
// the reference's solver loop compiles against these with a namespace
// alias.  Every operator forwards to one orion_hip_* entry point and turns a
// non-zero status into std::runtime_error, as the reference throws
// (src/core/OrionArray.hpp:25-27).  State lives on the device; FlowFieldSet
// owns the device context.
//
// The case readers (Plot3D grid, the 1976-format BC file, the param.dat
// namelist) read exactly the files the reference reads.
#pragma once

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/orion_hip.h"

namespace orion_amd {

// Throws std::runtime_error carrying orion_hip_last_error() when rc != 0.
void check(int rc, const char* what);

namespace core {

// The parameters the hot path reads (core::Params, src/core/Params.hpp:9-120).
struct Params {
  struct Inflow {
    double moo = 0.0, reynolds = 0.0, attack = 0.0, sideslip = 0.0;
    double tref = -1.0, twall = 0.0, pref = -1.0, rref = -1.0, vref = 0.0, height = -1.0;
    int ndim = 3;
  } inflow;
  struct Force {
    double lfref = 1.0, lref = 1.0;
  } force;
  struct Filename {
    std::string gridname = "grid.x", bcname = "grid.inp", flowname = "output/";
  } filename;
  struct Control {
    int nstart = 0, nomax = 0, method = 1;
  } control;
  struct Step {
    int ntmst = 0;
    double cfl = 1.0;
    double phydtime = 0.0;
  } step;
  struct FlowType {
    int nvis = 1;
  } flowtype;
  struct Technic {
    int nscheme = 41, csrv = 1;
  } technic;
  struct GclCic {
    int gcl = 1;
  } gcl_cic;

  // Params::load_namelist_file (src/core/Params.cpp:103-275): groups opened
  // by '&name' or '$name' and closed by '/', comma or newline separated
  // key = value items, '!' comments, Fortran 'd' exponents.  Keys the hot
  // path does not read are accepted and ignored.
  void load_namelist_file(const std::string& path);

  // The reference state as preprocess::init_inflow derives it
  // (src/preprocess/Inflow.cpp:300-410): with height >= 0 the standard
  // atmosphere at that height (km) gives temperature, pressure and speed of
  // sound, the density follows from the gas constant of the perfect-gas air
  // model, and the Reynolds number from the reference velocity moo * a and
  // the Sutherland viscosity; with height < 0 two of pref, tref and rref
  // give the third, and the Reynolds number is the input one.  (init_inflow
  // scales it by params.inflow.lref / lfref, which the namelist never sets,
  // so by 1: the $force lengths do not enter.)
  struct Reference {
    double reynolds, tref, pref, rref;
  };
  Reference derive_reference() const;

  // The inputs of the device context: the derived Reynolds number and
  // reference temperature (Sutherland constant 110.4 / tref) with the
  // namelist's other inputs.
  orion_inputs to_inputs() const;
};

}  // namespace core

namespace mesh {

struct Block {
  int idim = 0, jdim = 0, kdim = 0;
  std::vector<double> x, y, z;  // i fastest
  long long npoints() const { return (long long)idim * jdim * kdim; }
};

struct MultiBlockGrid {
  int nblocks = 0;
  std::vector<Block> blocks;
};

// Plot3DReader::read (src/mesh/Plot3DReader.cpp:17-224): Fortran
// sequential records with 32-bit markers -- block count, one (i,j,k) record
// per block, then one record of x, y, z per block in float or double; a
// file that does not read as that is read as the ASCII form (block count,
// the sizes, then every block's x, y and z as numbers separated by white
// space), as the reference falls back.
MultiBlockGrid read_plot3d(const std::string& path);

}  // namespace mesh

namespace bc {

struct BCRegion {  // bc::BCRegion (src/bc/BCData.hpp:8-50), 1-based windows
  int bctype = 0;
  std::array<int, 3> s_st{}, s_ed{}, t_st{}, t_ed{};
  int nbt = 0, ibcwin = 0;
};

struct BlockBC {
  std::string blockname;
  std::vector<BCRegion> regions;
};

struct BCData {
  int flow_solver_id = 1976;
  int nprocs_in_file = 1;
  std::vector<int> block_pid;
  std::vector<BlockBC> block_bc;
};

// BCReader::read_parallel_bc_root_only (src/bc/BCReader.cpp:31-182): the
// 1976 marker, process and block counts, then per block its owner, sizes
// (checked against the grid), name and regions; interface regions carry a
// second line with the target window, block and window number.
BCData read_bc(const std::string& path, const mesh::MultiBlockGrid& grid);

// bc::set_bc_index and bc::prepare_bc_topology (src/bc/BCPreprocess.hpp:8,12):
// the device context compiles the topology when it is created, so these
// only validate what the reference's versions would reject.
void set_bc_index(BCData& bc);
void prepare_bc_topology(BCData& bc);

}  // namespace bc

namespace preprocess {

// The flow field of every block, resident on the device.  Creating it does
// what OrionSolver::preprocess does before the step loop
// (src/core/OrionSolver.cpp:61-93): topology, allocation, metrics, inflow,
// initial field and temperature.
class FlowFieldSet {
 public:
  FlowFieldSet(const mesh::MultiBlockGrid& grid, const bc::BCData& bc, const core::Params& params, int device = 0,
               bool keep_q = false);
  ~FlowFieldSet();
  FlowFieldSet(const FlowFieldSet&) = delete;
  FlowFieldSet& operator=(const FlowFieldSet&) = delete;
  FlowFieldSet(FlowFieldSet&& o) noexcept;
  FlowFieldSet& operator=(FlowFieldSet&& o) noexcept;

  orion_hip_ctx* ctx() const { return ctx_; }
  int nblocks() const { return nblocks_; }
  long long interior_nodes() const;
  // whole field of one block in the reference's padded layout (ORION_FIELD_*)
  std::vector<double> download(int block, int field) const;
  void upload(int block, int field, const std::vector<double>& host);
  void synchronize() const;
  // restart state (every block's padded prim; no counterpart in the
  // reference): save after `step` steps, load returns the saved step
  void save_state(const std::string& path, long long step) const;
  long long load_state(const std::string& path);

 private:
  orion_hip_ctx* ctx_ = nullptr;
  int nblocks_ = 0;
};

// preprocess::update_temperature (src/preprocess/InitialCondition.cpp:91-147)
void update_temperature(const core::Params& params, FlowFieldSet& fs);

}  // namespace preprocess

namespace bc {
void apply_physical_bc(const BCData& bc, preprocess::FlowFieldSet& fs, const core::Params& params);
}

namespace solver {

struct ResidualStats {  // solver::ResidualStats (src/solver/ResidualMonitor.hpp:8-12)
  double rms_residual = 0.0;
  double max_residual = 0.0;
  std::array<int, 5> max_loc{};
};

class HaloExchanger {
 public:
  void exchange_bc(bc::BCData& bc, preprocess::FlowFieldSet& fs);
  void average_interface_residuals(bc::BCData& bc, preprocess::FlowFieldSet& fs);
};

class StateUpdater {
 public:
  static void update_flow_states(preprocess::FlowFieldSet& fs, const core::Params& params);
  static void update_solution(preprocess::FlowFieldSet& fs, const bc::BCData& bc, const core::Params& params);
};

class TimeIntegrator {
 public:
  static void calculate_time_step(preprocess::FlowFieldSet& fs, core::Params& params);
};

class FluxComputer {
 public:
  static void compute_viscous_rhs(preprocess::FlowFieldSet& fs, const core::Params& params);
};

class InviscidFluxComputer {
 public:
  static void compute_inviscid_rhs(preprocess::FlowFieldSet& fs, const core::Params& params);
};

class ResidualMonitor {
 public:
  static ResidualStats compute(const preprocess::FlowFieldSet& fs, const core::Params& params);
};

// One pseudo-time step through the operators above, in OrionSolver::solve's
// order (src/core/OrionSolver.cpp:139-209 without the debug block).
ResidualStats step(bc::BCData& bc, preprocess::FlowFieldSet& fs, core::Params& params, HaloExchanger& halo);

// n steps as one device sequence (captured once as a HIP graph);
// per-step residuals are read back at the end.
std::vector<orion_residual_stats> run(preprocess::FlowFieldSet& fs, int nsteps);

}  // namespace solver

namespace postprocess {

// postprocess::PostProcess (src/postprocess/PostProcess.hpp:10-33)
class PostProcess {
 public:
  // write_solution (src/postprocess/PostProcess.cpp:44-123): the fields are
  // formed on the device and written, with the grid's points, as one VTS
  // per block and the VTM index under params.filename.flowname
  static void write_solution(const preprocess::FlowFieldSet& fs, const mesh::MultiBlockGrid& grid,
                             const core::Params& params, int step);
};

}  // namespace postprocess
}  // namespace orion_amd
