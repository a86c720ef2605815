// ref_harness.cpp -- drives the reference solver's own operators, compiled
// from /root/reference/src by oracle/Makefile into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY.  It reproduces OrionSolver::preprocess
// (src/core/OrionSolver.cpp:32-100) and the step loop of OrionSolver::solve
// (src/core/OrionSolver.cpp:139-209) without the debug block at :156-201,
// which calls MPI_Abort after the first viscous RHS.  The only addition is
// an optional initial-state override (init.bin: prim[6][P] per block) applied
// after update_temperature, followed by update_temperature again, because
// the reference has no way to express a perturbed initial field
// (src/preprocess/InitialCondition.cpp:7-86).
//
// usage: ref_harness <casedir> <nsteps> <outdir> [dump_steps] [vts]
//   outdir/hist.bin   nsteps x 8 doubles: rms, max, max_loc[5], phydtime
//   outdir/s<step>_<stage>.bin   for step <= dump_steps, every block's
//                     [prim 6P][dq 5P][c P][mu P][dt P] after each stage
//   outdir/final.bin  same layout after the last step
//   with "vts": PostProcess::write_solution (src/postprocess/PostProcess.cpp:44-123)
//                     after the last step, as OrionSolver::solve ends
//                     (src/core/OrionSolver.cpp:225), into the case's
//                     flowname directory (relative to casedir)
#include "bc/BCDistributor.hpp"
#include "bc/BCPreprocess.hpp"
#include "bc/BCReader.hpp"
#include "bc/PhysicalBC.hpp"
#include "core/Params.hpp"
#include "core/Runtime.hpp"
#include "mesh/GridDistributor.hpp"
#include "mesh/MetricComputer.hpp"
#include "mesh/Plot3DReader.hpp"
#include "preprocess/FlowField.hpp"
#include "preprocess/Inflow.hpp"
#include "preprocess/InitialCondition.hpp"
#include "postprocess/PostProcess.hpp"
#include "solver/FluxComputer.hpp"
#include "solver/HaloExchanger.hpp"
#include "solver/InviscidFluxComputer.hpp"
#include "solver/ResidualMonitor.hpp"
#include "solver/StateUpdater.hpp"
#include "solver/TimeIntegrator.hpp"

#include <mpi.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace orion;

static void dump_state(const std::string& path, preprocess::FlowFieldSet& fs) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) { std::perror(path.c_str()); std::exit(2); }
  for (int nb : fs.local_block_ids) {
    auto& bf = fs.blocks[nb];
    std::fwrite(bf.prim.hostPtr(), sizeof(double), bf.prim.size(), f);
    std::fwrite(bf.dq.hostPtr(), sizeof(double), bf.dq.size(), f);
    std::fwrite(bf.c.hostPtr(), sizeof(double), bf.c.size(), f);
    std::fwrite(bf.mu.hostPtr(), sizeof(double), bf.mu.size(), f);
    std::fwrite(bf.dt.hostPtr(), sizeof(double), bf.dt.size(), f);
  }
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s casedir nsteps outdir [dump_steps [vts|time]]\n", argv[0]);
    return 2;
  }
  const std::string casedir = argv[1];
  const int nsteps = std::atoi(argv[2]);
  std::string outdir = argv[3];
  const int dump_steps = argc > 4 ? std::atoi(argv[4]) : 0;
  const bool vts = argc > 5 && std::string(argv[5]) == "vts";
  // time: the CPU baseline, on any number of MPI ranks -- the freestream
  // start, no dumps; rank 0 writes outdir/time.txt (steps, seconds of the
  // step loop between barriers, ranks) and nothing else is written
  const bool timing = argc > 5 && std::string(argv[5]) == "time";
  if (outdir[0] != '/') {
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd)) return 2;
    outdir = std::string(cwd) + "/" + outdir;
  }
  core::Runtime::init(argc, argv);
  if (chdir(casedir.c_str()) != 0) { std::perror("chdir"); return 2; }

  core::Params params;
  params.load_namelist_file("param.dat");
  mesh::MultiBlockGrid grid = mesh::Plot3DReader::read(params.filename.gridname, params.inflow.ndim);
  bc::BCData bcd = bc::BCReader::read_parallel_bc_root_only(params.filename.bcname, grid, core::Runtime::nproc,
                                                            {.mbsmarker = 1976, .verbose = false});
  mesh::distrib_grid_fast(grid, bcd.block_pid, core::Runtime::myid, core::Runtime::nproc,
                          {.master = 0, .max_chunk_bytes = 256ull * 1024 * 1024, .verbose = false});
  bc::distrib_bc_fast(bcd, core::Runtime::myid, core::Runtime::nproc, {.master = 0, .verbose = false});
  bc::set_bc_index(bcd);
  bc::prepare_bc_topology(bcd);
  preprocess::FlowFieldSet fs = preprocess::allocate_other_variable(grid, bcd, params, 5, 6);
  preprocess::fill_bc_flag(grid, bcd, fs);
  mesh::compute_grid_metrics(grid, fs, params);
  mesh::check_grid_metrics(grid, fs, bcd);
  preprocess::init_inflow(params, fs);
  preprocess::initialize_flow(params, fs);
  preprocess::update_temperature(params, fs);

  if (FILE* f = timing ? nullptr : std::fopen("init.bin", "rb")) {
    for (int nb : fs.local_block_ids) {
      auto& bf = fs.blocks[nb];
      if (std::fread(bf.prim.hostPtr(), sizeof(double), bf.prim.size(), f) != bf.prim.size()) {
        std::fprintf(stderr, "init.bin too short\n");
        return 2;
      }
    }
    std::fclose(f);
    preprocess::update_temperature(params, fs);
  }

  solver::HaloExchanger halo;
  std::vector<double> hist;
  MPI_Barrier(MPI_COMM_WORLD);
  const double t0 = MPI_Wtime();
  for (int step = 1; step <= nsteps; ++step) {
    const bool dump = !timing && step <= dump_steps;
    auto stage = [&](const char* name) {
      if (dump) dump_state(outdir + "/s" + std::to_string(step) + "_" + name + ".bin", fs);
    };
    halo.exchange_bc(bcd, fs);
    stage("halo");
    bc::apply_physical_bc(bcd, fs, params);
    stage("bc");
    solver::StateUpdater::update_flow_states(fs, params);
    stage("state");
    solver::TimeIntegrator::calculate_time_step(fs, params);
    stage("dt");
    if (params.flowtype.nvis == 1) {
      solver::FluxComputer::compute_viscous_rhs(fs, params);
      stage("visc");
      halo.average_interface_residuals(bcd, fs);
      stage("avg1");
    }
    solver::InviscidFluxComputer::compute_inviscid_rhs(fs, params);
    stage("inv");
    halo.average_interface_residuals(bcd, fs);
    stage("avg2");
    solver::StateUpdater::update_solution(fs, bcd, params);
    stage("upd");
    solver::ResidualStats st = solver::ResidualMonitor::compute(fs, params);
    hist.push_back(st.rms_residual);
    hist.push_back(st.max_residual);
    for (int m = 0; m < 5; ++m) hist.push_back((double)st.max_loc[m]);
    hist.push_back(params.step.phydtime);
  }
  MPI_Barrier(MPI_COMM_WORLD);
  const double t1 = MPI_Wtime();
  if (timing) {
    if (core::Runtime::myid == 0) {
      std::string p = outdir + "/time.txt";
      FILE* f = std::fopen(p.c_str(), "w");
      if (!f) { std::perror(p.c_str()); return 2; }
      std::fprintf(f, "%d %.6f %d\n", nsteps, t1 - t0, core::Runtime::nproc);
      std::fclose(f);
    }
    MPI_Finalize();
    return 0;
  }
  {
    std::string p = outdir + "/hist.bin";
    FILE* f = std::fopen(p.c_str(), "wb");
    std::fwrite(hist.data(), sizeof(double), hist.size(), f);
    std::fclose(f);
  }
  dump_state(outdir + "/final.bin", fs);
  if (vts) postprocess::PostProcess::write_solution(fs, grid, params, nsteps);
  MPI_Finalize();
  return 0;
}
