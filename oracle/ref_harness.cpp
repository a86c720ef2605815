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
// usage: ref_harness <casedir> <nsteps> <outdir> [dump_steps] [vts|time]
// (under mpiexec -n R for R ranks; the BC file's pids place the blocks)
//   outdir/hist.bin   nsteps x 9 doubles: rms, max, max_loc[5], phydtime,
//                     and the RMS again from a compensated sum of the same
//                     terms (the harness's own addition: the reference's
//                     loop-order sum is off the exact sum by up to ~1e-12
//                     relative at millions of nodes) (rank 0; reduced over
//                     all ranks)
//   outdir/s<step>_<stage>.bin   for step <= dump_steps, every block's
//                     [prim 6P][dq 5P][c P][mu P][dt P] after each stage
//                     (one rank only)
//   outdir/params.bin  moo, reynolds, tref, visc, poo, uoo, voo, woo, pref,
//                     rref as init_inflow derived them (rank 0)
//   outdir/final.bin  same layout after the last step (one rank), or
//   outdir/final_b<n>.bin  block n's fields, written by its owner (R > 1)
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

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace orion;

// ResidualMonitor::compute's RMS (src/solver/ResidualMonitor.cpp:38-102)
// with Neumaier's compensated summation of the same squared terms.
static double rms_compensated(preprocess::FlowFieldSet& fs) {
  double sum = 0.0, comp = 0.0, cnt = 0.0;
  for (int nb : fs.local_block_ids) {
    auto& bf = fs.blocks[nb];
    const int ng = fs.ng;
    const int nx = (int)bf.prim.dims()[0], ny = (int)bf.prim.dims()[1], nz = (int)bf.prim.dims()[2];
    for (int k = ng; k < nz - ng; ++k)
      for (int j = ng; j < ny - ng; ++j)
        for (int i = ng; i < nx - ng; ++i) {
          double vol = bf.vol(i, j, k), dt = bf.dt(i, j, k);
          if (vol < 1e-30) vol = 1e-30;
          if (dt < 1e-30) dt = 1e-30;
          const double inv = 1.0 / (dt * vol);
          for (int m = 0; m < 5; ++m) {
            const double r = std::fabs(bf.dq(i, j, k, m)) * inv;
            const double x = r * r, t = sum + x;
            comp += std::fabs(sum) >= std::fabs(x) ? (sum - t) + x : (x - t) + sum;
            sum = t;
          }
          cnt += 5.0;
        }
  }
  double loc[3] = {sum, comp, cnt}, tot[3];
  MPI_Allreduce(loc, tot, 3, MPI_DOUBLE, MPI_SUM, MPI_COMM_WORLD);
  return tot[2] > 0.0 ? std::sqrt((tot[0] + tot[1]) / tot[2]) : 0.0;
}

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
  // time: the CPU baseline, on any number of MPI ranks -- the same start
  // state (init.bin), no dumps; rank 0 writes outdir/time.txt (steps,
  // seconds of the step loop between barriers, ranks) and hist.bin
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
  if (core::Runtime::myid == 0) {
    // the inflow values the step reads, as init_inflow derived them
    const auto& F = params.inflow;
    const double vals[10] = {F.moo, F.reynolds, F.tref, F.visc, F.poo, F.uoo, F.voo, F.woo, F.pref, F.rref};
    std::string p = outdir + "/params.bin";
    if (FILE* f = std::fopen(p.c_str(), "wb")) {
      std::fwrite(vals, sizeof(double), 10, f);
      std::fclose(f);
    }
  }
  preprocess::initialize_flow(params, fs);
  preprocess::update_temperature(params, fs);

  if (FILE* f = std::fopen("init.bin", "rb")) {
    // init.bin holds every block in order; each rank reads its own blocks
    std::vector<long> first(grid.nblocks + 1, 0);
    for (int nb = 0; nb < grid.nblocks; ++nb) {
      const auto& g = grid.blocks[nb];
      first[nb + 1] = first[nb] + 6L * (g.idim + 2 * fs.ng) * (g.jdim + 2 * fs.ng) * (g.kdim + 2 * fs.ng);
    }
    for (int nb : fs.local_block_ids) {
      auto& bf = fs.blocks[nb];
      if (std::fseek(f, first[nb] * (long)sizeof(double), SEEK_SET) != 0 ||
          (long)bf.prim.size() != first[nb + 1] - first[nb] ||
          std::fread(bf.prim.hostPtr(), sizeof(double), bf.prim.size(), f) != bf.prim.size()) {
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
    hist.push_back(timing ? 0.0 : rms_compensated(fs));
  }
  MPI_Barrier(MPI_COMM_WORLD);
  const double t1 = MPI_Wtime();
  if (core::Runtime::myid == 0) {
    std::string p = outdir + "/hist.bin";
    FILE* f = std::fopen(p.c_str(), "wb");
    if (!f) { std::perror(p.c_str()); return 2; }
    std::fwrite(hist.data(), sizeof(double), hist.size(), f);
    std::fclose(f);
  }
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
  if (core::Runtime::nproc == 1) {
    dump_state(outdir + "/final.bin", fs);
  } else {
    for (int nb : fs.local_block_ids) {
      std::string p = outdir + "/final_b" + std::to_string(nb) + ".bin";
      FILE* f = std::fopen(p.c_str(), "wb");
      if (!f) { std::perror(p.c_str()); return 2; }
      auto& bf = fs.blocks[nb];
      std::fwrite(bf.prim.hostPtr(), sizeof(double), bf.prim.size(), f);
      std::fwrite(bf.dq.hostPtr(), sizeof(double), bf.dq.size(), f);
      std::fwrite(bf.c.hostPtr(), sizeof(double), bf.c.size(), f);
      std::fwrite(bf.mu.hostPtr(), sizeof(double), bf.mu.size(), f);
      std::fwrite(bf.dt.hostPtr(), sizeof(double), bf.dt.size(), f);
      std::fclose(f);
    }
  }
  if (vts) postprocess::PostProcess::write_solution(fs, grid, params, nsteps);
  MPI_Finalize();
  return 0;
}
