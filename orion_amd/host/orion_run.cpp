// orion_run -- run ORION's pseudo-time steps on the GPU from a case
// directory in the reference's formats (param.dat, the grid and BC files it
// names), like the reference's solver executable does from its working
// directory (src/main.cpp, src/core/OrionSolver.cpp:61-223).
//
//   orion_run CASEDIR [NSTEPS] [--ops] [--hist FILE] [--device N] [--check]
//             [--restart FILE] [--save FILE] [--no-output]
//
//   NSTEPS   steps to run (default: control.nomax of param.dat)
//   --ops    call the operators one by one (as OrionSolver::solve does)
//            instead of the captured step sequence
//   --hist   write one record of 8 doubles per step: rms, max, the 5 entries
//            of max_loc, phydtime
//   --check  parse the case only and print a summary (no device use): sizes,
//            the derived inflow reference state and a coordinate checksum
//   --restart  start from a state file written by --save (instead of the
//            initial field); the steps continue its step count
//   --save   write the state after the last step
//   --no-output  skip the solution files; by default, like the reference
//            (OrionSolver.cpp:225), the VTS/VTM files of the last step are
//            written under the case's flowname (relative to CASEDIR)
//
// An optional CASEDIR/init.bin overrides the initial primitive field
// (6 x padded points per block, doubles, the reference's layout), after
// which the temperature is recomputed (InitialCondition.cpp:91-147).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>

#include "orion_host.hpp"

using namespace orion_amd;

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s CASEDIR [NSTEPS] [--ops] [--hist FILE] [--device N] [--check]\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  int nsteps = -1, device = 0;
  bool ops = false, check_only = false, output = true;
  std::string hist, restart, save;
  for (int a = 2; a < argc; ++a) {
    const std::string s = argv[a];
    if (s == "--ops") ops = true;
    else if (s == "--check") check_only = true;
    else if (s == "--hist" && a + 1 < argc) hist = argv[++a];
    else if (s == "--device" && a + 1 < argc) device = std::atoi(argv[++a]);
    else if (s == "--restart" && a + 1 < argc) restart = argv[++a];
    else if (s == "--save" && a + 1 < argc) save = argv[++a];
    else if (s == "--no-output") output = false;
    else nsteps = std::atoi(s.c_str());
  }
  try {
    core::Params params;
    params.load_namelist_file(dir + "/param.dat");
    const mesh::MultiBlockGrid grid = mesh::read_plot3d(dir + "/" + params.filename.gridname);
    bc::BCData bcd = bc::read_bc(dir + "/" + params.filename.bcname, grid);
    bc::set_bc_index(bcd);
    bc::prepare_bc_topology(bcd);
    const orion_inputs in = params.to_inputs();
    if (nsteps < 0) nsteps = params.control.nomax;
    if (check_only) {
      long long nodes = 0;
      size_t nreg = 0;
      double xyz_sum = 0.0;  // a checksum of the coordinates as read
      for (const auto& b : grid.blocks) {
        nodes += b.npoints();
        for (const std::vector<double>* c : {&b.x, &b.y, &b.z})
          for (double v : *c) xyz_sum += v;
      }
      for (const auto& b : bcd.block_bc) nreg += b.regions.size();
      const core::Params::Reference ref = params.derive_reference();
      std::printf("{\"blocks\": %d, \"nodes\": %lld, \"regions\": %zu, \"moo\": %.17g, \"reynolds\": %.17g, "
                  "\"tref\": %.17g, \"pref\": %.17g, \"rref\": %.17g, \"cfl\": %.17g, \"nvis\": %d, \"gcl\": %d, "
                  "\"steps\": %d, \"xyz_sum\": %.17g}\n",
                  grid.nblocks, nodes, nreg, in.moo, in.reynolds, in.tref, ref.pref, ref.rref, in.cfl, in.nvis, in.gcl,
                  nsteps, xyz_sum);
      return 0;
    }
    preprocess::FlowFieldSet fs(grid, bcd, params, device);
    long long step0 = 0;
    if (!restart.empty()) {
      step0 = fs.load_state(restart);
    } else {
      std::ifstream init(dir + "/init.bin", std::ios::binary);
      if (init.good()) {
        for (int b = 0; b < fs.nblocks(); ++b) {
          std::vector<double> prim = fs.download(b, ORION_FIELD_PRIM);
          init.read(reinterpret_cast<char*>(prim.data()), (std::streamsize)(prim.size() * sizeof(double)));
          if (!init.good()) throw std::runtime_error("init.bin is shorter than the padded prim fields");
          fs.upload(b, ORION_FIELD_PRIM, prim);
        }
        preprocess::update_temperature(params, fs);
      }
    }
    std::vector<double> rec;
    if (ops) {
      solver::HaloExchanger halo;
      for (int n = 0; n < nsteps; ++n) {
        const solver::ResidualStats r = solver::step(bcd, fs, params, halo);
        std::printf("step %6lld  rms %.10e  max %.10e  at block %d (%d,%d,%d) var %d\n", step0 + n + 1, r.rms_residual,
                    r.max_residual, r.max_loc[0] + 1, r.max_loc[1], r.max_loc[2], r.max_loc[3], r.max_loc[4]);
        rec.insert(rec.end(), {r.rms_residual, r.max_residual, (double)r.max_loc[0], (double)r.max_loc[1],
                               (double)r.max_loc[2], (double)r.max_loc[3], (double)r.max_loc[4], params.step.phydtime});
      }
    } else {
      const auto h = solver::run(fs, nsteps);
      for (int n = 0; n < nsteps; ++n) {
        const auto& r = h[n];
        std::printf("step %6lld  rms %.10e  max %.10e  at block %d (%d,%d,%d) var %d\n", step0 + n + 1, r.rms_residual,
                    r.max_residual, r.max_loc[0] + 1, r.max_loc[1], r.max_loc[2], r.max_loc[3], r.max_loc[4]);
        rec.insert(rec.end(), {r.rms_residual, r.max_residual, (double)r.max_loc[0], (double)r.max_loc[1],
                               (double)r.max_loc[2], (double)r.max_loc[3], (double)r.max_loc[4], r.phydtime});
      }
    }
    const long long last = step0 + nsteps;
    if (!save.empty()) fs.save_state(save, last);
    if (output) {
      // relative flownames are relative to the case directory, the
      // reference's working directory
      core::Params out = params;
      if (!out.filename.flowname.empty() && out.filename.flowname[0] != '/')
        out.filename.flowname = dir + "/" + out.filename.flowname;
      postprocess::PostProcess::write_solution(fs, grid, out, (int)last);
    }
    if (!hist.empty()) {
      std::ofstream out(hist, std::ios::binary);
      out.write(reinterpret_cast<const char*>(rec.data()), (std::streamsize)(rec.size() * sizeof(double)));
      if (!out.good()) throw std::runtime_error("cannot write " + hist);
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "orion_run: %s\n", e.what());
    return 1;
  }
  return 0;
}
