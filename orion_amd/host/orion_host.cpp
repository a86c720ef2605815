// orion_host.cpp -- see orion_host.hpp.
#include "orion_host.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>

namespace orion_amd {

void check(int rc, const char* what) {
  if (rc != 0) throw std::runtime_error(std::string(what) + ": " + orion_hip_last_error());
}

static void require(bool ok, const std::string& msg) {
  if (!ok) throw std::runtime_error(msg);
}

// ---------------------------------------------------------------------------
// core::Params
// ---------------------------------------------------------------------------
namespace core {

static std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

static std::string lower(std::string s) {
  for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
  return s;
}

static double to_double(std::string v) {
  for (auto& ch : v)
    if (ch == 'd' || ch == 'D') ch = 'E';  // Fortran double exponent
  size_t used = 0;
  const double x = std::stod(v, &used);
  require(trim(v.substr(used)).empty(), "namelist: bad number '" + v + "'");
  return x;
}

static int to_int(const std::string& v) {
  size_t used = 0;
  const int x = std::stoi(v, &used);
  require(trim(v.substr(used)).empty(), "namelist: bad integer '" + v + "'");
  return x;
}

static std::string unquote(std::string v) {
  v = trim(v);
  if (v.size() >= 2 && (v.front() == '\'' || v.front() == '"') && v.back() == v.front()) return v.substr(1, v.size() - 2);
  return v;
}

// items of one line: comma separated outside quotes
static std::vector<std::string> items(const std::string& line) {
  std::vector<std::string> out;
  std::string cur;
  char q = 0;
  for (char ch : line) {
    if (q) {
      cur += ch;
      if (ch == q) q = 0;
    } else if (ch == '\'' || ch == '"') {
      q = ch;
      cur += ch;
    } else if (ch == ',') {
      if (!trim(cur).empty()) out.push_back(trim(cur));
      cur.clear();
    } else {
      cur += ch;
    }
  }
  if (!trim(cur).empty()) out.push_back(trim(cur));
  return out;
}

void Params::load_namelist_file(const std::string& path) {
  std::ifstream in(path);
  require(in.good(), "Failed to open param file: " + path);
  std::string group, line;
  auto assign = [&](const std::string& g, const std::string& k, const std::string& v) {
    if (g == "inflow") {
      if (k == "moo") inflow.moo = to_double(v);
      else if (k == "reynolds") inflow.reynolds = to_double(v);
      else if (k == "attack") inflow.attack = to_double(v);
      else if (k == "sideslip") inflow.sideslip = to_double(v);
      else if (k == "tref") inflow.tref = to_double(v);
      else if (k == "twall") inflow.twall = to_double(v);
      else if (k == "pref") inflow.pref = to_double(v);
      else if (k == "rref") inflow.rref = to_double(v);
      else if (k == "vref") inflow.vref = to_double(v);
      else if (k == "height") inflow.height = to_double(v);
      else if (k == "ndim") inflow.ndim = to_int(v);
    } else if (g == "force") {
      if (k == "lfref") force.lfref = to_double(v);
      else if (k == "lref") force.lref = to_double(v);
    } else if (g == "filename") {
      if (k == "gridname") filename.gridname = unquote(v);
      else if (k == "bcname") filename.bcname = unquote(v);
      else if (k == "flowname") filename.flowname = unquote(v);
    } else if (g == "control") {
      if (k == "nstart") control.nstart = to_int(v);
      else if (k == "nomax") control.nomax = to_int(v);
      else if (k == "method") control.method = to_int(v);
    } else if (g == "step") {
      if (k == "ntmst") step.ntmst = to_int(v);
      else if (k == "cfl") step.cfl = to_double(v);
    } else if (g == "flowtype") {
      if (k == "nvis") flowtype.nvis = to_int(v);
    } else if (g == "technic") {
      if (k == "nscheme") technic.nscheme = to_int(v);
      else if (k == "csrv") technic.csrv = to_int(v);
    } else if (g == "gcl_cic") {
      if (k == "gcl") gcl_cic.gcl = to_int(v);
    }
  };
  while (std::getline(in, line)) {
    const size_t bang = line.find('!');
    if (bang != std::string::npos) line.resize(bang);
    line = trim(line);
    if (line.empty()) continue;
    if (line[0] == '&' || line[0] == '$') {
      group = lower(trim(line.substr(1)));
      continue;
    }
    if (line == "/") {
      group.clear();
      continue;
    }
    if (group.empty()) continue;
    for (const auto& it : items(line)) {
      const size_t eq = it.find('=');
      if (eq == std::string::npos) continue;
      const std::string k = lower(trim(it.substr(0, eq))), v = trim(it.substr(eq + 1));
      try {
        assign(group, k, v);
      } catch (const std::exception& e) {
        throw std::runtime_error("param file " + path + ": " + group + "." + k + " = " + v + ": " + e.what());
      }
    }
  }
}

orion_inputs Params::to_inputs() const {
  require(inflow.height < 0.0, "height >= 0 selects the standard atmosphere, which this build does not implement");
  const bool p_given = inflow.pref > 0.0, t_given = inflow.tref > 0.0, r_given = inflow.rref > 0.0;
  require(!(p_given && t_given && r_given), "pressure, temperature and density cannot all be given (Inflow.cpp:367)");
  require(t_given, "this build needs tref > 0 (the temperature sets the Sutherland constant 110.4/tref)");
  orion_inputs in{};
  in.moo = inflow.moo;
  // Inflow.cpp:323-325
  in.reynolds = (std::fabs(force.lfref) > 1e-30) ? (inflow.reynolds / force.lfref) * force.lref : inflow.reynolds;
  in.tref = inflow.tref;
  in.twall = inflow.twall;
  in.cfl = step.cfl;
  in.attack = inflow.attack;
  in.sideslip = inflow.sideslip;
  in.csrv = technic.csrv;
  in.nvis = flowtype.nvis;
  in.method = control.method;
  in.ntmst = step.ntmst;
  in.gcl = gcl_cic.gcl;
  return in;
}

}  // namespace core

// ---------------------------------------------------------------------------
// mesh::read_plot3d
// ---------------------------------------------------------------------------
namespace mesh {

static std::vector<unsigned char> record(std::ifstream& in, const std::string& path) {
  uint32_t n1 = 0, n2 = 0;
  in.read(reinterpret_cast<char*>(&n1), 4);
  require(in.good(), path + ": end of file or bad record header");
  std::vector<unsigned char> buf(n1);
  in.read(reinterpret_cast<char*>(buf.data()), n1);
  in.read(reinterpret_cast<char*>(&n2), 4);
  require(in.good(), path + ": truncated record");
  require(n1 == n2, path + ": Fortran record length mismatch");
  return buf;
}

MultiBlockGrid read_plot3d(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  require(in.good(), "Failed to open grid file: " + path);
  MultiBlockGrid g;
  auto r = record(in, path);
  require(r.size() == 4, path + ": first record must hold the block count");
  std::memcpy(&g.nblocks, r.data(), 4);
  require(g.nblocks > 0, path + ": invalid block count");
  g.blocks.resize(g.nblocks);
  for (auto& b : g.blocks) {
    r = record(in, path);
    require(r.size() == 12, path + ": block size record must hold 3 integers");
    int d[3];
    std::memcpy(d, r.data(), 12);
    b.idim = d[0];
    b.jdim = d[1];
    b.kdim = d[2];
    require(b.idim > 0 && b.jdim > 0 && b.kdim > 0, path + ": invalid block dims");
  }
  for (auto& b : g.blocks) {
    const size_t n = (size_t)b.npoints();
    r = record(in, path);
    std::vector<double> xyz(3 * n);
    if (r.size() == 3 * n * sizeof(double)) {
      std::memcpy(xyz.data(), r.data(), r.size());
    } else if (r.size() == 3 * n * sizeof(float)) {
      std::vector<float> f(3 * n);
      std::memcpy(f.data(), r.data(), r.size());
      for (size_t i = 0; i < 3 * n; ++i) xyz[i] = (double)f[i];
    } else {
      throw std::runtime_error(path + ": unsupported coordinate record size " + std::to_string(r.size()));
    }
    b.x.assign(xyz.begin(), xyz.begin() + n);
    b.y.assign(xyz.begin() + n, xyz.begin() + 2 * n);
    b.z.assign(xyz.begin() + 2 * n, xyz.end());
  }
  return g;
}

}  // namespace mesh

// ---------------------------------------------------------------------------
// bc::read_bc and checks
// ---------------------------------------------------------------------------
namespace bc {

BCData read_bc(const std::string& path, const mesh::MultiBlockGrid& grid) {
  std::ifstream in(path);
  require(in.good(), "Failed to open bc file: " + path);
  BCData out;
  in >> out.flow_solver_id;
  require(in.good(), path + ": bad read of the solver marker");
  require(out.flow_solver_id == 1976, path + ": BC file marker " + std::to_string(out.flow_solver_id) + ", expect 1976");
  int nblocks = 0;
  in >> out.nprocs_in_file >> nblocks;
  require(in.good(), path + ": bad read of the process / block counts");
  require(nblocks == grid.nblocks, path + ": BC blocks != grid blocks");
  out.block_pid.resize(nblocks);
  out.block_bc.resize(nblocks);
  for (int b = 0; b < nblocks; ++b) {
    int dims[3];
    in >> out.block_pid[b] >> dims[0] >> dims[1] >> dims[2];
    require(in.good(), path + ": bad read of block " + std::to_string(b + 1));
    const auto& gb = grid.blocks[b];
    require(dims[0] == gb.idim && dims[1] == gb.jdim && dims[2] == gb.kdim,
            path + ": BC dims mismatch at block " + std::to_string(b + 1));
    std::string line;
    std::getline(in, line);  // rest of the size line
    while (std::getline(in, line) && core::trim(line).empty()) {
    }
    require(in.good(), path + ": missing block name of block " + std::to_string(b + 1));
    out.block_bc[b].blockname = core::trim(line);
    int nr = 0;
    in >> nr;
    require(in.good() && nr >= 0, path + ": bad region count of block " + std::to_string(b + 1));
    out.block_bc[b].regions.resize(nr);
    for (int r = 0; r < nr; ++r) {
      BCRegion& g = out.block_bc[b].regions[r];
      in >> g.s_st[0] >> g.s_ed[0] >> g.s_st[1] >> g.s_ed[1] >> g.s_st[2] >> g.s_ed[2] >> g.bctype;
      require(in.good(), path + ": bad region header, block " + std::to_string(b + 1) + " region " + std::to_string(r + 1));
      if (g.bctype < 0) {
        in >> g.t_st[0] >> g.t_ed[0] >> g.t_st[1] >> g.t_ed[1] >> g.t_st[2] >> g.t_ed[2] >> g.nbt >> g.ibcwin;
        require(in.good(), path + ": bad connect line, block " + std::to_string(b + 1) + " region " + std::to_string(r + 1));
        const bool flat = g.s_st[0] == g.s_ed[0] || g.s_st[1] == g.s_ed[1] || g.s_st[2] == g.s_ed[2];
        require(flat, "BC region " + std::to_string(r + 1) + " in block " + std::to_string(b + 1) +
                          " is an interface but not a flat plane (BCReader.cpp:162-166)");
        require(g.ibcwin > 0, "BC region " + std::to_string(r + 1) + " in block " + std::to_string(b + 1) +
                                  " has an invalid target window (BCReader.cpp:167-171)");
      }
    }
  }
  return out;
}

void set_bc_index(BCData& bc) {
  for (size_t b = 0; b < bc.block_bc.size(); ++b)
    for (const auto& r : bc.block_bc[b].regions) {
      const int t = r.bctype;
      require(t == -1 || t == 2 || t == 3 || t == 4 || t == 5 || t == 6,
              "block " + std::to_string(b + 1) + ": bctype " + std::to_string(t) +
                  " not among {-1,4,5,6,2,3} (BCPreprocess.cpp:196)");
    }
}

void prepare_bc_topology(BCData& bc) {
  const int nb = (int)bc.block_bc.size();
  for (int b = 0; b < nb; ++b)
    for (const auto& r : bc.block_bc[b].regions)
      if (r.bctype < 0) {
        require(r.nbt >= 1 && r.nbt <= nb, "block " + std::to_string(b + 1) + ": interface target block out of range");
        require(r.ibcwin <= (int)bc.block_bc[r.nbt - 1].regions.size(),
                "block " + std::to_string(b + 1) + ": interface target window out of range");
      }
}

}  // namespace bc

// ---------------------------------------------------------------------------
// preprocess::FlowFieldSet
// ---------------------------------------------------------------------------
namespace preprocess {

FlowFieldSet::FlowFieldSet(const mesh::MultiBlockGrid& grid, const bc::BCData& bc, const core::Params& params,
                           int device, bool keep_q) {
  require((int)bc.block_bc.size() == grid.nblocks, "grid and BC data disagree on the block count");
  std::vector<int> dims;
  std::vector<double> xyz;
  for (const auto& b : grid.blocks) {
    dims.insert(dims.end(), {b.idim, b.jdim, b.kdim});
    xyz.insert(xyz.end(), b.x.begin(), b.x.end());
    xyz.insert(xyz.end(), b.y.begin(), b.y.end());
    xyz.insert(xyz.end(), b.z.begin(), b.z.end());
  }
  std::vector<orion_region_desc> regs;
  for (int b = 0; b < grid.nblocks; ++b)
    for (const auto& r : bc.block_bc[b].regions) {
      orion_region_desc d{};
      d.block = b;
      d.bctype = r.bctype;
      for (int m = 0; m < 3; ++m) {
        d.s_st[m] = r.s_st[m];
        d.s_ed[m] = r.s_ed[m];
        d.t_st[m] = r.t_st[m];
        d.t_ed[m] = r.t_ed[m];
      }
      d.nbt = r.nbt;
      d.ibcwin = r.ibcwin;
      regs.push_back(d);
    }
  const orion_inputs in = params.to_inputs();
  check(orion_hip_backend_init(device), "backend init");
  check(orion_hip_ctx_create(grid.nblocks, dims.data(), xyz.data(), (int)regs.size(), regs.data(), &in, keep_q ? 1 : 0,
                             &ctx_),
        "FlowFieldSet");
  nblocks_ = grid.nblocks;
}

FlowFieldSet::~FlowFieldSet() {
  if (ctx_) orion_hip_ctx_destroy(ctx_);
}

FlowFieldSet::FlowFieldSet(FlowFieldSet&& o) noexcept : ctx_(o.ctx_), nblocks_(o.nblocks_) { o.ctx_ = nullptr; }

FlowFieldSet& FlowFieldSet::operator=(FlowFieldSet&& o) noexcept {
  if (this != &o) {
    if (ctx_) orion_hip_ctx_destroy(ctx_);
    ctx_ = o.ctx_;
    nblocks_ = o.nblocks_;
    o.ctx_ = nullptr;
  }
  return *this;
}

long long FlowFieldSet::interior_nodes() const { return orion_hip_interior_nodes(ctx_); }

std::vector<double> FlowFieldSet::download(int block, int field) const {
  double* p = nullptr;
  long long n = 0;
  check(orion_hip_field(ctx_, block, field, &p, &n), "field");
  std::vector<double> h((size_t)n);
  check(orion_hip_download(ctx_, block, field, h.data(), n), "download");
  return h;
}

void FlowFieldSet::upload(int block, int field, const std::vector<double>& host) {
  check(orion_hip_upload(ctx_, block, field, host.data(), (long long)host.size()), "upload");
}

void FlowFieldSet::synchronize() const { check(orion_hip_synchronize(ctx_), "synchronize"); }

void FlowFieldSet::save_state(const std::string& path, long long step) const {
  check(orion_hip_save_state(ctx_, path.c_str(), step), "save_state");
}

long long FlowFieldSet::load_state(const std::string& path) {
  long long step = 0;
  check(orion_hip_load_state(ctx_, path.c_str(), &step), "load_state");
  return step;
}

void update_temperature(const core::Params&, FlowFieldSet& fs) {
  check(orion_hip_update_temperature(fs.ctx()), "update_temperature");
}

}  // namespace preprocess

// ---------------------------------------------------------------------------
// operators
// ---------------------------------------------------------------------------
namespace bc {
void apply_physical_bc(const BCData&, preprocess::FlowFieldSet& fs, const core::Params&) {
  check(orion_hip_apply_physical_bc(fs.ctx()), "apply_physical_bc");
}
}  // namespace bc

namespace solver {

void HaloExchanger::exchange_bc(bc::BCData&, preprocess::FlowFieldSet& fs) {
  check(orion_hip_exchange_bc(fs.ctx()), "exchange_bc");
}
void HaloExchanger::average_interface_residuals(bc::BCData&, preprocess::FlowFieldSet& fs) {
  check(orion_hip_average_interface_residuals(fs.ctx()), "average_interface_residuals");
}
void StateUpdater::update_flow_states(preprocess::FlowFieldSet& fs, const core::Params&) {
  check(orion_hip_update_flow_states(fs.ctx()), "update_flow_states");
}
void StateUpdater::update_solution(preprocess::FlowFieldSet& fs, const bc::BCData&, const core::Params&) {
  check(orion_hip_update_solution(fs.ctx()), "update_solution");
}
void TimeIntegrator::calculate_time_step(preprocess::FlowFieldSet& fs, core::Params& params) {
  double phydtime = 0.0;
  check(orion_hip_calculate_time_step(fs.ctx(), &phydtime), "calculate_time_step");
  params.step.phydtime = phydtime;  // TimeIntegrator.cpp:175-177
}
void FluxComputer::compute_viscous_rhs(preprocess::FlowFieldSet& fs, const core::Params&) {
  check(orion_hip_compute_viscous_rhs(fs.ctx()), "compute_viscous_rhs");
}
void InviscidFluxComputer::compute_inviscid_rhs(preprocess::FlowFieldSet& fs, const core::Params&) {
  check(orion_hip_compute_inviscid_rhs(fs.ctx()), "compute_inviscid_rhs");
}
ResidualStats ResidualMonitor::compute(const preprocess::FlowFieldSet& fs, const core::Params&) {
  orion_residual_stats s{};
  check(orion_hip_residual(fs.ctx(), &s), "ResidualMonitor::compute");
  ResidualStats r;
  r.rms_residual = s.rms_residual;
  r.max_residual = s.max_residual;
  for (int i = 0; i < 5; ++i) r.max_loc[i] = s.max_loc[i];
  return r;
}

ResidualStats step(bc::BCData& bc, preprocess::FlowFieldSet& fs, core::Params& params, HaloExchanger& halo) {
  halo.exchange_bc(bc, fs);
  bc::apply_physical_bc(bc, fs, params);
  StateUpdater::update_flow_states(fs, params);
  TimeIntegrator::calculate_time_step(fs, params);
  if (params.flowtype.nvis == 1) {
    FluxComputer::compute_viscous_rhs(fs, params);
    halo.average_interface_residuals(bc, fs);
  }
  InviscidFluxComputer::compute_inviscid_rhs(fs, params);
  halo.average_interface_residuals(bc, fs);
  StateUpdater::update_solution(fs, bc, params);
  return ResidualMonitor::compute(fs, params);
}

std::vector<orion_residual_stats> run(preprocess::FlowFieldSet& fs, int nsteps) {
  std::vector<orion_residual_stats> h((size_t)std::max(nsteps, 0));
  check(orion_hip_run(fs.ctx(), nsteps, h.data()), "run");
  return h;
}

}  // namespace solver

namespace postprocess {

void PostProcess::write_solution(const preprocess::FlowFieldSet& fs, const mesh::MultiBlockGrid& grid,
                                 const core::Params& params, int step) {
  std::vector<double> xyz;
  for (const auto& b : grid.blocks)
    for (const auto* a : {&b.x, &b.y, &b.z}) xyz.insert(xyz.end(), a->begin(), a->end());
  check(orion_hip_write_solution(fs.ctx(), params.filename.flowname.c_str(), step, xyz.data()), "write_solution");
}

}  // namespace postprocess
}  // namespace orion_amd
