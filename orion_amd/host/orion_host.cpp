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

// call_air (src/preprocess/Inflow.cpp:54-265): the standard atmosphere's
// temperature (K), pressure (Pa), density (kg/m^3) and speed of sound at
// h_km.  The layer table and the expression forms are the reference's, so
// the values agree to the bit: layers with a temperature gradient
//   al = (t_top - t_lo) / span,  t = t_base + al (h - h_lo),
//   p = p_base (t / t_base)^(-g / (r al)),  rho = rho_base (t / t_base)^(-1 - g / (r al)),
// isothermal layers p = p_base exp(-g (h - h_lo) / (r t)) and the same for rho.
// Above 700 km no layer applies: t and p keep the values passed in and the
// density is 0, as in the reference.
static void standard_atmosphere(double h_km, double& t, double& p, double& den, double& a) {
  const double h = h_km * 1000.0;
  const double r = 287.053, g0 = 9.80665, rp = 6.37111e6;
  const double g = std::pow(rp / (rp + h), 2) * g0;
  struct Layer {
    double top;        // the layer holds h <= top
    bool iso;          // isothermal
    double t_lo, t_hi; // gradient: al = (t_hi - t_lo) / span
    double span, h_lo;
    double t_base;     // temperature at h_lo (the base of the power law)
    double p_base, rho_base;
  };
  static const Layer L[] = {
      {11019.0, false, 288.15, 216.65, 11019.0, 0.0, 288.15, 10.1325e2, 1.225},
      {20063.0, true, 216.65, 216.65, 0.0, 11019.0, 216.65, 2.2632e2, 3.6392e-1},
      {32162.0, false, 216.65, 228.65, 32162.0 - 20063.0, 20063.0, 216.65, 5.4747e1, 8.8035e-2},
      {47350.0, false, 228.65, 270.65, 47350.0 - 32162.0, 32162.0, 228.65, 8.6789, 1.3225e-2},
      {52429.0, true, 270.65, 270.65, 0.0, 47350.0, 270.65, 1.1090, 1.4275e-3},
      {61591.0, false, 270.65, 252.65, 61591.0 - 52429.0, 52429.0, 270.65, 5.8997e-1, 7.5943e-4},
      {79994.0, false, 252.65, 180.65, 79994.0 - 61591.0, 61591.0, 252.65, 1.8209e-1, 2.5109e-4},
      {90000.0, true, 180.65, 180.65, 0.0, 79994.0, 180.65, 1.0376e-2, 2.0010e-5},
      {100000.0, false, 180.65, 210.02, 10000.0, 90000.0, 180.65, 1.6437e-3, 3.4165e-6},
      {110000.0, false, 210.02, 257.00, 10000.0, 100000.0, 210.02, 3.0070e-4, 5.6044e-7},
      {120000.0, false, 257.00, 349.49, 10000.0, 110000.0, 257.00, 7.3527e-5, 9.7081e-8},
      {150000.0, false, 349.49, 892.79, 30000.0, 120000.0, 349.49, 2.5209e-5, 2.2222e-8},
      {160000.0, false, 892.79, 1022.20, 10000.0, 150000.0, 892.79, 5.0599e-6, 2.0752e-9},
      {170000.0, false, 1022.20, 1103.40, 10000.0, 160000.0, 1022.20, 3.6929e-6, 1.2336e-9},
      {190000.0, false, 1103.40, 1205.40, 20000.0, 170000.0, 1103.40, 2.7915e-6, 7.8155e-10},
      {230000.0, false, 1205.40, 1322.30, 40000.0, 190000.0, 1205.40, 1.6845e-6, 3.5807e-10},
      {300000.0, false, 1322.30, 1432.10, 70000.0, 230000.0, 1322.30, 6.7138e-7, 1.5640e-10},
      {400000.0, false, 1432.10, 1487.40, 100000.0, 300000.0, 1432.10, 1.8828e-7, 1.9159e-11},
      {500000.0, false, 1487.40, 1499.20, 100000.0, 400000.0, 1487.40, 4.0278e-8, 2.8028e-12},
      {600000.0, false, 1499.20, 1506.10, 100000.0, 500000.0, 1499.20, 1.0949e-8, 5.2148e-13},
      {700000.0, false, 1506.10, 1507.60, 100000.0, 600000.0, 1506.10, 3.4475e-9, 1.1367e-13},
  };
  double rho = 0.0;
  for (const Layer& y : L) {
    if (!(h <= y.top)) continue;
    if (y.iso) {
      t = y.t_base;
      p = y.p_base * std::exp(-g * (h - y.h_lo) / (r * y.t_base));
      rho = y.rho_base * std::exp(-g * (h - y.h_lo) / (r * y.t_base));
    } else {
      const double al = (y.t_hi - y.t_lo) / y.span;
      t = y.t_base + al * (h - y.h_lo);
      p = y.p_base * std::pow(t / y.t_base, -g / (r * al));
      rho = y.rho_base * std::pow(t / y.t_base, -1.0 - g / (r * al));
    }
    break;
  }
  a = std::sqrt(1.4 * r * t);
  p = p * 100.0;
  den = rho;
}

Params::Reference Params::derive_reference() const {
  // setup_perfect_gas_air and the species sums (Inflow.cpp:15-46,326-360):
  // N2 and O2 at 79 % and 21 %, molar masses in kg/mol
  const double ws[2] = {28.0 * 1.0e-3, 32.0 * 1.0e-3}, cn[2] = {0.79, 0.21};
  double mref1 = 0.0;
  for (int s = 0; s < 2; ++s) mref1 += cn[s] * (1.0 / ws[s]);
  const double rjmk = 8.31434;  // Params.hpp:19
  const double R_spec = rjmk * mref1;
  Reference o{inflow.reynolds, inflow.tref, inflow.pref, inflow.rref};
  if (inflow.height < 0.0) {
    // Inflow.cpp:362-390
    const bool p_given = o.pref > 0.0, t_given = o.tref > 0.0, r_given = o.rref > 0.0;
    require(!(p_given && t_given && r_given), "pressure, temperature and density cannot all be given (Inflow.cpp:369)");
    if (p_given && t_given && !r_given) o.rref = o.pref / (o.tref * R_spec);
    else if (r_given && t_given && !p_given) o.pref = o.rref * o.tref * R_spec;
    else if (p_given && r_given && !t_given) o.tref = o.pref / (o.rref * R_spec);
    else throw std::runtime_error("invalid combination of pressure, temperature and density (Inflow.cpp:388)");
    // Inflow.cpp:323-325 with params.inflow.lfref = lref = 1
    o.reynolds = (o.reynolds / 1.0) * 1.0;
  } else {
    // Inflow.cpp:393-407,411-415
    double a = 0.0, den = 0.0;
    standard_atmosphere(inflow.height, o.tref, o.pref, den, a);
    o.rref = o.pref / (o.tref * R_spec);
    const double ccon = (1.0 + 110.4 / 273.15) / (1.0 + 110.4 / o.tref);
    const double visloo = std::sqrt(o.tref / 273.15) * ccon * 1.715e-5;
    const double vref = a * inflow.moo;
    o.reynolds = o.rref * vref * 1.0 / visloo;
  }
  return o;
}

orion_inputs Params::to_inputs() const {
  const Reference ref = derive_reference();
  require(ref.tref > 0.0, "the reference temperature must be positive (the Sutherland constant is 110.4/tref)");
  orion_inputs in{};
  in.moo = inflow.moo;
  in.reynolds = ref.reynolds;
  in.tref = ref.tref;
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
  const std::streampos here = in.tellg();
  in.seekg(0, std::ios::end);
  const std::streamoff left = in.tellg() - here;
  in.seekg(here);
  require((std::streamoff)n1 + 4 <= left, path + ": record longer than the file");
  std::vector<unsigned char> buf(n1);
  in.read(reinterpret_cast<char*>(buf.data()), n1);
  in.read(reinterpret_cast<char*>(&n2), 4);
  require(in.good(), path + ": truncated record");
  require(n1 == n2, path + ": Fortran record length mismatch");
  return buf;
}

// read_ascii_simple (src/mesh/Plot3DReader.cpp:162-208)
static MultiBlockGrid read_plot3d_ascii(const std::string& path) {
  std::ifstream in(path);
  require(in.good(), "Failed to open ASCII grid file: " + path);
  MultiBlockGrid g;
  in >> g.nblocks;
  require(in.good() && g.nblocks > 0, path + ": invalid block count in ASCII");
  g.blocks.resize(g.nblocks);
  for (auto& b : g.blocks) {
    in >> b.idim >> b.jdim >> b.kdim;
    require(in.good() && b.idim > 0 && b.jdim > 0 && b.kdim > 0, path + ": bad block sizes in ASCII");
  }
  for (auto& b : g.blocks) {
    const size_t n = (size_t)b.npoints();
    for (std::vector<double>* c : {&b.x, &b.y, &b.z}) {
      c->resize(n);
      for (size_t i = 0; i < n; ++i) in >> (*c)[i];
    }
    require(!in.fail(), path + ": bad coordinates in ASCII");
  }
  return g;
}

static MultiBlockGrid read_plot3d_binary(const std::string& path);

MultiBlockGrid read_plot3d(const std::string& path) {
  // the binary form first, the ASCII form if that fails (Plot3DReader.cpp:210-224)
  try {
    return read_plot3d_binary(path);
  } catch (const std::exception&) {
    return read_plot3d_ascii(path);
  }
}

static MultiBlockGrid read_plot3d_binary(const std::string& path) {
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
