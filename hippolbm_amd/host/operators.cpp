// operators.cpp -- hippoLBM's hot-path operators over the C ABI; see operators.hpp.
#include "operators.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace hlbm_host {

namespace {

void check(int rc, const char* what) {
  if (rc != HLBM_OK) throw OperatorError(std::string(what) + ": " + hlbm_last_error());
}

const char* PLANES[6] = {"plan_xy_0", "plan_xy_l", "plan_xz_0", "plan_xz_l", "plan_yz_0", "plan_yz_l"};

double num_or(const msp::Node& args, const char* key, double dflt) {
  const msp::Node* n = args.find(key);
  return n && !n->is_null() ? n->number() : dflt;
}

std::vector<double> vec3(const msp::Node& n) {
  std::vector<double> v = n.numbers();
  if (v.size() != 3) throw OperatorError("expected 3 components");
  return v;
}

// [[x0,y0,z0],[x1,y1,z1]]
void aabb(const msp::Node& n, double out[6]) {
  if (!n.is_seq() || n.items.size() != 2) throw OperatorError("bounds must be [[x0,y0,z0],[x1,y1,z1]]");
  const std::vector<double> a = vec3(n.items[0]), b = vec3(n.items[1]);
  for (int d = 0; d < 3; d++) {
    out[d] = a[d];
    out[3 + d] = b[d];
  }
}

std::vector<std::string> regions(const msp::Node& args, const std::string& op) {
  const msp::Node* r = args.find("regions");
  if (!r || !r->is_seq()) throw OperatorError(op + ": the regions slot is required");
  std::vector<std::string> out;
  for (const auto& it : r->items) out.push_back(it.text);
  return out;
}

// ---- domain (src/grid/domain_lbm.cpp:47-48 and its checks) ----------------
struct Domain : OperatorNode {
  double lo[3], hi[3];
  int cells[3];
  bool per[3];
  explicit Domain(const msp::Node& a) {
    double b[6];
    aabb(a.at("bounds"), b);
    const std::vector<double> c = a.at("cell_dims").numbers();
    if (c.size() != 3) throw OperatorError("cell_dims needs 3 entries");
    const msp::Node* p = a.find("periodic");
    for (int d = 0; d < 3; d++) {
      lo[d] = b[d];
      hi[d] = b[3 + d];
      cells[d] = (int)c[d];
      per[d] = p ? p->items.at(d).boolean() : false;
    }
  }
  void execute(Slots& s) override {
    double res[3];
    for (int d = 0; d < 3; d++) res[d] = (hi[d] - lo[d]) / double(cells[d]);
    if (res[0] != res[1] || res[0] != res[2])
      throw OperatorError("[Error, domain], Dx is not the same for all dimension");
    for (int d = 0; d < 3; d++)
      if (lo[d] + cells[d] * res[0] != hi[d])
        throw OperatorError("[Error, domain], The resolution slot and bounds slot mismatch.");
    for (int d = 0; d < 3; d++) {
      s.bounds[0][d] = lo[d];
      s.bounds[1][d] = hi[d];
      s.cell_dims[d] = cells[d];
      s.periodic[d] = per[d];
      s.nodes[d] = cells[d] + 1;
    }
    s.dx = res[0];
    s.has_domain = true;
  }
};

// ---- lbm_parameters (src/grid/lbm_parameters.cpp:73-99) --------------------
struct LbmParametersOp : OperatorNode {
  std::map<std::string, double> in;
  std::vector<double> F = {0.0, 0.0, 0.0};
  explicit LbmParametersOp(const msp::Node& a) {
    for (const char* k : {"nuth", "celerity", "avg_rho", "tau", "dt"})
      if (const msp::Node* n = a.find(k)) in[k] = n->number();
    if (const msp::Node* n = a.find("Fext")) F = vec3(*n);
  }
  void execute(Slots& s) override {
    if (!s.has_domain) throw OperatorError("lbm_parameters: the domain is not defined");
    const double dx = s.dx;
    const double nuth = in.count("nuth") ? in["nuth"] : 1e-4;
    const double c = in.count("celerity") ? in["celerity"] : 1.0;
    const double dt = in.count("dt") ? in["dt"] : 0.0;
    LBMParameters& P = s.params;
    double c_;
    if (dt > 0.0) {
      P.dtLB = dt;
      P.tau = 3 * nuth * P.dtLB / (dx * dx) + 0.5;
      P.nu = (P.tau - 0.5) / 3.0;
      c_ = dx / P.dtLB;
    } else if (in.count("tau")) {
      P.tau = in["tau"];
      P.dtLB = dx * dx * (P.tau - 0.5) / (3.0 * nuth);
      P.nu = nuth * P.dtLB / (dx * dx);
      c_ = dx / P.dtLB;
    } else {
      c_ = c;
      P.dtLB = dx / c_;
      P.nu = nuth * P.dtLB / (dx * dx);
      P.tau = 3. * P.nu + 0.5;
    }
    P.celerity = c_;
    P.nuth = nuth;
    if (in.count("avg_rho")) P.avg_rho = in["avg_rho"];
    const double dxc = c_ * P.dtLB;
    for (int d = 0; d < 3; d++) P.Fext[d] = F[d] * P.dtLB * P.dtLB / dxc;
    s.has_params = true;
  }
};

// ---- print_domain (src/grid/print_domain.cpp:58-83) -------------------------
// Informational.  Boxes in the reference's global (Real) and local (Extend)
// coordinates of a single rank; the AABB is written as [(x,y,z),(x,y,z)]
// (Onika's stream format for it is not in the reference tree).
struct PrintDomain : OperatorNode {
  void execute(Slots& s) override {
    if (s.quiet) return;
    if (!s.has_domain) throw OperatorError("print_domain: the domain is not defined");
    const int* n = s.nodes;
    std::printf("=================================\n");
    std::printf("== Domain size: [(%g,%g,%g),(%g,%g,%g)]\n", s.bounds[0][0], s.bounds[0][1], s.bounds[0][2],
                s.bounds[1][0], s.bounds[1][1], s.bounds[1][2]);
    std::printf("== Dx: %g\n", s.dx);
    std::printf("== Grid size: (LX: %d, LY: %d, LZ %d)\n", n[0], n[1], n[2]);
    std::printf("=================================\n");
    std::printf("== Grid info \n= Rank 0 \n= Real Grid:\n");
    std::printf(" inf_: 0 0 0\n sup_: %d %d %d\n", n[0] - 1, n[1] - 1, n[2] - 1);
    std::printf("= Extended Grid: \n");
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {  // make_domain.hpp:178-188: All on periodic axes, Real otherwise (local)
      lo[d] = s.periodic[d] ? 0 : 2;
      hi[d] = s.periodic[d] ? n[d] + 3 : n[d] + 1;
    }
    std::printf(" inf_: %d %d %d\n sup_: %d %d %d\n", lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);
    std::printf("=================================\n");
    std::fflush(stdout);
  }
};

// ---- define_grid_3dq19 (src/grid/define_lbm_fields.cpp:58-86) --------------
struct DefineGrid : OperatorNode {
  void execute(Slots& s) override {
    if (!s.has_domain || !s.has_params) throw OperatorError("define_grid_3dq19: domain and lbm_parameters first");
    if (s.dom) return;
    hlbm_domain_desc d{};
    for (int k = 0; k < 3; k++) {
      d.nodes[k] = s.nodes[k];
      d.periodic[k] = s.periodic[k] ? 1 : 0;
      d.fext[k] = s.params.Fext[k];
    }
    d.storage = s.storage;
    d.pattern = s.pattern;
    d.collision = s.collision;
    d.tau = s.params.tau;
    d.dx = s.dx;
    if (!s.mrt_rates.empty()) {
      d.mrt_custom = 1;
      for (int k = 0; k < 10; k++) d.mrt_rates[k] = s.mrt_rates[k];
    }
    check(hlbm_domain_create(&d, &s.dom), "define_grid_3dq19");
  }
};

struct InitObst : OperatorNode {
  void execute(Slots& s) override { check(hlbm_init_obst(s.dom), "init_obst"); }
};

// ---- set_distribution (src/grid/include/hippoLBM/grid/set_distribution.hpp:38-45)
struct SetDistribution : OperatorNode {
  double value = 1.0;
  bool has_bounds = false;
  double b[6];
  explicit SetDistribution(const msp::Node& a) {
    value = num_or(a, "value", 1.0);
    if (const msp::Node* n = a.find("bounds")) {
      aabb(*n, b);
      has_bounds = true;
    }
  }
  void execute(Slots& s) override {
    if (s.in_loop && s.stage == HLBM_STAGE_START) {
      // start_iteration: the step kernel applies it to every step of the plan
      hlbm_bc e{};
      e.kind = HLBM_BC_SET_DISTRIBUTION;
      e.stage = HLBM_STAGE_START;
      e.plane = has_bounds ? 1 : 0;
      e.U[0] = value;
      if (has_bounds)
        for (int k = 0; k < 6; k++) e.bounds[k] = b[k];
      s.recording.push_back(e);
      return;
    }
    check(hlbm_set_distribution(s.dom, value, has_bounds ? b : nullptr), "set_distribution");
    if (!has_bounds) s.set_dist_value = value;
  }
};

// ---- init_double_couette (src/prepro/init_double_couette.cu:71-120) ----------
struct InitDoubleCouette : OperatorNode {
  std::vector<double> U;
  std::string dimension;
  explicit InitDoubleCouette(const msp::Node& a) {
    if (!a.find("U") || !a.find("dimension")) throw OperatorError("init_double_couette: U and dimension are required");
    U = vec3(a.at("U"));
    dimension = a.at("dimension").text;
  }
  void execute(Slots& s) override {
    const int dim = dimension == "X" ? 0 : dimension == "Y" ? 1 : dimension == "Z" ? 2 : -1;
    if (dim < 0) throw OperatorError("[init_double_couette] Please, select a valid dimension \"X\", \"Y\", or \"Z\".");
    const double Uc[3] = {U[0] / s.params.celerity, U[1] / s.params.celerity, U[2] / s.params.celerity};
    if (dim == 2 && !s.quiet) std::printf("Prepro double couette starting ... dim Z\n");
    check(hlbm_init_double_couette(s.dom, dim, Uc), "init_double_couette");
    if (dim == 2 && !s.quiet) std::printf("Prepro double couette ending ... dim Z \n");
  }
};

// ---- obstacles (src/obstacle/register_obstacles.cpp, update_obstacles.cu:32-72)
struct InitObstacles : OperatorNode {
  void execute(Slots& s) override { s.obstacles.clear(); }
};

struct RegisterBall : OperatorNode {
  long id;
  hlbm_obstacle o{};
  explicit RegisterBall(const msp::Node& a) {
    id = a.at("id").integer();
    const std::vector<double> c = vec3(a.at("center"));
    o.kind = 0;
    for (int d = 0; d < 3; d++) o.p0[d] = c[d];
    o.radius = a.at("radius").number();
  }
  void execute(Slots& s) override { s.obstacles[id] = o; }
};

struct RegisterWall : OperatorNode {
  long id;
  hlbm_obstacle o{};
  explicit RegisterWall(const msp::Node& a) {
    id = a.at("id").integer();
    double b[6];
    aabb(a.at("bounds"), b);
    o.kind = 1;
    for (int d = 0; d < 3; d++) {
      o.p0[d] = b[d];
      o.p1[d] = b[3 + d];
    }
  }
  void execute(Slots& s) override { s.obstacles[id] = o; }
};

// ---- register_quadrics (src/obstacle/register_quadrics.cpp:72-84) ----------
// Q' = M^-T Q M^-1 for the transform list M (applied in order).  The named
// shapes (sphere, cyl{x,y,z}, cone{x,y,z}, plane{x,y,z}) and the transforms
// (scale, translate, xrot/yrot/zrot in radians) restate Onika's YAML forms,
// which are not in the reference tree.  M^-1 is composed from each
// transform's inverse in reverse order; products sum k = 0..3 in order, as
// hippolbm_amd/quadrics.py does.
using Mat4 = std::array<std::array<double, 4>, 4>;

Mat4 mat_identity() {
  Mat4 m{};
  for (int i = 0; i < 4; i++) m[i][i] = 1.0;
  return m;
}

Mat4 mat_mul(const Mat4& a, const Mat4& b) {
  Mat4 o{};
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double acc = 0.0;
      for (int k = 0; k < 4; k++) acc += a[i][k] * b[k][j];
      o[i][j] = acc;
    }
  return o;
}

Mat4 named_quadric(const std::string& n) {
  Mat4 q{};
  const std::map<std::string, std::array<double, 4>> diag = {
      {"sphere", {1, 1, 1, -1}}, {"cylx", {0, 1, 1, -1}},  {"cyly", {1, 0, 1, -1}}, {"cylz", {1, 1, 0, -1}},
      {"conex", {-1, 1, 1, 0}},  {"coney", {1, -1, 1, 0}}, {"conez", {1, 1, -1, 0}}};
  auto it = diag.find(n);
  if (it != diag.end()) {
    for (int i = 0; i < 4; i++) q[i][i] = it->second[i];
    return q;
  }
  const int a = n == "planex" ? 0 : n == "planey" ? 1 : n == "planez" ? 2 : -1;
  if (a < 0) throw OperatorError("register_quadrics: unknown quadric '" + n + "'");
  q[a][3] = q[3][a] = 0.5;
  return q;
}

Mat4 transform_inverse(const std::string& op, const msp::Node& v) {
  Mat4 m = mat_identity();
  if (op == "scale") {
    const std::vector<double> s = vec3(v);
    for (int d = 0; d < 3; d++) m[d][d] = 1.0 / s[d];
  } else if (op == "translate") {
    const std::vector<double> t = vec3(v);
    for (int d = 0; d < 3; d++) m[d][3] = -t[d];
  } else if (op == "xrot" || op == "yrot" || op == "zrot") {
    const int axis = op[0] - 'x';
    const double c = std::cos(v.number()), s = -std::sin(v.number());
    const int i = axis == 0 ? 1 : axis == 1 ? 2 : 0, j = axis == 0 ? 2 : axis == 1 ? 0 : 1;
    m[i][i] = c;
    m[i][j] = -s;
    m[j][i] = s;
    m[j][j] = c;
  } else {
    throw OperatorError("register_quadrics: unknown transform '" + op + "'");
  }
  return m;
}

struct RegisterQuadric : OperatorNode {
  long id;
  hlbm_obstacle o{};
  explicit RegisterQuadric(const msp::Node& a) {
    if (!a.find("id") || !a.find("quadrics")) throw OperatorError("register_quadrics: the id and quadrics slots are required");
    id = a.at("id").integer();
    const msp::Node& qn = a.at("quadrics");
    Mat4 q{};
    if (qn.is_scalar()) {
      q = named_quadric(qn.text);
    } else {
      if (!qn.is_seq() || qn.items.size() != 4) throw OperatorError("register_quadrics: quadrics must be a name or a 4x4 matrix");
      for (int i = 0; i < 4; i++) {
        const std::vector<double> row = qn.items[i].numbers();
        if (row.size() != 4) throw OperatorError("register_quadrics: quadrics must be a name or a 4x4 matrix");
        for (int j = 0; j < 4; j++) q[i][j] = row[j];
      }
    }
    if (const msp::Node* t = a.find("transform")) {
      if (!t->is_seq()) throw OperatorError("register_quadrics: transform must be a list");
      Mat4 minv = mat_identity();
      for (const auto& step : t->items) {
        if (!step.is_map() || step.fields.size() != 1)
          throw OperatorError("register_quadrics: each transform is {scale|translate|xrot|yrot|zrot: value}");
        minv = mat_mul(minv, transform_inverse(step.fields[0].first, step.fields[0].second));
      }
      Mat4 mt{};
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) mt[i][j] = minv[j][i];
      q = mat_mul(mat_mul(mt, q), minv);
    }
    o.kind = 2;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) o.q[4 * i + j] = q[i][j];
  }
  void execute(Slots& s) override { s.obstacles[id] = o; }
};

// ---- display_obstacles (src/obstacle/display_obstacles.cpp:46-53,
//      obstacle.hpp:68-71,101-104,139-147) ----------------------------------------
struct DisplayObstacles : OperatorNode {
  void execute(Slots& s) override {
    if (s.quiet) return;
    for (const auto& kv : s.obstacles) {
      const hlbm_obstacle& o = kv.second;
      if (o.kind == 0) {
        std::printf("Ball center: (%g, %g, %g), radius: %g\n", o.p0[0], o.p0[1], o.p0[2], o.radius);
      } else if (o.kind == 1) {
        std::printf("Wall bounds: [(%g, %g, %g), (%g, %g, %g)]\n", o.p0[0], o.p0[1], o.p0[2], o.p1[0], o.p1[1],
                    o.p1[2]);
      } else {
        std::printf("Quadric matrix: \n");
        for (int i = 0; i < 4; i++)
          std::printf("%g %g %g %g \n", o.q[4 * i], o.q[4 * i + 1], o.q[4 * i + 2], o.q[4 * i + 3]);
      }
    }
    std::fflush(stdout);
  }
};

struct UpdateObstacles : OperatorNode {
  void execute(Slots& s) override {
    std::vector<hlbm_obstacle> v;
    for (const auto& kv : s.obstacles) v.push_back(kv.second);  // id order
    if (!v.empty()) check(hlbm_add_obstacles(s.dom, v.data(), (int32_t)v.size()), "update_obstacles");
  }
};

// ---- macro_variables (src/collision/macro_variables.cu:71-90) --------------
// Inside the time loop the fused step computes and stores rho/u itself.
struct MacroVariables : OperatorNode {
  void execute(Slots& s) override {
    if (s.in_loop) return;  // the step writes rho/u of the state it collides
    check(hlbm_macro(s.dom, nullptr, nullptr, HLBM_LAYOUT_NODES), "macro_variables");
  }
};

// ---- collision (src/collision/bgk.cu:77-94, mrt.cu:71-89) ------------------
struct Collision : OperatorNode {
  int32_t kind;
  explicit Collision(int32_t k) : kind(k) {}
  void execute(Slots& s) override {
    if (s.collision == kind) return;
    s.flush();
    s.collision = kind;
    check(hlbm_set_collision(s.dom, kind, kind == HLBM_MRT && !s.mrt_rates.empty() ? s.mrt_rates.data() : nullptr),
          name.c_str());
  }
};

// ---- boundary operators: each records its entry of the step plan ------------
struct BoundaryOp : OperatorNode {
  int32_t kind;
  std::vector<std::string> regs;  // planes (lid/neumann/pressure/outflow)
  std::vector<double> U = {0.0, 0.0, 0.0};
  bool lbm_units = false;  // U / celerity (lid_driven_cavity.cu:75, neuman.cu:74); cavity.cu:70 keeps U raw
  void execute(Slots& s) override {
    hlbm_bc b{};
    b.kind = kind;
    b.stage = s.stage;
    for (int d = 0; d < 3; d++) b.U[d] = lbm_units ? U[d] / s.params.celerity : U[d];
    if (regs.empty()) {
      s.recording.push_back(b);
      return;
    }
    for (const auto& r : regs) {
      b.plane = s.plane_of(r);
      s.recording.push_back(b);
    }
  }
};

std::unique_ptr<OperatorNode> make_bc(int32_t kind, const msp::Node& a, const std::string& name) {
  auto op = std::make_unique<BoundaryOp>();
  op->kind = kind;
  switch (kind) {
    case HLBM_BC_CAVITY_Z_0:
    case HLBM_BC_CAVITY_Z_L:
      op->U = vec3(a.at("U"));
      break;
    case HLBM_BC_LID_DRIVEN_CAVITY:
    case HLBM_BC_NEUMANN:
      if (!a.find("U") || !a.find("regions")) throw OperatorError(name + ": U and regions are required");
      op->U = vec3(a.at("U"));
      op->regs = regions(a, name);
      op->lbm_units = true;
      break;
    case HLBM_BC_PRESSURE_INLET:
      if (!a.find("rho")) throw OperatorError(name + ": rho is required");
      op->U = {a.at("rho").number(), 0.0, 0.0};
      op->regs = regions(a, name);
      break;
    case HLBM_BC_OUTFLOW:
      op->regs = regions(a, name);
      break;
    default:
      break;
  }
  return op;
}

// streaming (src/collision/streaming.cu:68-118).  Its `asynchrone: true` form
// runs nothing in the reference (the overlapped path is commented out,
// streaming.cu:84-106), i.e. a step without propagation; that is refused here
// rather than silently streaming.
struct Streaming : OperatorNode {
  explicit Streaming(const msp::Node& a) {
    if (const msp::Node* as = a.find("asynchrone"))
      if (as->boolean())
        throw OperatorError("streaming: asynchrone: true skips propagation in the reference "
                            "(streaming.cu:84-106); not supported");
  }
  void execute(Slots& s) override { s.streamed = true; }
};

struct Nop : OperatorNode {
  void execute(Slots&) override {}
};

// ---- simulation_state (src/io/simulation_state.cu:73-98) --------------------
struct SimulationState : OperatorNode {
  void execute(Slots& s) override {
    s.flush();
    double out[7];
    check(hlbm_reduce_stats(s.dom, out), "simulation_state");
    s.stats.sum_density = out[0];
    s.stats.min_velocity_norm = out[1];
    s.stats.max_velocity_norm = out[2];
  }
};

// ---- log (src/io/log.cpp:73-106) --------------------------------------------
struct Log : OperatorNode {
  void execute(Slots& s) override {
    s.flush();
    const long long size_xyz = (long long)s.nodes[0] * s.nodes[1] * s.nodes[2];
    const auto now = std::chrono::steady_clock::now();
    double mlups = 0.0;
    if (s.previous_step != 0 && s.timestep > s.previous_step) {
      const double T = std::chrono::duration<double>(now - s.previous_time).count();
      mlups = double(size_xyz * (s.timestep - s.previous_step)) / (1e6 * T);
    }
    s.previous_time = now;
    s.previous_step = s.timestep;
    if (s.quiet) return;
    const double c = s.params.celerity;
    if (s.print_log_header)
      std::printf("     Step     Time          Mesh Size   Sum(density)   min(||V||)   max(||V||)     MLUPS\n");
    s.print_log_header = false;
    std::printf("%9ld % .6e %13lld       %.2e     %.2e     %.2e   %.2e\n", s.timestep, s.physical_time, size_xyz,
                s.stats.sum_density, s.stats.min_velocity_norm * c, s.stats.max_velocity_norm * c, mlups);
    std::fflush(stdout);
  }
};

// ---- check_macro_quantities (src/check/check_marco_quantities.cpp:71-114) ---
struct CheckMacroQuantities : OperatorNode {
  double vmin = 0.0;
  bool has_vmax = false, has_density = false, has_tol = false;
  double vmax = 0.0, density = 0.0, tol = 0.0;
  explicit CheckMacroQuantities(const msp::Node& a) {
    vmin = num_or(a, "vmin", 0.0);
    if (const msp::Node* n = a.find("vmax")) has_vmax = true, vmax = n->number();
    if (const msp::Node* n = a.find("density")) has_density = true, density = n->number();
    if (const msp::Node* n = a.find("tol")) has_tol = true, tol = n->number();
  }
  void execute(Slots& s) override {
    const SimulationStatistics& st = s.stats;
    bool error = false;
    std::string msg = "Regression test checker [macro quantities]\n";
    if (st.min_velocity_norm < vmin - 1e-15) {
      msg += "min vel (simulation): " + std::to_string(st.min_velocity_norm) +
             " is below (user-defined check): " + std::to_string(vmin) + "\n";
      error = true;
    }
    if (has_vmax && st.max_velocity_norm > vmax + 1e-15) {
      msg += "max vel (simulation): " + std::to_string(st.max_velocity_norm) +
             " is upper (user-defined check): " + std::to_string(vmax) + "\n";
      error = true;
    }
    if (has_density) {
      if (!has_tol) throw OperatorError("Please add a tolerance if you use the density INPUT slot.");
      if (std::abs(st.sum_density - density) > tol) {
        msg += "Sum of densities (simulation): " + std::to_string(st.sum_density) +
               " is not closed to (user-defined check): " + std::to_string(density) + " +/- " +
               std::to_string(tol) + "\n";
        error = true;
      }
    }
    if (error) throw OperatorError(msg);
    std::printf("Test Pass.\n");
    std::fflush(stdout);
  }
};

struct Message : OperatorNode {
  std::string text;
  explicit Message(const msp::Node& a) : text(a.is_scalar() ? a.text : "") {}
  void execute(Slots& s) override {
    if (!s.quiet) std::printf("%s\n", text.c_str());
  }
};

template <class T>
OperatorFactoryFn plain() {
  return [](const msp::Node&) { return std::unique_ptr<OperatorNode>(new T()); };
}
template <class T>
OperatorFactoryFn with_args() {
  return [](const msp::Node& a) { return std::unique_ptr<OperatorNode>(new T(a)); };
}
OperatorFactoryFn bc(int32_t kind, const char* name) {
  return [kind, name](const msp::Node& a) { return make_bc(kind, a, name); };
}

// the registrations (ONIKA_AUTORUN_INIT blocks of the reference's operators)
struct Registrar {
  Registrar() {
    auto* f = OperatorNodeFactory::instance();
    f->register_factory("domain", with_args<Domain>());
    f->register_factory("lbm_parameters", with_args<LbmParametersOp>());
    f->register_factory("define_grid_3dq19", plain<DefineGrid>());
    f->register_factory("init_obst", plain<InitObst>());
    f->register_factory("set_distribution", with_args<SetDistribution>());
    f->register_factory("init_double_couette", with_args<InitDoubleCouette>());
    f->register_factory("init_obstacles", plain<InitObstacles>());
    f->register_factory("register_solid_ball", with_args<RegisterBall>());
    f->register_factory("register_solid_wall", with_args<RegisterWall>());
    f->register_factory("register_quadrics", with_args<RegisterQuadric>());
    f->register_factory("update_obstacles", plain<UpdateObstacles>());
    f->register_factory("display_obstacles", plain<DisplayObstacles>());
    f->register_factory("print_domain", plain<PrintDomain>());
    f->register_factory("macro_variables", plain<MacroVariables>());
    f->register_factory("bgk", [](const msp::Node&) { return std::unique_ptr<OperatorNode>(new Collision(HLBM_BGK)); });
    f->register_factory("mrt", [](const msp::Node&) { return std::unique_ptr<OperatorNode>(new Collision(HLBM_MRT)); });
    f->register_factory("pre_bounce_back", bc(HLBM_BC_PRE_BOUNCE_BACK, "pre_bounce_back"));
    f->register_factory("post_bounce_back", bc(HLBM_BC_POST_BOUNCE_BACK, "post_bounce_back"));
    f->register_factory("wall_bounce_back", bc(HLBM_BC_WALL_BOUNCE_BACK, "wall_bounce_back"));
    f->register_factory("cavity_z_0", bc(HLBM_BC_CAVITY_Z_0, "cavity_z_0"));
    f->register_factory("cavity_z_l", bc(HLBM_BC_CAVITY_Z_L, "cavity_z_l"));
    f->register_factory("lid_driven_cavity", bc(HLBM_BC_LID_DRIVEN_CAVITY, "lid_driven_cavity"));
    f->register_factory("neumann", bc(HLBM_BC_NEUMANN, "neumann"));
    f->register_factory("pressure_inlet", bc(HLBM_BC_PRESSURE_INLET, "pressure_inlet"));
    f->register_factory("outflow", bc(HLBM_BC_OUTFLOW, "outflow"));
    f->register_factory("streaming", with_args<Streaming>());
    // build_traversal (src/grid/build_traversal.cpp:68): the index lists it builds
    // are the kernels' own coordinate tests here, so it has nothing to do
    f->register_factory("build_traversal", plain<Nop>());
    f->register_factory("update_ghost", plain<Nop>());  // periodic wrap lives in the step kernel
    f->register_factory("nop", plain<Nop>());
    f->register_factory("simulation_state", plain<SimulationState>());
    f->register_factory("log", plain<Log>());
    f->register_factory("check_macro_quantities", with_args<CheckMacroQuantities>());
    f->register_factory("message", with_args<Message>());
  }
};
const Registrar registrar;

}  // namespace

// ---- Slots -------------------------------------------------------------------
Slots::~Slots() {
  if (dom) hlbm_domain_destroy(dom);
}

void Slots::flush() {
  if (pending <= 0) return;
  check(hlbm_step(dom, (int32_t)pending, HLBM_STEP_WRITE_MACRO, nullptr), "streaming");
  check(hlbm_synchronize(dom), "streaming");
  pending = 0;
}

int Slots::plane_of(const std::string& region) const {
  for (int p = 0; p < 6; p++)
    if (region == PLANES[p]) return p;
  throw OperatorError("Traversal name not allowed: " + region);  // neuman.cu:81-84
}

// ---- factory -------------------------------------------------------------------
OperatorNodeFactory* OperatorNodeFactory::instance() {
  static OperatorNodeFactory f;
  return &f;
}

void OperatorNodeFactory::register_factory(const std::string& name, OperatorFactoryFn fn) { fns_[name] = std::move(fn); }

bool OperatorNodeFactory::has(const std::string& name) const { return fns_.count(name) != 0; }

std::unique_ptr<OperatorNode> OperatorNodeFactory::make(const std::string& name, const msp::Node& args) const {
  auto it = fns_.find(name);
  if (it == fns_.end()) throw OperatorError("unsupported operator '" + name + "' (not on the hot path)");
  auto op = it->second(args);
  op->name = name;
  return op;
}

std::vector<std::string> OperatorNodeFactory::names() const {
  std::vector<std::string> v;
  for (const auto& kv : fns_) v.push_back(kv.first);
  return v;
}

std::vector<std::unique_ptr<OperatorNode>> make_list(const msp::Node* list) {
  std::vector<std::unique_ptr<OperatorNode>> out;
  if (!list || list->is_null()) return out;
  static const msp::Node empty{};
  auto one = [&](const msp::Node& item) {
    if (item.is_scalar()) {
      out.push_back(OperatorNodeFactory::instance()->make(item.text, empty));
    } else if (item.is_map() && item.fields.size() == 1) {
      out.push_back(OperatorNodeFactory::instance()->make(item.fields[0].first, item.fields[0].second));
    } else {
      throw OperatorError("an operator list item must be a name or {name: {slots}}");
    }
  };
  if (list->is_seq()) {
    for (const auto& item : list->items) one(item);
  } else {
    one(*list);  // a single operator given without a list (e.g. `collision: mrt`)
  }
  return out;
}

// ---- the simulation graph -------------------------------------------------------
// Top-level blocks the operator layer reads; everything else a deck could name
// (a redefinition of the default graph's batches, Onika's rebind / condition /
// loop control) is refused rather than silently ignored.
static void check_top_level(const msp::Node& deck) {
  static const char* known[] = {"do_domain", "set_lbm_parameters", "collision", "pre_stream_bcs", "post_stream_bcs",
                                "boundary_conditions", "set_distributions", "set_obstacles", "start_iteration",
                                "checker_op", "checker", "analysis", "global", "mrt_rates", "includes",
                                // a block named after the set_distribution operator (the reference's
                                // regression_lbm_diff_pressure.msp pin shows it leaves f unchanged)
                                "set_distribution",
                                // Onika's logging / profiling switches: no effect on the results
                                "configuration"};
  static const char* graph[] = {"compute_loop", "compute_loop_stop", "initialize", "simulation", "end_iteration",
                                "setup_obstacles", "grid_flavor", "default_grid", "trigger_dump_paraview",
                                "trigger_print_log", "trigger_analysis", "dump_paraview_if_triggered",
                                "print_log_if_triggered", "print_analysis_if_triggered",
                                "write_paraview_first_timestep", "streaming"};
  if (!deck.is_map()) return;
  for (const auto& kv : deck.fields) {
    const std::string& k = kv.first;
    for (const char* g : graph)
      if (k == g)
        throw OperatorError("the deck redefines '" + k + "' of the default graph (data/config/*.msp); the operator "
                            "layer runs that graph's fixed sequence and does not execute redefinitions");
    if (kv.second.is_map())
      for (const char* c : {"rebind", "condition", "loop", "body"})
        if (kv.second.find(c))
          throw OperatorError("block '" + k + "' uses Onika graph control ('" + c +
                              "'), which the operator layer does not execute");
    bool ok = false;
    for (const char* n : known) ok = ok || k == n;
    if (!ok) throw OperatorError("unsupported top-level block '" + k + "'");
  }
}

Simulation::Simulation(const msp::Node& deck) {
  check_top_level(deck);
  do_domain = make_list(deck.find("do_domain"));
  set_lbm_parameters = make_list(deck.find("set_lbm_parameters"));
  const msp::Node* coll = deck.find("collision");
  if (coll) {
    collision = make_list(coll);
  } else {
    collision.push_back(OperatorNodeFactory::instance()->make("bgk", msp::Node{}));  // default model
  }
  pre_stream_bcs = make_list(deck.find("pre_stream_bcs"));
  post_stream_bcs = make_list(deck.find("post_stream_bcs"));
  boundary_conditions = make_list(deck.find("boundary_conditions"));
  // set_distributions: items `set_distribution: {value, bounds}`.  A top-level
  // user block named `set_distribution` does not change the state (the
  // reference's regression_lbm_diff_pressure.msp pin, Sum(rho) = 29124, shows
  // f stays at rho = 1), so it is not read.
  set_distributions = make_list(deck.find("set_distributions"));
  set_obstacles = make_list(deck.find("set_obstacles"));
  start_iteration = make_list(deck.find("start_iteration"));
  checker_op = make_list(deck.find("checker_op"));
  analysis = make_list(deck.find("analysis"));
  if (const msp::Node* c = deck.find("checker")) {
    checker_list = make_list(c);
    checker_override = true;
  }
  paraview = OperatorNodeFactory::instance()->make("write_paraview", msp::Node{});
  if (const msp::Node* g = deck.find("global")) {
    slots.end_iteration = (long)num_or(*g, "simulation_end_iteration", 100);
    slots.print_log_freq = (long)num_or(*g, "simulation_print_log_freq", 1);
    slots.timestep = (long)num_or(*g, "timestep", 0);
    slots.physical_time = num_or(*g, "physical_time", 0.0);
    slots.paraview_freq = (long)num_or(*g, "simulation_paraview_freq", -1);
    slots.analysis_freq = (long)num_or(*g, "simulation_analysis_freq", -1);
    if (const msp::Node* n = g->find("display_paraview_first_timestep")) slots.display_paraview_first_timestep = n->boolean();
    if (const msp::Node* n = g->find("output_directory")) slots.output_directory = n->text;
  }
  // extension (mrt.hpp hard-codes its rates): mrt_rates: [s1 s2 s4 s9 s10 s13 s16 weps wepsj wxx]
  if (const msp::Node* r = deck.find("mrt_rates")) {
    slots.mrt_rates = r->numbers();
    if (slots.mrt_rates.size() != 10) throw OperatorError("mrt_rates needs 10 values");
  }
  for (const auto& op : collision)
    if (op->name == "mrt") slots.collision = HLBM_MRT;
}

static void run(std::vector<std::unique_ptr<OperatorNode>>& ops, Slots& s) {
  for (auto& op : ops) op->execute(s);
}

void Simulation::initialize() {
  Slots& s = slots;
  run(do_domain, s);
  if (set_lbm_parameters.empty()) set_lbm_parameters.push_back(OperatorNodeFactory::instance()->make("lbm_parameters", msp::Node{}));
  run(set_lbm_parameters, s);
  PrintDomain().execute(s);
  DefineGrid().execute(s);
  if (s.autotune) {
    hlbm_tune_info info{};
    check(hlbm_autotune(s.dom, 3, nullptr, &info), "autotune");
  }
  InitObst().execute(s);
  SetDistribution(msp::Node{}).execute(s);  // default f_iLB = w_i (rho = 1)
  run(set_distributions, s);
  InitObstacles().execute(s);
  run(set_obstacles, s);
  UpdateObstacles().execute(s);
  MacroVariables().execute(s);
  SimulationState().execute(s);
  if (s.display_paraview_first_timestep) {  // write_paraview_first_timestep (config_io.msp:20-27)
    MacroVariables().execute(s);
    paraview->execute(s);
  }
  s.previous_time = std::chrono::steady_clock::now();
  if (s.print_log_freq > 0) Log().execute(s);
  run(analysis, s);
  s.timestep += 1;  // next_time_step
  s.physical_time += s.params.dtLB;
}

void Simulation::compute_loop() {
  Slots& s = slots;
  s.in_loop = true;
  while (s.timestep <= s.end_iteration) {
    s.stage = HLBM_STAGE_START;
    run(start_iteration, s);
    MacroVariables().execute(s);
    run(collision, s);
    s.stage = HLBM_STAGE_PRE_STREAM;
    run(pre_stream_bcs, s);
    Streaming(msp::Node{}).execute(s);
    s.stage = HLBM_STAGE_POST_STREAM;
    run(post_stream_bcs, s);
    s.stage = HLBM_STAGE_BOUNDARY;
    run(boundary_conditions, s);
    // end_iteration: commit this step; a changed plan first runs the pending steps
    if (!s.plan_set || s.recording.size() != s.plan.size() ||
        !std::equal(s.recording.begin(), s.recording.end(), s.plan.begin(), [](const hlbm_bc& a, const hlbm_bc& b) {
          return a.kind == b.kind && a.stage == b.stage && a.plane == b.plane && a.U[0] == b.U[0] &&
                 a.U[1] == b.U[1] && a.U[2] == b.U[2] && std::equal(a.bounds, a.bounds + 6, b.bounds);
        })) {
      s.flush();
      check(hlbm_set_bcs(s.dom, s.recording.data(), (int32_t)s.recording.size()), "boundary conditions");
      s.plan = s.recording;
      s.plan_set = true;
    }
    s.recording.clear();
    s.streamed = false;
    s.pending++;
    // end_iteration (config_compute_loop.msp:8-16): dump_paraview_if_triggered
    // (update_ghost + macro_variables + write_paraview), print_log_if_triggered
    // (simulation_state + log), print_analysis_if_triggered
    if (s.paraview_freq > 0 && s.timestep % s.paraview_freq == 0) {
      // the dump's macro_variables also changes what the log below reports, so it
      // runs with --no-output too
      s.flush();
      check(hlbm_macro(s.dom, nullptr, nullptr, HLBM_LAYOUT_NODES), "macro_variables");
      paraview->execute(s);
    }
    if (s.print_log_freq > 0 && s.timestep % s.print_log_freq == 0) {
      SimulationState().execute(s);
      Log().execute(s);
    }
    if (s.analysis_freq > 0 && s.timestep % s.analysis_freq == 0) run(analysis, s);
    s.timestep += 1;  // next_time_step
    s.physical_time += s.params.dtLB;
  }
  s.flush();
  s.in_loop = false;
}

void Simulation::final_output() { paraview->execute(slots); }

void Simulation::checker() {
  if (checker_override) {
    run(checker_list, slots);
    return;
  }
  SimulationState().execute(slots);
  run(checker_op, slots);
}

}  // namespace hlbm_host
