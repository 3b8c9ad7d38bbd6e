// operators.hpp -- the host operator layer: hippoLBM's hot-path operators by
// their reference names, each an OperatorNode whose execute() drives the C ABI
// of include/hlbm.h (libhlbm.so).
//
// The reference registers Onika OperatorNodes by name
// (ONIKA_AUTORUN_INIT(...){ OperatorNodeFactory::instance()->register_factory("bgk", ...); },
// src/collision/bgk.cu:100-102) and wires them through named slots.  Onika is
// not available here, so this layer restates the part of that machinery the
// hot path needs: a factory keyed by operator name, the shared slots (Slots),
// and the initialize / compute_loop / checker sequence of
// data/config/config_initialize.msp:14-32, config_compute_loop.msp:23-38 and
// config_verification.msp.  Operator arguments come from the deck's .msp
// nodes under the reference's slot names (U, regions, bounds, cell_dims, ...).
//
// Boundary operators record their part of the step plan; `streaming` marks
// the step; end_iteration commits it.  Steps with the same plan are executed
// lazily, in one hlbm_step call, when a later operator needs the state (log,
// checker, a change of plan).
#pragma once

#include <chrono>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hlbm.h"
#include "msp.hpp"

namespace hlbm_host {

// LBMParameters (src/grid/include/hippoLBM/grid/lbm_parameters.hpp:41-142)
struct LBMParameters {
  double tau = 0.0, nu = 0.0, dtLB = 0.0, celerity = 1.0, nuth = 1e-4;
  double avg_rho = 1000.0;  // physical units (lbm_parameters.cpp:52)
  double Fext[3] = {0.0, 0.0, 0.0};  // LBM units
};

// SimulationStatistics (src/io/include/hippoLBM/io/simulation_state.hpp)
struct SimulationStatistics {
  double sum_density = 0.0, min_velocity_norm = 0.0, max_velocity_norm = 0.0;
};

// Error raised by an operator (the reference prints and exits with EXIT_FAILURE).
struct OperatorError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// The named slots the hot-path operators share.
struct Slots {
  // domain (src/grid/domain_lbm.cpp)
  bool has_domain = false;
  double bounds[2][3] = {{0, 0, 0}, {0, 0, 0}};
  int cell_dims[3] = {0, 0, 0};
  bool periodic[3] = {false, false, false};
  double dx = 1.0;
  int nodes[3] = {0, 0, 0};
  // lbm_parameters
  LBMParameters params;
  bool has_params = false;
  // collision model
  int32_t collision = HLBM_BGK;
  std::vector<double> mrt_rates;  // empty: the reference's rates
  // engine options (not reference slots)
  int32_t storage = HLBM_F64;
  int32_t pattern = HLBM_PATTERN_AB;
  bool autotune = false;
  bool io = true;  // false: write_paraview / profile / line operators write nothing
  // output (global block)
  std::string output_directory = "hippoLBMOutputDir";
  long paraview_freq = -1, analysis_freq = -1;
  bool display_paraview_first_timestep = true;
  double set_dist_value = 1.0;  // value of the last set_distribution without bounds
  // grid / fields
  hlbm_domain* dom = nullptr;
  std::map<long, hlbm_obstacle> obstacles;  // register_solid_*: by id
  // time loop (global block)
  long timestep = 0;
  double physical_time = 0.0;
  long end_iteration = 100;
  long print_log_freq = 1;
  bool print_log_header = true;
  bool quiet = false;
  bool in_loop = false;
  SimulationStatistics stats;
  // log timing
  long previous_step = 0;
  std::chrono::steady_clock::time_point previous_time;
  // step plan
  int32_t stage = HLBM_STAGE_PRE_STREAM;  // list being executed
  std::vector<hlbm_bc> recording;         // this iteration's boundary operators
  bool streamed = false;                  // `streaming` ran this iteration
  std::vector<hlbm_bc> plan;              // plan of the pending steps
  bool plan_set = false;
  long pending = 0;                       // committed iterations not yet executed

  ~Slots();
  // run the pending steps (the last one writes rho/u, as macro_variables would)
  void flush();
  int plane_of(const std::string& region) const;
};

class OperatorNode {
 public:
  virtual ~OperatorNode() = default;
  virtual void execute(Slots& s) = 0;
  std::string name;
};

using OperatorFactoryFn = std::function<std::unique_ptr<OperatorNode>(const msp::Node& args)>;

class OperatorNodeFactory {
 public:
  static OperatorNodeFactory* instance();
  void register_factory(const std::string& name, OperatorFactoryFn fn);
  bool has(const std::string& name) const;
  std::unique_ptr<OperatorNode> make(const std::string& name, const msp::Node& args) const;
  std::vector<std::string> names() const;

 private:
  std::map<std::string, OperatorFactoryFn> fns_;
};

// An operator list as the deck writes it: items are a name or {name: {args}}.
std::vector<std::unique_ptr<OperatorNode>> make_list(const msp::Node* list);

// The deck, resolved as the reference's configuration resolves it: the user
// blocks (do_domain, set_lbm_parameters, collision, pre_stream_bcs,
// post_stream_bcs, boundary_conditions, set_distributions, set_obstacles,
// checker_op, global) replace the defaults of data/config/*.msp.
struct Simulation {
  Slots slots;
  std::vector<std::unique_ptr<OperatorNode>> do_domain, set_lbm_parameters, collision, pre_stream_bcs,
      post_stream_bcs, boundary_conditions, set_distributions, set_obstacles, checker_op, analysis, checker_list,
      start_iteration;
  bool checker_override = false;  // the deck redefines `checker`
  std::unique_ptr<OperatorNode> paraview;  // write_paraview with the default slots

  explicit Simulation(const msp::Node& deck);
  void initialize();    // config_initialize.msp:14-32
  void compute_loop();  // config_compute_loop.msp:23-38
  void final_output();  // the write_paraview that follows compute_loop (main-config.msp:31)
  void checker();       // config_verification.msp
};

}  // namespace hlbm_host
