// hlbm_run -- runs a hippoLBM .msp deck through the operator layer on the GPU:
// initialize, compute_loop, checker (data/config/main-config.msp), printing
// the log operator's table; exits with EXIT_FAILURE when an operator or the
// regression checker fails, as the reference does.
//
//   hlbm_run deck.msp [--storage f64|f32|f16|bf16] [--pattern ab|aa] [--autotune]
//                     [--steps N] [--quiet] [--no-output] [--dump-deck] [--check-deck]
//
// --dump-deck prints the deck (with its includes merged) as JSON; --check-deck
// builds the operator graph (every operator and slot checked) and exits
// before any device work.
//
// Output operators (write_paraview, plane_velocity_profile,
// plot_line_velocity) write under the deck's output_directory, relative to
// the working directory, as the reference does; --no-output turns them off.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "operators.hpp"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s deck.msp [--storage f64|f32|f16|bf16] [--pattern ab|aa] [--autotune] "
                         "[--steps N] [--quiet] [--no-output] [--dump-deck] [--check-deck]\n", argv[0]);
    return 2;
  }
  std::string path;
  std::string storage = "f64", pattern = "ab";
  bool autotune = false, quiet = false, dump = false, io = true, check_only = false;
  long steps = -1;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    if (a == "--storage" && i + 1 < argc) storage = argv[++i];
    else if (a == "--pattern" && i + 1 < argc) pattern = argv[++i];
    else if (a == "--steps" && i + 1 < argc) steps = std::atol(argv[++i]);
    else if (a == "--autotune") autotune = true;
    else if (a == "--quiet") quiet = true;
    else if (a == "--no-output") io = false;
    else if (a == "--dump-deck") dump = true;
    else if (a == "--check-deck") check_only = true;
    else if (path.empty()) path = a;
    else {
      std::fprintf(stderr, "unknown argument %s\n", a.c_str());
      return 2;
    }
  }
  try {
    const msp::Node deck = msp::load_deck(path);
    if (dump) {
      std::printf("%s\n", msp::to_json(deck).c_str());
      return 0;
    }
    hlbm_host::Simulation sim(deck);
    if (check_only) {
      std::printf("deck ok\n");
      return 0;
    }
    hlbm_host::Slots& s = sim.slots;
    if (storage == "f64") s.storage = HLBM_F64;
    else if (storage == "f32") s.storage = HLBM_F32;
    else if (storage == "f16") s.storage = HLBM_F16;
    else if (storage == "bf16") s.storage = HLBM_BF16;
    else throw hlbm_host::OperatorError("unknown storage " + storage);
    if (pattern == "ab") s.pattern = HLBM_PATTERN_AB;
    else if (pattern == "aa") s.pattern = HLBM_PATTERN_AA;
    else throw hlbm_host::OperatorError("unknown pattern " + pattern);
    s.autotune = autotune;
    s.quiet = quiet;
    s.io = io;
    if (steps >= 0) s.end_iteration = steps;
    sim.initialize();
    sim.compute_loop();
    sim.final_output();
    sim.checker();
    if (quiet)
      std::printf("final step %ld sum_density %.17g min_velocity_norm %.17g max_velocity_norm %.17g\n",
                  s.timestep - 1, s.stats.sum_density, s.stats.min_velocity_norm, s.stats.max_velocity_norm);
  } catch (const msp::ParseError& e) {
    std::fprintf(stderr, "%s: %s\n", path.c_str(), e.what());
    return EXIT_FAILURE;
  } catch (const hlbm_host::OperatorError& e) {
    std::printf("%s\n", e.what());
    return EXIT_FAILURE;
  }
  return 0;
}
