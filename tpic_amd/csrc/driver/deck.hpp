// Input deck and run admission of the tpic_mi_run driver.
//
// The deck grammar is the reference's (src/config/simulation_parameters.cpp:
// 55-200, the grammar note of include/tpic/config/simulation_parameters.hpp:
// 209-216): one `key = value` per line, `#` starts a comment, surrounding
// blanks are dropped, the last assignment of a key wins, a value is read in
// full (a boolean is 0, 1, true or false; a list holds whitespace-separated
// values, at least as many as required, extras ignored), and keys nobody reads
// are reported and ignored. Defaults are those of config::SimulationParameters
// (include/tpic/config/simulation_parameters.hpp:26-199).
//
// Admission mirrors the reference's host checks (src/config/physics.cpp:26-119,
// src/config/domain.cpp:54-83, src/simulation/run_plan.cpp:100-135): a broken
// deck is an Invalid error (exit 1); a valid deck that asks for something this
// build does not run lands in `unsupported` (exit 2).
#pragma once

#include <cstdint>
#include <istream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace tpic_mi::deck {

// "TPIC <topic>: <message>" (include/tpic/config/detail/errors.hpp:13-28).
struct Invalid : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};

struct Species {
  double rho_init = 0.0;
  double thermal[3] = {0.0, 0.0, 0.0};  // uth, vth, wth
  double drift[3] = {0.0, 0.0, 0.0};    // u0, v0, w0
  double q_over_m = 0.0;
  int ppc[3] = {0, 0, 0};               // npcelx, npcely, npcelz
  bool center_initial_velocity = false;
};

// The keys of config::SimulationParameters that this path consumes; the
// other recipe keys are parsed (so that their values are checked) and kept
// only as "present".
struct Parameters {
  double dt = 0.0;
  int ncycles = 0;
  double th = 1.0;
  double c = 1.0;
  double smooth = 1.0;
  int smooth_niter = 6;
  std::string smoothing = "none";       // ECSimSmoothing: none | energy_conserving
  std::string smoothing_operator = "face7";
  int ns = 0;
  int ns_test = 0;
  double b0[3] = {0.0, 0.0, 0.0};
  double lengths[3] = {10.0, 10.0, 10.0};
  int cells[3] = {64, 64, 64};
  int processes[3] = {1, 1, 1};         // XLEN, YLEN, ZLEN
  bool periodic[3] = {true, true, true};
  bool particle_periodic[3] = {true, true, true};
  std::string case_name = "Uniform";
  bool quiet_start = false;
  std::string sim_name;
  double gmres_tol = 1e-3;
  int field_cycle = 100, particles_cycle = 0, test_particles_cycle = 0, restart_cycle = 0;
  bool call_finalize = true;
  int diagnostics_cycle = 100, energy_cycle = 1, gauss_cycle = 0;
  int em_face[6] = {0, 0, 0, 0, 0, 0};        // Xright Xleft Yright Yleft Zright Zleft
  int particle_face[6] = {1, 1, 1, 1, 1, 1};
  std::vector<Species> species;
  std::vector<Species> test_species;          // nsTestPart entries (refused at admission)
  std::vector<std::string> unknown_keys;      // "key (line n)", in deck order
  // Every key the parser reads with its resolved value (defaults included),
  // in read order; lists hold one entry per real and test species.
  std::vector<std::pair<std::string, std::string>> resolved;
};

Parameters parse(std::istream& in);
Parameters load(const std::string& path);

// Command-line choices (tpic::simulation::InputOptions, src/app/main.cpp:60-110).
struct Options {
  std::string order = "partial";     // full | partial | unsorted
  std::string mass_matrix = "scalar";
  std::string mass_apply = "full";
  std::string storage_precision = "double";   // --precision sets both levels
  std::string compute_precision = "double";
  int mass_operands = 0;             // TPIC_MI_MASS_*: --mass-precision fp64 | fp32 | bf16x3
  bool restart = false;
  bool overwrite_output = false;
  int rank = 0;
};

// One species' load and its plan (config::SpeciesParticlePlan,
// src/config/particle_plan.cpp:127-150).
struct SpeciesPlan {
  uint64_t per_cell = 0;
  uint64_t count = 0;
  uint64_t first_id = 0;
};

// The built initial cases (src/initialization/cases/*.cpp): what the load
// and the field images need.
struct CaseRecipe {
  // TwoStream: density seed 1 + amplitude cos(k x + phase) of the seeded
  // (electron) species, and E_x = e_amplitude sin(k x + phase) at nodes.
  // Weibel: B_z += b_amplitude cos(k x) at cells, E_y = e_amplitude sin(k x).
  double k = 0.0, phase = 0.0, density_amplitude = 0.0;
  double e_amplitude = 0.0, b_amplitude = 0.0;
  bool e_seed = false, b_seed = false;
  std::vector<bool> density_seeded;  // per species

  // Harris sheets (HarrisSheetCase, src/initialization/cases/harris_sheet.cpp:
  // 183-257): the seed of the sheet field, and per species the profile.
  bool harris = false, harris_double = false;
  double harris_b0x = 0.0, delta = 0.0;
  struct Seed {
    double mode_amplitude = 0, mode_kx = 0, mode_ky = 0, amplitude_x = 0, amplitude_y = 0;
    double width_x = 0, width_y = 0, period_x = 0, period_y = 0, slope_period_x = 0, slope_period_y = 0;
  } seed;
  struct Curl {
    double center_x = 0, center_y = 0, delta = 0, sheet = 0, mode = 0, mode_kx = 0, mode_ky = 0;
    double hump = 0, hump_kx = 0, hump_ky = 0, hump_inverse_wx2 = 0, hump_inverse_wy2 = 0;
  } curl;
  struct HarrisSpecies {
    bool density = false, varying_thermal = false;
    double ampere_drift = 0.0;
  };
  std::vector<HarrisSpecies> harris_species;

  // The seeded k-spectrum of B (Uniform case; MagneticSeedSpectrum* keys):
  // relative amplitude, modes and seed.
  double spectrum_amplitude = 0.0;
  int spectrum_modes = 0;
  uint64_t spectrum_seed = 7;
};

// The output this driver writes (config::PreparedOutput, src/config/output.cpp:
// 151-256): diagnostics.csv every diagnostics_cycle cycles and at the end, E
// and B snapshots (VTK ImageData) every field_cycle cycles and at the end.
struct OutputPlan {
  std::string save_dir = "data";
  bool diagnostics = false, fields = false, output_e = false, output_b = false, particles = false;
  int diagnostics_cycle = 0, field_cycle = 0, particles_cycle = 0;
  // Restart files (raw, tpic_mi_checkpoint_write) every restart_cycle cycles.
  std::string restart_dir = "data";
  int restart_cycle = 0;
};

struct Prepared {
  Parameters p;
  Options options;
  OutputPlan output;
  double spacing[3] = {0, 0, 0};
  uint64_t cells = 0;
  CaseRecipe recipe;
  std::vector<SpeciesPlan> plans;
  std::vector<std::string> notes, warnings, unsupported;
  // Energy-conserving smoothing is active (physics.cpp:65-71): SmoothNiter
  // passes of ECSimSmoothingOperator with factor Smooth.
  bool smoothing = false;
  bool admitted() const { return unsupported.empty(); }
};

// Throws Invalid on a broken deck or option.
Prepared prepare(const Parameters& p, const Options& options);

// The resolved value of a key the parser read (Invalid if it read none).
const std::string& resolved(const Parameters& p, const std::string& key);

// The initial field images of the case (make_host_field_image,
// include/tpic/initialization/fields/field_image.hpp:45-111, one rank,
// origin 0): E at the nodes x = i dx (i in [0, n], periodic duplicates
// sampled at their own coordinates), B at the cell centres x = (i + 1/2) dx,
// B0 added to B (initial_case.cpp:145-162). Layouts of include/tpic_mi.h.
void field_images(const Prepared& r, std::vector<double>& e_nodes, std::vector<double>& b_cells);

}  // namespace tpic_mi::deck
