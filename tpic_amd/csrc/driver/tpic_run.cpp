// tpic_mi_run: deck-driven ECSIM runs on one MI355X through the C-ABI of
// include/tpic_mi.h.
//
// The command line is the reference's `tpic` (src/app/main.cpp:17-127):
//   tpic_mi_run --check-input|--initialize|--run INPUT [--order full|partial|unsorted]
//               [--mass-matrix scalar|matrix-core] [--mass-apply full|split]
//               [--precision LEVEL] [--storage-precision LEVEL] [--compute-precision LEVEL]
//               [--rank N] [--overwrite-output] [--restart]
// plus this build's --mass-precision fp64|fp32|bf16x3 (operand precision of
// the per-cell mass and current moment products) and --timing (phase times).
// Exit codes: 0 done, 1 invalid input or command or a failed run, 2 a request
// this build does not run.
//
// A run follows Pipeline::run_cycles (include/tpic/simulation/driver/
// pipeline.hpp:437-541): the case's load of every species
// (prepared_initialization.cpp:37-85; the TwoStream density seed) with
// CenterInitialVelocity, the case's E and B images plus B0
// (field_image.hpp:45-111, initial_case.cpp:145-162), startup, then per cycle the
// cycle header, the energy of the state entering the cycle when
// DiagnosticEnergyCycle divides the cycle, one ECSIM cycle, and the solver
// and wall-time lines; then the loop summary and the final state. The report
// lines keep the reference's formats (src/simulation/run_reporter.cpp:10-103).
// RestartOutputCycle writes a restart file before every cycle it divides
// (not the run's first) and of the final state (with CallFinalize or when it
// divides ncycles); --restart continues from the newest one after the
// admission of src/simulation/run_plan.cpp:199-214, with the diagnostics
// file cut back to the restart cycle.
#include <chrono>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <string_view>
#include <vector>

#include "../../../include/tpic_mi.h"
#include "deck.hpp"
#include "deck_start.hpp"
#include "output.hpp"

namespace {

using tpic_mi::deck::Invalid;
using tpic_mi::deck::Options;
using tpic_mi::deck::Prepared;

constexpr std::string_view kUsage =
    "Usage: tpic_mi_run --check-input INPUT [--rank N] [--order full|partial|unsorted]\n"
    "       tpic_mi_run --initialize INPUT [options]\n"
    "       tpic_mi_run --run INPUT [options]\n"
    "       tpic_mi_run --help\n"
    "Options: --order full|partial|unsorted (default partial; both sorted orders run a full stable cell sort),\n"
    "         --mass-matrix scalar|matrix-core (both run the FP64 MFMA deposition),\n"
    "         --mass-apply full|split, --precision LEVEL (double is built),\n"
    "         --storage-precision LEVEL, --compute-precision LEVEL,\n"
    "         --mass-precision fp64|fp32|bf16x3 (operands of the per-cell moment products; default fp64),\n"
    "         --timing (print the phase times of the run), --overwrite-output,\n"
    "         --restart (continue from the newest restart file in RestartDirName).\n"
    "       tpic_mi_run --print-input INPUT prints every deck key with its resolved value.\n"
    "Exit codes: 0 done, 1 invalid input/command or a failed run, 2 unsupported requests.\n"
    "--check-input checks the deck without GPU initialization; --initialize completes startup and exits;\n"
    "--run advances ncycles and reports the energy every DiagnosticEnergyCycle cycles.\n"
    "Built: the Uniform, TwoStream, Weibel and Harris-sheet cases (GEMHarris, GEM, GEMnoPert, DoubleGEM,\n"
    "DoubleGEMnoPert, plasmoidChain), periodic or PEC/PMC field faces, periodic or mirror particle faces,\n"
    "one GPU.\n";

const char* kLevels[] = {"double", "single-position", "single-position-velocity",
                         "single-particles-moments-fields-double-mass", "single-particles-moments-fields"};

bool is_level(std::string_view v) {
  for (const char* l : kLevels)
    if (v == l) return true;
  return false;
}

[[noreturn]] void option_error(const std::string& message) { throw Invalid("TPIC options: " + message); }

void print_error(const std::string& message) {
  const char* prefix = message.rfind("TPIC ", 0) == 0 ? "" : "TPIC: ";
  std::fprintf(stderr, "%s%s\n", prefix, message.c_str());
}

const char* mass_operands_name(int m) {
  return m == TPIC_MI_MASS_FP32 ? "fp32" : m == TPIC_MI_MASS_BF16X3 ? "bf16x3" : "fp64";
}

void print_check(const Prepared& r) {
  const auto& p = r.p;
  std::ostringstream out;
  out << std::setprecision(17) << "TPIC host input check: " << (r.admitted() ? "passed" : "unsupported requests")
      << '\n'
      << "Case: " << p.case_name << "; particle recipe: "
      << (p.case_name == "Uniform" || p.case_name == "Weibel" ? "uniform"
          : p.case_name == "TwoStream"                        ? "TwoStream"
          : r.recipe.harris                                   ? (r.recipe.harris_double ? "DoubleGEM" : "GEM/Harris")
                                                              : "not built")
      << "; particle positions: " << (p.quiet_start ? "subcell centres" : "random in subcell") << '\n'
      << "Precision: " << r.options.storage_precision << "/" << r.options.compute_precision
      << "; particle order plan: " << r.options.order << "; species: " << p.ns << '\n'
      << "Mass matrix kernel: matrix-core (FP64 MFMA); moment operands: "
      << mass_operands_name(r.options.mass_operands) << '\n'
      << "Mass apply: " << r.options.mass_apply << '\n'
      << "Rank: " << r.options.rank << " / " << p.processes[0] * p.processes[1] * p.processes[2]
      << "; local cells: " << p.cells[0] << 'x' << p.cells[1] << 'x' << p.cells[2] << '\n';
  uint64_t total = 0;
  for (const auto& plan : r.plans) total += plan.count;
  out << "Particles: local=" << total << ", global=" << total << '\n';
  for (size_t s = 0; s < r.plans.size(); ++s)
    out << "  species " << s << ": count=" << r.plans[s].count << ", IDs=[" << r.plans[s].first_id << ','
        << r.plans[s].first_id + r.plans[s].count << ")\n";
  out << "EM faces [Xright Xleft Yright Yleft Zright Zleft]:";
  const char* walls[3] = {"PEC", "PMC", "open"};
  for (int f = 0; f < 6; ++f) out << ' ' << (p.periodic[f / 2] ? "periodic" : walls[p.em_face[f]]);
  out << "\nParticle faces:";
  for (int f = 0; f < 6; ++f) out << ' ' << (p.particle_periodic[f / 2] ? "periodic" : "mirror");
  out << "\nSmoothing: " << (r.smoothing ? "active" : "inactive")
      << "; GMRES restart=30, max iterations=100, tolerance=" << p.gmres_tol << '\n';
  for (const auto& n : r.notes) out << "Note: " << n << '\n';
  for (const auto& w : r.warnings) out << "Warning: " << w << '\n';
  for (const auto& u : r.unsupported) out << "Unsupported: " << u << '\n';
  out << "Input checking performs no simulation startup or ECSIM timestep.\n";
  std::cout << out.str() << std::flush;
}

struct Stats {
  uint64_t n = 0;
  double sum = 0, sum2 = 0, lo = 0, hi = 0;
  void add(double v) {
    lo = n ? std::min(lo, v) : v;
    hi = n ? std::max(hi, v) : v;
    ++n;
    sum += v;
    sum2 += v * v;
  }
  double mean() const { return n ? sum / double(n) : 0.0; }
  double stddev() const {
    if (n < 2) return 0.0;
    const double m = mean();
    return std::sqrt(std::max(0.0, (sum2 - double(n) * m * m) / double(n - 1)));
  }
};

// The library's failures end the run with its message (exit 1).
struct Failure : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// The restart file a --restart run continues from and its header.
struct Restart {
  std::string path;
  tpic_mi_checkpoint_header header{};
};

// FieldBoundary then ParticleBoundary codes of Xright..Zleft, as a restart
// file records them.
void deck_faces(const tpic_mi::deck::Parameters& p, int32_t faces[12]) {
  for (int k = 0; k < 6; ++k) {
    faces[k] = p.periodic[k / 2] ? -1 : p.em_face[k];
    faces[6 + k] = p.particle_periodic[k / 2] ? -1 : p.particle_face[k];
  }
}

// The newest restart file of the deck's RestartDirName, admitted against the
// deck (admit_restart, src/simulation/run_plan.cpp:199-214): what differs is
// an unsupported request.
Restart find_restart(const Prepared& r, std::vector<std::string>& unsupported) {
  Restart found;
  int cycle = -1;
  found.path = tpic_mi::output::newest_checkpoint(r.output.restart_dir, &cycle);
  if (found.path.empty()) throw Invalid("TPIC restart: no restart file in " + r.output.restart_dir);
  if (tpic_mi_checkpoint_read_header(found.path.c_str(), &found.header))
    throw Invalid(std::string("TPIC restart: ") + tpic_mi_last_error(nullptr));
  const auto& h = found.header;
  const auto& p = r.p;
  auto reject = [&](bool differs, const std::string& what) {
    if (differs) unsupported.push_back("Restart: " + what + " differs from the checkpoint's");
  };
  int32_t faces[12];
  deck_faces(p, faces);
  reject(h.global_cells[0] != p.cells[0] || h.global_cells[1] != p.cells[1] || h.global_cells[2] != p.cells[2],
         "the mesh (nxc, nyc, nzc)");
  reject(h.lengths[0] != r.spacing[0] * p.cells[0] || h.lengths[1] != r.spacing[1] * p.cells[1] ||
             h.lengths[2] != r.spacing[2] * p.cells[2],
         "the domain (Lx, Ly, Lz)");
  reject(std::memcmp(faces, h.faces, sizeof faces) != 0, "a field or particle boundary");
  reject(h.dt != p.dt, "dt");
  reject(h.c != p.c, "c");
  bool species = h.species == p.ns;
  for (int s = 0; species && s < p.ns; ++s) species = h.q_over_m[s] == p.species[size_t(s)].q_over_m;
  reject(!species, "the species (ns, qom)");
  for (int s = 0; species && s < p.ns; ++s)
    reject(h.global_count[s] != r.plans[size_t(s)].count, "the particle count of species " + std::to_string(s));
  if (p.ncycles <= h.cycle)
    unsupported.push_back("Restart: ncycles = " + std::to_string(p.ncycles) + " is not after the checkpoint cycle " +
                          std::to_string(h.cycle));
  return found;
}

class Run {
 public:
  explicit Run(const Prepared& r, const Restart* restart = nullptr) : r_(r), restart_(restart) {
    if (restart) first_cycle_ = restart->header.cycle;
  }
  ~Run() {
    if (ctx_) tpic_mi_destroy(ctx_);
  }

  void startup(bool timing) {
    ctx_ = tpic_mi::deck::start(r_, nullptr, timing, restart_ ? restart_->path.c_str() : nullptr);
    if (restart_) {
      initial_ = restart_->header.initial_energy;
      std::printf("Restart from %s (cycle %d)\n", restart_->path.c_str(), first_cycle_);
    }
  }

  int cycles() {
    const auto& p = r_.p;
    std::printf("\n============================================================\n"
                "  ECSIM Initialization complete\n"
                "  Starting ECSIM Main Simulation Loop\n"
                "============================================================\n");
    using Clock = std::chrono::steady_clock;
    const auto loop_start = Clock::now();
    Clock::time_point steady_start{};
    Stats all, ordinary;
    const auto& out = r_.output;
    const auto precision = tpic_mi::output::precision_metadata(r_.options.mass_operands != TPIC_MI_MASS_FP64);
    if (out.diagnostics || out.fields || out.particles) tpic_mi::output::ensure_directory(out.save_dir);
    if (out.restart_cycle > 0) tpic_mi::output::ensure_directory(out.restart_dir);
    if (out.diagnostics)
      csv_.open(out.save_dir, size_t(p.ns), precision, r_.options.overwrite_output, restart_ ? first_cycle_ : -1);
    for (int cycle = first_cycle_; cycle < p.ncycles; ++cycle) {
      const auto start = Clock::now();
      const double time = double(cycle) * p.dt;
      std::printf("\n ======= ECSIM Cycle %d  |  t = %.6e =======\n", cycle, time);
      const bool report = p.energy_cycle > 0 && cycle % p.energy_cycle == 0;
      const bool row = out.diagnostics && cycle % out.diagnostics_cycle == 0;
      if (report || row || (cycle == first_cycle_ && !restart_)) {
        sample();
        if (cycle == first_cycle_ && !restart_) initial_ = total_;
        if (report) energy();
        if (row) csv_.write(cycle, time, false, scalars_, initial_);
      }
      if (out.fields && cycle % out.field_cycle == 0) snapshot(cycle, time, false, precision);
      if (out.particles && cycle % out.particles_cycle == 0) particle_snapshot(cycle, time, false, precision);
      if (out.restart_cycle > 0 && cycle != first_cycle_ && cycle % out.restart_cycle == 0) checkpoint(cycle);
      tpic_mi_solve_result res{};
      const int rc = tpic_mi_cycle(ctx_, 1, &res);
      if (rc) throw Failure("TPIC cycle " + std::to_string(cycle) + ": " + tpic_mi_last_error(ctx_));
      check(tpic_mi_synchronize(ctx_));
      const double ms = std::chrono::duration<double, std::milli>(Clock::now() - start).count();
      all.add(ms);
      if (cycle > first_cycle_) ordinary.add(ms);
      std::printf("  Maxwell GMRES : converged in %d iterations; residual = %.6e\n"
                  "  Cycle time    : %.3f ms (rank 0 wall time)\n",
                  res.iterations, res.relative_residual, ms);
      std::fflush(stdout);
      if (cycle == first_cycle_) steady_start = Clock::now();
    }
    const auto loop_end = Clock::now();
    const int steps = p.ncycles - first_cycle_;
    const double total_s = std::chrono::duration<double>(loop_end - loop_start).count();
    const double steady_s = steps > 1 ? std::chrono::duration<double>(loop_end - steady_start).count() : 0.0;
    std::printf("\n=========================================\n"
                "ECSIM Main Loop Complete\n"
                "  Total cycles: %d\n"
                "  Total time: %.9f s (maximum rank)\n"
                "  Steady-state cycles: %d\n"
                "  Steady-state time: %.9f s (maximum rank; first cycle excluded)\n",
                steps, total_s, std::max(steps - 1, 0), steady_s);
    if (steps > 0)
      std::printf("  Average per cycle: %.3f ms (loop time / cycles)\n"
                  "  Std. dev. per cycle: %.3f ms (rank 0 samples)\n",
                  total_s * 1000.0 / steps, all.stddev());
    std::printf("  Ordinary after warm-up: %llu cycles; avg=%.3f std=%.3f min=%.3f max=%.3f ms\n",
                static_cast<unsigned long long>(ordinary.n), ordinary.mean(), ordinary.stddev(), ordinary.lo,
                ordinary.hi);
    std::printf("  Loop totals include cycle reporting; final-state diagnostics are excluded.\n"
                "=========================================\n");
    if (timing_enabled()) phases();
    sample();
    if (steps == 0 && !restart_) initial_ = total_;
    const int last = p.ncycles;
    std::printf("\n ======= ECSIM Final State  |  cycle = %d, t = %.6e =======\n", last, double(last) * p.dt);
    energy();
    if (out.diagnostics) csv_.write(last, double(last) * p.dt, true, scalars_, initial_);
    if (out.fields) snapshot(last, double(last) * p.dt, true, precision);
    if (out.particles) particle_snapshot(last, double(last) * p.dt, true, precision);
    if (out.restart_cycle > 0 && last != first_cycle_ && (p.call_finalize || last % out.restart_cycle == 0))
      checkpoint(last);
    unsigned flags = 0;
    check(tpic_mi_read_validation(ctx_, &flags));
    if (flags) throw Failure("TPIC run: device validation flags 0x" + hex(flags));
    return steps;
  }

 private:
  void check(int rc) {
    if (rc) throw Failure(std::string("TPIC: ") + (ctx_ ? tpic_mi_last_error(ctx_) : "library call failed"));
  }

  static std::string hex(unsigned v) {
    std::ostringstream o;
    o << std::hex << v;
    return o.str();
  }

  bool timing_enabled() const { return timing_; }

  // The restart file of the state entering cycle (after the previous cycle);
  // complete files of earlier cycles are then removed.
  void checkpoint(int cycle) {
    const auto& out = r_.output;
    const std::string path = tpic_mi::output::checkpoint_path(out.restart_dir, cycle);
    const std::string partial = path + ".partial";
    tpic_mi_checkpoint_header h{};
    h.cycle = cycle;
    h.initial_energy = initial_;
    deck_faces(r_.p, h.faces);
    check(tpic_mi_checkpoint_write(ctx_, partial.c_str(), &h));
    if (std::rename(partial.c_str(), path.c_str()) != 0)
      throw Failure("TPIC checkpoint: cannot rename " + partial + ": " + std::strerror(errno));
    tpic_mi::output::remove_older_checkpoints(out.restart_dir, cycle);
    last_checkpoint_ = path;
    ++checkpoints_;
  }

  // ScalarDiagnostics sample of the committed state (W_E, W_B, then per
  // species kinetic energy, charge and momentum).
  void sample() {
    scalars_.assign(2 + 5 * size_t(r_.p.ns), 0.0);
    check(tpic_mi_scalars(ctx_, scalars_.data()));
    electric_ = scalars_[0];
    magnetic_ = scalars_[1];
    kinetic_ = 0;
    for (int s = 0; s < r_.p.ns; ++s) kinetic_ += scalars_[2 + 5 * size_t(s)];
    total_ = electric_ + magnetic_ + kinetic_;
    if (!std::isfinite(kinetic_) || !std::isfinite(total_))
      throw Failure("TPIC reporter: nonfinite system energy in console diagnostics");
  }

  // E at the nodes and B at the cells, one .vti file each (FieldOutputTag).
  void snapshot(int cycle, double time, bool final_state, const tpic_mi::output::Precision& precision) {
    const auto& p = r_.p;
    const auto& out = r_.output;
    const size_t nodes = size_t(p.cells[0] + 1) * size_t(p.cells[1] + 1) * size_t(p.cells[2] + 1);
    if (out.output_e) {
      std::vector<double> e(3 * nodes);
      check(tpic_mi_get_field(ctx_, TPIC_MI_FIELD_E, e.data()));
      last_snapshot_ = tpic_mi::output::write_vti(out.save_dir, "E", cycle, time, final_state, precision, p.cells,
                                                  r_.spacing, false, e, r_.options.overwrite_output);
    }
    if (out.output_b) {
      std::vector<double> b(3 * size_t(r_.cells));
      check(tpic_mi_get_field(ctx_, TPIC_MI_FIELD_B, b.data()));
      last_snapshot_ = tpic_mi::output::write_vti(out.save_dir, "B", cycle, time, final_state, precision, p.cells,
                                                  r_.spacing, true, b, r_.options.overwrite_output);
    }
    ++snapshots_;
  }

  // Every species in its stored order, one .vtp piece and its .pvtp manifest.
  void particle_snapshot(int cycle, double time, bool final_state, const tpic_mi::output::Precision& precision) {
    const auto& p = r_.p;
    std::vector<tpic_mi::output::SpeciesParticles> species(size_t(p.ns));
    for (int s = 0; s < p.ns; ++s) {
      auto& h = species[size_t(s)];
      const size_t n = r_.plans[size_t(s)].count;
      for (auto* v : {&h.rx, &h.ry, &h.rz, &h.u, &h.v, &h.w, &h.q}) v->resize(n);
      h.cell.resize(n);
      h.id.resize(n);
      tpic_mi_particle_arrays a{h.rx.data(), h.ry.data(), h.rz.data(), h.u.data(), h.v.data(),
                                h.w.data(), h.q.data(), h.cell.data(), h.id.data()};
      check(tpic_mi_get_particles(ctx_, s, &a));
    }
    last_particles_ = tpic_mi::output::write_particles(r_.output.save_dir, cycle, time, final_state, precision,
                                                       p.cells, r_.spacing, species, r_.options.overwrite_output);
    ++particle_snapshots_;
  }

 public:
  // Output files written (print_completion, startup_report.cpp:395-427).
  void completion(int steps) const {
    auto label = [](const char* text) { std::printf("%-30s : ", text); };
    std::printf("\nSimulation complete\n%s\n", std::string(72, '-').c_str());
    label("Timesteps completed");
    std::printf("%d\n", steps);
    if (csv_.rows()) {
      label("Diagnostics rows");
      std::printf("%d\n", csv_.rows());
      label("Diagnostics file");
      std::printf("%s\n", csv_.path().c_str());
    }
    if (snapshots_) {
      label("Field snapshots");
      std::printf("%d\n", snapshots_);
      label("Last snapshot");
      std::printf("%s\n", last_snapshot_.c_str());
    }
    if (particle_snapshots_) {
      label("Particle snapshots");
      std::printf("%d\n", particle_snapshots_);
      label("Last particle snapshot");
      std::printf("%s\n", last_particles_.c_str());
    }
    if (checkpoints_) {
      label("Checkpoints");
      std::printf("%d\n", checkpoints_);
      label("Last checkpoint");
      std::printf("%s\n", last_checkpoint_.c_str());
    }
    if (!csv_.rows() && !snapshots_ && !particle_snapshots_ && !checkpoints_) std::printf("No output files written.\n");
    std::printf("%s\n", std::string(72, '=').c_str());
  }

 private:
  void energy() {
    const double change = total_ - initial_;
    std::printf("  Field energy  : electric = %.12e  magnetic = %.12e\n"
                "  Total energy  : system   = %.12e  kinetic  = %.12e\n",
                electric_, magnetic_, total_, kinetic_);
    if (initial_ != 0)
      std::printf("  Energy change : dE/E0 = %+.6e  dE = %+.6e\n", change / initial_, change);
    else
      std::printf("  Energy change : dE/E0 = n/a (zero initial energy)  dE = %+.6e\n", change);
    std::fflush(stdout);
  }

  void phases() {
    const char* names[] = {"sort", "move", "deposit", "assemble", "solve", "advance", "spmv"};
    std::printf("\nPhase times (HIP events on the context stream)\n");
    for (int ph = 0; ph < 7; ++ph) {
      double ms = 0;
      uint64_t calls = 0;
      check(tpic_mi_phase_time(ctx_, ph, &ms, &calls));
      std::printf("  %-9s: %12.3f ms in %llu calls\n", names[ph], ms, static_cast<unsigned long long>(calls));
    }
  }

 public:
  void set_timing(bool t) { timing_ = t; }

 private:
  const Prepared& r_;
  const Restart* restart_ = nullptr;
  int first_cycle_ = 0;
  int checkpoints_ = 0;
  std::string last_checkpoint_;
  tpic_mi_context* ctx_ = nullptr;
  bool timing_ = false;
  double electric_ = 0, magnetic_ = 0, kinetic_ = 0, total_ = 0, initial_ = 0;
  std::vector<double> scalars_;
  tpic_mi::output::DiagnosticsCsv csv_;
  int snapshots_ = 0, particle_snapshots_ = 0;
  std::string last_snapshot_, last_particles_;
};

}  // namespace

int main(int argc, char** argv) {
  try {
    Options options;
    bool check = false, initialize = false, run = false, rank_given = false, timing = false, print = false;
    std::string path;
    for (int i = 1; i < argc; ++i) {
      const std::string_view arg = argv[i];
      if (arg == "--help") {
        std::cout << kUsage << "Runtime default precision (storage/compute): double/double\n";
        return 0;
      }
      if (arg == "--check-input") { check = true; continue; }
      if (arg == "--initialize") { initialize = true; continue; }
      if (arg == "--run") { run = true; continue; }
      if (arg == "--overwrite-output") { options.overwrite_output = true; continue; }
      if (arg == "--restart") { options.restart = true; continue; }
      if (arg == "--timing") { timing = true; continue; }
      if (arg == "--print-input") { print = true; continue; }
      if (arg == "--rank" || arg == "--order" || arg == "--mass-matrix" || arg == "--mass-apply" ||
          arg == "--precision" || arg == "--storage-precision" || arg == "--compute-precision" ||
          arg == "--mass-precision") {
        if (++i == argc) option_error(std::string(arg) + " requires a value");
        const std::string_view v = argv[i];
        if (arg == "--rank") {
          char* end = nullptr;
          errno = 0;
          const long rank = std::strtol(argv[i], &end, 10);
          if (v.empty() || *end != '\0' || errno || rank < 0 || rank > 0x7fffffff)
            option_error("--rank requires a nonnegative signed integer");
          options.rank = int(rank);
          rank_given = true;
        } else if (arg == "--order") {
          if (v != "full" && v != "partial" && v != "unsorted")
            option_error("--order must be full, partial or unsorted");
          options.order = std::string(v);
        } else if (arg == "--mass-matrix") {
          if (v != "scalar" && v != "matrix-core") option_error("--mass-matrix must be scalar or matrix-core");
          options.mass_matrix = std::string(v);
        } else if (arg == "--mass-apply") {
          if (v != "full" && v != "split") option_error("--mass-apply must be full or split");
          options.mass_apply = std::string(v);
        } else if (arg == "--mass-precision") {
          if (v == "fp64") options.mass_operands = TPIC_MI_MASS_FP64;
          else if (v == "fp32") options.mass_operands = TPIC_MI_MASS_FP32;
          else if (v == "bf16x3") options.mass_operands = TPIC_MI_MASS_BF16X3;
          else option_error("--mass-precision must be fp64, fp32 or bf16x3");
        } else {
          if (!is_level(v))
            throw Invalid("TPIC precision: " + std::string(arg) +
                          " must be double, single-position, single-position-velocity, "
                          "single-particles-moments-fields-double-mass or single-particles-moments-fields");
          if (arg == "--precision") options.storage_precision = options.compute_precision = std::string(v);
          else if (arg == "--storage-precision") options.storage_precision = std::string(v);
          else options.compute_precision = std::string(v);
        }
        continue;
      }
      if (!arg.empty() && arg[0] == '-') option_error("unknown option: " + std::string(arg));
      if (!path.empty()) option_error("expected exactly one input file");
      path = std::string(arg);
    }
    if (print) {
      // The parsed deck alone: every key with its resolved value, then the
      // unknown keys (no admission, no GPU).
      if (int(check) + int(initialize) + int(run) != 0 || path.empty())
        option_error("--print-input takes one input file and no command");
      const auto parsed = tpic_mi::deck::load(path);
      for (const auto& [key, value] : parsed.resolved) std::printf("%s = %s\n", key.c_str(), value.c_str());
      for (const auto& key : parsed.unknown_keys) std::printf("unknown: %s\n", key.c_str());
      return 0;
    }
    if (int(check) + int(initialize) + int(run) != 1 || path.empty())
      option_error("give exactly one of --check-input, --initialize or --run and one input file\n" +
                   std::string(kUsage));
    if ((initialize || run) && rank_given)
      option_error("--rank is only valid with --check-input; this build runs rank 0 of one partition");
    if (check && options.restart) option_error("--restart is only valid with --run or --initialize");
    // A restart replaces the outputs of the run it continues (run_plan.cpp:92-93).
    if (options.restart) options.overwrite_output = true;

    Prepared prepared = tpic_mi::deck::prepare(tpic_mi::deck::load(path), options);
    if (check) {
      print_check(prepared);
      return prepared.admitted() ? 0 : 2;
    }
    Restart restart;
    if (options.restart && prepared.admitted()) restart = find_restart(prepared, prepared.unsupported);
    if (!prepared.admitted()) {
      for (const auto& u : prepared.unsupported) print_error("unsupported request: " + u);
      return 2;
    }
    for (const auto& w : prepared.warnings) std::printf("Warning: %s\n", w.c_str());
    const auto& p = prepared.p;
    std::printf("TPIC MI355X ECSIM: %s / %s, %dx%dx%d cells, %d species, mass operands %s\n",
                p.sim_name.empty() ? "(unnamed)" : p.sim_name.c_str(), p.case_name.c_str(), p.cells[0], p.cells[1],
                p.cells[2], p.ns, mass_operands_name(options.mass_operands));
    Run sim(prepared, options.restart ? &restart : nullptr);
    sim.set_timing(timing);
    const auto t0 = std::chrono::steady_clock::now();
    sim.startup(timing);
    const double boot = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (initialize) {
      std::printf("\nInitialization complete\n  Startup time: %.6f s\n", boot);
      return 0;
    }
    const int steps = sim.cycles();
    std::printf("  Bootstrap time: %.6f s (load, centring, first sort and deposition)\n", boot);
    sim.completion(steps);
    return 0;
  } catch (const std::exception& e) {
    print_error(e.what());
    return 1;
  }
}
