// Input deck parsing and run admission of the tpic_mi_run driver.
#include "deck.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <limits>
#include <sstream>
#include <string_view>
#include <type_traits>
#include <unordered_map>

namespace tpic_mi::deck {
namespace {

[[noreturn]] void invalid(std::string_view topic, const std::string& message) {
  throw Invalid("TPIC " + std::string(topic) + ": " + message);
}

std::string trim(std::string_view text) {
  size_t a = 0, b = text.size();
  while (a < b && std::isspace(static_cast<unsigned char>(text[a]))) ++a;
  while (b > a && std::isspace(static_cast<unsigned char>(text[b - 1]))) --b;
  return std::string(text.substr(a, b - a));
}

template <class T>
std::string kind() {
  if constexpr (std::is_same_v<T, bool>) return "a boolean (0, 1, true or false)";
  else if constexpr (std::is_floating_point_v<T>) return "a number";
  else return "an integer";
}

// One token read in full: stream extraction must consume it all
// (simulation_parameters.cpp:42-51).
template <class T>
bool token(const std::string& text, T& out) {
  if constexpr (std::is_same_v<T, bool>) {
    if (text == "1" || text == "true") { out = true; return true; }
    if (text == "0" || text == "false") { out = false; return true; }
    return false;
  } else {
    std::istringstream in(text);
    T value{};
    in >> value;
    if (in.fail()) return false;
    in >> std::ws;
    if (!in.eof()) return false;
    out = value;
    return true;
  }
}

class Deck {
 public:
  explicit Deck(std::istream& in) {
    std::string line;
    for (int number = 1; std::getline(in, line); ++number) {
      const std::string body = trim(std::string_view(line).substr(0, line.find('#')));
      if (body.empty()) continue;
      const size_t eq = body.find('=');
      std::string key = trim(std::string_view(body).substr(0, eq));
      if (eq == std::string::npos || key.empty())
        invalid("input", "line " + std::to_string(number) + ": expected 'key = value'");
      entries_[key] = Entry{trim(std::string_view(body).substr(eq + 1)), number, false};
    }
  }

  template <class T>
  void read(const std::string& key, T& target) {
    if (Entry* e = find(key)) target = value<T>(key, *e);
  }

  template <class T>
  T optional(const std::string& key, T fallback) {
    read(key, fallback);
    return fallback;
  }

  template <class T>
  T required(const std::string& key) {
    Entry* e = find(key);
    if (!e) invalid("input", "missing required key " + key);
    return value<T>(key, *e);
  }

  template <class T>
  std::vector<T> list(const std::string& key, size_t at_least) {
    Entry* e = find(key);
    if (!e) {
      if (at_least) invalid("input", "missing required key " + key);
      return {};
    }
    std::vector<T> values;
    std::istringstream in(e->value);
    for (std::string t; in >> t;) {
      T v{};
      if (!token(t, v))
        invalid("input", at(*e) + key + " = '" + e->value + "': '" + t + "' is not " + kind<T>());
      values.push_back(v);
    }
    if (values.size() < at_least)
      invalid("input", at(*e) + key + " lists " + std::to_string(values.size()) + " values, expected at least " +
                           std::to_string(at_least));
    return values;
  }

  // A value from a closed set of spellings.
  void choice(const std::string& key, std::string& target, const std::vector<std::string>& names) {
    Entry* e = find(key);
    if (!e) return;
    if (std::find(names.begin(), names.end(), e->value) == names.end()) reject(key, *e, names);
    target = e->value;
  }

  // An output method: empty, default or an engine spelling (src/config/output.cpp:21-25,77-89).
  std::string engine(const std::string& key) {
    static const std::vector<std::string> spellings = {"pvtk", "vtk", "adios2", "bp", "bp5", "hdf5", "h5"};
    Entry* e = find(key);
    if (!e) return "";
    if (e->value.empty() || e->value == "default") return e->value;
    if (std::find(spellings.begin(), spellings.end(), e->value) != spellings.end()) return e->value;
    std::string expected = "default";
    for (size_t i = 0; i < spellings.size(); ++i) expected += (i + 1 == spellings.size() ? " or " : ", ") + spellings[i];
    invalid("input", at(*e) + key + " = '" + e->value + "' is unknown (expected " + expected + ")");
  }

  void require_in(const std::string& key, const std::vector<std::string>& names) {
    Entry& e = entries_.at(key);
    if (std::find(names.begin(), names.end(), e.value) == names.end()) reject(key, e, names);
  }

  std::vector<std::string> unread() const {
    std::vector<std::pair<int, std::string>> keys;
    for (const auto& [key, e] : entries_)
      if (!e.read) keys.emplace_back(e.line, key);
    std::sort(keys.begin(), keys.end());
    std::vector<std::string> out;
    for (const auto& [line, key] : keys) out.push_back(key + " (line " + std::to_string(line) + ")");
    return out;
  }

 private:
  struct Entry {
    std::string value;
    int line;
    bool read;
  };

  Entry* find(const std::string& key) {
    auto it = entries_.find(key);
    if (it == entries_.end()) return nullptr;
    it->second.read = true;
    return &it->second;
  }

  static std::string at(const Entry& e) { return "line " + std::to_string(e.line) + ": "; }

  template <class T>
  static T value(const std::string& key, const Entry& e) {
    if constexpr (std::is_same_v<T, std::string>) {
      return e.value;
    } else {
      T v{};
      if (!token(e.value, v)) invalid("input", at(e) + key + " = '" + e.value + "' is not " + kind<T>());
      return v;
    }
  }

  [[noreturn]] static void reject(const std::string& key, const Entry& e, const std::vector<std::string>& names) {
    std::string expected;
    for (size_t i = 0; i < names.size(); ++i)
      expected += (i == 0 ? "" : i + 1 == names.size() ? " or " : ", ") + names[i];
    invalid("input", at(e) + key + " = '" + e.value + "' is unknown (expected " + expected + ")");
  }

  std::unordered_map<std::string, Entry> entries_;
};

// config::case_names (include/tpic/config/simulation_parameters.hpp:204-207).
const std::vector<std::string> kCaseNames = {
    "Uniform", "GEM", "GEMnoPert", "GEMHarris", "DoubleGEM", "DoubleGEMnoPert", "plasmoidChain", "ForceFree",
    "Weibel", "Whistler", "TaylorGreenVortex", "TwoStream", "fluxRope", "GaussianExpansion", "WhistlerTurbulence"};

constexpr int kMaxRealSpecies = 10;  // TPIC_CONFIG_MAX_SPECIES of a default reference build

}  // namespace

namespace {

std::string show(double v) {
  char buf[40];
  std::snprintf(buf, sizeof buf, "%.17g", v);
  return buf;
}
std::string show(int v) { return std::to_string(v); }
std::string show(uint64_t v) { return std::to_string(v); }
std::string show(bool v) { return v ? "true" : "false"; }
std::string show(const std::string& v) { return v; }

// The deck's resolved values, key by key, in read order.
class Resolved {
 public:
  explicit Resolved(Deck& d, Parameters& p) : d_(d), p_(p) {}

  template <class T>
  T get(const std::string& key, T fallback) {
    d_.read(key, fallback);
    put(key, show(fallback));
    return fallback;
  }
  template <class T>
  T need(const std::string& key) {
    const T v = d_.required<T>(key);
    put(key, show(v));
    return v;
  }
  std::string pick(const std::string& key, std::string fallback, const std::vector<std::string>& names) {
    d_.choice(key, fallback, names);
    put(key, fallback);
    return fallback;
  }
  void put(const std::string& key, const std::string& value) { p_.resolved.emplace_back(key, value); }

 private:
  Deck& d_;
  Parameters& p_;
};

}  // namespace

Parameters parse(std::istream& in) {
  Deck d(in);
  Parameters p;
  Resolved r(d, p);
  p.dt = r.need<double>("dt");
  p.ncycles = r.need<int>("ncycles");
  p.th = r.get("th", p.th);
  p.smooth = r.get("Smooth", p.smooth);
  p.smooth_niter = r.get("SmoothNiter", p.smooth_niter);
  p.smoothing = r.pick("ECSimSmoothing", p.smoothing, {"none", "energy_conserving"});
  p.smoothing_operator = r.pick("ECSimSmoothingOperator", p.smoothing_operator, {"face7", "binomial27"});
  r.get<std::string>("SaveDirName", "data");
  r.get<std::string>("RestartDirName", "data");
  p.ns = r.need<int>("ns");
  p.ns_test = r.get("nsTestPart", p.ns_test);
  if (p.ns < 1 || p.ns > kMaxRealSpecies)
    invalid("input", "unsupported ns=" + std::to_string(p.ns) + " (supported real species range: 1-" +
                         std::to_string(kMaxRealSpecies) + ")");
  if (p.ns_test < 0) invalid("input", "nsTestPart must be nonnegative");
  if (p.ns_test > std::numeric_limits<int>::max() - p.ns)
    invalid("input", "ns + nsTestPart exceeds the supported integer range");
  const size_t real = size_t(p.ns), total = size_t(p.ns) + size_t(p.ns_test);

  p.b0[0] = r.get("B0x", 0.0);
  p.b0[1] = r.get("B0y", 0.0);
  p.b0[2] = r.get("B0z", 0.0);
  // The other cases' recipe keys (defaults of config::SimulationParameters):
  // read, and so checked, but not used by the Uniform case.
  r.get("delta", 0.5);
  r.get("HarrisB0x", 0.0);
  r.get("pertGEM", 0.0);
  r.get("pertHump", 0.0);
  r.get("deltaxHump", 8.0);
  r.get("deltayHump", 4.0);
  r.get("kxHump", -1.0);
  r.get("kyHump", -1.0);
  r.get("currentFromAmpere", false);
  r.get("spatiallyVaryingThermal", false);
  r.get("ForceFreeB0x", 0.0);
  r.get("ForceFreeB0z", 0.0);
  r.get("FluxRopeB0", 0.1);
  r.get("FluxRopeWidth", 1.0);
  r.get("FluxRopePressureFloor", 9.0e-4);
  r.get("FluxRopeElectronPressureFraction", 0.1);
  r.get("WeibelSeedModeX", 0);
  r.get("WeibelSeedAmplitude", 0.0);
  r.get("WeibelSeedGrowthRate", 0.0);
  r.get("TwoStreamSeedModeX", 0);
  r.get("TwoStreamSeedDensityAmplitude", 0.0);
  r.get("TwoStreamSeedPhase", 0.0);
  r.get("WhistlerModeX", 0);
  r.get("WhistlerModeY", 0);
  r.get("WhistlerModeZ", 0);
  r.get("WhistlerSeedEpsilon", 0.0);
  r.get("WhistlerPhase", 0.0);
  r.get("WhistlerUseExactColdRoot", true);
  r.pick("WhistlerSeedMode", "eigenmode", {"eigenmode", "magnetic_only", "noise"});
  r.get<uint64_t>("WhistlerNoiseSeed", 12345);
  r.get("WhistlerNoiseModeMax", 0);
  r.get("WhistlerFieldTimeOffset", 0.0);
  r.get("WhistlerParticleTimeOffset", 0.0);
  // TaylorGreenVortexMode is the default of the three per-axis modes.
  const int tgv = r.get("TaylorGreenVortexMode", 1);
  r.get("TaylorGreenVortexModeX", tgv);
  r.get("TaylorGreenVortexModeY", tgv);
  r.get("TaylorGreenVortexModeZ", tgv);
  r.get("TaylorGreenVortexVelocityAmplitude", 0.0);
  r.get("TaylorGreenVortexMagneticAmplitude", 0.0);
  r.get("TaylorGreenVortexSeedMotionalE", false);
  r.get("GaussianExpansionSigmaX", 0.0);
  r.get("GaussianExpansionSigmaY", 0.0);
  r.get("GaussianExpansionSigmaZ", 0.0);
  r.get("GaussianExpansionCutoff", 0.0);
  r.get("WhistlerTurbulenceEpsilon", 0.0);
  r.get<uint64_t>("WhistlerTurbulenceSeed", 1);

  p.c = r.get("c", p.c);
  p.lengths[0] = r.get("Lx", p.lengths[0]);
  p.lengths[1] = r.get("Ly", p.lengths[1]);
  p.lengths[2] = r.get("Lz", p.lengths[2]);
  p.cells[0] = r.get("nxc", p.cells[0]);
  p.cells[1] = r.get("nyc", p.cells[1]);
  p.cells[2] = r.get("nzc", p.cells[2]);
  // Centres default to the middle of the box.
  r.get("FluxRopeCenterX", 0.5 * p.lengths[0]);
  r.get("FluxRopeCenterY", 0.5 * p.lengths[1]);
  r.get("GaussianExpansionCenterX", 0.5 * p.lengths[0]);
  r.get("GaussianExpansionCenterY", 0.5 * p.lengths[1]);
  r.get("GaussianExpansionCenterZ", 0.5 * p.lengths[2]);

  p.processes[0] = r.get("XLEN", p.processes[0]);
  p.processes[1] = r.get("YLEN", p.processes[1]);
  p.processes[2] = r.get("ZLEN", p.processes[2]);
  p.periodic[0] = r.get("PERIODICX", p.periodic[0]);
  p.periodic[1] = r.get("PERIODICY", p.periodic[1]);
  p.periodic[2] = r.get("PERIODICZ", p.periodic[2]);
  p.particle_periodic[0] = r.get("PERIODICX_P", p.periodic[0]);
  p.particle_periodic[1] = r.get("PERIODICY_P", p.periodic[1]);
  p.particle_periodic[2] = r.get("PERIODICZ_P", p.periodic[2]);

  p.case_name = r.need<std::string>("Case");
  d.require_in("Case", kCaseNames);
  p.quiet_start = r.get("QuietStart", p.quiet_start);
  for (const char* key : {"WriteMethod", "FieldWriteMethod", "ParticlesWriteMethod", "RestartWriteMethod"})
    r.put(key, d.engine(key));
  p.sim_name = r.get("SimulationName", p.sim_name);
  p.gmres_tol = r.get("GMREStol", p.gmres_tol);

  p.field_cycle = r.get("FieldOutputCycle", p.field_cycle);
  p.particles_cycle = r.get("ParticlesOutputCycle", p.particles_cycle);
  p.test_particles_cycle = r.get("TestParticlesOutputCycle", p.test_particles_cycle);
  p.restart_cycle = r.get("RestartOutputCycle", p.restart_cycle);
  p.call_finalize = r.get("CallFinalize", p.call_finalize);
  p.diagnostics_cycle = r.get("DiagnosticsOutputCycle", p.field_cycle);
  p.energy_cycle = r.get("DiagnosticEnergyCycle", p.energy_cycle);
  p.gauss_cycle = r.get("DiagnosticGaussCycle", p.gauss_cycle);
  for (const char* key : {"FieldOutputTag", "ParticlesOutputTag", "MomentsOutputTag"}) r.get<std::string>(key, "");

  const char* faces[6] = {"Xright", "Xleft", "Yright", "Yleft", "Zright", "Zleft"};
  for (int f = 0; f < 6; ++f) p.em_face[f] = r.get(std::string("bcEMface") + faces[f], p.em_face[f]);
  for (int f = 0; f < 6; ++f) p.particle_face[f] = r.get(std::string("bcPface") + faces[f], p.particle_face[f]);

  // Every list is read and checked before the species are sized. Real species
  // precede test species; a missing optional tail keeps the default 0.
  const auto rho = d.list<double>("rhoINIT", real);
  const auto qom = d.list<double>("qom", total);
  const auto px = d.list<int>("npcelx", total);
  const auto py = d.list<int>("npcely", total);
  const auto pz = d.list<int>("npcelz", total);
  const auto uth = d.list<double>("uth", real);
  const auto vth = d.list<double>("vth", real);
  const auto wth = d.list<double>("wth", real);
  const auto u0 = d.list<double>("u0", real);
  const auto v0 = d.list<double>("v0", real);
  const auto w0 = d.list<double>("w0", real);
  const auto center = d.list<bool>("CenterInitialVelocity", 0);
  std::vector<Species> all(total);
  auto assign = [&](auto member, const auto& values) {
    for (size_t s = 0; s < std::min(values.size(), total); ++s) member(all[s], values[s]);
  };
  assign([](Species& x, double v) { x.rho_init = v; }, rho);
  assign([](Species& x, double v) { x.q_over_m = v; }, qom);
  assign([](Species& x, int v) { x.ppc[0] = v; }, px);
  assign([](Species& x, int v) { x.ppc[1] = v; }, py);
  assign([](Species& x, int v) { x.ppc[2] = v; }, pz);
  assign([](Species& x, double v) { x.thermal[0] = v; }, uth);
  assign([](Species& x, double v) { x.thermal[1] = v; }, vth);
  assign([](Species& x, double v) { x.thermal[2] = v; }, wth);
  assign([](Species& x, double v) { x.drift[0] = v; }, u0);
  assign([](Species& x, double v) { x.drift[1] = v; }, v0);
  assign([](Species& x, double v) { x.drift[2] = v; }, w0);
  assign([](Species& x, bool v) { x.center_initial_velocity = v; }, center);
  auto column = [&](const char* key, auto field) {
    std::string text;
    for (size_t s = 0; s < total; ++s) text += (s ? " " : "") + show(field(all[s]));
    r.put(key, text);
  };
  column("rhoINIT", [](const Species& x) { return x.rho_init; });
  column("qom", [](const Species& x) { return x.q_over_m; });
  column("npcelx", [](const Species& x) { return x.ppc[0]; });
  column("npcely", [](const Species& x) { return x.ppc[1]; });
  column("npcelz", [](const Species& x) { return x.ppc[2]; });
  column("uth", [](const Species& x) { return x.thermal[0]; });
  column("vth", [](const Species& x) { return x.thermal[1]; });
  column("wth", [](const Species& x) { return x.thermal[2]; });
  column("u0", [](const Species& x) { return x.drift[0]; });
  column("v0", [](const Species& x) { return x.drift[1]; });
  column("w0", [](const Species& x) { return x.drift[2]; });
  column("CenterInitialVelocity", [](const Species& x) { return x.center_initial_velocity; });
  p.species.assign(all.begin(), all.begin() + std::ptrdiff_t(real));
  p.test_species.assign(all.begin() + std::ptrdiff_t(real), all.end());
  p.unknown_keys = d.unread();
  return p;
}

Parameters load(const std::string& path) {
  std::ifstream in(path);
  if (!in) invalid("input", "unable to open input file " + path);
  return parse(in);
}

const std::string& resolved(const Parameters& p, const std::string& key) {
  for (const auto& [k, v] : p.resolved)
    if (k == key) return v;
  invalid("input", "no resolved value for " + key);
}

namespace {

double number(const Parameters& p, const std::string& key) { return std::stod(resolved(p, key)); }

[[noreturn]] void recipe_error(const std::string& message) { throw Invalid("TPIC recipe: " + message); }
void require_recipe(bool ok, const std::string& message) {
  if (!ok) recipe_error(message);
}
void require_finite(double v, const std::string& name) {
  if (!std::isfinite(v)) recipe_error(name + " must be finite");
}
double representable(double v, const std::string& name) {
  if (!std::isfinite(v)) throw std::overflow_error("TPIC recipe: " + name + " is not representable");
  return v;
}
// initial_case.cpp:59-63.
double mode_wavenumber(int mode, double length) {
  const double k = representable(6.283185307179586476925286766559 * double(mode) / length, "wavenumber");
  require_recipe(mode == 0 || k != 0.0, "nonzero-mode wavenumber underflows");
  return k;
}
bool nearly_equal(double a, double b) {
  return std::abs(a - b) <= 1.0e-10 * std::max({1.0, std::abs(a), std::abs(b)});
}

// TwoStreamCase (src/initialization/cases/two_stream.cpp:16-73).
void two_stream(const Parameters& p, CaseRecipe& c) {
  const int mode = std::stoi(resolved(p, "TwoStreamSeedModeX"));
  const double amplitude = number(p, "TwoStreamSeedDensityAmplitude"), phase = number(p, "TwoStreamSeedPhase");
  require_recipe(mode >= 0, "TwoStreamSeedModeX must be non-negative");
  require_recipe(mode == 0 || 2LL * mode < p.cells[0], "TwoStreamSeedModeX must be below the x-grid Nyquist mode");
  require_finite(amplitude, "TwoStreamSeedDensityAmplitude");
  require_finite(phase, "TwoStreamSeedPhase");
  require_recipe(std::abs(amplitude) < 1.0, "abs(TwoStreamSeedDensityAmplitude) must be less than one");
  require_recipe(amplitude == 0.0 || mode > 0, "TwoStreamSeedModeX must be positive for nonzero seed amplitude");
  require_recipe(p.periodic[0] && p.particle_periodic[0],
                 "TwoStream requires periodic x boundaries for fields and particles");
  std::vector<const Species*> electrons;
  const Species* ion = nullptr;
  for (const auto& sp : p.species) {
    require_recipe(sp.rho_init > 0.0, "TwoStream species rhoINIT must be positive");
    if (sp.q_over_m < 0.0) electrons.push_back(&sp);
    else ion = &sp;
  }
  require_recipe(p.ns == 3 && electrons.size() == 2 && ion,
                 "TwoStream requires exactly two electron species and one ion species");
  const Species& a = *electrons[0];
  const Species& b = *electrons[1];
  require_recipe(nearly_equal(a.rho_init, b.rho_init), "TwoStream electron rhoINIT must match");
  require_recipe(nearly_equal(a.q_over_m, b.q_over_m), "TwoStream electron qom must match");
  require_recipe(nearly_equal(a.thermal[0], b.thermal[0]) && nearly_equal(a.thermal[1], b.thermal[1]) &&
                     nearly_equal(a.thermal[2], b.thermal[2]),
                 "TwoStream electron thermal velocities must match");
  require_recipe(a.drift[0] != 0.0 && nearly_equal(a.drift[0], -b.drift[0]),
                 "TwoStream electron u0 must be equal and opposite and nonzero");
  const double electrons_density = representable(a.rho_init + b.rho_init, "TwoStream electron density");
  require_recipe(nearly_equal(electrons_density, ion->rho_init),
                 "TwoStream ion rhoINIT must equal the sum of electron rhoINIT");
  require_recipe(nearly_equal(ion->drift[0], 0.0), "TwoStream ion drift must be zero");
  for (const auto& sp : p.species) {
    require_recipe(nearly_equal(sp.drift[1], 0.0) && nearly_equal(sp.drift[2], 0.0),
                   "TwoStream transverse drifts must be zero");
    require_recipe(sp.ppc[0] == a.ppc[0] && sp.ppc[1] == a.ppc[1] && sp.ppc[2] == a.ppc[2],
                   "TwoStream requires identical particles-per-cell dimensions");
  }
  c.k = mode_wavenumber(mode, p.lengths[0]);
  c.phase = phase;
  c.density_amplitude = amplitude;
  if (amplitude != 0.0) {
    c.e_amplitude = representable(-(electrons_density * amplitude / c.k), "TwoStream E amplitude");
    c.e_seed = true;
  }
  for (size_t s = 0; s < p.species.size(); ++s)
    c.density_seeded[s] = p.species[s].q_over_m < 0.0 && amplitude != 0.0;
}

// WeibelCase (src/initialization/cases/weibel.cpp:9-40).
void weibel(const Parameters& p, CaseRecipe& c) {
  const double amplitude = number(p, "WeibelSeedAmplitude"), growth = number(p, "WeibelSeedGrowthRate");
  const int mode = std::stoi(resolved(p, "WeibelSeedModeX"));
  require_finite(amplitude, "WeibelSeedAmplitude");
  require_finite(growth, "WeibelSeedGrowthRate");
  require_recipe(mode >= 0, "WeibelSeedModeX must be non-negative");
  require_recipe(amplitude == 0.0 || mode > 0, "WeibelSeedModeX must be positive for nonzero seed amplitude");
  c.k = mode_wavenumber(mode, p.lengths[0]);
  if (amplitude != 0.0) {
    c.e_amplitude = representable(-(growth / c.k) * amplitude, "Weibel E amplitude");
    c.b_amplitude = amplitude;
    c.b_seed = true;
    c.e_seed = growth != 0.0;
  }
}

constexpr double kPi = 3.1415926535897932384626433832795;
constexpr double kTwoPi = 6.283185307179586476925286766559;

// initial_case.cpp:32-57.
double nonzero_representable(double v, const std::string& name) {
  representable(v, name);
  if (v == 0.0) throw std::overflow_error("TPIC recipe: " + name + " underflows");
  return v;
}
double checked_product(double a, double b, const std::string& name) {
  return a != 0.0 && b != 0.0 ? nonzero_representable(a * b, name) : representable(a * b, name);
}
double checked_quotient(double n, double d, const std::string& name) {
  return n != 0.0 ? nonzero_representable(n / d, name) : representable(n / d, name);
}
double coordinate_extent(double length, int cells, double center) {
  const double spacing = nonzero_representable(length / cells, "recipe grid spacing");
  const double upper = representable(length + spacing, "recipe upper halo coordinate");
  return std::max(std::abs(representable(-spacing - center, "recipe lower displacement")),
                  std::abs(representable(upper - center, "recipe upper displacement")));
}

bool has_hump(const CaseRecipe::Seed& s) { return s.amplitude_x != 0.0 || s.amplitude_y != 0.0; }

// deck_seed (harris_sheet.cpp:19-45).
CaseRecipe::Seed deck_seed(const Parameters& p, double delta) {
  const double pert_gem = number(p, "pertGEM"), pert_hump = number(p, "pertHump");
  const double kx_hump = number(p, "kxHump"), ky_hump = number(p, "kyHump");
  const double deltax = number(p, "deltaxHump"), deltay = number(p, "deltayHump");
  require_finite(pert_gem, "pertGEM");
  require_finite(pert_hump, "pertHump");
  require_finite(kx_hump, "kxHump");
  require_finite(ky_hump, "kyHump");
  require_recipe(std::isfinite(deltax) && deltax > 0.0, "GEMHarris deltaxHump must be finite and positive");
  require_recipe(std::isfinite(deltay) && deltay > 0.0, "GEMHarris deltayHump must be finite and positive");
  CaseRecipe::Seed seed;
  seed.mode_amplitude = pert_gem;
  seed.mode_kx = mode_wavenumber(1, p.lengths[0]);
  seed.mode_ky = mode_wavenumber(1, p.lengths[1]);
  if (pert_hump == 0.0) return seed;
  const double kx = kx_hump < 0.0 ? kTwoPi / p.lengths[0] : kx_hump;
  const double ky = ky_hump < 0.0 ? .5 * kTwoPi / p.lengths[1] : ky_hump;
  require_recipe(ky != 0.0, "GEMHarris kyHump must be nonzero: the hump amplitude is pertHump*HarrisB0x/kyHump");
  seed.width_x = checked_product(deltax, delta, "GEMHarris hump x width");
  seed.width_y = checked_product(deltay, delta, "GEMHarris hump y width");
  seed.amplitude_x = checked_quotient(pert_hump, checked_product(ky, seed.width_y, "GEMHarris hump y slope"),
                                      "GEMHarris hump Bx amplitude");
  seed.amplitude_y = checked_quotient(pert_hump, checked_product(ky, seed.width_x, "GEMHarris hump x slope"),
                                      "GEMHarris hump By amplitude");
  seed.period_x = kPi / (kx * delta);
  seed.period_y = representable(kPi / (ky * delta), "GEMHarris hump y period");
  seed.slope_period_x = kPi / (kx * seed.width_x);
  seed.slope_period_y = kPi / (ky * seed.width_y);
  return seed;
}

// validate_harris_coefficients (harris_sheet.cpp:47-104).
void validate_harris(const Parameters& p, const CaseRecipe& c) {
  const auto& seed = c.seed;
  double bx_bound = std::abs(c.harris_b0x), by_bound = 0.0, extent_x = 0.0, extent_y = 0.0;
  if (has_hump(seed) || seed.mode_amplitude != 0.0) {
    const double center_fraction = c.harris_double ? .25 : .5;
    extent_x = coordinate_extent(p.lengths[0], p.cells[0], center_fraction * p.lengths[0]);
    extent_y = coordinate_extent(p.lengths[1], p.cells[1], center_fraction * p.lengths[1]);
  }
  if (has_hump(seed)) {
    const double axes[2][4] = {{extent_x, seed.period_x, seed.width_x, seed.slope_period_x},
                               {extent_y, seed.period_y, seed.width_y, seed.slope_period_y}};
    for (const auto& a : axes) {
      representable((checked_product(.5 * kTwoPi, a[0], "Harris phase numerator") / a[1]) / c.delta,
                    "Harris perturbation phase bound");
      const double gradient =
          representable(checked_product(2.0, a[0], "Harris gradient numerator") / a[2], "Harris perturbation gradient bound");
      representable(gradient + .5 * kTwoPi / a[3], "Harris perturbation gradient sum");
    }
    const double sheets = c.harris_double ? 2.0 : 1.0;
    bx_bound = representable(
        bx_bound + sheets * checked_product(checked_product(std::abs(c.harris_b0x), std::abs(seed.amplitude_x),
                                                            "Harris perturbation amplitude"),
                                            1.0 + .5 * kTwoPi / seed.slope_period_y, "Harris perturbation field bound"),
        "Harris total sheet field bound");
    by_bound = representable(
        sheets * checked_product(checked_product(std::abs(c.harris_b0x), std::abs(seed.amplitude_y),
                                                 "Harris perturbation amplitude"),
                                 1.0 + .5 * kTwoPi / seed.slope_period_x, "Harris perturbation field bound"),
        "Harris perturbation field bound");
  }
  if (seed.mode_amplitude != 0.0) {
    checked_product(seed.mode_kx, extent_x, "Harris mode phase bound");
    checked_product(seed.mode_ky, extent_y, "Harris mode phase bound");
    const double amplitude =
        checked_product(std::abs(c.harris_b0x), std::abs(seed.mode_amplitude), "Harris mode amplitude");
    bx_bound = representable(bx_bound + checked_product(amplitude, p.lengths[0] / p.lengths[1], "Harris mode field bound"),
                             "Harris total sheet field bound");
    by_bound = representable(by_bound + amplitude, "Harris mode field bound");
  }
  representable(bx_bound + std::abs(p.b0[0]), "Harris Bx plus background bound");
  representable(by_bound + std::abs(p.b0[1]), "Harris By plus background bound");
}

// admit_ampere_current (harris_sheet.cpp:127-181).
void admit_ampere_current(const Parameters& p, CaseRecipe& c) {
  double drift_sum = 0.0;
  for (const auto& sp : p.species) {
    require_recipe(sp.drift[0] == 0.0 && sp.drift[1] == 0.0,
                   "GEMHarris currentFromAmpere requires u0 = v0 = 0: the curl of the initial field is along z");
    drift_sum += (sp.q_over_m > 0.0 ? 1.0 : -1.0) * sp.drift[2];
  }
  require_recipe(std::isfinite(drift_sum) && drift_sum != 0.0,
                 "GEMHarris currentFromAmpere requires a nonzero sum of sign(qom)*w0");
  const auto& seed = c.seed;
  const double b0 = c.harris_b0x;
  auto& curl = c.curl;
  curl.center_x = 0.5 * p.lengths[0];
  curl.center_y = 0.5 * p.lengths[1];
  curl.delta = c.delta;
  curl.sheet = checked_quotient(-b0, c.delta, "GEMHarris sheet current");
  curl.mode = checked_product(checked_product(seed.mode_amplitude, b0, "GEMHarris mode current"),
                              seed.mode_kx + (p.lengths[0] / p.lengths[1]) * seed.mode_ky, "GEMHarris mode current");
  curl.mode_kx = seed.mode_kx;
  curl.mode_ky = seed.mode_ky;
  if (has_hump(seed)) {
    curl.hump = checked_product(checked_product(b0, seed.amplitude_x, "GEMHarris hump potential"), seed.width_y,
                                "GEMHarris hump potential");
    curl.hump_kx = kPi / (seed.period_x * c.delta);
    curl.hump_ky = kPi / (seed.period_y * c.delta);
    curl.hump_inverse_wx2 = representable(1.0 / (seed.width_x * seed.width_x), "GEMHarris hump x width");
    curl.hump_inverse_wy2 = representable(1.0 / (seed.width_y * seed.width_y), "GEMHarris hump y width");
  }
  constexpr double sqrt_2e = 2.3316439815971242;
  const double x_bound = (2.0 * curl.hump_inverse_wx2 + curl.hump_kx * curl.hump_kx) +
                         4.0 * std::abs(curl.hump_kx) * std::sqrt(curl.hump_inverse_wx2) / sqrt_2e;
  auto sech2 = [](double v) {
    const double ch = std::cosh(v);
    return 1.0 / (ch * ch);
  };
  const double dy = p.lengths[1] / p.cells[1];
  const bool varying = resolved(p, "spatiallyVaryingThermal") == "true";
  for (size_t s = 0; s < p.species.size(); ++s) {
    const auto& sp = p.species[s];
    auto& profile = c.harris_species[s];
    if (sp.drift[2] == 0.0) continue;
    profile.ampere_drift = checked_quotient(checked_product(sp.drift[2], p.c, "GEMHarris Ampere drift"),
                                            checked_product(drift_sum, sp.rho_init, "GEMHarris Ampere drift"),
                                            "GEMHarris Ampere drift coefficient");
    profile.varying_thermal = varying;
    for (int j = 0; j < p.cells[1]; ++j) {
      const double cell_shape = profile.density ? 0.5 * (sech2((j * dy - curl.center_y) / curl.delta) +
                                                         sech2((j * dy + dy - curl.center_y) / curl.delta))
                                                : 1.0;
      for (int row = 0; row < sp.ppc[1]; ++row) {
        const double ym = (j + (row + 0.5) / sp.ppc[1]) * dy - curl.center_y;
        const double sy = ym * curl.hump_inverse_wy2;
        const double cos_y = std::cos(curl.hump_ky * ym);
        const double bound = std::abs(curl.sheet) * sech2(ym / curl.delta) +
                             std::abs(curl.mode * std::cos(curl.mode_ky * ym)) +
                             std::abs(curl.hump) * std::exp(-ym * sy) *
                                 (std::abs(cos_y) * x_bound +
                                  std::abs((4.0 * sy * sy - 2.0 * curl.hump_inverse_wy2 - curl.hump_ky * curl.hump_ky) *
                                               cos_y +
                                           4.0 * sy * curl.hump_ky * std::sin(curl.hump_ky * ym)));
        require_recipe(std::abs(profile.ampere_drift) * bound / cell_shape <= p.c,
                       "GEMHarris currentFromAmpere drift exceeds c where the sheet density vanishes");
      }
    }
  }
}

// HarrisSheetCase (harris_sheet.cpp:183-227).
void harris_sheet(const Parameters& p, CaseRecipe& c) {
  c.harris = true;
  c.harris_double = p.case_name == "DoubleGEM" || p.case_name == "DoubleGEMnoPert";
  c.harris_b0x = number(p, "HarrisB0x");
  c.delta = number(p, "delta");
  require_finite(c.harris_b0x, "HarrisB0x");
  require_recipe(std::isfinite(c.delta) && c.delta > 0.0, "Harris delta must be finite and positive");
  if (p.case_name == "GEMHarris") {
    c.seed = deck_seed(p, c.delta);
  } else if (p.case_name == "GEM" || p.case_name == "DoubleGEM") {
    c.seed = CaseRecipe::Seed{0.0, 0.0, 0.0, 0.4, 0.4, c.delta, c.delta, 10.0, 10.0, 10.0, 10.0};
  }
  validate_harris(p, c);
  bool electron_seen = false, ion_seen = false;
  c.harris_species.assign(p.species.size(), {});
  for (size_t s = 0; s < p.species.size(); ++s) {
    const auto& sp = p.species[s];
    bool density;
    if (p.case_name == "plasmoidChain") {
      bool& seen = sp.q_over_m < 0.0 ? electron_seen : ion_seen;
      density = !seen;
      seen = true;
    } else {
      density = std::abs(sp.drift[0]) > 1.0e-10 || std::abs(sp.drift[1]) > 1.0e-10 || std::abs(sp.drift[2]) > 1.0e-10;
    }
    c.harris_species[s].density = density;
  }
  if (p.case_name == "GEMHarris" && resolved(p, "currentFromAmpere") == "true") admit_ampere_current(p, c);
}

// add_hump (harris_sheet.cpp:108-123).
void add_hump(const CaseRecipe::Seed& seed, double x, double y, double cx, double cy, double delta, double b0x,
              double sign, double b[3]) {
  const double xpert = x - cx;
  const double ypert = y - cy;
  const double envelope =
      std::exp(-(xpert / seed.width_x) * (xpert / seed.width_x) - (ypert / seed.width_y) * (ypert / seed.width_y));
  const double cos_x = std::cos(kPi * xpert / seed.period_x / delta);
  const double sin_x = std::sin(kPi * xpert / seed.period_x / delta);
  const double cos_y = std::cos(kPi * ypert / seed.period_y / delta);
  const double sin_y = std::sin(kPi * ypert / seed.period_y / delta);
  b[0] += sign * b0x * seed.amplitude_x * envelope *
          (-cos_x * cos_y * 2.0 * ypert / seed.width_y - cos_x * sin_y * kPi / seed.slope_period_y);
  b[1] += sign * b0x * seed.amplitude_y * envelope *
          (cos_x * cos_y * 2.0 * xpert / seed.width_x + sin_x * cos_y * kPi / seed.slope_period_x);
}

// HarrisSheetCase::magnetic (harris_sheet.cpp:229-255).
void harris_magnetic(const Parameters& p, const CaseRecipe& c, double x, double y, double b[3]) {
  b[0] = b[1] = b[2] = 0.0;
  const double lx = p.lengths[0], ly = p.lengths[1];
  if (c.harris_double) {
    const double bottom = (y - 0.25 * ly) / c.delta;
    const double top = (y - 0.75 * ly) / c.delta;
    b[0] = c.harris_b0x * (-1.0 + std::tanh(bottom) - std::tanh(top));
  } else {
    b[0] = c.harris_b0x * std::tanh((y - 0.5 * ly) / c.delta);
  }
  if (c.seed.mode_amplitude != 0.0) {
    const double xm = x - 0.5 * lx;
    const double ym = y - 0.5 * ly;
    b[0] += -c.seed.mode_amplitude * c.harris_b0x * (lx / ly) * std::cos(c.seed.mode_kx * xm) *
            std::sin(c.seed.mode_ky * ym);
    b[1] += c.seed.mode_amplitude * c.harris_b0x * std::sin(c.seed.mode_kx * xm) * std::cos(c.seed.mode_ky * ym);
  }
  if (has_hump(c.seed)) {
    if (c.harris_double) {
      add_hump(c.seed, x, y, 0.25 * lx, 0.25 * ly, c.delta, c.harris_b0x, 1.0, b);
      add_hump(c.seed, x, y, 0.75 * lx, 0.75 * ly, c.delta, c.harris_b0x, -1.0, b);
    } else {
      add_hump(c.seed, x, y, 0.5 * lx, 0.5 * ly, c.delta, c.harris_b0x, 1.0, b);
    }
  }
}

}  // namespace

void field_images(const Prepared& r, std::vector<double>& e, std::vector<double>& b) {
  const Parameters& p = r.p;
  const int nx = p.cells[0], ny = p.cells[1], nz = p.cells[2];
  const size_t nodes = size_t(nx + 1) * size_t(ny + 1) * size_t(nz + 1), cells = size_t(nx) * ny * nz;
  e.assign(3 * nodes, 0.0);
  b.assign(3 * cells, 0.0);
  const CaseRecipe& c = r.recipe;
  const bool two_stream = p.case_name == "TwoStream", weibel = p.case_name == "Weibel";
  auto check = [](double v) {
    if (!std::isfinite(v)) throw std::overflow_error("TPIC field profile: sampled field is not representable");
    return v;
  };
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        const size_t n = size_t(i) + size_t(nx + 1) * (size_t(j) + size_t(ny + 1) * size_t(k));
        const double x = (i + 0.0) * r.spacing[0];
        if (two_stream && c.e_seed) e[n] = check(c.e_amplitude * std::sin(c.k * x + c.phase));
        if (weibel && c.e_seed) e[nodes + n] = check(c.e_amplitude * std::sin(c.k * x));
      }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t n = size_t(i) + size_t(nx) * (size_t(j) + size_t(ny) * size_t(k));
        const double x = (i + 0.5) * r.spacing[0];
        double seeded[3] = {0.0, 0.0, 0.0};
        if (weibel && c.b_seed) seeded[2] = c.b_amplitude * std::cos(c.k * x);
        if (c.harris) harris_magnetic(p, c, x, (j + 0.5) * r.spacing[1], seeded);
        b[n] = check(seeded[0] + p.b0[0]);
        b[cells + n] = check(seeded[1] + p.b0[1]);
        b[2 * cells + n] = check(seeded[2] + p.b0[2]);
      }
}

Prepared prepare(const Parameters& p, const Options& o) {
  Prepared r;
  r.p = p;
  r.options = o;
  const char* axes[3] = {"x", "y", "z"};
  const char* faces[6] = {"Xright", "Xleft", "Yright", "Yleft", "Zright", "Zleft"};

  // Domain (src/config/domain.cpp:54-83,133-150).
  const int partitions = p.processes[0] * p.processes[1] * p.processes[2];
  for (int a = 0; a < 3; ++a) {
    if (p.processes[a] <= 0) invalid("domain", std::string(axes[a]) + " process extent must be positive");
    if (p.cells[a] <= 0) invalid("domain", std::string(axes[a]) + " global cell count must be positive");
    if (!std::isfinite(p.lengths[a]) || p.lengths[a] <= 0.0)
      invalid("domain", std::string(axes[a]) + " physical length must be finite and positive");
    const int minimum = p.processes[a] == 1 && p.periodic[a] ? 1 : 2;
    if (p.cells[a] / p.processes[a] < minimum)
      invalid("domain", std::string(axes[a]) + " requires at least " + std::to_string(minimum) + " cells per rank");
    r.spacing[a] = p.lengths[a] / p.cells[a];
    if (!(std::isfinite(r.spacing[a]) && r.spacing[a] > 0 && std::isfinite(1.0 / r.spacing[a])))
      invalid("domain", "cell spacing is not representable");
  }
  if (o.rank < 0 || o.rank >= partitions)
    invalid("domain", "rank must be in [0, communicator_size) with a positive size");
  r.cells = uint64_t(p.cells[0]) * uint64_t(p.cells[1]) * uint64_t(p.cells[2]);

  // Physics (src/config/physics.cpp:26-119).
  auto positive = [](double v, const std::string& name) {
    if (!(std::isfinite(v) && v > 0)) invalid("physics", name + " must be finite and positive");
  };
  auto finite = [](double v, const std::string& name) {
    if (!std::isfinite(v)) invalid("physics", name + " must be finite");
  };
  positive(p.dt, "dt");
  positive(p.c, "c");
  if (p.ncycles < 0) invalid("physics", "ncycles must be nonnegative");
  if (!(std::isfinite(p.th) && p.th >= 0.5 && p.th <= 1))
    invalid("physics", "th must be in [0.5, 1] for the admitted linear-ECSIM scheme");
  positive(p.gmres_tol, "GMREStol");
  if (!(p.gmres_tol < 1)) invalid("physics", "GMREStol must be less than 1");
  positive(p.th * p.dt, "th*dt");
  positive(p.c * (p.th * p.dt), "c*th*dt");
  positive(1 / (p.c * (p.th * p.dt)), "1/(c*th*dt)");
  finite(p.dt * p.ncycles, "dt*ncycles");
  finite(p.b0[0], "B0x");
  finite(p.b0[1], "B0y");
  finite(p.b0[2], "B0z");
  for (size_t s = 0; s < p.species.size(); ++s) {
    const Species& sp = p.species[s];
    const std::string prefix = "species " + std::to_string(s) + ": ";
    finite(sp.q_over_m, prefix + "qom");
    if (sp.q_over_m == 0) invalid("physics", prefix + "qom=0 has no defined neutral-species semantics");
    const std::pair<const char*, double> rates[4] = {
        {"rhoINIT", sp.rho_init}, {"uth", sp.thermal[0]}, {"vth", sp.thermal[1]}, {"wth", sp.thermal[2]}};
    for (const auto& [name, v] : rates)
      if (!(std::isfinite(v) && v >= 0)) invalid("physics", prefix + name + " must be finite and nonnegative");
    finite(sp.drift[0], prefix + "u0");
    finite(sp.drift[1], prefix + "v0");
    finite(sp.drift[2], prefix + "w0");
    const double beta = double(static_cast<long double>(sp.q_over_m) * p.dt / (2.0L * p.c));
    if (!(std::isfinite(beta) && beta != 0)) invalid("physics", prefix + "qom*dt/(2*c) is not representable");
  }
  bool smoothing = false;
  if (p.smoothing == "energy_conserving") {
    if (!(std::isfinite(p.smooth) && p.smooth >= 0 && p.smooth <= 1))
      invalid("physics", "energy-conserving smoothing requires 0 <= Smooth <= 1");
    if (p.smooth_niter < 0) invalid("physics", "energy-conserving smoothing requires SmoothNiter >= 0");
    smoothing = p.smooth_niter > 0 && p.smooth < 1;
  }
  r.smoothing = smoothing;
  bool open_boundary = false;
  for (int f = 0; f < 6; ++f) {
    open_boundary |= !p.periodic[f / 2] && p.em_face[f] == 2;
    if (p.em_face[f] < 0 || p.em_face[f] >= 3)
      invalid("physics", "bcEMface" + std::string(faces[f]) + "=" + std::to_string(p.em_face[f]) +
                             ": unknown code (expected 0 perfect electric conductor, 1 perfect magnetic "
                             "conductor or 2 open)");
    if (!p.particle_periodic[f / 2] && p.particle_face[f] != 1) {
      const std::string request = "bcPface" + std::string(faces[f]) + "=" + std::to_string(p.particle_face[f]);
      if (p.particle_face[f] == 3 || p.particle_face[f] == 4)
        r.unsupported.push_back(request + ": boundary particle injection is not implemented; an effective open "
                                          "particle face requires injection");
      else
        r.unsupported.push_back(request + ": only the mirror (1) is admitted on a physical particle face");
    }
  }
  if (smoothing && open_boundary)
    invalid("physics", "active energy-conserving smoothing cannot use an effective open EM face");
  if (p.c != 1)
    r.unsupported.push_back("c != 1 is not supported: the mass-matrix and mover normalization is qualified for c=1 only");
  if (p.ns_test != 0)
    r.unsupported.push_back("nsTestPart: test-particle initialization and evolution are not implemented");
  if (p.test_particles_cycle > 0)
    r.unsupported.push_back("TestParticlesOutputCycle: test-particle output is not implemented");
  r.notes.push_back("Linear ECSIM uses GMRES restart=30 and maximum iterations=100");

  // What this build runs: one partition on one GPU, periodic fields or
  // PEC/PMC walls with or without energy-conserving smoothing, the Uniform,
  // TwoStream, Weibel and Harris-sheet cases, FP64 particles and fields.
  const bool harris_case = p.case_name == "GEMHarris" || p.case_name == "GEM" || p.case_name == "GEMnoPert" ||
                           p.case_name == "DoubleGEM" || p.case_name == "DoubleGEMnoPert" ||
                           p.case_name == "plasmoidChain";
  if (p.case_name != "Uniform" && p.case_name != "TwoStream" && p.case_name != "Weibel" && !harris_case)
    r.unsupported.push_back("Case " + p.case_name +
                            ": this build initializes the Uniform, TwoStream, Weibel and Harris-sheet cases");
  if (partitions != 1)
    r.unsupported.push_back("XLEN*YLEN*ZLEN = " + std::to_string(partitions) +
                            ": this build runs one partition on one GPU");
  for (int f = 0; f < 6; ++f)
    if (!p.periodic[f / 2] && p.em_face[f] == 2)
      r.unsupported.push_back("bcEMface" + std::string(faces[f]) +
                              "=2: open electromagnetic boundaries are not implemented");
  if (p.ns > 8) r.unsupported.push_back("ns = " + std::to_string(p.ns) + ": this build holds at most 8 species");
  if (o.storage_precision != "double" || o.compute_precision != "double")
    r.unsupported.push_back("Precision pair " + o.storage_precision + "/" + o.compute_precision +
                            " is not built: particles and fields are stored in FP64 "
                            "(--mass-precision selects the moment-product operands)");
  if (o.order == "unsorted")
    r.unsupported.push_back("--order unsorted: deposition runs from cell-sorted particles (full or partial)");
  if (o.restart)
    r.unsupported.push_back("--restart: this build has no checkpoint backend (ADIOS2 or HDF5); there is no "
                            "checkpoint to continue from");
  if (o.mass_matrix == "scalar")
    r.notes.push_back("Mass matrix kernel scalar is served by the FP64 matrix-core deposition (same sums, MFMA order)");
  if (o.mass_apply == "split") r.notes.push_back("Mass apply split is served by the full-domain apply");
  if (o.order == "partial")
    r.notes.push_back("Particle order partial is served by a full stable cell sort every cycle");
  // Output (src/config/output.cpp:151-256; run_plan.cpp:124-135).
  {
    OutputPlan& out = r.output;
    const std::string save_dir = resolved(p, "SaveDirName");
    out.save_dir = save_dir.empty() ? "data" : save_dir;
    const std::string field_tag = resolved(p, "FieldOutputTag"), moments_tag = resolved(p, "MomentsOutputTag");
    const bool default_tags = field_tag.empty() && moments_tag.empty();
    out.output_b = out.output_e = default_tags;
    bool gauss_tag = false;
    for (size_t at = 0; at <= field_tag.size() && !field_tag.empty();) {
      const size_t plus = field_tag.find('+', at);
      const std::string token = field_tag.substr(at, plus == std::string::npos ? std::string::npos : plus - at);
      if (token == "B") out.output_b = true;
      else if (token == "E") out.output_e = true;
      else if (token == "Gauss") gauss_tag = true;
      else invalid("output", "FieldOutputTag = '" + field_tag + "': '" + token + "' is unknown (expected B, E or Gauss)");
      if (plus == std::string::npos) break;
      at = plus + 1;
    }
    std::string engine = resolved(p, "FieldWriteMethod");
    if (engine.empty()) engine = resolved(p, "WriteMethod");
    const bool vtk = engine.empty() || engine == "default" || engine == "pvtk" || engine == "vtk";
    out.field_cycle = p.field_cycle;
    out.diagnostics_cycle = p.diagnostics_cycle;
    out.diagnostics = p.diagnostics_cycle > 0;
    out.fields = p.field_cycle > 0 && (out.output_b || out.output_e);
    if (out.fields && !vtk)
      r.unsupported.push_back("FieldWriteMethod " + engine +
                              ": this build writes fields with pvtk only (no ADIOS2 or HDF5 backend)");
    if (p.field_cycle > 0 && (default_tags || !moments_tag.empty()))
      r.warnings.push_back("Moment output (rho, J, pressure) is not built: field snapshots hold E and B only");
    // Particle snapshots every ParticlesOutputCycle cycles and at the end
    // (PreparedOutput::should_write_particles, src/config/output.cpp:135-162).
    std::string particle_engine = resolved(p, "ParticlesWriteMethod");
    if (particle_engine.empty()) particle_engine = resolved(p, "WriteMethod");
    const bool particle_vtk = particle_engine.empty() || particle_engine == "default" || particle_engine == "pvtk" ||
                              particle_engine == "vtk";
    out.particles_cycle = p.particles_cycle;
    out.particles = p.particles_cycle > 0;
    if (out.particles && !particle_vtk)
      r.unsupported.push_back("ParticlesWriteMethod " + particle_engine +
                              ": this build writes particles with pvtk only (no ADIOS2 or HDF5 backend)");
    if (p.gauss_cycle > 0 || (gauss_tag && p.field_cycle > 0))
      r.warnings.push_back("DiagnosticGaussCycle: Gauss-law residual samples are not built; none will be written");
    if (p.restart_cycle > 0)
      r.warnings.push_back("RestartOutputCycle: this build has no checkpoint backend (TPIC_ENABLE_ADIOS2 or "
                           "TPIC_ENABLE_HDF5); no checkpoints will be written");
  }
  if (!p.unknown_keys.empty()) {
    std::string ignored = "Unknown input keys ignored:";
    for (size_t i = 0; i < p.unknown_keys.size(); ++i) ignored += (i == 0 ? " " : ", ") + p.unknown_keys[i];
    r.warnings.push_back(ignored);
  }

  r.recipe.density_seeded.assign(p.species.size(), false);
  if (p.case_name == "TwoStream") two_stream(p, r.recipe);
  if (p.case_name == "Weibel") weibel(p, r.recipe);
  if (harris_case) {
    harris_sheet(p, r.recipe);
    // Centring removes the mean of v - drift with a drift the same for every
    // particle; the Harris profile's z drift varies with position when it is
    // a double sheet's or the Ampere current's.
    for (size_t s = 0; s < p.species.size(); ++s)
      if (p.species[s].center_initial_velocity &&
          (r.recipe.harris_double || r.recipe.harris_species[s].ampere_drift != 0.0))
        r.unsupported.push_back("CenterInitialVelocity of species " + std::to_string(s) +
                                " with a position-dependent Harris drift is not built");
  }

  // Particle plan and load normalization (src/config/particle_plan.cpp:127-150,
  // src/initialization/prepared_initialization.cpp:37-66).
  uint64_t ids = 0;
  for (size_t s = 0; s < p.species.size(); ++s) {
    const Species& sp = p.species[s];
    SpeciesPlan plan;
    if (sp.ppc[0] <= 0 || sp.ppc[1] <= 0 || sp.ppc[2] <= 0)
      invalid("particles", "PPC factors must be positive and product representable");
    plan.per_cell = uint64_t(sp.ppc[0]) * uint64_t(sp.ppc[1]) * uint64_t(sp.ppc[2]);
    if (plan.per_cell > uint64_t(std::numeric_limits<int>::max()))
      invalid("particles", "PPC product exceeds signed int");
    plan.count = r.cells * plan.per_cell;
    if (plan.count > uint64_t(std::numeric_limits<int>::max()))
      invalid("particles", "local particle count exceeds signed int");
    plan.first_id = ids;
    ids += plan.count;
    const double volume = (r.spacing[0] * r.spacing[1]) * r.spacing[2];
    const double q_factor = (sp.q_over_m > 0.0 ? 1.0 : -1.0) * volume / double(plan.per_cell);
    const double density = sp.rho_init / 12.566370614359172953850573533118;
    if (!std::isfinite(q_factor) || q_factor == 0.0 || !std::isfinite(density) ||
        (sp.rho_init != 0.0 && density == 0.0) || !std::isfinite(q_factor * density) ||
        (density != 0.0 && q_factor * density == 0.0))
      throw std::overflow_error("TPIC particle profile: macrocharge normalization is not representable");
    r.plans.push_back(plan);
  }
  return r;
}

}  // namespace tpic_mi::deck
