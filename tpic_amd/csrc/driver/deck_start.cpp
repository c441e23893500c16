// Startup of an admitted deck through the C-ABI: the context, every species'
// load, the case's field images and the initialization (or a restart file in
// place of the load and the images). Shared by the tpic_mi_run driver and the
// library's tpic_mi_load_deck.
#include "deck_start.hpp"

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace tpic_mi::deck {
namespace {

class Starter {
 public:
  explicit Starter(const Prepared& r) : r_(r) {}
  ~Starter() {
    if (ctx_) tpic_mi_destroy(ctx_);
  }

  tpic_mi_context* run(void* stream, bool timing, const char* restart_path) {
    create(stream);
    if (timing) check(tpic_mi_enable_timing(ctx_, 1));
    if (restart_path) {
      // The restart file holds the particles, E and B of its cycle; the
      // initialization sorts, deposits and assembles them.
      check(tpic_mi_checkpoint_read(ctx_, restart_path, nullptr));
    } else {
      load();
      std::vector<double> e, b;
      field_images(r_, e, b);
      check(tpic_mi_set_field(ctx_, TPIC_MI_FIELD_E, e.data()));
      check(tpic_mi_set_field(ctx_, TPIC_MI_FIELD_B, b.data()));
    }
    check(tpic_mi_initialize(ctx_));
    check(tpic_mi_synchronize(ctx_));
    unsigned flags = 0;
    check(tpic_mi_read_validation(ctx_, &flags));
    if (flags) {
      char hex[16];
      std::snprintf(hex, sizeof hex, "%x", flags);
      throw StartupFailure(std::string("TPIC startup: device validation flags 0x") + hex, TPIC_MI_INVALID_STATE);
    }
    tpic_mi_context* out = ctx_;
    ctx_ = nullptr;
    return out;
  }

 private:
  void check(int rc) {
    if (rc)
      throw StartupFailure(std::string("TPIC: ") + (ctx_ ? tpic_mi_last_error(ctx_) : "library call failed"), rc);
  }

  // The grid, numerics and species of the deck (the Xright..Zleft bcEMface
  // codes 0 PEC, 1 PMC become the C-ABI's [axis][lower, upper] faces).
  void create(void* stream) {
    const auto& p = r_.p;
    tpic_mi_config cfg{};
    for (int a = 0; a < 3; ++a) {
      cfg.cells[a] = p.cells[a];
      cfg.lengths[a] = p.lengths[a];
      cfg.particle_periodic[a] = p.particle_periodic[a] ? 1 : 0;
    }
    cfg.dt = p.dt;
    cfg.theta = p.th;
    cfg.c = p.c;
    cfg.gmres_tolerance = p.gmres_tol;
    cfg.gmres_restart = 30;
    cfg.gmres_max_iterations = 100;
    cfg.mass_kernel = r_.options.mass_operands;
    cfg.species = p.ns;
    for (int s = 0; s < p.ns; ++s) {
      cfg.species_list[s].q_over_m = p.species[size_t(s)].q_over_m;
      cfg.species_list[s].count = uint32_t(r_.plans[size_t(s)].count);
    }
    cfg.stream = stream;
    cfg.smooth_operator = p.smoothing_operator == "binomial27" ? TPIC_MI_SMOOTH_BINOMIAL27 : TPIC_MI_SMOOTH_FACE7;
    cfg.smooth_iterations = r_.smoothing ? p.smooth_niter : 0;
    cfg.smooth_alpha = r_.smoothing ? p.smooth : 0.0;
    for (int a = 0; a < 3; ++a)
      for (int side = 0; side < 2; ++side) {
        const int code = p.em_face[2 * a + (side == 0 ? 1 : 0)];
        cfg.field_faces[a][side] =
            p.periodic[a] ? TPIC_MI_FACE_PERIODIC : (code == 0 ? TPIC_MI_FACE_PEC : TPIC_MI_FACE_PMC);
      }
    if (const int rc = tpic_mi_create(&cfg, &ctx_)) {
      ctx_ = nullptr;
      throw StartupFailure(std::string("TPIC startup: ") + tpic_mi_last_error(nullptr), rc);
    }
  }

  // Every species' load of the case (prepared_initialization.cpp:37-85):
  // Harris profiles, the TwoStream density seed or the uniform load, then
  // CenterInitialVelocity.
  void load() {
    const auto& p = r_.p;
    for (int s = 0; s < p.ns; ++s) {
      const auto& sp = p.species[size_t(s)];
      tpic_mi_uniform_load load{};
      for (int a = 0; a < 3; ++a) {
        load.ppc[a] = sp.ppc[a];
        load.drift[a] = sp.drift[a];
        load.thermal[a] = sp.thermal[a];
      }
      load.rho_init = sp.rho_init;
      load.quiet_start = p.quiet_start ? 1 : 0;
      load.rank = r_.options.rank;
      load.max_species = 10;
      load.first_id = r_.plans[size_t(s)].first_id;
      if (r_.recipe.harris) {
        const auto& c = r_.recipe;
        const auto& hs = c.harris_species[size_t(s)];
        tpic_mi_harris_profile prof{};
        prof.harris_density = hs.density ? 1 : 0;
        prof.is_double = c.harris_double ? 1 : 0;
        prof.varying_thermal = hs.varying_thermal ? 1 : 0;
        prof.ly = p.lengths[1];
        prof.delta = c.delta;
        prof.ampere_drift = hs.ampere_drift;
        prof.curl_center[0] = c.curl.center_x;
        prof.curl_center[1] = c.curl.center_y;
        prof.curl_delta = c.curl.delta;
        prof.curl_sheet = c.curl.sheet;
        prof.curl_mode = c.curl.mode;
        prof.curl_mode_k[0] = c.curl.mode_kx;
        prof.curl_mode_k[1] = c.curl.mode_ky;
        prof.curl_hump = c.curl.hump;
        prof.curl_hump_k[0] = c.curl.hump_kx;
        prof.curl_hump_k[1] = c.curl.hump_ky;
        prof.curl_hump_inverse_w2[0] = c.curl.hump_inverse_wx2;
        prof.curl_hump_inverse_w2[1] = c.curl.hump_inverse_wy2;
        check(tpic_mi_load_harris(ctx_, s, &load, &prof));
      } else if (r_.recipe.density_seeded[size_t(s)]) {
        const tpic_mi_density_seed mode{r_.recipe.k, r_.recipe.phase, r_.recipe.density_amplitude};
        check(tpic_mi_load_density_seeded(ctx_, s, &load, &mode));
      } else {
        check(tpic_mi_load_uniform(ctx_, s, &load));
      }
      if (sp.center_initial_velocity) check(tpic_mi_center_velocity(ctx_, s, sp.drift));
    }
  }

  const Prepared& r_;
  tpic_mi_context* ctx_ = nullptr;
};

}  // namespace

tpic_mi_context* start(const Prepared& r, void* stream, bool timing, const char* restart_path) {
  return Starter(r).run(stream, timing, restart_path);
}

}  // namespace tpic_mi::deck
