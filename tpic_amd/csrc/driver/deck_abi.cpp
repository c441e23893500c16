// tpic_mi_load_deck: the deck front end of the tpic_mi_run driver (parser,
// admission, startup) behind the C-ABI, linked into libtpic_mi.so.
#include <cstring>
#include <exception>
#include <string>

#include "../../../include/tpic_mi.h"
#include "deck.hpp"
#include "deck_start.hpp"

namespace tpic_mi {
// The error text tpic_mi_last_error(NULL) returns (context.hip).
void set_context_free_error(const std::string& message);
}  // namespace tpic_mi

extern "C" int tpic_mi_load_deck(const char* path, const tpic_mi_deck_options* options, void* stream,
                                 tpic_mi_context** ctx, tpic_mi_deck_run* run) {
  using namespace tpic_mi::deck;
  if (!ctx) {
    tpic_mi::set_context_free_error("tpic_mi_load_deck: null context pointer");
    return TPIC_MI_INVALID_ARGUMENT;
  }
  *ctx = nullptr;
  if (!path) {
    tpic_mi::set_context_free_error("tpic_mi_load_deck: null path");
    return TPIC_MI_INVALID_ARGUMENT;
  }
  Options o;
  if (options) {
    for (int32_t r : options->reserved)
      if (r != 0) {
        tpic_mi::set_context_free_error("tpic_mi_load_deck: reserved options must be zero");
        return TPIC_MI_INVALID_ARGUMENT;
      }
    if (options->mass_kernel < TPIC_MI_MASS_FP64 || options->mass_kernel > TPIC_MI_MASS_BF16X3) {
      tpic_mi::set_context_free_error("tpic_mi_load_deck: mass_kernel must be a TPIC_MI_MASS_* value");
      return TPIC_MI_INVALID_ARGUMENT;
    }
    o.mass_operands = options->mass_kernel;
  }
  try {
    const Prepared prepared = prepare(load(path), o);
    if (!prepared.admitted()) {
      std::string message = "TPIC run plan: unsupported requests";
      for (const auto& u : prepared.unsupported) message += "\n" + u;
      tpic_mi::set_context_free_error(message);
      return TPIC_MI_UNSUPPORTED;
    }
    tpic_mi_context* started = start(prepared, stream, false, nullptr);
    if (run) {
      const auto& p = prepared.p;
      std::memset(run, 0, sizeof *run);
      run->ncycles = p.ncycles;
      run->species = p.ns;
      run->energy_cycle = p.energy_cycle;
      for (int a = 0; a < 3; ++a) {
        run->cells[a] = p.cells[a];
        run->lengths[a] = p.lengths[a];
      }
      run->dt = p.dt;
      run->theta = p.th;
      run->c = p.c;
      for (int s = 0; s < p.ns; ++s) {
        run->count[s] = prepared.plans[size_t(s)].count;
        run->q_over_m[s] = p.species[size_t(s)].q_over_m;
      }
    }
    *ctx = started;
    return TPIC_MI_OK;
  } catch (const Invalid& e) {
    tpic_mi::set_context_free_error(e.what());
    return TPIC_MI_INVALID_ARGUMENT;
  } catch (const StartupFailure& e) {
    tpic_mi::set_context_free_error(e.what());
    return e.status;
  } catch (const std::exception& e) {
    tpic_mi::set_context_free_error(std::string("tpic_mi_load_deck: ") + e.what());
    return TPIC_MI_INVALID_ARGUMENT;
  }
}
