// Startup of an admitted deck through the C-ABI of include/tpic_mi.h.
#pragma once

#include <stdexcept>
#include <string>

#include "../../../include/tpic_mi.h"
#include "deck.hpp"

namespace tpic_mi::deck {

// A library failure during startup: the library's status and message.
struct StartupFailure : std::runtime_error {
  StartupFailure(const std::string& message, int status) : std::runtime_error(message), status(status) {}
  int status;
};

// Creates the context of an admitted deck on stream (NULL: the null stream),
// loads every species with the case's recipe, sets the case's E and B images
// (or reads restart_path in place of both) and initializes it: the state a
// run's first cycle starts from. Throws StartupFailure; nothing leaks.
tpic_mi_context* start(const Prepared& r, void* stream, bool timing, const char* restart_path);

}  // namespace tpic_mi::deck
