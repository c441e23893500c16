// The context behind the C-ABI: device owners sized once at creation and the
// host-side launchers of every kernel. Host code is C++20; every launcher
// enqueues on the context stream and never allocates.
#pragma once

#include "common.hpp"

#include <algorithm>
#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace tpic_mi {

enum class Phase : int { sort = 0, move, deposit, assemble, solve, advance, spmv, count };

struct Numerics {
  double dt, theta, c, gmres_tol;
  int restart, max_iterations;
  int mass_kernel;  // 0 = FP64
  // Energy-conserving smoothing: iterations passes of the Face7 (alpha, beta)
  // or Binomial27 stencil; active when iterations > 0 (smooth_alpha < 1).
  bool smoothing;
  int smooth_operator, smooth_iterations;
  double smooth_alpha, smooth_beta;
};

struct SpeciesState {
  double q_over_m = 0;
  uint32_t count = 0;
  ParticleView slot[2]{};
  int current = 0;
  uint32_t* counts = nullptr;   // [bins]
  uint32_t* offsets = nullptr;  // [bins + 1]
  // Source position of every destination of the species' last sort plan
  // (launch_sort_plan), read by the fused gather and deposition.
  uint32_t* source = nullptr;   // [count]
  // Crosser masks of the last block move that recorded them: bit l of word b
  // set when the move changed the key of particle 64 b + l.
  uint64_t* cross = nullptr;    // [ceil(count / 64)]
  bool sorted = false;
  ParticleView view() const { return slot[current]; }
  ParticleView scratch() const { return slot[current ^ 1]; }
};

// Workspace of the stable run-based counting sort; sized for the largest
// species at creation.
struct SortWorkspace {
  uint32_t capacity = 0;      // particles
  uint32_t tiles = 0;         // capacity / tile
  uint32_t* tile_runs = nullptr;   // [tiles + 1] run counts, then exclusive base
  uint32_t* wave_base = nullptr;   // [tiles * 4] heads before each wave span of a tile
  uint32_t* run_start = nullptr;   // [capacity + 1]
  int32_t* run_key = nullptr;      // [capacity]
  uint32_t* run_dest = nullptr;    // [capacity]
  uint32_t* bin_runs = nullptr;    // [capacity] run ids grouped by bin
  uint32_t* rcounts = nullptr;     // [bins] runs per bin
  uint32_t* roffsets = nullptr;    // [bins + 1]
  uint32_t* cursor = nullptr;      // [bins]
  uint32_t* large = nullptr;       // [bins + 1] bins with many runs; [0] = count
  uint32_t* scan_scratch = nullptr;
  size_t scan_scratch_words = 0;
  uint32_t* failure = nullptr;     // [1] device failure word
  uint32_t* permutation = nullptr; // [capacity] destination of each source (last sort)
  // The incremental plan (launch_sort_plan_incremental) also uses the run
  // arrays above for its crosser list and arrivals.
  uint32_t* block_base = nullptr;  // [ceil(capacity / 64) + 1] crossers before each 64-particle block
  uint32_t* bin_base = nullptr;    // [bins] stayer destination base of each bin
  uint32_t* bin_lower = nullptr;   // [bins] arrivals from lower bins
};

struct GmresWorkspace {
  unsigned count = 0;       // 3 * nodes
  unsigned blocks = 0;      // reduction blocks
  double* basis = nullptr;  // [(restart + 1) * count]
  double* x = nullptr;      // solution being updated
  double* b = nullptr;      // right-hand side
  double* r = nullptr;      // residual
  double* image = nullptr;  // operator image
  double* values = nullptr;    // [max(restart, 2)] reduced batch
  double* partials = nullptr;  // [max(restart, 2) * blocks]
  double* state = nullptr;     // Hessenberg, cos, sin, rhs, coefficients, status
  double* controls = nullptr;  // pinned host [8]
  double* curl_cells = nullptr;   // [3 * cells] scratch of the operator
};

struct Context {
  Grid grid{};
  Numerics num{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int device = 0;
  int cus = 0;
  std::vector<SpeciesState> species;
  double inverse_volume = 0, delta = 0, current_factor = 0;
  // Fields
  double* e = nullptr;       // E^n at nodes
  double* b = nullptr;       // B^n at cells
  double* bn = nullptr;      // B at nodes (old epoch for the mover, next for deposition)
  double* current = nullptr; // J hat at nodes
  double* etheta = nullptr;  // E_theta at nodes (GMRES solution)
  double* rhs = nullptr;     // solver right-hand side
  // Smoothing only: the two pass buffers and the movers' field S(E_theta).
  double* smooth_a = nullptr;
  double* smooth_b = nullptr;
  double* eparticle = nullptr;
  // Walls only: the projected input of tpic_mi_apply_operator, and (without
  // smoothing) the mass response between the two stages of the walled apply.
  double* projected = nullptr;
  double* mass_response = nullptr;
  // The node field the movers gather E from.
  const double* particle_field() const { return num.smoothing ? eparticle : etheta; }
  void* mass = nullptr;      // entry-major mass matrix: double, or float when mass_single
  bool mass_single = false;  // mixed forms store the assembled matrix in FP32
  // Per-cell deposition records, in the element type of the cell sums: double
  // for the FP64 form, float for the mixed forms (whose per-cell sums are FP32).
  void* cellbuf = nullptr;       // [cells][kCellRecord] mass moments (m * 9 + entry)
  double* corner_current = nullptr; // [cells][kCornerRecord]: corner currents (a * 3 + comp) of every cell
  uint32_t* bin_counts = nullptr;   // [species][bins]: SpeciesState::counts point here
  uint32_t* bin_offsets = nullptr;  // [species][bins + 1]: SpeciesState::offsets point here
  double* reduce = nullptr;  // diagnostics partials
  double* reduce_out = nullptr;
  unsigned* validation = nullptr;  // device word
  double* moment_cells = nullptr;  // [cells][80] output-moment corner sums, allocated at first use
  // Per species: nonzero when its incremental plan must be replaced by the
  // run-based one (device), and its pinned host copy.
  uint32_t* plan_status = nullptr;  // [kMaxSpecies]
  uint32_t* host_status = nullptr;  // pinned [kMaxSpecies]
  unsigned* host_word = nullptr;   // pinned
  double* host_scalars = nullptr;  // pinned
  SortWorkspace sort;
  // With two or more species, a second plan workspace and stream: the first
  // species' sort plan runs there while the others move (fork and join are
  // events on the two streams).
  SortWorkspace sort_side;
  hipStream_t side_stream = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  GmresWorkspace gmres;
  // Timing (hipEvents around phases when enabled)
  bool timing = false;
  hipEvent_t ev_begin[int(Phase::count)]{}, ev_end[int(Phase::count)]{};
  double phase_ms[int(Phase::count)]{};
  uint64_t phase_calls[int(Phase::count)]{};
  bool phase_open[int(Phase::count)]{};
  std::string error;
  void* owner = nullptr;  // the tpic_mi_context holding this context
  // The operator applies A^T (the adjoint solve; periodic fields only).
  bool transposed_operator = false;
  // In-device cycle checkpoints: the state a cycle starts from.
  struct Checkpoint {
    bool saved = false;
    std::vector<ParticleView> particles;  // one per species, the sorted order
    std::vector<bool> sorted;
    uint32_t* counts = nullptr;   // [species][bins]
    uint32_t* offsets = nullptr;  // [species][bins + 1]
    double* e = nullptr;          // E at nodes
    double* b = nullptr;          // B at cells
  };
  std::vector<Checkpoint> checkpoints;
};

// Returns 0 or a negative TPIC_MI_* status; sets ctx.error on failure.
int fail(Context& ctx, int status, const std::string& message);
// The error text of a call that has no context (tpic_mi_last_error(NULL)).
void set_context_free_error(const std::string& message);
int check_hip(Context& ctx, hipError_t e, const char* what);

// Phase timing: begin records an event before the phase's first launch and
// end after its last; end accumulates the previous measurement lazily.
void phase_begin(Context& ctx, Phase p);
void phase_end(Context& ctx, Phase p);

// ---- launchers (each returns 0 or a negative status) --------------------
int launch_sort(Context& ctx, int species);
// The plan of a species' sort without the copy: counts, offsets and the
// source of every destination (identity after a failed plan, whose particles
// are then copied unchanged into the scratch slot).
int launch_sort_plan(Context& ctx, int species);
// The same on stream `stream` with workspace `ws` (untimed off the context
// stream).
int launch_sort_plan(Context& ctx, int species, SortWorkspace& ws, hipStream_t stream);
// The plan of a species sorted before a block move that recorded its
// crossers, built from the crossers alone (the same counts, offsets and
// sources as the run-based plan). It sets ctx.plan_status[species] nonzero
// when the run-based plan must be run instead (a bin with many arrivals, or
// a previous plan that failed).
int launch_sort_plan_incremental(Context& ctx, int species, SortWorkspace& ws, hipStream_t stream);
// With record_crossers (a sorted species) the block mover also writes the
// species' crosser masks for launch_sort_plan_incremental.
int launch_move(Context& ctx, int species, bool record_crossers = false);
int launch_deposit(Context& ctx, int first_species, int last_species, bool accumulate);
// Every species' planned sort fused with the deposition: each destination
// record is gathered from its source, stored in the scratch slot and
// deposited in one pass; the slots then swap (the species are sorted).
int launch_sort_deposit(Context& ctx);
// The fused pass alone, after every species' plan.
int launch_gather_deposit(Context& ctx);
int launch_assemble(Context& ctx);
int launch_cell_to_node(Context& ctx);
int launch_zero_constrained(Context& ctx, double* v);  // no-op without walls
int launch_rhs(Context& ctx);
// stop: a device word that, once nonzero, makes the launch do nothing (the
// GMRES stop word; ignored with smoothing).
int launch_operator(Context& ctx, const double* in, double* out, const double* stop = nullptr);
// The configured smoothing S(src) of a node vector (ctx.num.smoothing must be
// set; constrained rows of src read as zero, those of every pass zeroed): the
// passes alternate between smooth_a and smooth_b, starting with the one src
// is not, and the last writes into dst when dst is given. *result: where the
// result is.
int launch_smooth(Context& ctx, const double* src, double* dst, const double** result);
int launch_next_fields(Context& ctx);
int launch_energy(Context& ctx, double* out);  // out: W_E, W_B, K per species
int launch_scalars(Context& ctx, double* out);
int launch_moments(Context& ctx, int species, unsigned selected, double* out);  // out: 10 x nodes, device  // out: W_E, W_B, then K, charge, momentum x y z per species
int run_gmres(Context& ctx, int* converged, int* iterations, double* initial_residual,
              double* residual, double* rhs_norm, double* relative);
// GMRES on A x = rhs (A^T with ctx.transposed_operator): x holds the initial
// guess and, when converged, the solution; neither may be a solver buffer.
int run_gmres_system(Context& ctx, const double* rhs, double* x, int* converged, int* iterations,
                     double* initial_residual, double* residual, double* rhs_norm, double* relative);
// curl_C2N of a cell vector at every node (periodic boxes) and curl_N2C of a
// node vector at every cell.
int launch_curl_c2n(Context& ctx, const double* cells_in, double* nodes_out);
int launch_curl_n2c(Context& ctx, const double* nodes_in, double* cells_out);
// The cycle checkpoints of tpic_mi_checkpoint_* (freed with the context).
void release_checkpoints(Context& ctx);
int launch_load(Context& ctx, int species, const int ppc[3], double rho, const double drift[3],
                const double thermal[3], uint64_t seed, uint64_t first_id);
int launch_center_velocity(Context& ctx, int species, const double drift[3]);
// The Harris profile's parameters (HarrisProfile and HarrisCurl,
// include/tpic/initialization/cases/harris_sheet.hpp:33-120).
struct HarrisLoad {
  int density, is_double, varying_thermal;
  double ly, delta, ampere_drift;
  double center_x, center_y, curl_delta, sheet, mode, mode_kx, mode_ky;
  double hump, hump_kx, hump_ky, hump_inverse_wx2, hump_inverse_wy2;
};
int launch_load_uniform(Context& ctx, int species, const int ppc[3], double rho, const double drift[3],
                        const double thermal[3], int quiet, uint64_t seed, uint64_t first_id,
                        const double* density_seed,  // {k, phase, amplitude} or null
                        const HarrisLoad* harris);   // or null
int launch_exclusive_scan(Context& ctx, const uint32_t* in, uint32_t* out, uint32_t n);

}  // namespace tpic_mi
