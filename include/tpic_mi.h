/*
 * tpic_mi.h - C-ABI of the MI355X ECSIM-cycle hot path (particle mover,
 * current and mass-matrix deposition, per-cell particle sort, implicit
 * Maxwell solve by GMRES) for one GPU, with periodic fields or PEC/PMC walls.
 *
 * The reference (lucak17/TPIC) exposes this path as compile-time C++
 * templates: free enqueue_* functions and the Gmres/MaxwellOperator owners
 * called by simulation::Pipeline. Each entry point below names the reference
 * interface it replaces (paths relative to the reference repository root).
 *
 * Conventions
 *  - Every function returns TPIC_MI_OK (0) or a negative status; no C++
 *    exception crosses this boundary. tpic_mi_last_error() gives the text.
 *  - All device work is ordered on the context's HIP stream. Step calls
 *    allocate nothing. Only tpic_mi_solve, tpic_mi_energy, tpic_mi_center_velocity,
 *    tpic_mi_read_*, tpic_mi_get_* and tpic_mi_synchronize block the host.
 *  - Layouts (FP64 unless stated; N = (nx+1)(ny+1)(nz+1) nodes including the
 *    periodic duplicates, C = nx*ny*nz cells):
 *      node vector  v[c*N + i + (nx+1)*(j + (ny+1)*k)],  c = x, y, z
 *      cell vector  b[c*C + i + nx*(j + ny*k)]
 *      particle key (i+2) + (nx+4)*((j+2) + (ny+4)*(k+2)), bins (nx+4)(ny+4)(nz+4)
 *      mass record  ((i+1) + (nx+3)*((j+1) + (ny+3)*(k+1)))*126 + stencil*9 + entry
 *  - A context is used by one host thread at a time.
 */
#ifndef TPIC_MI_H
#define TPIC_MI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (the error classes of lpfx::ErrorCode and the solver's
 * non-convergence, include/tpic/ecsim/gmres/gmres_device.hpp:427-454). */
#define TPIC_MI_OK 0
#define TPIC_MI_INVALID_ARGUMENT (-1)
#define TPIC_MI_INVALID_STATE (-2)
#define TPIC_MI_OVERFLOW (-3)
#define TPIC_MI_BACKEND (-4)
#define TPIC_MI_NOT_CONVERGED (-5)
/* A valid input deck that asks for something this build does not run (the
 * unsupported requests of run admission, src/simulation/run_plan.cpp:237-245). */
#define TPIC_MI_UNSUPPORTED (-6)

/* Mass-matrix deposition kernels (MassDepositionKernel, numerics_options.hpp:9-62,
 * with the precision pair of precision.hpp:18-160 narrowed to the moment
 * products): each cell tile's 27 x 12 moment matrix sum_p P_p T_p^T is formed
 * on MFMA with the given operands; tile totals are summed across cells in FP64. */
#define TPIC_MI_MASS_FP64 0       /* FP64 operands and accumulation (parity path) */
#define TPIC_MI_MASS_FP32 1       /* FP32 operands, FP32 accumulation within a cell */
#define TPIC_MI_MASS_BF16X3 2     /* BF16 hi/lo split (hi*hi + hi*lo + lo*hi), FP32 accumulation */
/* The mixed kernels (FP32, BF16X3) also store the assembled matrix in FP32
 * (StorageSingleParticlesMomentsFields::Mass, precision.hpp); its cell sums,
 * the operator's products and every other quantity stay FP64. */

/* Bits of the device validation word (the reference's `unsigned* invalid`
 * flags, include/tpic/device/device_validation.hpp:13-50, split by cause). */
#define TPIC_MI_INVALID_KEY 1u        /* key out of range or not an interior cell */
#define TPIC_MI_INVALID_FRACTION 2u   /* stored fraction outside [0, 1) */
#define TPIC_MI_INVALID_NONFINITE 4u  /* non-finite charge, velocity, field or response */
#define TPIC_MI_INVALID_FLIGHT 8u     /* cell shift outside the int32 range */
#define TPIC_MI_INVALID_HALO 16u      /* a particle left in a halo bin after a sort */
#define TPIC_MI_INVALID_SORT 32u      /* the sort rejected a key; particles unchanged */
#define TPIC_MI_INVALID_FIELD 64u     /* non-finite field or solver value */

#define TPIC_MI_MAX_SPECIES 8

/* Field selectors of tpic_mi_get_field / tpic_mi_set_field. */
#define TPIC_MI_FIELD_E 0      /* E^n at nodes (3N) */
#define TPIC_MI_FIELD_B 1      /* B^n at cells (3C) */
#define TPIC_MI_FIELD_BNODE 2  /* B at nodes (3N): cell_to_node_average of B */
#define TPIC_MI_FIELD_J 3      /* assembled J hat at nodes (3N) */
#define TPIC_MI_FIELD_ETHETA 4 /* E_theta at nodes (3N), the last solution */
#define TPIC_MI_FIELD_RHS 5    /* solver right-hand side (3N) */
#define TPIC_MI_FIELD_MASS 6   /* mass matrix, 126 scalars per vertex of (nx+3)(ny+3)(nz+3);
                                  stored in FP32 by the mixed mass kernels (get_field widens) */
#define TPIC_MI_FIELD_EPARTICLE 7 /* the movers' electric field at nodes (3N): S(E_theta) with
                                     smoothing (complete_particle_field, maxwell_device.hpp:
                                     213-228), formed by tpic_mi_solve; else E_theta itself */

/* Phase selectors of tpic_mi_phase_time. */
#define TPIC_MI_PHASE_SORT 0
#define TPIC_MI_PHASE_MOVE 1
#define TPIC_MI_PHASE_DEPOSIT 2
#define TPIC_MI_PHASE_ASSEMBLE 3
#define TPIC_MI_PHASE_SOLVE 4
#define TPIC_MI_PHASE_ADVANCE 5
#define TPIC_MI_PHASE_SPMV 6

typedef struct tpic_mi_context tpic_mi_context;

/* Species of one run (config::SpeciesParameters, include/tpic/config/
 * simulation_parameters.hpp: qom and the planned particle count). */
typedef struct {
  double q_over_m;
  uint32_t count;
} tpic_mi_species;

/* Field faces of tpic_mi_config.field_faces (config::FieldBoundary,
 * include/tpic/config/boundary_conditions.hpp:41-77; zero is periodic so that
 * a zeroed config is the periodic box). */
#define TPIC_MI_FACE_PERIODIC 0
#define TPIC_MI_FACE_PEC 1  /* perfect electric conductor: tangential E held at zero */
#define TPIC_MI_FACE_PMC 2  /* perfect magnetic conductor: normal E, tangential B at zero */

/* Geometry, numerics and species (the deck keys nxc/nyc/nzc, Lx/Ly/Lz,
 * PERIODIC_P*, dt, th, c, GMREStol; GMRES restart 30 and 100 iterations are
 * fixed in the reference, include/tpic/config/physics.hpp:20-21). Particles
 * are periodic or mirrored per axis; fields are periodic per axis or closed
 * by a PEC or PMC wall on both faces (field_faces, from PERIODICX/Y/Z and
 * bcEMface*). */
typedef struct {
  int cells[3];
  double lengths[3];
  int particle_periodic[3];
  double dt, theta, c;
  double gmres_tolerance;
  int gmres_restart;        /* <= 64 */
  int gmres_max_iterations;
  int mass_kernel;          /* TPIC_MI_MASS_*: operand precision of the cell moment products */
  int species;
  tpic_mi_species species_list[TPIC_MI_MAX_SPECIES];
  void* stream;             /* hipStream_t, or NULL for a context-owned stream */
  int field_faces[3][2];    /* [axis][lower, upper]: TPIC_MI_FACE_*; an axis is periodic on both
                               faces or walled on both; a walled axis needs at least 2 cells.
                               Walls: signed ghost images of B for the node curl and the
                               cell-to-node average (field_boundary_device.hpp:26-57), constrained
                               E rows held at zero (:60-120), J and the mass rows scaled by the
                               dual-volume factor f_g (wall_faces.hpp:27-30), no fold across a
                               wall, W_E weighted 1/f_g over the owned nodes. */
  /* Energy-conserving smoothing (ECSimSmoothing=energy_conserving with
   * ECSimSmoothingOperator, Smooth and SmoothNiter; admission src/config/
   * physics.cpp:65-71; the flow of MaxwellOperator, include/tpic/ecsim/maxwell/
   * maxwell_device.hpp:151-227,389-417,455-490). Active when smooth_iterations
   * > 0 and smooth_alpha < 1 (a zeroed config has none). S is smooth_iterations
   * passes of the node stencil (node_smoothing.hpp:27-58) over E-like node
   * ghosts; the solver uses S(J) in the right-hand side, S((1/V) f_g M S(v)) in
   * the operator, and the movers gather S(E_theta) (tpic_mi_smooth). */
  int smooth_operator;      /* TPIC_MI_SMOOTH_FACE7 or TPIC_MI_SMOOTH_BINOMIAL27 */
  int smooth_iterations;    /* SmoothNiter, >= 0 */
  double smooth_alpha;      /* Smooth in [0, 1]: Face7's centre weight */
} tpic_mi_config;

/* Node smoothing operators (config::EcsimSmoothingOperator,
 * include/tpic/config/simulation_parameters.hpp:22). */
#define TPIC_MI_SMOOTH_FACE7 0      /* alpha centre, (1 - alpha)/6 per face neighbour */
#define TPIC_MI_SMOOTH_BINOMIAL27 1 /* [1,2,1]^3 / 64 */

/* Host SoA arrays of one species (data::ParticleSchema,
 * include/tpic/data/particles/particle_schema.hpp:20-36). */
typedef struct {
  double *rx, *ry, *rz, *u, *v, *w, *q;
  int32_t* cell;
  uint64_t* id;
} tpic_mi_particle_arrays;

typedef struct {
  int converged;
  int iterations;
  double initial_residual, residual, rhs_norm, relative_residual;
} tpic_mi_solve_result;

/* Dense uniform Maxwellian load of one species (initial_particle_load_kernel,
 * include/tpic/initialization/particles/particle_loading_device.hpp:69-159):
 * ppc per axis, rhoINIT, drift u0/v0/w0 and thermal uth/vth/wth. */
typedef struct {
  int ppc[3];
  double rho;
  double drift[3];
  double thermal[3];
  uint64_t seed;
  uint64_t first_id;
} tpic_mi_load;

/* The reference's Uniform-case initial load of one species
 * (kernels::initial_particle_load_kernel with UniformProfile,
 * include/tpic/initialization/particles/particle_loading_device.hpp:69-159;
 * plan: src/initialization/prepared_initialization.cpp:37-85). */
typedef struct {
  int ppc[3];            /* npcelx, npcely, npcelz */
  double rho_init;       /* rhoINIT; the density is rho_init / (4 pi) */
  double drift[3];       /* u0, v0, w0 */
  double thermal[3];     /* uth, vth, wth */
  int quiet_start;       /* QuietStart: subcell centres, no offset draws */
  int rank;              /* partition rank in the seed (0 on one GPU) */
  int max_species;       /* TPIC_MAX_SPECIES of the reference build (10 by default) */
  uint64_t first_id;
} tpic_mi_uniform_load;

/* ---- lifetime -------------------------------------------------------- */

/* Allocates every device owner once (Simulation constructor,
 * include/tpic/simulation/driver/simulation_device.hpp:62-114). */
int tpic_mi_create(const tpic_mi_config* config, tpic_mi_context** out);
int tpic_mi_destroy(tpic_mi_context* ctx);
const char* tpic_mi_last_error(const tpic_mi_context* ctx);
int tpic_mi_synchronize(tpic_mi_context* ctx);
/* Device pointer of a field owner or a species' current SoA slot (valid until
 * the next sort or destroy); which = TPIC_MI_FIELD_* or, for particles, the
 * attribute index rx,ry,rz,u,v,w,q,cell,id = 0..8. */
int tpic_mi_device_field(tpic_mi_context* ctx, int which, void** out);
int tpic_mi_device_particles(tpic_mi_context* ctx, int species, int attribute, void** out);

/* ---- deck front end --------------------------------------------------- */
/* A context started from an input deck: the reference's deck grammar and
 * defaults (config::SimulationParameters, src/config/simulation_parameters.cpp:
 * 55-392), its admission (src/config/physics.cpp:26-119, src/config/domain.cpp:
 * 54-83, src/simulation/run_plan.cpp:100-135), then the startup of
 * run_simulation (include/tpic/simulation/driver/simulation_device.hpp:62-114,
 * src/initialization/prepared_initialization.cpp:37-85): every species'
 * load of the deck's case, the case's E and B images plus B0, and the
 * initialization, on stream (NULL: the null stream). The context is then
 * ready for tpic_mi_cycle. A broken deck returns TPIC_MI_INVALID_ARGUMENT, a
 * deck this build does not run TPIC_MI_UNSUPPORTED; either way *ctx stays
 * NULL and tpic_mi_last_error(NULL) holds the reason (every unsupported
 * request, one a line). Output keys are parsed and checked but write nothing
 * here (the tpic_mi_run driver writes them). */
typedef struct {
  int32_t mass_kernel;   /* TPIC_MI_MASS_* operands of the moment products */
  int32_t reserved[7];   /* zero */
} tpic_mi_deck_options;
typedef struct {
  int32_t ncycles, species;
  int32_t cells[3];
  int32_t energy_cycle;  /* DiagnosticEnergyCycle */
  double lengths[3];
  double dt, theta, c;
  uint64_t count[TPIC_MI_MAX_SPECIES];  /* particles of each species */
  double q_over_m[TPIC_MI_MAX_SPECIES];
} tpic_mi_deck_run;
int tpic_mi_load_deck(const char* path, const tpic_mi_deck_options* options, void* stream, tpic_mi_context** ctx,
                      tpic_mi_deck_run* run);

/* ---- state upload / download ----------------------------------------- */
int tpic_mi_set_particles(tpic_mi_context* ctx, int species, const tpic_mi_particle_arrays* host);
int tpic_mi_get_particles(tpic_mi_context* ctx, int species, tpic_mi_particle_arrays* host);
/* Device load of a uniform Maxwellian (replaces the hipRAND XORWOW draws of
 * the reference with Philox-4x32-10; positions, charges, keys and ids follow
 * the reference's dense path). */
int tpic_mi_load_maxwellian(tpic_mi_context* ctx, int species, const tpic_mi_load* load);
/* The reference's Uniform-case load, draw for draw: hipRAND XORWOW with seed
 * 12345 + 100 (rank max_species + species) and subsequence = load index;
 * three uniform offsets (unless quiet_start), a normal pair for u, v and a
 * normal for w; the exact-1 fraction carry; the startup boundary fold
 * (particle_loading_device.hpp:37-55,111-137; prepared_initialization.cpp:46-47;
 * initial_particle_classification_device.hpp:22-55). A non-finite velocity
 * or charge raises TPIC_MI_INVALID_NONFINITE. */
int tpic_mi_load_uniform(tpic_mi_context* ctx, int species, const tpic_mi_uniform_load* load);

/* The TwoStream case's density seed (TwoStreamProfile,
 * include/tpic/initialization/cases/two_stream.hpp:13-28): the load of
 * tpic_mi_load_uniform (same draws, positions, velocities, keys, ids) with
 * each charge multiplied by 1 + amplitude cos(wavenumber x + phase), x the
 * particle's position along x after the exact-1 carry (one-rank origin 0). */
typedef struct {
  double wavenumber;  /* 2 pi TwoStreamSeedModeX / Lx */
  double phase;       /* TwoStreamSeedPhase */
  double amplitude;   /* TwoStreamSeedDensityAmplitude, |amplitude| < 1 */
} tpic_mi_density_seed;
int tpic_mi_load_density_seeded(tpic_mi_context* ctx, int species, const tpic_mi_uniform_load* load,
                                const tpic_mi_density_seed* seed);

/* The Harris-sheet cases' profile (HarrisProfile and HarrisCurl,
 * include/tpic/initialization/cases/harris_sheet.hpp:33-120; built by
 * HarrisSheetCase, src/initialization/cases/harris_sheet.cpp:183-227): the
 * load of tpic_mi_load_uniform with, at the particle's position after the
 * carry, the density times the sheet shape averaged over the cell's two y
 * nodes (harris_density), the z drift times -tanh((y - ly/2)/delta) for a
 * double sheet, and with ampere_drift != 0 the z drift
 * ampere_drift curl_z(x, y) / cell shape (currentFromAmpere; with
 * varying_thermal the z thermal speed keeps rho (drift^2 + wth^2)). */
typedef struct {
  int harris_density;   /* the species carries the sheet density */
  int is_double;        /* DoubleGEM: sheets at ly/4 and 3 ly/4 */
  int varying_thermal;  /* spatiallyVaryingThermal */
  double ly, delta;     /* Ly and the sheet half-width delta */
  double ampere_drift;  /* w0 c / (rhoINIT sum sign(qom) w0), or 0 */
  double curl_center[2], curl_delta, curl_sheet, curl_mode, curl_mode_k[2];
  double curl_hump, curl_hump_k[2], curl_hump_inverse_w2[2];
} tpic_mi_harris_profile;
int tpic_mi_load_harris(tpic_mi_context* ctx, int species, const tpic_mi_uniform_load* load,
                        const tpic_mi_harris_profile* profile);
/* CenterInitialVelocity of a species loaded with a uniform drift (drift[3] =
 * u0, v0, w0): the charge-weighted mean of v - drift is removed, so the mean
 * velocity becomes the drift (VelocityCentering::host_blocking_center and
 * center_particle_velocity_kernel, include/tpic/initialization/particles/
 * particle_centering_device.hpp:57-215). Call after the load, before the
 * first sort. Blocks the host. TPIC_MI_OVERFLOW on a non-finite sum, mean or
 * corrected velocity (a velocity that would be non-finite is left as v - drift). */
int tpic_mi_center_velocity(tpic_mi_context* ctx, int species, const double* drift);
int tpic_mi_set_field(tpic_mi_context* ctx, int which, const double* host);
int tpic_mi_get_field(tpic_mi_context* ctx, int which, double* host);

/* Output moments of one sorted species at nodes, written to device_out as 10
 * node vectors in the order rho, Jx, Jy, Jz, PXX, PXY, PXZ, PYY, PYZ, PZZ:
 * charge density q/V, current q v / V and the raw mass-weighted second
 * moments (q / qom) v_a v_b / V, CIC-weighted, both copies of a periodic node
 * holding their sum, wall nodes scaled by the dual-volume factor (rho without
 * the 4 pi of its output scale). Bit c of selected picks component c; the
 * others are written as zero. Invalid particles are skipped and set bits of
 * the validation word. Returns after the moments are written.
 * (deposit_output_moments_kernel and enqueue_deposit_output_moments,
 * include/tpic/diagnostics/output_moment_deposition_device.hpp:35-182;
 * the capture, include/tpic/simulation/actions/output_actions_device.hpp:130-146;
 * component order, include/tpic/data/output/output_moment_quantities.hpp:26-34.) */
#define TPIC_MI_MOMENTS_ALL 0x3FFu
int tpic_mi_moments(tpic_mi_context* ctx, int species, unsigned selected, double* device_out);
/* counts[bins], offsets[bins + 1]; permutation[count] is the destination of
 * each source of the last sort (NULL to skip; needs record_permutation). */
int tpic_mi_get_bins(tpic_mi_context* ctx, int species, uint32_t* counts, uint32_t* offsets,
                     uint32_t* permutation);
int tpic_mi_record_permutation(tpic_mi_context* ctx, int enable);

/* ---- the hot path ----------------------------------------------------- */

/* Startup: B at nodes from B cells, sort every species, bootstrap deposition
 * and assembly (Pipeline::initialize and start_bootstrap,
 * include/tpic/simulation/driver/pipeline.hpp:331-435;
 * complete_initial_magnetic_field, magnetic_startup_device.hpp). */
int tpic_mi_initialize(tpic_mi_context* ctx);

/* Stable counting sort of one species by cell key (FullParticleSort::enqueue,
 * include/tpic/ecsim/particles/particle_sort_device.hpp:132-145, and
 * lpfx DeviceCellSort, include/lpfx/algorithm/particle/device/cuda_hip/sort.hpp:193-260;
 * order within a bin as lpfx::PreparedSerialParticleSort::run,
 * include/lpfx/algorithm/particle/host/sort.hpp:210-305). */
int tpic_mi_sort(tpic_mi_context* ctx, int species);

/* ECSIM push of one species with E_theta and B at nodes, periodic/mirror fold
 * (enqueue_local_particle_mover, include/tpic/ecsim/particles/particle_mover_device.hpp:310-328). */
int tpic_mi_move(tpic_mi_context* ctx, int species);

/* Current and mass deposition of species [first, last) from their sorted
 * particles with B at nodes into the per-cell moment buffer; accumulate = 0
 * overwrites it (enqueue_sorted_current, current_deposition_device.hpp:373-397;
 * enqueue_sorted_mass_matrix, mass_matrix_deposition_device.hpp:65-101;
 * enqueue_clear_moment_field / enqueue_clear_mass_matrix). */
int tpic_mi_deposit(tpic_mi_context* ctx, int first_species, int last_species, int accumulate);

/* Node assembly of J and M from the cell buffer, periodic fold included
 * (CurrentAssembly, current_assembly_device.hpp:17-84; MassHaloStorage
 * additive and ghost phases, mass_halo_device.hpp:162-399). */
int tpic_mi_assemble(tpic_mi_context* ctx);

/* rhs and GMRES solve for E_theta (MaxwellOperator::enqueue_rhs_fields and
 * complete_rhs, include/tpic/ecsim/maxwell/maxwell_device.hpp:134-171;
 * Gmres::host_blocking_solve, include/tpic/ecsim/gmres/gmres_device.hpp:355-454).
 * Returns TPIC_MI_NOT_CONVERGED and leaves E_theta unchanged when it fails. */
int tpic_mi_solve(tpic_mi_context* ctx, tpic_mi_solve_result* result);

/* Operator image out = A(in) on device node vectors (MaxwellOperator::complete_apply,
 * maxwell_device.hpp:184-433). With walls, in is taken with its constrained
 * rows zeroed (begin_apply_input, :270-272), and those rows of out are zero. */
int tpic_mi_apply_operator(tpic_mi_context* ctx, const double* device_in, double* device_out);

/* out = S(in), the configured smoothing (MaxwellOperator::complete_smoothing,
 * maxwell_device.hpp:452-490, whose passes are enqueue_node_smoothing,
 * include/tpic/ecsim/fields/node_smoothing_device.hpp:45-81): in with its
 * constrained rows zeroed, then smooth_iterations passes of the node stencil
 * over E-like wall ghosts, each result with its constrained rows zeroed.
 * Device node vectors; in may equal out. TPIC_MI_INVALID_ARGUMENT when
 * smoothing is not active. */
int tpic_mi_smooth(tpic_mi_context* ctx, const double* device_in, double* device_out);

/* E^{n+1}, B^{n+1} and B at nodes (enqueue_next_fields,
 * include/tpic/ecsim/maxwell/field_update_device.hpp:52-78; next-cell closure
 * and commit, maxwell_device.hpp:234-267). */
int tpic_mi_advance_fields(tpic_mi_context* ctx);

/* n full cycles: solve, move every species, advance fields, sort, deposit,
 * assemble (Pipeline::run_cycles, pipeline.hpp:437-541). Stops at the first
 * failure; last receives the last solve. */
int tpic_mi_cycle(tpic_mi_context* ctx, int cycles, tpic_mi_solve_result* last);

/* W_E, W_B, then kinetic energy per species (ScalarDiagnostics,
 * include/tpic/diagnostics/scalar_diagnostics_device.hpp:50-112). */
int tpic_mi_energy(tpic_mi_context* ctx, double* out);

/* The full scalar sample of ScalarDiagnostics (scalar_diagnostics_device.hpp:
 * 50-160; the columns of diagnostics.csv, src/io/diagnostics_writer.cpp:
 * 52-72): out[0] = W_E, out[1] = W_B, then per species s, out[2 + 5 s + r]
 * for r = kinetic energy sum (q/qom)|v|^2/2, charge sum q, momentum
 * sum (q/qom) v along x, y, z. Blocks the host. */
int tpic_mi_scalars(tpic_mi_context* ctx, double* out);

/* W_E and W_B of caller-owned device fields with no context and no
 * allocation. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops or
 * harnesses that forbid hipMalloc). It computes what tpic_mi_energy reports in
 * out[0..1] (the same kernel, ScalarDiagnostics' field rows).
 *  - config: only cells, lengths and field_faces are read.
 *  - device_e: E at nodes (3N), device_b: B at cells (3C), device memory.
 *  - device_out: device memory; [0] = W_E, [1] = W_B.
 *  - stream: hipStream_t, or NULL for the null stream. The work is enqueued
 *    there and the call returns without waiting.
 * Limit: the cell count and the owned-node count must each be at most
 * 16,777,216 (256^3), else TPIC_MI_INVALID_ARGUMENT. The per-block partial
 * sums live in one 1 MiB device array per GPU that every call on that GPU
 * shares; calls that overlap on one GPU (other streams, other host threads)
 * race, so the caller serialises them. Errors: tpic_mi_last_error(NULL). */
int tpic_mi_field_energy_static(const tpic_mi_config* config, const double* device_e, const double* device_b,
                                double* device_out, void* stream);

/* Validation word accumulated since creation or the last clear. */
int tpic_mi_read_validation(tpic_mi_context* ctx, unsigned* flags);
int tpic_mi_clear_validation(tpic_mi_context* ctx);

/* ---- cycle checkpoints ------------------------------------------------ */
/* In-device copies of the state a cycle starts from (sorted particles with
 * their bins, E at nodes, B at cells): the in-memory cycle checkpoints the
 * reference's checkpoint/restart keeps on files (CheckpointHeader and its E
 * and B pieces, include/tpic/data/output/checkpoint.hpp:14-60). Restoring
 * recomputes B at nodes, the deposition and the assembly with the kernels a
 * cycle ends with, so a run continued from a checkpoint repeats the original
 * bit for bit. Slots are device memory sized by the context's species. */
int tpic_mi_checkpoint_reserve(tpic_mi_context* ctx, int slots);
int tpic_mi_checkpoint_save(tpic_mi_context* ctx, int slot);
int tpic_mi_checkpoint_restore(tpic_mi_context* ctx, int slot);

/* Restart files. The header carries the fields of the reference's
 * CheckpointHeader (include/tpic/data/output/checkpoint.hpp:14-33, with
 * config::max_real_species = 10 entries per species array); after it come E
 * at the nodes, B at the cells and every species' particles in stored order,
 * raw little-endian (the reference writes ADIOS2 or HDF5 pieces). Written
 * after a cycle, read into a context of the same grid and species and
 * followed by tpic_mi_initialize, the run continues bit for bit. The writer
 * sets species, processes, global_cells, lengths, dt, c, storage_bytes,
 * q_over_m and global_count from the context and keeps the caller's cycle,
 * initial_energy and faces. */
#define TPIC_MI_CHECKPOINT_MAX_SPECIES 10
typedef struct {
  int32_t cycle, species;
  int32_t processes[3], global_cells[3];
  double lengths[3];
  double dt, c;
  double initial_energy;  /* total energy of the run's first state: the dE/E0 baseline */
  int32_t faces[12];      /* FieldBoundary of Xright..Zleft, then ParticleBoundary */
  int32_t storage_bytes[5]; /* position, velocity, weight, electric, magnetic */
  double q_over_m[TPIC_MI_CHECKPOINT_MAX_SPECIES];
  uint64_t global_count[TPIC_MI_CHECKPOINT_MAX_SPECIES];
} tpic_mi_checkpoint_header;
int tpic_mi_checkpoint_write(tpic_mi_context* ctx, const char* path, const tpic_mi_checkpoint_header* header);
/* Reads E, B and the particles (header may be NULL); the species are then
 * unsorted: call tpic_mi_initialize before the next cycle. */
int tpic_mi_checkpoint_read(tpic_mi_context* ctx, const char* path, tpic_mi_checkpoint_header* header);
/* The header alone, before any context exists (the restart admission of
 * src/simulation/run_plan.cpp:199-214 compares it with the deck); errors are
 * reported by tpic_mi_last_error(NULL). */
int tpic_mi_checkpoint_read_header(const char* path, tpic_mi_checkpoint_header* header);

/* ---- adjoint of the cycle (periodic fields and particles, no smoothing) - */
/* No counterpart in the reference: the discrete adjoint of the cycle
 * tpic_mi_cycle runs, for gradients of field energies with respect to the
 * initial particles and fields. */

/* out = A^T(in): the transposed Maxwell operator, the transposed 3 x 3 mass
 * blocks of mass_matrix_stencil.hpp:139-149 with the curl-curl term (symmetric
 * over the unique nodes of a periodic box). Device node vectors. */
int tpic_mi_apply_operator_transposed(tpic_mi_context* ctx, const double* device_in, double* device_out);
/* GMRES (the solver of tpic_mi_solve, gmres_device.hpp:355-454) on
 * A^T x = rhs; device_x holds the initial guess and receives the solution. */
int tpic_mi_solve_transposed(tpic_mi_context* ctx, const double* device_rhs, double* device_x,
                             tpic_mi_solve_result* result);

typedef struct tpic_mi_adjoint tpic_mi_adjoint;
int tpic_mi_adjoint_create(tpic_mi_context* ctx, tpic_mi_adjoint** out);
int tpic_mi_adjoint_destroy(tpic_mi_adjoint* adjoint);
/* The adjoint of L = weight_e W_E + weight_b W_B of the context's current
 * state: dL/dE = weight_e V/(4 pi) E, dL/dB = weight_b V/(4 pi) B, zero
 * particle adjoints, held in the context's particle order. first_ids[s]:
 * species s's ids must be a permutation of [first_ids[s], first_ids[s] + count). */
int tpic_mi_adjoint_seed(tpic_mi_context* ctx, tpic_mi_adjoint* adjoint, double weight_e, double weight_b,
                         const uint64_t* first_ids);
/* One cycle backwards: the context holds the (sorted) state the cycle starts
 * from (tpic_mi_checkpoint_restore), the adjoint that of the state the cycle
 * ends in. The cycle's E_theta is recomputed (forward), the adjoint becomes
 * that of the starting state, in its particle order (transposed: the
 * A^T solve). The context keeps the starting state. */
int tpic_mi_adjoint_cycle(tpic_mi_context* ctx, tpic_mi_adjoint* adjoint, tpic_mi_solve_result* forward,
                          tpic_mi_solve_result* transposed);
/* Ebar (TPIC_MI_FIELD_E, 3N) or Bbar (TPIC_MI_FIELD_B, 3C). */
int tpic_mi_adjoint_get_field(tpic_mi_adjoint* adjoint, int which, double* host);
/* Per particle, in the adjoint's order (ids): rbar[3][count] (adjoint of the
 * cell fractions), vbar[3][count], qbar[count]; any pointer may be NULL. */
int tpic_mi_adjoint_get_particles(tpic_mi_adjoint* adjoint, int species, double* rbar, double* vbar,
                                  double* qbar, uint64_t* ids);

/* ---- measurement ------------------------------------------------------ */
/* Phase timing with HIP events on the context stream (StageTiming,
 * include/tpic/simulation/timing/stage_timing.hpp). */
int tpic_mi_enable_timing(tpic_mi_context* ctx, int enable);
int tpic_mi_phase_time(tpic_mi_context* ctx, int phase, double* total_ms, uint64_t* calls);
int tpic_mi_reset_timing(tpic_mi_context* ctx);

#ifdef __cplusplus
}
#endif

#endif /* TPIC_MI_H */
