/*
 * CPU oracle of the ECSIM-cycle hot path: a plain C11 restatement of the
 * reference algorithm (lucak17/TPIC), used ONLY as test infrastructure.
 *
 * Nothing in the shipped library (tpic_amd/) links, loads or calls this code.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * and there only as the checker.
 *
 * Layouts (shared with the C-ABI in include/tpic_mi.h):
 *   compact node vector  v[c*N + n], n = i + (nx+1)*(j + (ny+1)*k), i in [0,nx]
 *   compact cell vector  b[c*C + n], n = i + nx*(j + ny*k),          i in [0,nx)
 *   particle key         (i+2) + (nx+4)*((j+2) + (ny+4)*(k+2))   (halo-2 map)
 *   mass record          vertex*126 + stencil*9 + comp,
 *                        vertex = (i+1) + (nx+3)*((j+1) + (ny+3)*(k+1)), i in [-1,nx+1]
 * Internally the oracle keeps one-ghost node/cell arrays and fills their
 * ghosts by periodic copies, as the reference's self-route halos do, or, on
 * a wall axis, by the wall's signed reflections.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).
 */
#ifndef TPIC_ORACLE_H
#define TPIC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Validation bits, identical to the C-ABI's TPIC_MI_INVALID_* values. */
enum {
  OR_INVALID_KEY = 1u,
  OR_INVALID_FRACTION = 2u,
  OR_INVALID_NONFINITE = 4u,
  OR_INVALID_FLIGHT = 8u,
  OR_INVALID_HALO = 16u,
  OR_INVALID_SORT = 32u,
  OR_INVALID_FIELD = 64u
};

/* Field faces (config::FieldBoundary kinds; 0 = periodic so that a zeroed
 * grid is the all-periodic box). An axis is periodic on both sides or has a
 * wall on both sides. */
enum { OR_FACE_PERIODIC = 0, OR_FACE_PEC = 1, OR_FACE_PMC = 2 };

typedef struct {
  int n[3];                 /* cells per axis */
  double spacing[3];        /* dx, dy, dz */
  int particle_periodic[3]; /* 1 periodic, 0 mirror */
  int field_faces[3][2];    /* [axis][lower, upper]: OR_FACE_* */
} or_grid;

typedef struct {
  double beta;              /* (q/m) dt / 2 */
  double beta_over_c;       /* beta / c */
  double dt;
  double inverse_spacing[3];
} or_mover_params;

typedef struct {
  double *rx, *ry, *rz, *u, *v, *w, *q;
  int32_t *cell;
  uint64_t *id;
  uint32_t count;
} or_particles;

typedef struct {
  int converged;
  int iterations;
  double initial_residual, residual, rhs_norm, relative_residual;
} or_gmres_result;

/* ---- geometry ---------------------------------------------------------- */
int or_node_count(const or_grid *g);
int or_cell_count(const or_grid *g);
int or_bin_count(const or_grid *g);
int or_vertex_count(const or_grid *g);
int32_t or_encode(const or_grid *g, int i, int j, int k);
void or_decode(const or_grid *g, int32_t key, int *ijk);

/* ---- per-particle primitives ------------------------------------------ */
void or_cic_weights(double rx, double ry, double rz, double w[8]);
void or_response(double bx, double by, double bz, double beta_over_c,
                 double a[9], double *norm_square);
int or_mass_pair_stencil(int a, int b);

/* ---- particle kernels ------------------------------------------------- */
unsigned or_move(const or_grid *g, const or_mover_params *p, or_particles *s,
                 const double *e_nodes, const double *b_nodes);
int or_sort(const or_grid *g, const or_particles *in, or_particles *out,
            uint32_t *permutation, uint32_t *counts, uint32_t *offsets);
unsigned or_deposit_current(const or_grid *g, const or_particles *s,
                            const uint32_t *counts, const uint32_t *offsets,
                            const double *b_nodes, double inverse_volume,
                            double beta_over_c, double *current);
unsigned or_deposit_mass(const or_grid *g, const or_particles *s,
                         const uint32_t *counts, const uint32_t *offsets,
                         const double *b_nodes, double beta_over_c, double *mass);
void or_fold_current(const or_grid *g, double *current);
unsigned or_deposit_moments(const or_grid *g, const or_particles *s, double q_over_m, double inverse_volume,
                            unsigned selected, double *out);
void or_fold_mass(const or_grid *g, double *mass);

/* ---- field operators (compact in, compact out) ------------------------- */
void or_curl_n2c(const or_grid *g, const double *nodes, double *cells);
void or_curl_c2n(const or_grid *g, const double *cells, double *nodes);
void or_cell_to_node(const or_grid *g, const double *cells, double *nodes);
void or_mass_apply(const or_grid *g, const double *mass, const double *v,
                   double inverse_volume, double *out);
void or_maxwell_apply(const or_grid *g, const double *mass, double delta,
                      double current_factor, double inverse_volume,
                      const double *v, double *out);
void or_maxwell_rhs(const or_grid *g, const double *e, const double *b,
                    const double *current, double delta, double current_factor,
                    double *rhs);

/* ---- energy-conserving smoothing ---------------------------------------- */
#define OR_SMOOTH_FACE7 0
#define OR_SMOOTH_BINOMIAL27 1
typedef struct {
  int op;          /* OR_SMOOTH_* */
  int iterations;  /* SmoothNiter */
  double alpha;    /* Smooth */
} or_smoothing;
/* Active when iterations > 0 and alpha < 1 (physics.cpp:70). */
int or_smoothing_active(const or_smoothing *sm);
/* out = S(in) = P (pass P)^n in, P zeroing the constrained rows (node
 * vectors, in may equal out). */
void or_smooth(const or_grid *g, const or_smoothing *sm, const double *in, double *out);
/* The smoothed operator on the projected v: 1*v + d^2 curl curl v +
 * k S((1/V) f_g M S(v)), constrained rows of v. */
void or_maxwell_apply_smoothed(const or_grid *g, const or_smoothing *sm, const double *mass, double delta,
                               double current_factor, double inverse_volume, const double *v, double *out);

/* ---- simulation (one GPU's whole domain, periodic fields) ------------- */
typedef struct or_sim or_sim;

or_sim *or_sim_create(const or_grid *g, int species, const double *q_over_m,
                      const uint32_t *counts, double dt, double theta, double c,
                      double gmres_tol, int restart, int max_iterations);
void or_sim_destroy(or_sim *s);
/* which: 0 E, 1 B (cells), 2 Bn, 3 J, 4 theta (E_theta), 5 mass */
double *or_sim_field(or_sim *s, int which);
or_particles *or_sim_particles(or_sim *s, int species);
uint32_t *or_sim_bins(or_sim *s, int species, int offsets);
int or_sim_initialize(or_sim *s);
/* Smoothing of the run (none by default); the movers then gather S(E_theta),
 * or_sim_field(s, 6). Returns -1 for an invalid setting. */
int or_sim_set_smoothing(or_sim *s, int op, int iterations, double alpha);
int or_sim_solve(or_sim *s, or_gmres_result *result);
unsigned or_sim_move(or_sim *s, int species);
void or_sim_advance_fields(or_sim *s);
int or_sim_sort(or_sim *s, int species);
unsigned or_sim_deposit(or_sim *s);
int or_sim_cycle(or_sim *s, int cycles, or_gmres_result *last);
/* out: W_E, W_B, then per species kinetic energy */
void or_sim_energy(const or_sim *s, double *out);
void or_sim_scalars(const or_sim *s, double *out);  /* W_E, W_B, 5 per species */
unsigned or_sim_validation(const or_sim *s);

/* ---- initial load ------------------------------------------------------ */
/* The backend's default XORWOW on ROCm (rocRAND's xorwow_engine, behind
 * hipRAND's hiprand_init / hiprand_uniform_double / hiprand_normal2_double /
 * hiprand_normal_double), restated: state, seeding, subsequence jump of
 * 2^67 steps, and the double conversions. */
typedef struct {
  uint32_t d, x[5];
  double cached_normal;   /* second Box-Muller value of hiprand_normal_double */
  int has_cached;
} or_xorwow;
void or_xorwow_init(or_xorwow *st, uint64_t seed, uint64_t subsequence);
uint32_t or_xorwow_next(or_xorwow *st);
double or_xorwow_uniform_double(or_xorwow *st);
void or_xorwow_normal2_double(or_xorwow *st, double out[2]);
double or_xorwow_normal_double(or_xorwow *st);
/* The reference's Uniform-case load of one species (particle_loading_device.hpp
 * :69-159 with UniformProfile, prepared_initialization.cpp:37-85, and the
 * startup boundary fold). out->count must be cells * ppc. */
void or_load_uniform(const or_grid *g, const int ppc[3], double rho_init, const double drift[3],
                     const double thermal[3], int quiet_start, uint64_t seed, uint64_t first_id,
                     double q_over_m, or_particles *out);
/* The load with the TwoStream density seed {k, phase, amplitude} (NULL: uniform). */
void or_load_seeded(const or_grid *g, const int ppc[3], double rho_init, const double drift[3],
                    const double thermal[3], int quiet_start, uint64_t seed, uint64_t first_id,
                    double q_over_m, const double *seed_mode, or_particles *out);
/* The Harris profile (HarrisProfile, HarrisCurl: harris_sheet.hpp:33-120). */
typedef struct {
  int density, is_double, varying_thermal;
  double ly, delta, ampere_drift;
  double center_x, center_y, curl_delta, sheet, mode, mode_kx, mode_ky;
  double hump, hump_kx, hump_ky, hump_inverse_wx2, hump_inverse_wy2;
} or_harris;
double or_harris_curl(const or_harris *h, double x, double y);
void or_load_profile(const or_grid *g, const int ppc[3], double rho_init, const double drift[3],
                     const double thermal[3], int quiet_start, uint64_t seed, uint64_t first_id,
                     double q_over_m, const double *seed_mode, const or_harris *harris, or_particles *out);
/* CenterInitialVelocity with a uniform drift; 1 on a non-finite sum/mean/velocity. */
int or_center_velocity(or_particles *s, const double drift[3]);

/* GMRES on a caller-supplied operator (tests of the solver alone). */
typedef void (*or_operator)(void *context, const double *in, double *out);
or_gmres_result or_gmres(int n, or_operator apply, void *context, const double *rhs,
                         double *solution, int restart, int max_iterations,
                         double tolerance);

#ifdef __cplusplus
}
#endif

#endif /* TPIC_ORACLE_H */
