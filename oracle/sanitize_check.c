/*
 * Host-sanitizer run of the CPU oracle (test infrastructure only): built with
 * AddressSanitizer and UndefinedBehaviorSanitizer by `make -C oracle
 * sanitize`, it drives every oracle stage on small grids -- the loads
 * (uniform, TwoStream seed, Harris profile), CenterInitialVelocity, the
 * cycle (solve, move, advance, sort, deposit) on a periodic box and on a
 * walled box with both smoothing operators, an empty species, a one-cell-thick
 * axis, and GMRES on a caller's operator -- and exits nonzero on a failed
 * check; the sanitizers abort on the first finding.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tpic_oracle.h"

static int failures = 0;

#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      fprintf(stderr, "check failed at %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

static or_grid make_grid(int nx, int ny, int nz, double lx, double ly, double lz) {
  or_grid g;
  memset(&g, 0, sizeof g);
  g.n[0] = nx;
  g.n[1] = ny;
  g.n[2] = nz;
  g.spacing[0] = lx / nx;
  g.spacing[1] = ly / ny;
  g.spacing[2] = lz / nz;
  for (int a = 0; a < 3; ++a) g.particle_periodic[a] = 1;
  return g;
}

/* A run of `cycles` cycles; every species loaded with ppc (a zero ppc makes
 * an empty species), species 0 with the Harris profile when harris != NULL. */
static void run(const or_grid *g, int species, const int ppc[][3], const double *qom, const or_harris *harris,
                const double *seed_mode, int smooth_op, int smooth_iterations, int cycles) {
  const int cells = or_cell_count(g);
  uint32_t counts[8];
  for (int s = 0; s < species; ++s) counts[s] = (uint32_t)(cells * ppc[s][0] * ppc[s][1] * ppc[s][2]);
  or_sim *sim = or_sim_create(g, species, qom, counts, 0.125, 0.5, 1.0, 1e-10, 30, 100);
  CHECK(sim != NULL);
  if (!sim) return;
  if (smooth_iterations > 0) CHECK(or_sim_set_smoothing(sim, smooth_op, smooth_iterations, 0.5) == 0);
  CHECK(or_sim_set_smoothing(sim, 7, 1, 0.5) == -1);
  uint64_t first = 0;
  for (int s = 0; s < species; ++s) {
    const double drift[3] = {0.05 * (s + 1), -0.02, 0.01};
    const double thermal[3] = {0.1, 0.08, 0.06};
    or_particles *p = or_sim_particles(sim, s);
    if (s == 0 && (harris || seed_mode))
      or_load_profile(g, ppc[s], 1.0, drift, thermal, 0, 12345u + 100u * (uint64_t)s, first, qom[s], seed_mode,
                      harris, p);
    else
      or_load_uniform(g, ppc[s], 1.0, drift, thermal, s == 1, 12345u + 100u * (uint64_t)s, first, qom[s], p);
    if (p->count) CHECK(or_center_velocity(p, drift) == 0);
    first += p->count;
  }
  const int nodes = or_node_count(g);
  double *e = or_sim_field(sim, 0), *b = or_sim_field(sim, 1);
  for (int n = 0; n < 3 * nodes; ++n) e[n] = 1e-3 * sin(0.37 * n);
  for (int n = 0; n < 3 * cells; ++n) b[n] = 0.2 + 1e-3 * cos(0.11 * n);
  CHECK(or_sim_initialize(sim) == 0);
  double before[2 + 5 * 8], after[2 + 5 * 8];
  or_sim_scalars(sim, before);
  or_gmres_result r;
  CHECK(or_sim_cycle(sim, cycles, &r) == 0);
  CHECK(r.converged);
  or_sim_scalars(sim, after);
  for (int k = 0; k < 2 + 5 * species; ++k) CHECK(isfinite(after[k]));
  /* charge is conserved by the cycle */
  for (int s = 0; s < species; ++s) CHECK(fabs(after[3 + 5 * s] - before[3 + 5 * s]) <= 1e-9 * (1 + fabs(before[3 + 5 * s])));
  /* the stages one at a time */
  CHECK(or_sim_solve(sim, &r) == 0);
  for (int s = 0; s < species; ++s) (void)or_sim_move(sim, s);
  or_sim_advance_fields(sim);
  for (int s = 0; s < species; ++s) CHECK(or_sim_sort(sim, s) == 0);
  (void)or_sim_deposit(sim);
  double energy[2 + 8];
  or_sim_energy(sim, energy);
  CHECK(isfinite(energy[0]) && isfinite(energy[1]));
  CHECK(or_sim_validation(sim) == 0);
  or_sim_destroy(sim);
}

/* A diagonal operator for GMRES alone: out_i = (2 + i % 3) in_i. */
static void diagonal(void *context, const double *in, double *out) {
  const int n = *(const int *)context;
  for (int i = 0; i < n; ++i) out[i] = (2.0 + i % 3) * in[i];
}

int main(void) {
  const double qom3[3] = {-1.0, 1.0, 0.5};
  const int ppc3[3][3] = {{2, 2, 2}, {1, 2, 2}, {0, 0, 0}};

  /* periodic box, three species (the last one empty) */
  or_grid g = make_grid(5, 4, 3, 5.0, 4.0, 3.0);
  run(&g, 3, ppc3, qom3, NULL, NULL, 0, 0, 3);

  /* one cell thick in z, TwoStream density seed on species 0 */
  or_grid thin = make_grid(16, 2, 1, 6.0, 0.75, 0.375);
  const double seed_mode[3] = {2.0 * 3.14159265358979 / 6.0, 0.25, 0.01};
  const int ppc2[2][3] = {{4, 2, 2}, {2, 2, 2}};
  run(&thin, 2, ppc2, qom3, NULL, seed_mode, 0, 0, 2);

  /* walled in y (PEC) and z (PMC), mirrored particles, Harris profile,
   * each smoothing operator */
  or_grid walled = make_grid(8, 6, 4, 12.8, 6.4, 1.6);
  walled.particle_periodic[1] = walled.particle_periodic[2] = 0;
  walled.field_faces[1][0] = walled.field_faces[1][1] = OR_FACE_PEC;
  walled.field_faces[2][0] = walled.field_faces[2][1] = OR_FACE_PMC;
  or_harris h;
  memset(&h, 0, sizeof h);
  h.density = 1;
  h.ly = 6.4;
  h.delta = 0.5;
  h.ampere_drift = 1;
  h.center_x = 6.4;
  h.center_y = 3.2;
  h.curl_delta = 0.5;
  h.sheet = 0.0195;
  h.mode = 0.001;
  h.mode_kx = 2 * 3.14159265358979 / 12.8;
  h.mode_ky = 3.14159265358979 / 6.4;
  for (int op = OR_SMOOTH_FACE7; op <= OR_SMOOTH_BINOMIAL27; ++op) run(&walled, 2, ppc2, qom3, &h, NULL, op, 2, 2);

  /* GMRES alone, restarted */
  int n = 50;
  double rhs[50], x[50];
  for (int i = 0; i < n; ++i) {
    rhs[i] = 1.0 + 0.1 * i;
    x[i] = 0.0;
  }
  or_gmres_result r = or_gmres(n, diagonal, &n, rhs, x, 5, 100, 1e-12);
  CHECK(r.converged);
  for (int i = 0; i < n; ++i) CHECK(fabs(x[i] * (2.0 + i % 3) - rhs[i]) <= 1e-9);

  /* the random stream */
  or_xorwow st;
  or_xorwow_init(&st, 12345u, 7u);
  double sum = 0.0, pair[2];
  for (int i = 0; i < 1000; ++i) sum += or_xorwow_uniform_double(&st);
  or_xorwow_normal2_double(&st, pair);
  CHECK(sum > 400.0 && sum < 600.0 && isfinite(pair[0]) && isfinite(pair[1]));

  if (failures) {
    fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  printf("oracle sanitizer run: all checks passed\n");
  return 0;
}
