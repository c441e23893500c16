/*
 * CPU oracle of the ECSIM-cycle hot path (test infrastructure only; see
 * tpic_oracle.h). Plain C11, compiled with -ffp-contract=off so that every
 * multiply and add rounds on its own: the mover kernel of the product uses the
 * same contraction-free expression order, which makes the two bit-identical.
 *
 * Reference paths are relative to the lucak17/TPIC repository root.
 */
#include "tpic_oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* OpenMP build (liboracle_omp.so, timing of the CPU baseline only): loops
 * whose iterations write disjoint outputs run in parallel and give the same
 * bits; the deposition colours z planes (even, then odd) and the solver's
 * dot products sum per-thread partials, so those results differ from the
 * serial build's in rounding. The serial build (liboracle.so) is the oracle. */
#ifdef _OPENMP
#define OR_OMP(directive) _Pragma(#directive)
#else
#define OR_OMP(directive)
#endif

#define OR_MAX_SPECIES 16
#define OR_PI 3.14159265358979323846264338327950288

/* ------------------------------------------------------------------------ */
/* Geometry: include/tpic/data/particles/particle_cell.hpp:22-60,            */
/* include/tpic/data/mass_matrix/mass_matrix_layout.hpp:28-68                 */
/* ------------------------------------------------------------------------ */

int or_node_count(const or_grid *g) { return (g->n[0] + 1) * (g->n[1] + 1) * (g->n[2] + 1); }
int or_cell_count(const or_grid *g) { return g->n[0] * g->n[1] * g->n[2]; }
int or_bin_count(const or_grid *g) { return (g->n[0] + 4) * (g->n[1] + 4) * (g->n[2] + 4); }
int or_vertex_count(const or_grid *g) { return (g->n[0] + 3) * (g->n[1] + 3) * (g->n[2] + 3); }

/* particle_cell.hpp:33-37: halo-2, x-fastest key. */
int32_t or_encode(const or_grid *g, int i, int j, int k) {
  const int ex = g->n[0] + 4, ey = g->n[1] + 4;
  return (i + 2) + ex * ((j + 2) + ey * (k + 2));
}

/* particle_cell.hpp:40-45 */
void or_decode(const or_grid *g, int32_t key, int *ijk) {
  const int ex = g->n[0] + 4, ey = g->n[1] + 4;
  const int yz = key / ex;
  ijk[0] = key % ex - 2;
  ijk[1] = yz % ey - 2;
  ijk[2] = yz / ey - 2;
}

static int interior(const or_grid *g, const int *c) {
  return c[0] >= 0 && c[0] < g->n[0] && c[1] >= 0 && c[1] < g->n[1] && c[2] >= 0 && c[2] < g->n[2];
}

static size_t node_c(const or_grid *g, int i, int j, int k) {
  return (size_t)i + (size_t)(g->n[0] + 1) * ((size_t)j + (size_t)(g->n[1] + 1) * (size_t)k);
}
static size_t cell_c(const or_grid *g, int i, int j, int k) {
  return (size_t)i + (size_t)g->n[0] * ((size_t)j + (size_t)g->n[1] * (size_t)k);
}
/* One-ghost node array, i in [-1, n+1]. */
static size_t node_g(const or_grid *g, int i, int j, int k) {
  return (size_t)(i + 1) + (size_t)(g->n[0] + 3) * ((size_t)(j + 1) + (size_t)(g->n[1] + 3) * (size_t)(k + 1));
}
static size_t node_g_count(const or_grid *g) {
  return (size_t)(g->n[0] + 3) * (size_t)(g->n[1] + 3) * (size_t)(g->n[2] + 3);
}
/* One-ghost cell array, i in [-1, n]. */
static size_t cell_g(const or_grid *g, int i, int j, int k) {
  return (size_t)(i + 1) + (size_t)(g->n[0] + 2) * ((size_t)(j + 1) + (size_t)(g->n[1] + 2) * (size_t)(k + 1));
}
static size_t cell_g_count(const or_grid *g) {
  return (size_t)(g->n[0] + 2) * (size_t)(g->n[1] + 2) * (size_t)(g->n[2] + 2);
}
/* mass_matrix_layout.hpp:36-42: vertex index of the one-node-halo layout. */
static size_t vertex(const or_grid *g, int i, int j, int k) { return node_g(g, i, j, k); }

/* ------------------------------------------------------------------------ */
/* Periodic self-route halos (the single-rank semantics of LPFx copy_ghosts  */
/* and sum_shared: include/lpfx/distributed/field/device.hpp:424-533) and    */
/* closed walls (include/tpic/ecsim/fields/field_boundary_device.hpp:26-57,  */
/* rules of include/tpic/config/boundary_conditions.hpp:68-77): a node ghost */
/* reflects through the wall node (ghost -1 = s node 1, ghost n+1 = s node  */
/* n-1), a cell ghost copies the adjacent cell (ghost -1 = s cell 0, ghost  */
/* n = s cell n-1), with the sign s of the wall, component and vector kind: */
/* PEC: E tangential odd, E normal even, B tangential even, B normal odd;    */
/* PMC the reverse. Axes are filled in x, y, z order, transverse ghosts     */
/* included, so a corner ghost carries the product of the axes' signs.       */
/* ------------------------------------------------------------------------ */

static int wall_axis(const or_grid *g, int axis) { return g->field_faces[axis][0] != OR_FACE_PERIODIC; }

static double ghost_sign(const or_grid *g, int axis, int side, int comp, int electric) {
  const int normal = comp == axis;
  const int odd = g->field_faces[axis][side] == OR_FACE_PEC ? (electric ? !normal : normal)
                                                             : (electric ? normal : !normal);
  return odd ? -1.0 : 1.0;
}

/* Rows held at zero (boundary_conditions.hpp:56-60): PEC tangential E, PMC normal E. */
static int constrained_row(const or_grid *g, int comp, int i, int j, int k) {
  const int at[3] = {i, j, k};
  for (int a = 0; a < 3; ++a) {
    if (!wall_axis(g, a)) continue;
    for (int side = 0; side < 2; ++side) {
      if (at[a] != (side ? g->n[a] : 0)) continue;
      const int pec = g->field_faces[a][side] == OR_FACE_PEC;
      if (pec ? comp != a : comp == a) return 1;
    }
  }
  return 0;
}

/* WallFaces::dual_volume_factor (include/tpic/ecsim/fields/wall_faces.hpp:27-30). */
static double dual_volume_factor(const or_grid *g, int i, int j, int k) {
  const int at[3] = {i, j, k};
  int walls = 0;
  for (int a = 0; a < 3; ++a)
    if (wall_axis(g, a)) walls += (at[a] == 0) + (at[a] == g->n[a]);
  return (double)(1u << walls);
}

/* Zero the constrained rows of a compact node vector (enqueue_constrained_rows
 * without input, field_boundary_device.hpp:100-120). */
static void zero_constrained_rows(const or_grid *g, double *v) {
  const size_t nn = (size_t)or_node_count(g);
  for (int k = 0; k <= g->n[2]; ++k)
    for (int j = 0; j <= g->n[1]; ++j)
      for (int i = 0; i <= g->n[0]; ++i)
        for (int c = 0; c < 3; ++c)
          if (constrained_row(g, c, i, j, k)) v[c * nn + node_c(g, i, j, k)] = 0.0;
}

/* Expand compact node component comp into a ghosted array and fill ghosts:
 * periodic ghost -1 holds node n-1 and ghost n+1 holds node 1. */
static void nodes_to_ghosted(const or_grid *g, int comp, const double *compact, double *out) {
  const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) out[node_g(g, i, j, k)] = compact[node_c(g, i, j, k)];
  const int wx = wall_axis(g, 0), wy = wall_axis(g, 1), wz = wall_axis(g, 2);
  const double sx0 = ghost_sign(g, 0, 0, comp, 1), sx1 = ghost_sign(g, 0, 1, comp, 1);
  const double sy0 = ghost_sign(g, 1, 0, comp, 1), sy1 = ghost_sign(g, 1, 1, comp, 1);
  const double sz0 = ghost_sign(g, 2, 0, comp, 1), sz1 = ghost_sign(g, 2, 1, comp, 1);
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j) {
      out[node_g(g, -1, j, k)] = wx ? sx0 * out[node_g(g, 1, j, k)] : out[node_g(g, nx - 1, j, k)];
      out[node_g(g, nx + 1, j, k)] = wx ? sx1 * out[node_g(g, nx - 1, j, k)] : out[node_g(g, 1, j, k)];
    }
  for (int k = 0; k <= nz; ++k)
    for (int i = -1; i <= nx + 1; ++i) {
      out[node_g(g, i, -1, k)] = wy ? sy0 * out[node_g(g, i, 1, k)] : out[node_g(g, i, ny - 1, k)];
      out[node_g(g, i, ny + 1, k)] = wy ? sy1 * out[node_g(g, i, ny - 1, k)] : out[node_g(g, i, 1, k)];
    }
  for (int j = -1; j <= ny + 1; ++j)
    for (int i = -1; i <= nx + 1; ++i) {
      out[node_g(g, i, j, -1)] = wz ? sz0 * out[node_g(g, i, j, 1)] : out[node_g(g, i, j, nz - 1)];
      out[node_g(g, i, j, nz + 1)] = wz ? sz1 * out[node_g(g, i, j, nz - 1)] : out[node_g(g, i, j, 1)];
    }
}

/* Cells of B-like component comp: periodic ghost -1 holds cell n-1 and ghost
 * n holds cell 0. */
static void cells_to_ghosted(const or_grid *g, int comp, const double *compact, double *out) {
  const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) out[cell_g(g, i, j, k)] = compact[cell_c(g, i, j, k)];
  const int wx = wall_axis(g, 0), wy = wall_axis(g, 1), wz = wall_axis(g, 2);
  const double sx0 = ghost_sign(g, 0, 0, comp, 0), sx1 = ghost_sign(g, 0, 1, comp, 0);
  const double sy0 = ghost_sign(g, 1, 0, comp, 0), sy1 = ghost_sign(g, 1, 1, comp, 0);
  const double sz0 = ghost_sign(g, 2, 0, comp, 0), sz1 = ghost_sign(g, 2, 1, comp, 0);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      out[cell_g(g, -1, j, k)] = wx ? sx0 * out[cell_g(g, 0, j, k)] : out[cell_g(g, nx - 1, j, k)];
      out[cell_g(g, nx, j, k)] = wx ? sx1 * out[cell_g(g, nx - 1, j, k)] : out[cell_g(g, 0, j, k)];
    }
  for (int k = 0; k < nz; ++k)
    for (int i = -1; i <= nx; ++i) {
      out[cell_g(g, i, -1, k)] = wy ? sy0 * out[cell_g(g, i, 0, k)] : out[cell_g(g, i, ny - 1, k)];
      out[cell_g(g, i, ny, k)] = wy ? sy1 * out[cell_g(g, i, ny - 1, k)] : out[cell_g(g, i, 0, k)];
    }
  for (int j = -1; j <= ny; ++j)
    for (int i = -1; i <= nx; ++i) {
      out[cell_g(g, i, j, -1)] = wz ? sz0 * out[cell_g(g, i, j, 0)] : out[cell_g(g, i, j, nz - 1)];
      out[cell_g(g, i, j, nz)] = wz ? sz1 * out[cell_g(g, i, j, nz - 1)] : out[cell_g(g, i, j, 0)];
    }
}

/* ------------------------------------------------------------------------ */
/* CIC and response: include/tpic/ecsim/particles/cic.hpp:24-36,            */
/* include/tpic/ecsim/particles/particle_response.hpp:17-34                 */
/* ------------------------------------------------------------------------ */

void or_cic_weights(double x, double y, double z, double w[8]) {
  const double cx = 1.0 - x, cy = 1.0 - y, cz = 1.0 - z;
  w[0] = cx * cy * cz;
  w[1] = x * cy * cz;
  w[2] = cx * y * cz;
  w[3] = x * y * cz;
  w[4] = cx * cy * z;
  w[5] = x * cy * z;
  w[6] = cx * y * z;
  w[7] = x * y * z;
}

void or_response(double bx, double by, double bz, double beta_over_c, double a[9], double *norm_square) {
  const double x = beta_over_c * bx, y = beta_over_c * by, z = beta_over_c * bz;
  const double n2 = x * x + y * y + z * z;
  const double inverse = 1.0 / (1.0 + n2);
  a[0] = (1.0 + x * x) * inverse;
  a[4] = (1.0 + y * y) * inverse;
  a[8] = (1.0 + z * z) * inverse;
  const double xy = x * y * inverse, xz = x * z * inverse, yz = y * z * inverse;
  const double sx = x * inverse, sy = y * inverse, sz = z * inverse;
  a[1] = xy + sz;
  a[2] = xz - sy;
  a[3] = xy - sz;
  a[5] = yz + sx;
  a[6] = xz + sy;
  a[7] = yz - sx;
  if (norm_square) *norm_square = n2;
}

/* Canonical offsets, mass_matrix_layout.hpp:17-25; pair table,
 * mass_matrix_stencil.hpp:15-36. */
static const int stencil_offset[14][3] = {
    {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, -1, 0}, {1, 0, 1},
    {1, 0, -1}, {0, 1, 1}, {0, 1, -1}, {1, -1, 1}, {1, 1, -1}, {-1, 1, 1}, {1, 1, 1}};

int or_mass_pair_stencil(int a, int b) {
  const int d[3] = {((b & 1) - (a & 1)), ((b >> 1) & 1) - ((a >> 1) & 1), ((b >> 2) & 1) - ((a >> 2) & 1)};
  for (int s = 0; s < 14; ++s)
    if (stencil_offset[s][0] == d[0] && stencil_offset[s][1] == d[1] && stencil_offset[s][2] == d[2]) return s;
  return -1;
}

static int in_unit_cell(double x, double y, double z) {
  return x >= 0.0 && x < 1.0 && y >= 0.0 && y < 1.0 && z >= 0.0 && z < 1.0;
}

static int finite_value(double v) { return v >= -DBL_MAX && v <= DBL_MAX; }

/* interpolate_cic (cic.hpp:61-73): corners x-fastest, value += w * f. */
static double gather(const or_grid *g, const double *component, const int *cell, const double w[8]) {
  double value = 0.0;
  for (int a = 0; a < 8; ++a) {
    const size_t n = node_c(g, cell[0] + (a & 1), cell[1] + ((a >> 1) & 1), cell[2] + ((a >> 2) & 1));
    value += w[a] * component[n];
  }
  return value;
}

/* ------------------------------------------------------------------------ */
/* Mover: detail::push_and_classify_particle                                 */
/* (include/tpic/ecsim/particles/particle_mover_device.hpp:35-122) with the  */
/* whole-domain arm of communication::classify_particle_point                */
/* (include/tpic/communication/particles/particle_boundaries.hpp:111-165).   */
/* ------------------------------------------------------------------------ */

static int64_t positive_modulo(int64_t cell, int64_t extent) {
  const int64_t r = cell % extent;
  return r < 0 ? r + extent : r;
}

unsigned or_move(const or_grid *g, const or_mover_params *p, or_particles *s,
                 const double *e_nodes, const double *b_nodes) {
  const size_t nodes = (size_t)or_node_count(g);
  const int bins = or_bin_count(g);
  unsigned flags = 0;
  OR_OMP(omp parallel for reduction(|:flags) schedule(static))
  for (uint32_t index = 0; index < s->count; ++index) {
    const int32_t key = s->cell[index];
    int cell[3];
    if (key < 0 || key >= bins) { flags |= OR_INVALID_KEY; continue; }
    or_decode(g, key, cell);
    if (!interior(g, cell)) { flags |= OR_INVALID_KEY; continue; }
    if (!finite_value(s->q[index])) { flags |= OR_INVALID_NONFINITE; continue; }
    const double fx = s->rx[index], fy = s->ry[index], fz = s->rz[index];
    if (!in_unit_cell(fx, fy, fz)) { flags |= OR_INVALID_FRACTION; continue; }
    double w[8];
    or_cic_weights(fx, fy, fz, w);
    const double bx = gather(g, b_nodes, cell, w);
    const double by = gather(g, b_nodes + nodes, cell, w);
    const double bz = gather(g, b_nodes + 2 * nodes, cell, w);
    const double ex = gather(g, e_nodes, cell, w);
    const double ey = gather(g, e_nodes + nodes, cell, w);
    const double ez = gather(g, e_nodes + 2 * nodes, cell, w);
    double a[9];
    or_response(bx, by, bz, p->beta_over_c, a, NULL);
    const double u = s->u[index], v = s->v[index], ww = s->w[index];
    const double hx = a[0] * u + a[1] * v + a[2] * ww;
    const double hy = a[3] * u + a[4] * v + a[5] * ww;
    const double hz = a[6] * u + a[7] * v + a[8] * ww;
    const double aex = a[0] * ex + a[1] * ey + a[2] * ez;
    const double aey = a[3] * ex + a[4] * ey + a[5] * ez;
    const double aez = a[6] * ex + a[7] * ey + a[8] * ez;
    const double px = hx + p->beta * aex;
    const double py = hy + p->beta * aey;
    const double pz = hz + p->beta * aez;
    int64_t pc[3] = {cell[0], cell[1], cell[2]};
    double frac[3] = {fx, fy, fz};
    double vel[3] = {2.0 * px - u, 2.0 * py - v, 2.0 * pz - ww};
    int bad = 0;
    for (int axis = 0; axis < 3 && !bad; ++axis) {
      if (!finite_value(vel[axis])) { flags |= OR_INVALID_NONFINITE; bad = 1; break; }
      const double displacement = p->dt * vel[axis] * p->inverse_spacing[axis];
      const double moved = frac[axis] + displacement;
      const double shift = floor(moved);
      if (!(shift >= (double)INT32_MIN && shift <= (double)INT32_MAX)) { flags |= OR_INVALID_FLIGHT; bad = 1; break; }
      pc[axis] += (int64_t)shift;
      frac[axis] = moved - shift;
      if (frac[axis] == 1.0) { ++pc[axis]; frac[axis] = 0.0; }
    }
    if (bad) continue;
    /* Whole-domain owner: fold any finite flight (particle_boundaries.hpp:116-160). */
    for (int axis = 0; axis < 3; ++axis) {
      const int64_t extent = g->n[axis];
      int64_t global = pc[axis];
      if (g->particle_periodic[axis]) {
        if (global < 0 || global >= extent) global = positive_modulo(global, extent);
        pc[axis] = global;
      } else if (global < 0 || global >= extent) {
        global = positive_modulo(global, 2 * extent);
        if (global >= extent) {
          global = 2 * extent - global - 1;
          frac[axis] = 1.0 - frac[axis];
          vel[axis] = -vel[axis];
          if (frac[axis] == 1.0) {
            if (global + 1 < extent) { ++global; frac[axis] = 0.0; }
            else frac[axis] = 1.0 - DBL_EPSILON / 2.0;
          }
        }
        pc[axis] = global;
      }
    }
    s->cell[index] = or_encode(g, (int)pc[0], (int)pc[1], (int)pc[2]);
    s->rx[index] = frac[0];
    s->ry[index] = frac[1];
    s->rz[index] = frac[2];
    s->u[index] = vel[0];
    s->v[index] = vel[1];
    s->w[index] = vel[2];
  }
  return flags;
}

/* ------------------------------------------------------------------------ */
/* Stable counting sort: lpfx::PreparedSerialParticleSort::run              */
/* (include/lpfx/algorithm/particle/host/sort.hpp:210-305). permutation is  */
/* the scatter form of the device sort (destination of each source,         */
/* include/lpfx/algorithm/particle/device/cuda_hip/sort.hpp:122-137).       */
/* ------------------------------------------------------------------------ */

int or_sort(const or_grid *g, const or_particles *in, or_particles *out, uint32_t *permutation,
            uint32_t *counts, uint32_t *offsets) {
  const int bins = or_bin_count(g);
  const uint32_t n = in->count;
  uint32_t *cursor = (uint32_t *)calloc((size_t)bins, sizeof(uint32_t));
  if (!cursor) return -4;
  for (uint32_t i = 0; i < n; ++i) {
    const int32_t key = in->cell[i];
    if (key < 0 || key >= bins) { free(cursor); return -1; }
    ++cursor[key];
  }
  uint32_t prefix = 0;
  for (int b = 0; b < bins; ++b) {
    const uint32_t c = cursor[b];
    counts[b] = c;
    offsets[b] = prefix;
    cursor[b] = prefix;
    prefix += c;
  }
  offsets[bins] = prefix;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t d = cursor[in->cell[i]]++;
    if (permutation) permutation[i] = d;
    out->rx[d] = in->rx[i];
    out->ry[d] = in->ry[i];
    out->rz[d] = in->rz[i];
    out->u[d] = in->u[i];
    out->v[d] = in->v[i];
    out->w[d] = in->w[i];
    out->q[d] = in->q[i];
    out->cell[d] = in->cell[i];
    out->id[d] = in->id[i];
  }
  out->count = n;
  free(cursor);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Sorted current deposition: detail::evaluate_current_particle and          */
/* deposit_product (include/tpic/ecsim/deposition/current_deposition_device.hpp:44-104), */
/* summed per cell and corner as kernels::sorted_current_kernel (:168-264)   */
/* does, then added to the node. Writes raw (unfolded) interior nodes.       */
/* ------------------------------------------------------------------------ */

/* One bin of the current deposition: the cell's 8 corner sums, in stored
 * order, flushed into the node field. */
static unsigned deposit_current_bin(const or_grid *g, const or_particles *s, const uint32_t *counts,
                                    const uint32_t *offsets, const double *b_nodes, double inverse_volume,
                                    double beta_over_c, double *current, int bin) {
  const size_t nodes = (size_t)or_node_count(g);
  unsigned flags = 0;
  if (counts[bin] == 0) return 0;
  int cell[3];
  or_decode(g, bin, cell);
  if (!interior(g, cell)) return OR_INVALID_HALO;
  double sum[8][3];
  memset(sum, 0, sizeof sum);
  for (uint32_t o = 0; o < counts[bin]; ++o) {
    const uint32_t index = offsets[bin] + o;
    if (s->cell[index] != bin) { flags |= OR_INVALID_KEY; continue; }
    const double fx = s->rx[index], fy = s->ry[index], fz = s->rz[index];
    if (!in_unit_cell(fx, fy, fz)) { flags |= OR_INVALID_FRACTION; continue; }
    double w[8], a[9];
    or_cic_weights(fx, fy, fz, w);
    const double bx = gather(g, b_nodes, cell, w);
    const double by = gather(g, b_nodes + nodes, cell, w);
    const double bz = gather(g, b_nodes + 2 * nodes, cell, w);
    or_response(bx, by, bz, beta_over_c, a, NULL);
    const double u = s->u[index], v = s->v[index], ww = s->w[index];
    const double hx = a[0] * u + a[1] * v + a[2] * ww;
    const double hy = a[3] * u + a[4] * v + a[5] * ww;
    const double hz = a[6] * u + a[7] * v + a[8] * ww;
    const double q = s->q[index] * inverse_volume;
    const double r[3] = {q * hx, q * hy, q * hz};
    if (!finite_value(r[0]) || !finite_value(r[1]) || !finite_value(r[2])) {
      flags |= OR_INVALID_NONFINITE;
      continue;
    }
    for (int c = 0; c < 8; ++c)
      for (int x = 0; x < 3; ++x) sum[c][x] += r[x] * w[c];
  }
  for (int c = 0; c < 8; ++c) {
    const size_t n = node_c(g, cell[0] + (c & 1), cell[1] + ((c >> 1) & 1), cell[2] + ((c >> 2) & 1));
    for (int x = 0; x < 3; ++x) current[x * nodes + n] += sum[c][x];
  }
  return flags;
}

/* Bins in ascending key order; the OpenMP build runs the z planes of one
 * parity at a time in parallel (a cell of plane k writes node planes k and
 * k + 1 only, so planes two apart never share a node). */
#ifdef _OPENMP
#define OR_FOR_BINS(g, flags, call)                                                      \
  do {                                                                                   \
    const int plane_bins_ = ((g)->n[0] + 4) * ((g)->n[1] + 4), planes_ = (g)->n[2] + 4; \
    for (int parity_ = 0; parity_ < 2; ++parity_) {                                      \
      OR_OMP(omp parallel for reduction(|:flags) schedule(dynamic, 1))                   \
      for (int plane_ = parity_; plane_ < planes_; plane_ += 2)                          \
        for (int bin = plane_ * plane_bins_; bin < (plane_ + 1) * plane_bins_; ++bin)    \
          flags |= (call);                                                               \
    }                                                                                    \
  } while (0)
#else
#define OR_FOR_BINS(g, flags, call)                                          \
  do {                                                                       \
    for (int bin = 0, bins_ = or_bin_count(g); bin < bins_; ++bin) flags |= (call); \
  } while (0)
#endif

unsigned or_deposit_current(const or_grid *g, const or_particles *s, const uint32_t *counts,
                            const uint32_t *offsets, const double *b_nodes, double inverse_volume,
                            double beta_over_c, double *current) {
  unsigned flags = 0;
  OR_FOR_BINS(g, flags, deposit_current_bin(g, s, counts, offsets, b_nodes, inverse_volume, beta_over_c,
                                            current, bin));
  return flags;
}

/* ------------------------------------------------------------------------ */
/* Sorted mass deposition: the host FP64 cell-sum oracle of the reference's  */
/* own test (tests/ecsim/sorted_mass_device.cpp:111-182), with the sample of */
/* ecsim::sample_mass_particle (mass_matrix_sorted_cell_device.hpp:79-119):  */
/* K[a,b,c] += (alpha_c * beta_over_c * q) * (W_a W_b), 36 canonical pairs    */
/* flushed per cell into the 126-scalar node records.                         */
/* ------------------------------------------------------------------------ */

static unsigned deposit_mass_bin(const or_grid *g, const or_particles *s, const uint32_t *counts,
                                 const uint32_t *offsets, const double *b_nodes, double beta_over_c, double *mass,
                                 int bin) {
  const size_t nodes = (size_t)or_node_count(g);
  unsigned flags = 0;
  if (counts[bin] == 0) return 0;
  int cell[3];
  or_decode(g, bin, cell);
  if (!interior(g, cell)) return OR_INVALID_HALO;
  double tensor[64 * 9];
  memset(tensor, 0, sizeof tensor);
  for (uint32_t o = 0; o < counts[bin]; ++o) {
    const uint32_t index = offsets[bin] + o;
    if (s->cell[index] != bin) { flags |= OR_INVALID_KEY; continue; }
    const double fx = s->rx[index], fy = s->ry[index], fz = s->rz[index];
    if (!in_unit_cell(fx, fy, fz)) { flags |= OR_INVALID_FRACTION; continue; }
    double w[8], a[9], n2;
    or_cic_weights(fx, fy, fz, w);
    const double bx = gather(g, b_nodes, cell, w);
    const double by = gather(g, b_nodes + nodes, cell, w);
    const double bz = gather(g, b_nodes + 2 * nodes, cell, w);
    or_response(bx, by, bz, beta_over_c, a, &n2);
    const double scale = beta_over_c * s->q[index];
    if (!finite_value(n2) || !finite_value(scale)) { flags |= OR_INVALID_NONFINITE; continue; }
    for (int ca = 0; ca < 8; ++ca)
      for (int cb = 0; cb < 8; ++cb)
        for (int c = 0; c < 9; ++c) tensor[(ca * 8 + cb) * 9 + c] += (a[c] * scale) * (w[ca] * w[cb]);
  }
  for (int ca = 0; ca < 8; ++ca) {
    const size_t v = vertex(g, cell[0] + (ca & 1), cell[1] + ((ca >> 1) & 1), cell[2] + ((ca >> 2) & 1));
    for (int cb = 0; cb < 8; ++cb) {
      const int st = or_mass_pair_stencil(ca, cb);
      if (st < 0) continue;
      for (int c = 0; c < 9; ++c) {
        const double value = tensor[(ca * 8 + cb) * 9 + c];
        if (!finite_value(value)) { flags |= OR_INVALID_NONFINITE; continue; }
        mass[v * 126 + (size_t)st * 9 + c] += value;
      }
    }
  }
  return flags;
}

unsigned or_deposit_mass(const or_grid *g, const or_particles *s, const uint32_t *counts,
                         const uint32_t *offsets, const double *b_nodes, double beta_over_c,
                         double *mass) {
  unsigned flags = 0;
  OR_FOR_BINS(g, flags, deposit_mass_bin(g, s, counts, offsets, b_nodes, beta_over_c, mass, bin));
  return flags;
}

/* ------------------------------------------------------------------------ */
/* Assembly on one periodic rank. Current: sum_shared of the "J.add" route   */
/* (include/tpic/ecsim/deposition/current_assembly_device.hpp:33-52) makes   */
/* both copies of a periodic node hold their sum; the dual-volume factor is  */
/* 1 without walls. Mass: additive then ghost phase of MassHaloStorage        */
/* (include/tpic/communication/mass_halo/mass_halo_device.hpp:162-399).      */
/* ------------------------------------------------------------------------ */

/* One node component: both copies of a periodic node hold their sum, then
 * the dual-volume scale. */
static void fold_node_component(const or_grid *g, double *f) {
  const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
  if (!wall_axis(g, 0))
    for (int k = 0; k <= nz; ++k)
      for (int j = 0; j <= ny; ++j) {
        const double s = f[node_c(g, 0, j, k)] + f[node_c(g, nx, j, k)];
        f[node_c(g, 0, j, k)] = f[node_c(g, nx, j, k)] = s;
      }
  if (!wall_axis(g, 1))
    for (int k = 0; k <= nz; ++k)
      for (int i = 0; i <= nx; ++i) {
        const double s = f[node_c(g, i, 0, k)] + f[node_c(g, i, ny, k)];
        f[node_c(g, i, 0, k)] = f[node_c(g, i, ny, k)] = s;
      }
  if (!wall_axis(g, 2))
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        const double s = f[node_c(g, i, j, 0)] + f[node_c(g, i, j, nz)];
        f[node_c(g, i, j, 0)] = f[node_c(g, i, j, nz)] = s;
      }
  /* enqueue_scale_moment_dual_volume (moment_field_device.hpp:84-96,137-154). */
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) f[node_c(g, i, j, k)] *= dual_volume_factor(g, i, j, k);
}

void or_fold_current(const or_grid *g, double *current) {
  const size_t nodes = (size_t)or_node_count(g);
  for (int c = 0; c < 3; ++c) fold_node_component(g, current + c * nodes);
}

/* ------------------------------------------------------------------------ */
/* Output moments of one species at nodes: deposit_output_moments_kernel     */
/* (include/tpic/diagnostics/output_moment_deposition_device.hpp:35-121),    */
/* components in output_moment_quantities order (rho, Jx, Jy, Jz, PXX, PXY,  */
/* PXZ, PYY, PYZ, PZZ; include/tpic/data/output/output_moment_quantities.hpp */
/* :26-34): charge = q / V, J = charge v, P_ab = (charge / qom) v_a v_b,     */
/* each times the CIC weight of the corner; then the fold and dual-volume    */
/* scale of the capture (output_actions_device.hpp:130-146). Unselected      */
/* components stay zero; an invalid particle is skipped and flagged.         */
/* ------------------------------------------------------------------------ */

unsigned or_deposit_moments(const or_grid *g, const or_particles *s, double q_over_m, double inverse_volume,
                            unsigned selected, double *out) {
  const size_t nodes = (size_t)or_node_count(g);
  const double inverse_qom = 1.0 / q_over_m;
  unsigned flags = 0;
  for (uint32_t index = 0; index < s->count; ++index) {
    const int32_t key = s->cell[index];
    int cell[3];
    if (key < 0 || key >= or_bin_count(g)) { flags |= OR_INVALID_KEY; continue; }
    or_decode(g, key, cell);
    if (!interior(g, cell)) { flags |= OR_INVALID_KEY; continue; }
    const double fx = s->rx[index], fy = s->ry[index], fz = s->rz[index];
    if (!in_unit_cell(fx, fy, fz)) { flags |= OR_INVALID_FRACTION; continue; }
    const double u = (selected & 0x72u) ? s->u[index] : 0.0; /* Jx PXX PXY PXZ */
    const double v = (selected & 0x1A4u) ? s->v[index] : 0.0; /* Jy PXY PYY PYZ */
    const double w = (selected & 0x348u) ? s->w[index] : 0.0; /* Jz PXZ PYZ PZZ */
    const double charge = s->q[index] * inverse_volume;
    const double mass = charge * inverse_qom;
    const double amount[10] = {charge, charge * u, charge * v, charge * w, mass * u * u,
                               mass * u * v, mass * u * w, mass * v * v, mass * v * w, mass * w * w};
    int finite = 1;
    for (int c = 0; c < 10; ++c)
      if ((selected >> c) & 1u) finite = finite && finite_value(amount[c]);
    if (!finite) { flags |= OR_INVALID_NONFINITE; continue; }
    double wt[8];
    or_cic_weights(fx, fy, fz, wt);
    for (int a = 0; a < 8; ++a) {
      const size_t n = node_c(g, cell[0] + (a & 1), cell[1] + ((a >> 1) & 1), cell[2] + ((a >> 2) & 1));
      for (int c = 0; c < 10; ++c)
        if ((selected >> c) & 1u) out[c * nodes + n] += amount[c] * wt[a];
    }
  }
  for (int c = 0; c < 10; ++c) fold_node_component(g, out + c * nodes);
  return flags;
}

static int wrap_vertex(int i, int n) { return i < 0 ? n - 1 : (i > n ? 1 : i); }

void or_fold_mass(const or_grid *g, double *mass) {
  const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
  if (!wall_axis(g, 0))
    for (int k = 0; k <= nz; ++k)
      for (int j = 0; j <= ny; ++j) {
        double *a = mass + vertex(g, 0, j, k) * 126, *b = mass + vertex(g, nx, j, k) * 126;
        for (int e = 0; e < 126; ++e) a[e] = b[e] = a[e] + b[e];
      }
  if (!wall_axis(g, 1))
    for (int k = 0; k <= nz; ++k)
      for (int i = 0; i <= nx; ++i) {
        double *a = mass + vertex(g, i, 0, k) * 126, *b = mass + vertex(g, i, ny, k) * 126;
        for (int e = 0; e < 126; ++e) a[e] = b[e] = a[e] + b[e];
      }
  if (!wall_axis(g, 2))
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        double *a = mass + vertex(g, i, j, 0) * 126, *b = mass + vertex(g, i, j, nz) * 126;
        for (int e = 0; e < 126; ++e) a[e] = b[e] = a[e] + b[e];
      }
  /* Ghost phase: a ghost record beyond a periodic face is a copy of its
   * periodic interior image; beyond a wall it holds no particle's moment
   * (no peer routes one there) and stays zero. */
  const int extent[3] = {nx, ny, nz};
  for (int k = -1; k <= nz + 1; ++k)
    for (int j = -1; j <= ny + 1; ++j)
      for (int i = -1; i <= nx + 1; ++i) {
        const int at[3] = {i, j, k};
        int ghost = 0, walled = 0;
        for (int a = 0; a < 3; ++a)
          if (at[a] < 0 || at[a] > extent[a]) {
            ghost = 1;
            walled |= wall_axis(g, a);
          }
        if (!ghost) continue;
        double *dst = mass + vertex(g, i, j, k) * 126;
        if (walled) {
          memset(dst, 0, 126 * sizeof(double));
          continue;
        }
        const double *src = mass + vertex(g, wrap_vertex(i, nx), wrap_vertex(j, ny), wrap_vertex(k, nz)) * 126;
        memcpy(dst, src, 126 * sizeof(double));
      }
}

/* ------------------------------------------------------------------------ */
/* Curls and cell-to-node: include/tpic/ecsim/fields/node_cell_curl.hpp:52-119, */
/* include/tpic/ecsim/fields/cell_to_node.hpp:19-32.                         */
/* ------------------------------------------------------------------------ */

typedef struct { double base, x, y, z, xy, xz, yz, xyz; } samples8;

/* Node sources extend forwards; difference = neighbor - anchor. */
static double d_node(double anchor, double neighbor) { return neighbor - anchor; }
/* Cell sources extend backwards; difference = anchor - neighbor. */
static double d_cell(double anchor, double neighbor) { return anchor - neighbor; }

static double derivative(const samples8 *s, int axis, double inverse, int node_source) {
  double (*d)(double, double) = node_source ? d_node : d_cell;
  double sum;
  if (axis == 0) sum = d(s->base, s->x) + d(s->z, s->xz) + d(s->y, s->xy) + d(s->yz, s->xyz);
  else if (axis == 1) sum = d(s->base, s->y) + d(s->z, s->yz) + d(s->x, s->xy) + d(s->xz, s->xyz);
  else sum = d(s->base, s->z) + d(s->x, s->xz) + d(s->y, s->yz) + d(s->xy, s->xyz);
  return 0.25 * inverse * sum;
}

static void inverse_spacing(const or_grid *g, double inv[3]) {
  for (int a = 0; a < 3; ++a) inv[a] = 1.0 / g->spacing[a];
}

/* Node (compact) -> interior cells. */
void or_curl_n2c(const or_grid *g, const double *nodes, double *cells) {
  const size_t nn = (size_t)or_node_count(g), nc = (size_t)or_cell_count(g);
  double inv[3];
  inverse_spacing(g, inv);
  OR_OMP(omp parallel for schedule(static))
  for (int k = 0; k < g->n[2]; ++k)
    for (int j = 0; j < g->n[1]; ++j)
      for (int i = 0; i < g->n[0]; ++i) {
        samples8 s[3];
        for (int c = 0; c < 3; ++c) {
          const double *f = nodes + c * nn;
          s[c].base = f[node_c(g, i, j, k)];
          s[c].x = f[node_c(g, i + 1, j, k)];
          s[c].y = f[node_c(g, i, j + 1, k)];
          s[c].z = f[node_c(g, i, j, k + 1)];
          s[c].xy = f[node_c(g, i + 1, j + 1, k)];
          s[c].xz = f[node_c(g, i + 1, j, k + 1)];
          s[c].yz = f[node_c(g, i, j + 1, k + 1)];
          s[c].xyz = f[node_c(g, i + 1, j + 1, k + 1)];
        }
        const size_t n = cell_c(g, i, j, k);
        cells[n] = derivative(&s[2], 1, inv[1], 1) - derivative(&s[1], 2, inv[2], 1);
        cells[nc + n] = derivative(&s[0], 2, inv[2], 1) - derivative(&s[2], 0, inv[0], 1);
        cells[2 * nc + n] = derivative(&s[1], 0, inv[0], 1) - derivative(&s[0], 1, inv[1], 1);
      }
}

/* Cells (compact, periodic ghosts) -> interior nodes. */
void or_curl_c2n(const or_grid *g, const double *cells, double *nodes) {
  const size_t nn = (size_t)or_node_count(g), nc = (size_t)or_cell_count(g);
  double *ghosted = (double *)malloc(3 * cell_g_count(g) * sizeof(double));
  const size_t cg = cell_g_count(g);
  for (int c = 0; c < 3; ++c) cells_to_ghosted(g, c, cells + c * nc, ghosted + c * cg);
  double inv[3];
  inverse_spacing(g, inv);
  OR_OMP(omp parallel for schedule(static))
  for (int k = 0; k <= g->n[2]; ++k)
    for (int j = 0; j <= g->n[1]; ++j)
      for (int i = 0; i <= g->n[0]; ++i) {
        samples8 s[3];
        for (int c = 0; c < 3; ++c) {
          const double *f = ghosted + c * cg;
          s[c].base = f[cell_g(g, i, j, k)];
          s[c].x = f[cell_g(g, i - 1, j, k)];
          s[c].y = f[cell_g(g, i, j - 1, k)];
          s[c].z = f[cell_g(g, i, j, k - 1)];
          s[c].xy = f[cell_g(g, i - 1, j - 1, k)];
          s[c].xz = f[cell_g(g, i - 1, j, k - 1)];
          s[c].yz = f[cell_g(g, i, j - 1, k - 1)];
          s[c].xyz = f[cell_g(g, i - 1, j - 1, k - 1)];
        }
        const size_t n = node_c(g, i, j, k);
        nodes[n] = derivative(&s[2], 1, inv[1], 0) - derivative(&s[1], 2, inv[2], 0);
        nodes[nn + n] = derivative(&s[0], 2, inv[2], 0) - derivative(&s[2], 0, inv[0], 0);
        nodes[2 * nn + n] = derivative(&s[1], 0, inv[0], 0) - derivative(&s[0], 1, inv[1], 0);
      }
  free(ghosted);
}

void or_cell_to_node(const or_grid *g, const double *cells, double *nodes) {
  const size_t nn = (size_t)or_node_count(g), nc = (size_t)or_cell_count(g), cg = cell_g_count(g);
  double *f = (double *)malloc(cg * sizeof(double));
  for (int c = 0; c < 3; ++c) {
    cells_to_ghosted(g, c, cells + c * nc, f);
    for (int k = 0; k <= g->n[2]; ++k)
      for (int j = 0; j <= g->n[1]; ++j)
        for (int i = 0; i <= g->n[0]; ++i)
          nodes[c * nn + node_c(g, i, j, k)] =
              (f[cell_g(g, i - 1, j - 1, k - 1)] + f[cell_g(g, i, j - 1, k - 1)] +
               f[cell_g(g, i - 1, j, k - 1)] + f[cell_g(g, i, j, k - 1)] +
               f[cell_g(g, i - 1, j - 1, k)] + f[cell_g(g, i, j - 1, k)] +
               f[cell_g(g, i - 1, j, k)] + f[cell_g(g, i, j, k)]) * 0.125;
  }
  free(f);
}

/* ------------------------------------------------------------------------ */
/* Mass apply: ecsim::apply_mass_matrix_node                                 */
/* (include/tpic/ecsim/mass_matrix/mass_matrix_stencil.hpp:106-149).         */
/* ------------------------------------------------------------------------ */

static void block_product(const double *b, double vx, double vy, double vz, double *x, double *y, double *z) {
  *x = vx * b[0] + vy * b[1] + vz * b[2];
  *y = vx * b[3] + vy * b[4] + vz * b[5];
  *z = vx * b[6] + vy * b[7] + vz * b[8];
}

static void mass_apply_ghosted(const or_grid *g, const double *mass, const double *vg, double inverse_volume,
                               double *out) {
  const size_t nn = (size_t)or_node_count(g), ng = node_g_count(g);

  const double *vx = vg, *vy = vg + ng, *vz = vg + 2 * ng;
  OR_OMP(omp parallel for schedule(static))
  for (int k = 0; k <= g->n[2]; ++k)
    for (int j = 0; j <= g->n[1]; ++j)
      for (int i = 0; i <= g->n[0]; ++i) {
        const size_t v0 = vertex(g, i, j, k);
        double x, y, z;
        size_t at = node_g(g, i, j, k);
        block_product(mass + v0 * 126, vx[at], vy[at], vz[at], &x, &y, &z);
        for (int s = 1; s < 14; ++s) {
          const int *o = stencil_offset[s];
          double bx, by, bz;
          at = node_g(g, i + o[0], j + o[1], k + o[2]);
          block_product(mass + v0 * 126 + (size_t)s * 9, vx[at], vy[at], vz[at], &bx, &by, &bz);
          x += bx; y += by; z += bz;
          const int ni = i - o[0], nj = j - o[1], nk = k - o[2];
          at = node_g(g, ni, nj, nk);
          block_product(mass + vertex(g, ni, nj, nk) * 126 + (size_t)s * 9, vx[at], vy[at], vz[at], &bx, &by, &bz);
          x += bx; y += by; z += bz;
        }
        /* mass_matrix_apply_device.hpp:34-35: row scale (1/V) f_g. */
        const double row = inverse_volume * dual_volume_factor(g, i, j, k);
        x *= row; y *= row; z *= row;
        const size_t n = node_c(g, i, j, k);
        out[n] = x;
        out[nn + n] = y;
        out[2 * nn + n] = z;
      }
}

void or_mass_apply(const or_grid *g, const double *mass, const double *v, double inverse_volume, double *out) {
  const size_t nn = (size_t)or_node_count(g), ng = node_g_count(g);
  double *vg = (double *)malloc(3 * ng * sizeof(double));
  for (int c = 0; c < 3; ++c) nodes_to_ghosted(g, c, v + c * nn, vg + c * ng);
  mass_apply_ghosted(g, mass, vg, inverse_volume, out);
  free(vg);
}

/* ------------------------------------------------------------------------ */
/* Maxwell operator: MaxwellOperator::complete_apply_unsmoothed and           */
/* combine_image (include/tpic/ecsim/maxwell/maxwell_device.hpp:353-433):    */
/* image = 1*v + (delta*delta)*curl_C2N(curl_N2C v) + k*M v; and complete_rhs */
/* (:152-171): rhs = 1*E + delta*curl_C2N(B) - k*J. Left folds.              */
/* ------------------------------------------------------------------------ */

void or_maxwell_apply(const or_grid *g, const double *mass, double delta, double current_factor,
                      double inverse_volume, const double *v_in, double *out) {
  const size_t nn = (size_t)or_node_count(g), nc = (size_t)or_cell_count(g);
  /* begin_apply_input (maxwell_device.hpp:270-272) projects the input. */
  double *v = (double *)malloc(3 * nn * sizeof(double));
  memcpy(v, v_in, 3 * nn * sizeof(double));
  zero_constrained_rows(g, v);
  double *cells = (double *)malloc(3 * nc * sizeof(double));
  double *curl = (double *)malloc(3 * nn * sizeof(double));
  double *m = (double *)malloc(3 * nn * sizeof(double));
  or_curl_n2c(g, v, cells);
  or_curl_c2n(g, cells, curl);
  or_mass_apply(g, mass, v, inverse_volume, m);
  const double d2 = delta * delta;
  for (size_t n = 0; n < 3 * nn; ++n) out[n] = 1.0 * v[n] + d2 * curl[n] + current_factor * m[n];
  /* Constrained image rows take the input's (already projected) rows
   * (enqueue_constrained_rows with input, field_boundary_device.hpp:100-120). */
  for (int k = 0; k <= g->n[2]; ++k)
    for (int j = 0; j <= g->n[1]; ++j)
      for (int i = 0; i <= g->n[0]; ++i)
        for (int c = 0; c < 3; ++c)
          if (constrained_row(g, c, i, j, k)) out[c * nn + node_c(g, i, j, k)] = v[c * nn + node_c(g, i, j, k)];
  free(v);
  free(cells);
  free(curl);
  free(m);
}

void or_maxwell_rhs(const or_grid *g, const double *e, const double *b, const double *current, double delta,
                    double current_factor, double *rhs) {
  const size_t nn = (size_t)or_node_count(g);
  double *curl = (double *)malloc(3 * nn * sizeof(double));
  or_curl_c2n(g, b, curl);
  for (size_t n = 0; n < 3 * nn; ++n) rhs[n] = 1.0 * e[n] + delta * curl[n] + (-current_factor) * current[n];
  zero_constrained_rows(g, rhs);
  free(curl);
}

/* ------------------------------------------------------------------------ */
/* Energy-conserving smoothing: MaxwellOperator::complete_smoothing           */
/* (include/tpic/ecsim/maxwell/maxwell_device.hpp:452-490): passes of the     */
/* node stencil node_smoothing_value (include/tpic/ecsim/fields/              */
/* node_smoothing.hpp:27-58) over E-like ghosts (periodic images; on a wall   */
/* the reflection through the wall node with the electric ghost sign). The   */
/* source and every pass's result have their constrained rows zeroed (the    */
/* projections of complete_electric_refresh, :114-121, and of the input,     */
/* :270-272): S = P (pass P)^n. Face7: alpha*c + beta*(six faces),           */
/* beta = (1 - alpha)/6; Binomial27: (8 c + 4 faces + 2 edges + corners)/64, */
/* each sum left to right.                                                    */
/* ------------------------------------------------------------------------ */

int or_smoothing_active(const or_smoothing *sm) {
  return sm && sm->iterations > 0 && sm->alpha < 1.0;
}

static void smooth_pass(const or_grid *g, const or_smoothing *sm, const double *in, double *out, double *ghosted) {
  const size_t nn = (size_t)or_node_count(g);
  const double alpha = sm->alpha, beta = (1.0 - sm->alpha) / 6.0;
  for (int c = 0; c < 3; ++c) {
    nodes_to_ghosted(g, c, in + c * nn, ghosted);
#define AT(di, dj, dk) ghosted[node_g(g, i + (di), j + (dj), k + (dk))]
    for (int k = 0; k <= g->n[2]; ++k)
      for (int j = 0; j <= g->n[1]; ++j)
        for (int i = 0; i <= g->n[0]; ++i) {
          const double face = AT(-1, 0, 0) + AT(1, 0, 0) + AT(0, -1, 0) + AT(0, 1, 0) + AT(0, 0, -1) + AT(0, 0, 1);
          double value;
          if (sm->op == OR_SMOOTH_FACE7) {
            value = alpha * AT(0, 0, 0) + beta * face;
          } else {
            const double edge = AT(-1, -1, 0) + AT(1, -1, 0) + AT(-1, 1, 0) + AT(1, 1, 0) + AT(-1, 0, -1) +
                                AT(1, 0, -1) + AT(-1, 0, 1) + AT(1, 0, 1) + AT(0, -1, -1) + AT(0, 1, -1) +
                                AT(0, -1, 1) + AT(0, 1, 1);
            const double corner = AT(-1, -1, -1) + AT(1, -1, -1) + AT(-1, 1, -1) + AT(1, 1, -1) + AT(-1, -1, 1) +
                                  AT(1, -1, 1) + AT(-1, 1, 1) + AT(1, 1, 1);
            value = (8.0 * AT(0, 0, 0) + 4.0 * face + 2.0 * edge + corner) * (1.0 / 64.0);
          }
          out[c * nn + node_c(g, i, j, k)] = value;
        }
#undef AT
  }
}

void or_smooth(const or_grid *g, const or_smoothing *sm, const double *in, double *out) {
  const size_t nn = (size_t)or_node_count(g);
  double *ghosted = (double *)malloc(node_g_count(g) * sizeof(double));
  double *a = (double *)malloc(3 * nn * sizeof(double));
  memcpy(a, in, 3 * nn * sizeof(double));
  zero_constrained_rows(g, a);
  for (int pass = 0; pass < sm->iterations; ++pass) {
    smooth_pass(g, sm, a, out, ghosted);
    zero_constrained_rows(g, out);
    memcpy(a, out, 3 * nn * sizeof(double));
  }
  memcpy(out, a, 3 * nn * sizeof(double));
  free(a);
  free(ghosted);
}

/* complete_apply_smoothed and combine_image (maxwell_device.hpp:389-433):
 * the input projected (begin_apply_input, :270-272), the mass acting on
 * S(v), its response smoothed before the combination. */
void or_maxwell_apply_smoothed(const or_grid *g, const or_smoothing *sm, const double *mass, double delta,
                               double current_factor, double inverse_volume, const double *v_in, double *out) {
  const size_t nn = (size_t)or_node_count(g), nc = (size_t)or_cell_count(g);
  double *v = (double *)malloc(3 * nn * sizeof(double));
  double *cells = (double *)malloc(3 * nc * sizeof(double));
  double *curl = (double *)malloc(3 * nn * sizeof(double));
  double *x = (double *)malloc(3 * nn * sizeof(double));
  double *m = (double *)malloc(3 * nn * sizeof(double));
  memcpy(v, v_in, 3 * nn * sizeof(double));
  zero_constrained_rows(g, v);
  or_curl_n2c(g, v, cells);
  or_curl_c2n(g, cells, curl);
  or_smooth(g, sm, v, x);
  or_mass_apply(g, mass, x, inverse_volume, m);
  or_smooth(g, sm, m, m);
  const double d2 = delta * delta;
  for (size_t n = 0; n < 3 * nn; ++n) out[n] = 1.0 * v[n] + d2 * curl[n] + current_factor * m[n];
  for (int k = 0; k <= g->n[2]; ++k)
    for (int j = 0; j <= g->n[1]; ++j)
      for (int i = 0; i <= g->n[0]; ++i)
        for (int c = 0; c < 3; ++c)
          if (constrained_row(g, c, i, j, k)) out[c * nn + node_c(g, i, j, k)] = v[c * nn + node_c(g, i, j, k)];
  free(v);
  free(cells);
  free(curl);
  free(x);
  free(m);
}

/* ------------------------------------------------------------------------ */
/* GMRES(m): Gmres::host_blocking_solve and its kernels                      */
/* (include/tpic/ecsim/gmres/gmres_device.hpp:65-98,164-282,355-454):        */
/* classical Gram-Schmidt in one pass, Givens on the Hessenberg column,      */
/* true residual after every restart, solution committed only on success.  */
/* ------------------------------------------------------------------------ */

static double dot(int n, const double *a, const double *b) {
  double s = 0.0;
  OR_OMP(omp parallel for reduction(+:s) schedule(static))
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

or_gmres_result or_gmres(int n, or_operator apply, void *context, const double *rhs, double *solution,
                         int restart, int max_iterations, double tolerance) {
  or_gmres_result result;
  memset(&result, 0, sizeof result);
  const int m = restart;
  double *basis = (double *)malloc((size_t)(m + 1) * (size_t)n * sizeof(double));
  double *x = (double *)malloc((size_t)n * sizeof(double));
  double *r = (double *)malloc((size_t)n * sizeof(double));
  double *w = (double *)malloc((size_t)n * sizeof(double));
  double *h = (double *)calloc((size_t)m * (size_t)(m + 1), sizeof(double));
  double *cs = (double *)calloc((size_t)m, sizeof(double));
  double *sn = (double *)calloc((size_t)m, sizeof(double));
  double *gv = (double *)calloc((size_t)m + 1, sizeof(double));
  double *y = (double *)calloc((size_t)m, sizeof(double));
  memcpy(x, solution, (size_t)n * sizeof(double));
  /* r = b - A x, with both norms. */
  apply(context, x, w);
  double rr = 0.0, bb = 0.0;
  for (int i = 0; i < n; ++i) {
    r[i] = rhs[i] - w[i];
    rr += r[i] * r[i];
    bb += rhs[i] * rhs[i];
  }
  result.residual = result.initial_residual = sqrt(rr);
  result.rhs_norm = sqrt(bb);
  const double scale = result.rhs_norm == 0.0 ? 1.0 : result.rhs_norm;
  const double absolute = scale * tolerance;
  result.relative_residual = result.residual / scale;
  result.converged = result.residual <= absolute;
  while (!result.converged && result.iterations < max_iterations) {
    memset(h, 0, (size_t)m * (size_t)(m + 1) * sizeof(double));
    memset(cs, 0, (size_t)m * sizeof(double));
    memset(sn, 0, (size_t)m * sizeof(double));
    memset(gv, 0, ((size_t)m + 1) * sizeof(double));
    memset(y, 0, (size_t)m * sizeof(double));
    /* Restart: v0 = r / |r|, rhs[0] = |r| (RestartRhsTail). */
    {
      const double norm2 = dot(n, r, r);
      gv[0] = sqrt(norm2);
      const double inv = norm2 > 0.0 ? 1.0 / sqrt(norm2) : 0.0;
      for (int i = 0; i < n; ++i) basis[i] = r[i] * inv;
    }
    const int columns = m < max_iterations - result.iterations ? m : max_iterations - result.iterations;
    int completed = 0, breakdown = 0;
    for (int col = 0; col < columns; ++col) {
      apply(context, basis + (size_t)col * n, w);
      double *hc = h + (size_t)col * (m + 1);
      for (int row = 0; row <= col; ++row) hc[row] = dot(n, basis + (size_t)row * n, w);
      double norm2 = 0.0;
      OR_OMP(omp parallel for reduction(+:norm2) schedule(static))
      for (int i = 0; i < n; ++i) {
        double value = w[i];
        for (int row = 0; row <= col; ++row) value -= hc[row] * basis[(size_t)row * n + i];
        w[i] = value;
        norm2 += value * value;
      }
      /* ArnoldiLeastSquaresUpdate (gmres_device.hpp:65-98). */
      const double sub = sqrt(norm2);
      hc[col + 1] = sub;
      for (int row = 0; row < col; ++row) {
        const double upper = hc[row], lower = hc[row + 1];
        hc[row] = cs[row] * upper + sn[row] * lower;
        hc[row + 1] = -sn[row] * upper + cs[row] * lower;
      }
      const double diagonal = hc[col];
      const double den = hypot(diagonal, sub);
      const double c = den != 0.0 ? diagonal / den : 1.0;
      const double s = den != 0.0 ? sub / den : 0.0;
      cs[col] = c;
      sn[col] = s;
      hc[col] = c * diagonal + s * sub;
      hc[col + 1] = 0.0;
      const double upper = gv[col], lower = gv[col + 1];
      gv[col] = c * upper + s * lower;
      gv[col + 1] = -s * upper + c * lower;
      const double estimate = fabs(gv[col + 1]);
      breakdown = sub == 0.0;
      const double inv = norm2 > 0.0 ? 1.0 / sqrt(norm2) : 0.0;
      for (int i = 0; i < n; ++i) basis[(size_t)(col + 1) * n + i] = w[i] * inv;
      completed = col + 1;
      if (estimate <= absolute || breakdown) break;
    }
    /* gmres_back_substitute_kernel */
    for (int row = completed - 1; row >= 0; --row) {
      double sum = 0.0;
      for (int col = row + 1; col < completed; ++col) sum += h[(size_t)col * (m + 1) + row] * y[col];
      const double d = h[(size_t)row * (m + 1) + row];
      y[row] = d != 0.0 ? (gv[row] - sum) / d : 0.0;
    }
    /* gmres_update_solution_kernel */
    for (int i = 0; i < n; ++i) {
      double value = x[i];
      for (int col = 0; col < completed; ++col) value += y[col] * basis[(size_t)col * n + i];
      x[i] = value;
    }
    result.iterations += completed;
    apply(context, x, w);
    rr = 0.0;
    for (int i = 0; i < n; ++i) {
      r[i] = rhs[i] - w[i];
      rr += r[i] * r[i];
    }
    result.residual = sqrt(rr);
    result.relative_residual = result.residual / scale;
    result.converged = result.residual <= absolute;
    if (breakdown && !result.converged) break;
  }
  if (result.converged) memcpy(solution, x, (size_t)n * sizeof(double));
  free(basis); free(x); free(r); free(w); free(h); free(cs); free(sn); free(gv); free(y);
  return result;
}

/* ------------------------------------------------------------------------ */
/* Simulation: the single-rank order of simulation::Pipeline                 */
/* (include/tpic/simulation/driver/pipeline.hpp:331-541,632-1000): solve,    */
/* move (old B at nodes), field advance (E, B, B at nodes), sort, deposit    */
/* (next B at nodes), assemble.                                              */
/* ------------------------------------------------------------------------ */

struct or_sim {
  or_grid g;
  int species;
  double q_over_m[OR_MAX_SPECIES];
  double dt, theta, c, tol;
  int restart, max_iterations;
  or_particles p[OR_MAX_SPECIES], scratch[OR_MAX_SPECIES];
  uint32_t *counts[OR_MAX_SPECIES], *offsets[OR_MAX_SPECIES];
  double *e, *b, *bn, *current, *etheta, *mass;
  double inverse_volume, delta, current_factor;
  unsigned validation;
  or_smoothing smoothing;
  double *eparticle; /* S(E_theta) when smoothing is active */
};

static int alloc_particles(or_particles *p, uint32_t n) {
  const size_t m = n ? n : 1;
  p->rx = calloc(m, sizeof(double)); p->ry = calloc(m, sizeof(double)); p->rz = calloc(m, sizeof(double));
  p->u = calloc(m, sizeof(double)); p->v = calloc(m, sizeof(double)); p->w = calloc(m, sizeof(double));
  p->q = calloc(m, sizeof(double));
  p->cell = calloc(m, sizeof(int32_t));
  p->id = calloc(m, sizeof(uint64_t));
  p->count = n;
  return p->rx && p->ry && p->rz && p->u && p->v && p->w && p->q && p->cell && p->id;
}

static void free_particles(or_particles *p) {
  free(p->rx); free(p->ry); free(p->rz); free(p->u); free(p->v); free(p->w); free(p->q);
  free(p->cell); free(p->id);
}

or_sim *or_sim_create(const or_grid *g, int species, const double *q_over_m, const uint32_t *counts, double dt,
                      double theta, double c, double gmres_tol, int restart, int max_iterations) {
  if (species <= 0 || species > OR_MAX_SPECIES) return NULL;
  or_sim *s = (or_sim *)calloc(1, sizeof(or_sim));
  if (!s) return NULL;
  s->g = *g;
  s->species = species;
  s->dt = dt; s->theta = theta; s->c = c; s->tol = gmres_tol;
  s->restart = restart; s->max_iterations = max_iterations;
  const size_t nn = (size_t)or_node_count(g), nc = (size_t)or_cell_count(g);
  const int bins = or_bin_count(g);
  for (int i = 0; i < species; ++i) {
    s->q_over_m[i] = q_over_m[i];
    alloc_particles(&s->p[i], counts[i]);
    alloc_particles(&s->scratch[i], counts[i]);
    s->counts[i] = calloc((size_t)bins, sizeof(uint32_t));
    s->offsets[i] = calloc((size_t)bins + 1, sizeof(uint32_t));
  }
  s->e = calloc(3 * nn, sizeof(double));
  s->b = calloc(3 * nc, sizeof(double));
  s->bn = calloc(3 * nn, sizeof(double));
  s->current = calloc(3 * nn, sizeof(double));
  s->etheta = calloc(3 * nn, sizeof(double));
  s->mass = calloc((size_t)or_vertex_count(g) * 126, sizeof(double));
  s->eparticle = calloc(3 * nn, sizeof(double));
  /* domain.cpp:140-146: V = (double)((long double)dx*dy*dz), 1/V in double. */
  const long double volume = (long double)g->spacing[0] * g->spacing[1] * g->spacing[2];
  s->inverse_volume = 1.0 / (double)volume;
  /* physics.cpp:35-36 and maxwell_device.hpp:502-510. */
  const double theta_dt = theta * dt;
  s->delta = c * theta_dt;
  s->current_factor = 4.0 * OR_PI * theta_dt;
  return s;
}

void or_sim_destroy(or_sim *s) {
  if (!s) return;
  for (int i = 0; i < s->species; ++i) {
    free_particles(&s->p[i]);
    free_particles(&s->scratch[i]);
    free(s->counts[i]);
    free(s->offsets[i]);
  }
  free(s->e); free(s->b); free(s->bn); free(s->current); free(s->etheta); free(s->mass); free(s->eparticle);
  free(s);
}

double *or_sim_field(or_sim *s, int which) {
  switch (which) {
    case 0: return s->e;
    case 1: return s->b;
    case 2: return s->bn;
    case 3: return s->current;
    case 4: return s->etheta;
    case 5: return s->mass;
    case 6: return or_smoothing_active(&s->smoothing) ? s->eparticle : s->etheta;
    default: return NULL;
  }
}

or_particles *or_sim_particles(or_sim *s, int species) { return &s->p[species]; }
uint32_t *or_sim_bins(or_sim *s, int species, int offsets) {
  return offsets ? s->offsets[species] : s->counts[species];
}
unsigned or_sim_validation(const or_sim *s) { return s->validation; }

int or_sim_sort(or_sim *s, int species) {
  const int rc = or_sort(&s->g, &s->p[species], &s->scratch[species], NULL, s->counts[species], s->offsets[species]);
  if (rc != 0) { s->validation |= OR_INVALID_SORT; return rc; }
  or_particles t = s->p[species];
  s->p[species] = s->scratch[species];
  s->scratch[species] = t;
  return 0;
}

unsigned or_sim_deposit(or_sim *s) {
  const size_t nn = (size_t)or_node_count(&s->g);
  memset(s->current, 0, 3 * nn * sizeof(double));
  memset(s->mass, 0, (size_t)or_vertex_count(&s->g) * 126 * sizeof(double));
  unsigned flags = 0;
  for (int i = 0; i < s->species; ++i) {
    /* deposition_actions_device.hpp:72-74: beta/c = (q/m) dt / (2c). */
    const double beta_over_c = s->q_over_m[i] * s->dt / (2 * s->c);
    flags |= or_deposit_current(&s->g, &s->p[i], s->counts[i], s->offsets[i], s->bn, s->inverse_volume,
                                beta_over_c, s->current);
    flags |= or_deposit_mass(&s->g, &s->p[i], s->counts[i], s->offsets[i], s->bn, beta_over_c, s->mass);
  }
  or_fold_current(&s->g, s->current);
  or_fold_mass(&s->g, s->mass);
  s->validation |= flags;
  return flags;
}

int or_sim_initialize(or_sim *s) {
  /* The committed field keeps its constrained rows at zero. */
  zero_constrained_rows(&s->g, s->e);
  or_cell_to_node(&s->g, s->b, s->bn);
  for (int i = 0; i < s->species; ++i)
    if (or_sim_sort(s, i) != 0) return -1;
  or_sim_deposit(s);
  return 0;
}

int or_sim_set_smoothing(or_sim *s, int op, int iterations, double alpha) {
  if ((op != OR_SMOOTH_FACE7 && op != OR_SMOOTH_BINOMIAL27) || iterations < 0 || !(alpha >= 0.0 && alpha <= 1.0))
    return -1;
  s->smoothing.op = op;
  s->smoothing.iterations = iterations;
  s->smoothing.alpha = alpha;
  return 0;
}

typedef struct { or_sim *s; } op_context;

static void sim_operator(void *context, const double *in, double *out) {
  or_sim *s = ((op_context *)context)->s;
  if (or_smoothing_active(&s->smoothing))
    or_maxwell_apply_smoothed(&s->g, &s->smoothing, s->mass, s->delta, s->current_factor, s->inverse_volume, in,
                              out);
  else
    or_maxwell_apply(&s->g, s->mass, s->delta, s->current_factor, s->inverse_volume, in, out);
}

/* With smoothing (complete_rhs and complete_particle_field,
 * maxwell_device.hpp:151-171,213-228): the right-hand side takes S(J), the
 * movers S(E_theta). */
int or_sim_solve(or_sim *s, or_gmres_result *result) {
  const size_t nn = (size_t)or_node_count(&s->g);
  const int smoothed = or_smoothing_active(&s->smoothing);
  double *rhs = (double *)malloc(3 * nn * sizeof(double));
  double *current = s->current;
  if (smoothed) {
    current = (double *)malloc(3 * nn * sizeof(double));
    or_smooth(&s->g, &s->smoothing, s->current, current);
  }
  or_maxwell_rhs(&s->g, s->e, s->b, current, s->delta, s->current_factor, rhs);
  if (smoothed) free(current);
  /* enqueue_rhs_fields (maxwell_device.hpp:140-142): theta starts as a
   * projected copy of E. */
  memcpy(s->etheta, s->e, 3 * nn * sizeof(double));
  zero_constrained_rows(&s->g, s->etheta);
  op_context ctx = {s};
  const or_gmres_result r = or_gmres((int)(3 * nn), sim_operator, &ctx, rhs, s->etheta, s->restart,
                                     s->max_iterations, s->tol);
  free(rhs);
  if (result) *result = r;
  if (r.converged && smoothed) or_smooth(&s->g, &s->smoothing, s->etheta, s->eparticle);
  return r.converged ? 0 : -5;
}

unsigned or_sim_move(or_sim *s, int species) {
  or_mover_params p;
  const double beta = s->q_over_m[species] * s->dt / 2;
  p.beta = beta;
  p.beta_over_c = beta / s->c;
  p.dt = s->dt;
  for (int a = 0; a < 3; ++a) p.inverse_spacing[a] = 1.0 / s->g.spacing[a];
  const double *field = or_smoothing_active(&s->smoothing) ? s->eparticle : s->etheta;
  const unsigned flags = or_move(&s->g, &p, &s->p[species], field, s->bn);
  s->validation |= flags;
  return flags;
}

/* field_update_device.hpp:52-78: E' = (1/th) Eth + (-(1-th)/th) E;
 * B' = 1*B + (-c dt) curl_N2C(Eth); Bn' = C2N(B'). */
void or_sim_advance_fields(or_sim *s) {
  const size_t nn = (size_t)or_node_count(&s->g), nc = (size_t)or_cell_count(&s->g);
  if (fabs(s->theta - 1.0) < 1e-10) {
    memcpy(s->e, s->etheta, 3 * nn * sizeof(double));
  } else {
    const double c0 = 1.0 / s->theta, c1 = -(1.0 - s->theta) / s->theta;
    for (size_t n = 0; n < 3 * nn; ++n) s->e[n] = c0 * s->etheta[n] + c1 * s->e[n];
  }
  double *curl = (double *)malloc(3 * nc * sizeof(double));
  or_curl_n2c(&s->g, s->etheta, curl);
  const double faraday = -s->c * s->dt;
  for (size_t n = 0; n < 3 * nc; ++n) s->b[n] = 1.0 * s->b[n] + faraday * curl[n];
  free(curl);
  or_cell_to_node(&s->g, s->b, s->bn);
}

int or_sim_cycle(or_sim *s, int cycles, or_gmres_result *last) {
  for (int n = 0; n < cycles; ++n) {
    or_gmres_result r;
    if (or_sim_solve(s, &r) != 0) {
      if (last) *last = r;
      return -5;
    }
    if (last) *last = r;
    for (int i = 0; i < s->species; ++i) or_sim_move(s, i);
    or_sim_advance_fields(s);
    for (int i = 0; i < s->species; ++i)
      if (or_sim_sort(s, i) != 0) return -1;
    or_sim_deposit(s);
  }
  return 0;
}

/* Scalar diagnostics (include/tpic/diagnostics/scalar_diagnostics_device.hpp:50-112):
 * W = V/(8 pi) sum over uniquely owned nodes of |E|^2/f, cells of |B|^2;
 * kinetic energy sum of 0.5 (q/qom) |v|^2. */
void or_sim_energy(const or_sim *s, double *out) {
  const or_grid *g = &s->g;
  const size_t nn = (size_t)or_node_count(g), nc = (size_t)or_cell_count(g);
  const long double volume = (long double)g->spacing[0] * g->spacing[1] * g->spacing[2];
  const double normalization = (double)volume / (8 * OR_PI);
  double we = 0.0, wb = 0.0;
  /* Owned nodes: the upper plane of a periodic axis is a duplicate; a wall
   * axis owns both wall planes, weighted 1/f_g. */
  const int upper[3] = {g->n[0] - !wall_axis(g, 0), g->n[1] - !wall_axis(g, 1), g->n[2] - !wall_axis(g, 2)};
  for (int k = 0; k <= upper[2]; ++k)
    for (int j = 0; j <= upper[1]; ++j)
      for (int i = 0; i <= upper[0]; ++i) {
        const size_t n = node_c(g, i, j, k);
        const double x = s->e[n], y = s->e[nn + n], z = s->e[2 * nn + n];
        we += (1.0 / dual_volume_factor(g, i, j, k)) * (x * x + y * y + z * z);
      }
  for (size_t n = 0; n < nc; ++n) {
    const double x = s->b[n], y = s->b[nc + n], z = s->b[2 * nc + n];
    wb += x * x + y * y + z * z;
  }
  out[0] = normalization * we;
  out[1] = normalization * wb;
  for (int i = 0; i < s->species; ++i) {
    const or_particles *p = &s->p[i];
    const double inverse_qom = 1.0 / s->q_over_m[i];
    double k = 0.0;
    for (uint32_t n = 0; n < p->count; ++n) {
      const double mass = p->q[n] * inverse_qom;
      k += 0.5 * mass * (p->u[n] * p->u[n] + p->v[n] * p->v[n] + p->w[n] * p->w[n]);
    }
    out[2 + i] = k;
  }
}

/* The full scalar sample (ScalarDiagnostics species rows, scalar_diagnostics_
 * device.hpp:50-160): W_E, W_B, then per species kinetic energy, charge sum q,
 * momentum sum (q/qom) v. */
void or_sim_scalars(const or_sim *s, double *out) {
  double e[2 + OR_MAX_SPECIES];
  or_sim_energy(s, e);
  out[0] = e[0];
  out[1] = e[1];
  for (int i = 0; i < s->species; ++i) {
    const or_particles *p = &s->p[i];
    const double inverse_qom = 1.0 / s->q_over_m[i];
    double q = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
    for (uint32_t n = 0; n < p->count; ++n) {
      const double mass = p->q[n] * inverse_qom;
      q += p->q[n];
      mx += mass * p->u[n];
      my += mass * p->v[n];
      mz += mass * p->w[n];
    }
    out[2 + 5 * i] = e[2 + i];
    out[3 + 5 * i] = q;
    out[4 + 5 * i] = mx;
    out[5 + 5 * i] = my;
    out[6 + 5 * i] = mz;
  }
}

/* ------------------------------------------------------------------------ */
/* XORWOW of rocRAND (rocrand_xorwow.h: xorwow_engine), the generator of    */
/* hipRAND's default state on ROCm, used by the reference's loader          */
/* (include/tpic/initialization/particles/particle_loading_device.hpp:37-55).*/
/* Marsaglia's xorwow: five-word xorshift plus a Weyl sequence d.           */
/* ------------------------------------------------------------------------ */
static void xorwow_shift(uint32_t x[5]) {
  const uint32_t t = x[0] ^ (x[0] >> 2);
  x[0] = x[1];
  x[1] = x[2];
  x[2] = x[3];
  x[3] = x[4];
  x[4] = (x[4] ^ (x[4] << 4)) ^ (t ^ (t << 1));
}

/* 160 x 160 matrices over GF(2), stored by columns: col[j] = M e_j. */
typedef struct { uint32_t col[160][5]; } gf2_matrix;

static void gf2_apply(const gf2_matrix *m, const uint32_t in[5], uint32_t out[5]) {
  uint32_t y[5] = {0, 0, 0, 0, 0};
  for (int j = 0; j < 160; ++j)
    if ((in[j >> 5] >> (j & 31)) & 1u)
      for (int w = 0; w < 5; ++w) y[w] ^= m->col[j][w];
  for (int w = 0; w < 5; ++w) out[w] = y[w];
}

static void gf2_square(const gf2_matrix *a, gf2_matrix *out) {
  for (int j = 0; j < 160; ++j) gf2_apply(a, a->col[j], out->col[j]);
}

/* jumps[k] = A^(2^(67 + k)), A the one-step xorshift transition. */
static gf2_matrix *xorwow_jumps = NULL;
enum { XORWOW_JUMPS = 48 };

static void xorwow_prepare(void) {
  if (xorwow_jumps) return;
  gf2_matrix *j = (gf2_matrix *)malloc(sizeof(gf2_matrix) * XORWOW_JUMPS);
  gf2_matrix *a = (gf2_matrix *)malloc(sizeof(gf2_matrix));
  gf2_matrix *b = (gf2_matrix *)malloc(sizeof(gf2_matrix));
  for (int c = 0; c < 160; ++c) {
    uint32_t e[5] = {0, 0, 0, 0, 0};
    e[c >> 5] = 1u << (c & 31);
    xorwow_shift(e);
    memcpy(a->col[c], e, sizeof e);
  }
  for (int s = 0; s < 67; ++s) {  /* A^(2^67) */
    gf2_square(a, b);
    gf2_matrix *t = a; a = b; b = t;
  }
  j[0] = *a;
  for (int k = 1; k < XORWOW_JUMPS; ++k) gf2_square(&j[k - 1], &j[k]);
  free(a);
  free(b);
  xorwow_jumps = j;
}

/* xorwow_engine(seed, subsequence, 0): the default state, seeded with two
 * products of the seed halves, then subsequence * 2^67 steps (the Weyl
 * value d is unchanged by a multiple of 2^32 steps). */
void or_xorwow_init(or_xorwow *st, uint64_t seed, uint64_t subsequence) {
  xorwow_prepare();
  st->x[0] = 123456789u;
  st->x[1] = 362436069u;
  st->x[2] = 521288629u;
  st->x[3] = 88675123u;
  st->x[4] = 5783321u;
  st->d = 6615241u;
  const uint32_t s0 = (uint32_t)seed ^ 0x2c7f967fu;
  const uint32_t s1 = (uint32_t)(seed >> 32) ^ 0xa03697cbu;
  const uint32_t t0 = 1228688033u * s0;
  const uint32_t t1 = 2073658381u * s1;
  st->x[0] += t0;
  st->x[1] ^= t0;
  st->x[2] += t1;
  st->x[3] ^= t1;
  st->x[4] += t0;
  st->d += t1 + t0;
  for (int k = 0; subsequence != 0 && k < XORWOW_JUMPS; ++k, subsequence >>= 1)
    if (subsequence & 1u) gf2_apply(&xorwow_jumps[k], st->x, st->x);
  st->has_cached = 0;
  st->cached_normal = 0.0;
}

uint32_t or_xorwow_next(or_xorwow *st) {
  xorwow_shift(st->x);
  st->d += 362437u;
  return st->d + st->x[4];
}

/* rocrand_uniform.h: 53 bits of two draws, (0, 1]. */
double or_xorwow_uniform_double(or_xorwow *st) {
  const uint32_t v1 = or_xorwow_next(st), v2 = or_xorwow_next(st);
  const uint64_t v = (uint64_t)v1 | ((uint64_t)(v2 >> 11) << 32);
  const double inv53 = 1.1102230246251565e-16; /* 2^-53 */
  return inv53 + (double)v * inv53;
}

/* rocrand_normal.h box_muller_double: u in (0, 1], w in (0, 2], radius
 * sqrt(-2 log u), angle pi w. */
static void box_muller(uint32_t a, uint32_t b, uint32_t c, uint32_t d, double out[2]) {
  const double inv53 = 1.1102230246251565e-16;
  const uint64_t v1 = (uint64_t)a ^ ((uint64_t)b << 21);
  const double u = inv53 + (double)v1 * inv53;
  const uint64_t v2 = (uint64_t)c ^ ((uint64_t)d << 21);
  const double w = inv53 * 2.0 + (double)v2 * (inv53 * 2.0);
  const double s = sqrt(-2.0 * log(u));
  const double pi = 3.141592653589793238462643383279502884;
  out[0] = sin(w * pi) * s;
  out[1] = cos(w * pi) * s;
}

void or_xorwow_normal2_double(or_xorwow *st, double out[2]) {
  const uint32_t a = or_xorwow_next(st), b = or_xorwow_next(st), c = or_xorwow_next(st), d = or_xorwow_next(st);
  box_muller(a, b, c, d, out);
}

/* One normal: the cached second value of the last pair, else a new pair
 * whose second value is cached (normal2 leaves the cache alone). */
double or_xorwow_normal_double(or_xorwow *st) {
  if (st->has_cached) {
    st->has_cached = 0;
    return st->cached_normal;
  }
  double pair[2];
  or_xorwow_normal2_double(st, pair);
  st->cached_normal = pair[1];
  st->has_cached = 1;
  return pair[0];
}

/* ------------------------------------------------------------------------ */
/* Uniform-case initial load: kernels::initial_particle_load_kernel with    */
/* UniformProfile (particle_loading_device.hpp:69-159), the plan of        */
/* prepare_particle_loads (prepared_initialization.cpp:37-85), the carry of */
/* canonicalize_particle_fraction (particle_cell.hpp:83-93) and the startup */
/* fold of classify_particle_point (particle_boundaries.hpp:111-165).       */
/* ------------------------------------------------------------------------ */
void or_load_uniform(const or_grid *g, const int ppc[3], double rho_init, const double drift[3],
                     const double thermal[3], int quiet_start, uint64_t seed, uint64_t first_id,
                     double q_over_m, or_particles *out) {
  or_load_seeded(g, ppc, rho_init, drift, thermal, quiet_start, seed, first_id, q_over_m, NULL, out);
}

/* The same load with the TwoStream profile's density seed
 * (TwoStreamProfile::evaluate, include/tpic/initialization/cases/
 * two_stream.hpp:17-23): charge = (q_factor density) (1 + A cos(k x + phase)),
 * x = (i + rx) dx after the carry, origin 0. seed_mode = {k, phase, A} or NULL. */
void or_load_seeded(const or_grid *g, const int ppc[3], double rho_init, const double drift[3],
                    const double thermal[3], int quiet_start, uint64_t seed, uint64_t first_id,
                    double q_over_m, const double *seed_mode, or_particles *out) {
  or_load_profile(g, ppc, rho_init, drift, thermal, quiet_start, seed, first_id, q_over_m, seed_mode, NULL, out);
}

static double sech_squared(double v) {
  const double ch = cosh(v);
  return 1.0 / (ch * ch);
}

/* HarrisProfile::harris_shape (include/tpic/initialization/cases/harris_sheet.hpp:110-117). */
static double harris_shape(const or_harris *h, double y) {
  if (h->is_double) return sech_squared((y - 0.25 * h->ly) / h->delta) + sech_squared((y - 0.75 * h->ly) / h->delta);
  return sech_squared((y - 0.5 * h->ly) / h->delta);
}

/* HarrisCurl::at (harris_sheet.hpp:33-51). */
double or_harris_curl(const or_harris *h, double x, double y) {
  const double xm = x - h->center_x;
  const double ym = y - h->center_y;
  const double ch = cosh(ym / h->curl_delta);
  const double sx = xm * h->hump_inverse_wx2;
  const double sy = ym * h->hump_inverse_wy2;
  const double cos_x = cos(h->hump_kx * xm);
  const double sin_x = sin(h->hump_kx * xm);
  const double cos_y = cos(h->hump_ky * ym);
  const double sin_y = sin(h->hump_ky * ym);
  return h->sheet / (ch * ch) + h->mode * cos(h->mode_kx * xm) * cos(h->mode_ky * ym) -
         h->hump * exp(-xm * sx - ym * sy) *
             (cos_y * ((4.0 * sx * sx - 2.0 * h->hump_inverse_wx2 - h->hump_kx * h->hump_kx) * cos_x +
                       4.0 * sx * h->hump_kx * sin_x) +
              cos_x * ((4.0 * sy * sy - 2.0 * h->hump_inverse_wy2 - h->hump_ky * h->hump_ky) * cos_y +
                       4.0 * sy * h->hump_ky * sin_y));
}

/* The load with a case profile: the TwoStream density seed {k, phase, A}
 * and/or the Harris profile (HarrisProfile::evaluate, harris_sheet.hpp:64-97):
 * density (density 0.5) nodes with nodes the sheet shape at the cell's two y
 * nodes; a double sheet's z drift times -tanh((y - ly/2)/delta); with
 * currentFromAmpere the z drift ampere_drift curl(x, y) / cell shape.
 * charge = (q_factor density) multiplier (particle_loading_device.hpp:141-142). */
void or_load_profile(const or_grid *g, const int ppc[3], double rho_init, const double drift[3],
                     const double thermal[3], int quiet_start, uint64_t seed, uint64_t first_id,
                     double q_over_m, const double *seed_mode, const or_harris *harris, or_particles *out) {
  const int per_cell = ppc[0] * ppc[1] * ppc[2];
  const double volume = (g->spacing[0] * g->spacing[1]) * g->spacing[2];
  const double q_factor = (q_over_m > 0.0 ? 1.0 : -1.0) * volume / (double)per_cell;
  const double density = rho_init / 12.566370614359172953850573533118;

  double inverse_ppc[3];
  for (int a = 0; a < 3; ++a) inverse_ppc[a] = 1.0 / (double)ppc[a];
  /* Subsequence i is the seeded state jumped i times by 2^67 steps; the
   * states are walked with one jump per particle (the same GF(2) products as
   * a fresh initialisation at subsequence i). */
  or_xorwow next;
  or_xorwow_init(&next, seed, 0);
  for (uint32_t index = 0; index < out->count; ++index) {
    const int cell = (int)(index / (uint32_t)per_cell), sub = (int)(index % (uint32_t)per_cell);
    int ijk[3] = {cell % g->n[0], (cell / g->n[0]) % g->n[1], cell / g->n[0] / g->n[1]};
    const int sxyz[3] = {sub % ppc[0], (sub / ppc[0]) % ppc[1], sub / ppc[0] / ppc[1]};
    or_xorwow st = next;
    gf2_apply(&xorwow_jumps[0], next.x, next.x);
    double offset[3];
    for (int a = 0; a < 3; ++a) offset[a] = quiet_start ? 0.5 : or_xorwow_uniform_double(&st);
    double frac[3];
    for (int a = 0; a < 3; ++a) {
      frac[a] = ((double)sxyz[a] + offset[a]) * inverse_ppc[a];
      if (frac[a] == 1.0) { ++ijk[a]; frac[a] = 0.0; }
    }
    double multiplier = 1.0, sample_density = density;
    double d[3] = {drift[0], drift[1], drift[2]}, t[3] = {thermal[0], thermal[1], thermal[2]};
    if (seed_mode) {
      const double x = ((double)ijk[0] + frac[0]) * g->spacing[0];
      multiplier = 1.0 + seed_mode[2] * cos(seed_mode[0] * x + seed_mode[1]);
    }
    if (harris) {
      const double x = ((double)ijk[0] + frac[0]) * g->spacing[0];
      const double y = ((double)ijk[1] + frac[1]) * g->spacing[1];
      double cell_shape = 1.0;
      if (harris->density) {
        const double lower_y = (double)ijk[1] * g->spacing[1];
        const double nodes = harris_shape(harris, lower_y) + harris_shape(harris, lower_y + g->spacing[1]);
        sample_density = sample_density * 0.5 * nodes;
        cell_shape = 0.5 * nodes;
      }
      if (harris->is_double) d[2] = d[2] * -tanh((y - 0.5 * harris->ly) / harris->delta);
      if (harris->ampere_drift != 0.0) {
        const double w = harris->ampere_drift * or_harris_curl(harris, x, y) / cell_shape;
        if (harris->varying_thermal) {
          const double square = w * w;
          const double variance = (square + t[2] * t[2]) - square;
          if (variance > 4.0 * DBL_EPSILON * square) t[2] = sqrt(variance);
        }
        d[2] = w;
      }
    }
    const double charge = (q_factor * sample_density) * multiplier;
    double pair[2];
    or_xorwow_normal2_double(&st, pair);
    const double nz = or_xorwow_normal_double(&st);
    double vel[3] = {d[0] + t[0] * pair[0], d[1] + t[1] * pair[1], d[2] + t[2] * nz};
    for (int a = 0; a < 3; ++a) {
      const int extent = g->n[a];
      if (ijk[a] < extent) continue;
      if (g->particle_periodic[a]) { ijk[a] = 0; continue; }
      int global = 2 * extent - extent - 1;
      frac[a] = 1.0 - frac[a];
      vel[a] = -vel[a];
      if (frac[a] == 1.0) {
        if (global + 1 < extent) { ++global; frac[a] = 0.0; }
        else frac[a] = 1.0 - DBL_EPSILON / 2.0;
      }
      ijk[a] = global;
    }
    out->rx[index] = frac[0];
    out->ry[index] = frac[1];
    out->rz[index] = frac[2];
    out->u[index] = vel[0];
    out->v[index] = vel[1];
    out->w[index] = vel[2];
    out->q[index] = charge;
    out->cell[index] = or_encode(g, ijk[0], ijk[1], ijk[2]);
    out->id[index] = first_id + index;
  }
}

/* ------------------------------------------------------------------------ */
/* CenterInitialVelocity for a uniform drift (center_particle_velocity_kernel */
/* and VelocityCentering::host_blocking_center, include/tpic/initialization/  */
/* particles/particle_centering_device.hpp:57-215): u' = u - drift, then      */
/* u = (u' - mean) + drift with mean = sum q u' / sum q (no shift when        */
/* sum q = 0). Sums in particle order. Returns 1 when a sum, the mean or a    */
/* corrected velocity is not finite (the velocity then stays u - drift).      */
/* ------------------------------------------------------------------------ */
int or_center_velocity(or_particles *s, const double drift[3]) {
  double sums[4] = {0.0, 0.0, 0.0, 0.0};
  double *vel[3] = {s->u, s->v, s->w};
  for (uint32_t i = 0; i < s->count; ++i) {
    sums[0] += s->q[i];
    for (int a = 0; a < 3; ++a) {
      vel[a][i] = vel[a][i] - drift[a];
      sums[a + 1] += s->q[i] * vel[a][i];
    }
  }
  double mean[3] = {0.0, 0.0, 0.0};
  for (int r = 0; r < 4; ++r)
    if (!isfinite(sums[r])) return 1;
  if (sums[0] != 0.0)
    for (int a = 0; a < 3; ++a) mean[a] = sums[a + 1] / sums[0];
  for (int a = 0; a < 3; ++a)
    if (!isfinite(mean[a])) return 1;
  int bad = 0;
  for (uint32_t i = 0; i < s->count; ++i) {
    double next[3];
    for (int a = 0; a < 3; ++a) next[a] = (vel[a][i] - mean[a]) + drift[a];
    if (!isfinite(next[0]) || !isfinite(next[1]) || !isfinite(next[2])) { bad = 1; continue; }
    for (int a = 0; a < 3; ++a) vel[a][i] = next[a];
  }
  return bad;
}
