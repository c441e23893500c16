// Shared device/host definitions of the MI355X ECSIM hot path.
//
// HBM layouts (all FP64 unless stated):
//   node vector   v[c*nodes + n],  n = i + (nx+1)*(j + (ny+1)*k), i in [0,nx]
//                 (periodic duplicates kept, as the reference's compact solver
//                 vectors do: include/tpic/ecsim/gmres/gmres_device.hpp:133-137)
//   cell vector   b[c*cells + n],  n = i + nx*(j + ny*k)
//   particle key  (i+2) + (nx+4)*((j+2) + (ny+4)*(k+2))   (halo-2 map,
//                 include/tpic/data/particles/particle_cell.hpp:33-45)
//   mass matrix   entry-major: m[e*vstride + vertex], e = stencil*9 + comp,
//                 vertex = (i+1) + (nx+3)*((j+1) + (ny+3)*(k+1)), i in [-1,nx+1]
//                 (the reference's node-major 126-scalar record,
//                 include/tpic/data/mass_matrix/mass_matrix_layout.hpp:28-68,
//                 transposed so that neighbouring nodes are neighbouring words)
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace tpic_mi {

// Validation bits (same values as TPIC_MI_INVALID_* in include/tpic_mi.h).
enum : unsigned {
  kInvalidKey = 1u,
  kInvalidFraction = 2u,
  kInvalidNonfinite = 4u,
  kInvalidFlight = 8u,
  kInvalidHalo = 16u,
  kInvalidSort = 32u,
  kInvalidField = 64u,
};

constexpr int kStencils = 14;
constexpr int kEntries = 126;          // 14 stencils x 9 tensor entries
constexpr int kMoments = 27;           // quadratic CIC pair moments per cell
constexpr int kMomentColumns = 12;     // per moment: 9 mass tensor entries, then 3 current components
constexpr int kMassColumns = 9;        // the mass tensor entries of a moment kept in the cell record
constexpr int kCellRecord = kMoments * kMassColumns;  // 243 mass quantities per cell, cell-major
constexpr int kCornerRecord = 24;                        // corner currents per cell: 8 corners x 3 components
constexpr int kMaxSpecies = 8;

// Operand precision of the per-cell mass/current moment products.
constexpr int kMassF64 = 0;     // FP64 MFMA (the parity path)
constexpr int kMassF32 = 1;     // FP32 operands and accumulation per cell, FP64 across cells
constexpr int kMassBf16x3 = 2;  // BF16 hi/lo split products, FP32 accumulation per cell

struct Grid {
  int nx, ny, nz;
  double dx, dy, dz;
  double ix, iy, iz;         // inverse spacing
  int particle_periodic[3];  // 1 periodic, 0 mirror
  unsigned vstride;          // padded vertex count of the mass matrix
  // Closed field walls (include/tpic/config/boundary_conditions.hpp:68-77).
  // An axis is periodic or has a wall on both faces; face = 2 axis + side,
  // side 0 the lower face. b_sign: the ghost sign of each component of a
  // cell (B-like) vector; constrained: bit c set when E component c is held
  // at zero on the face's nodes. (Node ghosts are read only against zero
  // mass blocks, so node signs are not needed.)
  int walls;                     // any wall axis
  int field_wall[3];
  signed char b_sign[3][6];
  unsigned b_flip;  // bit 6 c + face set when b_sign[c][face] is -1
  unsigned char constrained[6];

  // WallFaces::dual_volume_factor (include/tpic/ecsim/fields/wall_faces.hpp:27-30).
  __host__ __device__ unsigned dual_factor(int i, int j, int k) const {
    const unsigned w = (field_wall[0] ? unsigned(i == 0) + unsigned(i == nx) : 0u) +
                       (field_wall[1] ? unsigned(j == 0) + unsigned(j == ny) : 0u) +
                       (field_wall[2] ? unsigned(k == 0) + unsigned(k == nz) : 0u);
    return 1u << w;
  }
  // A row held at zero: E component c on a wall node whose face constrains it.
  __host__ __device__ bool constrained_row(int c, int i, int j, int k) const {
    const unsigned bit = 1u << c;
    return (field_wall[0] && ((i == 0 && (constrained[0] & bit)) || (i == nx && (constrained[1] & bit)))) ||
           (field_wall[1] && ((j == 0 && (constrained[2] & bit)) || (j == ny && (constrained[3] & bit)))) ||
           (field_wall[2] && ((k == 0 && (constrained[4] & bit)) || (k == nz && (constrained[5] & bit))));
  }

  __host__ __device__ unsigned nodes() const {
    return unsigned(nx + 1) * unsigned(ny + 1) * unsigned(nz + 1);
  }
  __host__ __device__ unsigned cells() const { return unsigned(nx) * unsigned(ny) * unsigned(nz); }
  __host__ __device__ unsigned bins() const {
    return unsigned(nx + 4) * unsigned(ny + 4) * unsigned(nz + 4);
  }
  __host__ __device__ unsigned vertices() const {
    return unsigned(nx + 3) * unsigned(ny + 3) * unsigned(nz + 3);
  }
  __host__ __device__ unsigned node(int i, int j, int k) const {
    return unsigned(i) + unsigned(nx + 1) * (unsigned(j) + unsigned(ny + 1) * unsigned(k));
  }
  __host__ __device__ unsigned cell(int i, int j, int k) const {
    return unsigned(i) + unsigned(nx) * (unsigned(j) + unsigned(ny) * unsigned(k));
  }
  __host__ __device__ unsigned vertex(int i, int j, int k) const {
    return unsigned(i + 1) + unsigned(nx + 3) * (unsigned(j + 1) + unsigned(ny + 3) * unsigned(k + 1));
  }
  __host__ __device__ int32_t encode(int i, int j, int k) const {
    return (i + 2) + (nx + 4) * ((j + 2) + (ny + 4) * (k + 2));
  }
  __host__ __device__ void decode(int32_t key, int& i, int& j, int& k) const {
    const int ex = nx + 4, ey = ny + 4;
    const int yz = key / ex;
    i = key % ex - 2;
    j = yz % ey - 2;
    k = yz / ey - 2;
  }
  __host__ __device__ bool interior(int i, int j, int k) const {
    return i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz;
  }
};

// Periodic neighbours on the compact node grid: node -1 is node n-1 and node
// n+1 is node 1 (the images the reference's self-route ghost copy holds).
__host__ __device__ inline int wrap_node(int i, int n) { return i < 0 ? i + n : (i > n ? i - n : i); }
// Periodic cells: cell -1 is cell n-1, cell n is cell 0.
__host__ __device__ inline int wrap_cell(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// The cell that holds the value of cell i (i in [-1, n]) along an axis: the
// periodic image, or on a wall axis the ghost's interior source (cell 0 or
// n-1) with face set to the wall's face (-1 for an interior cell).
__host__ __device__ inline int cell_image(const Grid& g, int axis, int i, int n, int& face) {
  face = -1;
  if (!g.field_wall[axis]) return wrap_cell(i, n);
  if (i < 0) {
    face = 2 * axis;
    return 0;
  }
  if (i >= n) {
    face = 2 * axis + 1;
    return n - 1;
  }
  return i;
}

// Product of the B-like ghost signs of component c over the faces crossed.
// fx is -1 (no face crossed) or face 0/1, fy -1 or 2/3, fz -1 or 4/5: each
// factor is selected among two per-face signs formed from the grid's flip
// bits (no indexing by a per-thread face, which the device would serve from
// memory).
__host__ __device__ inline double b_ghost_sign(const Grid& g, int c, int fx, int fy, int fz) {
  const unsigned row = g.b_flip >> (6 * c);
  const double s0 = (row & 1u) ? -1.0 : 1.0, s1 = (row & 2u) ? -1.0 : 1.0;
  const double s2 = (row & 4u) ? -1.0 : 1.0, s3 = (row & 8u) ? -1.0 : 1.0;
  const double s4 = (row & 16u) ? -1.0 : 1.0, s5 = (row & 32u) ? -1.0 : 1.0;
  const double sx = fx < 0 ? 1.0 : (fx == 0 ? s0 : s1);
  const double sy = fy < 0 ? 1.0 : (fy == 2 ? s2 : s3);
  const double sz = fz < 0 ? 1.0 : (fz == 4 ? s4 : s5);
  return sx * sy * sz;
}

// A node index i in [-1, n+1] along an axis: the periodic image, or on a
// wall axis the reflection through the wall node (the reads this serves
// multiply zero mass blocks there).
__host__ __device__ inline int node_image(const Grid& g, int axis, int i, int n) {
  if (!g.field_wall[axis]) return wrap_node(i, n);
  return i < 0 ? -i : (i > n ? 2 * n - i : i);
}

// One species' SoA particle slot.
struct ParticleView {
  double *rx, *ry, *rz, *u, *v, *w, *q;
  int32_t* cell;
  uint64_t* id;
};

// Blocks are dispatched round-robin over the 8 XCDs (each with its own L2).
// Remap so that XCD x receives one contiguous range of logical blocks: work on
// neighbouring cells then shares that XCD's L2. Bijective for any grid size.
__device__ inline unsigned xcd_logical_block(unsigned b, unsigned nb) {
  const unsigned q = nb / 8u, rem = nb % 8u, x = b % 8u, r = b / 8u;
  return x * q + (x < rem ? x : rem) + r;
}

__device__ inline bool finite_value(double v) { return v >= -1.7976931348623157e308 && v <= 1.7976931348623157e308; }

// Plan of the stable counting sort for destination-ordered passes. A run is
// a maximal sequence of input particles with one key: run r starts at source
// run_start[r] (run_start[r + 1] ends it) and lands at destination
// run_dest[r]. Bin b holds counts[b] particles from offsets[b] on, and its
// rcounts[b] runs are listed at bin_runs[roffsets[b] ...]. A nonzero
// *failure (an invalid key) voids the plan.
struct RunPlan {
  const uint32_t *counts, *offsets, *rcounts, *roffsets, *bin_runs, *run_start, *run_dest, *failure;
};

// The runs of one bin, lane t < nr (nr <= 64) holding run t: destination
// start, length and source start.
struct BinRuns {
  uint32_t nr, rfirst, rd, rl, rs;
};

__device__ inline void load_bin_runs(const RunPlan& p, uint32_t b, unsigned lane, BinRuns& br) {
  br.nr = p.rcounts[b];
  br.rfirst = p.roffsets[b];
  br.rd = br.rl = br.rs = 0;
  if (br.nr <= 64 && lane < br.nr) {
    const uint32_t r = p.bin_runs[br.rfirst + lane];
    br.rd = p.run_dest[r];
    br.rs = p.run_start[r];
    br.rl = p.run_start[r + 1] - br.rs;
  }
}

// Source index of destination d of the bin (a bin with more than 64 runs
// reads its run list from memory).
__device__ inline uint32_t run_source(const RunPlan& p, const BinRuns& br, uint32_t d) {
  uint32_t from = 0;
  if (br.nr <= 64) {
    for (uint32_t t = 0; t < br.nr; ++t) {
      const uint32_t td = uint32_t(__builtin_amdgcn_readlane(int(br.rd), int(t)));
      const uint32_t tl = uint32_t(__builtin_amdgcn_readlane(int(br.rl), int(t)));
      const uint32_t ts = uint32_t(__builtin_amdgcn_readlane(int(br.rs), int(t)));
      if (d - td < tl) from = ts + (d - td);
    }
  } else {
    for (uint32_t t = 0; t < br.nr; ++t) {
      const uint32_t r = p.bin_runs[br.rfirst + t];
      const uint32_t td = p.run_dest[r], ts = p.run_start[r], tl = p.run_start[r + 1] - ts;
      if (d - td < tl) from = ts + (d - td);
    }
  }
  return from;
}

}  // namespace tpic_mi
