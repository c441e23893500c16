// Implicit Maxwell operator, right-hand side and field advance on one GPU
// (compact node/cell vectors; periodic neighbours by index wrap, closed walls
// by signed ghost images, constrained rows and dual-volume row scales).
//
//   rhs  = 1*E + delta*curl_C2N(B) - k*J                 (maxwell_device.hpp:152-171)
//   A(v) = 1*v + delta^2*curl_C2N(curl_N2C v) + k*(1/V)*M v  (:353-433, mass apply
//          mass_matrix_stencil.hpp:120-149: centre, 13 forward and 13 backward
//          blocks, backward blocks not transposed)
//   E'   = (1/th) E_th + (-(1-th)/th) E;  B' = 1*B + (-c dt) curl_N2C(E_th)
//                                                        (field_update_device.hpp:25-78)
//   Bn   = cell_to_node_average(B)                      (cell_to_node.hpp:19-32)
// with delta = c*th*dt and k = 4*pi*th*dt. Curls average four edge differences
// with the factor 1/4 in the reference's order (node_cell_curl.hpp:60-119).
//
// The operator is two kernels: the node->cell curl (24 B per cell) and one
// fused node kernel doing the cell->node curl, the 27-block mass product from
// the entry-major matrix (coalesced: neighbouring nodes read neighbouring
// words) and the three-term combination. Bound: HBM, 1008 B of matrix per node.
#include "internal.hpp"
#include "../../include/tpic_mi.h"

namespace tpic_mi {
namespace {

constexpr unsigned kThreads = 256;

struct Samples {
  double base, x, y, z, xy, xz, yz, xyz;
};

// Node sources: difference = neighbour - anchor; cell sources: anchor - neighbour.
template <bool NodeSource>
__device__ __forceinline__ double diff(double anchor, double neighbour) {
  return NodeSource ? neighbour - anchor : anchor - neighbour;
}

template <bool NodeSource>
__device__ __forceinline__ double derivative(const Samples& s, int axis, double inverse) {
  double sum;
  if (axis == 0)
    sum = diff<NodeSource>(s.base, s.x) + diff<NodeSource>(s.z, s.xz) + diff<NodeSource>(s.y, s.xy) +
          diff<NodeSource>(s.yz, s.xyz);
  else if (axis == 1)
    sum = diff<NodeSource>(s.base, s.y) + diff<NodeSource>(s.z, s.yz) + diff<NodeSource>(s.x, s.xy) +
          diff<NodeSource>(s.xz, s.xyz);
  else
    sum = diff<NodeSource>(s.base, s.z) + diff<NodeSource>(s.x, s.xz) + diff<NodeSource>(s.y, s.yz) +
          diff<NodeSource>(s.xy, s.xyz);
  return 0.25 * inverse * sum;
}

template <bool NodeSource>
__device__ __forceinline__ void curl_of(const Samples* s, const Grid& g, double& x, double& y, double& z) {
  x = derivative<NodeSource>(s[2], 1, g.iy) - derivative<NodeSource>(s[1], 2, g.iz);
  y = derivative<NodeSource>(s[0], 2, g.iz) - derivative<NodeSource>(s[2], 0, g.ix);
  z = derivative<NodeSource>(s[1], 0, g.ix) - derivative<NodeSource>(s[0], 1, g.iy);
}

// Cell (i,j,k) from its 8 nodes (all interior, no wrap needed).
__device__ __forceinline__ void n2c_at(const Grid& g, const double* __restrict__ v, int i, int j, int k, double& x,
                                       double& y, double& z) {
  const unsigned nodes = g.nodes();
  const unsigned sy = unsigned(g.nx + 1), sz = sy * unsigned(g.ny + 1);
  const unsigned n0 = g.node(i, j, k);
  Samples s[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double* f = v + size_t(c) * nodes;
    s[c].base = f[n0];
    s[c].x = f[n0 + 1];
    s[c].y = f[n0 + sy];
    s[c].z = f[n0 + sz];
    s[c].xy = f[n0 + sy + 1];
    s[c].xz = f[n0 + sz + 1];
    s[c].yz = f[n0 + sz + sy];
    s[c].xyz = f[n0 + sz + sy + 1];
  }
  curl_of<true>(s, g, x, y, z);
}

// Node (i,j,k) from its 8 cells: periodic wrap, or with Walls the wall
// ghosts (the interior image times the B-like ghost sign of each crossed
// face, field_boundary_device.hpp:26-57).
template <bool Walls>
__device__ __forceinline__ void c2n_at(const Grid& g, const double* __restrict__ b, int i, int j, int k, double& x,
                                       double& y, double& z) {
  const unsigned cells = g.cells();
  Samples s[3];
  if constexpr (!Walls) {
    const int i0 = wrap_cell(i, g.nx), i1 = wrap_cell(i - 1, g.nx);
    const int j0 = wrap_cell(j, g.ny), j1 = wrap_cell(j - 1, g.ny);
    const int k0 = wrap_cell(k, g.nz), k1 = wrap_cell(k - 1, g.nz);
    const unsigned at[8] = {g.cell(i0, j0, k0), g.cell(i1, j0, k0), g.cell(i0, j1, k0), g.cell(i0, j0, k1),
                            g.cell(i1, j1, k0), g.cell(i1, j0, k1), g.cell(i0, j1, k1), g.cell(i1, j1, k1)};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double* f = b + size_t(c) * cells;
      s[c].base = f[at[0]];
      s[c].x = f[at[1]];
      s[c].y = f[at[2]];
      s[c].z = f[at[3]];
      s[c].xy = f[at[4]];
      s[c].xz = f[at[5]];
      s[c].yz = f[at[6]];
      s[c].xyz = f[at[7]];
    }
  } else {
    int fi0, fi1, fj0, fj1, fk0, fk1;
    const int i0 = cell_image(g, 0, i, g.nx, fi0), i1 = cell_image(g, 0, i - 1, g.nx, fi1);
    const int j0 = cell_image(g, 1, j, g.ny, fj0), j1 = cell_image(g, 1, j - 1, g.ny, fj1);
    const int k0 = cell_image(g, 2, k, g.nz, fk0), k1 = cell_image(g, 2, k - 1, g.nz, fk1);
    const unsigned at[8] = {g.cell(i0, j0, k0), g.cell(i1, j0, k0), g.cell(i0, j1, k0), g.cell(i0, j0, k1),
                            g.cell(i1, j1, k0), g.cell(i1, j0, k1), g.cell(i0, j1, k1), g.cell(i1, j1, k1)};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double* f = b + size_t(c) * cells;
      s[c].base = b_ghost_sign(g, c, fi0, fj0, fk0) * f[at[0]];
      s[c].x = b_ghost_sign(g, c, fi1, fj0, fk0) * f[at[1]];
      s[c].y = b_ghost_sign(g, c, fi0, fj1, fk0) * f[at[2]];
      s[c].z = b_ghost_sign(g, c, fi0, fj0, fk1) * f[at[3]];
      s[c].xy = b_ghost_sign(g, c, fi1, fj1, fk0) * f[at[4]];
      s[c].xz = b_ghost_sign(g, c, fi1, fj0, fk1) * f[at[5]];
      s[c].yz = b_ghost_sign(g, c, fi0, fj1, fk1) * f[at[6]];
      s[c].xyz = b_ghost_sign(g, c, fi1, fj1, fk1) * f[at[7]];
    }
  }
  curl_of<false>(s, g, x, y, z);
}

__global__ __launch_bounds__(kThreads) void k_curl_n2c(Grid g, const double* __restrict__ v,
                                                       double* __restrict__ out, const double* __restrict__ stop) {
  if (stop && *stop != 0.0) return;
  const unsigned cells = g.cells();
  for (unsigned c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += gridDim.x * blockDim.x) {
    const int i = int(c % unsigned(g.nx)), j = int((c / unsigned(g.nx)) % unsigned(g.ny)),
              k = int(c / (unsigned(g.nx) * unsigned(g.ny)));
    double x, y, z;
    n2c_at(g, v, i, j, k, x, y, z);
    out[c] = x;
    out[cells + c] = y;
    out[2 * size_t(cells) + c] = z;
  }
}

// curl_C2N of a cell vector at every node (periodic wrap).
__global__ __launch_bounds__(kThreads) void k_curl_c2n(Grid g, const double* __restrict__ b,
                                                       double* __restrict__ out) {
  const unsigned nodes = g.nodes();
  const unsigned sx = unsigned(g.nx + 1), sy = unsigned(g.ny + 1);
  for (unsigned n = blockIdx.x * blockDim.x + threadIdx.x; n < nodes; n += gridDim.x * blockDim.x) {
    const int i = int(n % sx), j = int((n / sx) % sy), k = int(n / (sx * sy));
    double x, y, z;
    c2n_at<false>(g, b, i, j, k, x, y, z);
    out[n] = x;
    out[nodes + n] = y;
    out[2 * size_t(nodes) + n] = z;
  }
}

template <bool Walls>
__global__ __launch_bounds__(kThreads) void k_cell_to_node(Grid g, const double* __restrict__ b,
                                                           double* __restrict__ out) {
  const unsigned nodes = g.nodes(), cells = g.cells();
  const unsigned sx = unsigned(g.nx + 1), sy = unsigned(g.ny + 1);
  for (unsigned n = blockIdx.x * blockDim.x + threadIdx.x; n < nodes; n += gridDim.x * blockDim.x) {
    const int i = int(n % sx), j = int((n / sx) % sy), k = int(n / (sx * sy));
    if constexpr (!Walls) {
      const int i0 = wrap_cell(i - 1, g.nx), i1 = wrap_cell(i, g.nx);
      const int j0 = wrap_cell(j - 1, g.ny), j1 = wrap_cell(j, g.ny);
      const int k0 = wrap_cell(k - 1, g.nz), k1 = wrap_cell(k, g.nz);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double* f = b + size_t(c) * cells;
        out[size_t(c) * nodes + n] = (f[g.cell(i0, j0, k0)] + f[g.cell(i1, j0, k0)] + f[g.cell(i0, j1, k0)] +
                                      f[g.cell(i1, j1, k0)] + f[g.cell(i0, j0, k1)] + f[g.cell(i1, j0, k1)] +
                                      f[g.cell(i0, j1, k1)] + f[g.cell(i1, j1, k1)]) *
                                     0.125;
      }
    } else {
      int fi0, fi1, fj0, fj1, fk0, fk1;
      const int i0 = cell_image(g, 0, i - 1, g.nx, fi0), i1 = cell_image(g, 0, i, g.nx, fi1);
      const int j0 = cell_image(g, 1, j - 1, g.ny, fj0), j1 = cell_image(g, 1, j, g.ny, fj1);
      const int k0 = cell_image(g, 2, k - 1, g.nz, fk0), k1 = cell_image(g, 2, k, g.nz, fk1);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double* f = b + size_t(c) * cells;
        out[size_t(c) * nodes + n] =
            (b_ghost_sign(g, c, fi0, fj0, fk0) * f[g.cell(i0, j0, k0)] +
             b_ghost_sign(g, c, fi1, fj0, fk0) * f[g.cell(i1, j0, k0)] +
             b_ghost_sign(g, c, fi0, fj1, fk0) * f[g.cell(i0, j1, k0)] +
             b_ghost_sign(g, c, fi1, fj1, fk0) * f[g.cell(i1, j1, k0)] +
             b_ghost_sign(g, c, fi0, fj0, fk1) * f[g.cell(i0, j0, k1)] +
             b_ghost_sign(g, c, fi1, fj0, fk1) * f[g.cell(i1, j0, k1)] +
             b_ghost_sign(g, c, fi0, fj1, fk1) * f[g.cell(i0, j1, k1)] +
             b_ghost_sign(g, c, fi1, fj1, fk1) * f[g.cell(i1, j1, k1)]) *
            0.125;
      }
    }
  }
}

// Zero the constrained rows of a node vector (enqueue_constrained_rows
// without input, field_boundary_device.hpp:60-120).
__global__ __launch_bounds__(kThreads) void k_zero_constrained(Grid g, double* __restrict__ v) {
  const unsigned nodes = g.nodes();
  const unsigned sx = unsigned(g.nx + 1), sy = unsigned(g.ny + 1);
  for (unsigned n = blockIdx.x * blockDim.x + threadIdx.x; n < nodes; n += gridDim.x * blockDim.x) {
    const int i = int(n % sx), j = int((n / sx) % sy), k = int(n / (sx * sy));
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (g.constrained_row(c, i, j, k)) v[size_t(c) * nodes + n] = 0.0;
  }
}

// ---- energy-conserving smoothing ------------------------------------------
// One pass of the node stencil (node_smoothing_value, include/tpic/ecsim/
// fields/node_smoothing.hpp:27-58; node_smoothing_kernel,
// node_smoothing_device.hpp:20-41) at every node, reading neighbours through
// E-like ghosts: the periodic image, or on a wall axis the reflection through
// the wall node, negated where the face constrains the component (an odd
// ghost, boundary_conditions.hpp:56-72; field_boundary_device.hpp:26-57). A
// corner ghost carries the product of its faces' signs. Sums run in the
// reference's order: the six faces, the twelve edges, the eight corners, each
// left to right.
// MaxwellOperator::complete_smoothing (maxwell_device.hpp:452-490) projects
// the source before the first pass (complete_electric_refresh, :114-121, or
// the input projection of begin_apply_input, :270-272) and every pass's
// result: each pass here reads the source with its constrained rows as zero
// and writes its result with them zeroed, so S = P (pass P)^n.
constexpr int kFace7 = TPIC_MI_SMOOTH_FACE7, kBinomial27 = TPIC_MI_SMOOTH_BINOMIAL27;

// The E-like images of the neighbours i - 1, i, i + 1 along one axis: the
// node index, and per component bit (1 << c) whether the image is negated
// (an odd ghost) or read as zero (a constrained row, the projected source).
struct AxisImages {
  int at[3];
  unsigned char negate[3], zero[3];
};

template <bool Walls>
__device__ __forceinline__ AxisImages axis_images(const Grid& g, int axis, int i, int n) {
  AxisImages r;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int x = i + d - 1;
    r.negate[d] = 0;
    r.zero[d] = 0;
    if (!Walls || !g.field_wall[axis]) {
      r.at[d] = wrap_node(x, n);
      continue;
    }
    int image = x;
    if (x < 0) {
      image = -x;
      r.negate[d] = g.constrained[2 * axis];
    } else if (x > n) {
      image = 2 * n - x;
      r.negate[d] = g.constrained[2 * axis + 1];
    }
    r.at[d] = image;
    r.zero[d] = image == 0 ? g.constrained[2 * axis] : (image == n ? g.constrained[2 * axis + 1] : 0);
  }
  return r;
}

template <int Op, bool Walls>
__global__ __launch_bounds__(kThreads) void k_smooth(Grid g, const double* __restrict__ src,
                                                     double* __restrict__ dst, double alpha, double beta,
                                                     unsigned* __restrict__ validation) {
  const unsigned nodes = g.nodes();
  const unsigned sx = unsigned(g.nx + 1), sy = unsigned(g.ny + 1);
  unsigned flags = 0;
  for (unsigned n = blockIdx.x * blockDim.x + threadIdx.x; n < nodes; n += gridDim.x * blockDim.x) {
    const int i = int(n % sx), j = int((n / sx) % sy), k = int(n / (sx * sy));
    // The images are computed once per node and shared by the components.
    const AxisImages ax = axis_images<Walls>(g, 0, i, g.nx), ay = axis_images<Walls>(g, 1, j, g.ny),
                     az = axis_images<Walls>(g, 2, k, g.nz);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double* f = src + size_t(c) * nodes;
      const unsigned bit = 1u << c;
      auto at = [&](int di, int dj, int dk) {
        const double v = f[unsigned(ax.at[di + 1]) + sx * (unsigned(ay.at[dj + 1]) + sy * unsigned(az.at[dk + 1]))];
        if (!Walls) return v;
        if ((ax.zero[di + 1] | ay.zero[dj + 1] | az.zero[dk + 1]) & bit) return 0.0;
        return ((ax.negate[di + 1] ^ ay.negate[dj + 1] ^ az.negate[dk + 1]) & bit) ? -v : v;
      };
      const double face = at(-1, 0, 0) + at(1, 0, 0) + at(0, -1, 0) + at(0, 1, 0) + at(0, 0, -1) + at(0, 0, 1);
      double value;
      if (Op == kFace7) {
        value = alpha * at(0, 0, 0) + beta * face;
      } else {
        const double edge = at(-1, -1, 0) + at(1, -1, 0) + at(-1, 1, 0) + at(1, 1, 0) + at(-1, 0, -1) +
                            at(1, 0, -1) + at(-1, 0, 1) + at(1, 0, 1) + at(0, -1, -1) + at(0, 1, -1) +
                            at(0, -1, 1) + at(0, 1, 1);
        const double corner = at(-1, -1, -1) + at(1, -1, -1) + at(-1, 1, -1) + at(1, 1, -1) + at(-1, -1, 1) +
                              at(1, -1, 1) + at(-1, 1, 1) + at(1, 1, 1);
        value = (8.0 * at(0, 0, 0) + 4.0 * face + 2.0 * edge + corner) * (1.0 / 64.0);
      }
      if (!finite_value(value)) flags |= kInvalidField;
      if (Walls && g.constrained_row(c, i, j, k)) value = 0.0;
      dst[size_t(c) * nodes + n] = value;
    }
  }
  if (flags) atomicOr(validation, flags);
}

template <bool Walls>
__global__ __launch_bounds__(kThreads) void k_rhs(Grid g, const double* __restrict__ e, const double* __restrict__ b,
                                                  const double* __restrict__ current, double delta,
                                                  double current_factor, double* __restrict__ rhs,
                                                  unsigned* __restrict__ validation) {
  const unsigned nodes = g.nodes();
  const unsigned sx = unsigned(g.nx + 1), sy = unsigned(g.ny + 1);
  unsigned flags = 0;
  for (unsigned n = blockIdx.x * blockDim.x + threadIdx.x; n < nodes; n += gridDim.x * blockDim.x) {
    const int i = int(n % sx), j = int((n / sx) % sy), k = int(n / (sx * sy));
    double cv[3];
    c2n_at<Walls>(g, b, i, j, k, cv[0], cv[1], cv[2]);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t at = size_t(c) * nodes + n;
      double value = 1.0 * e[at] + delta * cv[c] + (-current_factor) * current[at];
      if (Walls && g.constrained_row(c, i, j, k)) value = 0.0;
      rhs[at] = value;
      if (!finite_value(value)) flags |= kInvalidField;
    }
  }
  if (flags) atomicOr(validation, flags);
}

__constant__ int c_offsets[kStencils][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0},
                                            {1, -1, 0}, {1, 0, 1}, {1, 0, -1}, {0, 1, 1}, {0, 1, -1},
                                            {1, -1, 1}, {1, 1, -1}, {-1, 1, 1}, {1, 1, 1}};

// image = 1*v + d2*curl_C2N(curl_cells) + k*(1/V)*M v at every interior node.
// M is read in its storage type (double, or float for the mixed forms); the
// products and sums are FP64.
// Rows per traversal tile of the operator apply (measured at 128^3: 16 and
// 32 rows are 6% faster than plane order for the FP64 matrix, 4% for FP32).
constexpr unsigned kApplyTileRows = 16;

// (1/V) f_g (M v) at node (i, j, kk) = n: the centre block, then per stencil
// the forward block times v(i + o) and the backward block of the neighbour
// vertex (not transposed) times v(i - o); FP64 products and sums.
// With Transposed every 3 x 3 block enters transposed: (1/V) M^T v, the mass
// part of the adjoint operator (the backward blocks are the forward pair's,
// not transposed, so M^T needs the transpose of every block).
template <class MassT, bool Walls, bool Transposed = false>
__device__ __forceinline__ void mass_rows(const Grid& g, const double* __restrict__ v,
                                          const MassT* __restrict__ mass, double inverse_volume, int i, int j,
                                          int kk, unsigned n, double& mx, double& my, double& mz) {
  // Entry (row, column) of a block sits at (3 row + column) vs, or at
  // (3 column + row) vs transposed.
  constexpr int e01 = Transposed ? 3 : 1, e02 = Transposed ? 6 : 2, e10 = Transposed ? 1 : 3;
  constexpr int e12 = Transposed ? 7 : 5, e20 = Transposed ? 2 : 6, e21 = Transposed ? 5 : 7;
  const unsigned nodes = g.nodes();
  const unsigned ex = unsigned(g.nx + 3), exy = ex * unsigned(g.ny + 3);
  const size_t vs = g.vstride;
  const double* vx = v;
  const double* vy = v + nodes;
  const double* vz = v + 2 * size_t(nodes);
  const unsigned v0 = g.vertex(i, j, kk);
  // Node images of the neighbours along each axis (offsets -1, 0, +1): on a
  // wall axis the reflection through the wall node, else the periodic image.
  int xi[3], yi[3], zi[3];
  if (Walls) {
    xi[0] = node_image(g, 0, i - 1, g.nx), xi[1] = i, xi[2] = node_image(g, 0, i + 1, g.nx);
    yi[0] = node_image(g, 1, j - 1, g.ny), yi[1] = j, yi[2] = node_image(g, 1, j + 1, g.ny);
    zi[0] = node_image(g, 2, kk - 1, g.nz), zi[1] = kk, zi[2] = node_image(g, 2, kk + 1, g.nz);
  }
  {
    const double ax = vx[n], ay = vy[n], az = vz[n];
    const MassT* b = mass + v0;
    mx = ax * double(b[0]) + ay * double(b[e01 * vs]) + az * double(b[e02 * vs]);
    my = ax * double(b[e10 * vs]) + ay * double(b[4 * vs]) + az * double(b[e12 * vs]);
    mz = ax * double(b[e20 * vs]) + ay * double(b[e21 * vs]) + az * double(b[8 * vs]);
  }
#pragma unroll
  for (int s = 1; s < kStencils; ++s) {
    const int ox = c_offsets[s][0], oy = c_offsets[s][1], oz = c_offsets[s][2];
    // Forward: this vertex's block times v(i + o).
    {
      const unsigned at = Walls ? g.node(xi[1 + ox], yi[1 + oy], zi[1 + oz])
                                : g.node(wrap_node(i + ox, g.nx), wrap_node(j + oy, g.ny), wrap_node(kk + oz, g.nz));
      const double ax = vx[at], ay = vy[at], az = vz[at];
      const MassT* b = mass + size_t(s) * 9 * vs + v0;
      mx += ax * double(b[0]) + ay * double(b[e01 * vs]) + az * double(b[e02 * vs]);
      my += ax * double(b[e10 * vs]) + ay * double(b[4 * vs]) + az * double(b[e12 * vs]);
      mz += ax * double(b[e20 * vs]) + ay * double(b[e21 * vs]) + az * double(b[8 * vs]);
    }
    // Backward: the neighbour vertex's block (not transposed) times v(i - o).
    {
      const unsigned at = Walls ? g.node(xi[1 - ox], yi[1 - oy], zi[1 - oz])
                                : g.node(wrap_node(i - ox, g.nx), wrap_node(j - oy, g.ny), wrap_node(kk - oz, g.nz));
      const double ax = vx[at], ay = vy[at], az = vz[at];
      const unsigned vb = unsigned(int(v0) - ox - oy * int(ex) - oz * int(exy));
      const MassT* b = mass + size_t(s) * 9 * vs + vb;
      mx += ax * double(b[0]) + ay * double(b[e01 * vs]) + az * double(b[e02 * vs]);
      my += ax * double(b[e10 * vs]) + ay * double(b[4 * vs]) + az * double(b[e12 * vs]);
      mz += ax * double(b[e20 * vs]) + ay * double(b[e21 * vs]) + az * double(b[8 * vs]);
    }
  }
  // mass_matrix_apply_device.hpp:34-35: the row scale (1/V) f_g.
  const double row_scale = Walls ? inverse_volume * double(g.dual_factor(i, j, kk)) : inverse_volume;
  mx *= row_scale;
  my *= row_scale;
  mz *= row_scale;
}

// Stages of k_apply: the fused image; with smoothing, the mass response
// (1/V) f_g M v alone (enqueue_mass_response), then the combination with the
// smoothed response (combine_image, maxwell_device.hpp:419-433).
constexpr int kFused = 0, kMassResponse = 1, kCombine = 2;

template <class MassT, bool Walls, int Stage = kFused, bool Transposed = false>
__global__ __launch_bounds__(kThreads) void k_apply(Grid g, const double* __restrict__ v,
                                                    const double* __restrict__ curl_cells,
                                                    const MassT* __restrict__ mass,
                                                    const double* __restrict__ response, double d2, double k,
                                                    double inverse_volume, double* __restrict__ out,
                                                    unsigned tile_rows, const double* __restrict__ stop) {
  // A launch of a GMRES step after the solver stopped does nothing.
  if (stop && *stop != 0.0) return;
  const unsigned nodes = g.nodes();
  const unsigned sx = unsigned(g.nx + 1), sy = unsigned(g.ny + 1), sz = unsigned(g.nz + 1);
  const unsigned L = xcd_logical_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (L >= nodes) return;
  // Traversal: full x rows, grouped into tiles of tile_rows rows; a tile's
  // planes are visited in z order, so the blocks of the planes k - 1 and k of
  // a tile are close in time (within one XCD) and the backward reads of the
  // stencils with a z offset find the matrix lines in L2.
  const int i = int(L % sx);
  const unsigned row = L / sx;
  const unsigned full = sy / tile_rows, span = sz * tile_rows;
  int j, kk;
  if (row < full * span) {
    const unsigned t = row / span, rem = row % span;
    kk = int(rem / tile_rows);
    j = int(t * tile_rows + rem % tile_rows);
  } else {
    const unsigned last = sy - full * tile_rows, rem = row - full * span;
    kk = int(rem / last);
    j = int(full * tile_rows + rem % last);
  }
  const unsigned n = g.node(i, j, kk);
  const double* vx = v;
  const double* vy = v + nodes;
  const double* vz = v + 2 * size_t(nodes);
  double mx, my, mz;
  if (Stage == kCombine) {
    mx = response[n];
    my = response[nodes + n];
    mz = response[2 * size_t(nodes) + n];
  } else {
    mass_rows<MassT, Walls, Transposed>(g, v, mass, inverse_volume, i, j, kk, n, mx, my, mz);
  }
  double cx = 0, cy = 0, cz = 0;
  if (Stage != kMassResponse) c2n_at<Walls>(g, curl_cells, i, j, kk, cx, cy, cz);
  if (Stage == kMassResponse) {
    out[n] = mx;
    out[nodes + n] = my;
    out[2 * size_t(nodes) + n] = mz;
    return;
  }
  double ox = 1.0 * vx[n] + d2 * cx + k * mx;
  double oy = 1.0 * vy[n] + d2 * cy + k * my;
  double oz = 1.0 * vz[n] + d2 * cz + k * mz;
  if (Walls) {
    // Constrained image rows take the (projected) input rows.
    if (g.constrained_row(0, i, j, kk)) ox = vx[n];
    if (g.constrained_row(1, i, j, kk)) oy = vy[n];
    if (g.constrained_row(2, i, j, kk)) oz = vz[n];
  }
  out[n] = ox;
  out[nodes + n] = oy;
  out[2 * size_t(nodes) + n] = oz;
}

__global__ __launch_bounds__(kThreads) void k_next_e(unsigned count, const double* __restrict__ etheta,
                                                     double* __restrict__ e, double c0, double c1, int copy,
                                                     unsigned* __restrict__ validation) {
  unsigned flags = 0;
  for (unsigned n = blockIdx.x * blockDim.x + threadIdx.x; n < count; n += gridDim.x * blockDim.x) {
    const double value = copy ? etheta[n] : c0 * etheta[n] + c1 * e[n];
    e[n] = value;
    if (!finite_value(value)) flags |= kInvalidField;
  }
  if (flags) atomicOr(validation, flags);
}

__global__ __launch_bounds__(kThreads) void k_faraday(Grid g, const double* __restrict__ etheta,
                                                      double* __restrict__ b, double faraday,
                                                      unsigned* __restrict__ validation) {
  const unsigned cells = g.cells();
  unsigned flags = 0;
  for (unsigned c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += gridDim.x * blockDim.x) {
    const int i = int(c % unsigned(g.nx)), j = int((c / unsigned(g.nx)) % unsigned(g.ny)),
              k = int(c / (unsigned(g.nx) * unsigned(g.ny)));
    double r[3];
    n2c_at(g, etheta, i, j, k, r[0], r[1], r[2]);
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
      const size_t at = size_t(comp) * cells + c;
      const double value = 1.0 * b[at] + faraday * r[comp];
      b[at] = value;
      if (!finite_value(value)) flags |= kInvalidField;
    }
  }
  if (flags) atomicOr(validation, flags);
}

unsigned blocks_for(const Context& ctx, uint64_t work) {
  const uint64_t want = (work + kThreads - 1) / kThreads;
  const uint64_t cap = uint64_t(ctx.cus > 0 ? ctx.cus : 256) * 32;
  return unsigned(want < cap ? (want ? want : 1) : cap);
}

}  // namespace

int launch_cell_to_node(Context& ctx) {
  if (ctx.grid.walls)
    k_cell_to_node<true><<<blocks_for(ctx, ctx.grid.nodes()), kThreads, 0, ctx.stream>>>(ctx.grid, ctx.b, ctx.bn);
  else
    k_cell_to_node<false><<<blocks_for(ctx, ctx.grid.nodes()), kThreads, 0, ctx.stream>>>(ctx.grid, ctx.b, ctx.bn);
  return check_hip(ctx, hipGetLastError(), "cell to node");
}

int launch_zero_constrained(Context& ctx, double* v) {
  if (!ctx.grid.walls) return 0;
  k_zero_constrained<<<blocks_for(ctx, ctx.grid.nodes()), kThreads, 0, ctx.stream>>>(ctx.grid, v);
  return check_hip(ctx, hipGetLastError(), "constrained rows");
}

int launch_smooth(Context& ctx, const double* src, double* dst, const double** result) {
  const Grid& g = ctx.grid;
  const Numerics& k = ctx.num;
  if (!k.smoothing || !ctx.smooth_a || !ctx.smooth_b)
    return fail(ctx, TPIC_MI_INVALID_STATE, "smoothing is not active");
  double* buffers[2] = {ctx.smooth_a, ctx.smooth_b};
  int next = src == ctx.smooth_a ? 1 : 0;
  const double* from = src;
  const unsigned blocks = blocks_for(ctx, g.nodes());
  for (int pass = 0; pass < k.smooth_iterations; ++pass) {
    const bool last = pass + 1 == k.smooth_iterations;
    double* to = last && dst ? dst : buffers[next];
    if (k.smooth_operator == TPIC_MI_SMOOTH_FACE7 && g.walls)
      k_smooth<kFace7, true><<<blocks, kThreads, 0, ctx.stream>>>(g, from, to, k.smooth_alpha, k.smooth_beta,
                                                                 ctx.validation);
    else if (k.smooth_operator == TPIC_MI_SMOOTH_FACE7)
      k_smooth<kFace7, false><<<blocks, kThreads, 0, ctx.stream>>>(g, from, to, k.smooth_alpha, k.smooth_beta,
                                                                  ctx.validation);
    else if (g.walls)
      k_smooth<kBinomial27, true><<<blocks, kThreads, 0, ctx.stream>>>(
          g, from, to, k.smooth_alpha, k.smooth_beta, ctx.validation);
    else
      k_smooth<kBinomial27, false><<<blocks, kThreads, 0, ctx.stream>>>(
          g, from, to, k.smooth_alpha, k.smooth_beta, ctx.validation);
    if (int rc = check_hip(ctx, hipGetLastError(), "node smoothing")) return rc;
    from = to;
    next ^= 1;
  }
  *result = from;
  return 0;
}

int launch_rhs(Context& ctx) {
  // complete_rhs (maxwell_device.hpp:151-171): with smoothing the current
  // term is S(J).
  const double* current = ctx.current;
  if (ctx.num.smoothing)
    if (int rc = launch_smooth(ctx, ctx.current, nullptr, &current)) return rc;
  if (ctx.grid.walls)
    k_rhs<true><<<blocks_for(ctx, ctx.grid.nodes()), kThreads, 0, ctx.stream>>>(
        ctx.grid, ctx.e, ctx.b, current, ctx.delta, ctx.current_factor, ctx.rhs, ctx.validation);
  else
    k_rhs<false><<<blocks_for(ctx, ctx.grid.nodes()), kThreads, 0, ctx.stream>>>(
        ctx.grid, ctx.e, ctx.b, current, ctx.delta, ctx.current_factor, ctx.rhs, ctx.validation);
  return check_hip(ctx, hipGetLastError(), "rhs");
}

namespace {

// complete_apply_smoothed (maxwell_device.hpp:389-417): x = S(v),
// r = (1/V) f_g M x, y = S(r), image = 1*v + d2*curl_C2N(curl_cells) + k*y
// with the constrained rows of v.
template <class MassT>
int smoothed_apply(Context& ctx, const double* in, const MassT* mass, double* out, unsigned blocks, unsigned rows) {
  const Grid& g = ctx.grid;
  const double d2 = ctx.delta * ctx.delta;
  const double* x = nullptr;
  if (int rc = launch_smooth(ctx, in, nullptr, &x)) return rc;
  double* r = x == ctx.smooth_a ? ctx.smooth_b : ctx.smooth_a;
  if (g.walls)
    k_apply<MassT, true, kMassResponse><<<blocks, kThreads, 0, ctx.stream>>>(
        g, x, nullptr, mass, nullptr, d2, ctx.current_factor, ctx.inverse_volume, r, rows, nullptr);
  else
    k_apply<MassT, false, kMassResponse><<<blocks, kThreads, 0, ctx.stream>>>(
        g, x, nullptr, mass, nullptr, d2, ctx.current_factor, ctx.inverse_volume, r, rows, nullptr);
  if (int rc = check_hip(ctx, hipGetLastError(), "mass response")) return rc;
  const double* y = nullptr;
  if (int rc = launch_smooth(ctx, r, nullptr, &y)) return rc;
  if (g.walls)
    k_apply<double, true, kCombine><<<blocks, kThreads, 0, ctx.stream>>>(
        g, in, ctx.gmres.curl_cells, nullptr, y, d2, ctx.current_factor, ctx.inverse_volume, out, rows, nullptr);
  else
    k_apply<double, false, kCombine><<<blocks, kThreads, 0, ctx.stream>>>(
        g, in, ctx.gmres.curl_cells, nullptr, y, d2, ctx.current_factor, ctx.inverse_volume, out, rows, nullptr);
  return check_hip(ctx, hipGetLastError(), "operator combination");
}

}  // namespace

int launch_operator(Context& ctx, const double* in, double* out, const double* stop) {
  const Grid& g = ctx.grid;
  phase_begin(ctx, Phase::spmv);
  // The stop word reaches the kernels of the unsmoothed operator only.
  if (ctx.num.smoothing) stop = nullptr;
  k_curl_n2c<<<blocks_for(ctx, g.cells()), kThreads, 0, ctx.stream>>>(g, in, ctx.gmres.curl_cells, stop);
  int rc = check_hip(ctx, hipGetLastError(), "operator curl");
  if (rc) return rc;
  const unsigned blocks = (g.nodes() + kThreads - 1) / kThreads;
  const unsigned rows = std::min(kApplyTileRows, unsigned(g.ny + 1));
  const double d2 = ctx.delta * ctx.delta;
  const float* mf = static_cast<const float*>(ctx.mass);
  const double* md = static_cast<const double*>(ctx.mass);
  if (ctx.transposed_operator) {
    // The adjoint operator of a periodic box: curl-curl is symmetric over
    // the unique nodes, the mass blocks enter transposed.
    if (g.walls || ctx.num.smoothing) {
      phase_end(ctx, Phase::spmv);
      return fail(ctx, TPIC_MI_INVALID_STATE, "the transposed operator is built for periodic fields without smoothing");
    }
    if (ctx.mass_single)
      k_apply<float, false, kFused, true><<<blocks, kThreads, 0, ctx.stream>>>(
          g, in, ctx.gmres.curl_cells, mf, nullptr, d2, ctx.current_factor, ctx.inverse_volume, out, rows, stop);
    else
      k_apply<double, false, kFused, true><<<blocks, kThreads, 0, ctx.stream>>>(
          g, in, ctx.gmres.curl_cells, md, nullptr, d2, ctx.current_factor, ctx.inverse_volume, out, rows, stop);
    rc = check_hip(ctx, hipGetLastError(), "transposed operator apply");
    phase_end(ctx, Phase::spmv);
    return rc;
  }
  if (ctx.num.smoothing) {
    rc = ctx.mass_single ? smoothed_apply(ctx, in, mf, out, blocks, rows) : smoothed_apply(ctx, in, md, out, blocks, rows);
    phase_end(ctx, Phase::spmv);
    return rc;
  }
  if (g.walls) {
    // Two stages, the mass response (1/V) f_g M v into a node buffer, then
    // the image with the curl term: fused, the wall images and ghost signs of
    // both terms held the kernel to one wave per SIMD (256 registers). The
    // sums are those of the fused form.
    double* r = ctx.mass_response;
    if (!r) {
      phase_end(ctx, Phase::spmv);
      return fail(ctx, TPIC_MI_INVALID_STATE, "walled operator without its mass response buffer");
    }
    if (ctx.mass_single)
      k_apply<float, true, kMassResponse><<<blocks, kThreads, 0, ctx.stream>>>(
          g, in, nullptr, mf, nullptr, d2, ctx.current_factor, ctx.inverse_volume, r, rows, stop);
    else
      k_apply<double, true, kMassResponse><<<blocks, kThreads, 0, ctx.stream>>>(
          g, in, nullptr, md, nullptr, d2, ctx.current_factor, ctx.inverse_volume, r, rows, stop);
    if ((rc = check_hip(ctx, hipGetLastError(), "walled mass response"))) {
      phase_end(ctx, Phase::spmv);
      return rc;
    }
    k_apply<double, true, kCombine><<<blocks, kThreads, 0, ctx.stream>>>(
        g, in, ctx.gmres.curl_cells, nullptr, r, d2, ctx.current_factor, ctx.inverse_volume, out, rows, stop);
    rc = check_hip(ctx, hipGetLastError(), "walled operator combination");
    phase_end(ctx, Phase::spmv);
    return rc;
  }
  if (ctx.mass_single)
    k_apply<float, false><<<blocks, kThreads, 0, ctx.stream>>>(g, in, ctx.gmres.curl_cells, mf, nullptr, d2,
                                                               ctx.current_factor, ctx.inverse_volume, out, rows, stop);
  else
    k_apply<double, false><<<blocks, kThreads, 0, ctx.stream>>>(g, in, ctx.gmres.curl_cells, md, nullptr, d2,
                                                                ctx.current_factor, ctx.inverse_volume, out, rows, stop);
  rc = check_hip(ctx, hipGetLastError(), "operator apply");
  phase_end(ctx, Phase::spmv);
  return rc;
}

int launch_curl_c2n(Context& ctx, const double* cells_in, double* nodes_out) {
  if (ctx.grid.walls) return fail(ctx, TPIC_MI_INVALID_STATE, "the standalone node curl is built for periodic fields");
  k_curl_c2n<<<blocks_for(ctx, ctx.grid.nodes()), kThreads, 0, ctx.stream>>>(ctx.grid, cells_in, nodes_out);
  return check_hip(ctx, hipGetLastError(), "node curl");
}

int launch_curl_n2c(Context& ctx, const double* nodes_in, double* cells_out) {
  k_curl_n2c<<<blocks_for(ctx, ctx.grid.cells()), kThreads, 0, ctx.stream>>>(ctx.grid, nodes_in, cells_out,
                                                                              nullptr);
  return check_hip(ctx, hipGetLastError(), "cell curl");
}

int launch_next_fields(Context& ctx) {
  const Grid& g = ctx.grid;
  const double th = ctx.num.theta;
  const bool copy = std::abs(th - 1.0) < 1e-10;
  const double c0 = 1.0 / th, c1 = -(1.0 - th) / th;
  phase_begin(ctx, Phase::advance);
  k_next_e<<<blocks_for(ctx, 3ull * g.nodes()), kThreads, 0, ctx.stream>>>(3 * g.nodes(), ctx.etheta, ctx.e, c0,
                                                                           c1, copy ? 1 : 0, ctx.validation);
  int rc = check_hip(ctx, hipGetLastError(), "E update");
  if (rc) return rc;
  k_faraday<<<blocks_for(ctx, g.cells()), kThreads, 0, ctx.stream>>>(g, ctx.etheta, ctx.b, -ctx.num.c * ctx.num.dt,
                                                                     ctx.validation);
  if ((rc = check_hip(ctx, hipGetLastError(), "Faraday update"))) return rc;
  rc = launch_cell_to_node(ctx);
  phase_end(ctx, Phase::advance);
  return rc;
}

}  // namespace tpic_mi
