// Builds the reference's own host-callable functions of the hot path into
// oracle/_ref/libtpic_ref.so, so the oracle can be checked against them.
// Test infrastructure only. The headers are read from the reference checkout
// (REF in oracle/Makefile); nothing of the reference is copied into this
// repository. Only headers that compile without generated files are used:
//   include/tpic/ecsim/particles/cic.hpp                 make_cic_weights
//   include/tpic/ecsim/particles/particle_response.hpp   ParticleResponse
//   include/tpic/ecsim/mass_matrix/mass_matrix_stencil.hpp  mass_pair_stencil,
//                                                        apply_mass_matrix_node
//   include/tpic/ecsim/fields/node_cell_curl.hpp         node_cell_curl
//   include/tpic/ecsim/fields/cell_to_node.hpp           cell_to_node_average
// The mover, boundary fold and host sort include LPFx's generated
// lpfx/config.hpp and are therefore not built here.
// The per-node and per-cell loops run over z planes with OpenMP (every output
// is written by one iteration, so the results do not depend on the thread
// count); the CPU baseline times them on all host cores and on one.
#include <tpic/ecsim/fields/cell_to_node.hpp>
#include <tpic/ecsim/fields/node_cell_curl.hpp>
#include <tpic/ecsim/mass_matrix/mass_matrix_stencil.hpp>
#include <tpic/ecsim/particles/cic.hpp>
#include <tpic/ecsim/particles/particle_response.hpp>

#include <cstddef>

namespace {

// A one-ghost scalar component addressed by logical (i, j, k).
template <class Location> struct Component {
  using location_type = Location;
  using value_type = double;
  const double* base;
  int sx, sy;  // padded extents along x and y
  int halo;    // ghost layers before logical index 0
  double operator()(int i, int j, int k) const {
    return base[(i + halo) + static_cast<std::size_t>(sx) *
                                 ((j + halo) + static_cast<std::size_t>(sy) * (k + halo))];
  }
};

template <class Location> struct Vector {
  static constexpr int component_count = 3;
  Component<Location> c0, c1, c2;
  template <std::size_t I> Component<Location> component() const {
    if constexpr (I == 0) return c0;
    else if constexpr (I == 1) return c1;
    else return c2;
  }
};

tpic::ecsim::NodeCellSpacing<double> spacing_of(const double* spacing) {
  tpic::ecsim::NodeCellSpacing<double> s;
  (void)tpic::ecsim::try_make_node_cell_spacing<double>({spacing[0], spacing[1], spacing[2]}, s);
  return s;
}

}  // namespace

extern "C" {

void ref_cic_weights(double rx, double ry, double rz, double* w) {
  const auto weights = tpic::ecsim::make_cic_weights<double>(rx, ry, rz);
  for (int a = 0; a < 8; ++a) w[a] = weights.value[a];
}

void ref_response(double bx, double by, double bz, double beta_over_c, double* a, double* norm_square) {
  const tpic::ecsim::ParticleResponse<double> r(bx, by, bz, beta_over_c);
  for (int c = 0; c < 9; ++c) a[c] = r.a[c];
  *norm_square = r.norm_square;
}

int ref_mass_pair_stencil(int a, int b) { return tpic::ecsim::mass_pair_stencil(a, b); }

// Node arrays with one ghost layer: extents (n+3) per axis.
void ref_mass_apply(const int* n, const double* mass, const double* vx, const double* vy,
                    const double* vz, double inverse_volume, double* out) {
  const int ex = n[0] + 3, ey = n[1] + 3, ez = n[2] + 3;
  const tpic::data::MassMatrixLayout layout{n[0] + 1, n[1] + 1, n[2] + 1, ex, ey, ez, ex, ex * ey,
                                            (n[0] + 1) * (n[1] + 1) * (n[2] + 1), ex * ey * ez};
  const Component<lpfx::Node> x{vx, ex, ey, 1}, y{vy, ex, ey, 1}, z{vz, ex, ey, 1};
  const std::size_t nodes = static_cast<std::size_t>(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
#pragma omp parallel for schedule(static)
  for (int k = 0; k <= n[2]; ++k)
    for (int j = 0; j <= n[1]; ++j)
      for (int i = 0; i <= n[0]; ++i) {
        double ox, oy, oz;
        tpic::ecsim::apply_mass_matrix_node<double>(mass, layout, x, y, z, i, j, k, inverse_volume,
                                                    ox, oy, oz);
        const std::size_t at = i + static_cast<std::size_t>(n[0] + 1) * (j + static_cast<std::size_t>(n[1] + 1) * k);
        out[at] = ox;
        out[nodes + at] = oy;
        out[2 * nodes + at] = oz;
      }
}

// Node source (one ghost, extents n+3) to interior cells (compact).
void ref_curl_n2c(const int* n, const double* spacing, const double* fx, const double* fy,
                  const double* fz, double* out) {
  const int ex = n[0] + 3, ey = n[1] + 3;
  const Vector<lpfx::Node> v{{fx, ex, ey, 1}, {fy, ex, ey, 1}, {fz, ex, ey, 1}};
  const auto s = spacing_of(spacing);
  const std::size_t cells = static_cast<std::size_t>(n[0]) * n[1] * n[2];
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i) {
        const auto c = tpic::ecsim::node_cell_curl<double>(v, i, j, k, s);
        const std::size_t at = i + static_cast<std::size_t>(n[0]) * (j + static_cast<std::size_t>(n[1]) * k);
        out[at] = c.x;
        out[cells + at] = c.y;
        out[2 * cells + at] = c.z;
      }
}

// Cell source (one ghost, extents n+2) to interior nodes (compact).
void ref_curl_c2n(const int* n, const double* spacing, const double* fx, const double* fy,
                  const double* fz, double* out) {
  const int ex = n[0] + 2, ey = n[1] + 2;
  const Vector<lpfx::Cell> v{{fx, ex, ey, 1}, {fy, ex, ey, 1}, {fz, ex, ey, 1}};
  const auto s = spacing_of(spacing);
  const std::size_t nodes = static_cast<std::size_t>(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
#pragma omp parallel for schedule(static)
  for (int k = 0; k <= n[2]; ++k)
    for (int j = 0; j <= n[1]; ++j)
      for (int i = 0; i <= n[0]; ++i) {
        const auto c = tpic::ecsim::node_cell_curl<double>(v, i, j, k, s);
        const std::size_t at = i + static_cast<std::size_t>(n[0] + 1) * (j + static_cast<std::size_t>(n[1] + 1) * k);
        out[at] = c.x;
        out[nodes + at] = c.y;
        out[2 * nodes + at] = c.z;
      }
}

// Cell source (one ghost) to interior nodes, one component.
void ref_cell_to_node(const int* n, const double* f, double* out) {
  const Component<lpfx::Cell> c{f, n[0] + 2, n[1] + 2, 1};
  for (int k = 0; k <= n[2]; ++k)
    for (int j = 0; j <= n[1]; ++j)
      for (int i = 0; i <= n[0]; ++i)
        out[i + static_cast<std::size_t>(n[0] + 1) * (j + static_cast<std::size_t>(n[1] + 1) * k)] =
            tpic::ecsim::cell_to_node_average<double>(c, i, j, k);
}

}  // extern "C"
