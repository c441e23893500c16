// ECSIM particle mover: one coalesced SoA pass over a species.
//
// Per particle (detail::push_and_classify_particle,
// include/tpic/ecsim/particles/particle_mover_device.hpp:35-122):
//   W = CIC weights of the stored fractions; B, E_theta gathered at the 8 cell
//   corners with W; alpha = response(B) (particle_response.hpp:17-34);
//   v_new = 2*(alpha v + beta alpha E_theta) - v; per axis the displacement
//   dt*v_new/dx is added to the fraction, the integer shift floor(moved) moves
//   the cell and an exact 1.0 carries; then the whole-domain periodic/mirror
//   fold of communication::classify_particle_point
//   (include/tpic/communication/particles/particle_boundaries.hpp:111-165).
// A rejected record keeps every attribute and raises the validation word.
//
// Contraction is off in this translation unit (-ffp-contract=off and the
// pragma below): every product and sum rounds on its own, in the reference's
// expression order, so the result is bit-identical to the CPU oracle.
// Bound: HBM, 60 B read + 52 B written per particle; E_theta and B at nodes
// (2 x 3 x 8 B per node) stay in L2 because the particles arrive cell-sorted.
#include "internal.hpp"

#pragma clang fp contract(off)

namespace tpic_mi {
namespace {

constexpr unsigned kThreads = 256;

struct MoverParams {
  double beta, beta_over_c, dt, inverse_spacing[3];
};

__device__ __forceinline__ double gather8(const double* __restrict__ f, const unsigned* at, const double* w) {
  double value = 0.0;
#pragma unroll
  for (int a = 0; a < 8; ++a) value += w[a] * f[at[a]];
  return value;
}

__device__ __forceinline__ int64_t positive_modulo(int64_t cell, int64_t extent) {
  const int64_t r = cell % extent;
  return r < 0 ? r + extent : r;
}

__global__ __launch_bounds__(kThreads) void k_move(Grid g, MoverParams p, ParticleView s, uint32_t n,
                                                   const double* __restrict__ etheta,
                                                   const double* __restrict__ bnode,
                                                   unsigned* __restrict__ validation) {
  const unsigned nodes = g.nodes();
  const int32_t bins = int32_t(g.bins());
  unsigned flags = 0;
  for (uint32_t index = blockIdx.x * blockDim.x + threadIdx.x; index < n; index += gridDim.x * blockDim.x) {
    const int32_t key = s.cell[index];
    int ci, cj, ck;
    if (key < 0 || key >= bins) { flags |= kInvalidKey; continue; }
    g.decode(key, ci, cj, ck);
    if (!g.interior(ci, cj, ck)) { flags |= kInvalidKey; continue; }
    if (!finite_value(s.q[index])) { flags |= kInvalidNonfinite; continue; }
    const double fx = s.rx[index], fy = s.ry[index], fz = s.rz[index];
    if (!(fx >= 0.0 && fx < 1.0 && fy >= 0.0 && fy < 1.0 && fz >= 0.0 && fz < 1.0)) {
      flags |= kInvalidFraction;
      continue;
    }
    const double cx = 1.0 - fx, cy = 1.0 - fy, cz = 1.0 - fz;
    const double w[8] = {cx * cy * cz, fx * cy * cz, cx * fy * cz, fx * fy * cz,
                         cx * cy * fz, fx * cy * fz, cx * fy * fz, fx * fy * fz};
    const unsigned n0 = g.node(ci, cj, ck);
    const unsigned sy = unsigned(g.nx + 1), sz = unsigned(g.nx + 1) * unsigned(g.ny + 1);
    const unsigned at[8] = {n0, n0 + 1, n0 + sy, n0 + sy + 1, n0 + sz, n0 + sz + 1, n0 + sz + sy, n0 + sz + sy + 1};
    const double bx = gather8(bnode, at, w);
    const double by = gather8(bnode + nodes, at, w);
    const double bz = gather8(bnode + 2 * size_t(nodes), at, w);
    const double ex = gather8(etheta, at, w);
    const double ey = gather8(etheta + nodes, at, w);
    const double ez = gather8(etheta + 2 * size_t(nodes), at, w);
    // ParticleResponse(bx, by, bz, beta_over_c)
    const double x = p.beta_over_c * bx, y = p.beta_over_c * by, z = p.beta_over_c * bz;
    const double n2 = x * x + y * y + z * z;
    const double inverse = 1.0 / (1.0 + n2);
    const double a0 = (1.0 + x * x) * inverse, a4 = (1.0 + y * y) * inverse, a8 = (1.0 + z * z) * inverse;
    const double xy = x * y * inverse, xz = x * z * inverse, yz = y * z * inverse;
    const double sx = x * inverse, syy = y * inverse, szz = z * inverse;
    const double a1 = xy + szz, a2 = xz - syy, a3 = xy - szz, a5 = yz + sx, a6 = xz + syy, a7 = yz - sx;
    const double u = s.u[index], v = s.v[index], ww = s.w[index];
    const double hx = a0 * u + a1 * v + a2 * ww;
    const double hy = a3 * u + a4 * v + a5 * ww;
    const double hz = a6 * u + a7 * v + a8 * ww;
    const double aex = a0 * ex + a1 * ey + a2 * ez;
    const double aey = a3 * ex + a4 * ey + a5 * ez;
    const double aez = a6 * ex + a7 * ey + a8 * ez;
    const double px = hx + p.beta * aex, py = hy + p.beta * aey, pz = hz + p.beta * aez;
    int64_t pc[3] = {ci, cj, ck};
    double frac[3] = {fx, fy, fz};
    double vel[3] = {2.0 * px - u, 2.0 * py - v, 2.0 * pz - ww};
    bool bad = false;
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
      if (bad) break;
      if (!finite_value(vel[axis])) { flags |= kInvalidNonfinite; bad = true; break; }
      const double displacement = p.dt * vel[axis] * p.inverse_spacing[axis];
      const double moved = frac[axis] + displacement;
      const double shift = floor(moved);
      if (!(shift >= -2147483648.0 && shift <= 2147483647.0)) { flags |= kInvalidFlight; bad = true; break; }
      pc[axis] += int64_t(shift);
      frac[axis] = moved - shift;
      if (frac[axis] == 1.0) {
        ++pc[axis];
        frac[axis] = 0.0;
      }
    }
    if (bad) continue;
    const int extents[3] = {g.nx, g.ny, g.nz};
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
      const int64_t extent = extents[axis];
      int64_t global = pc[axis];
      if (g.particle_periodic[axis]) {
        if (global < 0 || global >= extent) global = positive_modulo(global, extent);
        pc[axis] = global;
      } else if (global < 0 || global >= extent) {
        global = positive_modulo(global, 2 * extent);
        if (global >= extent) {
          global = 2 * extent - global - 1;
          frac[axis] = 1.0 - frac[axis];
          vel[axis] = -vel[axis];
          if (frac[axis] == 1.0) {
            if (global + 1 < extent) {
              ++global;
              frac[axis] = 0.0;
            } else {
              frac[axis] = 1.0 - 1.1102230246251565e-16;  // 1 - epsilon/2
            }
          }
        }
        pc[axis] = global;
      }
    }
    s.cell[index] = g.encode(int(pc[0]), int(pc[1]), int(pc[2]));
    s.rx[index] = frac[0];
    s.ry[index] = frac[1];
    s.rz[index] = frac[2];
    s.u[index] = vel[0];
    s.v[index] = vel[1];
    s.w[index] = vel[2];
  }
  if (flags) atomicOr(validation, flags);
}

}  // namespace

int launch_move(Context& ctx, int s) {
  SpeciesState& sp = ctx.species[size_t(s)];
  if (sp.count == 0) return 0;
  MoverParams p;
  // prepare_particle_mover_parameters (particle_mover_parameters.hpp:23-53):
  // beta = (q/m) dt / 2, beta / c and 1 / spacing formed in double.
  const double beta = sp.q_over_m * ctx.num.dt / 2;
  p.beta = beta;
  p.beta_over_c = beta / ctx.num.c;
  p.dt = ctx.num.dt;
  p.inverse_spacing[0] = 1.0 / ctx.grid.dx;
  p.inverse_spacing[1] = 1.0 / ctx.grid.dy;
  p.inverse_spacing[2] = 1.0 / ctx.grid.dz;
  const unsigned blocks = std::min<uint64_t>((uint64_t(sp.count) + kThreads - 1) / kThreads,
                                             uint64_t(ctx.cus > 0 ? ctx.cus : 256) * 32);
  phase_begin(ctx, Phase::move);
  k_move<<<blocks, kThreads, 0, ctx.stream>>>(ctx.grid, p, sp.view(), sp.count, ctx.etheta, ctx.bn, ctx.validation);
  const int rc = check_hip(ctx, hipGetLastError(), "mover");
  phase_end(ctx, Phase::move);
  sp.sorted = false;
  return rc;
}

}  // namespace tpic_mi
