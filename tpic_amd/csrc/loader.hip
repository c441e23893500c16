// The reference's Uniform-case initial particle load, draw for draw.
//
// kernels::initial_particle_load_kernel with UniformProfile
// (include/tpic/initialization/particles/particle_loading_device.hpp:69-159),
// its plan (src/initialization/prepared_initialization.cpp:37-85) and the
// startup fold of enqueue_local_initial_particle_boundaries:
//   load index i -> cell i / ppc, subcell i % ppc (both x fastest);
//   generator: the backend's default XORWOW (hipRAND, i.e. rocRAND's
//   xorwow_engine) with seed = 12345 + 100 (rank * max_species + species),
//   subsequence = i, offset 0 (particle_loading_device.hpp:37-55,
//   prepared_initialization.cpp:46-47);
//   draws, in order: three uniform doubles for the subcell offsets (none under
//   QuietStart, which uses 0.5), one normal pair for u and v, one normal for w;
//   fraction = (sub + offset) / ppc with the exact-1 carry of
//   canonicalize_particle_fraction (include/tpic/data/particles/
//   particle_cell.hpp:83-93); velocity = drift + thermal * N;
//   charge = (sign(q/m) V / ppc) (sampled density) times the profile's density
//   multiplier: the density rho / 4 pi and multiplier 1 for the Uniform
//   profile, 1 + A cos(k x + phase) for the TwoStream density seed
//   (include/tpic/initialization/cases/two_stream.hpp:13-28); the Harris
//   profile (harris_sheet.hpp:56-120) averages the sheet shape over the two
//   y nodes of the cell into the density, scales the z drift of a double
//   sheet and sets the z drift of currentFromAmpere from the field's curl;
//   x, y the position after the carry; key = halo-2 encoding of the
//   (carried) cell;
//   id = first_id + i.
// A carry out of the last cell of an axis is folded as the startup
// boundaries fold any point (classify_particle_point,
// include/tpic/communication/particles/particle_boundaries.hpp:111-165):
// periodic wraps to cell 0; a mirror reflects the point, which then sits on
// the wall. Contraction is off in this translation unit so positions are
// bit-identical to the oracle.
#include "../../include/tpic_mi.h"
#include "internal.hpp"

#include <hiprand/hiprand_kernel.h>

#include <cmath>

#pragma clang fp contract(off)

namespace tpic_mi {
namespace {

constexpr unsigned kThreads = 256;

struct UniformLoadPlan {
  uint64_t seed, first_id;
  int ppc[3];
  double inverse_ppc[3];
  double q_factor, density;  // charge = (q_factor * sampled density) * multiplier
  double drift[3], thermal[3];
  int quiet;
  // TwoStream density seed: multiplier 1 + amplitude cos(k x + phase).
  int density_seed;
  double seed_k, seed_phase, seed_amplitude;
  // Harris sheet profile (HarrisProfile::evaluate).
  int harris;
  HarrisLoad h;
};

__device__ inline double sech_squared(double v) {
  const double ch = cosh(v);
  return 1.0 / (ch * ch);
}

__device__ inline double harris_shape(const HarrisLoad& h, double y) {
  if (h.is_double) return sech_squared((y - 0.25 * h.ly) / h.delta) + sech_squared((y - 0.75 * h.ly) / h.delta);
  return sech_squared((y - 0.5 * h.ly) / h.delta);
}

// HarrisCurl::at (include/tpic/initialization/cases/harris_sheet.hpp:33-51):
// the z component of the curl of the sheet field with its seed.
__device__ inline double harris_curl(const HarrisLoad& h, double x, double y) {
  const double xm = x - h.center_x;
  const double ym = y - h.center_y;
  const double ch = cosh(ym / h.curl_delta);
  const double sx = xm * h.hump_inverse_wx2;
  const double sy = ym * h.hump_inverse_wy2;
  const double cos_x = cos(h.hump_kx * xm);
  const double sin_x = sin(h.hump_kx * xm);
  const double cos_y = cos(h.hump_ky * ym);
  const double sin_y = sin(h.hump_ky * ym);
  return h.sheet / (ch * ch) + h.mode * cos(h.mode_kx * xm) * cos(h.mode_ky * ym) -
         h.hump * exp(-xm * sx - ym * sy) *
             (cos_y * ((4.0 * sx * sx - 2.0 * h.hump_inverse_wx2 - h.hump_kx * h.hump_kx) * cos_x +
                       4.0 * sx * h.hump_kx * sin_x) +
              cos_x * ((4.0 * sy * sy - 2.0 * h.hump_inverse_wy2 - h.hump_ky * h.hump_ky) * cos_y +
                       4.0 * sy * h.hump_ky * sin_y));
}

// The fold of a point one cell past the upper end of an axis (the only
// out-of-range position a load can produce): periodic -> cell 0; mirror ->
// reflected onto the wall (classify_particle_point's endpoint rule).
__device__ inline void fold_carry(int extent, int periodic, int& cell, double& frac, double& vel) {
  if (cell < extent) return;
  if (periodic) {
    cell = 0;
    return;
  }
  // global = extent: mod 2 extent leaves extent >= extent, so reflect.
  int global = 2 * extent - extent - 1;
  frac = 1.0 - frac;
  vel = -vel;
  if (frac == 1.0) {
    if (global + 1 < extent) {
      ++global;
      frac = 0.0;
    } else {
      frac = 1.0 - 1.1102230246251565e-16;  // 1 - epsilon/2
    }
  }
  cell = global;
}

__global__ __launch_bounds__(kThreads) void k_load_uniform(Grid g, UniformLoadPlan plan, ParticleView p, uint32_t n,
                                                           unsigned* __restrict__ validation) {
  const uint32_t per_cell = uint32_t(plan.ppc[0] * plan.ppc[1] * plan.ppc[2]);
  unsigned flags = 0;
  for (uint32_t index = blockIdx.x * blockDim.x + threadIdx.x; index < n; index += gridDim.x * blockDim.x) {
    const uint32_t cell = index / per_cell, sub = index % per_cell;
    int ijk[3] = {int(cell % unsigned(g.nx)), int((cell / unsigned(g.nx)) % unsigned(g.ny)),
                  int(cell / unsigned(g.nx) / unsigned(g.ny))};
    const int sxyz[3] = {int(sub % unsigned(plan.ppc[0])), int((sub / unsigned(plan.ppc[0])) % unsigned(plan.ppc[1])),
                         int(sub / unsigned(plan.ppc[0]) / unsigned(plan.ppc[1]))};
    hiprandState state;
    hiprand_init(plan.seed, index, 0, &state);
    double offset[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) offset[a] = plan.quiet ? 0.5 : hiprand_uniform_double(&state);
    double frac[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      frac[a] = (double(sxyz[a]) + offset[a]) * plan.inverse_ppc[a];
      if (frac[a] == 1.0) {
        ++ijk[a];
        frac[a] = 0.0;
      }
    }
    // The profile samples the stored position and cell after the carry (the
    // one-rank origin is 0).
    double multiplier = 1.0, density = plan.density;
    double drift[3] = {plan.drift[0], plan.drift[1], plan.drift[2]};
    double thermal[3] = {plan.thermal[0], plan.thermal[1], plan.thermal[2]};
    if (plan.density_seed) {
      const double x = (double(ijk[0]) + frac[0]) * g.dx;
      multiplier = 1.0 + plan.seed_amplitude * cos(plan.seed_k * x + plan.seed_phase);
    }
    if (plan.harris) {
      const HarrisLoad& h = plan.h;
      const double x = (double(ijk[0]) + frac[0]) * g.dx, y = (double(ijk[1]) + frac[1]) * g.dy;
      double cell_shape = 1.0;
      if (h.density) {
        // The two y-node densities of the containing cell, averaged.
        const double lower_y = double(ijk[1]) * g.dy;
        const double nodes = harris_shape(h, lower_y) + harris_shape(h, lower_y + g.dy);
        density = density * 0.5 * nodes;
        cell_shape = 0.5 * nodes;
      }
      if (h.is_double) drift[2] = drift[2] * -tanh((y - 0.5 * h.ly) / h.delta);
      if (h.ampere_drift != 0.0) {
        const double d = h.ampere_drift * harris_curl(h, x, y) / cell_shape;
        if (h.varying_thermal) {
          const double square = d * d;
          const double variance = (square + thermal[2] * thermal[2]) - square;
          if (variance > 4.0 * 2.220446049250313e-16 * square) thermal[2] = sqrt(variance);
        }
        drift[2] = d;
      }
    }
    const double charge = (plan.q_factor * density) * multiplier;
    const double2 xy = hiprand_normal2_double(&state);
    const double nz = hiprand_normal_double(&state);
    double vel[3] = {drift[0] + thermal[0] * xy.x, drift[1] + thermal[1] * xy.y, drift[2] + thermal[2] * nz};
    const int extents[3] = {g.nx, g.ny, g.nz};
#pragma unroll
    for (int a = 0; a < 3; ++a) fold_carry(extents[a], g.particle_periodic[a], ijk[a], frac[a], vel[a]);
    if (!finite_value(vel[0]) || !finite_value(vel[1]) || !finite_value(vel[2]) || !finite_value(charge))
      flags |= kInvalidNonfinite;
    p.rx[index] = frac[0];
    p.ry[index] = frac[1];
    p.rz[index] = frac[2];
    p.u[index] = vel[0];
    p.v[index] = vel[1];
    p.w[index] = vel[2];
    p.q[index] = charge;
    p.cell[index] = g.encode(ijk[0], ijk[1], ijk[2]);
    p.id[index] = plan.first_id + index;
  }
  if (flags) atomicOr(validation, flags);
}

// Velocity centering (kernels::center_particle_velocity_kernel and
// VelocityCentering::host_blocking_center, include/tpic/initialization/
// particles/particle_centering_device.hpp:57-215) for a profile whose drift is
// the same everywhere (the Uniform case): first every velocity loses the
// drift, u' = (u - 0) - drift, and the four sums sum q, sum q u', sum q v',
// sum q w' are formed; then u = (u' - mean) + drift with mean = sums[1..3] /
// sums[0] (no shift when sums[0] = 0). A non-finite corrected velocity is not
// stored and raises the flag. The sums are FP64, in a fixed order: per-thread
// strided partials, then waves, then blocks in index order.
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__global__ __launch_bounds__(kThreads) void k_center_rest(ParticleView p, uint32_t n, double dx, double dy,
                                                          double dz, double* __restrict__ partials) {
  __shared__ double lds[4][kThreads / 64];
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double q = p.q[i];
    const double u = p.u[i] - dx, v = p.v[i] - dy, w = p.w[i] - dz;
    p.u[i] = u;
    p.v[i] = v;
    p.w[i] = w;
    s[0] += q;
    s[1] += q * u;
    s[2] += q * v;
    s[3] += q * w;
  }
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const double t = wave_sum(s[r]);
    if (lane == 0) lds[r][wave] = t;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0.0;
    for (unsigned k = 0; k < kThreads / 64; ++k) t += lds[threadIdx.x][k];
    partials[size_t(threadIdx.x) * gridDim.x + blockIdx.x] = t;
  }
}

// One block: out[r] = sum over the blocks' partials of row r, r < 4.
__global__ __launch_bounds__(kThreads) void k_center_sums(const double* __restrict__ partials, unsigned count,
                                                          double* __restrict__ out) {
  __shared__ double lds[4][kThreads / 64];
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    double s = 0.0;
    for (unsigned i = threadIdx.x; i < count; i += blockDim.x) s += partials[size_t(r) * count + i];
    s = wave_sum(s);
    if (lane == 0) lds[r][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0.0;
    for (unsigned k = 0; k < kThreads / 64; ++k) t += lds[threadIdx.x][k];
    out[threadIdx.x] = t;
  }
}

__global__ __launch_bounds__(kThreads) void k_center_apply(ParticleView p, uint32_t n, double mx, double my,
                                                           double mz, double dx, double dy, double dz,
                                                           unsigned* __restrict__ flag) {
  unsigned bad = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double u = (p.u[i] - mx) + dx, v = (p.v[i] - my) + dy, w = (p.w[i] - mz) + dz;
    if (!finite_value(u) || !finite_value(v) || !finite_value(w)) {
      bad = 1;
    } else {
      p.u[i] = u;
      p.v[i] = v;
      p.w[i] = w;
    }
  }
  if (bad) atomicOr(flag, 1u);
}

}  // namespace

int launch_center_velocity(Context& ctx, int s, const double drift[3]) {
  SpeciesState& sp = ctx.species[size_t(s)];
  hipStream_t st = ctx.stream;
  const unsigned blocks = unsigned(std::max(1, std::min(1024, (ctx.cus > 0 ? ctx.cus : 256) * 4)));
  k_center_rest<<<blocks, kThreads, 0, st>>>(sp.view(), sp.count, drift[0], drift[1], drift[2], ctx.reduce);
  int rc = check_hip(ctx, hipGetLastError(), "velocity centering");
  if (rc) return rc;
  k_center_sums<<<1, kThreads, 0, st>>>(ctx.reduce, blocks, ctx.reduce_out);
  if ((rc = check_hip(ctx, hipGetLastError(), "velocity centering"))) return rc;
  if ((rc = check_hip(ctx, hipMemcpyAsync(ctx.host_scalars, ctx.reduce_out, 4 * sizeof(double),
                                          hipMemcpyDeviceToHost, st),
                      "velocity centering readback")))
    return rc;
  if ((rc = check_hip(ctx, hipStreamSynchronize(st), "velocity centering readback"))) return rc;
  const double* sums = ctx.host_scalars;
  double mean[3] = {0.0, 0.0, 0.0};
  bool finite = true;
  for (int r = 0; r < 4; ++r) finite = finite && std::isfinite(sums[r]);
  if (finite && sums[0] != 0.0)
    for (int a = 0; a < 3; ++a) mean[a] = sums[a + 1] / sums[0];
  for (int a = 0; a < 3; ++a) finite = finite && std::isfinite(mean[a]);
  if (!finite) return fail(ctx, TPIC_MI_OVERFLOW, "velocity centering: nonfinite global sum or mean");
  unsigned* flag = reinterpret_cast<unsigned*>(ctx.reduce_out + 4);
  if ((rc = check_hip(ctx, hipMemsetAsync(flag, 0, sizeof(unsigned), st), "velocity centering"))) return rc;
  k_center_apply<<<blocks, kThreads, 0, st>>>(sp.view(), sp.count, mean[0], mean[1], mean[2], drift[0], drift[1],
                                              drift[2], flag);
  if ((rc = check_hip(ctx, hipGetLastError(), "velocity centering"))) return rc;
  if ((rc = check_hip(ctx, hipMemcpyAsync(ctx.host_word, flag, sizeof(unsigned), hipMemcpyDeviceToHost, st),
                      "velocity centering readback")))
    return rc;
  if ((rc = check_hip(ctx, hipStreamSynchronize(st), "velocity centering readback"))) return rc;
  if (*ctx.host_word) return fail(ctx, TPIC_MI_OVERFLOW, "velocity centering: nonfinite corrected velocity");
  return TPIC_MI_OK;
}

int launch_load_uniform(Context& ctx, int s, const int ppc[3], double rho, const double drift[3],
                        const double thermal[3], int quiet, uint64_t seed, uint64_t first_id,
                        const double* density_seed, const HarrisLoad* harris) {
  SpeciesState& sp = ctx.species[size_t(s)];
  const Grid& g = ctx.grid;
  UniformLoadPlan plan{};
  for (int a = 0; a < 3; ++a) {
    plan.ppc[a] = ppc[a];
    plan.inverse_ppc[a] = 1.0 / double(ppc[a]);
    plan.drift[a] = drift[a];
    plan.thermal[a] = thermal[a];
  }
  const uint64_t per_cell = uint64_t(ppc[0]) * ppc[1] * ppc[2];
  if (per_cell == 0 || per_cell * g.cells() != sp.count)
    return fail(ctx, -1, "load: ppc x cells must equal the species count");
  // prepared_initialization.cpp:49-54: V = (dx*dy)*dz in double,
  // q = (sign(q/m) V/ppc) * (rho / 4pi), times the profile's multiplier 1.
  const double volume = (g.dx * g.dy) * g.dz;
  const double q_factor = (sp.q_over_m > 0.0 ? 1.0 : -1.0) * volume / double(per_cell);
  const double density = rho / 12.566370614359172953850573533118;
  plan.q_factor = q_factor;
  plan.density = density;
  if (harris) {
    plan.harris = 1;
    plan.h = *harris;
  }
  if (density_seed) {
    plan.density_seed = 1;
    plan.seed_k = density_seed[0];
    plan.seed_phase = density_seed[1];
    plan.seed_amplitude = density_seed[2];
  }
  plan.seed = seed;
  plan.first_id = first_id;
  plan.quiet = quiet ? 1 : 0;
  const unsigned blocks = unsigned(std::min<uint64_t>((uint64_t(sp.count) + kThreads - 1) / kThreads,
                                                      uint64_t(ctx.cus > 0 ? ctx.cus : 256) * 64));
  k_load_uniform<<<blocks ? blocks : 1, kThreads, 0, ctx.stream>>>(g, plan, sp.view(), sp.count, ctx.validation);
  sp.sorted = false;
  return check_hip(ctx, hipGetLastError(), "uniform load");
}

}  // namespace tpic_mi
