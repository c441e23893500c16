// Scalar energies of the committed state and the seeded particle loader.
//
// Energy (ScalarDiagnostics, include/tpic/diagnostics/scalar_diagnostics_device.hpp:50-112):
//   W_E = V/(8 pi) * sum over uniquely owned nodes of |E|^2 / f_g (periodic
//   upper planes excluded; f_g = 2^(walls through the node)), W_B = V/(8 pi) * sum over
//   cells of |B|^2, K_s = sum 0.5 (q/qom) |v|^2. Deterministic block partials.
//
// Loader (the dense uniform path of kernels::initial_particle_load_kernel,
// include/tpic/initialization/particles/particle_loading_device.hpp:69-159):
// load index -> cell = index / ppc (x fastest), subcell = index % ppc;
// fraction = (sub + U) / ppc with the exact-1 carry of
// canonicalize_particle_fraction (particle_cell.hpp:83-93); velocity =
// drift + thermal * N; charge = (sign(q/m) V/ppc) (rho/4pi); id = first + index.
// The random source is Philox-4x32-10 keyed by the species seed and
// countered by the load index (the reference uses hipRAND XORWOW, which this
// loader does not reproduce).
#include "../../include/tpic_mi.h"
#include "internal.hpp"

#include <cmath>
#include <mutex>

namespace tpic_mi {
namespace {

constexpr unsigned kThreads = 256;

__device__ inline double block_reduce(double v, double* lds) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double sum = 0.0;
  if (threadIdx.x == 0)
    for (unsigned w = 0; w < kThreads / 64; ++w) sum += lds[w];
  __syncthreads();
  return sum;
}

__global__ __launch_bounds__(kThreads) void k_field_energy(Grid g, const double* __restrict__ e,
                                                           const double* __restrict__ b,
                                                           double* __restrict__ partials) {
  __shared__ double lds[kThreads / 64];
  const unsigned nodes = g.nodes(), cells = g.cells();
  double we = 0.0, wb = 0.0;
  if (!g.walls) {
    // Uniquely owned nodes of a periodic box are the cell-anchored ones.
    for (unsigned c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += gridDim.x * blockDim.x) {
      const int i = int(c % unsigned(g.nx)), j = int((c / unsigned(g.nx)) % unsigned(g.ny)),
                k = int(c / (unsigned(g.nx) * unsigned(g.ny)));
      const unsigned n = g.node(i, j, k);
      const double x = e[n], y = e[nodes + n], z = e[2 * size_t(nodes) + n];
      we += x * x + y * y + z * z;
      const double bx = b[c], by = b[cells + c], bz = b[2 * size_t(cells) + c];
      wb += bx * bx + by * by + bz * bz;
    }
  } else {
    // A wall axis owns both wall planes, weighted 1/f_g; a periodic axis
    // drops its upper (duplicate) plane.
    const unsigned ox = unsigned(g.nx + g.field_wall[0]), oy = unsigned(g.ny + g.field_wall[1]),
                   oz = unsigned(g.nz + g.field_wall[2]);
    const unsigned owned = ox * oy * oz;
    for (unsigned o = blockIdx.x * blockDim.x + threadIdx.x; o < owned; o += gridDim.x * blockDim.x) {
      const int i = int(o % ox), j = int((o / ox) % oy), k = int(o / (ox * oy));
      const unsigned n = g.node(i, j, k);
      const double x = e[n], y = e[nodes + n], z = e[2 * size_t(nodes) + n];
      we += (1.0 / double(g.dual_factor(i, j, k))) * (x * x + y * y + z * z);
    }
    for (unsigned c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += gridDim.x * blockDim.x) {
      const double bx = b[c], by = b[cells + c], bz = b[2 * size_t(cells) + c];
      wb += bx * bx + by * by + bz * bz;
    }
  }
  const double t0 = block_reduce(we, lds);
  const double t1 = block_reduce(wb, lds);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = t0;
    partials[gridDim.x + blockIdx.x] = t1;
  }
}

// The five scalars of a species (ScalarDiagnostics species row): kinetic
// energy sum 0.5 (q/qom)|v|^2, charge sum q, momentum sum (q/qom) v.
__global__ __launch_bounds__(kThreads) void k_species_scalars(ParticleView p, uint32_t n, double inverse_qom,
                                                              double* __restrict__ partials) {
  __shared__ double lds[kThreads / 64];
  double k = 0.0, q = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double charge = p.q[i];
    const double mass = charge * inverse_qom;
    const double u = p.u[i], v = p.v[i], w = p.w[i];
    k += 0.5 * mass * (u * u + v * v + w * w);
    q += charge;
    mx += mass * u;
    my += mass * v;
    mz += mass * w;
  }
  const double sums[5] = {k, q, mx, my, mz};
#pragma unroll
  for (int r = 0; r < 5; ++r) {
    const double t = block_reduce(sums[r], lds);
    if (threadIdx.x == 0) partials[size_t(r) * gridDim.x + blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kThreads) void k_kinetic(ParticleView p, uint32_t n, double inverse_qom,
                                                      double* __restrict__ partials) {
  __shared__ double lds[kThreads / 64];
  double k = 0.0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double mass = p.q[i] * inverse_qom;
    const double u = p.u[i], v = p.v[i], w = p.w[i];
    k += 0.5 * mass * (u * u + v * v + w * w);
  }
  const double t = block_reduce(k, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ __launch_bounds__(kThreads) void k_sum_rows(const double* __restrict__ partials,
                                                       double* __restrict__ out, unsigned count) {
  __shared__ double lds[kThreads / 64];
  double s = 0.0;
  for (unsigned i = threadIdx.x; i < count; i += blockDim.x) s += partials[size_t(blockIdx.x) * count + i];
  const double t = block_reduce(s, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// Row sums scaled by one factor: out[row] = scale * sum of partials[row][0, count).
__global__ __launch_bounds__(kThreads) void k_sum_rows_scaled(const double* __restrict__ partials,
                                                              double* __restrict__ out, unsigned count,
                                                              double scale) {
  __shared__ double lds[kThreads / 64];
  double s = 0.0;
  for (unsigned i = threadIdx.x; i < count; i += blockDim.x) s += partials[size_t(blockIdx.x) * count + i];
  const double t = block_reduce(s, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = scale * t;
}

// ---- Philox-4x32-10 ------------------------------------------------------
struct Philox {
  uint32_t c[4];
  uint32_t k[2];
};

__device__ inline void philox_round(Philox& s) {
  const uint64_t p0 = uint64_t(0xD2511F53u) * s.c[0];
  const uint64_t p1 = uint64_t(0xCD9E8D57u) * s.c[2];
  const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
  const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
  const uint32_t c0 = hi1 ^ s.c[1] ^ s.k[0];
  const uint32_t c2 = hi0 ^ s.c[3] ^ s.k[1];
  s.c[0] = c0;
  s.c[1] = lo1;
  s.c[2] = c2;
  s.c[3] = lo0;
  s.k[0] += 0x9E3779B9u;
  s.k[1] += 0xBB67AE85u;
}

// Four 32-bit words of block `draw` of stream (seed, index).
__device__ inline void philox4(uint64_t seed, uint64_t index, uint32_t draw, uint32_t out[4]) {
  Philox s{{uint32_t(index), uint32_t(index >> 32), draw, 0u}, {uint32_t(seed), uint32_t(seed >> 32)}};
#pragma unroll
  for (int r = 0; r < 10; ++r) philox_round(s);
  out[0] = s.c[0];
  out[1] = s.c[1];
  out[2] = s.c[2];
  out[3] = s.c[3];
}

// 53-bit uniform strictly inside (0, 1).
__device__ inline double uniform53(uint32_t hi, uint32_t lo) {
  const uint64_t bits = (uint64_t(hi >> 5) << 26) | uint64_t(lo >> 6);
  return (double(bits) + 0.5) * 1.1102230246251565e-16;
}

struct LoadPlan {
  int ppc[3];
  double inverse_ppc[3];
  double charge;
  double drift[3], thermal[3];
  uint64_t seed, first_id;
};

__global__ __launch_bounds__(kThreads) void k_load(Grid g, LoadPlan plan, ParticleView p, uint32_t n) {
  const uint32_t per_cell = uint32_t(plan.ppc[0] * plan.ppc[1] * plan.ppc[2]);
  for (uint32_t index = blockIdx.x * blockDim.x + threadIdx.x; index < n; index += gridDim.x * blockDim.x) {
    const uint32_t cell = index / per_cell, sub = index % per_cell;
    int ijk[3] = {int(cell % unsigned(g.nx)), int((cell / unsigned(g.nx)) % unsigned(g.ny)),
                  int(cell / (unsigned(g.nx) * unsigned(g.ny)))};
    const int sxyz[3] = {int(sub % unsigned(plan.ppc[0])), int((sub / unsigned(plan.ppc[0])) % unsigned(plan.ppc[1])),
                         int(sub / unsigned(plan.ppc[0] * plan.ppc[1]))};
    uint32_t r0[4], r1[4], r2[4], r3[4];
    philox4(plan.seed, index, 0u, r0);
    philox4(plan.seed, index, 1u, r1);
    philox4(plan.seed, index, 2u, r2);
    philox4(plan.seed, index, 3u, r3);
    const double offset[3] = {uniform53(r0[0], r0[1]), uniform53(r0[2], r0[3]), uniform53(r1[0], r1[1])};
    double frac[3];
    const int extents[3] = {g.nx, g.ny, g.nz};
    for (int a = 0; a < 3; ++a) {
      frac[a] = (double(sxyz[a]) + offset[a]) * plan.inverse_ppc[a];
      if (frac[a] == 1.0) {  // canonical carry, folded back into the box
        frac[a] = 0.0;
        ijk[a] = ijk[a] + 1 == extents[a] ? 0 : ijk[a] + 1;
      }
    }
    // Box-Muller: two pairs of uniforms give four normals, three are used.
    const double u1 = uniform53(r1[2], r1[3]), u2 = uniform53(r2[0], r2[1]);
    const double u3 = uniform53(r2[2], r2[3]), u4 = uniform53(r3[0], r3[1]);
    const double rad1 = sqrt(-2.0 * log(u1)), rad2 = sqrt(-2.0 * log(u3));
    const double two_pi = 6.283185307179586476925286766559;
    const double nrm[3] = {rad1 * cos(two_pi * u2), rad1 * sin(two_pi * u2), rad2 * cos(two_pi * u4)};
    p.rx[index] = frac[0];
    p.ry[index] = frac[1];
    p.rz[index] = frac[2];
    p.u[index] = plan.drift[0] + plan.thermal[0] * nrm[0];
    p.v[index] = plan.drift[1] + plan.thermal[1] * nrm[1];
    p.w[index] = plan.drift[2] + plan.thermal[2] * nrm[2];
    p.q[index] = plan.charge;
    p.cell[index] = g.encode(ijk[0], ijk[1], ijk[2]);
    p.id[index] = plan.first_id + index;
  }
}

unsigned reduction_blocks(const Context& ctx) {
  return unsigned(std::max(1, std::min(1024, (ctx.cus > 0 ? ctx.cus : 256) * 4)));
}

// ---- output moments --------------------------------------------------------
//
// deposit_output_moments_kernel (include/tpic/diagnostics/
// output_moment_deposition_device.hpp:35-121): per particle, charge = q / V,
// J = charge v and the raw mass-weighted second moments (charge / qom) v_a v_b,
// each times the CIC weight of a corner node. The reference scatters them with
// atomics; here the species is sorted, so a block of 128 threads takes one
// cell: its particles pass through LDS in chunks of 128 (thread t stages
// particle t's 8 weights and 10 amounts), then thread 8 c + a sums component c
// at corner a over the chunk in particle order. k_assemble_moments pulls the 8
// cells around each node, as the current assembly does.
constexpr int kMomentComponents = 10;
constexpr int kMomentRecord = 8 * kMomentComponents;  // doubles per cell: c * 8 + a
constexpr int kMomentThreads = 128;

__global__ __launch_bounds__(kMomentThreads) void k_cell_moments(Grid g, ParticleView p,
                                                                 const uint32_t* __restrict__ counts,
                                                                 const uint32_t* __restrict__ offsets,
                                                                 double inverse_volume, double inverse_qom,
                                                                 unsigned selected, double* __restrict__ cm,
                                                                 unsigned* __restrict__ validation) {
  __shared__ double s_weight[8][kMomentThreads];
  __shared__ double s_amount[kMomentComponents][kMomentThreads];
  const unsigned t = threadIdx.x;
  const int comp = int(t >> 3), corner = int(t & 7);
  const bool uses_u = selected & 0x72u, uses_v = selected & 0x1A4u, uses_w = selected & 0x348u;
  const unsigned cells = g.cells();
  for (unsigned cell = blockIdx.x; cell < cells; cell += gridDim.x) {
    const int ci = int(cell % unsigned(g.nx)), cj = int((cell / unsigned(g.nx)) % unsigned(g.ny)),
              ck = int(cell / (unsigned(g.nx) * unsigned(g.ny)));
    const int32_t key = g.encode(ci, cj, ck);
    const uint32_t begin = offsets[key], n = counts[key];
    double acc = 0.0;
    for (uint32_t base = 0; base < n; base += kMomentThreads) {
      const uint32_t m = min(uint32_t(kMomentThreads), n - base);
      __syncthreads();  // the previous chunk's sums are done with LDS
      if (t < m) {
        const uint32_t i = begin + base + t;
        unsigned bad = 0;
        if (p.cell[i] != key) bad |= TPIC_MI_INVALID_KEY;
        const double x = p.rx[i], y = p.ry[i], z = p.rz[i];
        if (!((x >= 0.0) & (x < 1.0) & (y >= 0.0) & (y < 1.0) & (z >= 0.0) & (z < 1.0)))
          bad |= TPIC_MI_INVALID_FRACTION;
        const double u = uses_u ? p.u[i] : 0.0, v = uses_v ? p.v[i] : 0.0, w = uses_w ? p.w[i] : 0.0;
        const double charge = p.q[i] * inverse_volume;
        const double mass = charge * inverse_qom;
        double amount[kMomentComponents] = {charge,       charge * u,   charge * v,   charge * w,   mass * u * u,
                                            mass * u * v, mass * u * w, mass * v * v, mass * v * w, mass * w * w};
        bool finite = true;
#pragma unroll
        for (int c = 0; c < kMomentComponents; ++c)
          if ((selected >> c) & 1u) finite = finite && finite_value(amount[c]);
        if (!bad && !finite) bad |= TPIC_MI_INVALID_NONFINITE;
        if (bad) atomicOr(validation, bad);
        const double cx = 1.0 - x, cy = 1.0 - y, cz = 1.0 - z;
#pragma unroll
        for (int a = 0; a < 8; ++a)
          s_weight[a][t] = ((a & 1) ? x : cx) * ((a & 2) ? y : cy) * ((a & 4) ? z : cz);
#pragma unroll
        for (int c = 0; c < kMomentComponents; ++c)
          s_amount[c][t] = (bad || !((selected >> c) & 1u)) ? 0.0 : amount[c];
      }
      __syncthreads();
      if (comp < kMomentComponents)
        for (uint32_t j = 0; j < m; ++j) acc += s_amount[comp][j] * s_weight[corner][j];
    }
    if (comp < kMomentComponents) cm[size_t(cell) * kMomentRecord + t] = acc;
  }
}

__global__ __launch_bounds__(kThreads) void k_assemble_moments(Grid g, const double* __restrict__ cm,
                                                               double* __restrict__ out) {
  const unsigned nodes = g.nodes();
  const unsigned sx = unsigned(g.nx + 1), sy = unsigned(g.ny + 1);
  for (unsigned n = blockIdx.x * blockDim.x + threadIdx.x; n < nodes; n += gridDim.x * blockDim.x) {
    const int i = int(n % sx), j = int((n / sx) % sy), k = int(n / (sx * sy));
    double sum[kMomentComponents] = {};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      int ci = i - (a & 1), cj = j - ((a >> 1) & 1), ck = k - ((a >> 2) & 1);
      // No cell beyond a wall; periodic axes wrap.
      if (g.field_wall[0] && (ci < 0 || ci >= g.nx)) continue;
      if (g.field_wall[1] && (cj < 0 || cj >= g.ny)) continue;
      if (g.field_wall[2] && (ck < 0 || ck >= g.nz)) continue;
      ci = wrap_cell(ci, g.nx);
      cj = wrap_cell(cj, g.ny);
      ck = wrap_cell(ck, g.nz);
      const double* rec = cm + size_t(g.cell(ci, cj, ck)) * kMomentRecord + a;
#pragma unroll
      for (int c = 0; c < kMomentComponents; ++c) sum[c] += rec[c * 8];
    }
    // Dual-volume scale (enqueue_scale_moment_dual_volume, moment_field_device.hpp:84-96).
    const double f = double(g.dual_factor(i, j, k));
#pragma unroll
    for (int c = 0; c < kMomentComponents; ++c) out[size_t(c) * nodes + n] = sum[c] * f;
  }
}

}  // namespace

int launch_moments(Context& ctx, int species, unsigned selected, double* out) {
  if (species < 0 || size_t(species) >= ctx.species.size() || !out || (selected & ~0x3FFu))
    return TPIC_MI_INVALID_ARGUMENT;
  const SpeciesState& sp = ctx.species[size_t(species)];
  if (!std::isfinite(sp.q_over_m) || sp.q_over_m == 0.0) return TPIC_MI_INVALID_ARGUMENT;
  if (!sp.sorted) return TPIC_MI_INVALID_STATE;
  const Grid& g = ctx.grid;
  const size_t record_bytes = size_t(g.cells()) * kMomentRecord * sizeof(double);
  int rc;
  if (!ctx.moment_cells &&
      (rc = check_hip(ctx, hipMalloc((void**)&ctx.moment_cells, record_bytes), "moment cell records")))
    return rc;
  hipStream_t st = ctx.stream;
  const unsigned cell_blocks = std::max(1u, std::min(g.cells(), 8192u));
  k_cell_moments<<<cell_blocks, kMomentThreads, 0, st>>>(g, sp.view(), sp.counts, sp.offsets, ctx.inverse_volume,
                                                         1.0 / sp.q_over_m, selected, ctx.moment_cells,
                                                         ctx.validation);
  if ((rc = check_hip(ctx, hipGetLastError(), "output moments"))) return rc;
  const unsigned node_blocks = std::max(1u, std::min((g.nodes() + kThreads - 1) / kThreads, 8192u));
  k_assemble_moments<<<node_blocks, kThreads, 0, st>>>(g, ctx.moment_cells, out);
  if ((rc = check_hip(ctx, hipGetLastError(), "output moment assembly"))) return rc;
  return check_hip(ctx, hipStreamSynchronize(st), "output moments");
}

int launch_scalars(Context& ctx, double* out) {
  const Grid& g = ctx.grid;
  const unsigned blocks = reduction_blocks(ctx);
  hipStream_t st = ctx.stream;
  const size_t ns = ctx.species.size(), rows = 2 + 5 * ns;
  k_field_energy<<<blocks, kThreads, 0, st>>>(g, ctx.e, ctx.b, ctx.reduce);
  int rc = check_hip(ctx, hipGetLastError(), "field energy");
  if (rc) return rc;
  for (size_t s = 0; s < ns; ++s) {
    const SpeciesState& sp = ctx.species[s];
    k_species_scalars<<<blocks, kThreads, 0, st>>>(sp.view(), sp.count, 1.0 / sp.q_over_m,
                                                   ctx.reduce + (2 + 5 * s) * blocks);
    if ((rc = check_hip(ctx, hipGetLastError(), "species scalars"))) return rc;
  }
  k_sum_rows<<<unsigned(rows), kThreads, 0, st>>>(ctx.reduce, ctx.reduce_out, blocks);
  if ((rc = check_hip(ctx, hipGetLastError(), "scalars reduce"))) return rc;
  if ((rc = check_hip(ctx, hipMemcpyAsync(ctx.host_scalars, ctx.reduce_out, rows * sizeof(double),
                                          hipMemcpyDeviceToHost, st),
                      "scalars readback")))
    return rc;
  if ((rc = check_hip(ctx, hipStreamSynchronize(st), "scalars readback"))) return rc;
  // scalar_diagnostics_device.hpp:138: V/(8 pi), V formed with a wider product.
  const long double volume = (long double)g.dx * g.dy * g.dz;
  const double normalization = double(volume) / (8 * 3.14159265358979323846);
  out[0] = normalization * ctx.host_scalars[0];
  out[1] = normalization * ctx.host_scalars[1];
  for (size_t r = 2; r < rows; ++r) out[r] = ctx.host_scalars[r];
  return 0;
}

int launch_energy(Context& ctx, double* out) {
  const Grid& g = ctx.grid;
  const unsigned blocks = reduction_blocks(ctx);
  hipStream_t st = ctx.stream;
  const size_t ns = ctx.species.size();
  k_field_energy<<<blocks, kThreads, 0, st>>>(g, ctx.e, ctx.b, ctx.reduce);
  int rc = check_hip(ctx, hipGetLastError(), "field energy");
  if (rc) return rc;
  for (size_t s = 0; s < ns; ++s) {
    const SpeciesState& sp = ctx.species[s];
    k_kinetic<<<blocks, kThreads, 0, st>>>(sp.view(), sp.count, 1.0 / sp.q_over_m,
                                           ctx.reduce + (2 + s) * blocks);
    if ((rc = check_hip(ctx, hipGetLastError(), "kinetic energy"))) return rc;
  }
  k_sum_rows<<<unsigned(2 + ns), kThreads, 0, st>>>(ctx.reduce, ctx.reduce_out, blocks);
  if ((rc = check_hip(ctx, hipGetLastError(), "energy reduce"))) return rc;
  if ((rc = check_hip(ctx, hipMemcpyAsync(ctx.host_scalars, ctx.reduce_out, (2 + ns) * sizeof(double),
                                          hipMemcpyDeviceToHost, st),
                      "energy readback")))
    return rc;
  if ((rc = check_hip(ctx, hipStreamSynchronize(st), "energy readback"))) return rc;
  // scalar_diagnostics_device.hpp:138: V/(8 pi), V formed with a wider product.
  const long double volume = (long double)g.dx * g.dy * g.dz;
  const double normalization = double(volume) / (8 * 3.14159265358979323846);
  out[0] = normalization * ctx.host_scalars[0];
  out[1] = normalization * ctx.host_scalars[1];
  for (size_t s = 0; s < ns; ++s) out[2 + s] = ctx.host_scalars[2 + s];
  return 0;
}

int launch_load(Context& ctx, int s, const int ppc[3], double rho, const double drift[3], const double thermal[3],
                uint64_t seed, uint64_t first_id) {
  SpeciesState& sp = ctx.species[size_t(s)];
  const Grid& g = ctx.grid;
  LoadPlan plan{};
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
  // q = (sign(q/m) V/ppc) * (rho / 4pi).
  const double volume = (g.dx * g.dy) * g.dz;
  const double q_factor = (sp.q_over_m > 0.0 ? 1.0 : -1.0) * volume / double(per_cell);
  const double density = rho / 12.566370614359172953850573533118;
  plan.charge = q_factor * density * 1.0;
  plan.seed = seed;
  plan.first_id = first_id;
  const unsigned blocks = unsigned(std::min<uint64_t>((uint64_t(sp.count) + kThreads - 1) / kThreads,
                                                      uint64_t(ctx.cus > 0 ? ctx.cus : 256) * 32));
  k_load<<<blocks ? blocks : 1, kThreads, 0, ctx.stream>>>(g, plan, sp.view(), sp.count);
  sp.sorted = false;
  return check_hip(ctx, hipGetLastError(), "load");
}

}  // namespace tpic_mi

// ---- field energy without allocation --------------------------------------
//
// tpic_mi_field_energy_static: W_E and W_B of caller-owned device fields
// without a context and without any allocation. Its one scratch buffer, the
// per-block partial sums of k_field_energy (one block of kThreads per
// kThreads cells or owned nodes, row 0 W_E, row 1 W_B), is this file-scope
// device array. Each GPU holds its own copy until the process ends; host code
// reaches it only through the address hipGetSymbolAddress returns on that
// device (the array's name is not a device address on the host).
namespace tpic_mi {
namespace {

constexpr unsigned kStaticEnergyBlocks = 65536;  // grids up to 65536 x 256 = 16,777,216 cells
static __device__ __align__(256) double scratch[2 * kStaticEnergyBlocks];  // 1 MiB

constexpr int kMaxStaticDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch[kMaxStaticDevices] = {};

// The scratch array's address on the current device, looked up once per device.
int static_scratch(double** out) {
  int device = -1;
  hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess) {
    set_context_free_error(std::string("tpic_mi_field_energy_static: no HIP device: ") + hipGetErrorString(e));
    return TPIC_MI_BACKEND;
  }
  if (device < 0 || device >= kMaxStaticDevices) {
    set_context_free_error("tpic_mi_field_energy_static: device index beyond 63");
    return TPIC_MI_BACKEND;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[device]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_context_free_error(std::string("tpic_mi_field_energy_static: scratch address lookup failed: ") +
                             hipGetErrorString(e));
      return TPIC_MI_BACKEND;
    }
    g_scratch[device] = static_cast<double*>(ptr);
  }
  *out = g_scratch[device];
  return TPIC_MI_OK;
}

}  // namespace
}  // namespace tpic_mi

extern "C" int tpic_mi_field_energy_static(const tpic_mi_config* config, const double* device_e,
                                           const double* device_b, double* device_out, void* stream) {
  using namespace tpic_mi;
  auto reject = [](const std::string& why) {
    set_context_free_error("tpic_mi_field_energy_static: " + why);
    return TPIC_MI_INVALID_ARGUMENT;
  };
  if (!config) return reject("null config");
  if (!device_e || !device_b || !device_out) return reject("null field or output pointer");
  const tpic_mi_config& k = *config;
  Grid g{};
  for (int a = 0; a < 3; ++a) {
    if (k.cells[a] <= 0) return reject("cells must be positive");
    if (!(std::isfinite(k.lengths[a]) && k.lengths[a] > 0)) return reject("lengths must be finite and positive");
    const int lo = k.field_faces[a][0], hi = k.field_faces[a][1];
    for (int f : {lo, hi})
      if (f != TPIC_MI_FACE_PERIODIC && f != TPIC_MI_FACE_PEC && f != TPIC_MI_FACE_PMC)
        return reject("field faces must be TPIC_MI_FACE_PERIODIC, TPIC_MI_FACE_PEC or TPIC_MI_FACE_PMC");
    if ((lo == TPIC_MI_FACE_PERIODIC) != (hi == TPIC_MI_FACE_PERIODIC))
      return reject("an axis is periodic on both faces or has a wall on both");
    if (lo != TPIC_MI_FACE_PERIODIC && k.cells[a] < 2) return reject("a wall axis needs at least 2 cells");
    g.field_wall[a] = lo != TPIC_MI_FACE_PERIODIC ? 1 : 0;
    g.walls |= g.field_wall[a];
  }
  // The partials: one block per kThreads of the larger of the cell count and
  // the owned-node count (walled axes own both wall planes).
  uint64_t cells = 1, owned = 1;
  for (int a = 0; a < 3; ++a) {
    cells *= uint64_t(k.cells[a]);
    owned *= uint64_t(k.cells[a] + g.field_wall[a]);
  }
  const uint64_t units = std::max(cells, owned);
  const uint64_t blocks = (units + kThreads - 1) / kThreads;
  if (blocks > kStaticEnergyBlocks)
    return reject("grid of " + std::to_string(cells) + " cells exceeds the scratch limit of " +
                  std::to_string(uint64_t(kStaticEnergyBlocks) * kThreads) + " cells and owned nodes");
  g.nx = k.cells[0];
  g.ny = k.cells[1];
  g.nz = k.cells[2];
  g.dx = k.lengths[0] / k.cells[0];
  g.dy = k.lengths[1] / k.cells[1];
  g.dz = k.lengths[2] / k.cells[2];
  double* partials = nullptr;
  int rc = static_scratch(&partials);
  if (rc) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  // k_field_energy writes every one of its 2 x blocks partials, so nothing
  // left by an earlier call is read.
  k_field_energy<<<unsigned(blocks), kThreads, 0, st>>>(g, device_e, device_b, partials);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    // scalar_diagnostics_device.hpp:138: V/(8 pi), V formed with a wider product.
    const long double volume = (long double)g.dx * g.dy * g.dz;
    const double normalization = double(volume) / (8 * 3.14159265358979323846);
    k_sum_rows_scaled<<<2u, kThreads, 0, st>>>(partials, device_out, unsigned(blocks), normalization);
    e = hipGetLastError();
  }
  if (e != hipSuccess) {
    set_context_free_error(std::string("tpic_mi_field_energy_static: launch failed: ") + hipGetErrorString(e));
    return TPIC_MI_BACKEND;
  }
  return TPIC_MI_OK;
}
