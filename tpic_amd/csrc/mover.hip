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
// Two kernels share the push: the sorted kernel gives one wave to each
// interior cell (the bins of the last sort), stages the cell's 48 corner field
// values once in LDS and reads them back as broadcasts, so a particle costs
// only its own 60 B read and 52 B write; the unsorted kernel gathers the
// corners per particle (input that was never sorted).
//
// Contraction is off in this translation unit (-ffp-contract=off and the
// pragma below): every product and sum rounds on its own, in the reference's
// expression order, so the result is bit-identical to the CPU oracle.
// Bound: HBM, 112 B per particle.
#include "internal.hpp"

#pragma clang fp contract(off)

namespace tpic_mi {
namespace {

constexpr unsigned kThreads = 256;
constexpr unsigned kWaves = kThreads / 64;

struct MoverParams {
  double beta, beta_over_c, dt, inverse_spacing[3];
};

__device__ __forceinline__ int64_t positive_modulo(int64_t cell, int64_t extent) {
  const int64_t r = cell % extent;
  return r < 0 ? r + extent : r;
}

// The record of one particle as the mover reads it (60 B).
struct Raw {
  double rx, ry, rz, u, v, w, q;
  int32_t key;
};

__device__ __forceinline__ void load_raw(const ParticleView& s, uint32_t index, Raw& r) {
  r.rx = s.rx[index];
  r.ry = s.ry[index];
  r.rz = s.rz[index];
  r.u = s.u[index];
  r.v = s.v[index];
  r.w = s.w[index];
  r.q = s.q[index];
  r.key = s.cell[index];
}

// The pushed state of one particle: new key, fractions and velocity.
struct Pushed {
  int32_t key;
  double r[3], v[3];
};

// Record checks and the 8 CIC weights of the stored fractions (corners
// x-fastest). Returns the validation flags of a rejection, or 0.
__device__ __forceinline__ unsigned particle_weights(const Raw& r, double (&w)[8]) {
  if (!finite_value(r.q)) return kInvalidNonfinite;
  const double fx = r.rx, fy = r.ry, fz = r.rz;
  if (!(fx >= 0.0 && fx < 1.0 && fy >= 0.0 && fy < 1.0 && fz >= 0.0 && fz < 1.0)) return kInvalidFraction;
  const double cx = 1.0 - fx, cy = 1.0 - fy, cz = 1.0 - fz;
  w[0] = cx * cy * cz;
  w[1] = fx * cy * cz;
  w[2] = cx * fy * cz;
  w[3] = fx * fy * cz;
  w[4] = cx * cy * fz;
  w[5] = fx * cy * fz;
  w[6] = cx * fy * fz;
  w[7] = fx * fy * fz;
  return 0u;
}

// B and E_theta at the particle: component c (Bx, By, Bz, Ex, Ey, Ez) is the
// weighted sum of corner(c, a) over a = 0..7 in that order.
template <class Corner>
__device__ __forceinline__ void interpolate(const double (&w)[8], const Corner& corner, double (&f)[6]) {
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    double value = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) value += w[a] * corner(c, a);
    f[c] = value;
  }
}

// Corner values staged per cell: 8 consecutive values per component.
struct StagedCorners {
  const double* v;
  __device__ double operator()(int c, int a) const { return v[c * 8 + a]; }
};

// Corner values read from the node fields (a particle outside the staged cell).
struct NodeCorners {
  const double* bnode;
  const double* etheta;
  unsigned nodes;
  unsigned at[8];
  __device__ double operator()(int c, int a) const {
    return c < 3 ? bnode[size_t(c) * nodes + at[a]] : etheta[size_t(c - 3) * nodes + at[a]];
  }
};

__device__ __forceinline__ unsigned push_fields(const Grid& g, const MoverParams& p, const Raw& r, int ci, int cj,
                                                int ck, const double (&f)[6], Pushed& out);

// Push and fold one particle of the validated interior cell (ci, cj, ck).
// `corner` gives the cell's 8 corner values of Bx, By, Bz, Ex, Ey, Ez.
// Returns the validation flags of a rejection (the record must then stay
// unchanged); on acceptance `out` holds the new state.
template <class Corner>
__device__ __forceinline__ unsigned push_particle(const Grid& g, const MoverParams& p, const Raw& r, int ci, int cj,
                                                  int ck, const Corner& corner, Pushed& out) {
  double w[8];
  const unsigned bad = particle_weights(r, w);
  if (bad) return bad;
  double f[6];
  interpolate(w, corner, f);
  return push_fields(g, p, r, ci, cj, ck, f, out);
}

// The push from the interpolated fields: response, velocity update,
// displacement with carry, then the whole-domain fold.
__device__ __forceinline__ unsigned push_fields(const Grid& g, const MoverParams& p, const Raw& r, int ci, int cj,
                                                int ck, const double (&f)[6], Pushed& out) {
  const double fx = r.rx, fy = r.ry, fz = r.rz;
  const double bx = f[0], by = f[1], bz = f[2], ex = f[3], ey = f[4], ez = f[5];
  // ParticleResponse(bx, by, bz, beta_over_c)
  const double x = p.beta_over_c * bx, y = p.beta_over_c * by, z = p.beta_over_c * bz;
  const double n2 = x * x + y * y + z * z;
  const double inverse = 1.0 / (1.0 + n2);
  const double a0 = (1.0 + x * x) * inverse, a4 = (1.0 + y * y) * inverse, a8 = (1.0 + z * z) * inverse;
  const double xy = x * y * inverse, xz = x * z * inverse, yz = y * z * inverse;
  const double sx = x * inverse, syy = y * inverse, szz = z * inverse;
  const double a1 = xy + szz, a2 = xz - syy, a3 = xy - szz, a5 = yz + sx, a6 = xz + syy, a7 = yz - sx;
  const double u = r.u, v = r.v, ww = r.w;
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
#pragma unroll
  for (int axis = 0; axis < 3; ++axis) {
    if (!finite_value(vel[axis])) return kInvalidNonfinite;
    const double displacement = p.dt * vel[axis] * p.inverse_spacing[axis];
    const double moved = frac[axis] + displacement;
    const double shift = floor(moved);
    if (!(shift >= -2147483648.0 && shift <= 2147483647.0)) return kInvalidFlight;
    pc[axis] += int64_t(shift);
    frac[axis] = moved - shift;
    if (frac[axis] == 1.0) {
      ++pc[axis];
      frac[axis] = 0.0;
    }
  }
  const int extents[3] = {g.nx, g.ny, g.nz};
#pragma unroll
  for (int axis = 0; axis < 3; ++axis) {
    const int64_t extent = extents[axis];
    int64_t global = pc[axis];
    if (global >= 0 && global < extent) continue;
    if (g.particle_periodic[axis]) {
      pc[axis] = positive_modulo(global, extent);
    } else {
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
  out.key = g.encode(int(pc[0]), int(pc[1]), int(pc[2]));
  out.r[0] = frac[0];
  out.r[1] = frac[1];
  out.r[2] = frac[2];
  out.v[0] = vel[0];
  out.v[1] = vel[1];
  out.v[2] = vel[2];
  return 0u;
}

// The accepted state overwrites the record in place (a rejected record keeps
// every attribute).
__device__ __forceinline__ void store_pushed(const ParticleView& s, uint32_t index, const Pushed& out) {
  s.cell[index] = out.key;
  s.rx[index] = out.r[0];
  s.ry[index] = out.r[1];
  s.rz[index] = out.r[2];
  s.u[index] = out.v[0];
  s.v[index] = out.v[1];
  s.w[index] = out.v[2];
}

__device__ __forceinline__ void corner_nodes(const Grid& g, int ci, int cj, int ck, unsigned at[8]) {
  const unsigned n0 = g.node(ci, cj, ck);
  const unsigned sy = unsigned(g.nx + 1), sz = unsigned(g.nx + 1) * unsigned(g.ny + 1);
  at[0] = n0;
  at[1] = n0 + 1;
  at[2] = n0 + sy;
  at[3] = n0 + sy + 1;
  at[4] = n0 + sz;
  at[5] = n0 + sz + 1;
  at[6] = n0 + sz + sy;
  at[7] = n0 + sz + sy + 1;
}

// Wave-uniform metadata of one cell: bin, particle count and first index.
struct CellMeta {
  int32_t bin;
  uint32_t count, first;
};

__device__ __forceinline__ void fetch_meta(const Grid& g, const uint32_t* __restrict__ counts,
                                           const uint32_t* __restrict__ offsets, unsigned cell, CellMeta& m) {
  const int ci = int(cell % unsigned(g.nx)), cj = int((cell / unsigned(g.nx)) % unsigned(g.ny)),
            ck = int(cell / (unsigned(g.nx) * unsigned(g.ny)));
  m.bin = g.encode(ci, cj, ck);
  m.count = counts[m.bin];
  m.first = offsets[m.bin];
}

// The cell's corner field value of this lane (lanes < 48) and the first 64
// particle records (lanes < count).
struct CellData {
  double corner;
  Raw raw;
};

__device__ __forceinline__ void fetch_data(const Grid& g, const ParticleView& s, const double* __restrict__ etheta,
                                           const double* __restrict__ bnode, unsigned cell, uint32_t count,
                                           uint32_t first, unsigned lane, CellData& d) {
  if (lane < 48) {
    const int ci = int(cell % unsigned(g.nx)), cj = int((cell / unsigned(g.nx)) % unsigned(g.ny)),
              ck = int(cell / (unsigned(g.nx) * unsigned(g.ny)));
    unsigned at[8];
    corner_nodes(g, ci, cj, ck, at);
    const unsigned nodes = g.nodes();
    const unsigned comp = lane >> 3, a = lane & 7u;
    const double* f = comp < 3 ? bnode + size_t(comp) * nodes : etheta + size_t(comp - 3) * nodes;
    d.corner = f[at[a]];
  }
  if (lane < count) load_raw(s, first + lane, d.raw);
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return uint32_t(__builtin_amdgcn_readfirstlane(int(v))); }

// Each wave owns a contiguous range of interior cells (blocks remapped so an
// XCD's waves own neighbouring cells) and walks it with a two-stage pipeline:
// the metadata of cell c+2 and the corner fields and first 64 records of cell
// c+1 are in flight while cell c is pushed. A record whose key is not the
// cell's bin is flagged and left unchanged.
__global__ __launch_bounds__(kThreads) void k_move_cells(Grid g, MoverParams p, ParticleView s,
                                                         const uint32_t* __restrict__ counts,
                                                         const uint32_t* __restrict__ offsets,
                                                         const double* __restrict__ etheta,
                                                         const double* __restrict__ bnode,
                                                         unsigned* __restrict__ validation) {
  __shared__ double s_corner[kWaves][48];
  const unsigned lane = threadIdx.x & 63u;
  const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double* corner = s_corner[wave];
  const unsigned cells = g.cells();
  const unsigned waves_total = gridDim.x * kWaves;
  const unsigned wid = xcd_logical_block(blockIdx.x, gridDim.x) * kWaves + wave;
  const unsigned per = (cells + waves_total - 1) / waves_total;
  const unsigned c_begin = wid * per;
  const unsigned c_end = c_begin + per < cells ? c_begin + per : cells;
  if (c_begin >= c_end) return;
  unsigned flags = 0;
  CellMeta m0, m1{}, m2{};
  CellData d0, d1;
  fetch_meta(g, counts, offsets, c_begin, m0);
  m0.count = uniform(m0.count);
  m0.first = uniform(m0.first);
  if (c_begin + 1 < c_end) fetch_meta(g, counts, offsets, c_begin + 1, m1);
  fetch_data(g, s, etheta, bnode, c_begin, m0.count, m0.first, lane, d0);
  for (unsigned c = c_begin; c < c_end; ++c) {
    const bool more1 = c + 1 < c_end, more2 = c + 2 < c_end;
    if (more1) {
      m1.count = uniform(m1.count);
      m1.first = uniform(m1.first);
    }
    if (more2) fetch_meta(g, counts, offsets, c + 2, m2);
    if (more1) fetch_data(g, s, etheta, bnode, c + 1, m1.count, m1.first, lane, d1);
    const uint32_t count = m0.count, first = m0.first;
    if (count != 0) {
      const int ci = int(c % unsigned(g.nx)), cj = int((c / unsigned(g.nx)) % unsigned(g.ny)),
                ck = int(c / (unsigned(g.nx) * unsigned(g.ny)));
      if (lane < 48) corner[lane] = d0.corner;
      __builtin_amdgcn_wave_barrier();
      asm volatile("" ::: "memory");
      for (uint32_t o = lane; o < count; o += 64) {
        const uint32_t index = first + o;
        Raw r;
        if (o < 64)
          r = d0.raw;
        else
          load_raw(s, index, r);
        Pushed out;
        unsigned f = kInvalidKey;
        if (r.key == m0.bin) f = push_particle(g, p, r, ci, cj, ck, StagedCorners{corner}, out);
        flags |= f;
        if (f == 0u) store_pushed(s, index, out);
      }
      __builtin_amdgcn_wave_barrier();
      asm volatile("" ::: "memory");
    }
    m0 = m1;
    m1 = m2;
    d0 = d1;
  }
  if (flags) atomicOr(validation, flags);
}

// Grid-stride over particles in any order (gathers per particle).
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
    if (key < 0 || key >= bins) {
      flags |= kInvalidKey;
      continue;
    }
    g.decode(key, ci, cj, ck);
    if (!g.interior(ci, cj, ck)) {
      flags |= kInvalidKey;
      continue;
    }
    NodeCorners corner{bnode, etheta, nodes, {}};
    corner_nodes(g, ci, cj, ck, corner.at);
    Raw r;
    load_raw(s, index, r);
    Pushed out;
    const unsigned f = push_particle(g, p, r, ci, cj, ck, corner, out);
    flags |= f;
    if (f == 0u) store_pushed(s, index, out);
  }
  if (flags) atomicOr(validation, flags);
}

// prepare_particle_mover_parameters (particle_mover_parameters.hpp:23-53):
// beta = (q/m) dt / 2, beta / c and 1 / spacing formed in double.
MoverParams mover_params(const Context& ctx, const SpeciesState& sp) {
  MoverParams p;
  const double beta = sp.q_over_m * ctx.num.dt / 2;
  p.beta = beta;
  p.beta_over_c = beta / ctx.num.c;
  p.dt = ctx.num.dt;
  p.inverse_spacing[0] = 1.0 / ctx.grid.dx;
  p.inverse_spacing[1] = 1.0 / ctx.grid.dy;
  p.inverse_spacing[2] = 1.0 / ctx.grid.dz;
  return p;
}

unsigned cell_walk_blocks(const Context& ctx) {
  const unsigned want = (ctx.grid.cells() + kWaves - 1) / kWaves;
  const unsigned cap = unsigned(ctx.cus > 0 ? ctx.cus : 256) * 32;
  return want < cap ? want : cap;
}

}  // namespace

int launch_move(Context& ctx, int s) {
  SpeciesState& sp = ctx.species[size_t(s)];
  if (sp.count == 0) return 0;
  const MoverParams p = mover_params(ctx, sp);
  phase_begin(ctx, Phase::move);
  if (sp.sorted) {
    const unsigned blocks = cell_walk_blocks(ctx);
    k_move_cells<<<blocks, kThreads, 0, ctx.stream>>>(ctx.grid, p, sp.view(), sp.counts, sp.offsets, ctx.particle_field(),
                                                      ctx.bn, ctx.validation);
  } else {
    const unsigned blocks = unsigned(std::min<uint64_t>((uint64_t(sp.count) + kThreads - 1) / kThreads,
                                                        uint64_t(ctx.cus > 0 ? ctx.cus : 256) * 32));
    k_move<<<blocks, kThreads, 0, ctx.stream>>>(ctx.grid, p, sp.view(), sp.count, ctx.particle_field(), ctx.bn,
                                                ctx.validation);
  }
  const int rc = check_hip(ctx, hipGetLastError(), "mover");
  phase_end(ctx, Phase::move);
  sp.sorted = false;
  return rc;
}

}  // namespace tpic_mi
