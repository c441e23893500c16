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
// Two kernels share the push: the sorted kernel walks the sorted species in
// aligned 64-particle blocks, stages the 48 corner field values of the block's
// bins once in LDS and reads them back, so a particle costs only its own 60 B
// read and 52 B write; the unsorted kernel gathers the corners per particle
// (input that was never sorted).
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

// The sorted mover walks the species array in aligned blocks of 64
// particles, not cell by cell: every record line is read once and every
// block is one full pass of the wave (a cell walk reads the line a cell
// shares with the next one twice and spends a second, nearly empty pass on a
// cell holding more than 64 particles). The particles of a block lie in a few
// consecutive bins; a pass stages the 48 corner field values of kSlots
// consecutive bins in LDS and pushes the lanes whose key is one of them.
constexpr unsigned kSlots = 2;
constexpr unsigned kSlotValues = kSlots * 48;
constexpr unsigned kSlotLoads = (kSlotValues + 63) / 64;

// The bin whose range holds sorted position `index` (the last bin whose
// first index is at most `index`), or -1 when no bin's range holds it.
__device__ int32_t holding_bin(const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ counts,
                               uint32_t bins, uint32_t index) {
  if (bins == 0 || offsets[0] > index) return -1;
  uint32_t lo = 0, hi = bins;
  while (hi - lo > 1) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (offsets[mid] <= index)
      lo = mid;
    else
      hi = mid;
  }
  return index - offsets[lo] < counts[lo] ? int32_t(lo) : -1;
}

// A record the block walk does not push is flagged when the cell walk of the
// bins would have visited it and found a key other than the bin: its position
// lies in an interior bin's range and the key differs. Records in halo bin
// ranges, or in no range (a failed sort leaves every range empty), are left
// as they are without a flag.
__device__ unsigned unpushed_flags(const Grid& g, const uint32_t* __restrict__ offsets,
                                   const uint32_t* __restrict__ counts, uint32_t index, int32_t key) {
  const int32_t hb = holding_bin(offsets, counts, g.bins(), index);
  if (hb < 0 || hb == key) return 0u;
  int i, j, k;
  g.decode(hb, i, j, k);
  return g.interior(i, j, k) ? unsigned(kInvalidKey) : 0u;
}

// Lane values of a pass's staged corners: value v = lane + 64 t is corner
// v % 8 of component (v % 48) / 8 of the cell of bin base + v / 48, loaded
// when that bin is an interior cell in base's row (others are never read).
__device__ __forceinline__ void load_slots(const Grid& g, int32_t base, const double* __restrict__ bnode,
                                           const double* __restrict__ etheta, unsigned lane,
                                           double (&cv)[kSlotLoads]) {
  int bi, bj, bk;
  g.decode(base, bi, bj, bk);
  const bool row = bj >= 0 && bj < g.ny && bk >= 0 && bk < g.nz;
  const unsigned nodes = g.nodes();
  const unsigned sy = unsigned(g.nx + 1), sz = unsigned(g.nx + 1) * unsigned(g.ny + 1);
#pragma unroll
  for (unsigned t = 0; t < kSlotLoads; ++t) {
    const unsigned v = lane + 64u * t;
    cv[t] = 0.0;
    const int i = bi + int(v / 48u);
    if (v < kSlotValues && row && i >= 0 && i < g.nx) {
      const unsigned jv = v % 48u, comp = jv >> 3, a = jv & 7u;
      const unsigned at = g.node(i, bj, bk) + (a & 1u) + ((a >> 1) & 1u) * sy + ((a >> 2) & 1u) * sz;
      cv[t] = comp < 3 ? bnode[size_t(comp) * nodes + at] : etheta[size_t(comp - 3) * nodes + at];
    }
  }
}

__device__ __forceinline__ void stage_slots(double* corner, unsigned lane, const double (&cv)[kSlotLoads]) {
#pragma unroll
  for (unsigned t = 0; t < kSlotLoads; ++t)
    if (lane + 64u * t < kSlotValues) corner[lane + 64u * t] = cv[t];
}

// The first bin of a block's passes: the key of its first lane with a key in
// range (wave-uniform), or -1 when it has none.
__device__ __forceinline__ int32_t first_base(bool candidate, int32_t key) {
  const uint64_t m = __builtin_amdgcn_ballot_w64(candidate);
  if (m == 0) return -1;
  return __builtin_amdgcn_readlane(key, int(__builtin_ctzll(m)));
}

__device__ __forceinline__ bool key_in_range(int32_t key, uint32_t bins) { return key >= 0 && uint32_t(key) < bins; }

// Each wave owns a contiguous range of blocks (remapped so an XCD's waves own
// neighbouring blocks) and walks it with a two-stage pipeline: the keys of
// block b+2, and the records, key ranges and first-pass corner fields of
// block b+1, are in flight while block b is pushed; block b+1's corners are
// staged in the second LDS buffer once b is done.
__global__ __launch_bounds__(kThreads) void k_move_blocks(Grid g, MoverParams p, ParticleView s, uint32_t n,
                                                          const uint32_t* __restrict__ counts,
                                                          const uint32_t* __restrict__ offsets,
                                                          const double* __restrict__ etheta,
                                                          const double* __restrict__ bnode,
                                                          unsigned* __restrict__ validation,
                                                          uint64_t* __restrict__ cross) {
  __shared__ double s_corner[kWaves][2][kSlotValues];
  const unsigned lane = threadIdx.x & 63u;
  const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t bins = g.bins();
  const uint32_t nblocks = (n + 63u) / 64u;
  const unsigned waves_total = gridDim.x * kWaves;
  const unsigned wid = xcd_logical_block(blockIdx.x, gridDim.x) * kWaves + wave;
  const uint32_t per = (nblocks + waves_total - 1) / waves_total;
  const uint32_t b_begin = wid * per;
  const uint32_t b_end = b_begin + per < nblocks ? b_begin + per : nblocks;
  if (b_begin >= b_end) return;
  unsigned flags = 0;
  unsigned buf = 0;
  // Block b_begin: records, whether each lane lies in its key's bin range,
  // first-pass corners staged; block b_begin + 1: keys.
  Raw cur;
  cur.key = -1;
  bool cur_in_range = false;
  int32_t base0;
  {
    const uint32_t index = b_begin * 64u + lane;
    if (index < n) load_raw(s, index, cur);
    if (index < n && key_in_range(cur.key, bins)) cur_in_range = index - offsets[cur.key] < counts[cur.key];
    base0 = first_base(index < n && key_in_range(cur.key, bins), cur.key);
    if (base0 >= 0) {
      double cv[kSlotLoads];
      load_slots(g, base0, bnode, etheta, lane, cv);
      stage_slots(s_corner[wave][buf], lane, cv);
    }
  }
  int32_t key1 = -1;
  if (b_begin + 1 < b_end && (b_begin + 1) * 64u + lane < n) key1 = s.cell[(b_begin + 1) * 64u + lane];
  for (uint32_t b = b_begin; b < b_end; ++b) {
    // Block b + 1: records, key ranges and first-pass corners; block b + 2: keys.
    const bool more = b + 1 < b_end;
    const uint32_t index1 = (b + 1) * 64u + lane;
    Raw nxt;
    nxt.key = -1;
    uint32_t first1 = 0, count1 = 0;
    double cv_next[kSlotLoads];
    int32_t base1 = -1;
    if (more) {
      // The key-dependent requests first (their lane reads of key1 then wait
      // only for key1's own load, issued a block earlier), then the records:
      // all of them stay in flight while block b is pushed.
      const bool ranged = index1 < n && key_in_range(key1, bins);
      base1 = first_base(ranged, key1);
      if (ranged) {
        first1 = offsets[key1];
        count1 = counts[key1];
      }
      if (base1 >= 0) load_slots(g, base1, bnode, etheta, lane, cv_next);
      if (index1 < n) load_raw(s, index1, nxt);
    }
    int32_t key2 = -1;
    if (b + 2 < b_end && (b + 2) * 64u + lane < n) key2 = s.cell[(b + 2) * 64u + lane];
    // Push block b: the lanes whose key is one of the kSlots staged bins from
    // base0 in base0's row read the staged corners; any other lane (a block
    // spanning more bins, or a row's end) gathers its corners from the node
    // fields.
    const double* corner = s_corner[wave][buf];
    const uint32_t index = b * 64u + lane;
    const int32_t key = cur.key;
    const bool candidate = index < n && key_in_range(key, bins);
    bool unpushed = index < n && !candidate;
    bool gather = false;
    bool crossed = false;  // pushed into another cell
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
    if (candidate) {
      int bi, bj, bk;
      g.decode(base0, bi, bj, bk);
      const int d = key - base0;
      if (d >= 0 && d < int(kSlots) && bi + d <= g.nx + 1) {
        const int ci = bi + d;
        const bool cell_ok = ci >= 0 && ci < g.nx && bj >= 0 && bj < g.ny && bk >= 0 && bk < g.nz;
        if (cell_ok && cur_in_range) {
          Pushed out;
          const unsigned f = push_particle(g, p, cur, ci, bj, bk, StagedCorners{corner + d * 48}, out);
          flags |= f;
          if (f == 0u) {
            store_pushed(s, index, out);
            crossed = out.key != key;
          }
        } else {
          unpushed = true;
        }
      } else {
        gather = true;
      }
    }
    if (__builtin_amdgcn_ballot_w64(gather) != 0 && gather) {
      int ci, cj, ck;
      g.decode(key, ci, cj, ck);
      if (g.interior(ci, cj, ck) && cur_in_range) {
        NodeCorners nc{bnode, etheta, g.nodes(), {}};
        corner_nodes(g, ci, cj, ck, nc.at);
        Pushed out;
        const unsigned f = push_particle(g, p, cur, ci, cj, ck, nc, out);
        flags |= f;
        if (f == 0u) {
          store_pushed(s, index, out);
          crossed = out.key != key;
        }
      } else {
        unpushed = true;
      }
    }
    if (cross) {
      const uint64_t m = __builtin_amdgcn_ballot_w64(crossed);
      if (lane == 0) cross[b] = m;
    }
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
    if (unpushed) flags |= unpushed_flags(g, offsets, counts, index, key);
    // Block b + 1 becomes current: its corners go to the other buffer.
    if (more && base1 >= 0) stage_slots(s_corner[wave][buf ^ 1u], lane, cv_next);
    cur = nxt;
    cur_in_range = index1 - first1 < count1;
    base0 = base1;
    key1 = key2;
    buf ^= 1u;
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

// About 32 blocks of 4 waves per CU, each wave a contiguous range of
// 64-particle blocks.
unsigned block_walk_blocks(const Context& ctx, uint32_t n) {
  const uint64_t want = (uint64_t((n + 63u) / 64u) + kWaves - 1) / kWaves;
  const uint64_t cap = uint64_t(ctx.cus > 0 ? ctx.cus : 256) * 32;
  return unsigned(std::max<uint64_t>(1, want < cap ? want : cap));
}

}  // namespace

int launch_move(Context& ctx, int s, bool record_crossers) {
  SpeciesState& sp = ctx.species[size_t(s)];
  if (sp.count == 0) return 0;
  const MoverParams p = mover_params(ctx, sp);
  phase_begin(ctx, Phase::move);
  if (sp.sorted) {
    const unsigned blocks = block_walk_blocks(ctx, sp.count);
    k_move_blocks<<<blocks, kThreads, 0, ctx.stream>>>(ctx.grid, p, sp.view(), sp.count, sp.counts, sp.offsets,
                                                       ctx.particle_field(), ctx.bn, ctx.validation,
                                                       record_crossers ? sp.cross : nullptr);
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
