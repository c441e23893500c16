// Current and mass-matrix deposition from cell-sorted particles, and the
// deterministic node assembly of both.
//
// What is computed (FP64): for every particle p of a cell, with CIC weights
// W_a (8 corners), gathered B^{n+1} at nodes, response alpha
// (include/tpic/ecsim/particles/particle_response.hpp:17-34) and
// s = beta/c * q_p,
//   J_a   += (q_p/V) (alpha v)  W_a          (current_deposition_device.hpp:61-104)
//   K_abc += (s alpha_c) W_a W_b             (mass_matrix_scalar_device.hpp:25-97,
//                                             tests/ecsim/sorted_mass_device.cpp:111-182)
// MI355X formulation: W_a W_b depends only on m = a + b (per axis 0, 1, 2), so
// a cell's whole mass contribution is the 27 x 9 moment matrix
//   Q[m][c] = sum_p P_p[m] (s alpha_c)_p,   P_p[m] = X2[mx] Y2[my] Z2[mz],
//   X2 = ((1-x)^2, x(1-x), x^2),
// and the current is J_a = sum_b Qh[a+b] with Qh[m] = sum_p P_p[m] h_p
// (partition of unity: sum_b W_b = 1). One wave owns a cell; each group of four
// particles is two v_mfma_f64_16x16x4 products (rows m, columns the 9 alpha
// entries and 3 currents). The cell's 243 mass moments and 24 corner currents
// go to a q-major buffer; no atomics anywhere.
//
// Assembly: each mass record (vertex, stencil s, entry) is the sum over the
// (1, 2, 4 or 8) cells sharing the node pair of Q_cell[2a + s][entry], cells
// taken with periodic wrap in a fixed order, which also writes the ghost records
// the apply reads (MassHaloStorage fold, mass_halo_device.hpp:162-399); each
// current node sums its 8 surrounding cells (sum_shared of "J.add").
#include "internal.hpp"

namespace tpic_mi {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr unsigned kThreads = 256;
constexpr unsigned kWaves = kThreads / 64;

struct SpeciesDeposit {
  ParticleView p;
  const uint32_t* counts;
  const uint32_t* offsets;
  double beta_over_c;
};

struct DepositArgs {
  SpeciesDeposit sp[kMaxSpecies];
  int species;
};

// LDS written by a wave and read back by the same wave needs no wait: a
// wave's LDS operations execute in order. Only compiler reordering must be
// prevented.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}

// The raw record of one particle for the deposition (60 B).
struct Raw {
  double rx, ry, rz, u, v, w, q;
  int32_t key;
};

__device__ __forceinline__ void load_raw(const ParticleView& p, uint32_t index, Raw& r) {
  r.rx = p.rx[index];
  r.ry = p.ry[index];
  r.rz = p.rz[index];
  r.u = p.u[index];
  r.v = p.v[index];
  r.w = p.w[index];
  r.q = p.q[index];
  r.key = p.cell[index];
}

// Per-cell metadata held by the wave: lane s < species holds the count and
// first index of species s; lanes < 24 hold the cell's B corner values
// (component lane/8, corner lane%8). The bin is wave-uniform.
struct CellMeta {
  int32_t bin;
  uint32_t count, first;
  double b;
};

__device__ __forceinline__ void fetch_cell(const Grid& g, const DepositArgs& args, const double* __restrict__ bnode,
                                           unsigned cell, unsigned lane, CellMeta& m) {
  const int ci = int(cell % unsigned(g.nx)), cj = int((cell / unsigned(g.nx)) % unsigned(g.ny)),
            ck = int(cell / (unsigned(g.nx) * unsigned(g.ny)));
  m.bin = g.encode(ci, cj, ck);
  m.count = 0;
  m.first = 0;
#pragma unroll
  for (int s = 0; s < kMaxSpecies; ++s) {
    if (s < args.species && lane == unsigned(s)) {
      m.count = args.sp[s].counts[m.bin];
      m.first = args.sp[s].offsets[m.bin];
    }
  }
  m.b = 0.0;
  if (lane < 24) {
    const unsigned nodes = g.nodes();
    const unsigned n0 = g.node(ci, cj, ck);
    const unsigned sy = unsigned(g.nx + 1), sz = unsigned(g.nx + 1) * unsigned(g.ny + 1);
    const unsigned a = lane & 7u, comp = lane >> 3;
    const unsigned at = n0 + (a & 1u) + ((a >> 1) & 1u) * sy + ((a >> 2) & 1u) * sz;
    m.b = bnode[size_t(comp) * nodes + at];
  }
}

// The value of lane s (wave-uniform s) of a per-species register.
__device__ __forceinline__ uint32_t pick(uint32_t v, int s) { return uint32_t(__builtin_amdgcn_readlane(int(v), s)); }

// Each wave owns a contiguous range of interior cells (blocks remapped so an
// XCD's waves own neighbouring cells). Per 64-particle chunk every lane
// evaluates one particle (weights, B gather, response, currents) and stages
// its nine axis pair factors X2, Y2, Z2 and twelve MFMA columns (9 scaled
// alpha entries, 3 currents) in LDS; then each group of four particles is two
// FP64 MFMAs (rows m of P, columns of T), alternating between two accumulator
// sets. The next cell's metadata and B corners, and the next chunk's particle
// records, are loaded one step ahead so their latency overlaps the matrix work.
__global__ __launch_bounds__(kThreads) void k_deposit(Grid g, DepositArgs args, const double* __restrict__ bnode,
                                                      double inverse_volume, double* __restrict__ cellbuf,
                                                      int accumulate, unsigned* __restrict__ validation) {
  __shared__ double s_b[kWaves][24];
  __shared__ double s_f[kWaves][64 * 9];
  __shared__ double s_t[kWaves][64 * 12];
  const unsigned lane = threadIdx.x & 63u;
  const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double* bcorner = s_b[wave];
  double* ff = s_f[wave];
  double* tt = s_t[wave];
  const unsigned cells = g.cells();
  // MFMA operand roles of this lane: A row m (two 16-row tiles), K = particle
  // lane/16 of the group, B column = lane%16.
  const int col = int(lane & 15u);
  const int kp = int(lane >> 4);
  const int m0 = col, m1 = 16 + col;
  const int f0x = m0 % 3, f0y = 3 + (m0 / 3) % 3, f0z = 6 + m0 / 9;
  const bool m1_valid = m1 < kMoments;
  const int f1x = m1_valid ? m1 % 3 : 0, f1y = m1_valid ? 3 + (m1 / 3) % 3 : 3, f1z = m1_valid ? 6 + m1 / 9 : 6;
  const double m1_scale = m1_valid ? 1.0 : 0.0;
  const int species = args.species;
  const unsigned waves_total = gridDim.x * kWaves;
  const unsigned wid = xcd_logical_block(blockIdx.x, gridDim.x) * kWaves + wave;
  const unsigned per = (cells + waves_total - 1) / waves_total;
  const unsigned c_begin = wid * per;
  const unsigned c_end = c_begin + per < cells ? c_begin + per : cells;
  if (c_begin >= c_end) return;
  unsigned flags = 0;
  CellMeta cur_cell, next_cell;
  fetch_cell(g, args, bnode, c_begin, lane, cur_cell);
  // Particle records of the chunk to be processed next (prefetched).
  Raw next;
  bool have_next = false;
  unsigned next_c = 0;
  int next_s = 0;
  uint32_t next_chunk = 0;
  for (unsigned c = c_begin; c < c_end; ++c) {
    const bool more = c + 1 < c_end;
    if (more) fetch_cell(g, args, bnode, c + 1, lane, next_cell);
    const int32_t bin = cur_cell.bin;
    if (lane < 24) bcorner[lane] = cur_cell.b;
    wave_sync();
    d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    d4 acc2 = {0.0, 0.0, 0.0, 0.0}, acc3 = {0.0, 0.0, 0.0, 0.0};
    for (int s = 0; s < species; ++s) {
      const uint32_t count = pick(cur_cell.count, s);
      const uint32_t first = pick(cur_cell.first, s);
      const ParticleView pv = args.sp[s].p;
      const double beta_over_c = args.sp[s].beta_over_c;
      for (uint32_t chunk = 0; chunk < count; chunk += 64) {
        Raw cur;
        if (have_next && next_c == c && next_s == s && next_chunk == chunk) {
          cur = next;
        } else if (chunk + lane < count) {
          load_raw(pv, first + chunk + lane, cur);
        }
        // Successor: next chunk of this species, first chunk of a later
        // species of this cell, or the first chunk of the next cell.
        {
          bool found = false;
          unsigned sc = c;
          int ss = s;
          uint32_t sk = chunk + 64, scount = count, sfirst = first;
          if (sk < count) {
            found = true;
          } else {
            sk = 0;
            for (int s2 = s + 1; s2 < species && !found; ++s2) {
              const uint32_t k2 = pick(cur_cell.count, s2);
              if (k2) {
                ss = s2;
                scount = k2;
                sfirst = pick(cur_cell.first, s2);
                found = true;
              }
            }
            for (int s2 = 0; more && s2 < species && !found; ++s2) {
              const uint32_t k2 = pick(next_cell.count, s2);
              if (k2) {
                sc = c + 1;
                ss = s2;
                scount = k2;
                sfirst = pick(next_cell.first, s2);
                found = true;
              }
            }
          }
          have_next = found;
          if (found) {
            next_c = sc;
            next_s = ss;
            next_chunk = sk;
            if (sk + lane < scount) load_raw(args.sp[ss].p, sfirst + sk + lane, next);
          }
        }
        double f[9], t[12];
#pragma unroll
        for (int e = 0; e < 9; ++e) f[e] = 0.0;
#pragma unroll
        for (int e = 0; e < 12; ++e) t[e] = 0.0;
        if (chunk + lane < count) {
          const double fx = cur.rx, fy = cur.ry, fz = cur.rz;
          bool ok = cur.key == bin;
          if (!ok) flags |= kInvalidKey;
          if (ok && !(fx >= 0.0 && fx < 1.0 && fy >= 0.0 && fy < 1.0 && fz >= 0.0 && fz < 1.0)) {
            flags |= kInvalidFraction;
            ok = false;
          }
          if (ok) {
            const double cx = 1.0 - fx, cy = 1.0 - fy, cz = 1.0 - fz;
            const double w[8] = {cx * cy * cz, fx * cy * cz, cx * fy * cz, fx * fy * cz,
                                 cx * cy * fz, fx * cy * fz, cx * fy * fz, fx * fy * fz};
            double bx = 0.0, by = 0.0, bz = 0.0;
#pragma unroll
            for (int a = 0; a < 8; ++a) {
              bx += w[a] * bcorner[a];
              by += w[a] * bcorner[8 + a];
              bz += w[a] * bcorner[16 + a];
            }
            const double X = beta_over_c * bx, Y = beta_over_c * by, Z = beta_over_c * bz;
            const double n2 = X * X + Y * Y + Z * Z;
            const double inv = 1.0 / (1.0 + n2);
            double al[9];
            al[0] = (1.0 + X * X) * inv;
            al[4] = (1.0 + Y * Y) * inv;
            al[8] = (1.0 + Z * Z) * inv;
            const double xy = X * Y * inv, xz = X * Z * inv, yz = Y * Z * inv;
            const double sx = X * inv, syy = Y * inv, szz = Z * inv;
            al[1] = xy + szz;
            al[2] = xz - syy;
            al[3] = xy - szz;
            al[5] = yz + sx;
            al[6] = xz + syy;
            al[7] = yz - sx;
            const double q = cur.q;
            const double scale = beta_over_c * q;
            const double qv = q * inverse_volume;
            const double h0 = qv * (al[0] * cur.u + al[1] * cur.v + al[2] * cur.w);
            const double h1 = qv * (al[3] * cur.u + al[4] * cur.v + al[5] * cur.w);
            const double h2 = qv * (al[6] * cur.u + al[7] * cur.v + al[8] * cur.w);
            if (!finite_value(n2) || !finite_value(scale) || !finite_value(h0) || !finite_value(h1) ||
                !finite_value(h2)) {
              flags |= kInvalidNonfinite;
            } else {
              // Axis pair factors: (1-x)^2, x(1-x), x^2 per axis.
              f[0] = cx * cx; f[1] = fx * cx; f[2] = fx * fx;
              f[3] = cy * cy; f[4] = fy * cy; f[5] = fy * fy;
              f[6] = cz * cz; f[7] = fz * cz; f[8] = fz * fz;
#pragma unroll
              for (int e = 0; e < 9; ++e) t[e] = al[e] * scale;
              t[9] = h0;
              t[10] = h1;
              t[11] = h2;
            }
          }
        }
#pragma unroll
        for (int e = 0; e < 9; ++e) ff[lane * 9 + e] = f[e];
#pragma unroll
        for (int e = 0; e < 12; ++e) tt[lane * 12 + e] = t[e];
        wave_sync();
        const uint32_t populated = count - chunk < 64u ? count - chunk : 64u;
        // Rows at or beyond `populated` hold zeros.
        for (uint32_t g8 = 0; g8 < populated; g8 += 8) {
          {
            const unsigned p = g8 + unsigned(kp);
            const double* fp = ff + p * 9;
            const double tv = col < 12 ? tt[p * 12 + col] : 0.0;
            const double a0 = fp[f0x] * fp[f0y] * fp[f0z];
            const double a1 = fp[f1x] * fp[f1y] * fp[f1z] * m1_scale;
            acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, tv, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, tv, acc1, 0, 0, 0);
          }
          {
            const unsigned p = g8 + 4u + unsigned(kp);
            const double* fp = ff + p * 9;
            const double tv = col < 12 ? tt[p * 12 + col] : 0.0;
            const double a0 = fp[f0x] * fp[f0y] * fp[f0z];
            const double a1 = fp[f1x] * fp[f1y] * fp[f1z] * m1_scale;
            acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, tv, acc2, 0, 0, 0);
            acc3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, tv, acc3, 0, 0, 0);
          }
        }
        wave_sync();
      }
    }
    acc0 += acc2;
    acc1 += acc3;
    // v_mfma_f64_16x16x4f64 result D[i][j]: lane j + 16*(i%4), element i/4
    // (A[i][k] is lane i + 16k, B[k][j] lane j + 16k).
    double* qm = tt;  // reuse: Q[m][col], 27 x 12
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = 4 * r + kp;
      if (col < 12) {
        qm[m * 12 + col] = acc0[r];
        if (16 + m < kMoments) qm[(16 + m) * 12 + col] = acc1[r];
      }
    }
    wave_sync();
    for (unsigned qi = lane; qi < unsigned(kCellQuantities); qi += 64) {
      double value;
      if (qi < unsigned(kCellMass)) {
        value = qm[(qi / 9) * 12 + (qi % 9)];
      } else {
        const unsigned e = qi - unsigned(kCellCurrent);
        const int a = int(e / 3), comp = int(e % 3);
        const int ax = a & 1, ay = (a >> 1) & 1, az = (a >> 2) & 1;
        value = 0.0;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          const int m = (ax + (b & 1)) + 3 * (ay + ((b >> 1) & 1)) + 9 * (az + ((b >> 2) & 1));
          value += qm[m * 12 + 9 + comp];
        }
      }
      double* out = cellbuf + size_t(qi) * cells + c;
      if (accumulate)
        *out += value;
      else
        *out = value;
    }
    wave_sync();
    if (more) cur_cell = next_cell;
  }
  if (flags) atomicOr(validation, flags);
}

// Pair table of each canonical stencil: corners a with a + offset a corner,
// and the moment m = 2a + offset; ascending a (fixed summation order).
struct PairTable {
  int count[kStencils];
  int a[kStencils][8];
  int m[kStencils][8];
};
__constant__ PairTable c_pairs;

__device__ __forceinline__ int wrap_vertex(int i, int n) { return i < 0 ? n - 1 : (i > n ? 1 : i); }

__global__ __launch_bounds__(kThreads) void k_assemble_mass(Grid g, const double* __restrict__ cellbuf,
                                                            double* __restrict__ mass,
                                                            unsigned* __restrict__ validation) {
  const unsigned verts = g.vertices();
  const unsigned cells = g.cells();
  const unsigned ex = unsigned(g.nx + 3), ey = unsigned(g.ny + 3);
  const uint64_t total = uint64_t(verts) * kEntries;
  unsigned flags = 0;
  for (uint64_t idx = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
       idx += uint64_t(gridDim.x) * blockDim.x) {
    const unsigned e = unsigned(idx / verts), v = unsigned(idx % verts);
    const int s = int(e / 9u), comp = int(e % 9u);
    const int i = wrap_vertex(int(v % ex) - 1, g.nx), j = wrap_vertex(int((v / ex) % ey) - 1, g.ny),
              k = wrap_vertex(int(v / (ex * ey)) - 1, g.nz);
    double sum = 0.0;
    const int np = c_pairs.count[s];
    for (int t = 0; t < np; ++t) {
      const int a = c_pairs.a[s][t];
      const int cx = wrap_cell(i - (a & 1), g.nx), cy = wrap_cell(j - ((a >> 1) & 1), g.ny),
                cz = wrap_cell(k - ((a >> 2) & 1), g.nz);
      sum += cellbuf[size_t(c_pairs.m[s][t] * 9 + comp) * cells + g.cell(cx, cy, cz)];
    }
    if (!finite_value(sum)) flags |= kInvalidNonfinite;
    mass[size_t(e) * g.vstride + v] = sum;
  }
  if (flags) atomicOr(validation, flags);
}

__global__ __launch_bounds__(kThreads) void k_assemble_current(Grid g, const double* __restrict__ cellbuf,
                                                               double* __restrict__ current) {
  const unsigned nodes = g.nodes(), cells = g.cells();
  const unsigned sx = unsigned(g.nx + 1), sy = unsigned(g.ny + 1);
  for (unsigned idx = blockIdx.x * blockDim.x + threadIdx.x; idx < 3 * nodes; idx += gridDim.x * blockDim.x) {
    const unsigned comp = idx / nodes, n = idx % nodes;
    const int i = int(n % sx), j = int((n / sx) % sy), k = int(n / (sx * sy));
    double sum = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int cx = wrap_cell(i - (a & 1), g.nx), cy = wrap_cell(j - ((a >> 1) & 1), g.ny),
                cz = wrap_cell(k - ((a >> 2) & 1), g.nz);
      sum += cellbuf[size_t(kCellCurrent + a * 3 + int(comp)) * cells + g.cell(cx, cy, cz)];
    }
    current[idx] = sum;
  }
}

// Entry-major mass -> the reference's node-major records (vertex*126 + e).
__global__ __launch_bounds__(kThreads) void k_mass_records(Grid g, const double* __restrict__ mass,
                                                           double* __restrict__ out) {
  const unsigned verts = g.vertices();
  const uint64_t total = uint64_t(verts) * kEntries;
  for (uint64_t idx = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
       idx += uint64_t(gridDim.x) * blockDim.x) {
    const unsigned v = unsigned(idx / kEntries), e = unsigned(idx % kEntries);
    out[idx] = mass[size_t(e) * g.vstride + v];
  }
}

const int kOffsets[kStencils][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, -1, 0}, {1, 0, 1},
                                    {1, 0, -1}, {0, 1, 1}, {0, 1, -1}, {1, -1, 1}, {1, 1, -1}, {-1, 1, 1}, {1, 1, 1}};

bool g_pairs_ready[64] = {};

int ensure_pairs(Context& ctx) {
  if (ctx.device >= 0 && ctx.device < 64 && g_pairs_ready[ctx.device]) return 0;
  PairTable t{};
  for (int s = 0; s < kStencils; ++s) {
    int n = 0;
    for (int a = 0; a < 8; ++a) {
      const int ax = a & 1, ay = (a >> 1) & 1, az = (a >> 2) & 1;
      const int bx = ax + kOffsets[s][0], by = ay + kOffsets[s][1], bz = az + kOffsets[s][2];
      if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
      t.a[s][n] = a;
      t.m[s][n] = (ax + bx) + 3 * (ay + by) + 9 * (az + bz);
      ++n;
    }
    t.count[s] = n;
  }
  const int rc = check_hip(ctx, hipMemcpyToSymbol(HIP_SYMBOL(c_pairs), &t, sizeof t), "pair table");
  if (rc == 0 && ctx.device >= 0 && ctx.device < 64) g_pairs_ready[ctx.device] = true;
  return rc;
}

unsigned grid_blocks(const Context& ctx, uint64_t work) {
  const uint64_t want = (work + kThreads - 1) / kThreads;
  const uint64_t cap = uint64_t(ctx.cus > 0 ? ctx.cus : 256) * 64;
  return unsigned(want < cap ? (want ? want : 1) : cap);
}

}  // namespace

int launch_deposit(Context& ctx, int first, int last, bool accumulate) {
  DepositArgs args{};
  args.species = 0;
  for (int s = first; s < last; ++s) {
    const SpeciesState& sp = ctx.species[size_t(s)];
    if (!sp.sorted) return fail(ctx, -2, "deposition needs a sorted species (call tpic_mi_sort first)");
    SpeciesDeposit& d = args.sp[args.species++];
    d.p = sp.view();
    d.counts = sp.counts;
    d.offsets = sp.offsets;
    // deposition_actions_device.hpp:72-74: beta/c = (q/m) dt / (2c), in double.
    d.beta_over_c = sp.q_over_m * ctx.num.dt / (2 * ctx.num.c);
  }
  const unsigned cells = ctx.grid.cells();
  // Each wave owns a contiguous range of cells; about 8 waves per SIMD slot.
  const unsigned want = (cells + kWaves - 1) / kWaves;
  const unsigned cap = unsigned(ctx.cus > 0 ? ctx.cus : 256) * 16;
  const unsigned blocks = want < cap ? want : cap;
  phase_begin(ctx, Phase::deposit);
  k_deposit<<<blocks, kThreads, 0, ctx.stream>>>(ctx.grid, args, ctx.bn, ctx.inverse_volume, ctx.cellbuf,
                                                 accumulate ? 1 : 0, ctx.validation);
  const int rc = check_hip(ctx, hipGetLastError(), "deposit");
  phase_end(ctx, Phase::deposit);
  return rc;
}

int launch_assemble(Context& ctx) {
  int rc = ensure_pairs(ctx);
  if (rc) return rc;
  phase_begin(ctx, Phase::assemble);
  const Grid& g = ctx.grid;
  k_assemble_mass<<<grid_blocks(ctx, uint64_t(g.vertices()) * kEntries), kThreads, 0, ctx.stream>>>(
      g, ctx.cellbuf, ctx.mass, ctx.validation);
  if ((rc = check_hip(ctx, hipGetLastError(), "mass assembly"))) return rc;
  k_assemble_current<<<grid_blocks(ctx, 3ull * g.nodes()), kThreads, 0, ctx.stream>>>(g, ctx.cellbuf,
                                                                                      ctx.current);
  rc = check_hip(ctx, hipGetLastError(), "current assembly");
  phase_end(ctx, Phase::assemble);
  return rc;
}

int launch_mass_to_records(Context& ctx, double* out) {
  k_mass_records<<<grid_blocks(ctx, uint64_t(ctx.grid.vertices()) * kEntries), kThreads, 0, ctx.stream>>>(
      ctx.grid, ctx.mass, out);
  return check_hip(ctx, hipGetLastError(), "mass records");
}

}  // namespace tpic_mi
