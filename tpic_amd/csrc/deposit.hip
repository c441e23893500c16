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
// a cell's whole contribution is the 27 x 12 moment matrix
//   Q[m][c] = sum_p P_p[m] T_p[c],   P_p[m] = X2[mx] Y2[my] Z2[mz],
//   X2 = ((1-x)^2, x(1-x), x^2),     T_p = (s alpha_0..8, (q/V) alpha v),
// and the current of corner a is J_a = sum_b Q[a+b][9..11] (partition of
// unity: sum_b W_b = 1). One wave owns a cell at a time; each group of four
// particles is two v_mfma_f64_16x16x4 products (rows m, columns c). The cell
// mass record [m * 9 + c] (the 9 mass entries of each moment m) is cell-major
// and contiguous, 1944 B per cell: one cell's store instructions write it
// whole, so lines shared by two consecutive cells of a wave are completed
// within a few microseconds (no partial-line write backs). The cell's eight
// corner currents J_a = sum_b Q[a+b][9 + comp] are summed through LDS from the
// accumulators' current columns and written as the contiguous corner record
// [a * 3 + comp] (192 B per cell). No atomics anywhere.
//
// Assembly: each mass record (vertex, stencil s, entry) is the sum over the
// (1, 2, 4 or 8) cells sharing the node pair of Q_cell[2a + s][entry], cells
// taken with periodic wrap in a fixed order, which also writes the ghost records
// the apply reads (MassHaloStorage fold, mass_halo_device.hpp:162-399); each
// current node sums, over its 8 surrounding cells, the corner current
// J_a = sum_b Q[a+b][9 + comp] (sum_shared of "J.add").
#include "internal.hpp"

#include <mutex>

namespace tpic_mi {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Operand precision of the per-cell moment products. A mode owns the LDS
// layout the evaluated particles are staged in (`stage`) and the product
// of a staged chunk into the two 16 x 16 accumulators (`contract`, rows m of
// P, columns c of T); `row(r, kp)` names the product row held by element r
// of lane group kp = lane / 16 (column = lane % 16).
constexpr unsigned kThreads = 256;
constexpr unsigned kWaves = kThreads / 64;
constexpr int kChunk = 64;     // particles evaluated per wave pass

struct SpeciesDeposit {
  ParticleView p;           // the particles deposited (sorted), or the sources of a fused gather
  ParticleView dst;         // fused gather: the sorted records are written here
  const uint32_t* source;   // fused gather: source of every destination
  const uint32_t* counts;
  const uint32_t* offsets;
  double beta_over_c;
};

struct DepositArgs {
  SpeciesDeposit sp[kMaxSpecies];
  int species;
  // Bin counts and offsets of the deposited species, species-major
  // (species s of the launch at rows s of [species][bins] / [species][bins+1]).
  const uint32_t* counts;
  const uint32_t* offsets;
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

// With Key false the key is not read: a record gathered by the sort plan
// into a bin's range has that bin as its key.
template <bool Key = true>
__device__ __forceinline__ void load_raw(const ParticleView& p, uint32_t index, Raw& r) {
  r.rx = p.rx[index];
  r.ry = p.ry[index];
  r.rz = p.rz[index];
  r.u = p.u[index];
  r.v = p.v[index];
  r.w = p.w[index];
  r.q = p.q[index];
  if (Key) r.key = p.cell[index];
}

// The particle at sorted position d. With Gather the record comes from its
// source position of the species' sort plan, together with its id, and d is
// kept as the destination it is written to once evaluated.
template <bool Gather>
__device__ __forceinline__ void fetch_particle(const SpeciesDeposit& sp, uint32_t d, Raw& r, uint64_t& id,
                                               uint32_t& dest) {
  if (Gather) {
    const uint32_t from = sp.source[d];
    load_raw<false>(sp.p, from, r);
    id = sp.p.id[from];
    dest = d;
  } else {
    load_raw(sp.p, d, r);
  }
}

// The nine attributes of a gathered particle written at its destination.
__device__ __forceinline__ void store_particle(const ParticleView& p, uint32_t d, const Raw& r, uint64_t id) {
  p.rx[d] = r.rx;
  p.ry[d] = r.ry;
  p.rz[d] = r.rz;
  p.u[d] = r.u;
  p.v[d] = r.v;
  p.w[d] = r.w;
  p.q[d] = r.q;
  p.cell[d] = r.key;
  p.id[d] = id;
}

// Per-cell metadata held by the wave: lane s < species of `cf` holds the
// count of species s, lane kMaxSpecies + s its first index; lanes < 24 of b
// hold the cell's B corner values (component lane/8, corner lane%8) in the
// evaluation arithmetic B. The bin is wave-uniform.
template <class B>
struct CellMeta {
  int32_t bin;
  uint32_t cf;
  B b;
};

// The lane's part of every cell's metadata, fixed for the launch: its row of
// the species counts (lanes < species) or first indices (lanes 8 + s), and
// its B corner value's offset (lanes < 24: component lane / 8, corner
// lane % 8) from the cell's lower node.
struct CellLanes {
  const uint32_t* row;  // or null
  unsigned b_off;
  bool b_lane;
};

__device__ __forceinline__ CellLanes cell_lanes(const Grid& g, const DepositArgs& args, unsigned lane) {
  CellLanes l{nullptr, 0u, lane < 24u};
  const unsigned bins = g.bins();
  if (lane < unsigned(args.species))
    l.row = args.counts + size_t(lane) * bins;
  else if (lane >= unsigned(kMaxSpecies) && lane < unsigned(kMaxSpecies + args.species))
    l.row = args.offsets + size_t(lane - kMaxSpecies) * (bins + 1);
  const unsigned sy = unsigned(g.nx + 1), sz = unsigned(g.nx + 1) * unsigned(g.ny + 1);
  const unsigned a = lane & 7u, comp = lane >> 3;
  if (l.b_lane) l.b_off = comp * g.nodes() + (a & 1u) + ((a >> 1) & 1u) * sy + ((a >> 2) & 1u) * sz;
  return l;
}

template <class B>
__device__ __forceinline__ void fetch_cell(const Grid& g, const CellLanes& l, const double* __restrict__ bnode,
                                           unsigned cell, CellMeta<B>& m) {
  const int ci = int(cell % unsigned(g.nx)), cj = int((cell / unsigned(g.nx)) % unsigned(g.ny)),
            ck = int(cell / (unsigned(g.nx) * unsigned(g.ny)));
  m.bin = g.encode(ci, cj, ck);
  m.cf = l.row ? l.row[unsigned(m.bin)] : 0u;
  m.b = l.b_lane ? B(bnode[l.b_off + g.node(ci, cj, ck)]) : B(0);
}

// The value of lane s (wave-uniform s) of a per-species register.
__device__ __forceinline__ uint32_t pick(uint32_t v, int s) { return uint32_t(__builtin_amdgcn_readlane(int(v), s)); }

// Count and first index of species s (wave-uniform) in a cell.
template <class B>
__device__ __forceinline__ uint32_t count_of(const CellMeta<B>& m, int s) { return pick(m.cf, s); }
template <class B>
__device__ __forceinline__ uint32_t first_of(const CellMeta<B>& m, int s) { return pick(m.cf, kMaxSpecies + s); }

// Index of column j of moment m of a cell in the deposition buffer.
__host__ __device__ __forceinline__ size_t record_index(int m, unsigned cell, int j) {
  return size_t(cell) * kCellRecord + unsigned(m * kMassColumns + j);
}

__device__ __forceinline__ void store_row(double* row, const double* v, int n) {
#pragma unroll
  for (int e = 0; e < n; e += 2) *reinterpret_cast<d2*>(row + e) = d2{v[e], v[e + 1]};
}

__device__ __forceinline__ void store_row(float* row, const float* v, int n) {
#pragma unroll
  for (int e = 0; e < n; e += 4) *reinterpret_cast<f4*>(row + e) = f4{v[e], v[e + 1], v[e + 2], v[e + 3]};
}

__device__ __forceinline__ void store_row(float* row, const double* v, int n) {
#pragma unroll
  for (int e = 0; e < n; e += 4)
    *reinterpret_cast<f4*>(row + e) = f4{float(v[e]), float(v[e + 1]), float(v[e + 2]), float(v[e + 3])};
}

// K = 4 modes (FP64, FP32): each lane owns an LDS row per particle, Pxy[9],
// Z2[3], T[12] at a padded stride; each group of four particles is two
// 16x16x4 MFMAs in which lane (col, kp) supplies A = P[col] and P[16 + col]
// of particle kp (Pxy * Z2, one multiply) and B = T[col]. Rows m >= 27 and
// columns >= 12 read neighbouring row data and are discarded.
template <class Elem, int Stride>
struct RowModeBase {
  using T = Elem;
  // Arithmetic of the particle evaluation: the operand type (FP64 for the
  // parity path; FP32 for the FP32 form, whose operands are FP32 anyway).
  using Compute = Elem;
  // Stride in elements: 52 dwords (FP64) or 28 dwords (FP32), so the
  // 16-byte row writes of 8 consecutive lanes cover distinct banks.
  static constexpr int kStride = Stride;
  static constexpr int kLdsElems = kChunk * kStride + 8;
  // Byte offset of corner slot (comp, m) in the wave's rows: the two doubles
  // after the 24 staged values of row m (comp 0, 1) or of row 32 + m (comp 2),
  // read by the products only as the discarded columns 12..15.
  __device__ static constexpr int corner_slot(int comp, int m) {
    return (m + 32 * (comp >> 1)) * kStride * int(sizeof(T)) + 24 * int(sizeof(T)) + 8 * (comp & 1);
  }

  __device__ static void stage(T* lds, unsigned lane, const Compute (&p)[12], const Compute (&t)[12]) {
    store_row(lds + lane * kStride, p, 12);
    store_row(lds + lane * kStride + 12, t, 12);
  }

  template <class Acc, class Mfma>
  __device__ static void contract_rows(const T* lds, unsigned lane, uint32_t populated, Acc& acc0, Acc& acc1,
                                       Mfma mfma) {
    const int col = int(lane & 15u), kp = int(lane >> 4);
    const uint32_t groups = (populated + 3u) / 4u;
    const T* p0x = lds + kp * kStride + col % 9;
    const T* p0z = lds + kp * kStride + 9 + col / 9;
    const T* p1x = lds + kp * kStride + (16 + col) % 9;
    const T* p1z = lds + kp * kStride + 9 + (16 + col) / 9;
    const T* pb = lds + kp * kStride + 12 + col;
    // Software pipeline over two register sets, groups taken in pairs (a
    // pair past the populated rows multiplies zero rows): the operands of
    // the next group are requested from LDS before the two MFMAs of the
    // current group issue, so their latency hides behind the matrix pipe
    // (scheduling barriers keep that order). Every read stays within the
    // 64 rows; the last prefetch re-reads the current pair.
    constexpr int kG = 4 * kStride;  // one group of four rows
    if (populated == uint32_t(kChunk)) {
      // A full chunk (the common case): unrolled, every operand address is
      // the lane's base plus an immediate offset.
      constexpr int kGroups = kChunk / 4;
      T x0[2], z0[2], x1[2], z1[2], bv[2];
      x0[0] = p0x[0];
      z0[0] = p0z[0];
      x1[0] = p1x[0];
      z1[0] = p1z[0];
      bv[0] = pb[0];
#pragma unroll
      for (int gi = 0; gi < kGroups; ++gi) {
        const int cur = gi & 1, nxt = cur ^ 1;
        if (gi + 1 < kGroups) {
          x0[nxt] = p0x[(gi + 1) * kG];
          z0[nxt] = p0z[(gi + 1) * kG];
          x1[nxt] = p1x[(gi + 1) * kG];
          z1[nxt] = p1z[(gi + 1) * kG];
          bv[nxt] = pb[(gi + 1) * kG];
        }
        const T a0 = x0[cur] * z0[cur], a1 = x1[cur] * z1[cur];
        __builtin_amdgcn_sched_barrier(0);
        acc0 = mfma(a0, bv[cur], acc0);
        acc1 = mfma(a1, bv[cur], acc1);
        __builtin_amdgcn_sched_barrier(0);
      }
      return;
    }
    const uint32_t pairs = (groups + 1u) / 2u;
    T xa0 = p0x[0], za0 = p0z[0], xa1 = p1x[0], za1 = p1z[0], ba = pb[0];
    for (uint32_t pi = 0; pi < pairs; ++pi) {
      const T xb0 = p0x[kG], zb0 = p0z[kG], xb1 = p1x[kG], zb1 = p1z[kG], bb = pb[kG];
      {
        const T a0 = xa0 * za0, a1 = xa1 * za1;
        __builtin_amdgcn_sched_barrier(0);
        acc0 = mfma(a0, ba, acc0);
        acc1 = mfma(a1, ba, acc1);
        __builtin_amdgcn_sched_barrier(0);
      }
      const int step = pi + 1 < pairs ? 2 * kG : 0;
      p0x += step;
      p0z += step;
      p1x += step;
      p1z += step;
      pb += step;
      xa0 = p0x[0];
      za0 = p0z[0];
      xa1 = p1x[0];
      za1 = p1z[0];
      ba = pb[0];
      {
        const T a0 = xb0 * zb0, a1 = xb1 * zb1;
        __builtin_amdgcn_sched_barrier(0);
        acc0 = mfma(a0, bb, acc0);
        acc1 = mfma(a1, bb, acc1);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
};

struct MomentsF64 : RowModeBase<double, 26> {
  using Acc = d4;
  using Record = double;  // element type of the stored cell sums
  __device__ static int row(int r, int kp) { return 4 * r + kp; }
  __device__ static void contract(const T* lds, unsigned lane, uint32_t populated, Acc& acc0, Acc& acc1) {
    contract_rows(lds, lane, populated, acc0, acc1,
                  [](T a, T b, Acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); });
  }
};

// FP32 operands, FP32 accumulation within a cell; the cell total is widened
// to FP64 before it is stored and summed across cells.
struct MomentsF32 : RowModeBase<float, 28> {
  using Acc = f4;
  using Record = float;
  __device__ static int row(int r, int kp) { return 4 * kp + r; }
  __device__ static void contract(const T* lds, unsigned lane, uint32_t populated, Acc& acc0, Acc& acc1) {
    contract_rows(lds, lane, populated, acc0, acc1,
                  [](T a, T b, Acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); });
  }
};

// BF16 hi/lo split products with FP32 accumulation (K = 32 particles per
// v_mfma_f32_16x16x32_bf16). The chunk is staged factor-major in FP32:
// F[f][p] (Pxy[0..8], Z2[0..2]) then T[j][p] (12 used of 16 rows), rows
// padded to 68 floats so lanes reading different rows at one column use
// different banks. For each 32-particle block, lane (i, kq) forms its 8
// values of row m = i (and 16 + i) as Pxy[m % 9] * Z2[m / 9] and of column
// j = i of T, splits each x into hi = bf16(x), lo = bf16(x - hi), and issues
// hi*hi + hi*lo + lo*hi per tile (the lo*lo term, 2^-16 relative, dropped).
struct MomentsBf16x3 {
  using T = float;
  using Compute = float;  // operands are split from FP32 values
  using Acc = f4;
  using Record = float;
  static constexpr int kPad = 68;                 // floats per staged row
  static constexpr int kRows = 12 + 16;           // factors, then T columns
  static constexpr int kLdsElems = kRows * kPad;
  // Byte offset of corner slot (comp, m): the rows of T columns 12..15, which
  // the products read only as discarded columns.
  __device__ static constexpr int corner_slot(int comp, int m) { return 24 * kPad * 4 + (comp * 32 + m) * 8; }
  __device__ static int row(int r, int kp) { return 4 * kp + r; }

  __device__ static void stage(T* lds, unsigned lane, const float (&p)[12], const float (&t)[12]) {
#pragma unroll
    for (int f = 0; f < 12; ++f) lds[f * kPad + lane] = p[f];
#pragma unroll
    for (int j = 0; j < 12; ++j) lds[(12 + j) * kPad + lane] = t[j];
  }

  typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

  __device__ static void split(const float (&x)[8], bf8& hi, bf8& lo) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const __bf16 h = __bf16(x[e]);
      hi[e] = h;
      lo[e] = __bf16(x[e] - float(h));
    }
  }

  __device__ static void load8(const T* at, float (&x)[8]) {
    const f4 a = *reinterpret_cast<const f4*>(at);
    const f4 b = *reinterpret_cast<const f4*>(at + 4);
    x[0] = a[0]; x[1] = a[1]; x[2] = a[2]; x[3] = a[3];
    x[4] = b[0]; x[5] = b[1]; x[6] = b[2]; x[7] = b[3];
  }

  __device__ static void contract(const T* lds, unsigned lane, uint32_t populated, Acc& acc0, Acc& acc1) {
    const int i = int(lane & 15u), kq = int(lane >> 4);
    const int m1 = 16 + i;
    // Row 16 + i >= 27 reads the first T rows: finite, and discarded.
    const T* r0x = lds + (i % 9) * kPad;
    const T* r0z = lds + (9 + i / 9) * kPad;
    const T* r1x = lds + (m1 % 9) * kPad;
    const T* r1z = lds + (9 + m1 / 9) * kPad;
    const T* rb = lds + (12 + i) * kPad;
    const uint32_t blocks = (populated + 31u) / 32u;
    for (uint32_t kb = 0; kb < blocks; ++kb) {
      const int o = int(kb) * 32 + 8 * kq;
      bf8 bh, bl;
      {
        float b[8];
        load8(rb + o, b);
        split(b, bh, bl);
      }
      {
        float x[8], z[8], a[8];
        load8(r0x + o, x);
        load8(r0z + o, z);
#pragma unroll
        for (int e = 0; e < 8; ++e) a[e] = x[e] * z[e];
        bf8 ah, al;
        split(a, ah, al);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc0, 0, 0, 0);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc0, 0, 0, 0);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc0, 0, 0, 0);
      }
      {
        float x[8], z[8], a[8];
        load8(r1x + o, x);
        load8(r1z + o, z);
#pragma unroll
        for (int e = 0; e < 8; ++e) a[e] = x[e] * z[e];
        bf8 ah, al;
        split(a, ah, al);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc1, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc1, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc1, 0, 0, 0);
      }
    }
  }
};

// 1 / d for d >= 1: hardware reciprocal and Newton steps (two in FP64, one
// in FP32), accurate to about one unit in the last place.
__device__ __forceinline__ double reciprocal(double d) {
  double inv = __builtin_amdgcn_rcp(d);
  inv = __builtin_fma(inv, __builtin_fma(-d, inv, 1.0), inv);
  return __builtin_fma(inv, __builtin_fma(-d, inv, 1.0), inv);
}
__device__ __forceinline__ float reciprocal(float d) {
  const float inv = __builtin_amdgcn_rcpf(d);
  return __builtin_fmaf(inv, __builtin_fmaf(-d, inv, 1.0f), inv);
}

// One particle (one lane): validate, gather B, form the response and stage
// Pxy[mx + 3 my] = X2[mx] Y2[my], Z2[0..2], T[0..11] in the mode's layout, in
// the mode's evaluation arithmetic C. An absent or rejected particle enters
// with zero fractions, charge and velocity, so its T (and its MFMA
// contribution) is zero; non-finite products (a present particle with a
// non-finite charge or velocity, or a non-finite field) are flagged and zeroed
// on a rarely taken branch.
template <class M>
__device__ __forceinline__ void evaluate(const Raw& r, bool present, int32_t bin,
                                         const typename M::Compute* __restrict__ bc, double beta_over_c_d,
                                         double inverse_volume_d, typename M::T* __restrict__ lds, unsigned lane,
                                         unsigned& flags) {
  using C = typename M::Compute;
  const bool key_ok = r.key == bin;
  const bool frac_ok = (r.rx >= 0.0) & (r.rx < 1.0) & (r.ry >= 0.0) & (r.ry < 1.0) & (r.rz >= 0.0) & (r.rz < 1.0);
  if (present & !key_ok) flags |= kInvalidKey;
  if (present & key_ok & !frac_ok) flags |= kInvalidFraction;
  const bool ok = present & key_ok & frac_ok;
  const C zero = C(0), one = C(1);
  C fx = C(r.rx), fy = C(r.ry), fz = C(r.rz), q = C(r.q), u = C(r.u), v = C(r.v), w = C(r.w);
  // Wave-uniform test: a full chunk of valid particles skips the zeroing.
  if (__builtin_amdgcn_ballot_w64(!ok) != 0) {
    if (!ok) fx = fy = fz = q = u = v = w = zero;
  }
  const C beta_over_c = C(beta_over_c_d), inverse_volume = C(inverse_volume_d);
  const C cx = one - fx, cy = one - fy, cz = one - fz;
  C bx, by, bz;
  {
    const C cxy = cx * cy, fxy = fx * cy, cfy = cx * fy, ffy = fx * fy;
    const C wt[8] = {cxy * cz, fxy * cz, cfy * cz, ffy * cz, cxy * fz, fxy * fz, cfy * fz, ffy * fz};
    bx = wt[0] * bc[0];
    by = wt[0] * bc[8];
    bz = wt[0] * bc[16];
#pragma unroll
    for (int a = 1; a < 8; ++a) {
      bx += wt[a] * bc[a];
      by += wt[a] * bc[8 + a];
      bz += wt[a] * bc[16 + a];
    }
  }
  const C X = beta_over_c * bx, Y = beta_over_c * by, Z = beta_over_c * bz;
  const C xx = X * X, yy = Y * Y, zz = Z * Z;
  const C n2 = xx + yy + zz;
  const C inv = reciprocal(one + n2);
  const C sx = X * inv, syy = Y * inv, szz = Z * inv;
  const C xy = sx * Y, xz = sx * Z, yz = syy * Z;
  C al[9];
  al[0] = xx * inv + inv;
  al[4] = yy * inv + inv;
  al[8] = zz * inv + inv;
  al[1] = xy + szz;
  al[2] = xz - syy;
  al[3] = xy - szz;
  al[5] = yz + sx;
  al[6] = xz + syy;
  al[7] = yz - sx;
  const C scale = beta_over_c * q;
  const C qv = q * inverse_volume;
  C t[12];
#pragma unroll
  for (int e = 0; e < 9; ++e) t[e] = al[e] * scale;
  t[9] = qv * (al[0] * u + al[1] * v + al[2] * w);
  t[10] = qv * (al[3] * u + al[4] * v + al[5] * w);
  t[11] = qv * (al[6] * u + al[7] * v + al[8] * w);
  // Wave-uniform test, so the zeroing below stays a skipped branch.
  if (__builtin_amdgcn_ballot_w64(!__builtin_isfinite(n2 + scale + t[9] + t[10] + t[11])) != 0) {
    const bool finite = __builtin_isfinite(n2) & __builtin_isfinite(scale) & __builtin_isfinite(t[9]) &
                        __builtin_isfinite(t[10]) & __builtin_isfinite(t[11]);
    if (!finite) {
      if (ok) flags |= kInvalidNonfinite;
#pragma unroll
      for (int e = 0; e < 12; ++e) t[e] = zero;
    }
  }
  const C x2[3] = {cx * cx, fx * cx, fx * fx};
  const C y2[3] = {cy * cy, fy * cy, fy * fy};
  C p[12];
#pragma unroll
  for (int my = 0; my < 3; ++my)
#pragma unroll
    for (int mx = 0; mx < 3; ++mx) p[mx + 3 * my] = x2[mx] * y2[my];
  p[9] = cz * cz;
  p[10] = fz * cz;
  p[11] = fz * fz;
  M::stage(lds, lane, p, t);
}

// Store (or add) one cell's moment matrix: lane (j, kp) with j < 9 holds
// mass column j of the rows m of its accumulator elements; the rows of
// element v are affine in v, so the eight addresses are one base plus
// immediate offsets; rows m >= 27 are skipped. The cell's corner currents
// J_a = sum_b Q[a+b][9 + comp]: lanes (9 + comp, kp) put the current columns
// of their rows into the corner slots of the wave's LDS (M::corner_slot,
// space the staged rows leave unused), then lane 3a + comp (a < 8) sums the
// eight moments of its corner in ascending corner b and writes the
// contiguous corner record.
template <class M>
__device__ __forceinline__ void store_cell(typename M::Record* __restrict__ cellbuf, double* __restrict__ corner,
                                           unsigned cell, int col, int kp, const typename M::Acc& acc0,
                                           const typename M::Acc& acc1, int accumulate,
                                           typename M::T* __restrict__ rows, unsigned lane) {
  using Rec = typename M::Record;
  char* const slots = reinterpret_cast<char*>(rows);
  const int comp = col - kMassColumns;  // current component, or negative
  if (comp >= 0 && comp < 3) {
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int m = t < 4 ? M::row(t, kp) : 16 + M::row(t - 4, kp);
      *reinterpret_cast<double*>(slots + M::corner_slot(comp, m)) = double(t < 4 ? acc0[t] : acc1[t - 4]);
    }
  }
  wave_sync();
  if (lane < 24) {
    const int a = int(lane) / 3, c = int(lane) % 3;
    const int ma = (a & 1) + 3 * ((a >> 1) & 1) + 9 * ((a >> 2) & 1);
    double j = 0.0;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int mb = (b & 1) + 3 * ((b >> 1) & 1) + 9 * ((b >> 2) & 1);
      j += *reinterpret_cast<const double*>(slots + M::corner_slot(c, ma + mb));
    }
    double* const cc = corner + size_t(cell) * kCornerRecord + lane;
    if (accumulate) j += *cc;
    *cc = j;
  }
  if (comp >= 0) return;
  Rec* const rec = cellbuf + size_t(cell) * kCellRecord + col;
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const int m = v < 4 ? M::row(v, kp) : 16 + M::row(v - 4, kp);
    if (m < kMoments) {
      Rec value = Rec(v < 4 ? acc0[v] : acc1[v - 4]);
      if (accumulate) value += rec[m * kMassColumns];
      rec[m * kMassColumns] = value;
    }
  }
}

// Species are visited in ascending order in even cells and descending order
// in odd cells, so a species' last chunk of one cell and its first chunk of
// the next are adjacent items for every other cell boundary: the record lines
// the two chunks share are read again while still in L2.
// The first species of cell c with particles (lane t of `count` holds species
// t's count), or `species` when there is none.
__device__ __forceinline__ int first_species(uint32_t count, unsigned c, int species) {
  if (c & 1u) {
    for (int s = species - 1; s >= 0; --s)
      if (pick(count, s) != 0u) return s;
    return species;
  }
  for (int s = 0; s < species; ++s)
    if (pick(count, s) != 0u) return s;
  return species;
}

// The species of cell c with particles visited after species s, or `species`.
__device__ __forceinline__ int species_after(uint32_t count, unsigned c, int s, int species) {
  if (c & 1u) {
    for (--s; s >= 0; --s)
      if (pick(count, s) != 0u) return s;
    return species;
  }
  for (++s; s < species; ++s)
    if (pick(count, s) != 0u) return s;
  return species;
}

// A work item of a wave's walk: chunk `chunk` of species s of cell c (s ==
// species marks a cell without particles, one empty item). The fields are
// wave-uniform.
struct Item {
  unsigned c;
  int s;
  uint32_t chunk;
};

// The item after `it`: the next chunk of the species, a later species of the
// cell, or the first item of the next cell (cell c_end ends the walk).
// `here` is the metadata of it.c, `next` that of it.c + 1.
template <class B>
__device__ __forceinline__ Item successor(const Item& it, const CellMeta<B>& here, const CellMeta<B>& next,
                                          int species, unsigned c_end) {
  const uint32_t count = it.s < species ? count_of(here, it.s) : 0u;
  Item n{it.c, it.s, it.chunk + uint32_t(kChunk)};
  if (n.chunk >= count) {
    n.chunk = 0;
    n.s = it.s < species ? species_after(here.cf, it.c, it.s, species) : species;
    if (n.s >= species) {
      n.c = it.c + 1;
      n.s = n.c < c_end ? first_species(next.cf, n.c, species) : species;
    }
  }
  return n;
}

// This lane's particle of item `it` (metadata `m` of its cell): whether it
// exists and its (sorted) position.
template <class B>
__device__ __forceinline__ bool item_lane(const Item& it, const CellMeta<B>& m, int species, unsigned c_end,
                                          unsigned lane, uint32_t& at) {
  at = 0;
  if (it.c >= c_end || it.s >= species) return false;
  at = first_of(m, it.s) + it.chunk + lane;
  return it.chunk + lane < count_of(m, it.s);
}

// Each wave owns a contiguous range of interior cells (blocks remapped so an
// XCD's waves own neighbouring cells) and walks it as one flat sequence of
// work items (cell, species, 64-particle chunk); a cell without particles is
// one empty item, so its record is still written. Per item every lane
// evaluates one particle into its LDS row; then each group of four particles
// is two FP64 MFMAs: lane (col, kp) supplies A = P[col] and P[16 + col] of
// particle kp (Pxy * Z2, one multiply) and B = T[col], in the operand
// precision of M (FP64 for the plain path). Rows m >= 27 and columns >= 12 of
// the products read neighbouring row data and are discarded.
// The particle records of the next item are loaded before the MFMAs of the
// current one; the metadata and B corners of a cell are loaded two cells
// ahead. With Gather (the sort fused in) each record is loaded from its
// source in the species' sort plan and, once evaluated, stored at its
// destination, so the sorted particles are written by the pass that deposits
// them; the source positions are loaded two items ahead, so the record loads
// of the next item never wait for them. The sums are those of the deposition
// of the sorted particles, bit for bit.
template <class M, bool Gather = false>
__device__ __forceinline__ void deposit_cells(Grid g, const DepositArgs& args, const double* __restrict__ bnode,
                                              double inverse_volume, typename M::Record* __restrict__ cellbuf,
                                              double* __restrict__ corner, int accumulate,
                                              unsigned* __restrict__ validation) {
  using T = typename M::T;
  using Acc = typename M::Acc;
  __shared__ __align__(16) T s_rows[kWaves][M::kLdsElems];
  using Cmp = typename M::Compute;
  __shared__ __align__(16) Cmp s_b[kWaves][24];
  const unsigned lane = threadIdx.x & 63u;
  const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  T* rows = s_rows[wave];
  Cmp* bc = s_b[wave];
  const unsigned cells = g.cells();
  const int col = int(lane & 15u);
  const int kp = int(lane >> 4);
  const int species = args.species;
  const unsigned waves_total = gridDim.x * kWaves;
  const unsigned wid = xcd_logical_block(blockIdx.x, gridDim.x) * kWaves + wave;
  const unsigned per = (cells + waves_total - 1) / waves_total;
  const unsigned c_begin = wid * per;
  const unsigned c_end = c_begin + per < cells ? c_begin + per : cells;
  if (c_begin >= c_end) return;
  unsigned flags = 0;
  // The current item (cell c, metadata m0), the metadata of cells c + 1 and
  // c + 2 (m1, m2; not fetched past the range, where the walk never reads
  // them), and the next item with its lane's position (and source).
  CellMeta<Cmp> m0{}, m1{}, m2{};
  const CellLanes cl = cell_lanes(g, args, lane);
  fetch_cell(g, cl, bnode, c_begin, m0);
  if (c_begin + 1 < c_end) fetch_cell(g, cl, bnode, c_begin + 1, m1);
  if (c_begin + 2 < c_end) fetch_cell(g, cl, bnode, c_begin + 2, m2);
  Item it{c_begin, first_species(m0.cf, c_begin, species), 0};
  Raw raw;
  uint64_t raw_id = 0;
  uint32_t raw_dest = 0;
  if (it.s < species && lane < count_of(m0, it.s))
    fetch_particle<Gather>(args.sp[it.s], first_of(m0, it.s) + lane, raw, raw_id, raw_dest);
  Item it1 = successor(it, m0, m1, species, c_end);
  uint32_t at1 = 0, from1 = 0;
  bool present1 = item_lane(it1, it1.c == it.c ? m0 : m1, species, c_end, lane, at1);
  if (Gather && present1) from1 = args.sp[it1.s].source[at1];
  if (lane < 24) bc[lane] = m0.b;
  wave_sync();
  Acc acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
  // A finished cell's matrix stays in the accumulators and is stored one item
  // later, after that item's particle records have arrived: the vector memory
  // counter is in order, so waiting for loads issued after a store would also
  // wait for the store.
  bool pending = false;
  unsigned pending_cell = 0;
  while (true) {
    const uint32_t count = it.s < species ? count_of(m0, it.s) : 0u;
    const double beta_over_c = it.s < species ? args.sp[it.s].beta_over_c : 0.0;
    const bool same_cell = it1.c == it.c;
    const bool done = it1.c >= c_end;
    // The item after the next one: its cell is it1.c or it1.c + 1, that is
    // c, c + 1 or c + 2.
    Item it2{c_end, species, 0};
    uint32_t at2 = 0, from2 = 0;
    bool present2 = false;
    if (!done) {
      it2 = successor(it1, same_cell ? m0 : m1, same_cell ? m1 : m2, species, c_end);
      const unsigned d2 = it2.c - it.c;
      present2 = item_lane(it2, d2 == 0 ? m0 : (d2 == 1 ? m1 : m2), species, c_end, lane, at2);
      // Requested two items ahead: its latency hides behind this item's
      // evaluation and contraction.
      if (Gather && present2) from2 = args.sp[it2.s].source[at2];
    }
    // When the cell changes after this item, the metadata of the cell that
    // then becomes c + 2 is requested now, so it has arrived when the
    // metadata rotate at the end of the item (copying a value still in flight
    // would wait for every load of the wave).
    CellMeta<Cmp> m3{};
    if (!same_cell && !done && it1.c + 2 < c_end) fetch_cell(g, cl, bnode, it1.c + 2, m3);
    // A gathered record lies in its bin's range by the plan: its key is the bin.
    if (Gather) raw.key = m0.bin;
    evaluate<M>(raw, it.chunk + lane < count, m0.bin, bc, beta_over_c, inverse_volume, rows, lane, flags);
    if (Gather && it.chunk + lane < count) store_particle(args.sp[it.s].dst, raw_dest, raw, raw_id);
    wave_sync();
    if (pending) {
      store_cell<M>(cellbuf, corner, pending_cell, col, kp, acc0, acc1, accumulate, rows, lane);
      acc0 = Acc{0, 0, 0, 0};
      acc1 = Acc{0, 0, 0, 0};
      pending = false;
    }
    if (present1) {
      if (Gather) {
        load_raw<false>(args.sp[it1.s].p, from1, raw);
        raw_id = args.sp[it1.s].p.id[from1];
        raw_dest = at1;
      } else {
        load_raw(args.sp[it1.s].p, at1, raw);
      }
    }
    const uint32_t populated =
        count > it.chunk ? (count - it.chunk < uint32_t(kChunk) ? count - it.chunk : uint32_t(kChunk)) : 0u;
    if (populated) M::contract(rows, lane, populated, acc0, acc1);
    wave_sync();
    if (!same_cell) {
      pending_cell = it.c;
      pending = true;
      if (done) {
        store_cell<M>(cellbuf, corner, pending_cell, col, kp, acc0, acc1, accumulate, rows, lane);
        break;
      }
      m0 = m1;
      m1 = m2;
      m2 = m3;
      if (lane < 24) bc[lane] = m0.b;
      wave_sync();
    }
    it = it1;
    it1 = it2;
    present1 = present2;
    at1 = at2;
    from1 = from2;
  }
  if (flags) atomicOr(validation, flags);
}

// FP64 form: its LDS rows (13 KB per wave) hold it to 3 waves per SIMD.
__global__ __launch_bounds__(kThreads) void k_deposit_fp64(Grid g, DepositArgs args, const double* __restrict__ bnode,
                                                           double inverse_volume, double* __restrict__ cellbuf,
                                                           double* __restrict__ corner, int accumulate,
                                                           unsigned* __restrict__ validation) {
  deposit_cells<MomentsF64>(g, args, bnode, inverse_volume, cellbuf, corner, accumulate, validation);
}

// Mixed forms: 4 waves per SIMD (registers held to 128 per lane), which
// measured 2.6% faster than 3 for the BF16 form.
template <class M>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_deposit_mixed(
    Grid g, DepositArgs args, const double* __restrict__ bnode, double inverse_volume, float* __restrict__ cellbuf,
    double* __restrict__ corner, int accumulate, unsigned* __restrict__ validation) {
  deposit_cells<M>(g, args, bnode, inverse_volume, cellbuf, corner, accumulate, validation);
}

// The sort's gather fused into the deposition (same occupancy as above).
__global__ __launch_bounds__(kThreads) void k_gather_deposit_fp64(Grid g, DepositArgs args,
                                                                  const double* __restrict__ bnode,
                                                                  double inverse_volume, double* __restrict__ cellbuf,
                                                                  double* __restrict__ corner,
                                                                  unsigned* __restrict__ validation) {
  deposit_cells<MomentsF64, true>(g, args, bnode, inverse_volume, cellbuf, corner, 0, validation);
}

template <class M>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_gather_deposit_mixed(
    Grid g, DepositArgs args, const double* __restrict__ bnode, double inverse_volume, float* __restrict__ cellbuf,
    double* __restrict__ corner, unsigned* __restrict__ validation) {
  deposit_cells<M, true>(g, args, bnode, inverse_volume, cellbuf, corner, 0, validation);
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

// One block per (strip of 64 consecutive vertices, stencil s); thread
// (vertex, part) sums entries 3 part .. 3 part + 2 of its vertex's record over
// the stencil's pairs t, reading each pair's moment (m_t, cell) straight from
// the cell mass records (the three parts of a vertex read one moment's nine
// entries together), in a fixed order in FP64, and writes them as rows of the
// entry-major matrix in its storage type (double, or float for the mixed
// forms). The cell records are double, or float in the mixed forms (their
// per-cell sums are FP32, so widening them here gives the same FP64 sums).
constexpr int kStrip = 64;
constexpr unsigned kStripThreads = 3 * kStrip;  // (vertex, part)
// Vertex rows per traversal tile of the assembly (measured at 128^3, FP64:
// tiles of 1, 2, 4, 8 and 16 rows take 2.16, 1.87, 1.75, 1.74 and 1.75 ms
// against 1.89 ms in plain vertex order).
constexpr unsigned kAssembleTileRows = 8;

// The traversal of the padded vertex array: its rows (ey per plane, ex
// vertices each) in tiles of kAssembleTileRows rows (the last tile holds the
// remaining `last` rows); a tile's planes are visited in z order, and each
// tile plane, a contiguous range of vertices, is cut into 64-vertex strips.
struct StripWalk {
  unsigned full, last;          // full tiles, rows of the last tile (0: none)
  unsigned per_full, per_last;  // strips of a full tile plane, of a last-tile plane
};

StripWalk strip_walk(const Grid& g) {
  const unsigned ex = unsigned(g.nx + 3), ey = unsigned(g.ny + 3);
  StripWalk w{};
  w.full = ey / kAssembleTileRows;
  w.last = ey - w.full * kAssembleTileRows;
  w.per_full = (kAssembleTileRows * ex + kStrip - 1) / kStrip;
  w.per_last = (w.last * ex + kStrip - 1) / kStrip;
  return w;
}

template <class RecT, class MassT>
__global__ __launch_bounds__(kStripThreads) void k_assemble_mass(Grid g, StripWalk walk,
                                                                 const RecT* __restrict__ cellbuf,
                                                                 MassT* __restrict__ mass,
                                                                 unsigned* __restrict__ validation) {
  const unsigned ex = unsigned(g.nx + 3), ey = unsigned(g.ny + 3), ez = unsigned(g.nz + 3);
  // Strip-major block order (the 14 stencil blocks of a strip are adjacent
  // logical blocks, on one XCD), so a strip's cell moments are fetched once
  // into that XCD's L2 and reused by all stencils; the tiles' z order keeps
  // the moments of the cells between two vertex planes in L2 until the
  // second plane reads them.
  const unsigned lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int s = int(lb % kStencils);
  unsigned q = lb / kStencils, first, length, u;
  const unsigned full_strips = walk.full * ez * walk.per_full;
  if (q < full_strips) {
    u = q % walk.per_full;
    q /= walk.per_full;
    first = ((q % ez) * ey + (q / ez) * kAssembleTileRows) * ex;
    length = kAssembleTileRows * ex;
  } else {
    q -= full_strips;
    u = q % walk.per_last;
    first = ((q / walk.per_last) * ey + walk.full * kAssembleTileRows) * ex;
    length = walk.last * ex;
  }
  const unsigned lv = threadIdx.x / 3u, part = threadIdx.x % 3u;
  if (u * kStrip + lv >= length) return;
  const unsigned v = first + u * kStrip + lv;
  const int np = c_pairs.count[s];
  const int vi = int(v % ex) - 1, vj = int((v / ex) % ey) - 1, vk = int(v / (ex * ey)) - 1;
  double sum[3] = {0.0, 0.0, 0.0};
  if (!g.walls) {
    const int i = wrap_vertex(vi, g.nx), j = wrap_vertex(vj, g.ny), k = wrap_vertex(vk, g.nz);
    for (int t = 0; t < np; ++t) {
      const int a = c_pairs.a[s][t];
      const int cx = wrap_cell(i - (a & 1), g.nx), cy = wrap_cell(j - ((a >> 1) & 1), g.ny),
                cz = wrap_cell(k - ((a >> 2) & 1), g.nz);
      const RecT* q = cellbuf + record_index(c_pairs.m[s][t], g.cell(cx, cy, cz), int(3 * part));
#pragma unroll
      for (int e = 0; e < 3; ++e) sum[e] += double(q[e]);
    }
  } else {
    // Closed walls: no fold across a wall axis; a vertex beyond a wall and a
    // pair whose cell lies beyond one hold no moment (zero).
    const bool outside = (g.field_wall[0] && (vi < 0 || vi > g.nx)) ||
                         (g.field_wall[1] && (vj < 0 || vj > g.ny)) ||
                         (g.field_wall[2] && (vk < 0 || vk > g.nz));
    const int i = g.field_wall[0] ? vi : wrap_vertex(vi, g.nx);
    const int j = g.field_wall[1] ? vj : wrap_vertex(vj, g.ny);
    const int k = g.field_wall[2] ? vk : wrap_vertex(vk, g.nz);
    for (int t = 0; t < np; ++t) {
      const int a = c_pairs.a[s][t];
      int cx = i - (a & 1), cy = j - ((a >> 1) & 1), cz = k - ((a >> 2) & 1);
      bool none = outside;
      if (g.field_wall[0]) none = none || cx < 0 || cx >= g.nx;
      else cx = wrap_cell(cx, g.nx);
      if (g.field_wall[1]) none = none || cy < 0 || cy >= g.ny;
      else cy = wrap_cell(cy, g.ny);
      if (g.field_wall[2]) none = none || cz < 0 || cz >= g.nz;
      else cz = wrap_cell(cz, g.nz);
      double q[3] = {0.0, 0.0, 0.0};
      if (!none) {
        const RecT* r = cellbuf + record_index(c_pairs.m[s][t], g.cell(cx, cy, cz), int(3 * part));
#pragma unroll
        for (int e = 0; e < 3; ++e) q[e] = double(r[e]);
      }
#pragma unroll
      for (int e = 0; e < 3; ++e) sum[e] += q[e];
    }
  }
  unsigned flags = 0;
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    if (!finite_value(sum[e])) flags |= kInvalidNonfinite;
    mass[size_t(s * 9 + int(3 * part) + e) * g.vstride + v] = MassT(sum[e]);
  }
  if (flags) atomicOr(validation, flags);
}

// Current assembly: per node, 3 components, the corner current of the
// node's corner a in each of the 8 surrounding cells (the cell-major corner
// records [cell][a * 3 + comp] the deposition writes), in a fixed order (sum_shared of "J.add").
__global__ __launch_bounds__(kThreads) void k_assemble_current(Grid g, const double* __restrict__ corner,
                                                               double* __restrict__ current) {
  const unsigned nodes = g.nodes();
  const unsigned sx = unsigned(g.nx + 1), sy = unsigned(g.ny + 1);
  for (unsigned n = blockIdx.x * blockDim.x + threadIdx.x; n < nodes; n += gridDim.x * blockDim.x) {
    const int i = int(n % sx), j = int((n / sx) % sy), k = int(n / (sx * sy));
    double jx = 0.0, jy = 0.0, jz = 0.0;
    if (!g.walls) {
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        const int ax = a & 1, ay = (a >> 1) & 1, az = (a >> 2) & 1;
        const unsigned cell = g.cell(wrap_cell(i - ax, g.nx), wrap_cell(j - ay, g.ny), wrap_cell(k - az, g.nz));
        jx += corner[size_t(cell) * kCornerRecord + a * 3 + 0];
        jy += corner[size_t(cell) * kCornerRecord + a * 3 + 1];
        jz += corner[size_t(cell) * kCornerRecord + a * 3 + 2];
      }
    } else {
      // No cell beyond a wall; then the dual-volume scale f_g
      // (enqueue_scale_moment_dual_volume, moment_field_device.hpp:84-96).
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        const int ax = a & 1, ay = (a >> 1) & 1, az = (a >> 2) & 1;
        int ci = i - ax, cj = j - ay, ck = k - az;
        if (g.field_wall[0] ? (ci < 0 || ci >= g.nx) : false) continue;
        if (g.field_wall[1] ? (cj < 0 || cj >= g.ny) : false) continue;
        if (g.field_wall[2] ? (ck < 0 || ck >= g.nz) : false) continue;
        if (!g.field_wall[0]) ci = wrap_cell(ci, g.nx);
        if (!g.field_wall[1]) cj = wrap_cell(cj, g.ny);
        if (!g.field_wall[2]) ck = wrap_cell(ck, g.nz);
        const unsigned cell = g.cell(ci, cj, ck);
        jx += corner[size_t(cell) * kCornerRecord + a * 3 + 0];
        jy += corner[size_t(cell) * kCornerRecord + a * 3 + 1];
        jz += corner[size_t(cell) * kCornerRecord + a * 3 + 2];
      }
      const double f = double(g.dual_factor(i, j, k));
      jx *= f;
      jy *= f;
      jz *= f;
    }
    current[n] = jx;
    current[nodes + n] = jy;
    current[2 * size_t(nodes) + n] = jz;
  }
}

const int kOffsets[kStencils][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, -1, 0}, {1, 0, 1},
                                    {1, 0, -1}, {0, 1, 1}, {0, 1, -1}, {1, -1, 1}, {1, 1, -1}, {-1, 1, 1}, {1, 1, 1}};

// The constant pair table is uploaded once per device; contexts created on
// several host threads serialise on the mutex.
std::mutex g_pairs_mutex;
bool g_pairs_ready[64] = {};

int ensure_pairs(Context& ctx) {
  std::lock_guard<std::mutex> lock(g_pairs_mutex);
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
  args.counts = ctx.species[size_t(first)].counts;
  args.offsets = ctx.species[size_t(first)].offsets;
  const unsigned cells = ctx.grid.cells();
  // Each wave owns a contiguous range of cells; about 8 waves per SIMD slot.
  const unsigned want = (cells + kWaves - 1) / kWaves;
  const unsigned cap = unsigned(ctx.cus > 0 ? ctx.cus : 256) * 16;
  const unsigned blocks = want < cap ? want : cap;
  phase_begin(ctx, Phase::deposit);
  float* records32 = static_cast<float*>(ctx.cellbuf);
  if (ctx.num.mass_kernel == kMassF32)
    k_deposit_mixed<MomentsF32><<<blocks, kThreads, 0, ctx.stream>>>(ctx.grid, args, ctx.bn, ctx.inverse_volume,
                                                                     records32, ctx.corner_current,
                                                                     accumulate ? 1 : 0, ctx.validation);
  else if (ctx.num.mass_kernel == kMassBf16x3)
    k_deposit_mixed<MomentsBf16x3><<<blocks, kThreads, 0, ctx.stream>>>(ctx.grid, args, ctx.bn, ctx.inverse_volume,
                                                                        records32, ctx.corner_current,
                                                                        accumulate ? 1 : 0, ctx.validation);
  else
    k_deposit_fp64<<<blocks, kThreads, 0, ctx.stream>>>(ctx.grid, args, ctx.bn, ctx.inverse_volume,
                                                        static_cast<double*>(ctx.cellbuf),
                                                        ctx.corner_current, accumulate ? 1 : 0,
                                                        ctx.validation);
  const int rc = check_hip(ctx, hipGetLastError(), "deposit");
  phase_end(ctx, Phase::deposit);
  return rc;
}

int launch_sort_deposit(Context& ctx) {
  for (int s = 0; s < int(ctx.species.size()); ++s)
    if (int rc = launch_sort_plan(ctx, s)) return rc;
  return launch_gather_deposit(ctx);
}

int launch_gather_deposit(Context& ctx) {
  const int ns = int(ctx.species.size());
  if (ns > kMaxSpecies) return fail(ctx, -1, "too many species for one deposition");
  int rc = 0;
  DepositArgs args{};
  args.species = ns;
  for (int s = 0; s < ns; ++s) {
    const SpeciesState& sp = ctx.species[size_t(s)];
    SpeciesDeposit& d = args.sp[s];
    d.p = sp.view();
    d.dst = sp.scratch();
    d.source = sp.source;
    d.counts = sp.counts;
    d.offsets = sp.offsets;
    d.beta_over_c = sp.q_over_m * ctx.num.dt / (2 * ctx.num.c);
  }
  args.counts = ctx.species[0].counts;
  args.offsets = ctx.species[0].offsets;
  const unsigned cells = ctx.grid.cells();
  const unsigned want = (cells + kWaves - 1) / kWaves;
  const unsigned cap = unsigned(ctx.cus > 0 ? ctx.cus : 256) * 16;
  const unsigned blocks = want < cap ? want : cap;
  phase_begin(ctx, Phase::deposit);
  float* records32 = static_cast<float*>(ctx.cellbuf);
  if (ctx.num.mass_kernel == kMassF32)
    k_gather_deposit_mixed<MomentsF32><<<blocks, kThreads, 0, ctx.stream>>>(ctx.grid, args, ctx.bn, ctx.inverse_volume,
                                                                            records32, ctx.corner_current,
                                                                            ctx.validation);
  else if (ctx.num.mass_kernel == kMassBf16x3)
    k_gather_deposit_mixed<MomentsBf16x3><<<blocks, kThreads, 0, ctx.stream>>>(
        ctx.grid, args, ctx.bn, ctx.inverse_volume, records32, ctx.corner_current, ctx.validation);
  else
    k_gather_deposit_fp64<<<blocks, kThreads, 0, ctx.stream>>>(ctx.grid, args, ctx.bn, ctx.inverse_volume,
                                                               static_cast<double*>(ctx.cellbuf), ctx.corner_current,
                                                               ctx.validation);
  rc = check_hip(ctx, hipGetLastError(), "gather and deposit");
  phase_end(ctx, Phase::deposit);
  if (rc) return rc;
  // Every destination of a species holds its record now (a failed plan
  // copied the particles unchanged); the slots swap.
  for (auto& sp : ctx.species) {
    if (sp.count) sp.current ^= 1;
    sp.sorted = true;
  }
  return 0;
}

int launch_assemble(Context& ctx) {
  int rc = ensure_pairs(ctx);
  if (rc) return rc;
  phase_begin(ctx, Phase::assemble);
  const Grid& g = ctx.grid;
  const StripWalk walk = strip_walk(g);
  const unsigned ez = unsigned(g.nz + 3);
  const unsigned mass_blocks = (walk.full * walk.per_full + walk.per_last) * ez * kStencils;
  if (ctx.mass_single)
    k_assemble_mass<<<mass_blocks, kStripThreads, 0, ctx.stream>>>(
        g, walk, static_cast<const float*>(ctx.cellbuf), static_cast<float*>(ctx.mass), ctx.validation);
  else
    k_assemble_mass<<<mass_blocks, kStripThreads, 0, ctx.stream>>>(
        g, walk, static_cast<const double*>(ctx.cellbuf), static_cast<double*>(ctx.mass), ctx.validation);
  if ((rc = check_hip(ctx, hipGetLastError(), "mass assembly"))) return rc;
  k_assemble_current<<<grid_blocks(ctx, g.nodes()), kThreads, 0, ctx.stream>>>(g, ctx.corner_current,
                                                                               ctx.current);
  rc = check_hip(ctx, hipGetLastError(), "current assembly");
  phase_end(ctx, Phase::assemble);
  return rc;
}

}  // namespace tpic_mi
