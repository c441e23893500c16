// Stable counting sort of one species' particles by cell key.
//
// Semantics: the stable counting sort of LPFx's host sort
// (include/lpfx/algorithm/particle/host/sort.hpp:210-305): counts[bin] is the
// key histogram over the (nx+4)(ny+4)(nz+4) halo-2 bins, offsets its exclusive
// scan with offsets[bins] = N, and the particles of a bin keep their input
// order. The reference's device sort
// (include/lpfx/algorithm/particle/device/cuda_hip/sort.hpp:27-260) gives the
// same counts and offsets but an atomic, run-dependent order inside a bin;
// this sort is deterministic and equals the stable order.
//
// MI355X design: after a mover the keys are nearly sorted, so the input is a
// short sequence of runs of equal keys. The sort counts and orders RUNS, not
// particles: (1) run heads per 4096-particle tile, (2) tile scan, (3) run
// table, (4) run histogram with run lengths, (5) bin scans, (6) runs scattered
// to their bins, (7) inside a bin, runs ordered by run id (= input order) give
// each run its destination, (8) one destination-ordered pass (a wave per bin)
// copies the nine SoA attributes, so every output line is written whole and
// only the particles that changed cell are scattered loads. Particle traffic
// is two key passes for the runs and the attribute copy (about 144 B per FP64
// particle). In the cycle, a species sorted before a block move is planned
// from the particles that move changed instead (launch_sort_plan_incremental).
#include "internal.hpp"

namespace tpic_mi {
namespace {

constexpr unsigned kThreads = 256;
constexpr unsigned kItems = 16;
constexpr unsigned kTile = kThreads * kItems;

// Exclusive prefix of a per-thread flag over a 256-thread block, in thread
// order; also returns the block total.
__device__ inline uint32_t scan_flags(bool flag, uint32_t* lds, uint32_t& total) {
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const unsigned long long mask = __ballot(flag);
  const unsigned long long below = lane == 0 ? 0ull : (mask & ((1ull << lane) - 1ull));
  const uint32_t within = uint32_t(__popcll(below));
  if (lane == 0) lds[wave] = uint32_t(__popcll(mask));
  __syncthreads();
  uint32_t before = 0, sum = 0;
#pragma unroll
  for (unsigned w = 0; w < kThreads / 64; ++w) {
    const uint32_t v = lds[w];
    before += w < wave ? v : 0u;
    sum += v;
  }
  __syncthreads();
  total = sum;
  return before + within;
}

__device__ inline uint32_t block_sum(uint32_t v, uint32_t* lds) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  uint32_t sum = 0;
  for (unsigned w = 0; w < blockDim.x / 64; ++w) sum += lds[w];
  __syncthreads();
  return sum;
}

__global__ __launch_bounds__(kThreads) void k_tile_heads(const int32_t* __restrict__ keys, uint32_t n,
                                                         uint32_t bins, uint32_t* __restrict__ tile_runs,
                                                         uint32_t* __restrict__ failure) {
  __shared__ uint32_t lds[kThreads / 64];
  const uint32_t base = blockIdx.x * kTile;
  uint32_t heads = 0;
  bool bad = false;
#pragma unroll 4
  for (unsigned r = 0; r < kItems; ++r) {
    const uint32_t i = base + r * kThreads + threadIdx.x;
    if (i < n) {
      const int32_t key = keys[i];
      bad |= key < 0 || uint32_t(key) >= bins;
      heads += (i == 0 || keys[i - 1] != key) ? 1u : 0u;
    }
  }
  if (bad) atomicOr(failure, 1u);
  const uint32_t total = block_sum(heads, lds);
  if (threadIdx.x == 0) tile_runs[blockIdx.x] = total;
}

// Each wave of a tile takes a contiguous span of kWaveSpan keys, 64 at a
// time, and keeps the head flags of its span as ballot words.
constexpr unsigned kWaveSpan = kTile / (kThreads / 64);
constexpr unsigned kWaveItems = kWaveSpan / 64;

__device__ inline uint64_t lanes_below(unsigned lane) { return lane == 0 ? 0ull : ((1ull << lane) - 1ull); }

// Set bits of the words (lo, hi) below this lane (v_mbcnt_lo/hi).
__device__ __forceinline__ uint32_t bits_below(uint32_t lo, uint32_t hi) {
  return __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
}

// Waits for every vector memory load of the wave (s_waitcnt vmcnt(0)): the
// stores that follow, each under its own lane mask, then need no further waits
// on the loaded values.
__device__ __forceinline__ void wait_loads() { __builtin_amdgcn_s_waitcnt(0x0F70); }


// Run table and histogram in one pass over the keys, after the scan of the
// tile head counts: run_start[r] and run_key[r] for every head in input
// order; counts[key] gains each run's length and rcounts[key] one per run,
// as integer atomics per run segment (a run continued from an earlier wave
// span adds its leading segment there), so the sums do not depend on their
// order. wave_base[] keeps the number of heads before each wave span for
// k_tile_sources; the tile holding the last key closes the run table.
__global__ __launch_bounds__(kThreads) void k_tile_runs(const int32_t* __restrict__ keys, uint32_t n,
                                                        uint32_t tiles, const uint32_t* __restrict__ tile_base,
                                                        uint32_t* __restrict__ run_start,
                                                        int32_t* __restrict__ run_key,
                                                        uint32_t* __restrict__ counts,
                                                        uint32_t* __restrict__ rcounts,
                                                        uint32_t* __restrict__ wave_base,
                                                        const uint32_t* __restrict__ failure) {
  if (*failure) return;
  __shared__ uint32_t lds[kThreads / 64];
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t sub = blockIdx.x * kTile + wave * kWaveSpan;
  uint64_t heads[kWaveItems];
  int32_t key[kWaveItems];
  uint32_t count = 0;
#pragma unroll
  for (unsigned r = 0; r < kWaveItems; ++r) {
    const uint32_t i = sub + r * 64 + lane;
    bool head = false;
    key[r] = 0;
    if (i < n) {
      key[r] = keys[i];
      head = i == 0 || keys[i - 1] != key[r];
    }
    heads[r] = __ballot(head);
    count += uint32_t(__popcll(heads[r]));
  }
  if (lane == 0) lds[wave] = count;
  __syncthreads();
  uint32_t before = tile_base[blockIdx.x];
  uint32_t tile_total = before;
#pragma unroll
  for (unsigned w = 0; w < kThreads / 64; ++w) {
    before += w < wave ? lds[w] : 0u;
    tile_total += lds[w];
  }
  if (lane == 0) wave_base[blockIdx.x * (kThreads / 64) + wave] = before;
  if (blockIdx.x == tiles - 1 && threadIdx.x == 0) run_start[tile_total] = n;
  const uint32_t end = sub + kWaveSpan < n ? sub + kWaveSpan : n;
  uint32_t rank = before;
#pragma unroll
  for (unsigned r = 0; r < kWaveItems; ++r) {
    const uint64_t m = heads[r];
    if ((m >> lane) & 1ull) {
      const uint32_t i = sub + r * 64 + lane;
      const uint32_t id = rank + uint32_t(__popcll(m & lanes_below(lane)));
      run_start[id] = i;
      run_key[id] = key[r];
      // The run's segment in this span ends at the span's next head.
      uint32_t next = end;
      const uint64_t above = lane == 63 ? 0ull : (m >> (lane + 1));
      if (above) {
        next = i + 1u + uint32_t(__builtin_ctzll(above));
      } else {
#pragma unroll
        for (unsigned r2 = r + 1; r2 < kWaveItems; ++r2)
          if (next == end && heads[r2]) next = sub + r2 * 64 + uint32_t(__builtin_ctzll(heads[r2]));
      }
      atomicAdd(counts + key[r], next - i);
      atomicAdd(rcounts + key[r], 1u);
    }
    rank += uint32_t(__popcll(m));
  }
  // The span's keys before its first head continue a run of an earlier span.
  if (lane == 0 && sub < n && !(heads[0] & 1ull)) {
    uint32_t first = end;
#pragma unroll
    for (unsigned r2 = 0; r2 < kWaveItems; ++r2)
      if (first == end && heads[r2]) first = sub + r2 * 64 + uint32_t(__builtin_ctzll(heads[r2]));
    atomicAdd(counts + key[0], first - sub);
  }
}

// The plan alone, in source order: particle i of run r lands at
// run_dest[r] + (i - run_start[r]), so source[that] = i (and, when recorded,
// permutation[i] = that). A particle's run is the last head at or before it:
// its rank among the heads, counted from the wave span's base.
__global__ __launch_bounds__(kThreads) void k_tile_sources(const int32_t* __restrict__ keys, uint32_t n,
                                                           const uint32_t* __restrict__ wave_base,
                                                           const uint32_t* __restrict__ run_start,
                                                           const uint32_t* __restrict__ run_dest,
                                                           uint32_t* __restrict__ source,
                                                           uint32_t* __restrict__ permutation,
                                                           const uint32_t* __restrict__ failure) {
  if (*failure) return;
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t sub = blockIdx.x * kTile + wave * kWaveSpan;
  if (sub >= n) return;
  // All keys of the span first (the key before each one comes from the lane
  // below, or the item before), then every run's table entries, then the
  // stores: the loads of a phase are in flight together.
  int32_t key[kWaveItems];
#pragma unroll
  for (unsigned r = 0; r < kWaveItems; ++r) {
    const uint32_t i = sub + r * 64 + lane;
    key[r] = i < n ? keys[i] : -1;
  }
  const int32_t before_span = (lane == 0 && sub > 0) ? keys[sub - 1] : -1;
  uint32_t before = wave_base[blockIdx.x * (kThreads / 64) + wave];
  const uint64_t upto = lanes_below(lane) | (1ull << lane);
  uint32_t run[kWaveItems];
#pragma unroll
  for (unsigned r = 0; r < kWaveItems; ++r) {
    const uint32_t i = sub + r * 64 + lane;
    const int32_t below = __shfl_up(key[r], 1u, 64);
    const int32_t carried = __shfl(r == 0 ? before_span : key[r > 0 ? r - 1 : 0], 63, 64);
    const int32_t prev = lane == 0 ? (r == 0 ? before_span : carried) : below;
    const bool head = i < n && (i == 0 || prev != key[r]);
    const uint64_t m = __ballot(head);
    run[r] = before + uint32_t(__popcll(m & upto)) - 1u;
    before += uint32_t(__popcll(m));
  }
  uint32_t dest[kWaveItems], start[kWaveItems];
#pragma unroll
  for (unsigned r = 0; r < kWaveItems; ++r) {
    const uint32_t i = sub + r * 64 + lane;
    dest[r] = i < n ? run_dest[run[r]] : 0u;
    start[r] = i < n ? run_start[run[r]] : 0u;
  }
  wait_loads();
#pragma unroll
  for (unsigned r = 0; r < kWaveItems; ++r) {
    const uint32_t i = sub + r * 64 + lane;
    if (i < n) {
      const uint32_t d = dest[r] + (i - start[r]);
      source[d] = i;
      if (permutation) permutation[i] = d;
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_run_scatter(const uint32_t* __restrict__ tile_base, uint32_t tiles,
                                                          const int32_t* __restrict__ run_key,
                                                          const uint32_t* __restrict__ roffsets,
                                                          uint32_t* __restrict__ cursor,
                                                          uint32_t* __restrict__ bin_runs,
                                                          const uint32_t* __restrict__ failure) {
  if (*failure) return;
  const uint32_t runs = tile_base[tiles];
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < runs; r += gridDim.x * blockDim.x) {
    const int32_t key = run_key[r];
    bin_runs[roffsets[key] + atomicAdd(cursor + key, 1u)] = r;
  }
}

// Destination of every run: the bin offset plus the lengths of the bin's runs
// with a smaller run id (earlier in the input). Bins with at most kSmall runs
// are done by one thread; others are listed for k_bin_large.
constexpr uint32_t kSmall = 24;

__global__ __launch_bounds__(kThreads) void k_bin_small(Grid g, const uint32_t* __restrict__ rcounts,
                                                        const uint32_t* __restrict__ roffsets,
                                                        const uint32_t* __restrict__ offsets,
                                                        const uint32_t* __restrict__ bin_runs,
                                                        const uint32_t* __restrict__ run_start,
                                                        uint32_t* __restrict__ run_dest,
                                                        uint32_t* __restrict__ large,
                                                        unsigned* __restrict__ validation,
                                                        const uint32_t* __restrict__ failure) {
  if (*failure) return;
  const uint32_t bins = g.bins();
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < bins; b += gridDim.x * blockDim.x) {
    const uint32_t nr = rcounts[b];
    if (nr == 0) continue;
    int i, j, k;
    g.decode(int32_t(b), i, j, k);
    if (!g.interior(i, j, k)) atomicOr(validation, kInvalidHalo);
    const uint32_t first = roffsets[b];
    if (nr > kSmall) {
      large[1 + atomicAdd(large, 1u)] = b;
      continue;
    }
    uint32_t id[kSmall], len[kSmall];
    for (uint32_t t = 0; t < nr; ++t) {
      id[t] = bin_runs[first + t];
      len[t] = run_start[id[t] + 1] - run_start[id[t]];
    }
    const uint32_t base = offsets[b];
    for (uint32_t t = 0; t < nr; ++t) {
      uint32_t d = base;
      for (uint32_t f = 0; f < nr; ++f) d += id[f] < id[t] ? len[f] : 0u;
      run_dest[id[t]] = d;
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_bin_large(const uint32_t* __restrict__ large,
                                                        const uint32_t* __restrict__ rcounts,
                                                        const uint32_t* __restrict__ roffsets,
                                                        const uint32_t* __restrict__ offsets,
                                                        const uint32_t* __restrict__ bin_runs,
                                                        const uint32_t* __restrict__ run_start,
                                                        uint32_t* __restrict__ run_dest,
                                                        const uint32_t* __restrict__ failure) {
  if (*failure) return;
  __shared__ uint32_t s_id[kThreads], s_len[kThreads];
  const uint32_t listed = large[0];
  for (uint32_t l = blockIdx.x; l < listed; l += gridDim.x) {
    const uint32_t b = large[1 + l];
    const uint32_t nr = rcounts[b], first = roffsets[b], base = offsets[b];
    for (uint32_t e0 = 0; e0 < nr; e0 += kThreads) {
      const uint32_t e = e0 + threadIdx.x;
      const uint32_t mine = e < nr ? bin_runs[first + e] : 0u;
      uint32_t d = base;
      for (uint32_t f0 = 0; f0 < nr; f0 += kThreads) {
        __syncthreads();
        const uint32_t f = f0 + threadIdx.x;
        if (f < nr) {
          const uint32_t r = bin_runs[first + f];
          s_id[threadIdx.x] = r;
          s_len[threadIdx.x] = run_start[r + 1] - run_start[r];
        }
        __syncthreads();
        const uint32_t m = nr - f0 < kThreads ? nr - f0 : kThreads;
        for (uint32_t t = 0; t < m; ++t) d += s_id[t] < mine ? s_len[t] : 0u;
      }
      if (e < nr) run_dest[mine] = d;
    }
    __syncthreads();
  }
}

// The nine attributes of one particle (68 B).
struct Record {
  double rx, ry, rz, u, v, w, q;
  int32_t cell;
  uint64_t id;
};

__device__ __forceinline__ void load_record(const ParticleView& p, uint32_t i, Record& r) {
  r.rx = p.rx[i];
  r.ry = p.ry[i];
  r.rz = p.rz[i];
  r.u = p.u[i];
  r.v = p.v[i];
  r.w = p.w[i];
  r.q = p.q[i];
  r.cell = p.cell[i];
  r.id = p.id[i];
}

__device__ __forceinline__ void store_record(const ParticleView& p, uint32_t i, const Record& r) {
  p.rx[i] = r.rx;
  p.ry[i] = r.ry;
  p.rz[i] = r.rz;
  p.u[i] = r.u;
  p.v[i] = r.v;
  p.w[i] = r.w;
  p.q[i] = r.q;
  p.cell[i] = r.cell;
  p.id[i] = r.id;
}

// Destination-ordered copy: one wave per bin writes the bin's destination
// range [offsets[b], offsets[b] + count) in order, each lane taking the next
// particle from the run that covers its destination (run r covers
// [run_dest[r], run_dest[r] + length) and starts at source run_start[r]).
// Stores are contiguous per bin, so every line of the output is written whole
// by one or two waves close in time; the particles that changed cell are the
// only scattered accesses, and they are loads. Bins with more than 64 runs
// take the general loop.
__global__ __launch_bounds__(kThreads) void k_gather_bins(Grid g, ParticleView src, ParticleView dst, RunPlan plan,
                                                          uint32_t* __restrict__ permutation) {
  if (*plan.failure) return;
  const uint32_t bins = g.bins();
  const unsigned lane = threadIdx.x & 63u;
  const uint32_t waves = gridDim.x * (kThreads / 64);
  // XCD-aware: an XCD's waves take neighbouring bins, so the output lines two
  // adjacent bins share are completed in one L2.
  const uint32_t first_bin = xcd_logical_block(blockIdx.x, gridDim.x) * (kThreads / 64) + (threadIdx.x >> 6);
  for (uint32_t b = __builtin_amdgcn_readfirstlane(first_bin); b < bins; b += waves) {
    if (plan.rcounts[b] == 0) continue;
    BinRuns br;
    load_bin_runs(plan, b, lane, br);
    const uint32_t count = plan.counts[b], base = plan.offsets[b];
    for (uint32_t i = lane; i < count; i += 64) {
      const uint32_t d = base + i;
      const uint32_t from = run_source(plan, br, d);
      Record rec;
      load_record(src, from, rec);
      store_record(dst, d, rec);
      if (permutation) permutation[from] = d;
    }
  }
}

// The records of the halo bins of a plan (particles left outside the
// interior, flagged by k_bin_small): the fused gather and deposition visits
// interior cells only, so these are copied here, one thread per bin.
__global__ __launch_bounds__(kThreads) void k_copy_halo_bins(Grid g, ParticleView src, ParticleView dst,
                                                             const uint32_t* __restrict__ counts,
                                                             const uint32_t* __restrict__ offsets,
                                                             const uint32_t* __restrict__ source,
                                                             const uint32_t* __restrict__ failure) {
  if (*failure) return;
  const uint32_t bins = g.bins();
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < bins; b += gridDim.x * blockDim.x) {
    const uint32_t count = counts[b];
    if (count == 0) continue;
    int i, j, k;
    g.decode(int32_t(b), i, j, k);
    if (g.interior(i, j, k)) continue;
    for (uint32_t d = offsets[b]; d < offsets[b] + count; ++d) {
      Record rec;
      load_record(src, source[d], rec);
      store_record(dst, d, rec);
    }
  }
}

// Identity copy after a failed plan (invalid keys): particles keep their
// order (and, when given, every destination is its own source).
__global__ __launch_bounds__(kThreads) void k_copy_on_failure(ParticleView src, ParticleView dst, uint32_t n,
                                                              uint32_t* __restrict__ permutation,
                                                              const uint32_t* __restrict__ failure,
                                                              uint32_t* __restrict__ source) {
  if (*failure == 0) return;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    Record rec;
    load_record(src, i, rec);
    store_record(dst, i, rec);
    if (permutation) permutation[i] = i;
    if (source) source[i] = i;
  }
}

__global__ void k_sort_failed(const uint32_t* failure, unsigned* validation) {
  if (*failure) atomicOr(validation, kInvalidSort);
}

// ---- device-wide exclusive scan of uint32 (out has n + 1 entries) --------
constexpr unsigned kScanItems = 8;
constexpr unsigned kScanBlock = kThreads * kScanItems;

__global__ __launch_bounds__(kThreads) void k_scan_reduce(const uint32_t* __restrict__ in, uint32_t n,
                                                          uint32_t* __restrict__ sums) {
  __shared__ uint32_t lds[kThreads / 64];
  const uint32_t base = blockIdx.x * kScanBlock;
  uint32_t s = 0;
#pragma unroll
  for (unsigned r = 0; r < kScanItems; ++r) {
    const uint32_t i = base + r * kThreads + threadIdx.x;
    s += i < n ? in[i] : 0u;
  }
  const uint32_t total = block_sum(s, lds);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive scan of one block of kScanBlock items with an added offset;
// writes out[n] (the total) from the block holding the last item.
__global__ __launch_bounds__(kThreads) void k_scan_apply(const uint32_t* __restrict__ in, uint32_t n,
                                                         const uint32_t* __restrict__ block_offset,
                                                         uint32_t* __restrict__ out) {
  __shared__ uint32_t lds[kThreads / 64];
  const uint32_t base = blockIdx.x * kScanBlock;
  // Blocked arrangement: thread t owns items t*8 .. t*8+7 of the block.
  uint32_t v[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (unsigned r = 0; r < kScanItems; ++r) {
    const uint32_t i = base + threadIdx.x * kScanItems + r;
    v[r] = i < n ? in[i] : 0u;
    s += v[r];
  }
  // Exclusive scan of per-thread sums across the block.
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t inc = s;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(inc, off, 64);
    if (int(lane) >= off) inc += t;
  }
  if (lane == 63) lds[wave] = inc;
  __syncthreads();
  uint32_t before = 0, total = 0;
  for (unsigned w = 0; w < kThreads / 64; ++w) {
    before += w < wave ? lds[w] : 0u;
    total += lds[w];
  }
  uint32_t running = (block_offset ? block_offset[blockIdx.x] : 0u) + before + inc - s;
#pragma unroll
  for (unsigned r = 0; r < kScanItems; ++r) {
    const uint32_t i = base + threadIdx.x * kScanItems + r;
    if (i < n) out[i] = running;
    running += v[r];
  }
  const uint32_t last_block = n == 0 ? 0 : (n - 1) / kScanBlock;
  if (blockIdx.x == last_block && threadIdx.x == 0)
    out[n] = (block_offset ? block_offset[blockIdx.x] : 0u) + total;
}

// ---- The incremental plan: a sorted species after the block mover ----
//
// The block mover records, per aligned 64-particle block, the lanes whose key
// it changed (crossers). Every other particle (a stayer) still lies in its
// key's bin range of the previous plan, in input order. The stable order of a
// bin is therefore: its arrivals (crossers with the bin as new key) from
// earlier positions (lower bins), in position order; its stayers, in position
// order; its arrivals from later positions. Every stayer's destination is its
// bin's base plus its position less the crossers before it; each bin orders
// its few arrivals by position. A bin with more than kArrivalsMax arrivals,
// or an old state in which the stayers and crossers do not add up to the
// particles (a previous plan that failed), sets the species' plan status: the
// run-based plan is then run instead.
constexpr uint32_t kArrivalsMax = 64;
constexpr uint32_t kArrivalsInRegisters = 16;

__device__ __forceinline__ uint32_t crossers_before(const uint64_t* __restrict__ cross,
                                                    const uint32_t* __restrict__ block_base, uint32_t blocks,
                                                    uint32_t p) {
  const uint32_t blk = p >> 6, bit = p & 63u;
  if (blk >= blocks) return block_base[blocks];
  const uint32_t below = bit ? uint32_t(__popcll(cross[blk] & ((1ull << bit) - 1ull))) : 0u;
  return block_base[blk] + below;
}

__global__ __launch_bounds__(kThreads) void k_cross_counts(const uint64_t* __restrict__ cross, uint32_t blocks,
                                                           uint32_t* __restrict__ out) {
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < blocks; b += gridDim.x * blockDim.x)
    out[b] = uint32_t(__popcll(cross[b]));
}

// Per crosser one arrival in its new bin (integer atomics: crosser t of the
// position-ordered crosser list keeps the arrival's index j within the bin in
// slot[t]; the bins order their arrivals by position afterwards), counted as
// from a lower bin when its position precedes the bin's old range. Tiles of
// kTile particles: the crosser words of a wave span and every crosser's key
// are loaded first, then the atomics and range starts, then the stores.
__global__ __launch_bounds__(kThreads) void k_cross_arrivals(const uint64_t* __restrict__ cross, uint32_t n,
                                                             const uint32_t* __restrict__ block_base,
                                                             const int32_t* __restrict__ keys, uint32_t bins,
                                                             const uint32_t* __restrict__ old_offsets,
                                                             uint32_t* __restrict__ arrivals,
                                                             uint32_t* __restrict__ lower,
                                                             uint32_t* __restrict__ slot,
                                                             uint32_t* __restrict__ status) {
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t sub = blockIdx.x * kTile + wave * kWaveSpan;
  if (sub >= n) return;
  const uint32_t blk0 = sub >> 6;  // sub is a multiple of 64
  const bool item_lane = lane < kWaveItems && sub + lane * 64u < n;
  const uint64_t words = item_lane ? cross[blk0 + lane] : 0ull;
  const uint32_t bases = item_lane ? block_base[blk0 + lane] : 0u;
  uint32_t crossers = 0;  // bit r: the lane's particle of item r is a crosser
  uint32_t list[kWaveItems];  // its index in the crosser list
  int32_t key[kWaveItems];
#pragma unroll
  for (unsigned r = 0; r < kWaveItems; ++r) {
    const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(words)), int(r)));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(words >> 32)), int(r)));
    const bool crosser = ((lane < 32u ? lo : hi) >> (lane & 31u)) & 1u;
    crossers |= uint32_t(crosser) << r;
    list[r] = uint32_t(__builtin_amdgcn_readlane(int(bases), int(r))) + bits_below(lo, hi);
    key[r] = crosser ? keys[sub + r * 64 + lane] : -1;
  }
  if (crossers == 0) return;
  wait_loads();
  uint32_t j[kWaveItems], start[kWaveItems];
#pragma unroll
  for (unsigned r = 0; r < kWaveItems; ++r) {
    const int32_t k = key[r];
    const bool ok = ((crossers >> r) & 1u) && k >= 0 && uint32_t(k) < bins;
    j[r] = ok ? atomicAdd(arrivals + k, 1u) : 0u;
    start[r] = ok ? old_offsets[k] : 0u;
  }
  wait_loads();
#pragma unroll
  for (unsigned r = 0; r < kWaveItems; ++r) {
    if (!((crossers >> r) & 1u)) continue;
    const int32_t k = key[r];
    if (k < 0 || uint32_t(k) >= bins) {
      atomicOr(status, 4u);
      continue;
    }
    slot[list[r]] = j[r];
    if (sub + r * 64 + lane < start[r]) atomicAdd(lower + k, 1u);
  }
}

// Per bin: new count (old count less the bin's crossers plus its arrivals)
// and `shift`, the crossers before the old range less the range's start (the
// stayer base less the new offset and the arrivals from lower bins). Reads
// the previous plan's offsets, before the scan of the new counts replaces
// them.
__global__ __launch_bounds__(kThreads) void k_cross_counts_new(uint32_t bins, const uint32_t* __restrict__ old_offsets,
                                                               const uint64_t* __restrict__ cross,
                                                               const uint32_t* __restrict__ block_base,
                                                               uint32_t blocks,
                                                               const uint32_t* __restrict__ arrivals,
                                                               uint32_t* __restrict__ counts,
                                                               uint32_t* __restrict__ shift,
                                                               uint32_t* __restrict__ status) {
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < bins; b += gridDim.x * blockDim.x) {
    const uint32_t o = old_offsets[b], e = old_offsets[b + 1];
    const uint32_t lo = crossers_before(cross, block_base, blocks, o);
    const uint32_t hi = crossers_before(cross, block_base, blocks, e);
    const uint32_t arrive = arrivals[b];
    if (arrive > kArrivalsMax) atomicOr(status, 1u);
    counts[b] = (e - o) - (hi - lo) + arrive;
    shift[b] = lo - o;
  }
}

// Per bin: the stayer base (stayer p of the bin lands at base + p -
// crossers_before(p)), after the scans. With a consistent old state (the new
// counts add up to the particles) a non-empty halo bin is flagged as by the
// run-based plan.
__global__ __launch_bounds__(kThreads) void k_cross_base(Grid g, uint32_t n, const uint32_t* __restrict__ counts,
                                                         const uint32_t* __restrict__ offsets,
                                                         const uint32_t* __restrict__ shift,
                                                         const uint32_t* __restrict__ lower,
                                                         uint32_t* __restrict__ bin_base,
                                                         uint32_t* __restrict__ status,
                                                         unsigned* __restrict__ validation) {
  const uint32_t bins = g.bins();
  const bool consistent = offsets[bins] == n;
  if (blockIdx.x == 0 && threadIdx.x == 0 && !consistent) atomicOr(status, 2u);
  if (!consistent) return;
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < bins; b += gridDim.x * blockDim.x) {
    if (counts[b] != 0) {
      int i, j, k;
      g.decode(int32_t(b), i, j, k);
      if (!g.interior(i, j, k)) atomicOr(validation, kInvalidHalo);
    }
    bin_base[b] = offsets[b] + lower[b] + shift[b];
  }
}

// Source-ordered pass over tiles of kTile particles: a stayer writes its own
// source entry; a crosser is listed among its new bin's arrivals (at its
// index j within the bin).
__global__ __launch_bounds__(kThreads) void k_cross_sources(const int32_t* __restrict__ keys, uint32_t n, uint32_t bins,
                                                            const uint64_t* __restrict__ cross,
                                                            const uint32_t* __restrict__ block_base,
                                                            const uint32_t* __restrict__ bin_base,
                                                            const uint32_t* __restrict__ arr_off,
                                                            const uint32_t* __restrict__ slot,
                                                            uint32_t* __restrict__ bin_arrivals,
                                                            const uint32_t* __restrict__ status,
                                                            uint32_t* __restrict__ source) {
  if (*status) return;
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t sub = blockIdx.x * kTile + wave * kWaveSpan;
  if (sub >= n) return;
  // Lane r < kWaveItems holds the crosser word and crossers-before of item r.
  const uint32_t blk0 = sub >> 6;  // sub is a multiple of 64
  const bool item_lane = lane < kWaveItems && sub + lane * 64u < n;
  const uint64_t words = item_lane ? cross[blk0 + lane] : 0ull;
  const uint32_t bases = item_lane ? block_base[blk0 + lane] : 0u;
  int32_t key[kWaveItems];
#pragma unroll
  for (unsigned r = 0; r < kWaveItems; ++r) {
    const uint32_t i = sub + r * 64 + lane;
    key[r] = i < n ? keys[i] : -1;
  }
  // Every item's crosser bit and crossers-before first (lane reads only), so
  // the gathers below are issued back to back.
  uint32_t crossers = 0;  // bit r: the lane's particle of item r is a crosser
  uint32_t before[kWaveItems];
#pragma unroll
  for (unsigned r = 0; r < kWaveItems; ++r) {
    const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(words)), int(r)));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(words >> 32)), int(r)));
    crossers |= (((lane < 32u ? lo : hi) >> (lane & 31u)) & 1u) << r;
    before[r] = uint32_t(__builtin_amdgcn_readlane(int(bases), int(r))) + bits_below(lo, hi);
  }
  // A crosser's arrival slot (its bin's arrival offset plus its index j), or
  // a stayer's bin base less the crossers before it.
  uint32_t target[kWaveItems];
#pragma unroll
  for (unsigned r = 0; r < kWaveItems; ++r) {
    const bool valid = key[r] >= 0 && uint32_t(key[r]) < bins;
    const bool crosser = (crossers >> r) & 1u;
    const uint32_t* table = crosser ? arr_off : bin_base;
    const uint32_t base_or_offset = valid ? table[key[r]] : 0u;
    const uint32_t j = (valid && crosser) ? slot[before[r]] : 0u;
    target[r] = crosser ? base_or_offset + j : base_or_offset - before[r];
  }
  wait_loads();
#pragma unroll
  for (unsigned r = 0; r < kWaveItems; ++r) {
    const uint32_t i = sub + r * 64 + lane;
    if (key[r] < 0 || uint32_t(key[r]) >= bins) continue;
    if ((crossers >> r) & 1u) {
      bin_arrivals[target[r]] = i;
    } else {
      const uint32_t dest = target[r] + i;
      if (dest < n) source[dest] = i;
    }
  }
}

// Per bin: the arrivals in position order (rank = arrivals at smaller
// positions; positions are distinct) at their destinations: before the
// stayers those from lower bins, after them the others.
__global__ __launch_bounds__(kThreads) void k_cross_bins(uint32_t bins, uint32_t n,
                                                         const uint32_t* __restrict__ counts,
                                                         const uint32_t* __restrict__ offsets,
                                                         const uint32_t* __restrict__ arrivals,
                                                         const uint32_t* __restrict__ arr_off,
                                                         const uint32_t* __restrict__ bin_arrivals,
                                                         const uint32_t* __restrict__ lower,
                                                         const uint32_t* __restrict__ status,
                                                         uint32_t* __restrict__ source) {
  if (*status) return;
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < bins; b += gridDim.x * blockDim.x) {
    const uint32_t k = arrivals[b];
    if (k == 0 || k > kArrivalsMax) continue;
    const uint32_t a0 = arr_off[b], d0 = offsets[b], below = lower[b], stay = counts[b] - k;
    if (k <= kArrivalsInRegisters) {
      uint32_t pos[kArrivalsInRegisters];
#pragma unroll
      for (uint32_t r = 0; r < kArrivalsInRegisters; ++r) pos[r] = r < k ? bin_arrivals[a0 + r] : 0xffffffffu;
#pragma unroll
      for (uint32_t r = 0; r < kArrivalsInRegisters; ++r) {
        if (r >= k) break;
        uint32_t rank = 0;
#pragma unroll
        for (uint32_t q = 0; q < kArrivalsInRegisters; ++q) rank += pos[q] < pos[r] ? 1u : 0u;
        const uint32_t dest = d0 + rank + (rank >= below ? stay : 0u);
        if (dest < n) source[dest] = pos[r];
      }
    } else {
      for (uint32_t r = 0; r < k; ++r) {
        const uint32_t p = bin_arrivals[a0 + r];
        uint32_t rank = 0;
        for (uint32_t q = 0; q < k; ++q) rank += bin_arrivals[a0 + q] < p ? 1u : 0u;
        const uint32_t dest = d0 + rank + (rank >= below ? stay : 0u);
        if (dest < n) source[dest] = p;
      }
    }
  }
}

}  // namespace

// Device-wide exclusive scan on stream st with the scratch of workspace ws.
static int exclusive_scan(Context& ctx, SortWorkspace& ws, hipStream_t st, const uint32_t* in, uint32_t* out,
                          uint32_t n) {
  if (n == 0) {
    return check_hip(ctx, hipMemsetAsync(out, 0, sizeof(uint32_t), st), "scan");
  }
  const uint32_t blocks = (n + kScanBlock - 1) / kScanBlock;
  if (blocks == 1) {
    k_scan_apply<<<1, kThreads, 0, st>>>(in, n, nullptr, out);
    return check_hip(ctx, hipGetLastError(), "scan");
  }
  // Level scratch: block sums [blocks], their exclusive scan [blocks + 1].
  uint32_t* sums = ws.scan_scratch;
  uint32_t* scanned = sums + blocks;
  if (size_t(2 * blocks + 1) > ws.scan_scratch_words)
    return fail(ctx, -3, "scan scratch too small");
  k_scan_reduce<<<blocks, kThreads, 0, st>>>(in, n, sums);
  // Recurse on the block sums with the remaining scratch.
  {
    uint32_t* saved = ws.scan_scratch;
    const size_t saved_words = ws.scan_scratch_words;
    ws.scan_scratch = scanned + blocks + 1;
    ws.scan_scratch_words = saved_words - (2 * size_t(blocks) + 1);
    const int rc = exclusive_scan(ctx, ws, st, sums, scanned, blocks);
    ws.scan_scratch = saved;
    ws.scan_scratch_words = saved_words;
    if (rc) return rc;
  }
  k_scan_apply<<<blocks, kThreads, 0, st>>>(in, n, scanned, out);
  return check_hip(ctx, hipGetLastError(), "scan");
}

int launch_exclusive_scan(Context& ctx, const uint32_t* in, uint32_t* out, uint32_t n) {
  return exclusive_scan(ctx, ctx.sort, ctx.stream, in, out, n);
}

// Runs, histogram and run destinations of `keys` (stage 1-7 of the sort):
// counts/offsets receive the bin histogram and its scan; returns the tile
// base of the run scan in *tile_base_out.
static int build_runs(Context& ctx, SortWorkspace& ws, hipStream_t st, const int32_t* keys, uint32_t n,
                      uint32_t* counts, uint32_t* offsets, uint32_t** tile_base_out) {
  const Grid& g = ctx.grid;
  const uint32_t bins = g.bins();
  int rc = 0;
#define TPIC_TRY(expr, what)                         \
  do {                                              \
    rc = check_hip(ctx, (expr), what);              \
    if (rc) return rc;                              \
  } while (0)
  TPIC_TRY(hipMemsetAsync(counts, 0, bins * sizeof(uint32_t), st), "sort clear");
  const uint32_t tiles = (n + kTile - 1) / kTile;
  const unsigned grid_stride_blocks = unsigned(ctx.cus > 0 ? ctx.cus * 8 : 2048);
  TPIC_TRY(hipMemsetAsync(ws.failure, 0, sizeof(uint32_t), st), "sort clear");
  TPIC_TRY(hipMemsetAsync(ws.rcounts, 0, bins * sizeof(uint32_t), st), "sort clear");
  TPIC_TRY(hipMemsetAsync(ws.cursor, 0, bins * sizeof(uint32_t), st), "sort clear");
  TPIC_TRY(hipMemsetAsync(ws.large, 0, sizeof(uint32_t), st), "sort clear");
  k_tile_heads<<<tiles, kThreads, 0, st>>>(keys, n, bins, ws.tile_runs, ws.failure);
  TPIC_TRY(hipGetLastError(), "sort heads");
  // tile_runs holds the counts; scan into tile_runs + (tiles + 1).
  uint32_t* tile_base = ws.tile_runs + (ws.tiles + 1);
  if ((rc = exclusive_scan(ctx, ws, st, ws.tile_runs, tile_base, tiles))) return rc;
  k_tile_runs<<<tiles, kThreads, 0, st>>>(keys, n, tiles, tile_base, ws.run_start, ws.run_key, counts, ws.rcounts,
                                          ws.wave_base, ws.failure);
  TPIC_TRY(hipGetLastError(), "sort runs");
  if ((rc = exclusive_scan(ctx, ws, st, counts, offsets, bins))) return rc;
  if ((rc = exclusive_scan(ctx, ws, st, ws.rcounts, ws.roffsets, bins))) return rc;
  k_run_scatter<<<grid_stride_blocks, kThreads, 0, st>>>(tile_base, tiles, ws.run_key, ws.roffsets, ws.cursor,
                                                         ws.bin_runs, ws.failure);
  TPIC_TRY(hipGetLastError(), "sort run scatter");
  k_bin_small<<<(bins + kThreads - 1) / kThreads, kThreads, 0, st>>>(
      g, ws.rcounts, ws.roffsets, offsets, ws.bin_runs, ws.run_start, ws.run_dest, ws.large, ctx.validation,
      ws.failure);
  TPIC_TRY(hipGetLastError(), "sort bins");
  k_bin_large<<<grid_stride_blocks, kThreads, 0, st>>>(ws.large, ws.rcounts, ws.roffsets, offsets, ws.bin_runs,
                                                       ws.run_start, ws.run_dest, ws.failure);
  TPIC_TRY(hipGetLastError(), "sort large bins");
#undef TPIC_TRY
  *tile_base_out = tile_base;
  return 0;
}

static RunPlan sort_run_plan(const SortWorkspace& ws, const uint32_t* counts, const uint32_t* offsets) {
  return RunPlan{counts, offsets, ws.rcounts, ws.roffsets, ws.bin_runs, ws.run_start, ws.run_dest, ws.failure};
}

int launch_sort(Context& ctx, int s) {
  SpeciesState& sp = ctx.species[size_t(s)];
  SortWorkspace& ws = ctx.sort;
  const uint32_t n = sp.count;
  const uint32_t bins = ctx.grid.bins();
  hipStream_t st = ctx.stream;
  phase_begin(ctx, Phase::sort);
  int rc = 0;
  if (n == 0) {
    rc = check_hip(ctx, hipMemsetAsync(sp.counts, 0, bins * sizeof(uint32_t), st), "sort clear");
    if (!rc) rc = check_hip(ctx, hipMemsetAsync(sp.offsets, 0, (bins + 1) * sizeof(uint32_t), st), "sort clear");
    if (rc) return rc;
    sp.sorted = true;
    phase_end(ctx, Phase::sort);
    return 0;
  }
  const ParticleView src = sp.view(), dst = sp.scratch();
  uint32_t* tile_base = nullptr;
  if ((rc = build_runs(ctx, ws, st, src.cell, n, sp.counts, sp.offsets, &tile_base))) return rc;
  (void)tile_base;
  const Grid& g = ctx.grid;
  k_gather_bins<<<(bins + kThreads / 64 - 1) / (kThreads / 64), kThreads, 0, st>>>(g, src, dst, sort_run_plan(ws, sp.counts, sp.offsets),
                                                                                  ws.permutation);
  if ((rc = check_hip(ctx, hipGetLastError(), "sort gather"))) return rc;
  k_copy_on_failure<<<unsigned(ctx.cus > 0 ? ctx.cus * 8 : 2048), kThreads, 0, st>>>(src, dst, n, ws.permutation,
                                                                                     ws.failure, nullptr);
  if ((rc = check_hip(ctx, hipGetLastError(), "sort fallback copy"))) return rc;
  k_sort_failed<<<1, 1, 0, st>>>(ws.failure, ctx.validation);
  if ((rc = check_hip(ctx, hipGetLastError(), "sort status"))) return rc;
  sp.current ^= 1;
  sp.sorted = true;
  phase_end(ctx, Phase::sort);
  return 0;
}

int launch_sort_plan(Context& ctx, int s, SortWorkspace& ws, hipStream_t st) {
  SpeciesState& sp = ctx.species[size_t(s)];
  const uint32_t n = sp.count;
  const uint32_t bins = ctx.grid.bins();
  // Phase times are events on the context stream: a plan on another stream
  // is not timed.
  const bool timed = st == ctx.stream;
  if (timed) phase_begin(ctx, Phase::sort);
  int rc = 0;
  if (n == 0) {
    rc = check_hip(ctx, hipMemsetAsync(sp.counts, 0, bins * sizeof(uint32_t), st), "sort clear");
    if (!rc) rc = check_hip(ctx, hipMemsetAsync(sp.offsets, 0, (bins + 1) * sizeof(uint32_t), st), "sort clear");
    if (rc) return rc;
    if (timed) phase_end(ctx, Phase::sort);
    return 0;
  }
  const ParticleView src = sp.view(), dst = sp.scratch();
  uint32_t* tile_base = nullptr;
  if ((rc = build_runs(ctx, ws, st, src.cell, n, sp.counts, sp.offsets, &tile_base))) return rc;
  const Grid& g = ctx.grid;
  const uint32_t tiles = (n + kTile - 1) / kTile;
  k_tile_sources<<<tiles, kThreads, 0, st>>>(src.cell, n, ws.wave_base, ws.run_start, ws.run_dest, sp.source,
                                             ws.permutation, ws.failure);
  if ((rc = check_hip(ctx, hipGetLastError(), "sort plan"))) return rc;
  k_copy_halo_bins<<<(bins + kThreads - 1) / kThreads, kThreads, 0, st>>>(g, src, dst, sp.counts, sp.offsets,
                                                                          sp.source, ws.failure);
  if ((rc = check_hip(ctx, hipGetLastError(), "sort halo copy"))) return rc;
  k_copy_on_failure<<<unsigned(ctx.cus > 0 ? ctx.cus * 8 : 2048), kThreads, 0, st>>>(src, dst, n, ws.permutation,
                                                                                     ws.failure, sp.source);
  if ((rc = check_hip(ctx, hipGetLastError(), "sort fallback copy"))) return rc;
  k_sort_failed<<<1, 1, 0, st>>>(ws.failure, ctx.validation);
  if ((rc = check_hip(ctx, hipGetLastError(), "sort status"))) return rc;
  if (timed) phase_end(ctx, Phase::sort);
  return 0;
}

int launch_sort_plan_incremental(Context& ctx, int s, SortWorkspace& ws, hipStream_t st) {
  SpeciesState& sp = ctx.species[size_t(s)];
  const uint32_t n = sp.count;
  const Grid& g = ctx.grid;
  const uint32_t bins = g.bins();
  uint32_t* status = ctx.plan_status + s;
  const bool timed = st == ctx.stream;
  if (timed) phase_begin(ctx, Phase::sort);
  int rc = 0;
#define TPIC_TRY(expr, what)            \
  do {                                 \
    rc = check_hip(ctx, (expr), what); \
    if (rc) return rc;                 \
  } while (0)
  TPIC_TRY(hipMemsetAsync(status, 0, sizeof(uint32_t), st), "plan clear");
  if (n == 0) {
    TPIC_TRY(hipMemsetAsync(sp.counts, 0, bins * sizeof(uint32_t), st), "sort clear");
    TPIC_TRY(hipMemsetAsync(sp.offsets, 0, (bins + 1) * sizeof(uint32_t), st), "sort clear");
    if (timed) phase_end(ctx, Phase::sort);
    return 0;
  }
  if (n > ws.capacity || !sp.cross) return fail(ctx, -3, "plan workspace too small");
  const uint32_t blocks = (n + 63u) / 64u;
  TPIC_TRY(hipMemsetAsync(ws.rcounts, 0, bins * sizeof(uint32_t), st), "plan clear");
  TPIC_TRY(hipMemsetAsync(ws.bin_lower, 0, bins * sizeof(uint32_t), st), "plan clear");
  const unsigned bin_blocks = (bins + kThreads - 1) / kThreads;
  const ParticleView src = sp.view(), dst = sp.scratch();
  // Run-table arrays reused: run_dest the crossers per block, run_start (by
  // crosser list index) each crosser's index among its bin's arrivals,
  // rcounts and roffsets the
  // arrivals per bin and their offsets, large each bin's shift, bin_runs the
  // arrivals by bin.
  const uint32_t tiles = (n + kTile - 1) / kTile;
  k_cross_counts<<<(blocks + kThreads - 1) / kThreads, kThreads, 0, st>>>(sp.cross, blocks, ws.run_dest);
  TPIC_TRY(hipGetLastError(), "plan crossers");
  if ((rc = exclusive_scan(ctx, ws, st, ws.run_dest, ws.block_base, blocks))) return rc;
  k_cross_arrivals<<<tiles, kThreads, 0, st>>>(sp.cross, n, ws.block_base, src.cell, bins, sp.offsets, ws.rcounts,
                                               ws.bin_lower, ws.run_start, status);
  TPIC_TRY(hipGetLastError(), "plan arrivals");
  k_cross_counts_new<<<bin_blocks, kThreads, 0, st>>>(bins, sp.offsets, sp.cross, ws.block_base, blocks,
                                                      ws.rcounts, sp.counts, ws.large, status);
  TPIC_TRY(hipGetLastError(), "plan counts");
  if ((rc = exclusive_scan(ctx, ws, st, sp.counts, sp.offsets, bins))) return rc;
  if ((rc = exclusive_scan(ctx, ws, st, ws.rcounts, ws.roffsets, bins))) return rc;
  k_cross_base<<<bin_blocks, kThreads, 0, st>>>(g, n, sp.counts, sp.offsets, ws.large, ws.bin_lower, ws.bin_base,
                                                status, ctx.validation);
  TPIC_TRY(hipGetLastError(), "plan bases");
  k_cross_sources<<<tiles, kThreads, 0, st>>>(src.cell, n, bins, sp.cross, ws.block_base, ws.bin_base, ws.roffsets,
                                               ws.run_start, ws.bin_runs, status, sp.source);
  TPIC_TRY(hipGetLastError(), "plan sources");
  k_cross_bins<<<bin_blocks, kThreads, 0, st>>>(bins, n, sp.counts, sp.offsets, ws.rcounts, ws.roffsets, ws.bin_runs,
                                                ws.bin_lower, status, sp.source);
  TPIC_TRY(hipGetLastError(), "plan arrival order");
  k_copy_halo_bins<<<bin_blocks, kThreads, 0, st>>>(g, src, dst, sp.counts, sp.offsets, sp.source, status);
  TPIC_TRY(hipGetLastError(), "plan halo copy");
#undef TPIC_TRY
  if (timed) phase_end(ctx, Phase::sort);
  return 0;
}

int launch_sort_plan(Context& ctx, int s) { return launch_sort_plan(ctx, s, ctx.sort, ctx.stream); }

}  // namespace tpic_mi
