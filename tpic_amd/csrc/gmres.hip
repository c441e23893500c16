// Restarted GMRES(m) for A(E_theta) = rhs on the device, FP64 throughout
// (run_gmres_system solves A x = b for any device b and initial x, and with
// Context::transposed_operator the adjoint system A^T x = b).
//
// Algorithm of Gmres::host_blocking_solve
// (include/tpic/ecsim/gmres/gmres_device.hpp:355-454): x0 = E^n, r = b - A x,
// tolerance relative to |b| (scale 1 for b = 0); each restart normalises r;
// each Arnoldi step applies A, forms the projection dots against the basis in
// one batch, orthogonalises by one classical Gram-Schmidt pass fused with the
// squared norm, then one kernel normalises the new basis vector while block 0
// thread 0 appends the column to the Hessenberg matrix and applies the Givens
// rotations (ArnoldiLeastSquaresUpdate, :65-98). The reference's host reads
// two control scalars per step; here that thread also tests them (the
// estimate against the tolerance, the breakdown) and raises a stop word with
// the step count, every kernel of a later step returns at once when the word
// is up, and the host enqueues several steps between reads of the word: the
// same steps run and stop where they would, with fewer host round trips.
// After each restart the true residual |b - A x| decides
// convergence; the solution is committed only when converged. Vectors hold
// every interior node, periodic duplicates included (:133-137), so dots weight
// shared nodes as the reference's compact vectors do.
//
// Reductions are deterministic: fixed block partials reduced by one block per
// row in a fixed tree, no atomics.
#include "internal.hpp"

#include <cmath>

namespace tpic_mi {
namespace {

constexpr unsigned kThreads = 256;
constexpr unsigned kTile = 4;

__device__ inline double block_reduce(double v, double* lds) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double sum = 0.0;
  if (threadIdx.x == 0)
    for (unsigned w = 0; w < kThreads / 64; ++w) sum += lds[w];
  __syncthreads();
  return sum;  // valid in thread 0
}

// Stop words of one Arnoldi restart cycle (after the Hessenberg state):
// [0] residual estimate, [1] breakdown, [2] stop, [3] steps done at the stop,
// [4] a non-finite recurrence.
constexpr int kStatusWords = 5;

__device__ inline bool stopped(const double* stop) { return stop && *stop != 0.0; }

__global__ __launch_bounds__(kThreads) void k_dot_batch(const double* __restrict__ basis,
                                                        const double* __restrict__ image,
                                                        double* __restrict__ partials, unsigned count,
                                                        int columns, const double* __restrict__ stop) {
  if (stopped(stop)) return;
  __shared__ double lds[kThreads / 64];
  for (unsigned first = blockIdx.y * kTile; first < unsigned(columns); first += gridDim.y * kTile) {
    double sum[kTile] = {0.0, 0.0, 0.0, 0.0};
    for (unsigned n = blockIdx.x * blockDim.x + threadIdx.x; n < count; n += gridDim.x * blockDim.x) {
      const double v = image[n];
#pragma unroll
      for (unsigned r = 0; r < kTile; ++r)
        if (first + r < unsigned(columns)) sum[r] += basis[size_t(first + r) * count + n] * v;
    }
#pragma unroll
    for (unsigned r = 0; r < kTile; ++r)
      if (first + r < unsigned(columns)) {
        const double total = block_reduce(sum[r], lds);
        if (threadIdx.x == 0) partials[size_t(first + r) * gridDim.x + blockIdx.x] = total;
      }
  }
}

// One block per row: sum of the row's partials; rows >= filled are zeroed.
__global__ __launch_bounds__(kThreads) void k_reduce_rows(const double* __restrict__ partials,
                                                          double* __restrict__ values, unsigned stride,
                                                          unsigned count, unsigned filled,
                                                          const double* __restrict__ stop) {
  if (stopped(stop)) return;
  __shared__ double lds[kThreads / 64];
  const unsigned row = blockIdx.x;
  if (row >= filled) {
    if (threadIdx.x == 0) values[row] = 0.0;
    return;
  }
  double s = 0.0;
  for (unsigned i = threadIdx.x; i < count; i += blockDim.x) s += partials[size_t(row) * stride + i];
  const double total = block_reduce(s, lds);
  if (threadIdx.x == 0) values[row] = total;
}

// r = b - image; partial sums of |r|^2 (and |b|^2).
__global__ __launch_bounds__(kThreads) void k_residual(const double* __restrict__ b,
                                                       const double* __restrict__ image,
                                                       double* __restrict__ r, double* __restrict__ partials,
                                                       unsigned count, int with_rhs) {
  __shared__ double lds[kThreads / 64];
  double s0 = 0.0, s1 = 0.0;
  for (unsigned n = blockIdx.x * blockDim.x + threadIdx.x; n < count; n += gridDim.x * blockDim.x) {
    const double bv = b[n];
    const double v = bv - image[n];
    r[n] = v;
    s0 += v * v;
    s1 += bv * bv;
  }
  const double t0 = block_reduce(s0, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = t0;
  if (with_rhs) {
    const double t1 = block_reduce(s1, lds);
    if (threadIdx.x == 0) partials[gridDim.x + blockIdx.x] = t1;
  }
}

// Classical Gram-Schmidt against the first `columns` basis vectors with the
// reduced dots; stores the image and its squared-norm partial; block 0 copies
// the dots into the Hessenberg column.
__global__ __launch_bounds__(kThreads) void k_orthogonalize(const double* __restrict__ basis,
                                                            const double* __restrict__ dots,
                                                            double* __restrict__ image,
                                                            double* __restrict__ hcol,
                                                            double* __restrict__ partials, unsigned count,
                                                            int columns, const double* __restrict__ stop) {
  if (stopped(stop)) return;
  __shared__ double lds[kThreads / 64];
  __shared__ double s_dots[64];
  if (threadIdx.x < unsigned(columns) && threadIdx.x < 64) s_dots[threadIdx.x] = dots[threadIdx.x];
  __syncthreads();
  double norm2 = 0.0;
  for (unsigned n = blockIdx.x * blockDim.x + threadIdx.x; n < count; n += gridDim.x * blockDim.x) {
    double value = image[n];
    for (int c = 0; c < columns; ++c) value -= s_dots[c] * basis[size_t(c) * count + n];
    image[n] = value;
    norm2 += value * value;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int row = 0; row < columns; ++row) hcol[row] = s_dots[row];
  const double total = block_reduce(norm2, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

struct GivensTail {
  int mode;  // 0: restart (rhs[0] = |r|), 1: Arnoldi update of column `step`
  double* column;
  double* cosine;
  double* sine;
  double* rhs;
  double* status;  // kStatusWords stop words
  int step;
  double absolute;  // the tolerance on the residual estimate
};

__device__ void run_tail(const GivensTail& t, double norm2) {
  if (t.mode == 0) {
    t.rhs[0] = sqrt(norm2);
    return;
  }
  const int step = t.step;
  double* column = t.column;
  const double sub = sqrt(norm2);
  column[step + 1] = sub;
  for (int row = 0; row < step; ++row) {
    const double upper = column[row], lower = column[row + 1];
    column[row] = t.cosine[row] * upper + t.sine[row] * lower;
    column[row + 1] = -t.sine[row] * upper + t.cosine[row] * lower;
  }
  const double diagonal = column[step];
  const double den = hypot(diagonal, sub);
  const double c = den != 0.0 ? diagonal / den : 1.0;
  const double s = den != 0.0 ? sub / den : 0.0;
  t.cosine[step] = c;
  t.sine[step] = s;
  column[step] = c * diagonal + s * sub;
  column[step + 1] = 0.0;
  const double upper = t.rhs[step], lower = t.rhs[step + 1];
  t.rhs[step] = c * upper + s * lower;
  t.rhs[step + 1] = -s * upper + c * lower;
  t.status[0] = fabs(t.rhs[step + 1]);
  t.status[1] = sub == 0.0 ? 1.0 : 0.0;
  // The host's test of the two controls (gmres_device.hpp:431-441).
  const bool finite = t.status[0] <= 1.7976931348623157e308;
  if (!finite || t.status[0] <= t.absolute || t.status[1] != 0.0) {
    t.status[4] = finite ? 0.0 : 1.0;
    t.status[3] = double(step + 1);
    t.status[2] = 1.0;
  }
}

// target = source / sqrt(norm2) (zero for a zero norm); block 0 thread 0 runs
// the Givens tail first.
__global__ __launch_bounds__(kThreads) void k_normalize(const double* __restrict__ source,
                                                        const double* __restrict__ norm2,
                                                        double* __restrict__ target, unsigned count,
                                                        GivensTail tail, const double* __restrict__ stop) {
  // A block that finds the word already up (raised by an earlier step, or by
  // this step's tail) skips: the vector of a step past the stop is unused.
  if (stopped(stop)) return;
  const double n2 = norm2[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) run_tail(tail, n2);
  const double scale = n2 > 0.0 ? 1.0 / sqrt(n2) : 0.0;
  for (unsigned n = blockIdx.x * blockDim.x + threadIdx.x; n < count; n += gridDim.x * blockDim.x)
    target[n] = source[n] * scale;
}

__global__ void k_back_substitute(const double* __restrict__ h, const double* __restrict__ rhs,
                                  double* __restrict__ y, int leading, int columns) {
  for (int row = columns - 1; row >= 0; --row) {
    double sum = 0.0;
    for (int col = row + 1; col < columns; ++col) sum += h[size_t(col) * leading + row] * y[col];
    const double d = h[size_t(row) * leading + row];
    y[row] = d != 0.0 ? (rhs[row] - sum) / d : 0.0;
  }
}

__global__ __launch_bounds__(kThreads) void k_update(const double* __restrict__ basis,
                                                     const double* __restrict__ y, double* __restrict__ x,
                                                     unsigned count, int columns) {
  __shared__ double s_y[64];
  if (threadIdx.x < unsigned(columns) && threadIdx.x < 64) s_y[threadIdx.x] = y[threadIdx.x];
  __syncthreads();
  for (unsigned n = blockIdx.x * blockDim.x + threadIdx.x; n < count; n += gridDim.x * blockDim.x) {
    double value = x[n];
    for (int c = 0; c < columns; ++c) value += s_y[c] * basis[size_t(c) * count + n];
    x[n] = value;
  }
}

}  // namespace

int run_gmres(Context& ctx, int* converged_out, int* iterations_out, double* initial_out, double* residual_out,
              double* rhs_norm_out, double* relative_out) {
  return run_gmres_system(ctx, ctx.rhs, ctx.etheta, converged_out, iterations_out, initial_out, residual_out,
                          rhs_norm_out, relative_out);
}

int run_gmres_system(Context& ctx, const double* rhs_in, double* solution, int* converged_out, int* iterations_out,
                     double* initial_out, double* residual_out, double* rhs_norm_out, double* relative_out) {
  GmresWorkspace& w = ctx.gmres;
  hipStream_t st = ctx.stream;
  const unsigned n = w.count;
  const int m = ctx.num.restart;
  const unsigned blocks = w.blocks;
  const unsigned batch = unsigned(m > 2 ? m : 2);
  double* h = w.state;
  double* cs = h + size_t(m) * (m + 1);
  double* sn = cs + m;
  double* g = sn + m;
  double* y = g + (m + 1);
  double* status = y + m;
  int rc = 0;
#define TPIC_TRY(expr, what)             \
  do {                                  \
    rc = check_hip(ctx, (expr), what);  \
    if (rc) return rc;                  \
  } while (0)
  auto read_words = [&](const double* src, int words, double* out) -> int {
    int e = check_hip(ctx, hipMemcpyAsync(w.controls, src, size_t(words) * sizeof(double), hipMemcpyDeviceToHost, st),
                      "gmres readback");
    if (e) return e;
    e = check_hip(ctx, hipStreamSynchronize(st), "gmres readback");
    if (e) return e;
    for (int k = 0; k < words; ++k) out[k] = w.controls[k];
    return 0;
  };
  auto read_two = [&](const double* src, double out[2]) -> int { return read_words(src, 2, out); };
  auto residual = [&](bool with_rhs, double out[2]) -> int {
    int e = launch_operator(ctx, w.x, w.image);
    if (e) return e;
    k_residual<<<blocks, kThreads, 0, st>>>(w.b, w.image, w.r, w.partials, n, with_rhs ? 1 : 0);
    if ((e = check_hip(ctx, hipGetLastError(), "gmres residual"))) return e;
    k_reduce_rows<<<batch, kThreads, 0, st>>>(w.partials, w.values, blocks, blocks, with_rhs ? 2u : 1u, nullptr);
    if ((e = check_hip(ctx, hipGetLastError(), "gmres reduce"))) return e;
    return read_two(w.values, out);
  };
  phase_begin(ctx, Phase::solve);
  TPIC_TRY(hipMemcpyAsync(w.x, solution, n * sizeof(double), hipMemcpyDeviceToDevice, st), "gmres x");
  TPIC_TRY(hipMemcpyAsync(w.b, rhs_in, n * sizeof(double), hipMemcpyDeviceToDevice, st), "gmres b");
  double two[2];
  if ((rc = residual(true, two))) return rc;
  if (!std::isfinite(two[0]) || !std::isfinite(two[1]) || two[0] < 0 || two[1] < 0)
    return fail(ctx, -3, "GMRES: nonfinite initial residual");
  double res = std::sqrt(two[0]);
  const double initial = res, rhs_norm = std::sqrt(two[1]);
  const double scale = rhs_norm == 0.0 ? 1.0 : rhs_norm;
  const double absolute = scale * ctx.num.gmres_tol;
  bool converged = res <= absolute;
  int iterations = 0;
  const size_t state_words = size_t(m) * (m + 1) + 4 * size_t(m) + 1 + kStatusWords;
  // Steps enqueued between two reads of the stop word (one by one with
  // smoothing, whose kernels do not see the word).
  const int steps_per_read = ctx.num.smoothing ? 1 : 8;
  double* basis = w.basis;
  while (!converged && iterations < ctx.num.max_iterations) {
    TPIC_TRY(hipMemsetAsync(w.state, 0, state_words * sizeof(double), st), "gmres state");
    // Restart: basis 0 = r / |r| with |r|^2 still in values[0]; rhs[0] = |r|.
    GivensTail restart{0, nullptr, nullptr, nullptr, g, nullptr, 0, 0.0};
    k_normalize<<<blocks, kThreads, 0, st>>>(w.r, w.values, basis, n, restart, nullptr);
    TPIC_TRY(hipGetLastError(), "gmres restart");
    const int columns = std::min(m, ctx.num.max_iterations - iterations);
    const double* stop = steps_per_read > 1 ? status + 2 : nullptr;
    int completed = 0;
    bool breakdown = false;
    int enqueued = 0;
    while (enqueued < columns) {
      const int upto = std::min(columns, enqueued + steps_per_read);
      for (int col = enqueued; col < upto; ++col) {
        if ((rc = launch_operator(ctx, basis + size_t(col) * n, w.image, stop))) return rc;
        const dim3 dgrid(blocks, unsigned((col + 1 + kTile - 1) / kTile));
        k_dot_batch<<<dgrid, kThreads, 0, st>>>(basis, w.image, w.partials, n, col + 1, stop);
        TPIC_TRY(hipGetLastError(), "gmres dots");
        k_reduce_rows<<<batch, kThreads, 0, st>>>(w.partials, w.values, blocks, blocks, unsigned(col + 1), stop);
        TPIC_TRY(hipGetLastError(), "gmres reduce");
        double* hcol = h + size_t(col) * (m + 1);
        k_orthogonalize<<<blocks, kThreads, 0, st>>>(basis, w.values, w.image, hcol, w.partials, n, col + 1, stop);
        TPIC_TRY(hipGetLastError(), "gmres orthogonalize");
        k_reduce_rows<<<batch, kThreads, 0, st>>>(w.partials, w.values, blocks, blocks, 1u, stop);
        TPIC_TRY(hipGetLastError(), "gmres reduce");
        GivensTail arnoldi{1, hcol, cs, sn, g, status, col, absolute};
        k_normalize<<<blocks, kThreads, 0, st>>>(w.image, w.values, basis + size_t(col + 1) * n, n, arnoldi, stop);
        TPIC_TRY(hipGetLastError(), "gmres normalize");
      }
      enqueued = upto;
      double word[kStatusWords];
      if ((rc = read_words(status, kStatusWords, word))) return rc;
      if (word[4] != 0.0) return fail(ctx, -3, "GMRES: nonfinite Arnoldi recurrence");
      if (word[2] != 0.0) {
        completed = int(word[3]);
        breakdown = word[1] != 0.0;
        break;
      }
      completed = enqueued;
    }
    k_back_substitute<<<1, 1, 0, st>>>(h, g, y, m + 1, completed);
    TPIC_TRY(hipGetLastError(), "gmres back substitution");
    k_update<<<blocks, kThreads, 0, st>>>(basis, y, w.x, n, completed);
    TPIC_TRY(hipGetLastError(), "gmres update");
    iterations += completed;
    // Always verify b - A x after an accepted update.
    if ((rc = residual(false, two))) return rc;
    if (!std::isfinite(two[0]) || two[0] < 0) return fail(ctx, -3, "GMRES: nonfinite residual");
    res = std::sqrt(two[0]);
    converged = res <= absolute;
    if (breakdown && !converged) break;
  }
  if (converged)
    TPIC_TRY(hipMemcpyAsync(solution, w.x, n * sizeof(double), hipMemcpyDeviceToDevice, st), "gmres commit");
  phase_end(ctx, Phase::solve);
  *converged_out = converged ? 1 : 0;
  *iterations_out = iterations;
  *initial_out = initial;
  *residual_out = res;
  *rhs_norm_out = rhs_norm;
  *relative_out = res / scale;
  return 0;
#undef TPIC_TRY
}

}  // namespace tpic_mi
