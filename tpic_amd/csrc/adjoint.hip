// Discrete adjoint of one ECSIM cycle on a periodic box and the transposed
// field solve (the cycle checkpoints it recomputes from are checkpoint.hip).
//
// The reference has no adjoint; this is the exact transpose of the cycle this
// library runs (the reference's cycle: include/tpic/simulation/driver/
// pipeline.hpp:437-541), written for periodic fields and particles. Per
// cycle n the forward map is, with W_a the CIC weights of particle p at its
// fraction r (corners a, x fastest), Bn = C2N(B^n) and beta_c = (q/m) dt/(2c):
//   B_p = sum_a W_a Bn_a,  omega = beta_c B_p,
//   alpha v = (v + (omega.v) omega + v x omega) / (1 + |omega|^2)
//     (ParticleResponse, include/tpic/ecsim/particles/particle_response.hpp:10-44)
//   J_a += (q/V) W_a alpha v,  M_ab += beta_c q W_a W_b alpha  (same alpha for
//     both orders of a pair: mass_matrix_stencil.hpp:139-149)
//   rhs = E + delta curl_C2N(B) - k J,  A E_theta = rhs,
//     A = I + delta^2 curl_C2N curl_N2C + (k/V) M        (maxwell_device.hpp:134-433)
//   v' = 2 alpha (v + beta E_p) - v, r' = r + dt v'/dx  (E_p = sum_a W_a E_theta_a;
//     particle_mover_device.hpp:35-122)
//   E^{n+1} = E_theta/theta - (1 - theta)/theta E^n, B^{n+1} = B^n - c dt curl_N2C(E_theta)
//     (field_update_device.hpp:25-78).
// Given the adjoints (r', v', E^{n+1}, B^{n+1}) of the next state, the
// adjoint cycle forms, in order:
//   Ebar_theta = Ebar'/theta - c dt curl_N2C^T Bbar',  Ebar^n = -(1-theta)/theta Ebar',
//   the push's part per particle (g = vbar' + dt/dx rbar'):
//     vbar = 2 alpha^T g - g,  Ebar_p = 2 beta alpha^T g (scattered into Ebar_theta),
//     omegabar = 2 grad_omega(g . alpha (v + beta E_p)),  rbar = rbar' + dW . (Ebar_p . E_theta),
//   the adjoint solve A^T mu = Ebar_theta (GMRES on the transposed operator),
//   Ebar^n += mu,  Bbar^n += delta curl_C2N^T mu,  and with m = gather(mu):
//     J: vbar += (q/V) alpha^T (-k m), qbar += (-k/V) m . alpha v, omegabar, rbar likewise;
//     M: qbar += -(k/V) beta_c m . alpha E_p, omegabar += -(k/V) beta_c q grad(m . alpha E_p),
//        rbar += -(k/V) beta_c q dW . (mu . alpha E_p + alpha^T m . E_theta);
//   Bbar_p = beta_c omegabar, scattered into Bnbar; rbar += dW . (Bbar_p . Bn);
//   Bbar^n += C2N^T Bnbar.
// grad_omega(a . alpha b) = [(a.omega) b + (b.omega) a + a x b - 2 (a . alpha b) omega] / D.
// On a periodic box curl_C2N = curl_N2C^T over the unique nodes, so both
// curl transposes are the other curl. Node vectors keep the periodic
// duplicates with the value of their unique node.
//
// Particle order: the sort permutes particles each cycle; the adjoint keeps
// the ids of the order it is held in and permutes by id (ids of a species
// must be a permutation of [first, first + count)).
#include "../../include/tpic_mi.h"
#include "internal.hpp"

#include <cmath>
#include <new>
#include <string>
#include <vector>

namespace tpic_mi {

Context* context_of(tpic_mi_context* h);

namespace {

constexpr unsigned kThreads = 256;
constexpr unsigned kWaves = kThreads / 64;

// One species' adjoint arrays.
struct AdjointArrays {
  uint32_t count = 0;
  uint64_t first_id = 0;
  double* rbar = nullptr;  // [3][count] dL/d(fraction)
  double* vbar = nullptr;  // [3][count]
  double* qbar = nullptr;  // [count]
  double* wbar = nullptr;  // [3][count] omega adjoint between the two particle passes
  uint64_t* ids = nullptr; // [count] ids of the order the adjoint is held in
  double* stage = nullptr; // [7][count] permutation staging
  uint32_t* inverse = nullptr;  // [count] position of id first + d in the context's order
};

// Per-particle parameters of a species in the particle passes.
struct SpeciesPass {
  ParticleView p;
  uint32_t count;  // stride of the [3][count] adjoint arrays
  const uint32_t* counts;
  const uint32_t* offsets;
  double* rbar;
  double* vbar;
  double* qbar;
  double* wbar;
  double beta, beta_over_c;
};

struct PassConstants {
  double dtx[3];          // dt / dx per axis
  double inverse_volume;  // 1 / V
  double k;               // 4 pi theta dt
};

__device__ __forceinline__ void cic(double fx, double fy, double fz, double (&w)[8], double (&dw)[8][3]) {
  const double wx[2] = {1.0 - fx, fx}, wy[2] = {1.0 - fy, fy}, wz[2] = {1.0 - fz, fz};
  const double sx[2] = {-1.0, 1.0};
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int ax = a & 1, ay = (a >> 1) & 1, az = (a >> 2) & 1;
    w[a] = (wx[ax] * wy[ay]) * wz[az];
    dw[a][0] = sx[ax] * wy[ay] * wz[az];
    dw[a][1] = sx[ay] * wx[ax] * wz[az];
    dw[a][2] = sx[az] * wx[ax] * wy[ay];
  }
}

// alpha (row-major) of omega.
__device__ __forceinline__ void response(const double (&om)[3], double (&al)[9], double& d) {
  const double x = om[0], y = om[1], z = om[2];
  d = 1.0 + (x * x + y * y + z * z);
  const double inv = 1.0 / d;
  al[0] = (1.0 + x * x) * inv;
  al[1] = (x * y + z) * inv;
  al[2] = (x * z - y) * inv;
  al[3] = (x * y - z) * inv;
  al[4] = (1.0 + y * y) * inv;
  al[5] = (y * z + x) * inv;
  al[6] = (x * z + y) * inv;
  al[7] = (y * z - x) * inv;
  al[8] = (1.0 + z * z) * inv;
}

__device__ __forceinline__ void times(const double (&al)[9], const double (&v)[3], double (&out)[3]) {
#pragma unroll
  for (int r = 0; r < 3; ++r) out[r] = al[3 * r] * v[0] + al[3 * r + 1] * v[1] + al[3 * r + 2] * v[2];
}

__device__ __forceinline__ void times_transposed(const double (&al)[9], const double (&v)[3], double (&out)[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) out[c] = al[c] * v[0] + al[3 + c] * v[1] + al[6 + c] * v[2];
}

__device__ __forceinline__ double dot3(const double (&a)[3], const double (&b)[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// scale * grad_omega (a . alpha b) added to acc.
__device__ __forceinline__ void add_omega_gradient(const double (&om)[3], double d, const double (&al)[9],
                                                   const double (&a)[3], const double (&b)[3], double scale,
                                                   double (&acc)[3]) {
  double ab[3];
  times(al, b, ab);
  const double form = dot3(a, ab), ao = dot3(a, om), bo = dot3(b, om);
  const double cross[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  const double s = scale / d;
#pragma unroll
  for (int c = 0; c < 3; ++c) acc[c] += s * (ao * b[c] + bo * a[c] + cross[c] - 2.0 * form * om[c]);
}

// The wave's cell (i, j, k), its bin range, and its corner values of up to
// three node fields staged in LDS as [field][component][corner].
struct CellWalk {
  int i, j, k;
  uint32_t first, count;
};

__device__ __forceinline__ void stage_corners(const Grid& g, const CellWalk& c, const double* const* fields,
                                              int nfields, unsigned lane, double* lds) {
  const unsigned nodes = g.nodes();
  for (unsigned t = lane; t < unsigned(24 * nfields); t += 64) {
    const unsigned f = t / 24, comp = (t / 8) % 3, a = t % 8;
    const unsigned n = g.node(c.i + int(a & 1), c.j + int((a >> 1) & 1), c.k + int((a >> 2) & 1));
    lds[t] = fields[f][size_t(comp) * nodes + n];
  }
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}

__device__ __forceinline__ void gather(const double (&w)[8], const double* lds, int field, double (&out)[3]) {
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) s += w[a] * lds[field * 24 + comp * 8 + a];
    out[comp] = s;
  }
}

// Sum over the wave of the 24 corner accumulators; lane 0 writes the cell's
// record [corner * 3 + component][cells].
__device__ __forceinline__ void flush_corners(double (&acc)[24], unsigned lane, unsigned cell, unsigned cells,
                                              double* __restrict__ record) {
#pragma unroll
  for (int t = 0; t < 24; ++t) {
    double v = acc[t];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    acc[t] = v;
  }
  if (lane == 0) {
#pragma unroll
    for (int t = 0; t < 24; ++t) record[size_t(t) * cells + cell] = acc[t];
  }
}

__device__ __forceinline__ bool next_cell(const Grid& g, const uint32_t* counts, const uint32_t* offsets,
                                          unsigned cell, CellWalk& c) {
  c.i = int(cell % unsigned(g.nx));
  c.j = int((cell / unsigned(g.nx)) % unsigned(g.ny));
  c.k = int(cell / (unsigned(g.nx) * unsigned(g.ny)));
  const int32_t bin = g.encode(c.i, c.j, c.k);
  c.count = counts[bin];
  c.first = offsets[bin];
  return c.count != 0;
}

// Pass 1: the push's part. Reads rbar', vbar' (this order) and writes the
// partial rbar, vbar, the omega adjoint, and per cell the scattered
// E_theta adjoint.
__global__ __launch_bounds__(kThreads) void k_adjoint_push(Grid g, SpeciesPass sp, PassConstants pc,
                                                           const double* __restrict__ bnode,
                                                           const double* __restrict__ etheta,
                                                           double* __restrict__ record) {
  __shared__ double s_stage[kWaves][48];
  const unsigned lane = threadIdx.x & 63u;
  const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double* lds = s_stage[wave];
  const unsigned cells = g.cells();
  const double* fields[2] = {bnode, etheta};
  for (unsigned cell = blockIdx.x * kWaves + wave; cell < cells; cell += gridDim.x * kWaves) {
    CellWalk c;
    double acc[24];
#pragma unroll
    for (int t = 0; t < 24; ++t) acc[t] = 0.0;
    if (next_cell(g, sp.counts, sp.offsets, cell, c)) {
      stage_corners(g, c, fields, 2, lane, lds);
      for (uint32_t o = lane; o < c.count; o += 64) {
        const uint32_t at = c.first + o;
        const size_t n = sp.count;
        double w[8], dw[8][3];
        cic(sp.p.rx[at], sp.p.ry[at], sp.p.rz[at], w, dw);
        double bp[3], ep[3];
        gather(w, lds, 0, bp);
        gather(w, lds, 1, ep);
        const double om[3] = {sp.beta_over_c * bp[0], sp.beta_over_c * bp[1], sp.beta_over_c * bp[2]};
        double al[9], d;
        response(om, al, d);
        const double v[3] = {sp.p.u[at], sp.p.v[at], sp.p.w[at]};
        double gv[3], rb[3];
#pragma unroll
        for (int c3 = 0; c3 < 3; ++c3) {
          rb[c3] = sp.rbar[c3 * n + at];
          gv[c3] = sp.vbar[c3 * n + at] + pc.dtx[c3] * rb[c3];
        }
        double atg[3];
        times_transposed(al, gv, atg);
        const double ebar_p[3] = {2.0 * sp.beta * atg[0], 2.0 * sp.beta * atg[1], 2.0 * sp.beta * atg[2]};
        const double b[3] = {v[0] + sp.beta * ep[0], v[1] + sp.beta * ep[1], v[2] + sp.beta * ep[2]};
        double wb[3] = {0.0, 0.0, 0.0};
        add_omega_gradient(om, d, al, gv, b, 2.0, wb);
#pragma unroll
        for (int a = 0; a < 8; ++a) {
          const double ea[3] = {lds[24 + a], lds[32 + a], lds[40 + a]};
          const double s = dot3(ebar_p, ea);
          rb[0] += dw[a][0] * s;
          rb[1] += dw[a][1] * s;
          rb[2] += dw[a][2] * s;
          acc[a * 3 + 0] += w[a] * ebar_p[0];
          acc[a * 3 + 1] += w[a] * ebar_p[1];
          acc[a * 3 + 2] += w[a] * ebar_p[2];
        }
#pragma unroll
        for (int c3 = 0; c3 < 3; ++c3) {
          sp.rbar[c3 * n + at] = rb[c3];
          sp.vbar[c3 * n + at] = 2.0 * atg[c3] - gv[c3];
          sp.wbar[c3 * n + at] = wb[c3];
        }
      }
      __builtin_amdgcn_wave_barrier();
      asm volatile("" ::: "memory");
    }
    flush_corners(acc, lane, cell, cells, record);
  }
}

// Pass 2: the deposition's part (J and M) with mu, then the whole omega
// adjoint scattered into Bnbar; writes the state-n rbar, vbar, qbar.
__global__ __launch_bounds__(kThreads) void k_adjoint_deposit(Grid g, SpeciesPass sp, PassConstants pc,
                                                              const double* __restrict__ bnode,
                                                              const double* __restrict__ etheta,
                                                              const double* __restrict__ mu,
                                                              double* __restrict__ record) {
  __shared__ double s_stage[kWaves][72];
  const unsigned lane = threadIdx.x & 63u;
  const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double* lds = s_stage[wave];
  const unsigned cells = g.cells();
  const double* fields[3] = {bnode, etheta, mu};
  const double t_mass = -(pc.k * pc.inverse_volume) * sp.beta_over_c;
  for (unsigned cell = blockIdx.x * kWaves + wave; cell < cells; cell += gridDim.x * kWaves) {
    CellWalk c;
    double acc[24];
#pragma unroll
    for (int t = 0; t < 24; ++t) acc[t] = 0.0;
    if (next_cell(g, sp.counts, sp.offsets, cell, c)) {
      stage_corners(g, c, fields, 3, lane, lds);
      for (uint32_t o = lane; o < c.count; o += 64) {
        const uint32_t at = c.first + o;
        const size_t n = sp.count;
        double w[8], dw[8][3];
        cic(sp.p.rx[at], sp.p.ry[at], sp.p.rz[at], w, dw);
        double bp[3], ep[3], m[3];
        gather(w, lds, 0, bp);
        gather(w, lds, 1, ep);
        gather(w, lds, 2, m);
        const double om[3] = {sp.beta_over_c * bp[0], sp.beta_over_c * bp[1], sp.beta_over_c * bp[2]};
        double al[9], d;
        response(om, al, d);
        const double v[3] = {sp.p.u[at], sp.p.v[at], sp.p.w[at]};
        const double q = sp.p.q[at];
        double rb[3], vb[3], wb[3];
#pragma unroll
        for (int c3 = 0; c3 < 3; ++c3) {
          rb[c3] = sp.rbar[c3 * n + at];
          vb[c3] = sp.vbar[c3 * n + at];
          wb[c3] = sp.wbar[c3 * n + at];
        }
        double qb = sp.qbar[at];
        // J: jbar_p = -k m.
        const double jb[3] = {-pc.k * m[0], -pc.k * m[1], -pc.k * m[2]};
        const double sj = q * pc.inverse_volume;
        double av[3], atj[3], ae[3], atm[3];
        times(al, v, av);
        times_transposed(al, jb, atj);
        times(al, ep, ae);
        times_transposed(al, m, atm);
#pragma unroll
        for (int c3 = 0; c3 < 3; ++c3) vb[c3] += sj * atj[c3];
        qb += pc.inverse_volume * dot3(jb, av);
        add_omega_gradient(om, d, al, jb, v, sj, wb);
        // M: -(k/V) beta_c q m . alpha E_p.
        qb += t_mass * dot3(m, ae);
        add_omega_gradient(om, d, al, m, ep, t_mass * q, wb);
        // omega = beta_c B_p.
        const double bbar[3] = {sp.beta_over_c * wb[0], sp.beta_over_c * wb[1], sp.beta_over_c * wb[2]};
#pragma unroll
        for (int a = 0; a < 8; ++a) {
          const double mua[3] = {lds[48 + a], lds[56 + a], lds[64 + a]};
          const double ea[3] = {lds[24 + a], lds[32 + a], lds[40 + a]};
          const double ba[3] = {lds[a], lds[8 + a], lds[16 + a]};
          const double s = sj * (-pc.k) * dot3(mua, av) + t_mass * q * (dot3(mua, ae) + dot3(atm, ea)) +
                           dot3(bbar, ba);
          rb[0] += dw[a][0] * s;
          rb[1] += dw[a][1] * s;
          rb[2] += dw[a][2] * s;
          acc[a * 3 + 0] += w[a] * bbar[0];
          acc[a * 3 + 1] += w[a] * bbar[1];
          acc[a * 3 + 2] += w[a] * bbar[2];
        }
#pragma unroll
        for (int c3 = 0; c3 < 3; ++c3) {
          sp.rbar[c3 * n + at] = rb[c3];
          sp.vbar[c3 * n + at] = vb[c3];
        }
        sp.qbar[at] = qb;
      }
      __builtin_amdgcn_wave_barrier();
      asm volatile("" ::: "memory");
    }
    flush_corners(acc, lane, cell, cells, record);
  }
}

// Node sums of the per-cell corner records (the 8 cells around a node, with
// periodic wrap, corners in a fixed order): out (+)= sum.
__global__ __launch_bounds__(kThreads) void k_corner_sum(Grid g, const double* __restrict__ record,
                                                         double* __restrict__ out, int accumulate) {
  const unsigned nodes = g.nodes(), cells = g.cells();
  const unsigned sx = unsigned(g.nx + 1), sy = unsigned(g.ny + 1);
  for (unsigned n = blockIdx.x * blockDim.x + threadIdx.x; n < nodes; n += gridDim.x * blockDim.x) {
    const int i = int(n % sx), j = int((n / sx) % sy), k = int(n / (sx * sy));
    double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int ax = a & 1, ay = (a >> 1) & 1, az = (a >> 2) & 1;
      const unsigned cell = g.cell(wrap_cell(i - ax, g.nx), wrap_cell(j - ay, g.ny), wrap_cell(k - az, g.nz));
#pragma unroll
      for (int c = 0; c < 3; ++c) s[c] += record[size_t(a * 3 + c) * cells + cell];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t at = size_t(c) * nodes + n;
      out[at] = accumulate ? out[at] + s[c] : s[c];
    }
  }
}

// out = a x + b y over count values (y may be null: out = a x).
// out may alias x or y (element-wise).
__global__ __launch_bounds__(kThreads) void k_combine(unsigned count, double a, const double* x, double b,
                                                      const double* y, double* out) {
  for (unsigned n = blockIdx.x * blockDim.x + threadIdx.x; n < count; n += gridDim.x * blockDim.x)
    out[n] = y ? a * x[n] + b * y[n] : a * x[n];
}

// Bbar_c += 1/8 sum of the cell's 8 corner values of Bnbar (C2N^T).
__global__ __launch_bounds__(kThreads) void k_c2n_transposed(Grid g, const double* __restrict__ bnbar,
                                                             double* __restrict__ bbar) {
  const unsigned nodes = g.nodes(), cells = g.cells();
  for (unsigned c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += gridDim.x * blockDim.x) {
    const int i = int(c % unsigned(g.nx)), j = int((c / unsigned(g.nx)) % unsigned(g.ny)),
              k = int(c / (unsigned(g.nx) * unsigned(g.ny)));
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 8; ++a)
        s += bnbar[size_t(comp) * nodes + g.node(i + (a & 1), j + ((a >> 1) & 1), k + ((a >> 2) & 1))];
      bbar[size_t(comp) * cells + c] += 0.125 * s;
    }
  }
}

// inverse[id - first] = position of that id in the context's order; flags
// an id out of range.
__global__ __launch_bounds__(kThreads) void k_id_inverse(const uint64_t* __restrict__ ids, uint32_t count,
                                                         uint64_t first, uint32_t* __restrict__ inverse,
                                                         unsigned* __restrict__ flag) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const uint64_t d = ids[i] - first;
    if (d < count) inverse[d] = i;
    else atomicOr(flag, 1u);
  }
}

// The adjoint held in order `from_ids` moved to the context's order: value
// j goes to position inverse[from_ids[j] - first], which must hold the same
// id (ids a permutation of [first, first + count)).
__global__ __launch_bounds__(kThreads) void k_permute(const uint64_t* __restrict__ from_ids,
                                                      const uint64_t* __restrict__ to_ids, uint32_t count,
                                                      uint64_t first, const uint32_t* __restrict__ inverse,
                                                      const double* __restrict__ rbar,
                                                      const double* __restrict__ vbar,
                                                      const double* __restrict__ qbar, double* __restrict__ stage,
                                                      unsigned* __restrict__ flag) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < count; j += gridDim.x * blockDim.x) {
    const uint64_t d = from_ids[j] - first;
    if (d >= count) {
      atomicOr(flag, 1u);
      continue;
    }
    const uint32_t i = inverse[d];
    if (i >= count || to_ids[i] != from_ids[j]) {
      atomicOr(flag, 1u);
      continue;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      stage[size_t(c) * count + i] = rbar[size_t(c) * count + j];
      stage[size_t(3 + c) * count + i] = vbar[size_t(c) * count + j];
    }
    stage[size_t(6) * count + i] = qbar[j];
  }
}

unsigned grid_for(const Context& ctx, uint64_t work) {
  const uint64_t want = (work + kThreads - 1) / kThreads;
  const uint64_t cap = uint64_t(ctx.cus > 0 ? ctx.cus : 256) * 32;
  return unsigned(want < cap ? (want ? want : 1) : cap);
}

unsigned cell_blocks(const Context& ctx) {
  const unsigned want = (ctx.grid.cells() + kWaves - 1) / kWaves;
  const unsigned cap = unsigned(ctx.cus > 0 ? ctx.cus : 256) * 16;
  return want < cap ? want : cap;
}

template <class T>
int device_alloc(Context& c, T** p, size_t n, const char* what) {
  *p = nullptr;
  return check_hip(c, hipMalloc(reinterpret_cast<void**>(p), (n ? n : 1) * sizeof(T)), what);
}

template <class T>
void device_free(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace

}  // namespace tpic_mi

using namespace tpic_mi;

struct tpic_mi_adjoint {
  tpic_mi_context* owner = nullptr;
  std::vector<AdjointArrays> species;
  double* ebar = nullptr;        // [3N] E adjoint (Ebar' on entry of a cycle, Ebar^n on exit)
  double* bbar = nullptr;        // [3C] B adjoint
  double* ethetabar = nullptr;   // [3N] E_theta adjoint (the transposed solve's right-hand side)
  double* mu = nullptr;          // [3N] the transposed solve's solution
  double* bnbar = nullptr;       // [3N] adjoint of B at nodes
  double* curl = nullptr;        // [3N] node scratch
  double* curl_cells = nullptr;  // [3C] cell scratch
  double* record = nullptr;      // [24][cells] per-cell corner sums
  unsigned* flag = nullptr;      // id check word
  unsigned* host_flag = nullptr; // pinned
};

namespace {

void free_adjoint(tpic_mi_adjoint* a) {
  for (auto& s : a->species) {
    device_free(s.rbar); device_free(s.vbar); device_free(s.qbar); device_free(s.wbar);
    device_free(s.ids); device_free(s.stage); device_free(s.inverse);
  }
  device_free(a->ebar); device_free(a->bbar); device_free(a->ethetabar); device_free(a->mu);
  device_free(a->bnbar); device_free(a->curl); device_free(a->curl_cells); device_free(a->record);
  device_free(a->flag);
  if (a->host_flag) (void)hipHostFree(a->host_flag);
  a->host_flag = nullptr;
}

// The adjoint is built for periodic fields and particles, no smoothing.
int adjoint_supported(Context& c) {
  if (c.grid.walls) return fail(c, TPIC_MI_INVALID_STATE, "the adjoint is built for periodic fields");
  for (int a = 0; a < 3; ++a)
    if (!c.grid.particle_periodic[a]) return fail(c, TPIC_MI_INVALID_STATE, "the adjoint is built for periodic particles");
  if (c.num.smoothing) return fail(c, TPIC_MI_INVALID_STATE, "the adjoint is built without smoothing");
  return 0;
}

int read_flag(Context& c, tpic_mi_adjoint* a, unsigned* out) {
  int rc = check_hip(c, hipMemcpyAsync(a->host_flag, a->flag, sizeof(unsigned), hipMemcpyDeviceToHost, c.stream),
                     "adjoint flag");
  if (!rc) rc = check_hip(c, hipStreamSynchronize(c.stream), "adjoint flag");
  if (!rc) *out = *a->host_flag;
  return rc;
}

// Moves the particle adjoints into the context's current particle order.
int permute_to_context(Context& c, tpic_mi_adjoint* a) {
  int rc = check_hip(c, hipMemsetAsync(a->flag, 0, sizeof(unsigned), c.stream), "adjoint flag");
  if (rc) return rc;
  for (size_t s = 0; s < a->species.size(); ++s) {
    AdjointArrays& ad = a->species[s];
    if (ad.count == 0) continue;
    const ParticleView v = c.species[s].view();
    const unsigned blocks = grid_for(c, ad.count);
    k_id_inverse<<<blocks, kThreads, 0, c.stream>>>(v.id, ad.count, ad.first_id, ad.inverse, a->flag);
    k_permute<<<blocks, kThreads, 0, c.stream>>>(ad.ids, v.id, ad.count, ad.first_id, ad.inverse, ad.rbar, ad.vbar,
                                                 ad.qbar, ad.stage, a->flag);
    if ((rc = check_hip(c, hipGetLastError(), "adjoint permutation"))) return rc;
    // The staged values become the arrays; ids follow the context.
    const size_t n = ad.count;
    if ((rc = check_hip(c, hipMemcpyAsync(ad.rbar, ad.stage, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, c.stream),
                        "adjoint permutation")) ||
        (rc = check_hip(c, hipMemcpyAsync(ad.vbar, ad.stage + 3 * n, 3 * n * sizeof(double), hipMemcpyDeviceToDevice,
                                          c.stream),
                        "adjoint permutation")) ||
        (rc = check_hip(c, hipMemcpyAsync(ad.qbar, ad.stage + 6 * n, n * sizeof(double), hipMemcpyDeviceToDevice,
                                          c.stream),
                        "adjoint permutation")) ||
        (rc = check_hip(c, hipMemcpyAsync(ad.ids, v.id, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, c.stream),
                        "adjoint permutation")))
      return rc;
  }
  unsigned bad = 0;
  if ((rc = read_flag(c, a, &bad))) return rc;
  if (bad) return fail(c, TPIC_MI_INVALID_STATE, "particle ids are not a permutation of each species' first ids");
  return 0;
}

SpeciesPass species_pass(Context& c, AdjointArrays& ad, size_t s) {
  SpeciesPass p;
  p.p = c.species[s].view();
  p.count = c.species[s].count;
  p.counts = c.species[s].counts;
  p.offsets = c.species[s].offsets;
  p.rbar = ad.rbar;
  p.vbar = ad.vbar;
  p.qbar = ad.qbar;
  p.wbar = ad.wbar;
  // prepare_particle_mover_parameters: beta = (q/m) dt / 2; the response
  // and the mass scale use beta / c.
  p.beta = c.species[s].q_over_m * c.num.dt / 2;
  p.beta_over_c = p.beta / c.num.c;
  return p;
}

}  // namespace

extern "C" {

int tpic_mi_apply_operator_transposed(tpic_mi_context* h, const double* in, double* out) {
  if (!h || !in || !out || in == out) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = *context_of(h);
  if (int rc = adjoint_supported(c)) return rc;
  c.transposed_operator = true;
  const int rc = launch_operator(c, in, out);
  c.transposed_operator = false;
  return rc;
}

int tpic_mi_solve_transposed(tpic_mi_context* h, const double* device_rhs, double* device_x,
                             tpic_mi_solve_result* r) {
  if (!h || !device_rhs || !device_x) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = *context_of(h);
  if (int rc = adjoint_supported(c)) return rc;
  tpic_mi_solve_result out{};
  c.transposed_operator = true;
  const int rc = run_gmres_system(c, device_rhs, device_x, &out.converged, &out.iterations, &out.initial_residual,
                                  &out.residual, &out.rhs_norm, &out.relative_residual);
  c.transposed_operator = false;
  if (r) *r = out;
  if (rc) return rc;
  if (!out.converged) return fail(c, TPIC_MI_NOT_CONVERGED, "GMRES on the transposed operator did not converge");
  return TPIC_MI_OK;
}

int tpic_mi_adjoint_create(tpic_mi_context* h, tpic_mi_adjoint** out) {
  if (!h || !out) return TPIC_MI_INVALID_ARGUMENT;
  *out = nullptr;
  Context& c = *context_of(h);
  if (int rc = adjoint_supported(c)) return rc;
  tpic_mi_adjoint* a = new (std::nothrow) tpic_mi_adjoint();
  if (!a) return TPIC_MI_BACKEND;
  a->owner = h;
  const size_t nodes = c.grid.nodes(), cells = c.grid.cells();
  int rc = 0;
  a->species.resize(c.species.size());
  for (size_t s = 0; s < c.species.size() && !rc; ++s) {
    AdjointArrays& ad = a->species[s];
    ad.count = c.species[s].count;
    const size_t n = ad.count;
    if ((rc = device_alloc(c, &ad.rbar, 3 * n, "adjoint")) || (rc = device_alloc(c, &ad.vbar, 3 * n, "adjoint")) ||
        (rc = device_alloc(c, &ad.qbar, n, "adjoint")) || (rc = device_alloc(c, &ad.wbar, 3 * n, "adjoint")) ||
        (rc = device_alloc(c, &ad.ids, n, "adjoint")) || (rc = device_alloc(c, &ad.stage, 7 * n, "adjoint")))
      break;
    rc = device_alloc(c, &ad.inverse, n, "adjoint");
  }
  if (!rc && !(rc = device_alloc(c, &a->ebar, 3 * nodes, "adjoint")) &&
      !(rc = device_alloc(c, &a->bbar, 3 * cells, "adjoint")) &&
      !(rc = device_alloc(c, &a->ethetabar, 3 * nodes, "adjoint")) &&
      !(rc = device_alloc(c, &a->mu, 3 * nodes, "adjoint")) && !(rc = device_alloc(c, &a->bnbar, 3 * nodes, "adjoint")) &&
      !(rc = device_alloc(c, &a->curl, 3 * nodes, "adjoint")) &&
      !(rc = device_alloc(c, &a->curl_cells, 3 * cells, "adjoint")) &&
      !(rc = device_alloc(c, &a->record, 24 * cells, "adjoint")) && !(rc = device_alloc(c, &a->flag, 1, "adjoint")))
    rc = check_hip(c, hipHostMalloc(reinterpret_cast<void**>(&a->host_flag), sizeof(unsigned)), "adjoint pinned");
  if (rc) {
    free_adjoint(a);
    delete a;
    return rc;
  }
  *out = a;
  return TPIC_MI_OK;
}

int tpic_mi_adjoint_destroy(tpic_mi_adjoint* a) {
  if (!a) return TPIC_MI_INVALID_ARGUMENT;
  if (a->owner) (void)hipStreamSynchronize(context_of(a->owner)->stream);
  free_adjoint(a);
  delete a;
  return TPIC_MI_OK;
}

int tpic_mi_adjoint_seed(tpic_mi_context* h, tpic_mi_adjoint* a, double weight_e, double weight_b,
                         const uint64_t* first_ids) {
  if (!h || !a || a->owner != h || !first_ids) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = *context_of(h);
  if (!std::isfinite(weight_e) || !std::isfinite(weight_b))
    return fail(c, TPIC_MI_INVALID_ARGUMENT, "adjoint seed weights must be finite");
  const size_t nodes = c.grid.nodes(), cells = c.grid.cells();
  // W_E = V/(8 pi) sum over unique nodes of |E|^2: dW_E/dE = V/(4 pi) E (each
  // periodic duplicate holds its unique node's value); W_B likewise over cells.
  const double volume = double((long double)c.grid.dx * c.grid.dy * c.grid.dz);
  const double scale = volume / (4 * 3.14159265358979323846);
  k_combine<<<grid_for(c, 3 * nodes), kThreads, 0, c.stream>>>(unsigned(3 * nodes), weight_e * scale, c.e, 0.0,
                                                                 nullptr, a->ebar);
  k_combine<<<grid_for(c, 3 * cells), kThreads, 0, c.stream>>>(unsigned(3 * cells), weight_b * scale, c.b, 0.0,
                                                                 nullptr, a->bbar);
  int rc = check_hip(c, hipGetLastError(), "adjoint seed");
  for (size_t s = 0; s < a->species.size() && !rc; ++s) {
    AdjointArrays& ad = a->species[s];
    ad.first_id = first_ids[s];
    const size_t n = ad.count;
    if (!n) continue;
    if ((rc = check_hip(c, hipMemsetAsync(ad.rbar, 0, 3 * n * sizeof(double), c.stream), "adjoint seed")) ||
        (rc = check_hip(c, hipMemsetAsync(ad.vbar, 0, 3 * n * sizeof(double), c.stream), "adjoint seed")) ||
        (rc = check_hip(c, hipMemsetAsync(ad.qbar, 0, n * sizeof(double), c.stream), "adjoint seed")))
      break;
    rc = check_hip(c, hipMemcpyAsync(ad.ids, c.species[s].view().id, n * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                                     c.stream),
                   "adjoint seed");
  }
  if (rc) return rc;
  return check_hip(c, hipStreamSynchronize(c.stream), "adjoint seed");
}

int tpic_mi_adjoint_cycle(tpic_mi_context* h, tpic_mi_adjoint* a, tpic_mi_solve_result* forward,
                          tpic_mi_solve_result* transposed) {
  if (!h || !a || a->owner != h) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = *context_of(h);
  if (int rc = adjoint_supported(c)) return rc;
  for (const auto& sp : c.species)
    if (!sp.sorted) return fail(c, TPIC_MI_INVALID_STATE, "the adjoint cycle starts from a sorted state");
  const Grid& g = c.grid;
  const size_t nodes = g.nodes(), cells = g.cells();
  // 1. The particle adjoints into this state's order.
  int rc = permute_to_context(c, a);
  if (rc) return rc;
  // 2. E_theta of this cycle, recomputed as the forward cycle computed it.
  if ((rc = tpic_mi_solve(h, forward))) return rc;
  // 3. Field update: Ebar_theta = Ebar'/theta - c dt curl_C2N(Bbar');
  //    Ebar^n = -(1 - theta)/theta Ebar'; Bbar^n = Bbar' (in place).
  const double theta = c.num.theta;
  const bool copy = std::abs(theta - 1.0) < 1e-10;
  const double c0 = copy ? 1.0 : 1.0 / theta, c1 = copy ? 0.0 : -(1.0 - theta) / theta;
  if ((rc = launch_curl_c2n(c, a->bbar, a->curl))) return rc;
  k_combine<<<grid_for(c, 3 * nodes), kThreads, 0, c.stream>>>(unsigned(3 * nodes), c0, a->ebar, -c.num.c * c.num.dt,
                                                                 a->curl, a->ethetabar);
  k_combine<<<grid_for(c, 3 * nodes), kThreads, 0, c.stream>>>(unsigned(3 * nodes), c1, a->ebar, 0.0, nullptr,
                                                                 a->ebar);
  if ((rc = check_hip(c, hipGetLastError(), "adjoint field update"))) return rc;
  // 4. The push's part per species; E_theta adjoint scattered to the nodes.
  PassConstants pc{};
  pc.dtx[0] = c.num.dt / g.dx;
  pc.dtx[1] = c.num.dt / g.dy;
  pc.dtx[2] = c.num.dt / g.dz;
  pc.inverse_volume = c.inverse_volume;
  pc.k = c.current_factor;
  const unsigned cblocks = cell_blocks(c), nblocks = grid_for(c, nodes);
  for (size_t s = 0; s < c.species.size(); ++s) {
    if (c.species[s].count == 0) continue;
    SpeciesPass sp = species_pass(c, a->species[s], s);
    k_adjoint_push<<<cblocks, kThreads, 0, c.stream>>>(g, sp, pc, c.bn, c.particle_field(), a->record);
    k_corner_sum<<<nblocks, kThreads, 0, c.stream>>>(g, a->record, a->ethetabar, 1);
    if ((rc = check_hip(c, hipGetLastError(), "adjoint push"))) return rc;
  }
  // 5. A^T mu = Ebar_theta from mu = 0.
  if ((rc = check_hip(c, hipMemsetAsync(a->mu, 0, 3 * nodes * sizeof(double), c.stream), "adjoint solve"))) return rc;
  tpic_mi_solve_result tr{};
  c.transposed_operator = true;
  rc = run_gmres_system(c, a->ethetabar, a->mu, &tr.converged, &tr.iterations, &tr.initial_residual, &tr.residual,
                        &tr.rhs_norm, &tr.relative_residual);
  c.transposed_operator = false;
  if (transposed) *transposed = tr;
  if (rc) return rc;
  if (!tr.converged) return fail(c, TPIC_MI_NOT_CONVERGED, "GMRES on the transposed operator did not converge");
  // 6. rhs: Ebar^n += mu; Bbar^n += delta curl_N2C(mu).
  k_combine<<<grid_for(c, 3 * nodes), kThreads, 0, c.stream>>>(unsigned(3 * nodes), 1.0, a->ebar, 1.0, a->mu,
                                                                 a->ebar);
  if ((rc = check_hip(c, hipGetLastError(), "adjoint rhs"))) return rc;
  if ((rc = launch_curl_n2c(c, a->mu, a->curl_cells))) return rc;
  k_combine<<<grid_for(c, 3 * cells), kThreads, 0, c.stream>>>(unsigned(3 * cells), 1.0, a->bbar, c.delta,
                                                                 a->curl_cells, a->bbar);
  // 7. The deposition's part; B-at-nodes adjoint scattered and folded into Bbar^n.
  bool first = true;
  for (size_t s = 0; s < c.species.size(); ++s) {
    if (c.species[s].count == 0) continue;
    SpeciesPass sp = species_pass(c, a->species[s], s);
    k_adjoint_deposit<<<cblocks, kThreads, 0, c.stream>>>(g, sp, pc, c.bn, c.particle_field(), a->mu, a->record);
    k_corner_sum<<<nblocks, kThreads, 0, c.stream>>>(g, a->record, a->bnbar, first ? 0 : 1);
    if ((rc = check_hip(c, hipGetLastError(), "adjoint deposit"))) return rc;
    first = false;
  }
  if (!first) {
    k_c2n_transposed<<<grid_for(c, cells), kThreads, 0, c.stream>>>(g, a->bnbar, a->bbar);
    if ((rc = check_hip(c, hipGetLastError(), "adjoint cell to node"))) return rc;
  }
  return check_hip(c, hipStreamSynchronize(c.stream), "adjoint cycle");
}

int tpic_mi_adjoint_get_field(tpic_mi_adjoint* a, int which, double* host) {
  if (!a || !host) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = *context_of(a->owner);
  const double* src = nullptr;
  size_t n = 0;
  if (which == TPIC_MI_FIELD_E) {
    src = a->ebar;
    n = 3 * size_t(c.grid.nodes());
  } else if (which == TPIC_MI_FIELD_B) {
    src = a->bbar;
    n = 3 * size_t(c.grid.cells());
  } else {
    return fail(c, TPIC_MI_INVALID_ARGUMENT, "the adjoint holds E (TPIC_MI_FIELD_E) and B (TPIC_MI_FIELD_B)");
  }
  int rc = check_hip(c, hipMemcpyAsync(host, src, n * sizeof(double), hipMemcpyDeviceToHost, c.stream), "adjoint field");
  if (!rc) rc = check_hip(c, hipStreamSynchronize(c.stream), "adjoint field");
  return rc;
}

int tpic_mi_adjoint_get_particles(tpic_mi_adjoint* a, int species, double* rbar, double* vbar, double* qbar,
                                  uint64_t* ids) {
  if (!a) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = *context_of(a->owner);
  if (species < 0 || size_t(species) >= a->species.size()) return fail(c, TPIC_MI_INVALID_ARGUMENT, "species");
  AdjointArrays& ad = a->species[size_t(species)];
  const size_t n = ad.count;
  int rc = 0;
  if (n && rbar) rc = check_hip(c, hipMemcpyAsync(rbar, ad.rbar, 3 * n * 8, hipMemcpyDeviceToHost, c.stream), "adjoint");
  if (!rc && n && vbar)
    rc = check_hip(c, hipMemcpyAsync(vbar, ad.vbar, 3 * n * 8, hipMemcpyDeviceToHost, c.stream), "adjoint");
  if (!rc && n && qbar) rc = check_hip(c, hipMemcpyAsync(qbar, ad.qbar, n * 8, hipMemcpyDeviceToHost, c.stream), "adjoint");
  if (!rc && n && ids) rc = check_hip(c, hipMemcpyAsync(ids, ad.ids, n * 8, hipMemcpyDeviceToHost, c.stream), "adjoint");
  if (!rc) rc = check_hip(c, hipStreamSynchronize(c.stream), "adjoint");
  return rc;
}

}  // extern "C"
