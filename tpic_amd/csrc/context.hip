// C-ABI of the MI355X ECSIM hot path (declared in include/tpic_mi.h): the
// context owns every device buffer, sized once at creation, and sequences the
// kernels on one in-order HIP stream (the single-GPU replacement of the
// reference's multi-stream Pipeline, include/tpic/simulation/driver/pipeline.hpp).
#include "../../include/tpic_mi.h"
#include "internal.hpp"

#include <cmath>
#include <cstring>
#include <new>
#include <utility>

struct tpic_mi_context {
  tpic_mi::Context ctx;
  // Timing: pending (begin, end) event pairs per phase, resolved on query.
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[int(tpic_mi::Phase::count)];
  std::vector<hipEvent_t> pool;
  hipEvent_t open_begin[int(tpic_mi::Phase::count)]{};
  bool record_permutation = false;
};

namespace tpic_mi {

namespace {
thread_local std::string g_create_error;

tpic_mi_context* owner_of(Context& ctx) { return static_cast<tpic_mi_context*>(ctx.owner); }

hipEvent_t take_event(tpic_mi_context* h) {
  if (!h->pool.empty()) {
    hipEvent_t e = h->pool.back();
    h->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

int fail(Context& ctx, int status, const std::string& message) {
  ctx.error = message;
  return status;
}

void set_context_free_error(const std::string& message) { g_create_error = message; }

Context* context_of(tpic_mi_context* h) { return &h->ctx; }

int check_hip(Context& ctx, hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  ctx.error = std::string(what) + ": " + hipGetErrorString(e);
  return TPIC_MI_BACKEND;
}

void phase_begin(Context& ctx, Phase p) {
  if (!ctx.timing) return;
  tpic_mi_context* h = owner_of(ctx);
  hipEvent_t e = take_event(h);
  if (!e) return;
  (void)hipEventRecord(e, ctx.stream);
  h->open_begin[int(p)] = e;
}

void phase_end(Context& ctx, Phase p) {
  if (!ctx.timing) return;
  tpic_mi_context* h = owner_of(ctx);
  hipEvent_t b = h->open_begin[int(p)];
  if (!b) return;
  hipEvent_t e = take_event(h);
  if (!e) return;
  (void)hipEventRecord(e, ctx.stream);
  h->pending[int(p)].emplace_back(b, e);
  h->open_begin[int(p)] = nullptr;
}

}  // namespace tpic_mi

using namespace tpic_mi;

namespace {

template <class T>
int alloc(Context& ctx, T** p, size_t count, const char* what) {
  *p = nullptr;
  if (count == 0) count = 1;
  return check_hip(ctx, hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)), what);
}

template <class T>
void release(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

size_t scan_scratch_words(size_t n) {
  size_t words = 0;
  while (n > 2048) {
    const size_t blocks = (n + 2047) / 2048;
    words += 2 * blocks + 1;
    n = blocks;
  }
  return words + 16;
}

void free_all(tpic_mi_context* h) {
  Context& c = h->ctx;
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  release_checkpoints(c);
  for (auto& sp : c.species) {
    for (auto& s : sp.slot) {
      release(s.rx); release(s.ry); release(s.rz); release(s.u); release(s.v); release(s.w); release(s.q);
      release(s.cell); release(s.id);
    }
    sp.counts = nullptr;
    sp.offsets = nullptr;
    release(sp.source);
    release(sp.cross);
  }
  release(c.bin_counts);
  release(c.bin_offsets);
  release(c.e); release(c.b); release(c.bn); release(c.current); release(c.etheta); release(c.rhs);
  release(c.smooth_a); release(c.smooth_b); release(c.eparticle); release(c.projected); release(c.mass_response);
  release(c.mass); release(c.cellbuf); release(c.corner_current); release(c.reduce); release(c.reduce_out); release(c.validation);
  release(c.moment_cells);
  for (SortWorkspace* pw : {&c.sort, &c.sort_side}) {
    SortWorkspace& w = *pw;
    release(w.tile_runs); release(w.wave_base); release(w.run_start); release(w.run_key); release(w.run_dest);
    release(w.bin_runs); release(w.rcounts); release(w.roffsets); release(w.cursor); release(w.large);
    release(w.scan_scratch); release(w.failure); release(w.permutation);
    release(w.block_base); release(w.bin_base); release(w.bin_lower);
  }
  release(c.plan_status);
  if (c.fork) (void)hipEventDestroy(c.fork);
  if (c.join) (void)hipEventDestroy(c.join);
  if (c.side_stream) (void)hipStreamDestroy(c.side_stream);
  c.fork = c.join = nullptr;
  c.side_stream = nullptr;
  GmresWorkspace& g = c.gmres;
  release(g.basis); release(g.x); release(g.b); release(g.r); release(g.image); release(g.values);
  release(g.partials); release(g.state); release(g.curl_cells);
  if (g.controls) (void)hipHostFree(g.controls);
  if (c.host_word) (void)hipHostFree(c.host_word);
  if (c.host_scalars) (void)hipHostFree(c.host_scalars);
  if (c.host_status) (void)hipHostFree(c.host_status);
  c.host_status = nullptr;
  g.controls = nullptr;
  c.host_word = nullptr;
  c.host_scalars = nullptr;
  for (auto& v : h->pending)
    for (auto& pr : v) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  for (auto e : h->pool) (void)hipEventDestroy(e);
  for (auto e : h->open_begin)
    if (e) (void)hipEventDestroy(e);
  if (c.own_stream && c.stream) (void)hipStreamDestroy(c.stream);
  c.stream = nullptr;
}

int allocate(tpic_mi_context* h) {
  Context& c = h->ctx;
  const Grid& g = c.grid;
  const size_t nodes = g.nodes(), cells = g.cells(), bins = g.bins();
  int rc = 0;
#define TPIC_ALLOC(ptr, n) \
  if ((rc = alloc(c, &(ptr), (n), "allocation " #ptr))) return rc
  uint32_t capacity = 0;
  for (auto& sp : c.species) {
    for (auto& s : sp.slot) {
      TPIC_ALLOC(s.rx, sp.count); TPIC_ALLOC(s.ry, sp.count); TPIC_ALLOC(s.rz, sp.count);
      TPIC_ALLOC(s.u, sp.count); TPIC_ALLOC(s.v, sp.count); TPIC_ALLOC(s.w, sp.count);
      TPIC_ALLOC(s.q, sp.count); TPIC_ALLOC(s.cell, sp.count); TPIC_ALLOC(s.id, sp.count);
    }
    TPIC_ALLOC(sp.source, sp.count);
    TPIC_ALLOC(sp.cross, (size_t(sp.count) + 63) / 64);
    capacity = std::max(capacity, sp.count);
  }
  TPIC_ALLOC(c.bin_counts, c.species.size() * bins);
  TPIC_ALLOC(c.bin_offsets, c.species.size() * (bins + 1));
  for (size_t s = 0; s < c.species.size(); ++s) {
    c.species[s].counts = c.bin_counts + s * bins;
    c.species[s].offsets = c.bin_offsets + s * (bins + 1);
  }
  TPIC_ALLOC(c.e, 3 * nodes);
  TPIC_ALLOC(c.b, 3 * cells);
  TPIC_ALLOC(c.bn, 3 * nodes);
  TPIC_ALLOC(c.current, 3 * nodes);
  TPIC_ALLOC(c.etheta, 3 * nodes);
  TPIC_ALLOC(c.rhs, 3 * nodes);
  if (g.walls) TPIC_ALLOC(c.projected, 3 * nodes);
  if (g.walls && !c.num.smoothing) TPIC_ALLOC(c.mass_response, 3 * nodes);
  if (c.num.smoothing) {
    TPIC_ALLOC(c.smooth_a, 3 * nodes);
    TPIC_ALLOC(c.smooth_b, 3 * nodes);
    TPIC_ALLOC(c.eparticle, 3 * nodes);
  }
  c.mass_single = c.num.mass_kernel != TPIC_MI_MASS_FP64;
  // The matrix is allocated as doubles in both forms (FP32 uses half of it).
  {
    double* m = nullptr;
    TPIC_ALLOC(m, size_t(kEntries) * g.vstride);
    c.mass = m;
  }
  // Cell records are allocated as doubles in both forms (float uses half).
  {
    double* records = nullptr;
    TPIC_ALLOC(records, size_t(kCellRecord) * cells);
    c.cellbuf = records;
  }
  TPIC_ALLOC(c.corner_current, size_t(kCornerRecord) * cells);
  const size_t diag_blocks = size_t(std::max(1, std::min(1024, (c.cus > 0 ? c.cus : 256) * 4)));
  // Rows: the scalars (2 + 5 per species) or the four velocity-centering sums;
  // reduce_out also holds the centering's flag word after its four sums.
  const size_t diag_rows = std::max<size_t>(2 + 5 * c.species.size(), 5);
  TPIC_ALLOC(c.reduce, diag_rows * diag_blocks);
  TPIC_ALLOC(c.reduce_out, diag_rows);
  TPIC_ALLOC(c.validation, 1);
  TPIC_ALLOC(c.plan_status, kMaxSpecies);
  // The main plan workspace (sized for the largest species, with the
  // permutation record) and, with two or more species, the side one for
  // species 0.
  const bool side = c.species.size() >= 2;
  for (int which = 0; which < (side ? 2 : 1); ++which) {
    SortWorkspace& w = which == 0 ? c.sort : c.sort_side;
    w.capacity = which == 0 ? capacity : c.species[0].count;
    w.tiles = (w.capacity + 4095) / 4096;
    TPIC_ALLOC(w.tile_runs, 2 * (size_t(w.tiles) + 1));
    TPIC_ALLOC(w.wave_base, 4 * (size_t(w.tiles) + 1));
    TPIC_ALLOC(w.run_start, size_t(w.capacity) + 1);
    TPIC_ALLOC(w.run_key, w.capacity);
    TPIC_ALLOC(w.run_dest, w.capacity);
    TPIC_ALLOC(w.bin_runs, w.capacity);
    if (which == 0) TPIC_ALLOC(w.permutation, w.capacity);
    TPIC_ALLOC(w.rcounts, bins);
    TPIC_ALLOC(w.roffsets, bins + 1);
    TPIC_ALLOC(w.cursor, bins);
    TPIC_ALLOC(w.large, bins + 1);
    TPIC_ALLOC(w.failure, 1);
    const size_t blocks = (size_t(w.capacity) + 63) / 64;
    TPIC_ALLOC(w.block_base, blocks + 1);
    TPIC_ALLOC(w.bin_base, bins);
    TPIC_ALLOC(w.bin_lower, bins);
    w.scan_scratch_words = scan_scratch_words(std::max({bins, size_t(w.tiles) + 1, blocks + 1}));
    TPIC_ALLOC(w.scan_scratch, w.scan_scratch_words);
  }
  if (side) {
    // Default priority, like the context stream.
    if ((rc = check_hip(c, hipStreamCreateWithFlags(&c.side_stream, hipStreamNonBlocking), "side stream")) ||
        (rc = check_hip(c, hipEventCreateWithFlags(&c.fork, hipEventDisableTiming), "fork event")) ||
        (rc = check_hip(c, hipEventCreateWithFlags(&c.join, hipEventDisableTiming), "join event")))
      return rc;
  }
  GmresWorkspace& gm = c.gmres;
  gm.count = unsigned(3 * nodes);
  gm.blocks = unsigned(std::max<size_t>(1, std::min<size_t>((gm.count + 255) / 256, size_t(c.cus > 0 ? c.cus : 256) * 4)));
  const size_t m = size_t(c.num.restart), batch = std::max<size_t>(m, 2);
  TPIC_ALLOC(gm.basis, (m + 1) * gm.count);
  TPIC_ALLOC(gm.x, gm.count);
  TPIC_ALLOC(gm.b, gm.count);
  TPIC_ALLOC(gm.r, gm.count);
  TPIC_ALLOC(gm.image, gm.count);
  TPIC_ALLOC(gm.values, batch);
  TPIC_ALLOC(gm.partials, batch * gm.blocks);
  TPIC_ALLOC(gm.state, m * (m + 1) + 4 * m + 6);  // Hessenberg, rotations, rhs, coefficients, 5 stop words
  TPIC_ALLOC(gm.curl_cells, 3 * cells);
#undef TPIC_ALLOC
  if ((rc = check_hip(c, hipHostMalloc(reinterpret_cast<void**>(&gm.controls), 8 * sizeof(double)), "pinned")))
    return rc;
  if ((rc = check_hip(c, hipHostMalloc(reinterpret_cast<void**>(&c.host_word), sizeof(unsigned)), "pinned")))
    return rc;
  if ((rc = check_hip(c, hipHostMalloc(reinterpret_cast<void**>(&c.host_status), kMaxSpecies * sizeof(uint32_t)),
                      "pinned")))
    return rc;
  if ((rc = check_hip(c, hipHostMalloc(reinterpret_cast<void**>(&c.host_scalars),
                                       diag_rows * sizeof(double)),
                      "pinned")))
    return rc;
  // Zero every state owner once.
  hipStream_t st = c.stream;
  for (auto& sp : c.species) {
    for (auto& s : sp.slot) {
      (void)hipMemsetAsync(s.cell, 0, std::max<size_t>(sp.count, 1) * sizeof(int32_t), st);
      (void)hipMemsetAsync(s.rx, 0, std::max<size_t>(sp.count, 1) * sizeof(double), st);
    }
    (void)hipMemsetAsync(sp.counts, 0, bins * sizeof(uint32_t), st);
    (void)hipMemsetAsync(sp.offsets, 0, (bins + 1) * sizeof(uint32_t), st);
  }
  for (double* f : {c.e, c.bn, c.current, c.etheta, c.rhs, c.smooth_a, c.smooth_b, c.eparticle})
    if (f) (void)hipMemsetAsync(f, 0, 3 * nodes * sizeof(double), st);
  (void)hipMemsetAsync(c.b, 0, 3 * cells * sizeof(double), st);
  (void)hipMemsetAsync(c.mass, 0, size_t(kEntries) * g.vstride * sizeof(double), st);
  (void)hipMemsetAsync(c.cellbuf, 0, size_t(kCellRecord) * cells * sizeof(double), st);
  (void)hipMemsetAsync(c.validation, 0, sizeof(unsigned), st);
  return check_hip(c, hipStreamSynchronize(st), "initial clear");
}

bool valid_species(const Context& c, int s) { return s >= 0 && size_t(s) < c.species.size(); }

size_t field_size(const Grid& g, int which) {
  switch (which) {
    case TPIC_MI_FIELD_B: return 3 * size_t(g.cells());
    case TPIC_MI_FIELD_MASS: return size_t(g.vertices()) * kEntries;
    default: return 3 * size_t(g.nodes());
  }
}

double* field_pointer(Context& c, int which) {
  switch (which) {
    case TPIC_MI_FIELD_E: return c.e;
    case TPIC_MI_FIELD_B: return c.b;
    case TPIC_MI_FIELD_BNODE: return c.bn;
    case TPIC_MI_FIELD_J: return c.current;
    case TPIC_MI_FIELD_ETHETA: return c.etheta;
    case TPIC_MI_FIELD_RHS: return c.rhs;
    case TPIC_MI_FIELD_MASS: return static_cast<double*>(c.mass);
    case TPIC_MI_FIELD_EPARTICLE: return c.num.smoothing ? c.eparticle : c.etheta;
    default: return nullptr;
  }
}

}  // namespace

extern "C" {

int tpic_mi_create(const tpic_mi_config* config, tpic_mi_context** out) {
  if (!out) return TPIC_MI_INVALID_ARGUMENT;
  *out = nullptr;
  if (!config) {
    g_create_error = "tpic_mi_create: null config";
    return TPIC_MI_INVALID_ARGUMENT;
  }
  const tpic_mi_config& k = *config;
  auto reject = [](const char* why) {
    g_create_error = std::string("tpic_mi_create: ") + why;
    return TPIC_MI_INVALID_ARGUMENT;
  };
  for (int a = 0; a < 3; ++a) {
    if (k.cells[a] <= 0) return reject("cells must be positive");
    if (!(std::isfinite(k.lengths[a]) && k.lengths[a] > 0)) return reject("lengths must be finite and positive");
  }
  // physics.cpp:33-39: dt > 0, c > 0, 0.5 <= th <= 1, 0 < GMREStol < 1.
  if (!(k.dt > 0 && std::isfinite(k.dt))) return reject("dt must be finite and positive");
  if (!(k.c > 0 && std::isfinite(k.c))) return reject("c must be finite and positive");
  if (!(k.theta >= 0.5 && k.theta <= 1.0)) return reject("th must lie in [0.5, 1]");
  if (!(k.gmres_tolerance > 0 && k.gmres_tolerance < 1)) return reject("GMREStol must lie in (0, 1)");
  if (k.gmres_restart <= 0 || k.gmres_restart > 64) return reject("GMRES restart must lie in [1, 64]");
  if (k.gmres_max_iterations <= 0) return reject("GMRES iterations must be positive");
  if (k.mass_kernel != TPIC_MI_MASS_FP64 && k.mass_kernel != TPIC_MI_MASS_FP32 &&
      k.mass_kernel != TPIC_MI_MASS_BF16X3)
    return reject("mass_kernel must be TPIC_MI_MASS_FP64, TPIC_MI_MASS_FP32 or TPIC_MI_MASS_BF16X3");
  if (k.species <= 0 || k.species > TPIC_MI_MAX_SPECIES) return reject("species count out of range");
  for (int a = 0; a < 3; ++a) {
    const int lo = k.field_faces[a][0], hi = k.field_faces[a][1];
    for (int f : {lo, hi})
      if (f != TPIC_MI_FACE_PERIODIC && f != TPIC_MI_FACE_PEC && f != TPIC_MI_FACE_PMC)
        return reject("field faces must be TPIC_MI_FACE_PERIODIC, TPIC_MI_FACE_PEC or TPIC_MI_FACE_PMC");
    if ((lo == TPIC_MI_FACE_PERIODIC) != (hi == TPIC_MI_FACE_PERIODIC))
      return reject("an axis is periodic on both faces or has a wall on both");
    // domain.cpp:74-78: a wall axis needs at least two cells.
    if (lo != TPIC_MI_FACE_PERIODIC && k.cells[a] < 2) return reject("a wall axis needs at least 2 cells");
  }
  // physics.cpp:65-71 and node_smoothing_device.hpp:56-60.
  if (k.smooth_operator != TPIC_MI_SMOOTH_FACE7 && k.smooth_operator != TPIC_MI_SMOOTH_BINOMIAL27)
    return reject("smooth_operator must be TPIC_MI_SMOOTH_FACE7 or TPIC_MI_SMOOTH_BINOMIAL27");
  if (k.smooth_iterations < 0) return reject("energy-conserving smoothing requires SmoothNiter >= 0");
  if (!(std::isfinite(k.smooth_alpha) && k.smooth_alpha >= 0 && k.smooth_alpha <= 1))
    return reject("energy-conserving smoothing requires 0 <= Smooth <= 1");
  const uint64_t nodes = uint64_t(k.cells[0] + 1) * (k.cells[1] + 1) * (k.cells[2] + 1);
  const uint64_t bins = uint64_t(k.cells[0] + 4) * (k.cells[1] + 4) * (k.cells[2] + 4);
  const uint64_t verts = uint64_t(k.cells[0] + 3) * (k.cells[1] + 3) * (k.cells[2] + 3);
  if (3 * nodes > 0x7fffffffull || bins > 0x7fffffffull || verts * kEntries > 0xffffffffull * 8)
    return reject("grid too large for 32-bit indices");
  for (int s = 0; s < k.species; ++s) {
    if (!(std::isfinite(k.species_list[s].q_over_m) && k.species_list[s].q_over_m != 0))
      return reject("q/m must be finite and nonzero");
    if (k.species_list[s].count >= 0xffffff00u) return reject("species count exceeds 32-bit indices");
  }
  tpic_mi_context* h = new (std::nothrow) tpic_mi_context();
  if (!h) return TPIC_MI_BACKEND;
  Context& c = h->ctx;
  c.owner = h;
  Grid& g = c.grid;
  g.nx = k.cells[0];
  g.ny = k.cells[1];
  g.nz = k.cells[2];
  // domain.cpp:132-146: spacing = L / n; V = (double)((long double)dx*dy*dz).
  g.dx = k.lengths[0] / k.cells[0];
  g.dy = k.lengths[1] / k.cells[1];
  g.dz = k.lengths[2] / k.cells[2];
  g.ix = 1.0 / g.dx;
  g.iy = 1.0 / g.dy;
  g.iz = 1.0 / g.dz;
  for (int a = 0; a < 3; ++a) g.particle_periodic[a] = k.particle_periodic[a] ? 1 : 0;
  g.vstride = (g.vertices() + 63u) & ~63u;
  // Closed walls (config/boundary_conditions.hpp:68-77; field_boundary.hpp:
  // 43-80): PEC holds tangential E at zero, ghost signs E tangential odd, E
  // normal even, B tangential even, B normal odd; PMC the reverse.
  g.walls = 0;
  g.b_flip = 0;
  for (int a = 0; a < 3; ++a) {
    g.field_wall[a] = k.field_faces[a][0] != TPIC_MI_FACE_PERIODIC ? 1 : 0;
    g.walls |= g.field_wall[a];
    for (int side = 0; side < 2; ++side) {
      const int face = 2 * a + side;
      const bool pec = k.field_faces[a][side] == TPIC_MI_FACE_PEC;
      g.constrained[face] = 0;
      for (int comp = 0; comp < 3; ++comp) {
        const bool normal = comp == a;
        g.b_sign[comp][face] = (pec ? normal : !normal) ? -1 : 1;
        if (g.b_sign[comp][face] < 0) g.b_flip |= 1u << (6 * comp + face);
        if (g.field_wall[a] && (pec ? !normal : normal)) g.constrained[face] |= (unsigned char)(1u << comp);
      }
    }
  }
  c.num = Numerics{k.dt, k.theta, k.c, k.gmres_tolerance, k.gmres_restart, k.gmres_max_iterations, k.mass_kernel,
                   false, k.smooth_operator, k.smooth_iterations, k.smooth_alpha, 0.0};
  // physics.cpp:70: smoothing_enabled = SmoothNiter > 0 && Smooth < 1; Face7's
  // face weight (1 - alpha) / 6 is formed once (node_smoothing_device.hpp:70-73).
  c.num.smoothing = k.smooth_iterations > 0 && k.smooth_alpha < 1;
  c.num.smooth_beta = (1.0 - k.smooth_alpha) / 6.0;
  const long double volume = (long double)g.dx * g.dy * g.dz;
  c.inverse_volume = 1.0 / double(volume);
  // physics.cpp:35-36; maxwell_device.hpp:502-510.
  const double theta_dt = k.theta * k.dt;
  c.delta = k.c * theta_dt;
  c.current_factor = 4 * 3.14159265358979323846 * theta_dt;
  c.species.resize(size_t(k.species));
  for (int s = 0; s < k.species; ++s) {
    c.species[size_t(s)].q_over_m = k.species_list[s].q_over_m;
    c.species[size_t(s)].count = k.species_list[s].count;
  }
  if (hipGetDevice(&c.device) != hipSuccess) {
    g_create_error = "tpic_mi_create: no HIP device";
    delete h;
    return TPIC_MI_BACKEND;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c.device) != hipSuccess) {
    g_create_error = "tpic_mi_create: device query failed";
    delete h;
    return TPIC_MI_BACKEND;
  }
  c.cus = prop.multiProcessorCount;
  if (k.stream) {
    c.stream = reinterpret_cast<hipStream_t>(k.stream);
  } else {
    if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) {
      g_create_error = "tpic_mi_create: stream creation failed";
      delete h;
      return TPIC_MI_BACKEND;
    }
    c.own_stream = true;
  }
  const int rc = allocate(h);
  if (rc) {
    g_create_error = c.error;
    free_all(h);
    delete h;
    return rc;
  }
  *out = h;
  return TPIC_MI_OK;
}

int tpic_mi_destroy(tpic_mi_context* h) {
  if (!h) return TPIC_MI_INVALID_ARGUMENT;
  free_all(h);
  delete h;
  return TPIC_MI_OK;
}

const char* tpic_mi_last_error(const tpic_mi_context* h) {
  return h ? h->ctx.error.c_str() : g_create_error.c_str();
}

int tpic_mi_synchronize(tpic_mi_context* h) {
  if (!h) return TPIC_MI_INVALID_ARGUMENT;
  return check_hip(h->ctx, hipStreamSynchronize(h->ctx.stream), "synchronize");
}

int tpic_mi_device_field(tpic_mi_context* h, int which, void** out) {
  if (!h || !out) return TPIC_MI_INVALID_ARGUMENT;
  double* p = field_pointer(h->ctx, which);
  if (!p) return fail(h->ctx, TPIC_MI_INVALID_ARGUMENT, "unknown field");
  *out = p;
  return TPIC_MI_OK;
}

int tpic_mi_device_particles(tpic_mi_context* h, int s, int attribute, void** out) {
  if (!h || !out || !valid_species(h->ctx, s)) return TPIC_MI_INVALID_ARGUMENT;
  const ParticleView v = h->ctx.species[size_t(s)].view();
  void* table[9] = {v.rx, v.ry, v.rz, v.u, v.v, v.w, v.q, v.cell, v.id};
  if (attribute < 0 || attribute > 8) return fail(h->ctx, TPIC_MI_INVALID_ARGUMENT, "unknown attribute");
  *out = table[attribute];
  return TPIC_MI_OK;
}

int tpic_mi_set_particles(tpic_mi_context* h, int s, const tpic_mi_particle_arrays* a) {
  if (!h || !a || !valid_species(h->ctx, s)) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = h->ctx;
  SpeciesState& sp = c.species[size_t(s)];
  const ParticleView v = sp.view();
  const size_t n = sp.count;
  if (n == 0) return TPIC_MI_OK;
  const void* src[9] = {a->rx, a->ry, a->rz, a->u, a->v, a->w, a->q, a->cell, a->id};
  void* dst[9] = {v.rx, v.ry, v.rz, v.u, v.v, v.w, v.q, v.cell, v.id};
  const size_t bytes[9] = {8, 8, 8, 8, 8, 8, 8, 4, 8};
  for (int i = 0; i < 9; ++i) {
    if (!src[i]) return fail(c, TPIC_MI_INVALID_ARGUMENT, "null particle attribute");
    const int rc = check_hip(c, hipMemcpyAsync(dst[i], src[i], n * bytes[i], hipMemcpyHostToDevice, c.stream),
                             "particle upload");
    if (rc) return rc;
  }
  sp.sorted = false;
  return check_hip(c, hipStreamSynchronize(c.stream), "particle upload");
}

int tpic_mi_get_particles(tpic_mi_context* h, int s, tpic_mi_particle_arrays* a) {
  if (!h || !a || !valid_species(h->ctx, s)) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = h->ctx;
  SpeciesState& sp = c.species[size_t(s)];
  const ParticleView v = sp.view();
  const size_t n = sp.count;
  if (n == 0) return TPIC_MI_OK;
  void* dst[9] = {a->rx, a->ry, a->rz, a->u, a->v, a->w, a->q, a->cell, a->id};
  const void* src[9] = {v.rx, v.ry, v.rz, v.u, v.v, v.w, v.q, v.cell, v.id};
  const size_t bytes[9] = {8, 8, 8, 8, 8, 8, 8, 4, 8};
  for (int i = 0; i < 9; ++i) {
    if (!dst[i]) continue;
    const int rc = check_hip(c, hipMemcpyAsync(dst[i], src[i], n * bytes[i], hipMemcpyDeviceToHost, c.stream),
                             "particle download");
    if (rc) return rc;
  }
  return check_hip(c, hipStreamSynchronize(c.stream), "particle download");
}

int tpic_mi_load_maxwellian(tpic_mi_context* h, int s, const tpic_mi_load* l) {
  if (!h || !l || !valid_species(h->ctx, s)) return TPIC_MI_INVALID_ARGUMENT;
  for (int a = 0; a < 3; ++a)
    if (l->ppc[a] <= 0) return fail(h->ctx, TPIC_MI_INVALID_ARGUMENT, "ppc must be positive");
  return launch_load(h->ctx, s, l->ppc, l->rho, l->drift, l->thermal, l->seed, l->first_id);
}

int tpic_mi_load_uniform(tpic_mi_context* h, int s, const tpic_mi_uniform_load* l) {
  if (!h || !l || !valid_species(h->ctx, s)) return TPIC_MI_INVALID_ARGUMENT;
  for (int a = 0; a < 3; ++a)
    if (l->ppc[a] <= 0) return fail(h->ctx, TPIC_MI_INVALID_ARGUMENT, "ppc must be positive");
  if (l->rank < 0 || l->max_species <= s) return fail(h->ctx, TPIC_MI_INVALID_ARGUMENT, "rank or max_species");
  // prepared_initialization.cpp:46-47.
  const uint64_t seed = 12345ULL + 100ULL * (uint64_t(l->rank) * uint64_t(l->max_species) + uint64_t(s));
  return launch_load_uniform(h->ctx, s, l->ppc, l->rho_init, l->drift, l->thermal, l->quiet_start, seed,
                             l->first_id, nullptr, nullptr);
}

int tpic_mi_load_density_seeded(tpic_mi_context* h, int s, const tpic_mi_uniform_load* l,
                                const tpic_mi_density_seed* seed_mode) {
  if (!h || !l || !seed_mode || !valid_species(h->ctx, s)) return TPIC_MI_INVALID_ARGUMENT;
  for (int a = 0; a < 3; ++a)
    if (l->ppc[a] <= 0) return fail(h->ctx, TPIC_MI_INVALID_ARGUMENT, "ppc must be positive");
  if (l->rank < 0 || l->max_species <= s) return fail(h->ctx, TPIC_MI_INVALID_ARGUMENT, "rank or max_species");
  if (!std::isfinite(seed_mode->wavenumber) || !std::isfinite(seed_mode->phase) ||
      !(std::fabs(seed_mode->amplitude) < 1.0))
    return fail(h->ctx, TPIC_MI_INVALID_ARGUMENT, "density seed: finite k and phase, |amplitude| < 1");
  const uint64_t seed = 12345ULL + 100ULL * (uint64_t(l->rank) * uint64_t(l->max_species) + uint64_t(s));
  const double mode[3] = {seed_mode->wavenumber, seed_mode->phase, seed_mode->amplitude};
  return launch_load_uniform(h->ctx, s, l->ppc, l->rho_init, l->drift, l->thermal, l->quiet_start, seed,
                             l->first_id, mode, nullptr);
}

int tpic_mi_load_harris(tpic_mi_context* h, int s, const tpic_mi_uniform_load* l, const tpic_mi_harris_profile* p) {
  if (!h || !l || !p || !valid_species(h->ctx, s)) return TPIC_MI_INVALID_ARGUMENT;
  for (int a = 0; a < 3; ++a)
    if (l->ppc[a] <= 0) return fail(h->ctx, TPIC_MI_INVALID_ARGUMENT, "ppc must be positive");
  if (l->rank < 0 || l->max_species <= s) return fail(h->ctx, TPIC_MI_INVALID_ARGUMENT, "rank or max_species");
  if (!(std::isfinite(p->delta) && p->delta > 0) || !(std::isfinite(p->ly) && p->ly > 0))
    return fail(h->ctx, TPIC_MI_INVALID_ARGUMENT, "Harris profile: delta and ly must be finite and positive");
  HarrisLoad load{};
  load.density = p->harris_density ? 1 : 0;
  load.is_double = p->is_double ? 1 : 0;
  load.varying_thermal = p->varying_thermal ? 1 : 0;
  load.ly = p->ly;
  load.delta = p->delta;
  load.ampere_drift = p->ampere_drift;
  load.center_x = p->curl_center[0];
  load.center_y = p->curl_center[1];
  load.curl_delta = p->curl_delta;
  load.sheet = p->curl_sheet;
  load.mode = p->curl_mode;
  load.mode_kx = p->curl_mode_k[0];
  load.mode_ky = p->curl_mode_k[1];
  load.hump = p->curl_hump;
  load.hump_kx = p->curl_hump_k[0];
  load.hump_ky = p->curl_hump_k[1];
  load.hump_inverse_wx2 = p->curl_hump_inverse_w2[0];
  load.hump_inverse_wy2 = p->curl_hump_inverse_w2[1];
  const uint64_t seed = 12345ULL + 100ULL * (uint64_t(l->rank) * uint64_t(l->max_species) + uint64_t(s));
  return launch_load_uniform(h->ctx, s, l->ppc, l->rho_init, l->drift, l->thermal, l->quiet_start, seed,
                             l->first_id, nullptr, &load);
}

int tpic_mi_center_velocity(tpic_mi_context* h, int s, const double* drift) {
  if (!h || !drift || !valid_species(h->ctx, s)) return TPIC_MI_INVALID_ARGUMENT;
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(drift[a])) return fail(h->ctx, TPIC_MI_INVALID_ARGUMENT, "drift must be finite");
  return launch_center_velocity(h->ctx, s, drift);
}

int tpic_mi_set_field(tpic_mi_context* h, int which, const double* host) {
  if (!h || !host) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = h->ctx;
  if (which == TPIC_MI_FIELD_MASS) return fail(c, TPIC_MI_INVALID_ARGUMENT, "the mass matrix is deposited, not set");
  double* p = field_pointer(c, which);
  if (!p) return fail(c, TPIC_MI_INVALID_ARGUMENT, "unknown field");
  int rc = check_hip(c, hipMemcpyAsync(p, host, field_size(c.grid, which) * sizeof(double), hipMemcpyHostToDevice,
                                       c.stream),
                     "field upload");
  if (rc) return rc;
  return check_hip(c, hipStreamSynchronize(c.stream), "field upload");
}

int tpic_mi_get_field(tpic_mi_context* h, int which, double* host) {
  if (!h || !host) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = h->ctx;
  const size_t n = field_size(c.grid, which);
  int rc;
  if (which == TPIC_MI_FIELD_MASS) {
    // The entry-major device matrix is downloaded as it is and reordered
    // into the reference's node-major records on the host (no device staging).
    const Grid& g = c.grid;
    const size_t words = size_t(kEntries) * g.vstride;
    std::vector<double> wide;
    std::vector<float> narrow;
    if (c.mass_single) narrow.resize(words);
    else wide.resize(words);
    void* staging = c.mass_single ? static_cast<void*>(narrow.data()) : static_cast<void*>(wide.data());
    rc = check_hip(c, hipMemcpyAsync(staging, c.mass, words * (c.mass_single ? sizeof(float) : sizeof(double)),
                                     hipMemcpyDeviceToHost, c.stream),
                   "mass download");
    if (!rc) rc = check_hip(c, hipStreamSynchronize(c.stream), "mass download");
    if (rc) return rc;
    const size_t verts = g.vertices();
    for (size_t v = 0; v < verts; ++v)
      for (size_t e = 0; e < size_t(kEntries); ++e)
        host[v * kEntries + e] = c.mass_single ? double(narrow[e * g.vstride + v]) : wide[e * g.vstride + v];
    return TPIC_MI_OK;
  }
  double* p = field_pointer(c, which);
  if (!p) return fail(c, TPIC_MI_INVALID_ARGUMENT, "unknown field");
  rc = check_hip(c, hipMemcpyAsync(host, p, n * sizeof(double), hipMemcpyDeviceToHost, c.stream), "field download");
  if (rc) return rc;
  return check_hip(c, hipStreamSynchronize(c.stream), "field download");
}

int tpic_mi_record_permutation(tpic_mi_context* h, int enable) {
  if (!h) return TPIC_MI_INVALID_ARGUMENT;
  h->record_permutation = enable != 0;
  return TPIC_MI_OK;
}

int tpic_mi_get_bins(tpic_mi_context* h, int s, uint32_t* counts, uint32_t* offsets, uint32_t* permutation) {
  if (!h || !valid_species(h->ctx, s)) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = h->ctx;
  const SpeciesState& sp = c.species[size_t(s)];
  const size_t bins = c.grid.bins();
  int rc = 0;
  if (counts && (rc = check_hip(c, hipMemcpyAsync(counts, sp.counts, bins * 4, hipMemcpyDeviceToHost, c.stream),
                                "bins download")))
    return rc;
  if (offsets && (rc = check_hip(c, hipMemcpyAsync(offsets, sp.offsets, (bins + 1) * 4, hipMemcpyDeviceToHost,
                                                   c.stream),
                                 "bins download")))
    return rc;
  if (permutation) {
    if (!h->record_permutation) return fail(c, TPIC_MI_INVALID_STATE, "permutation recording is off");
    if (sp.count &&
        (rc = check_hip(c, hipMemcpyAsync(permutation, c.sort.permutation, size_t(sp.count) * 4,
                                          hipMemcpyDeviceToHost, c.stream),
                        "permutation download")))
      return rc;
  }
  return check_hip(c, hipStreamSynchronize(c.stream), "bins download");
}

int tpic_mi_sort(tpic_mi_context* h, int s) {
  if (!h || !valid_species(h->ctx, s)) return TPIC_MI_INVALID_ARGUMENT;
  uint32_t* saved = h->ctx.sort.permutation;
  if (!h->record_permutation) h->ctx.sort.permutation = nullptr;
  const int rc = launch_sort(h->ctx, s);
  h->ctx.sort.permutation = saved;
  return rc;
}

int tpic_mi_move(tpic_mi_context* h, int s) {
  if (!h || !valid_species(h->ctx, s)) return TPIC_MI_INVALID_ARGUMENT;
  return launch_move(h->ctx, s);
}

int tpic_mi_deposit(tpic_mi_context* h, int first, int last, int accumulate) {
  if (!h) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = h->ctx;
  if (first < 0 || last > int(c.species.size()) || first >= last)
    return fail(c, TPIC_MI_INVALID_ARGUMENT, "species range");
  if (last - first > kMaxSpecies) return fail(c, TPIC_MI_INVALID_ARGUMENT, "too many species");
  return launch_deposit(c, first, last, accumulate != 0);
}

int tpic_mi_assemble(tpic_mi_context* h) {
  if (!h) return TPIC_MI_INVALID_ARGUMENT;
  return launch_assemble(h->ctx);
}

int tpic_mi_initialize(tpic_mi_context* h) {
  if (!h) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = h->ctx;
  // The committed field keeps its constrained rows at zero.
  int rc = launch_zero_constrained(c, c.e);
  if (rc) return rc;
  rc = launch_cell_to_node(c);
  if (rc) return rc;
  for (size_t s = 0; s < c.species.size(); ++s)
    if ((rc = tpic_mi_sort(h, int(s)))) return rc;
  if ((rc = launch_deposit(c, 0, int(c.species.size()), false))) return rc;
  return launch_assemble(c);
}

int tpic_mi_solve(tpic_mi_context* h, tpic_mi_solve_result* r) {
  if (!h) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = h->ctx;
  int rc = launch_rhs(c);
  if (rc) return rc;
  // The initial guess is E^n projected (enqueue_rhs_fields copies E into
  // theta and projects it, maxwell_device.hpp:140-142): a field set with
  // nonzero constrained rows still starts the solve in the constrained subspace.
  if ((rc = check_hip(c, hipMemcpyAsync(c.etheta, c.e, 3 * size_t(c.grid.nodes()) * sizeof(double),
                                        hipMemcpyDeviceToDevice, c.stream),
                      "theta initial copy")))
    return rc;
  if ((rc = launch_zero_constrained(c, c.etheta))) return rc;
  tpic_mi_solve_result out{};
  rc = run_gmres(c, &out.converged, &out.iterations, &out.initial_residual, &out.residual, &out.rhs_norm,
                 &out.relative_residual);
  if (r) *r = out;
  if (rc) return rc;
  if (!out.converged) return fail(c, TPIC_MI_NOT_CONVERGED, "GMRES did not converge");
  // complete_particle_field (maxwell_device.hpp:213-228): the movers gather
  // S(E_theta).
  if (c.num.smoothing) {
    const double* smoothed = nullptr;
    if ((rc = launch_smooth(c, c.etheta, c.eparticle, &smoothed))) return rc;
  }
  return TPIC_MI_OK;
}

int tpic_mi_apply_operator(tpic_mi_context* h, const double* in, double* out) {
  if (!h || !in || !out || in == out) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = h->ctx;
  if (!c.grid.walls) return launch_operator(c, in, out);
  // begin_apply_input (maxwell_device.hpp:270-272) projects the input; the
  // solver's own vectors already have zero constrained rows.
  int rc = check_hip(c, hipMemcpyAsync(c.projected, in, 3 * size_t(c.grid.nodes()) * sizeof(double),
                                       hipMemcpyDeviceToDevice, c.stream),
                     "operator input copy");
  if (!rc) rc = launch_zero_constrained(c, c.projected);
  if (!rc) rc = launch_operator(c, c.projected, out);
  return rc;
}

int tpic_mi_smooth(tpic_mi_context* h, const double* in, double* out) {
  if (!h || !in || !out) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = h->ctx;
  if (!c.num.smoothing) return fail(c, TPIC_MI_INVALID_ARGUMENT, "smoothing is not active");
  const double* smoothed = nullptr;
  // A pass does not run in place: with in == out the result stays in a pass
  // buffer and is copied.
  int rc = launch_smooth(c, in, in == out ? nullptr : out, &smoothed);
  if (rc || smoothed == out) return rc;
  return check_hip(c, hipMemcpyAsync(out, smoothed, 3 * size_t(c.grid.nodes()) * sizeof(double),
                                     hipMemcpyDeviceToDevice, c.stream),
                   "smoothing copy");
}

int tpic_mi_advance_fields(tpic_mi_context* h) {
  if (!h) return TPIC_MI_INVALID_ARGUMENT;
  return launch_next_fields(h->ctx);
}

int tpic_mi_cycle(tpic_mi_context* h, int cycles, tpic_mi_solve_result* last) {
  if (!h || cycles < 0) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = h->ctx;
  const int ns = int(c.species.size());
  for (int n = 0; n < cycles; ++n) {
    int rc = tpic_mi_solve(h, last);
    if (rc) return rc;
    // Move, advance, sort (the reference's order); the sort's record copy is
    // fused into the deposition, whose sums are those of the separate calls.
    // With two or more species (and no permutation record), species 0's sort
    // plan runs on the side stream while the other species move: the plan
    // from the crossers is light in HBM traffic, the mover bound by it.
    // A species sorted before its move gets the incremental plan (from the
    // crossers the block mover records); a species whose incremental plan
    // reports a bin it cannot order, or an inconsistent previous plan, then
    // gets the run-based plan (one host read of the plan words per cycle).
    uint32_t* saved = c.sort.permutation;
    if (!h->record_permutation) c.sort.permutation = nullptr;
    const bool fork = c.side_stream && !h->record_permutation;
    bool incremental[kMaxSpecies] = {};
    bool any_incremental = false;
    for (int s = 0; s < ns; ++s) {
      incremental[s] = !h->record_permutation && c.species[size_t(s)].sorted && c.species[size_t(s)].count > 0;
      any_incremental = any_incremental || incremental[s];
    }
    auto plan = [&](int s, SortWorkspace& ws, hipStream_t st) {
      return incremental[s] ? launch_sort_plan_incremental(c, s, ws, st) : launch_sort_plan(c, s, ws, st);
    };
    for (int s = 0; s < ns && !rc; ++s) {
      rc = launch_move(c, s, incremental[s]);
      if (!rc && s == 0 && fork) {
        if (!(rc = check_hip(c, hipEventRecord(c.fork, c.stream), "fork")) &&
            !(rc = check_hip(c, hipStreamWaitEvent(c.side_stream, c.fork, 0), "fork")) &&
            !(rc = plan(0, c.sort_side, c.side_stream)))
          rc = check_hip(c, hipEventRecord(c.join, c.side_stream), "join");
      }
    }
    if (!rc) rc = launch_next_fields(c);
    for (int s = fork ? 1 : 0; s < ns && !rc; ++s) rc = plan(s, c.sort, c.stream);
    if (!rc && fork) rc = check_hip(c, hipStreamWaitEvent(c.stream, c.join, 0), "join");
    if (!rc && any_incremental) {
      rc = check_hip(c, hipMemcpyAsync(c.host_status, c.plan_status, size_t(ns) * sizeof(uint32_t),
                                       hipMemcpyDeviceToHost, c.stream),
                     "plan status");
      if (!rc) rc = check_hip(c, hipStreamSynchronize(c.stream), "plan status");
      for (int s = 0; s < ns && !rc; ++s)
        if (incremental[s] && c.host_status[s] != 0) rc = launch_sort_plan(c, s);
    }
    if (!rc) rc = launch_gather_deposit(c);
    c.sort.permutation = saved;
    if (rc) return rc;
    if ((rc = launch_assemble(c))) return rc;
  }
  return TPIC_MI_OK;
}

int tpic_mi_scalars(tpic_mi_context* h, double* out) {
  if (!h || !out) return TPIC_MI_INVALID_ARGUMENT;
  return launch_scalars(h->ctx, out);
}

int tpic_mi_moments(tpic_mi_context* h, int species, unsigned selected, double* device_out) {
  if (!h) return TPIC_MI_INVALID_ARGUMENT;
  return launch_moments(h->ctx, species, selected, device_out);
}

int tpic_mi_energy(tpic_mi_context* h, double* out) {
  if (!h || !out) return TPIC_MI_INVALID_ARGUMENT;
  return launch_energy(h->ctx, out);
}

int tpic_mi_read_validation(tpic_mi_context* h, unsigned* flags) {
  if (!h || !flags) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = h->ctx;
  int rc = check_hip(c, hipMemcpyAsync(c.host_word, c.validation, sizeof(unsigned), hipMemcpyDeviceToHost, c.stream),
                     "validation readback");
  if (rc) return rc;
  rc = check_hip(c, hipStreamSynchronize(c.stream), "validation readback");
  if (rc) return rc;
  *flags = *c.host_word;
  return TPIC_MI_OK;
}

int tpic_mi_clear_validation(tpic_mi_context* h) {
  if (!h) return TPIC_MI_INVALID_ARGUMENT;
  return check_hip(h->ctx, hipMemsetAsync(h->ctx.validation, 0, sizeof(unsigned), h->ctx.stream), "validation clear");
}

int tpic_mi_enable_timing(tpic_mi_context* h, int enable) {
  if (!h) return TPIC_MI_INVALID_ARGUMENT;
  h->ctx.timing = enable != 0;
  return TPIC_MI_OK;
}

int tpic_mi_phase_time(tpic_mi_context* h, int phase, double* total_ms, uint64_t* calls) {
  if (!h || phase < 0 || phase >= int(Phase::count)) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = h->ctx;
  int rc = check_hip(c, hipStreamSynchronize(c.stream), "timing");
  if (rc) return rc;
  for (int p = 0; p < int(Phase::count); ++p) {
    for (auto& pr : h->pending[p]) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
        c.phase_ms[p] += double(ms);
        c.phase_calls[p] += 1;
      }
      h->pool.push_back(pr.first);
      h->pool.push_back(pr.second);
    }
    h->pending[p].clear();
  }
  if (total_ms) *total_ms = c.phase_ms[phase];
  if (calls) *calls = c.phase_calls[phase];
  return TPIC_MI_OK;
}

int tpic_mi_reset_timing(tpic_mi_context* h) {
  if (!h) return TPIC_MI_INVALID_ARGUMENT;
  double t;
  uint64_t n;
  const int rc = tpic_mi_phase_time(h, 0, &t, &n);
  for (int p = 0; p < int(Phase::count); ++p) {
    h->ctx.phase_ms[p] = 0;
    h->ctx.phase_calls[p] = 0;
  }
  return rc;
}

}  // extern "C"
