// Cycle checkpoints: in-device copies of the state a cycle starts from, and
// restart files of that state.
//
// A cycle starts from sorted particles (with their bins), E at nodes and B at
// cells; B at nodes, the deposition and the assembly follow from them by the
// kernels a cycle ends with (launch_cell_to_node, launch_deposit,
// launch_assemble), so they are recomputed rather than stored and a run
// continued from a checkpoint repeats the original bit for bit.
//
// The restart file holds the fields of the reference's CheckpointHeader
// (include/tpic/data/output/checkpoint.hpp:14-60: cycle, species, process
// grid, global cells, lengths, dt, c, the first state's energy, faces,
// storage bytes, q/m and particle count per species), then E at the nodes
// (x fastest, the reference's checkpoint E piece at the local nodes with both
// ends of every axis), B at the cells, and every species' particles in stored
// order (rx, ry, rz, u, v, w, q, cell, id as arrays), raw little-endian. The
// reference writes ADIOS2 or HDF5, neither of which this image has.
#include "../../include/tpic_mi.h"
#include "internal.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace tpic_mi {

Context* context_of(tpic_mi_context* h);

namespace {

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

// The stored state after tpic_mi_checkpoint_restore or a file read:
// B at nodes, deposition and assembly of the sorted particles.
int refresh(Context& c) {
  int rc = launch_cell_to_node(c);
  if (rc) return rc;
  for (size_t s = 0; s < c.species.size(); ++s)
    if (!c.species[s].sorted) return fail(c, TPIC_MI_INVALID_STATE, "a checkpoint holds an unsorted species");
  if ((rc = launch_deposit(c, 0, int(c.species.size()), false))) return rc;
  return launch_assemble(c);
}

constexpr char kMagic[8] = {'T', 'P', 'I', 'C', 'M', 'I', 'C', 'K'};
constexpr uint32_t kVersion = 1;

// The magic, the version and the header of a restart file.
bool read_header(std::FILE* f, tpic_mi_checkpoint_header& hd) {
  char magic[8];
  uint32_t version = 0;
  return std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
         std::fread(&version, 1, 4, f) == 4 && version == kVersion && std::fread(&hd, 1, sizeof hd, f) == sizeof hd;
}

}  // namespace

void release_checkpoints(Context& c) {
  for (auto& cp : c.checkpoints) {
    for (auto& v : cp.particles) {
      device_free(v.rx); device_free(v.ry); device_free(v.rz); device_free(v.u); device_free(v.v);
      device_free(v.w); device_free(v.q); device_free(v.cell); device_free(v.id);
    }
    device_free(cp.counts);
    device_free(cp.offsets);
    device_free(cp.e);
    device_free(cp.b);
  }
  c.checkpoints.clear();
}

}  // namespace tpic_mi

using namespace tpic_mi;

extern "C" {

int tpic_mi_checkpoint_reserve(tpic_mi_context* h, int slots) {
  if (!h || slots < 0) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = *context_of(h);
  const size_t have = c.checkpoints.size();
  if (size_t(slots) <= have) return TPIC_MI_OK;
  c.checkpoints.resize(size_t(slots));
  const size_t nodes = c.grid.nodes(), cells = c.grid.cells(), bins = c.grid.bins(), ns = c.species.size();
  for (size_t k = have; k < size_t(slots); ++k) {
    auto& cp = c.checkpoints[k];
    cp.particles.resize(ns);
    cp.sorted.assign(ns, false);
    int rc = 0;
    for (size_t s = 0; s < ns; ++s) {
      ParticleView& v = cp.particles[s];
      const size_t n = c.species[s].count;
      if ((rc = device_alloc(c, &v.rx, n, "checkpoint")) || (rc = device_alloc(c, &v.ry, n, "checkpoint")) ||
          (rc = device_alloc(c, &v.rz, n, "checkpoint")) || (rc = device_alloc(c, &v.u, n, "checkpoint")) ||
          (rc = device_alloc(c, &v.v, n, "checkpoint")) || (rc = device_alloc(c, &v.w, n, "checkpoint")) ||
          (rc = device_alloc(c, &v.q, n, "checkpoint")) || (rc = device_alloc(c, &v.cell, n, "checkpoint")) ||
          (rc = device_alloc(c, &v.id, n, "checkpoint")))
        break;
    }
    if (!rc && !(rc = device_alloc(c, &cp.counts, ns * bins, "checkpoint")) &&
        !(rc = device_alloc(c, &cp.offsets, ns * (bins + 1), "checkpoint")) &&
        !(rc = device_alloc(c, &cp.e, 3 * nodes, "checkpoint")))
      rc = device_alloc(c, &cp.b, 3 * cells, "checkpoint");
    if (rc) {
      release_checkpoints(c);
      return rc;
    }
  }
  return TPIC_MI_OK;
}

int tpic_mi_checkpoint_save(tpic_mi_context* h, int slot) {
  if (!h) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = *context_of(h);
  if (slot < 0 || size_t(slot) >= c.checkpoints.size())
    return fail(c, TPIC_MI_INVALID_ARGUMENT, "checkpoint slot out of range (reserve slots first)");
  auto& cp = c.checkpoints[size_t(slot)];
  const size_t nodes = c.grid.nodes(), cells = c.grid.cells(), bins = c.grid.bins(), ns = c.species.size();
  auto copy = [&](void* dst, const void* src, size_t bytes) {
    return check_hip(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c.stream), "checkpoint save");
  };
  int rc = 0;
  for (size_t s = 0; s < ns && !rc; ++s) {
    const ParticleView from = c.species[s].view();
    const ParticleView& to = cp.particles[s];
    const size_t n = c.species[s].count;
    if ((rc = copy(to.rx, from.rx, 8 * n)) || (rc = copy(to.ry, from.ry, 8 * n)) || (rc = copy(to.rz, from.rz, 8 * n)) ||
        (rc = copy(to.u, from.u, 8 * n)) || (rc = copy(to.v, from.v, 8 * n)) || (rc = copy(to.w, from.w, 8 * n)) ||
        (rc = copy(to.q, from.q, 8 * n)) || (rc = copy(to.cell, from.cell, 4 * n)) || (rc = copy(to.id, from.id, 8 * n)))
      break;
    cp.sorted[s] = c.species[s].sorted;
  }
  if (!rc) rc = copy(cp.counts, c.bin_counts, 4 * ns * bins);
  if (!rc) rc = copy(cp.offsets, c.bin_offsets, 4 * ns * (bins + 1));
  if (!rc) rc = copy(cp.e, c.e, 24 * nodes);
  if (!rc) rc = copy(cp.b, c.b, 24 * cells);
  if (!rc) cp.saved = true;
  return rc;
}

int tpic_mi_checkpoint_restore(tpic_mi_context* h, int slot) {
  if (!h) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = *context_of(h);
  if (slot < 0 || size_t(slot) >= c.checkpoints.size() || !c.checkpoints[size_t(slot)].saved)
    return fail(c, TPIC_MI_INVALID_ARGUMENT, "checkpoint slot out of range or never saved");
  auto& cp = c.checkpoints[size_t(slot)];
  const size_t nodes = c.grid.nodes(), cells = c.grid.cells(), bins = c.grid.bins(), ns = c.species.size();
  auto copy = [&](void* dst, const void* src, size_t bytes) {
    return check_hip(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c.stream), "checkpoint restore");
  };
  int rc = 0;
  for (size_t s = 0; s < ns && !rc; ++s) {
    const ParticleView to = c.species[s].view();
    const ParticleView& from = cp.particles[s];
    const size_t n = c.species[s].count;
    if ((rc = copy(to.rx, from.rx, 8 * n)) || (rc = copy(to.ry, from.ry, 8 * n)) || (rc = copy(to.rz, from.rz, 8 * n)) ||
        (rc = copy(to.u, from.u, 8 * n)) || (rc = copy(to.v, from.v, 8 * n)) || (rc = copy(to.w, from.w, 8 * n)) ||
        (rc = copy(to.q, from.q, 8 * n)) || (rc = copy(to.cell, from.cell, 4 * n)) || (rc = copy(to.id, from.id, 8 * n)))
      break;
    c.species[s].sorted = cp.sorted[s];
  }
  if (!rc) rc = copy(c.bin_counts, cp.counts, 4 * ns * bins);
  if (!rc) rc = copy(c.bin_offsets, cp.offsets, 4 * ns * (bins + 1));
  if (!rc) rc = copy(c.e, cp.e, 24 * nodes);
  if (!rc) rc = copy(c.b, cp.b, 24 * cells);
  if (rc) return rc;
  return refresh(c);
}


int tpic_mi_checkpoint_write(tpic_mi_context* h, const char* path, const tpic_mi_checkpoint_header* header) {
  if (!h || !path || !header) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = *context_of(h);
  const Grid& g = c.grid;
  tpic_mi_checkpoint_header hd = *header;
  // The context's own facts override the caller's.
  hd.species = int32_t(c.species.size());
  for (int a = 0; a < 3; ++a) hd.processes[a] = 1;
  hd.global_cells[0] = g.nx;
  hd.global_cells[1] = g.ny;
  hd.global_cells[2] = g.nz;
  hd.lengths[0] = g.dx * g.nx;
  hd.lengths[1] = g.dy * g.ny;
  hd.lengths[2] = g.dz * g.nz;
  hd.dt = c.num.dt;
  hd.c = c.num.c;
  for (int k = 0; k < 5; ++k) hd.storage_bytes[k] = 8;
  for (int s = 0; s < TPIC_MI_CHECKPOINT_MAX_SPECIES; ++s) {
    hd.q_over_m[s] = size_t(s) < c.species.size() ? c.species[size_t(s)].q_over_m : 0.0;
    hd.global_count[s] = size_t(s) < c.species.size() ? c.species[size_t(s)].count : 0;
  }
  const size_t nodes = g.nodes(), cells = g.cells();
  std::vector<double> field(3 * std::max(nodes, cells));
  std::unique_ptr<std::FILE, int (*)(std::FILE*)> f(std::fopen(path, "wb"), std::fclose);
  if (!f) return fail(c, TPIC_MI_INVALID_ARGUMENT, std::string("cannot create checkpoint file ") + path);
  auto put = [&](const void* p, size_t bytes) { return std::fwrite(p, 1, bytes, f.get()) == bytes; };
  bool ok = put(kMagic, 8) && put(&kVersion, 4) && put(&hd, sizeof hd);
  int rc = 0;
  for (int which : {0, 1}) {
    const size_t n = which == 0 ? 3 * nodes : 3 * cells;
    rc = check_hip(c, hipMemcpyAsync(field.data(), which == 0 ? c.e : c.b, n * 8, hipMemcpyDeviceToHost, c.stream),
                   "checkpoint write");
    if (!rc) rc = check_hip(c, hipStreamSynchronize(c.stream), "checkpoint write");
    if (rc) return rc;
    ok = ok && put(field.data(), n * 8);
  }
  for (size_t s = 0; s < c.species.size() && ok; ++s) {
    const ParticleView v = c.species[s].view();
    const size_t n = c.species[s].count;
    const void* src[9] = {v.rx, v.ry, v.rz, v.u, v.v, v.w, v.q, v.cell, v.id};
    const size_t bytes[9] = {8, 8, 8, 8, 8, 8, 8, 4, 8};
    std::vector<unsigned char> buffer(8 * std::max<size_t>(n, 1));
    for (int k = 0; k < 9 && ok; ++k) {
      rc = check_hip(c, hipMemcpyAsync(buffer.data(), src[k], n * bytes[k], hipMemcpyDeviceToHost, c.stream),
                     "checkpoint write");
      if (!rc) rc = check_hip(c, hipStreamSynchronize(c.stream), "checkpoint write");
      if (rc) return rc;
      ok = put(buffer.data(), n * bytes[k]);
    }
  }
  if (!ok) return fail(c, TPIC_MI_BACKEND, std::string("checkpoint write failed: ") + path);
  return TPIC_MI_OK;
}

int tpic_mi_checkpoint_read_header(const char* path, tpic_mi_checkpoint_header* header) {
  if (!path || !header) {
    set_context_free_error("tpic_mi_checkpoint_read_header: null path or header");
    return TPIC_MI_INVALID_ARGUMENT;
  }
  std::unique_ptr<std::FILE, int (*)(std::FILE*)> f(std::fopen(path, "rb"), std::fclose);
  if (!f) {
    set_context_free_error(std::string("cannot open checkpoint file ") + path);
    return TPIC_MI_INVALID_ARGUMENT;
  }
  if (!read_header(f.get(), *header)) {
    set_context_free_error(std::string("not a checkpoint of this build: ") + path);
    return TPIC_MI_INVALID_ARGUMENT;
  }
  return TPIC_MI_OK;
}

int tpic_mi_checkpoint_read(tpic_mi_context* h, const char* path, tpic_mi_checkpoint_header* header) {
  if (!h || !path) return TPIC_MI_INVALID_ARGUMENT;
  Context& c = *context_of(h);
  const Grid& g = c.grid;
  std::unique_ptr<std::FILE, int (*)(std::FILE*)> f(std::fopen(path, "rb"), std::fclose);
  if (!f) return fail(c, TPIC_MI_INVALID_ARGUMENT, std::string("cannot open checkpoint file ") + path);
  auto get = [&](void* p, size_t bytes) { return std::fread(p, 1, bytes, f.get()) == bytes; };
  tpic_mi_checkpoint_header hd{};
  if (!read_header(f.get(), hd))
    return fail(c, TPIC_MI_INVALID_ARGUMENT, std::string("not a checkpoint of this build: ") + path);
  // The file must describe this context's grid and species.
  bool same = hd.species == int32_t(c.species.size()) && hd.global_cells[0] == g.nx && hd.global_cells[1] == g.ny &&
              hd.global_cells[2] == g.nz && hd.storage_bytes[0] == 8;
  for (size_t s = 0; same && s < c.species.size(); ++s)
    same = hd.q_over_m[s] == c.species[s].q_over_m && hd.global_count[s] == c.species[s].count;
  if (!same) return fail(c, TPIC_MI_INVALID_ARGUMENT, "the checkpoint's grid or species differ from the context's");
  const size_t nodes = g.nodes(), cells = g.cells();
  std::vector<double> field(3 * std::max(nodes, cells));
  int rc = 0;
  for (int which : {0, 1}) {
    const size_t n = which == 0 ? 3 * nodes : 3 * cells;
    if (!get(field.data(), n * 8)) return fail(c, TPIC_MI_INVALID_ARGUMENT, "truncated checkpoint file");
    rc = check_hip(c, hipMemcpyAsync(which == 0 ? c.e : c.b, field.data(), n * 8, hipMemcpyHostToDevice, c.stream),
                   "checkpoint read");
    if (!rc) rc = check_hip(c, hipStreamSynchronize(c.stream), "checkpoint read");
    if (rc) return rc;
  }
  for (size_t s = 0; s < c.species.size(); ++s) {
    const ParticleView v = c.species[s].view();
    const size_t n = c.species[s].count;
    void* dst[9] = {v.rx, v.ry, v.rz, v.u, v.v, v.w, v.q, v.cell, v.id};
    const size_t bytes[9] = {8, 8, 8, 8, 8, 8, 8, 4, 8};
    std::vector<unsigned char> buffer(8 * std::max<size_t>(n, 1));
    for (int k = 0; k < 9; ++k) {
      if (!get(buffer.data(), n * bytes[k])) return fail(c, TPIC_MI_INVALID_ARGUMENT, "truncated checkpoint file");
      rc = check_hip(c, hipMemcpyAsync(dst[k], buffer.data(), n * bytes[k], hipMemcpyHostToDevice, c.stream),
                     "checkpoint read");
      if (!rc) rc = check_hip(c, hipStreamSynchronize(c.stream), "checkpoint read");
      if (rc) return rc;
    }
    // Stored after a cycle's sort; tpic_mi_initialize sorts (the stable
    // sort of a sorted species keeps its order) and deposits.
    c.species[s].sorted = false;
  }
  if (header) *header = hd;
  return TPIC_MI_OK;
}

}  // extern "C"
