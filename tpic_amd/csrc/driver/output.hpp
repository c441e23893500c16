// Output files of the tpic_mi_run driver, in the reference's formats:
//   diagnostics.csv - the scalar diagnostics rows (src/io/diagnostics_writer.cpp:52-212)
//   <name>_t<cycle>.vti - a field snapshot as VTK XML ImageData with one
//                         appended raw Float64 array (src/io/vtk_field_engine.cpp:22-99,
//                         include/tpic/io/detail/vtk_output.hpp)
//   particles_t<cycle>.pvtp + particles_t<cycle>_r000000.vtp - a particle
//                         snapshot (src/io/vtk_particle_engine.cpp)
// Files are created exclusively unless overwrite is set (--overwrite-output).
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace tpic_mi::output {

// The precision profile name and roles recorded in every file
// (config::precision_metadata, include/tpic/config/precision.hpp:295-325).
struct Precision {
  std::string name, roles;
};
// FP64 storage and compute; a mixed mass kernel stores and forms the mass
// matrix in FP32.
Precision precision_metadata(bool mixed_mass);

// Creates the directory (and parents) if missing.
void ensure_directory(const std::string& dir);

class DiagnosticsCsv {
 public:
  DiagnosticsCsv() = default;
  DiagnosticsCsv(const DiagnosticsCsv&) = delete;
  DiagnosticsCsv& operator=(const DiagnosticsCsv&) = delete;
  ~DiagnosticsCsv();
  // restart_cycle >= 0 continues an existing file: its rows from that cycle
  // on (and an unterminated last line) are dropped, later rows are appended
  // (detail::CycleCsvFile::open, src/io/diagnostics_writer.cpp:85-140).
  void open(const std::string& dir, size_t species, const Precision& precision, bool overwrite,
            int restart_cycle = -1);
  bool is_open() const { return file_ != nullptr; }
  // sample: W_E, W_B, then 5 scalars per species (tpic_mi_scalars).
  void write(int cycle, double time, bool final_state, const std::vector<double>& sample, double initial_energy);
  const std::string& path() const { return path_; }
  int rows() const { return rows_; }

 private:
  std::FILE* file_ = nullptr;
  std::string path_;
  size_t species_ = 0;
  int rows_ = 0, last_cycle_ = -1;
  double last_time_ = 0.0;
};

// One field snapshot: values in the C-ABI's SoA layout (component-major),
// written with components interleaved per point, x fastest. cell = false: E
// at the (n+1)^3 nodes (PointData); true: B at the n^3 cells (CellData).
// Returns the path written.
std::string write_vti(const std::string& dir, const std::string& name, int cycle, double time, bool final_state,
                      const Precision& precision, const int cells[3], const double spacing[3], bool cell,
                      const std::vector<double>& soa, bool overwrite);

// One species' particles in the C-ABI's SoA form (fractions and cell keys).
struct SpeciesParticles {
  std::vector<double> rx, ry, rz, u, v, w, q;
  std::vector<int32_t> cell;
  std::vector<uint64_t> id;
};

// One particle snapshot (src/io/vtk_particle_engine.cpp:22-185): the piece
// particles_t<cycle>_r000000.vtp as VTK XML PolyData (velocity, charge, id and
// species point arrays, positions (cell + fraction) * spacing, one vertex per
// particle; appended raw data, species consecutively in stored order) and its
// manifest particles_t<cycle>.pvtp. Throws on a key outside the interior, a
// fraction outside [0, 1) or a nonfinite value. Returns the manifest's path.
std::string write_particles(const std::string& dir, int cycle, double time, bool final_state,
                            const Precision& precision, const int cells[3], const double spacing[3],
                            const std::vector<SpeciesParticles>& species, bool overwrite);

// Restart files of one rank: <dir>/checkpoint_t<cycle>_n1.tpicmi, written
// under a temporary name and renamed when complete (the checkpoint_t<cycle>_n
// <ranks> naming of src/io/checkpoint.cpp:27-45).
std::string checkpoint_path(const std::string& dir, int cycle);
// The newest complete restart file of dir and its cycle; "" when it has none.
std::string newest_checkpoint(const std::string& dir, int* cycle);
// Removes the restart files of dir older than cycle (remove_old_checkpoints,
// src/io/checkpoint.cpp:88-108); a failure is a warning on stderr.
void remove_older_checkpoints(const std::string& dir, int cycle);

}  // namespace tpic_mi::output
