// Output files of the tpic_mi_run driver (diagnostics.csv, field snapshots).
#include "output.hpp"

#include <array>
#include <bit>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <iomanip>
#include <limits>
#include <locale>
#include <sstream>
#include <stdexcept>
#include <string_view>
#include <system_error>

#include <cstdio>
#include <iostream>

#include <fcntl.h>
#include <unistd.h>

namespace tpic_mi::output {
namespace {

[[noreturn]] void output_error(const std::string& context, const char* operation, const std::string& path,
                               int error) {
  throw std::system_error(error == 0 ? EIO : error, std::generic_category(),
                          "TPIC " + context + ": " + operation + " " + path);
}

// Exclusive creation unless overwrite (then truncation).
std::FILE* create(const std::string& path, bool overwrite, const std::string& context) {
  const int flags = O_WRONLY | O_CREAT | (overwrite ? O_TRUNC : O_EXCL);
  const int fd = ::open(path.c_str(), flags, 0644);
  if (fd < 0) output_error(context, "create", path, errno);
  std::FILE* f = ::fdopen(fd, "wb");
  if (!f) {
    const int e = errno;
    ::close(fd);
    output_error(context, "open", path, e);
  }
  return f;
}

void write_all(std::FILE* f, const void* data, size_t bytes, const std::string& context, const std::string& path) {
  if (bytes && std::fwrite(data, 1, bytes, f) != bytes) output_error(context, "write", path, errno);
}

// The cycle padded to six digits (output_file.hpp:20-27).
std::string padded(int cycle) {
  auto digits = std::to_string(cycle);
  return std::string(digits.size() < 6 ? 6 - digits.size() : 0, '0') + digits;
}

void append(std::string& out, double v) {
  char buf[64];
  const auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::general,
                               std::numeric_limits<double>::max_digits10);
  out.append(buf, r.ptr);
}

}  // namespace

Precision precision_metadata(bool mixed_mass) {
  const char* mass = mixed_mass ? "float32" : "float64";
  return Precision{"double/double",
                   std::string("storage: position=float64 velocity=float64 weight=float64 E=float64 B=float64 "
                               "J=float64 mass=") +
                       mass +
                       " cell=int32 ID=uint64; compute: position=float64 velocity=float64 E=float64 B=float64 "
                       "J=float64 mass=" +
                       mass + " solver-reduction=float64 reduction=float64 output-moment-accumulation=float64"};
}

void ensure_directory(const std::string& dir) {
  std::error_code ec;
  std::filesystem::create_directories(dir, ec);
  if (ec) throw std::system_error(ec, "TPIC output: create directory " + dir);
}

DiagnosticsCsv::~DiagnosticsCsv() {
  if (file_) std::fclose(file_);
}

// The length of the part of a diagnostics file a restart at cycle keeps:
// everything before the first row whose cycle is at least cycle, and before an
// unterminated last line; -1 when the file has no header line.
long kept_bytes(const std::string& text, int cycle) {
  size_t at = 0;
  bool header = false;
  while (at < text.size()) {
    const size_t end = text.find('\n', at);
    if (end == std::string::npos) break;
    const std::string_view line(text.data() + at, end - at);
    if (line.substr(0, 6) == "cycle,") {
      header = true;
    } else if (!line.empty() && line[0] != '#') {
      int row = 0;
      const auto r = std::from_chars(line.data(), line.data() + line.size(), row);
      if (r.ec != std::errc{} || row >= cycle) break;
    }
    at = end + 1;
  }
  return header ? long(at) : -1;
}

void DiagnosticsCsv::open(const std::string& dir, size_t species, const Precision& precision, bool overwrite,
                          int restart_cycle) {
  path_ = (std::filesystem::path(dir) / "diagnostics.csv").string();
  species_ = species;
  if (restart_cycle >= 0) {
    std::string text;
    if (std::FILE* old = std::fopen(path_.c_str(), "rb")) {
      char buf[65536];
      for (size_t got; (got = std::fread(buf, 1, sizeof buf, old)) > 0;) text.append(buf, got);
      std::fclose(old);
      const long kept = kept_bytes(text, restart_cycle);
      if (kept >= 0) {
        std::error_code ec;
        std::filesystem::resize_file(path_, uintmax_t(kept), ec);
        if (ec) throw std::system_error(ec, "TPIC diagnostics output: truncate " + path_);
        file_ = std::fopen(path_.c_str(), "ab");
        if (!file_) output_error("diagnostics output", "open", path_, errno);
        last_cycle_ = restart_cycle - 1;
        last_time_ = -1.0;
        return;
      }
    }
    overwrite = true;
  }
  file_ = create(path_, overwrite, "diagnostics output");
  std::string text = "# precision_profile = " + precision.name + "\n# precision_roles = " + precision.roles + "\n";
  text +=
      "# TPIC scalar diagnostics v1\n"
      "# Values are globally reduced in double precision with no visualization rescaling.\n"
      "# electric_energy = cell_volume/(8*pi) * sum_unique_nodes(dual_volume_weight * |E_node|^2).\n"
      "# Node ownership excludes upper shared/periodic duplicates; each physical boundary axis contributes a "
      "factor 1/2.\n"
      "# magnetic_energy = cell_volume/(8*pi) * sum_owned_cells(|B_cell|^2).\n"
      "# Species kinetic_energy = sum((q/qom)*|v|^2/2); charge = sum(q); momentum = sum((q/qom)*v).\n"
      "# state=pre_solve identifies the state before the cycle's solve/push; state=final identifies the completed "
      "run state.\n"
      "# time is physical simulation time at the recorded state; total_energy_change = total_energy - the total "
      "energy of the run's first state, kept across restarts (absolute units).\n"
      "cycle,time,state,electric_energy,magnetic_energy,kinetic_energy,total_energy,total_energy_change,charge,"
      "momentum_x,momentum_y,momentum_z";
  const char* names[5] = {"kinetic_energy", "charge", "momentum_x", "momentum_y", "momentum_z"};
  for (size_t s = 0; s < species; ++s)
    for (const char* n : names) text += std::string(",") + n + "_" + std::to_string(s);
  text += "\n";
  write_all(file_, text.data(), text.size(), "diagnostics output", path_);
  std::fflush(file_);
}

void DiagnosticsCsv::write(int cycle, double time, bool final_state, const std::vector<double>& sample,
                           double initial_energy) {
  if (sample.size() != 2 + 5 * species_)
    throw std::invalid_argument("TPIC diagnostics: sample has a different species count");
  if (cycle < 0 || cycle <= last_cycle_ || !std::isfinite(time) || time < 0.0 || (last_cycle_ >= 0 && time <= last_time_))
    throw std::invalid_argument("TPIC diagnostics: scalar CSV cycles and times must increase");
  double totals[5] = {0, 0, 0, 0, 0};
  for (size_t s = 0; s < species_; ++s)
    for (int r = 0; r < 5; ++r) totals[r] += sample[2 + 5 * s + size_t(r)];
  const double total = sample[0] + sample[1] + totals[0];
  const double change = total - initial_energy;
  for (double v : totals)
    if (!std::isfinite(v)) throw std::overflow_error("TPIC diagnostics: species total is nonfinite");
  if (!std::isfinite(total) || !std::isfinite(change))
    throw std::overflow_error("TPIC diagnostics: total energy or change is nonfinite");
  std::string row = std::to_string(cycle) + ",";
  append(row, time);
  row += final_state ? ",final" : ",pre_solve";
  auto column = [&row](double v) {
    row += ",";
    append(row, v);
  };
  column(sample[0]);
  column(sample[1]);
  column(totals[0]);
  column(total);
  column(change);
  for (int r = 1; r < 5; ++r) column(totals[r]);
  for (size_t i = 2; i < sample.size(); ++i) column(sample[i]);
  row += "\n";
  write_all(file_, row.data(), row.size(), "diagnostics output", path_);
  std::fflush(file_);
  last_cycle_ = cycle;
  last_time_ = time;
  ++rows_;
}

namespace {

// The FieldData block every VTK file carries (vtk_output.hpp:27-47).
void field_data(std::ostringstream& xml, int cycle, double time, bool final_state, const Precision& precision) {
  xml << "    <FieldData>\n"
      << "      <DataArray type=\"Float64\" Name=\"TimeValue\" NumberOfTuples=\"1\" format=\"ascii\">" << time
      << "</DataArray>\n"
      << "      <DataArray type=\"Int32\" Name=\"cycle\" NumberOfTuples=\"1\" format=\"ascii\">" << cycle
      << "</DataArray>\n"
      << "      <DataArray type=\"UInt8\" Name=\"final_state\" NumberOfTuples=\"1\" format=\"ascii\">"
      << (final_state ? 1 : 0) << "</DataArray>\n";
  auto string_array = [&xml](const std::string& key, const std::string& value) {
    xml << "      <Array type=\"String\" Name=\"" << key << "\" NumberOfTuples=\"1\" format=\"ascii\">";
    for (const unsigned char ch : value) xml << unsigned(ch) << ' ';
    xml << "0</Array>\n";
  };
  string_array("state", final_state ? "final" : "pre_solve");
  string_array("precision_profile", precision.name);
  string_array("precision_roles", precision.roles);
  xml << "    </FieldData>\n";
}

std::ostringstream xml_stream() {
  std::ostringstream xml;
  xml.imbue(std::locale::classic());
  xml << std::setprecision(std::numeric_limits<double>::max_digits10);
  return xml;
}

const char* byte_order() { return std::endian::native == std::endian::little ? "LittleEndian" : "BigEndian"; }

}  // namespace

std::string write_particles(const std::string& dir, int cycle, double time, bool final_state,
                            const Precision& precision, const int cells[3], const double spacing[3],
                            const std::vector<SpeciesParticles>& species, bool overwrite) {
  uint64_t count = 0;
  for (const auto& s : species) count += s.id.size();
  const uint64_t scalar_bytes = count * sizeof(uint64_t);
  // Positions first: a bad record throws before any file is created.
  std::vector<double> position(3 * count), velocity(3 * count), charge(count);
  const int ex = cells[0] + 4, ey = cells[1] + 4;
  const auto invalid = [] {
    throw std::invalid_argument("TPIC particle output: invalid particle coordinate or nonfinite value");
  };
  uint64_t at = 0;
  for (const auto& s : species) {
    for (size_t p = 0; p < s.id.size(); ++p, ++at) {
      const int32_t key = s.cell[p];
      const int cell[3] = {key % ex - 2, (key / ex) % ey - 2, key / ex / ey - 2};
      if (key < 0) invalid();
      const double frac[3] = {s.rx[p], s.ry[p], s.rz[p]};
      const double vel[3] = {s.u[p], s.v[p], s.w[p]};
      for (int a = 0; a < 3; ++a) {
        if (cell[a] < 0 || cell[a] >= cells[a]) invalid();
        if (!(frac[a] >= 0 && frac[a] < 1) || !std::isfinite(vel[a])) invalid();
        position[3 * at + size_t(a)] = (0.0 + double(cell[a]) + frac[a]) * spacing[a];
        velocity[3 * at + size_t(a)] = vel[a];
      }
      if (!std::isfinite(s.q[p])) invalid();
      charge[at] = s.q[p];
    }
  }
  auto xml = xml_stream();
  xml << "<?xml version=\"1.0\"?>\n<VTKFile type=\"PolyData\" version=\"1.0\" byte_order=\"" << byte_order()
      << "\" header_type=\"UInt64\">\n  <PolyData>\n";
  field_data(xml, cycle, time, final_state, precision);
  xml << "    <Piece NumberOfPoints=\"" << count << "\" NumberOfVerts=\"" << count
      << "\" NumberOfLines=\"0\" NumberOfStrips=\"0\" NumberOfPolys=\"0\">\n      <PointData>\n";
  uint64_t offset = 0;
  const auto array = [&](const char* type, const char* name, int components) {
    xml << "        <DataArray type=\"" << type << "\" Name=\"" << name << "\" NumberOfComponents=\"" << components
        << "\" format=\"appended\" offset=\"" << offset << "\"/>\n";
    offset += sizeof(uint64_t) + scalar_bytes * uint64_t(components);
  };
  array("Float64", "velocity", 3);
  array("Float64", "charge", 1);
  array("UInt64", "id", 1);
  array("UInt64", "species", 1);
  xml << "      </PointData>\n      <CellData/>\n      <Points>\n";
  array("Float64", "position", 3);
  xml << "      </Points>\n      <Verts>\n";
  array("Int64", "connectivity", 1);
  array("Int64", "offsets", 1);
  xml << "      </Verts>\n    </Piece>\n  </PolyData>\n  <AppendedData encoding=\"raw\">\n_";
  const std::string head = xml.str();
  const std::string stem = "particles_t" + padded(cycle);
  const std::string piece = stem + "_r" + padded(0) + ".vtp";
  const std::string path = (std::filesystem::path(dir) / piece).string();
  std::FILE* f = create(path, overwrite, "particle output");
  try {
    const auto value = [&](uint64_t v) { write_all(f, &v, sizeof v, "particle output", path); };
    write_all(f, head.data(), head.size(), "particle output", path);
    value(3 * scalar_bytes);
    write_all(f, velocity.data(), 3 * scalar_bytes, "particle output", path);
    value(scalar_bytes);
    write_all(f, charge.data(), scalar_bytes, "particle output", path);
    value(scalar_bytes);
    for (const auto& s : species) write_all(f, s.id.data(), s.id.size() * sizeof(uint64_t), "particle output", path);
    value(scalar_bytes);
    for (size_t s = 0; s < species.size(); ++s) {
      const std::vector<uint64_t> tag(species[s].id.size(), uint64_t(s));
      write_all(f, tag.data(), tag.size() * sizeof(uint64_t), "particle output", path);
    }
    value(3 * scalar_bytes);
    write_all(f, position.data(), 3 * scalar_bytes, "particle output", path);
    for (uint64_t first : {uint64_t{0}, uint64_t{1}}) {
      value(scalar_bytes);
      std::vector<uint64_t> index(count);
      for (uint64_t i = 0; i < count; ++i) index[i] = i + first;
      write_all(f, index.data(), scalar_bytes, "particle output", path);
    }
    static const char close_text[] = "\n  </AppendedData>\n</VTKFile>\n";
    write_all(f, close_text, sizeof close_text - 1, "particle output", path);
  } catch (...) {
    std::fclose(f);
    throw;
  }
  if (std::fclose(f) != 0) output_error("particle output", "close", path, errno);

  // The manifest (VtkParticleEngine::publish, vtk_particle_engine.cpp:51-71).
  auto manifest = xml_stream();
  manifest << "<?xml version=\"1.0\"?>\n<VTKFile type=\"PPolyData\" version=\"1.0\" byte_order=\"" << byte_order()
           << "\" header_type=\"UInt64\">\n  <PPolyData GhostLevel=\"0\">\n";
  field_data(manifest, cycle, time, final_state, precision);
  const auto parallel = [&manifest](const char* type, const char* name, int components) {
    manifest << "      <PDataArray type=\"" << type << "\" Name=\"" << name << "\" NumberOfComponents=\"" << components
             << "\"/>\n";
  };
  manifest << "    <PPointData>\n";
  parallel("Float64", "velocity", 3);
  parallel("Float64", "charge", 1);
  parallel("UInt64", "id", 1);
  parallel("UInt64", "species", 1);
  manifest << "    </PPointData>\n    <PCellData/>\n    <PPoints>\n";
  parallel("Float64", "position", 3);
  manifest << "    </PPoints>\n    <Piece Source=\"" << piece << "\"/>\n  </PPolyData>\n</VTKFile>\n";
  const std::string text = manifest.str();
  const std::string manifest_path = (std::filesystem::path(dir) / (stem + ".pvtp")).string();
  std::FILE* m = create(manifest_path, overwrite, "particle output");
  try {
    write_all(m, text.data(), text.size(), "particle output", manifest_path);
  } catch (...) {
    std::fclose(m);
    throw;
  }
  if (std::fclose(m) != 0) output_error("particle output", "close", manifest_path, errno);
  return manifest_path;
}

std::string write_vti(const std::string& dir, const std::string& name, int cycle, double time, bool final_state,
                      const Precision& precision, const int cells[3], const double spacing[3], bool cell,
                      const std::vector<double>& soa, bool overwrite) {
  const size_t ext[3] = {size_t(cells[0]) + (cell ? 0 : 1), size_t(cells[1]) + (cell ? 0 : 1),
                         size_t(cells[2]) + (cell ? 0 : 1)};
  const size_t points = ext[0] * ext[1] * ext[2];
  if (soa.size() != 3 * points) throw std::invalid_argument("TPIC field output: array size does not match the grid");
  auto xml = xml_stream();
  const std::string whole = "0 " + std::to_string(cells[0]) + " 0 " + std::to_string(cells[1]) + " 0 " +
                            std::to_string(cells[2]);
  xml << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"" << byte_order()
      << "\" header_type=\"UInt64\">\n"
      << "  <ImageData WholeExtent=\"" << whole << "\" Origin=\"0 0 0\" Spacing=\"" << spacing[0] << ' '
      << spacing[1] << ' ' << spacing[2] << "\">\n";
  field_data(xml, cycle, time, final_state, precision);
  xml << "    <Piece Extent=\"" << whole << "\">\n"
      << (cell ? "      <PointData/>\n      <CellData>\n" : "      <PointData>\n")
      << "        <DataArray type=\"Float64\" Name=\"" << name
      << "\" NumberOfComponents=\"3\" format=\"appended\" offset=\"0\"/>\n"
      << (cell ? "      </CellData>\n" : "      </PointData>\n      <CellData/>\n")
      << "    </Piece>\n  </ImageData>\n  <AppendedData encoding=\"raw\">\n_";
  const std::string head = xml.str();
  std::vector<double> interleaved(3 * points);
  for (size_t n = 0; n < points; ++n)
    for (size_t c = 0; c < 3; ++c) interleaved[3 * n + c] = soa[c * points + n];
  const uint64_t bytes = uint64_t(interleaved.size()) * sizeof(double);
  const std::string path = (std::filesystem::path(dir) / (name + "_t" + padded(cycle) + ".vti")).string();
  std::FILE* f = create(path, overwrite, "field output");
  try {
    write_all(f, head.data(), head.size(), "field output", path);
    write_all(f, &bytes, sizeof bytes, "field output", path);
    write_all(f, interleaved.data(), bytes, "field output", path);
    static const char close_text[] = "\n  </AppendedData>\n</VTKFile>\n";
    write_all(f, close_text, sizeof close_text - 1, "field output", path);
  } catch (...) {
    std::fclose(f);
    throw;
  }
  if (std::fclose(f) != 0) output_error("field output", "close", path, errno);
  return path;
}

std::string checkpoint_path(const std::string& dir, int cycle) {
  return (std::filesystem::path(dir) / ("checkpoint_t" + std::to_string(cycle) + "_n1.tpicmi")).string();
}

namespace {

// The cycle of a complete restart file name checkpoint_t<cycle>_n1.tpicmi.
bool checkpoint_cycle(const std::string& name, int& cycle) {
  const std::string prefix = "checkpoint_t", suffix = "_n1.tpicmi";
  if (name.size() <= prefix.size() + suffix.size() || name.compare(0, prefix.size(), prefix) != 0 ||
      name.compare(name.size() - suffix.size(), suffix.size(), suffix) != 0)
    return false;
  const char* first = name.data() + prefix.size();
  const char* last = name.data() + name.size() - suffix.size();
  const auto r = std::from_chars(first, last, cycle);
  return r.ec == std::errc{} && r.ptr == last && cycle >= 0;
}

}  // namespace

std::string newest_checkpoint(const std::string& dir, int* cycle) {
  std::string best;
  int best_cycle = -1;
  std::error_code ec;
  for (std::filesystem::directory_iterator it(dir, ec), end; !ec && it != end; it.increment(ec)) {
    int c = 0;
    if (it->is_regular_file() && checkpoint_cycle(it->path().filename().string(), c) && c > best_cycle) {
      best_cycle = c;
      best = it->path().string();
    }
  }
  if (ec && ec != std::errc::no_such_file_or_directory)
    throw std::system_error(ec, "TPIC checkpoint: cannot list " + dir);
  if (cycle) *cycle = best_cycle;
  return best;
}

void remove_older_checkpoints(const std::string& dir, int cycle) {
  try {
    std::vector<std::filesystem::path> old;
    for (const auto& entry : std::filesystem::directory_iterator(dir)) {
      int c = 0;
      if (entry.is_regular_file() && checkpoint_cycle(entry.path().filename().string(), c) && c < cycle)
        old.push_back(entry.path());
    }
    for (const auto& path : old) std::filesystem::remove(path);
  } catch (const std::exception& e) {
    std::cerr << "TPIC checkpoint: warning: old checkpoints were not removed: " << e.what() << '\n';
  }
}

}  // namespace tpic_mi::output
