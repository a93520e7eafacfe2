// triumvirate_gpu.hpp -- C++ drop-in for Triumvirate's three-point API,
// implemented over the C ABI of include/trv_gpu.h (header only).
//
// The names, member layouts and function signatures follow the reference:
//   trv::ParticleData / trv::ParticleCatalogue  (include/particles.hpp:63-78)
//   trv::LineOfSight                             (include/dataobjs.hpp:186)
//   trv::Binning                                 (include/dataobjs.hpp:58)
//   trv::BispecMeasurements / ThreePCFMeasurements (include/dataobjs.hpp:249-283)
//   trv::ParameterSet (numerical subset + validate semantics,
//                      include/parameters.hpp:58-240)
//   trv::compute_bispec, compute_3pcf, compute_bispec_in_gpp_box,
//   compute_3pcf_in_gpp_box, calc_bispec_normalisation_from_particles,
//   calc_bispec_normalisation_from_mesh      (include/threept.hpp:138-294)
//   trv::compute_powspec*, compute_corrfunc*, the two-point normalisations
//                                            (include/twopt.hpp:62-320)
//   ParameterSet::read_from_file / print_to_file, ParticleCatalogue::
//   load_catalogue_file (plain text), trv::io measurement-file writers
//                                            (include/io.hpp:36-230)
// Errors are thrown as trv::sys::InvalidParameterError /
// trv::sys::InvalidDataError (include/monitor.hpp:731-779), or
// std::runtime_error for device failures.
#ifndef TRIUMVIRATE_GPU_HPP_INCLUDED
#define TRIUMVIRATE_GPU_HPP_INCLUDED

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <regex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <system_error>
#include <vector>

#include "trv_gpu.h"

namespace trv {

namespace sys {
class InvalidParameterError : public std::invalid_argument {
 public:
  explicit InvalidParameterError(const std::string& m) : std::invalid_argument(m) {}
};
class InvalidDataError : public std::runtime_error {
 public:
  explicit InvalidDataError(const std::string& m) : std::runtime_error(m) {}
};
// trv::sys::IOError (include/monitor.hpp:781-800).
class IOError : public std::runtime_error {
 public:
  explicit IOError(const std::string& m) : std::runtime_error(m) {}
};

// trv::sys::expand_envar_in_path (src/monitor.cpp:1012-1037): each
// "${NAME}" whose variable is set is replaced by its value; unset ones
// are left as they are.
inline void expand_envar_in_path(std::string& path) {
  static const std::regex pattern(R"(\$\{([^}]+)\})");
  std::string out;
  std::smatch m;
  auto it = path.cbegin();
  while (std::regex_search(it, path.cend(), m, pattern)) {
    out.append(it, m[0].first);
    const char* v = std::getenv(m[1].str().c_str());
    out += v ? std::string(v) : m[0].str();
    it = m[0].second;
  }
  out.append(it, path.cend());
  path = out;
}

inline void check_status(int st) {
  if (st == TRVG_OK) return;
  std::string msg = trvg_last_error();
  if (st == TRVG_ERR_PARAM) throw InvalidParameterError(msg);
  if (st == TRVG_ERR_DATA) throw InvalidDataError(msg);
  throw std::runtime_error("[trvg status " + std::to_string(st) + "] " + msg);
}
}  // namespace sys

struct ParticleData {
  double pos[3];
  double nz;
  double ws;
  double wc;
  double w;
};

struct LineOfSight {
  double pos[3];
};

class ParticleCatalogue {
 public:
  std::string source;
  ParticleData* pdata = nullptr;
  int ntotal = 0;
  double wtotal = 0.;
  double wstotal = 0.;
  double pos_min[3] = {0., 0., 0.};
  double pos_max[3] = {0., 0., 0.};
  double pos_span[3] = {0., 0., 0.};

  ParticleCatalogue() = default;
  ParticleCatalogue(const ParticleCatalogue&) = delete;
  ParticleCatalogue& operator=(const ParticleCatalogue&) = delete;
  ~ParticleCatalogue() { delete[] pdata; }

  ParticleData& operator[](int pid) { return pdata[pid]; }

  // ParticleCatalogue::load_particle_data (src/particles.cpp:512-565).
  int load_particle_data(std::vector<double> x, std::vector<double> y,
                         std::vector<double> z, std::vector<double> nz,
                         std::vector<double> ws, std::vector<double> wc) {
    source = "extdata";
    std::size_t n = x.size();
    if (!(n == y.size() && n == z.size() && n == nz.size() && n == ws.size() &&
          n == wc.size()))
      throw sys::InvalidDataError(
          "Inconsistent particle data dimensions (source=extdata).");
    delete[] pdata;
    ntotal = static_cast<int>(n);
    pdata = new ParticleData[n];
    for (std::size_t i = 0; i < n; i++) {
      pdata[i].pos[0] = x[i];
      pdata[i].pos[1] = y[i];
      pdata[i].pos[2] = z[i];
      pdata[i].nz = nz[i];
      pdata[i].ws = ws[i];
      pdata[i].wc = wc[i];
      pdata[i].w = ws[i] * wc[i];
    }
    calc_totals_and_extents();
    return 0;
  }

  // ParticleCatalogue::load_catalogue_file (src/particles.cpp:94-510) for
  // plain-text files: '#' and empty lines are skipped; `catalogue_columns`
  // names the file's columns, comma-separated, among which x, y, z are
  // required and nz, ws, wc optional (nz defaults to ntotal / volume when
  // volume > 0, else 0; weights to 1).  Several files may be joined with
  // "::".  HDF5 files raise, as in a reference build without HDF5.
  int load_catalogue_file(const std::string& catalogue_filepath,
                          const std::string& catalogue_columns,
                          const std::string& catalogue_dataset = "",
                          double volume = 0.) {
    (void)catalogue_dataset;   // names an HDF5 dataset only
    if (!source.empty())
      throw sys::InvalidDataError(
          "Catalogue already loaded from another source: " + source + ".");
    auto ends_with = [](const std::string& t, const std::string& e) {
      return t.size() >= e.size() && t.compare(t.size() - e.size(), e.size(), e) == 0;
    };
    if (ends_with(catalogue_filepath, ".h5") || ends_with(catalogue_filepath, ".hdf5"))
      throw sys::InvalidDataError(
          "HDF5 file format is not supported in this build: " + catalogue_filepath);
    source = "extfile:" + catalogue_filepath;

    const char* names[6] = {"x", "y", "z", "nz", "ws", "wc"};
    std::vector<std::string> colnames;
    {
      std::istringstream iss(catalogue_columns);
      std::string name;
      while (std::getline(iss, name, ',')) colnames.push_back(name);
    }
    int col[6];
    for (int k = 0; k < 6; k++) {
      auto it = std::find(colnames.begin(), colnames.end(), names[k]);
      col[k] = (it == colnames.end()) ? -1 : int(it - colnames.begin());
    }
    for (int k = 0; k < 3; k++)
      if (col[k] < 0)
        throw sys::InvalidDataError(std::string("Mandatory field ") + names[k] +
                                    " is missing.");
    std::vector<std::string> files;
    {
      std::string rest = catalogue_filepath;
      std::size_t pos;
      while ((pos = rest.find("::")) != std::string::npos) {
        if (pos) files.push_back(rest.substr(0, pos));
        rest.erase(0, pos + 2);
      }
      if (!rest.empty()) files.push_back(rest);
    }

    std::vector<std::vector<double>> rows;
    for (const std::string& f : files) {
      std::ifstream fin(f);
      if (!fin) throw sys::IOError("Failed to open file: " + f);
      std::string line;
      while (std::getline(fin, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ls(line);
        std::vector<double> row;
        double v;
        while (ls >> v) row.push_back(v);
        int need = 0;
        for (int k = 0; k < 6; k++) need = std::max(need, col[k] + 1);
        if (int(row.size()) < need)
          throw sys::InvalidDataError("Catalogue row has fewer columns than "
                                      "named in: " + f);
        rows.push_back(std::move(row));
      }
    }
    delete[] pdata;
    ntotal = static_cast<int>(rows.size());
    pdata = new ParticleData[rows.size()];
    double nz_default = (volume > 0.) ? ntotal / volume : 0.;
    for (int i = 0; i < ntotal; i++) {
      const std::vector<double>& r = rows[i];
      for (int a = 0; a < 3; a++) pdata[i].pos[a] = r[col[a]];
      pdata[i].nz = (col[3] >= 0) ? r[col[3]] : nz_default;
      pdata[i].ws = (col[4] >= 0) ? r[col[4]] : 1.;
      pdata[i].wc = (col[5] >= 0) ? r[col[5]] : 1.;
      pdata[i].w = pdata[i].ws * pdata[i].wc;
    }
    calc_totals_and_extents();
    return 0;
  }

  // ParticleCatalogue::calc_total_weights / calc_pos_extents
  // (src/particles.cpp:598-656).
  void calc_totals_and_extents() {
    wtotal = wstotal = 0.;
    for (int i = 0; i < ntotal; i++) {
      wtotal += pdata[i].w;
      wstotal += pdata[i].ws;
    }
    for (int a = 0; a < 3; a++) {
      pos_min[a] = pos_max[a] = ntotal ? pdata[0].pos[a] : 0.;
      for (int i = 0; i < ntotal; i++) {
        pos_min[a] = std::fmin(pos_min[a], pdata[i].pos[a]);
        pos_max[a] = std::fmax(pos_max[a], pdata[i].pos[a]);
      }
      pos_span[a] = pos_max[a] - pos_min[a];
    }
  }

  // ParticleCatalogue::offset_coords (src/particles.cpp:658-676).
  void offset_coords(const double dvec[3]) {
    for (int i = 0; i < ntotal; i++)
      for (int a = 0; a < 3; a++) pdata[i].pos[a] -= dvec[a];
    for (int a = 0; a < 3; a++) {
      pos_min[a] -= dvec[a];
      pos_max[a] -= dvec[a];
    }
  }

  // C-ABI view of the 56-byte records in place (stride 7 doubles).
  trvg_catalogue abi_view() const {
    if (pdata == nullptr)
      throw sys::InvalidDataError("Particle data are uninitialised.");
    trvg_catalogue v;
    v.ntotal = ntotal;
    v.stride = sizeof(ParticleData) / sizeof(double);
    v.x = &pdata[0].pos[0];
    v.y = &pdata[0].pos[1];
    v.z = &pdata[0].pos[2];
    v.nz = &pdata[0].nz;
    v.ws = &pdata[0].ws;
    v.wc = &pdata[0].wc;
    return v;
  }
};

// Numerical subset of trv::ParameterSet with the validation rules the
// three-point path depends on (src/parameters.cpp:466-1270).
class ParameterSet {
 public:
  // I/O (include/parameters.hpp:65-78) and the catalogue files they name,
  // split at "::" (src/parameters.cpp:488-600).
  std::string catalogue_dir;
  std::string measurement_dir;
  std::string data_catalogue_file;
  std::string rand_catalogue_file;
  std::string catalogue_columns;
  std::string catalogue_dataset;
  std::string output_tag;
  std::vector<std::string> data_catalogue_files;
  std::vector<std::string> rand_catalogue_files;
  double expand = 1.;
  double cutoff_nyq = 0.;
  std::string fftw_scheme = "measure";
  std::string use_fftw_wisdom = "false";
  std::string save_binned_vectors = "false";
  std::string progbar = "false";
  double boxsize[3] = {0., 0., 0.};
  int ngrid[3] = {0, 0, 0};
  double volume = 0.;
  long long nmesh = 0;
  std::string assignment = "tsc";
  int assignment_order = 3;
  std::string interlace = "false";
  std::string catalogue_type;
  std::string statistic_type;
  std::string npoint;
  std::string space;
  int ell1 = 0, ell2 = 0, ELL = 0;
  std::string form = "diag";
  std::string shape = "diag";
  int idx_bin = 0;
  int num_bins = 0;
  double bin_min = 0., bin_max = 0.;
  std::string binning = "lin";
  std::string norm_convention = "particle";
  int i_wa = 0, j_wa = 0;   // wide-angle orders of the 3PCF window
  std::string alignment = "centre";
  std::string padscale = "box";
  double padfactor = 0.;
  int verbose = 20;
  int precision = TRVG_PREC_FP64;   // engine extension: 64 or 32

  // ParameterSet::read_from_file (src/parameters.cpp:94-393): lines
  // "name = value"; comment lines ('#'), lines without a value and
  // unknown names are skipped; then validate(true).
  int read_from_file(const char* parameter_filepath) {
    std::ifstream fin(parameter_filepath);
    if (!fin)
      throw sys::IOError(std::string("Failed to open parameter file: ") +
                         parameter_filepath);
    std::string line;
    while (std::getline(fin, line)) {
      if (line.rfind("#", 0) == 0) continue;
      std::istringstream iss(line);
      std::string name, eq, value;
      if (!(iss >> name >> eq >> value) || eq != "=") continue;
      assign_parameter(name, value);
    }
    return validate(true);
  }
  int read_from_file(const std::string& parameter_filepath) {
    return read_from_file(parameter_filepath.c_str());
  }

  // ParameterSet::print_to_file (src/parameters.cpp:1272-1390): the
  // validated parameters as "name = value" lines, by default to
  // <measurement_dir>/parameters_used<output_tag>.
  int print_to_file(const char* out_parameter_filepath) const {
    std::FILE* f = std::fopen(out_parameter_filepath, "w");
    if (f == nullptr)
      throw sys::IOError("Non-existent or unwritable output directory: " +
                         measurement_dir);
    std::fprintf(f, "catalogue_dir = %s\n", catalogue_dir.c_str());
    std::fprintf(f, "measurement_dir = %s\n", measurement_dir.c_str());
    std::fprintf(f, "data_catalogue_file = %s\n", data_catalogue_file.c_str());
    std::fprintf(f, "rand_catalogue_file = %s\n", rand_catalogue_file.c_str());
    std::fprintf(f, "catalogue_columns = %s\n", catalogue_columns.c_str());
    std::fprintf(f, "catalogue_dataset = %s\n", catalogue_dataset.c_str());
    std::fprintf(f, "output_tag = %s\n", output_tag.c_str());
    const char ax[3] = {'x', 'y', 'z'};
    for (int a = 0; a < 3; a++) std::fprintf(f, "boxsize_%c = %.3f\n", ax[a], boxsize[a]);
    for (int a = 0; a < 3; a++) std::fprintf(f, "ngrid_%c = %d\n", ax[a], ngrid[a]);
    std::fprintf(f, "volume = %.6e\n", volume);
    std::fprintf(f, "nmesh = %lld\n", nmesh);
    std::fprintf(f, "expand = %.4f\n", expand);
    std::fprintf(f, "alignment = %s\n", alignment.c_str());
    std::fprintf(f, "padscale = %s\n", padscale.c_str());
    std::fprintf(f, "padfactor = %.4f\n", padfactor);
    std::fprintf(f, "assignment = %s\n", assignment.c_str());
    std::fprintf(f, "interlace = %s\n", interlace.c_str());
    std::fprintf(f, "assignment_order = %d\n", assignment_order);
    std::fprintf(f, "catalogue_type = %s\n", catalogue_type.c_str());
    std::fprintf(f, "statistic_type = %s\n", statistic_type.c_str());
    std::fprintf(f, "npoint = %s\n", npoint.c_str());
    std::fprintf(f, "space = %s\n", space.c_str());
    std::fprintf(f, "ell1 = %d\nell2 = %d\nELL = %d\n", ell1, ell2, ELL);
    std::fprintf(f, "i_wa = %d\nj_wa = %d\n", i_wa, j_wa);
    std::fprintf(f, "form = %s\n", form.c_str());
    std::fprintf(f, "norm_convention = %s\n", norm_convention.c_str());
    std::fprintf(f, "binning = %s\n", binning.c_str());
    std::fprintf(f, "shape = %s\n", shape.c_str());
    std::fprintf(f, "bin_min = %.4f\nbin_max = %.4f\n", bin_min, bin_max);
    std::fprintf(f, "num_bins = %d\nidx_bin = %d\n", num_bins, idx_bin);
    std::fprintf(f, "fftw_scheme = %s\n", fftw_scheme.c_str());
    std::fprintf(f, "use_fftw_wisdom = %s\n", use_fftw_wisdom.c_str());
    std::fprintf(f, "save_binned_vectors = %s\n", save_binned_vectors.c_str());
    std::fprintf(f, "progbar = %s\n", progbar.c_str());
    std::fprintf(f, "verbose = %d\n", verbose);
    std::fclose(f);
    return 0;
  }
  int print_to_file() const {
    std::string path = measurement_dir + "parameters_used" + output_tag;
    return print_to_file(path.c_str());
  }

  int validate(bool init = false) {
    validate_paths(init);
    if (assignment == "ngp") assignment_order = 1;
    else if (assignment == "cic") assignment_order = 2;
    else if (assignment == "tsc") assignment_order = 3;
    else if (assignment == "pcs") assignment_order = 4;
    else
      throw sys::InvalidParameterError(
          "Mesh assignment scheme must be 'ngp', 'cic', 'tsc' or 'pcs': "
          "`assignment` = '" + assignment + "'.");
    if (interlace == "true" || interlace == "on") interlace = "true";
    else if (interlace == "false" || interlace == "off") interlace = "false";
    else
      throw sys::InvalidParameterError(
          "Interlacing must be 'true'/'on' or 'false'/'off'.");
    if (statistic_type == "bispec") { npoint = "3pt"; space = "fourier"; }
    else if (statistic_type == "3pcf" || statistic_type == "3pcf-win" ||
             statistic_type == "3pcf-win-wa") {
      npoint = "3pt";
      space = "config";
    } else if (statistic_type == "powspec") { npoint = "2pt"; space = "fourier"; }
    else if (statistic_type == "2pcf" || statistic_type == "2pcf-win") {
      npoint = "2pt";
      space = "config";
    }
    if (!(form == "full" || form == "diag" || form == "off-diag" || form == "row"))
      throw sys::InvalidParameterError(
          "Three-point statistic form is not recognised: `form` = '" + form + "'.");
    shape = (form == "full" && ell1 == ell2) ? "triu" : form;
    volume = boxsize[0] * boxsize[1] * boxsize[2];
    nmesh = static_cast<long long>(ngrid[0]) * ngrid[1] * ngrid[2];
    if (npoint == "3pt" && interlace == "true") interlace = "false";
    return 0;
  }

 private:
  void assign_parameter(const std::string& name, const std::string& value) {
    auto real = [&](double& dst) { dst = std::strtod(value.c_str(), nullptr); };
    auto integer = [&](int& dst) {
      dst = static_cast<int>(std::strtol(value.c_str(), nullptr, 10));
    };
    if (name == "catalogue_dir") catalogue_dir = value;
    else if (name == "measurement_dir") measurement_dir = value;
    else if (name == "data_catalogue_file") data_catalogue_file = value;
    else if (name == "rand_catalogue_file") rand_catalogue_file = value;
    else if (name == "catalogue_columns") catalogue_columns = value;
    else if (name == "catalogue_dataset") catalogue_dataset = value;
    else if (name == "output_tag") output_tag = value;
    else if (name == "boxsize_x") real(boxsize[0]);
    else if (name == "boxsize_y") real(boxsize[1]);
    else if (name == "boxsize_z") real(boxsize[2]);
    else if (name == "ngrid_x") integer(ngrid[0]);
    else if (name == "ngrid_y") integer(ngrid[1]);
    else if (name == "ngrid_z") integer(ngrid[2]);
    else if (name == "expand") real(expand);
    else if (name == "cutoff_nyq") real(cutoff_nyq);
    else if (name == "alignment") alignment = value;
    else if (name == "padscale") padscale = value;
    else if (name == "padfactor") real(padfactor);
    else if (name == "assignment") assignment = value;
    else if (name == "interlace") interlace = value;
    else if (name == "catalogue_type") catalogue_type = value;
    else if (name == "statistic_type") statistic_type = value;
    else if (name == "form") form = value;
    else if (name == "norm_convention") norm_convention = value;
    else if (name == "binning") binning = value;
    else if (name == "ell1") integer(ell1);
    else if (name == "ell2") integer(ell2);
    else if (name == "ELL") integer(ELL);
    else if (name == "i_wa") integer(i_wa);
    else if (name == "j_wa") integer(j_wa);
    else if (name == "bin_min") real(bin_min);
    else if (name == "bin_max") real(bin_max);
    else if (name == "num_bins") integer(num_bins);
    else if (name == "idx_bin") integer(idx_bin);
    else if (name == "fftw_scheme") fftw_scheme = value;
    else if (name == "use_fftw_wisdom") use_fftw_wisdom = value;
    else if (name == "save_binned_vectors") save_binned_vectors = value;
    else if (name == "progbar") progbar = value;
    else if (name == "verbose") integer(verbose);
  }

  // Directory and catalogue-file transmutations of ParameterSet::validate
  // (src/parameters.cpp:466-620): directories gain a trailing '/', an empty
  // measurement directory becomes "./", ${VAR} expands from the
  // environment, and each "::"-separated catalogue file not starting with
  // '/' is taken relative to the catalogue directory.
  void validate_paths(bool init) {
    auto blank = [](const std::string& t) {
      return t.find_first_not_of(" \t\n\r\v\f") == std::string::npos;
    };
    if (!blank(catalogue_dir) && init) catalogue_dir += "/";
    if (blank(measurement_dir)) measurement_dir = "./";
    else if (init) measurement_dir += "/";
    sys::expand_envar_in_path(catalogue_dir);
    sys::expand_envar_in_path(measurement_dir);
    auto split = [&](std::string& spec, std::vector<std::string>& out) {
      out.clear();
      if (blank(spec)) return;
      const std::string delim = "::";
      if (spec.size() < delim.size() ||
          spec.compare(spec.size() - delim.size(), delim.size(), delim) != 0)
        spec += delim;
      std::size_t start = 0, pos;
      while ((pos = spec.find(delim, start)) != std::string::npos) {
        std::string f = spec.substr(start, pos - start);
        start = pos + delim.size();
        if (f.rfind("/", 0) != 0 && init) f = catalogue_dir + f;
        sys::expand_envar_in_path(f);
        out.push_back(f);
      }
    };
    data_catalogue_files.clear();
    rand_catalogue_files.clear();
    if (catalogue_type == "survey" || catalogue_type == "sim")
      split(data_catalogue_file, data_catalogue_files);
    if (catalogue_type == "survey" || catalogue_type == "random")
      split(rand_catalogue_file, rand_catalogue_files);
    auto join = [](const std::vector<std::string>& parts) {
      std::string out;
      for (std::size_t i = 0; i < parts.size(); i++)
        out += (i ? "::" : "") + parts[i];
      return out;
    };
    data_catalogue_file = join(data_catalogue_files);
    rand_catalogue_file = join(rand_catalogue_files);
  }

 public:
  trvg_params abi_view() const {
    trvg_params p{};
    for (int a = 0; a < 3; a++) {
      p.boxsize[a] = boxsize[a];
      p.ngrid[a] = ngrid[a];
    }
    p.assignment_order = assignment_order;
    p.interlace = (interlace == "true") ? 1 : 0;
    p.ell1 = ell1;
    p.ell2 = ell2;
    p.ELL = ELL;
    if (shape == "diag") p.shape = TRVG_SHAPE_DIAG;
    else if (shape == "off-diag") p.shape = TRVG_SHAPE_OFFDIAG;
    else if (shape == "row") p.shape = TRVG_SHAPE_ROW;
    else if (shape == "full") p.shape = TRVG_SHAPE_FULL;
    else if (shape == "triu") p.shape = TRVG_SHAPE_TRIU;
    else throw sys::InvalidParameterError("Unknown shape '" + shape + "'.");
    p.idx_bin = idx_bin;
    p.precision = precision;
    p.verbose = verbose;
    p.device = -1;
    return p;
  }
};

// trv::set_ngrid_from_cutoff (src/parameters.cpp:1434-1463): mesh numbers
// from the Nyquist cutoff, ceil(L k_cut / pi) in Fourier space and
// ceil(2 L / r_cut) in configuration space, rounded up to even.
inline void set_ngrid_from_cutoff(ParameterSet& params) {
  for (int a = 0; a < 3; a++) {
    int n;
    if (params.space == "fourier")
      n = static_cast<int>(std::ceil(params.boxsize[a] * params.cutoff_nyq / M_PI));
    else if (params.space == "config")
      n = static_cast<int>(std::ceil(2. * params.boxsize[a] / params.cutoff_nyq));
    else
      throw sys::InvalidParameterError("Space must be 'fourier' or 'config': `space` = '" +
                                       params.space + "'.");
    params.ngrid[a] = n + (n % 2);
  }
  params.validate();
}

// trv::Binning (include/dataobjs.hpp:58-150): linear, logarithmic and the
// padded schemes (src/dataobjs.cpp:134-250 arithmetic), the grid-scale
// default binning and custom edges.
class Binning {
 public:
  std::string space, scheme;
  double bin_min = 0., bin_max = 0.;
  int num_bins = 0;
  std::vector<double> bin_edges, bin_centres, bin_widths;
  int nbin_pad = 5;                 // padded bins of the *pad schemes
  double dbin_pad_fourier = 1.e-3;  // their width in Fourier space
  double dbin_pad_config = 10.;     // and in configuration space

  Binning(std::string space_, std::string scheme_)
      : space(std::move(space_)), scheme(std::move(scheme_)) {}
  // Range from the parameters; padding at the grid scale (dataobjs.cpp:42-56).
  explicit Binning(const ParameterSet& p) : Binning(p.space, p.binning) {
    bin_min = p.bin_min;
    bin_max = p.bin_max;
    num_bins = p.num_bins;
    const double lmax = std::max(p.boxsize[0], std::max(p.boxsize[1], p.boxsize[2]));
    const int nmin = std::min(p.ngrid[0], std::min(p.ngrid[1], p.ngrid[2]));
    dbin_pad_config = (1 + 5.e-3) * lmax / double(nmin);
    dbin_pad_fourier = (1 + 5.e-3) * (2 * M_PI) / lmax;
  }

  void set_bins(double coord_min, double coord_max, int nbin) {
    if (coord_min < 0.)
      throw sys::InvalidParameterError("Bin range must be non-negative.");
    if (nbin <= 0)
      throw sys::InvalidParameterError("Number of bins must be positive.");
    bin_min = coord_min;
    bin_max = coord_max;
    num_bins = nbin;
    set_bins();
  }

  // Recompute the bins from the stored range and scheme.
  void set_bins() {
    bin_edges.clear();
    bin_centres.clear();
    bin_widths.clear();
    compute_binning();
  }

  // Linear bins at the grid scale: half the grid number of bins from zero
  // (dataobjs.cpp:80-104).
  void set_bins(double boxsize_max, int ngrid_min) {
    scheme = "lin";
    bin_min = 0.;
    double width = 0.;
    if (space == "config") {
      width = boxsize_max / double(ngrid_min);
      bin_max = boxsize_max / 2.;
    } else if (space == "fourier") {
      width = 2 * M_PI / boxsize_max;
      bin_max = M_PI * double(ngrid_min) / boxsize_max;
    }
    bin_max += width / 2.;
    num_bins = ngrid_min / 2;
    set_bins();
  }

  // Custom edges (dataobjs.cpp:106-126).
  void set_bins(const std::vector<double>& edges) {
    bin_min = edges.front();
    bin_max = edges.back();
    num_bins = int(edges.size()) - 1;
    scheme = "custom";
    bin_edges = edges;
    bin_centres.clear();
    bin_widths.clear();
    for (int i = 0; i < num_bins; i++) {
      bin_centres.push_back((edges[i] + edges[i + 1]) / 2.);
      bin_widths.push_back(edges[i + 1] - edges[i]);
    }
  }

 private:
  void compute_binning() {
    const double pad = (space == "fourier") ? dbin_pad_fourier : dbin_pad_config;
    if (scheme == "lin") {
      double dbin = (bin_max - bin_min) / double(num_bins);
      for (int i = 0; i < num_bins; i++) push_bin(bin_min + dbin * i, dbin);
    } else if (scheme == "log") {
      if (bin_min == 0.)
        throw sys::InvalidParameterError(
            "Cannot use logarithmic binning when the lowest edge is zero.");
      double dln = (std::log(bin_max) - std::log(bin_min)) / double(num_bins);
      for (int i = 0; i < num_bins; i++)
        push_log_bin(bin_min * std::exp(dln * i), bin_min * std::exp(dln * (i + 1)));
    } else if (scheme == "linpad" || scheme == "logpad") {
      // nbin_pad linear bins of width `pad` from zero, then the rest of the
      // range linear or logarithmic.
      for (int i = 0; i < nbin_pad; i++) push_bin(pad * i, pad);
      const double lo = pad * nbin_pad;
      const double nrest = double(num_bins - nbin_pad);
      if (scheme == "linpad") {
        double dbin = (bin_max - lo) / nrest;
        for (int i = nbin_pad; i < num_bins; i++) push_bin(lo + dbin * (i - nbin_pad), dbin);
      } else {
        double dln = (std::log(bin_max) - std::log(lo)) / nrest;
        for (int i = nbin_pad; i < num_bins; i++)
          push_log_bin(lo * std::exp(dln * (i - nbin_pad)),
                       lo * std::exp(dln * (i - nbin_pad + 1)));
      }
    } else {
      throw sys::InvalidParameterError(
          "Unrecognised/unsupported binning `scheme`: " + scheme + ".");
    }
    bin_edges.push_back(bin_max);
  }
  void push_bin(double left, double width) {
    bin_edges.push_back(left);
    bin_centres.push_back(left + width / 2.);
    bin_widths.push_back(width);
  }
  void push_log_bin(double left, double right) {
    bin_edges.push_back(left);
    bin_centres.push_back((left + right) / 2.);
    bin_widths.push_back(right - left);
  }

 public:
  trvg_binning abi_view() const {
    trvg_binning b;
    b.num_bins = num_bins;
    b.bin_edges = bin_edges.data();
    b.bin_centres = bin_centres.data();
    return b;
  }
};

struct PowspecMeasurements {
  int dim = 0;
  std::vector<double> kbin, keff;
  std::vector<int> nmodes;
  std::vector<std::complex<double>> pk_raw, pk_shot;
};

struct TwoPCFMeasurements {
  int dim = 0;
  std::vector<double> rbin, reff;
  std::vector<int> npairs;
  std::vector<std::complex<double>> xi;
};

struct TwoPCFWindowMeasurements {
  int dim = 0;
  std::vector<double> rbin, reff;
  std::vector<int> npairs;
  std::vector<std::complex<double>> xi;
};

struct BispecMeasurements {
  int dim = 0;
  std::vector<double> k1_bin, k2_bin, k1_eff, k2_eff;
  std::vector<int> nmodes_1, nmodes_2;
  std::vector<std::complex<double>> bk_raw, bk_shot;
};

struct ThreePCFMeasurements {
  int dim = 0;
  std::vector<double> r1_bin, r2_bin, r1_eff, r2_eff;
  std::vector<int> npairs_1, npairs_2;
  std::vector<std::complex<double>> zeta_raw, zeta_shot;
};

// trv::ThreePCFWindowMeasurements (include/dataobjs.hpp:288): the fields
// of ThreePCFMeasurements.
struct ThreePCFWindowMeasurements : ThreePCFMeasurements {};

namespace detail {

inline int dim_for(const trvg_params& p, const Binning& b) {
  int dim = 0;
  sys::check_status(trvg_result_dim(p.shape, b.num_bins, p.idx_bin, &dim));
  return dim;
}

template <class Call>
BispecMeasurements run_bispec(const trvg_params& p, const Binning& b,
                              Call&& call) {
  BispecMeasurements m;
  int dim = dim_for(p, b);
  m.k1_bin.resize(dim); m.k2_bin.resize(dim);
  m.k1_eff.resize(dim); m.k2_eff.resize(dim);
  m.nmodes_1.resize(dim); m.nmodes_2.resize(dim);
  m.bk_raw.resize(dim); m.bk_shot.resize(dim);
  trvg_bispec_result r{dim, m.k1_bin.data(), m.k2_bin.data(), m.k1_eff.data(),
                       m.k2_eff.data(), m.nmodes_1.data(), m.nmodes_2.data(),
                       reinterpret_cast<double*>(m.bk_raw.data()),
                       reinterpret_cast<double*>(m.bk_shot.data())};
  sys::check_status(call(&r));
  m.dim = r.dim;
  return m;
}

template <class Call>
ThreePCFMeasurements run_3pcf(const trvg_params& p, const Binning& b,
                              Call&& call) {
  ThreePCFMeasurements m;
  int dim = dim_for(p, b);
  m.r1_bin.resize(dim); m.r2_bin.resize(dim);
  m.r1_eff.resize(dim); m.r2_eff.resize(dim);
  m.npairs_1.resize(dim); m.npairs_2.resize(dim);
  m.zeta_raw.resize(dim); m.zeta_shot.resize(dim);
  trvg_3pcf_result r{dim, m.r1_bin.data(), m.r2_bin.data(), m.r1_eff.data(),
                     m.r2_eff.data(), m.npairs_1.data(), m.npairs_2.data(),
                     reinterpret_cast<double*>(m.zeta_raw.data()),
                     reinterpret_cast<double*>(m.zeta_shot.data())};
  sys::check_status(call(&r));
  m.dim = r.dim;
  return m;
}

template <class Call>
PowspecMeasurements run_powspec(const Binning& b, Call&& call) {
  PowspecMeasurements m;
  const int n = b.num_bins;
  m.kbin.resize(n); m.keff.resize(n); m.nmodes.resize(n);
  m.pk_raw.resize(n); m.pk_shot.resize(n);
  trvg_powspec_result r{n, m.kbin.data(), m.keff.data(), m.nmodes.data(),
                        reinterpret_cast<double*>(m.pk_raw.data()),
                        reinterpret_cast<double*>(m.pk_shot.data())};
  sys::check_status(call(&r));
  m.dim = r.dim;
  return m;
}

template <class M, class Call>
M run_2pcf(const Binning& b, Call&& call) {
  M m;
  const int n = b.num_bins;
  m.rbin.resize(n); m.reff.resize(n); m.npairs.resize(n); m.xi.resize(n);
  trvg_2pcf_result r{n, m.rbin.data(), m.reff.data(), m.npairs.data(),
                     reinterpret_cast<double*>(m.xi.data())};
  sys::check_status(call(&r));
  m.dim = r.dim;
  return m;
}

}  // namespace detail

inline double calc_bispec_normalisation_from_particles(
    ParticleCatalogue& particles, double alpha = 1.) {
  trvg_catalogue c = particles.abi_view();
  double nf = 0.;
  sys::check_status(trvg_calc_bispec_normalisation_from_particles(&c, alpha, &nf));
  return nf;
}

inline double calc_bispec_normalisation_from_mesh(ParticleCatalogue& particles,
                                                  ParameterSet& params,
                                                  double alpha = 1.) {
  trvg_catalogue c = particles.abi_view();
  trvg_params p = params.abi_view();
  double nf = 0.;
  sys::check_status(trvg_calc_bispec_normalisation_from_mesh(&c, &p, alpha, &nf));
  return nf;
}

inline BispecMeasurements compute_bispec(
    ParticleCatalogue& catalogue_data, ParticleCatalogue& catalogue_rand,
    LineOfSight* los_data, LineOfSight* los_rand, ParameterSet& params,
    Binning& kbinning, double norm_factor) {
  trvg_catalogue d = catalogue_data.abi_view(), r = catalogue_rand.abi_view();
  trvg_params p = params.abi_view();
  trvg_binning b = kbinning.abi_view();
  return detail::run_bispec(p, kbinning, [&](trvg_bispec_result* out) {
    return trvg_compute_bispec(&d, &r, &los_data[0].pos[0], &los_rand[0].pos[0],
                               &p, &b, norm_factor, out);
  });
}

inline ThreePCFMeasurements compute_3pcf(
    ParticleCatalogue& catalogue_data, ParticleCatalogue& catalogue_rand,
    LineOfSight* los_data, LineOfSight* los_rand, ParameterSet& params,
    Binning& rbinning, double norm_factor) {
  trvg_catalogue d = catalogue_data.abi_view(), r = catalogue_rand.abi_view();
  trvg_params p = params.abi_view();
  trvg_binning b = rbinning.abi_view();
  return detail::run_3pcf(p, rbinning, [&](trvg_3pcf_result* out) {
    return trvg_compute_3pcf(&d, &r, &los_data[0].pos[0], &los_rand[0].pos[0],
                             &p, &b, norm_factor, out);
  });
}

inline BispecMeasurements compute_bispec_in_gpp_box(
    ParticleCatalogue& catalogue_data, ParameterSet& params, Binning kbinning,
    double norm_factor) {
  trvg_catalogue d = catalogue_data.abi_view();
  trvg_params p = params.abi_view();
  trvg_binning b = kbinning.abi_view();
  return detail::run_bispec(p, kbinning, [&](trvg_bispec_result* out) {
    return trvg_compute_bispec_in_gpp_box(&d, &p, &b, norm_factor, out);
  });
}

inline ThreePCFMeasurements compute_3pcf_in_gpp_box(
    ParticleCatalogue& catalogue_data, ParameterSet& params, Binning& rbinning,
    double norm_factor) {
  trvg_catalogue d = catalogue_data.abi_view();
  trvg_params p = params.abi_view();
  trvg_binning b = rbinning.abi_view();
  return detail::run_3pcf(p, rbinning, [&](trvg_3pcf_result* out) {
    return trvg_compute_3pcf_in_gpp_box(&d, &p, &b, norm_factor, out);
  });
}

// threept.cpp:2621-3077 (include/threept.hpp:310-314).
inline ThreePCFWindowMeasurements compute_3pcf_window(
    ParticleCatalogue& catalogue_rand, LineOfSight* los_rand, ParameterSet& params,
    Binning& rbinning, double alpha, double norm_factor, bool wide_angle = false) {
  trvg_catalogue r = catalogue_rand.abi_view();
  trvg_params p = params.abi_view();
  trvg_binning b = rbinning.abi_view();
  ThreePCFWindowMeasurements w;
  static_cast<ThreePCFMeasurements&>(w) =
      detail::run_3pcf(p, rbinning, [&](trvg_3pcf_result* out) {
        return trvg_compute_3pcf_window(&r, &los_rand[0].pos[0], &p, &b, alpha,
                                        norm_factor, wide_angle ? 1 : 0, params.i_wa,
                                        params.j_wa, out);
      });
  return w;
}

// ---------------------------------------------------------------------------
// Two-point statistics (include/twopt.hpp).
// ---------------------------------------------------------------------------

inline double calc_powspec_normalisation_from_particles(ParticleCatalogue& particles,
                                                        double alpha = 1.) {
  trvg_catalogue c = particles.abi_view();
  double nf = 0.;
  sys::check_status(trvg_calc_powspec_normalisation_from_particles(&c, alpha, &nf));
  return nf;
}

inline double calc_powspec_normalisation_from_mesh(ParticleCatalogue& particles,
                                                   ParameterSet& params,
                                                   double alpha = 1.) {
  trvg_catalogue c = particles.abi_view();
  trvg_params p = params.abi_view();
  double nf = 0.;
  sys::check_status(trvg_calc_powspec_normalisation_from_mesh(&c, &p, alpha, &nf));
  return nf;
}

// twopt.cpp:111-228: both catalogues aligned in the box of `params` as the
// measurement would be (centre, or pad by box or grid), the mixed-mesh
// normalisation computed, and the original positions restored.
inline double calc_powspec_normalisation_from_meshes(ParticleCatalogue& particles_data,
                                                     ParticleCatalogue& particles_rand,
                                                     ParameterSet& params, double alpha) {
  double dvec[3];
  for (int a = 0; a < 3; a++) {
    if (params.alignment == "pad") {
      const double pad = (params.padscale == "grid")
                             ? params.padfactor * params.boxsize[a] / params.ngrid[a]
                             : params.padfactor * params.boxsize[a];
      dvec[a] = particles_rand.pos_min[a] - pad;
    } else {
      dvec[a] = (particles_rand.pos_min[a] + particles_rand.pos_max[a]) / 2. -
                params.boxsize[a] / 2.;
    }
  }
  particles_data.offset_coords(dvec);
  particles_rand.offset_coords(dvec);
  trvg_catalogue d = particles_data.abi_view(), r = particles_rand.abi_view();
  trvg_params p = params.abi_view();
  double nf = 0.;
  const int st = trvg_calc_powspec_normalisation_from_meshes(&d, &r, &p, alpha, &nf);
  const double back[3] = {-dvec[0], -dvec[1], -dvec[2]};
  particles_data.offset_coords(back);
  particles_rand.offset_coords(back);
  sys::check_status(st);
  return nf;
}

// twopt.cpp:230-268: the mixed-mesh geometry from the catalogue spans, a
// cell size and a scheme.
inline double calc_powspec_normalisation_from_meshes(ParticleCatalogue& particles_data,
                                                     ParticleCatalogue& particles_rand,
                                                     ParameterSet& params, double alpha,
                                                     double padding, double cellsize,
                                                     const std::string& assignment) {
  ParameterSet pn = params;
  double span = 0.;
  for (int a = 0; a < 3; a++)
    span = std::max(span, std::max(particles_data.pos_span[a], particles_rand.pos_span[a]));
  double L = (1. + padding) * span;
  int n = int(std::ceil(L / cellsize));
  n += n % 2;
  L = n * cellsize;
  for (int a = 0; a < 3; a++) {
    pn.boxsize[a] = L;
    pn.ngrid[a] = n;
  }
  pn.assignment = assignment;
  pn.validate();
  return calc_powspec_normalisation_from_meshes(particles_data, particles_rand, pn, alpha);
}

inline double calc_powspec_shotnoise_from_particles(ParticleCatalogue& particles,
                                                    double alpha = 1.) {
  trvg_catalogue c = particles.abi_view();
  double sn = 0.;
  sys::check_status(trvg_calc_powspec_shotnoise_from_particles(&c, alpha, &sn));
  return sn;
}

inline PowspecMeasurements compute_powspec(
    ParticleCatalogue& catalogue_data, ParticleCatalogue& catalogue_rand,
    LineOfSight* los_data, LineOfSight* los_rand, ParameterSet& params,
    Binning& kbinning, double norm_factor) {
  trvg_catalogue d = catalogue_data.abi_view(), r = catalogue_rand.abi_view();
  trvg_params p = params.abi_view();
  trvg_binning b = kbinning.abi_view();
  return detail::run_powspec(kbinning, [&](trvg_powspec_result* out) {
    return trvg_compute_powspec(&d, &r, &los_data[0].pos[0], &los_rand[0].pos[0], &p,
                                &b, norm_factor, out);
  });
}

inline TwoPCFMeasurements compute_corrfunc(
    ParticleCatalogue& catalogue_data, ParticleCatalogue& catalogue_rand,
    LineOfSight* los_data, LineOfSight* los_rand, ParameterSet& params,
    Binning& rbinning, double norm_factor) {
  trvg_catalogue d = catalogue_data.abi_view(), r = catalogue_rand.abi_view();
  trvg_params p = params.abi_view();
  trvg_binning b = rbinning.abi_view();
  return detail::run_2pcf<TwoPCFMeasurements>(rbinning, [&](trvg_2pcf_result* out) {
    return trvg_compute_corrfunc(&d, &r, &los_data[0].pos[0], &los_rand[0].pos[0], &p,
                                 &b, norm_factor, out);
  });
}

inline PowspecMeasurements compute_powspec_in_gpp_box(
    ParticleCatalogue& catalogue_data, ParameterSet& params, Binning kbinning,
    double norm_factor) {
  trvg_catalogue d = catalogue_data.abi_view();
  trvg_params p = params.abi_view();
  trvg_binning b = kbinning.abi_view();
  return detail::run_powspec(kbinning, [&](trvg_powspec_result* out) {
    return trvg_compute_powspec_in_gpp_box(&d, &p, &b, norm_factor, out);
  });
}

inline TwoPCFMeasurements compute_corrfunc_in_gpp_box(
    ParticleCatalogue& catalogue_data, ParameterSet& params, Binning& rbinning,
    double norm_factor) {
  trvg_catalogue d = catalogue_data.abi_view();
  trvg_params p = params.abi_view();
  trvg_binning b = rbinning.abi_view();
  return detail::run_2pcf<TwoPCFMeasurements>(rbinning, [&](trvg_2pcf_result* out) {
    return trvg_compute_corrfunc_in_gpp_box(&d, &p, &b, norm_factor, out);
  });
}

inline TwoPCFWindowMeasurements compute_corrfunc_window(
    ParticleCatalogue& catalogue_rand, LineOfSight* los_rand, ParameterSet& params,
    Binning rbinning, double alpha, double norm_factor) {
  trvg_catalogue r = catalogue_rand.abi_view();
  trvg_params p = params.abi_view();
  trvg_binning b = rbinning.abi_view();
  return detail::run_2pcf<TwoPCFWindowMeasurements>(rbinning, [&](trvg_2pcf_result* out) {
    return trvg_compute_corrfunc_window(&r, &los_rand[0].pos[0], &p, &b, alpha,
                                        norm_factor, out);
  });
}

// ---------------------------------------------------------------------------
// Measurement files (include/io.hpp:36-230, src/io.cpp:39-558): the header
// lines of the catalogues, sampling parameters and normalisations, then the
// column names and the tab-separated table.
// ---------------------------------------------------------------------------

namespace sys {

// trv::sys::if_filepath_is_set (src/io.cpp:39-57): a path is set when it
// is non-blank and does not name a directory (no trailing '/').
inline bool if_filepath_is_set(const std::string& pathstr) {
  if (pathstr.empty() || pathstr.back() == '/') return false;
  return pathstr.find_first_not_of(" \t\n\r\v\f") != std::string::npos;
}

// trv::sys::make_write_dir (src/io.cpp:59-91): create the directory and
// its parents if absent ("", ".", "./" and "/" are left alone).
inline void make_write_dir(std::string dirstr) {
  if (dirstr.empty() || dirstr == "." || dirstr == "./" || dirstr == "/") return;
  while (!dirstr.empty() && dirstr.back() == '/') dirstr.pop_back();
  std::error_code ec;
  std::filesystem::create_directories(std::filesystem::path(dirstr), ec);
  if (ec) throw IOError("Failed to create directory: " + dirstr);
}

}  // namespace sys

namespace io {

const char comment_delimiter[] = "#";

namespace detail {

inline void print_catalogue_lines(std::FILE* f, const char* source_label,
                                  const char* size_label, const char* extent_label,
                                  const ParticleCatalogue& c) {
  std::fprintf(f, "%s %s: %s\n", comment_delimiter, source_label, c.source.c_str());
  std::fprintf(f, "%s %s: ntotal = %d, wtotal = %.3f, wstotal = %.3f\n",
               comment_delimiter, size_label, c.ntotal, c.wtotal, c.wstotal);
  std::fprintf(f, "%s %s: [(%.3f, %.3f), (%.3f, %.3f), (%.3f, %.3f)]\n",
               comment_delimiter, extent_label, c.pos_min[0], c.pos_max[0],
               c.pos_min[1], c.pos_max[1], c.pos_min[2], c.pos_max[2]);
}

inline void print_sampling_lines(std::FILE* f, const ParameterSet& p,
                                 double part, double mesh, double meshes) {
  std::fprintf(f, "%s Box size: [%.3f, %.3f, %.3f]\n", comment_delimiter,
               p.boxsize[0], p.boxsize[1], p.boxsize[2]);
  std::fprintf(f, "%s Box alignment: %s\n", comment_delimiter, p.alignment.c_str());
  std::fprintf(f, "%s Mesh number: [%d, %d, %d]\n", comment_delimiter,
               p.ngrid[0], p.ngrid[1], p.ngrid[2]);
  std::fprintf(f, "%s Mesh assignment and interlacing: %s, %s\n", comment_delimiter,
               p.assignment.c_str(), p.interlace.c_str());
  double used = 1.;
  if (p.norm_convention == "particle") used = part;
  else if (p.norm_convention == "mesh") used = mesh;
  else if (p.norm_convention == "mesh-mixed") used = meshes;
  std::fprintf(f, "%s Normalisation factor: %.9e (%s)\n", comment_delimiter, used,
               p.norm_convention.c_str());
  std::fprintf(f,
               "%s Normalisation factor alternatives: "
               "%.9e (particle), %.9e (mesh), %.9e (mesh-mixed)\n",
               comment_delimiter, part, mesh, meshes);
}

inline std::string multipole_str(const ParameterSet& p) {
  return std::to_string(p.ell1) + std::to_string(p.ell2) + std::to_string(p.ELL);
}

template <class C>
inline void print_threept_rows(std::FILE* f, int dim, const std::vector<double>& b1,
                               const std::vector<double>& e1, const std::vector<int>& n1,
                               const std::vector<double>& b2,
                               const std::vector<double>& e2, const std::vector<int>& n2,
                               const C& raw, const C& shot) {
  for (int i = 0; i < dim; i++)
    std::fprintf(f, "%.9e\t%.9e\t%10d\t%.9e\t%.9e\t%10d\t% .9e\t% .9e\t% .9e\t% .9e\n",
                 b1[i], e1[i], n1[i], b2[i], e2[i], n2[i], raw[i].real(),
                 raw[i].imag(), shot[i].real(), shot[i].imag());
}

}  // namespace detail

// Header of a paired data/random measurement (src/io.cpp:102-210).
inline void print_measurement_header_to_file(
    std::FILE* fileptr, ParameterSet& params, ParticleCatalogue& catalogue_data,
    ParticleCatalogue& catalogue_rand, double norm_factor_part,
    double norm_factor_mesh, double norm_factor_meshes) {
  detail::print_catalogue_lines(fileptr, "Data catalogue source", "Data catalogue size",
                                "Data-source particle extents", catalogue_data);
  detail::print_catalogue_lines(fileptr, "Random catalogue source",
                                "Random catalogue size",
                                "Random-source particle extents", catalogue_rand);
  detail::print_sampling_lines(fileptr, params, norm_factor_part, norm_factor_mesh,
                               norm_factor_meshes);
}

// Header of a single-catalogue measurement (src/io.cpp:212-298).
inline void print_measurement_header_to_file(
    std::FILE* fileptr, ParameterSet& params, ParticleCatalogue& catalogue,
    double norm_factor_part, double norm_factor_mesh, double norm_factor_meshes) {
  detail::print_catalogue_lines(fileptr, "Catalogue source", "Catalogue size",
                                "Catalogue particle extents", catalogue);
  detail::print_sampling_lines(fileptr, params, norm_factor_part, norm_factor_mesh,
                               norm_factor_meshes);
}

// Data tables (src/io.cpp:360-555).
inline void print_measurement_datatab_to_file(std::FILE* fileptr, ParameterSet& params,
                                              PowspecMeasurements& m) {
  const int L = params.ELL;
  std::fprintf(fileptr,
               "%s [0] k_cen, [1] k_eff, [2] nmodes, [3] Re{pk%d_raw}, "
               "[4] Im{pk%d_raw}, [5] Re{pk%d_shot}, [6] Im{pk%d_shot}\n",
               comment_delimiter, L, L, L, L);
  for (int i = 0; i < m.dim; i++)
    std::fprintf(fileptr, "%.9e\t%.9e\t%10d\t% .9e\t% .9e\t% .9e\t% .9e\n", m.kbin[i],
                 m.keff[i], m.nmodes[i], m.pk_raw[i].real(), m.pk_raw[i].imag(),
                 m.pk_shot[i].real(), m.pk_shot[i].imag());
}

namespace detail {
template <class M>
inline void print_corrfunc_table(std::FILE* f, const ParameterSet& p, const M& m) {
  std::fprintf(f, "%s [0] r_cen, [1] r_eff, [2] npairs, [3] Re{xi%d}, [4] Im{xi%d}\n",
               comment_delimiter, p.ELL, p.ELL);
  for (int i = 0; i < m.dim; i++)
    std::fprintf(f, "%.9e\t%.9e\t%10d\t% .9e\t% .9e\n", m.rbin[i], m.reff[i],
                 m.npairs[i], m.xi[i].real(), m.xi[i].imag());
}
template <class M>
inline void print_zeta_table(std::FILE* f, const ParameterSet& p, const M& m) {
  const std::string mp = multipole_str(p);
  std::fprintf(f,
               "%s [0] r1_cen, [1] r1_eff, [2] npairs_1, [3] r2_cen, [4] r2_eff, "
               "[5] npairs_2, [6] Re{zeta%s_raw}, [7] Im{zeta%s_raw}, "
               "[8] Re{zeta%s_shot}, [9] Im{zeta%s_shot}\n",
               comment_delimiter, mp.c_str(), mp.c_str(), mp.c_str(), mp.c_str());
  print_threept_rows(f, m.dim, m.r1_bin, m.r1_eff, m.npairs_1, m.r2_bin, m.r2_eff,
                     m.npairs_2, m.zeta_raw, m.zeta_shot);
}
}  // namespace detail

inline void print_measurement_datatab_to_file(std::FILE* fileptr, ParameterSet& params,
                                              TwoPCFMeasurements& m) {
  detail::print_corrfunc_table(fileptr, params, m);
}

inline void print_measurement_datatab_to_file(std::FILE* fileptr, ParameterSet& params,
                                              TwoPCFWindowMeasurements& m) {
  detail::print_corrfunc_table(fileptr, params, m);
}

inline void print_measurement_datatab_to_file(std::FILE* fileptr, ParameterSet& params,
                                              BispecMeasurements& m) {
  const std::string mp = detail::multipole_str(params);
  std::fprintf(fileptr,
               "%s [0] k1_cen, [1] k1_eff, [2] nmodes_1, [3] k2_cen, [4] k2_eff, "
               "[5] nmodes_2, [6] Re{bk%s_raw}, [7] Im{bk%s_raw}, "
               "[8] Re{bk%s_shot}, [9] Im{bk%s_shot}\n",
               comment_delimiter, mp.c_str(), mp.c_str(), mp.c_str(), mp.c_str());
  detail::print_threept_rows(fileptr, m.dim, m.k1_bin, m.k1_eff, m.nmodes_1, m.k2_bin,
                             m.k2_eff, m.nmodes_2, m.bk_raw, m.bk_shot);
}

inline void print_measurement_datatab_to_file(std::FILE* fileptr, ParameterSet& params,
                                              ThreePCFMeasurements& m) {
  detail::print_zeta_table(fileptr, params, m);
}

inline void print_measurement_datatab_to_file(std::FILE* fileptr, ParameterSet& params,
                                              ThreePCFWindowMeasurements& m) {
  detail::print_zeta_table(fileptr, params, m);
}

}  // namespace io

}  // namespace trv

#endif  // TRIUMVIRATE_GPU_HPP_INCLUDED
