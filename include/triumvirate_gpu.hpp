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
// Errors are thrown as trv::sys::InvalidParameterError /
// trv::sys::InvalidDataError (include/monitor.hpp:731-779), or
// std::runtime_error for device failures.
#ifndef TRIUMVIRATE_GPU_HPP_INCLUDED
#define TRIUMVIRATE_GPU_HPP_INCLUDED

#include <algorithm>
#include <cmath>
#include <complex>
#include <stdexcept>
#include <string>
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
    return 0;
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

  int validate() {
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

}  // namespace trv

#endif  // TRIUMVIRATE_GPU_HPP_INCLUDED
