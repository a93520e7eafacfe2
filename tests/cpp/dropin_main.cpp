// Drop-in check: calls the trv:: C++ API of include/triumvirate_gpu.hpp the
// way the reference CLI calls its own (main/triumvirate.cpp:711-726) on the
// reference's periodic-box test catalogue, and prints the data vector: the
// bispectrum B_000 (default) or, with "powspec", the power spectrum P_0.
// Usage: dropin_main <data_catalogue.txt> [powspec]
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "triumvirate_gpu.hpp"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: dropin_main <catalogue.txt>\n";
    return 2;
  }
  std::ifstream in(argv[1]);
  std::vector<double> x, y, z, nz;
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ss(line);
    double a, b, c, d;
    ss >> a >> b >> c >> d;
    x.push_back(a); y.push_back(b); z.push_back(c); nz.push_back(d);
  }
  const double L = 1000.;
  // Periodise about the extents' mid-point (catalogue.py:648-676).
  for (auto* col : {&x, &y, &z}) {
    double lo = (*col)[0], hi = (*col)[0];
    for (double v : *col) { lo = std::min(lo, v); hi = std::max(hi, v); }
    for (double& v : *col) {
      v = std::fmod(v + L / 2. - (hi + lo) / 2., L);
      if (v < 0.) v += L;
    }
  }
  std::vector<double> ones(x.size(), 1.);
  trv::ParticleCatalogue cat;
  cat.load_particle_data(x, y, z, nz, ones, ones);

  trv::ParameterSet params;
  for (int a = 0; a < 3; a++) { params.boxsize[a] = L; params.ngrid[a] = 64; }
  params.assignment = "tsc";
  params.statistic_type = "bispec";
  params.catalogue_type = "sim";
  params.ell1 = params.ell2 = params.ELL = 0;
  params.form = "diag";
  params.num_bins = 4;
  params.bin_min = 0.005;
  params.bin_max = 0.105;
  params.validate();

  trv::Binning kbin("fourier", "lin");
  kbin.set_bins(0.005, 0.105, 4);

  if (argc > 2 && std::string(argv[2]) == "powspec") {
    params.statistic_type = "powspec";
    params.validate();
    try {
      double norm = trv::calc_powspec_normalisation_from_particles(cat, 1.);
      trv::PowspecMeasurements m =
          trv::compute_powspec_in_gpp_box(cat, params, kbin, norm);
      for (int i = 0; i < m.dim; i++)
        std::printf("%.9e %.9e %d %.12e %.12e %.12e %.12e\n", m.kbin[i], m.keff[i],
                    m.nmodes[i], m.pk_raw[i].real(), m.pk_raw[i].imag(),
                    m.pk_shot[i].real(), m.pk_shot[i].imag());
    } catch (const std::exception& e) {
      std::cerr << "error: " << e.what() << "\n";
      return 1;
    }
    return 0;
  }
  try {
    double norm = trv::calc_bispec_normalisation_from_particles(cat, 1.);
    trv::BispecMeasurements m =
        trv::compute_bispec_in_gpp_box(cat, params, kbin, norm);
    for (int i = 0; i < m.dim; i++) {
      std::printf("%.9e %.9e %d %.9e %.9e %d %.12e %.12e %.12e %.12e\n",
                  m.k1_bin[i], m.k1_eff[i], m.nmodes_1[i], m.k2_bin[i],
                  m.k2_eff[i], m.nmodes_2[i], m.bk_raw[i].real(),
                  m.bk_raw[i].imag(), m.bk_shot[i].real(), m.bk_shot[i].imag());
    }
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
