// File I/O of the C++ drop-in (parameter files, catalogue files,
// measurement files), driven by tests/test_io_formats.py.
//
//   io_main params <in.ini> <out>         read, print, write, read back
//   io_main catalogue <file> <columns> <volume>
//   io_main table <fixture.txt> <catalogue.txt> <out.txt> <bispec|powspec|2pcf|3pcf>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "triumvirate_gpu.hpp"

static void print_params(const trv::ParameterSet& p) {
  std::printf("catalogue_dir=%s\n", p.catalogue_dir.c_str());
  std::printf("measurement_dir=%s\n", p.measurement_dir.c_str());
  std::printf("data_catalogue_file=%s\n", p.data_catalogue_file.c_str());
  std::printf("boxsize=%.3f,%.3f,%.3f\n", p.boxsize[0], p.boxsize[1], p.boxsize[2]);
  std::printf("ngrid=%d,%d,%d\n", p.ngrid[0], p.ngrid[1], p.ngrid[2]);
  std::printf("volume=%.6e\n", p.volume);
  std::printf("nmesh=%lld\n", p.nmesh);
  std::printf("expand=%.4f\n", p.expand);
  std::printf("alignment=%s\n", p.alignment.c_str());
  std::printf("assignment=%s order=%d\n", p.assignment.c_str(), p.assignment_order);
  std::printf("interlace=%s\n", p.interlace.c_str());
  std::printf("norm_convention=%s\n", p.norm_convention.c_str());
  std::printf("binning=%s num_bins=%d\n", p.binning.c_str(), p.num_bins);
  std::printf("form=%s shape=%s\n", p.form.c_str(), p.shape.c_str());
  std::printf("verbose=%d progbar=%s\n", p.verbose, p.progbar.c_str());
}

static std::vector<std::vector<double>> read_columns(const char* path) {
  std::ifstream f(path);
  std::string line;
  std::vector<std::vector<double>> cols;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ls(line);
    double v;
    std::size_t c = 0;
    while (ls >> v) {
      if (cols.size() <= c) cols.emplace_back();
      cols[c++].push_back(v);
    }
  }
  return cols;
}

template <class T>
static std::vector<T> as(const std::vector<double>& v) {
  return std::vector<T>(v.begin(), v.end());
}

static std::vector<std::complex<double>> cplx(const std::vector<double>& re,
                                              const std::vector<double>& im) {
  std::vector<std::complex<double>> out(re.size());
  for (std::size_t i = 0; i < re.size(); i++) out[i] = {re[i], im[i]};
  return out;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  std::string mode = argv[1];
  try {
    if (mode == "params" && argc == 4) {
      trv::ParameterSet p;
      p.read_from_file(argv[2]);
      print_params(p);
      p.print_to_file(argv[3]);
      trv::ParameterSet q;
      q.read_from_file(argv[3]);
      std::printf("---\n");
      print_params(q);
      return 0;
    }
    if (mode == "catalogue" && argc == 5) {
      trv::ParticleCatalogue c;
      c.load_catalogue_file(argv[2], argv[3], "", std::strtod(argv[4], nullptr));
      std::printf("source=%s\n", c.source.c_str());
      std::printf("ntotal=%d wtotal=%.6f wstotal=%.6f\n", c.ntotal, c.wtotal, c.wstotal);
      for (int i = 0; i < c.ntotal && i < 3; i++)
        std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", c[i].pos[0],
                    c[i].pos[1], c[i].pos[2], c[i].nz, c[i].ws, c[i].wc, c[i].w);
      std::printf("extents=%.6f %.6f %.6f %.6f %.6f %.6f\n", c.pos_min[0], c.pos_max[0],
                  c.pos_min[1], c.pos_max[1], c.pos_min[2], c.pos_max[2]);
      return 0;
    }
    if (mode == "table" && argc == 6) {
      auto cols = read_columns(argv[2]);
      trv::ParticleCatalogue cat;
      cat.load_catalogue_file(argv[3], "x,y,z,nz");
      trv::ParameterSet p;
      p.boxsize[0] = p.boxsize[1] = p.boxsize[2] = 1000.;
      p.ngrid[0] = p.ngrid[1] = p.ngrid[2] = 64;
      std::string kind = argv[5];
      p.statistic_type = kind;
      if (kind == "powspec" || kind == "2pcf") p.ELL = 2;
      else { p.ell1 = 2; p.ell2 = 0; p.ELL = 2; }
      p.validate();
      std::FILE* f = std::fopen(argv[4], "w");
      if (f == nullptr) return 3;
      trv::io::print_measurement_header_to_file(f, p, cat, 3.5e7, 4.25e3, 0.);
      int dim = static_cast<int>(cols[0].size());
      if (kind == "bispec") {
        trv::BispecMeasurements m;
        m.dim = dim;
        m.k1_bin = cols[0]; m.k1_eff = cols[1]; m.nmodes_1 = as<int>(cols[2]);
        m.k2_bin = cols[3]; m.k2_eff = cols[4]; m.nmodes_2 = as<int>(cols[5]);
        m.bk_raw = cplx(cols[6], cols[7]);
        m.bk_shot = cplx(cols[8], cols[9]);
        trv::io::print_measurement_datatab_to_file(f, p, m);
      } else if (kind == "3pcf") {
        trv::ThreePCFWindowMeasurements m;
        m.dim = dim;
        m.r1_bin = cols[0]; m.r1_eff = cols[1]; m.npairs_1 = as<int>(cols[2]);
        m.r2_bin = cols[3]; m.r2_eff = cols[4]; m.npairs_2 = as<int>(cols[5]);
        m.zeta_raw = cplx(cols[6], cols[7]);
        m.zeta_shot = cplx(cols[8], cols[9]);
        trv::io::print_measurement_datatab_to_file(f, p, m);
      } else if (kind == "powspec") {
        trv::PowspecMeasurements m;
        m.dim = dim;
        m.kbin = cols[0]; m.keff = cols[1]; m.nmodes = as<int>(cols[2]);
        m.pk_raw = cplx(cols[3], cols[4]);
        m.pk_shot = cplx(cols[5], cols[6]);
        trv::io::print_measurement_datatab_to_file(f, p, m);
      } else {
        trv::TwoPCFMeasurements m;
        m.dim = dim;
        m.rbin = cols[0]; m.reff = cols[1]; m.npairs = as<int>(cols[2]);
        m.xi = cplx(cols[3], cols[4]);
        trv::io::print_measurement_datatab_to_file(f, p, m);
      }
      std::fclose(f);
      return 0;
    }
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    return 1;
  }
  return 2;
}
