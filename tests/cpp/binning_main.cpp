// Prints the bins of the C++ drop-in trv::Binning for every scheme: one
// line per scheme, "scheme edges... | centres... | widths...".
#include <cstdio>

#include "triumvirate_gpu.hpp"

static void dump(const trv::Binning& b) {
  std::printf("%s", b.scheme.c_str());
  for (double v : b.bin_edges) std::printf(" %.17g", v);
  std::printf(" |");
  for (double v : b.bin_centres) std::printf(" %.17g", v);
  std::printf(" |");
  for (double v : b.bin_widths) std::printf(" %.17g", v);
  std::printf("\n");
}

int main() {
  const char* schemes[] = {"lin", "log", "linpad", "logpad"};
  for (const char* sch : schemes) {
    trv::Binning f("fourier", sch);
    f.set_bins(0.01, 0.2, 12);
    dump(f);
    trv::Binning c("config", sch);
    c.set_bins(20., 180., 10);
    dump(c);
  }
  trv::ParameterSet ps;
  ps.boxsize[0] = 1000.; ps.boxsize[1] = 1500.; ps.boxsize[2] = 800.;
  ps.ngrid[0] = 64; ps.ngrid[1] = 128; ps.ngrid[2] = 96;
  ps.space = "fourier";
  ps.binning = "linpad";
  ps.bin_min = 0.;
  ps.bin_max = 0.2;
  ps.num_bins = 15;
  trv::Binning g(ps);
  g.set_bins();
  dump(g);
  trv::Binning h("config", "lin");
  h.set_bins(1500., 64);
  dump(h);
  return 0;
}
