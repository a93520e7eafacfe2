// engine_twopt.hip -- two-point statistics: power spectrum and correlation
// function multipoles of survey-type catalogue pairs (local plane-parallel),
// of periodic boxes (global plane-parallel), and the correlation-function
// window of a random catalogue (src/twopt.cpp:388-901 over
// src/field.cpp:1208-1655 and 2511-2945).
//
// Schedule: dn_00(k) once; for each M = -ELL..ELL the field dn_LM(k), its
// shot-noise amplitude (all M in one pass per catalogue) and the coupled
// statistic of dn_LM against dn_00 (only m1 = -M couples).  Interlaced
// fields add the half-cell-shifted shadow assignment, its transform and
// the phase-weighted average.  Fields are FP64 (the shot-noise subtraction
// and the mean subtraction of a box are cancellations).
#include <cmath>

#include "engine.hpp"

namespace trvg {

namespace {

// Average of a transformed field and its transformed shadow, the shadow
// rotated by the half-cell phase pi (m_x + m_y + m_z), m = i / n below
// n / 2 and i / n - 1 from it (field.cpp:1609-1650).
__global__ void k_interlace(double2* __restrict__ f, const double2* __restrict__ fs,
                            Geom g) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < g.nmesh;
       c += (long long)gridDim.x * blockDim.x) {
    const int k = int(c % g.n[2]);
    const long long ij = c / g.n[2];
    const int j = int(ij % g.n[1]), i = int(ij / g.n[1]);
    const double m0 = (i < g.n[0] / 2) ? double(i) / g.n[0] : double(i) / g.n[0] - 1;
    const double m1 = (j < g.n[1] / 2) ? double(j) / g.n[1] : double(j) / g.n[1] - 1;
    const double m2 = (k < g.n[2] / 2) ? double(k) / g.n[2] : double(k) / g.n[2] - 1;
    double sn, cs;
    sincos(M_PI * (m0 + m1 + m2), &sn, &cs);
    const double2 a = f[c], b = fs[c];
    const double re = a.x + (cs * b.x - sn * b.y);
    const double im = a.y + (sn * b.x + cs * b.y);
    f[c] = dmk(re / 2., im / 2.);
  }
}

template <int POW>
__global__ void k_pow_sum(const double2* __restrict__ f, const double2* __restrict__ h,
                          long long n, double* __restrict__ part) {
  __shared__ double red[RED_THREADS];
  double s = 0.;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x)
    s += POW == 2 ? f[c].x * f[c].x : f[c].x * h[c].x;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int q = RED_THREADS / 2; q > 0; q >>= 1) {
    if (threadIdx.x < q) red[threadIdx.x] += red[threadIdx.x + q];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

void zero_mesh(Context& ctx, double2* f) {
  hipLaunchKernelGGL((k_zero<double>), dim3(grid_for(ctx, ctx.g.nmesh)), dim3(RED_THREADS),
                     0, ctx.stream, f, ctx.g.nmesh);
  TRVG_LAUNCH_CHECK();
}

// The half-cell-shifted pass of field.cpp's interlacing, added into fs.
template <int ORDER>
void shadow_order(Context& ctx, double2* fs, const DeviceCatalogue& c, int mode, int ell,
                  int m, double scale) {
  const int cplx_w = (mode == WM_YLM_W && m != 0) ? 1 : 0;
  hipLaunchKernelGGL((k_assign<double, ORDER>), dim3(grid_for(ctx, c.n, 16)), dim3(256), 0,
                     ctx.stream, fs, c.view(), ctx.g, mode, ell, m, scale, cplx_w, 1,
                     nullptr);
  TRVG_LAUNCH_CHECK();
}

void shadow_add(Context& ctx, double2* fs, const DeviceCatalogue& c, int mode, int ell,
                int m, double scale) {
  if (mode == WM_YLM_W && c.los.get() == nullptr)
    throw Error(TRVG_ERR_DATA, "Lines of sight are required for survey-type fields.");
  ProfScope ps(ctx, KC_ASSIGN, double(c.n) * 56. + double(ctx.g.nmesh) * 16.);
  switch (ctx.p.assignment_order) {
    case 1: shadow_order<1>(ctx, fs, c, mode, ell, m, scale); break;
    case 2: shadow_order<2>(ctx, fs, c, mode, ell, m, scale); break;
    case 3: shadow_order<3>(ctx, fs, c, mode, ell, m, scale); break;
    default: shadow_order<4>(ctx, fs, c, mode, ell, m, scale); break;
  }
}

// Which field a two-point call builds.
enum FieldKind {
  FK_PAIR,     // sum_d y w - alpha sum_r y w       (field.cpp:1246-1323)
  FK_SINGLE,   // alpha sum_r y w; shadow unscaled  (field.cpp:1325-1362)
  FK_BOX       // n - N / V on the main mesh only   (field.cpp:1225-1244)
};

// One transformed (and, with interlacing, shadow-averaged) field of (ell, m)
// into f; fs is a scratch mesh used when interlacing.
void twopt_field(Context& ctx, double2* f, double2* fs, FieldKind kind,
                 const DeviceCatalogue* d, const DeviceCatalogue& r, double alpha,
                 int ell, int m) {
  const bool il = ctx.p.interlace != 0;
  zero_mesh(ctx, f);
  double shift = 0.;
  if (kind == FK_PAIR) {
    assign_add<double>(ctx, f, *d, WM_YLM_W, ell, m, 1.);
    assign_add<double>(ctx, f, r, WM_YLM_W, ell, m, -alpha);
  } else if (kind == FK_SINGLE) {
    assign_add<double>(ctx, f, r, WM_YLM_W, ell, m, alpha);
  } else {
    assign_add<double>(ctx, f, r, WM_UNIT, 0, 0, 1.);
    shift = -(double(r.n) / ctx.g.vol);
  }
  fft3d<double>(ctx, f, LoadScaleShift<double>{f, shift, ctx.g.vol_cell}, -1);
  if (!il) return;
  zero_mesh(ctx, fs);
  if (kind == FK_PAIR) {
    shadow_add(ctx, fs, *d, WM_YLM_W, ell, m, 1.);
    shadow_add(ctx, fs, r, WM_YLM_W, ell, m, -alpha);
  } else if (kind == FK_SINGLE) {
    shadow_add(ctx, fs, r, WM_YLM_W, ell, m, 1.);
  } else {
    shadow_add(ctx, fs, r, WM_UNIT, 0, 0, 1.);
  }
  fft3d<double>(ctx, fs, LoadScaleShift<double>{fs, 0., ctx.g.vol_cell}, -1);
  ProfScope ps(ctx, KC_MISC, 3. * double(ctx.g.nmesh) * 16.);
  hipLaunchKernelGGL(k_interlace, dim3(grid_for(ctx, ctx.g.nmesh)), dim3(RED_THREADS), 0,
                     ctx.stream, f, fs, ctx.g);
  TRVG_LAUNCH_CHECK();
}

// sum_p y_lm(los_p) w_p^2 of one catalogue for (ell, M), M = -ell..ell.
std::vector<cplx> sn_sums(Context& ctx, const DeviceCatalogue& c, int ell) {
  std::vector<cplx> out(2 * ell + 1, cplx(0., 0.));
  if (c.los.get() == nullptr)
    throw Error(TRVG_ERR_DATA, "Lines of sight are required for survey-type fields.");
  for (int M0 = -ell; M0 <= ell; M0 += SBAR_MAX_KEYS) {
    SbarKeys sk{};
    sk.wpow = 2;
    sk.n = std::min(SBAR_MAX_KEYS, ell - M0 + 1);
    for (int q = 0; q < sk.n; q++) {
      sk.ell[q] = ell;
      sk.m[q] = M0 + q;
    }
    const unsigned grid = grid_for(ctx, c.n, 4);
    DevArray<double2> part((size_t)grid * sk.n);
    hipLaunchKernelGGL(k_sbar_multi, dim3(grid), dim3(RED_THREADS), 0, ctx.stream,
                       c.view(), sk, part.get());
    TRVG_LAUNCH_CHECK();
    const double2* h = ctx.download(part.get(), (size_t)grid * sk.n);
    ctx.keep_until_flush(std::move(part));
    ctx.flush();
    for (int q = 0; q < sk.n; q++) {
      double sx = 0., sy = 0.;
      for (unsigned b = 0; b < grid; b++) {
        sx += h[b * sk.n + q].x;
        sy += h[b * sk.n + q].y;
      }
      out[M0 + q + ell] = cplx(sx, sy);
    }
  }
  return out;
}

struct BinnedStat {
  std::vector<int> count;
  std::vector<double> eff;
  std::vector<cplx> val, sn;
};

// FieldStats::compute_ylm_wgtd_2pt_stats_in_fourier (field.cpp:2511-2703):
// per bin n_modes, mean |k|, mean (fa fb^* / W) y and mean (S C1 / W) y.
BinnedStat stats_fourier(Context& ctx, const double2* fa, const double2* fb, cplx sbar,
                         int ell, int m, const trvg_binning& b) {
  const int nb = b.num_bins;
  std::shared_ptr<const TwoptMapData> map = twopt_map(ctx, b.bin_edges, nb);
  const unsigned grid = grid_for(ctx, ctx.g.nmesh, 8);
  DevArray<unsigned long long> cnt(nb);
  DevArray<double> part((size_t)grid * nb * 5), sum((size_t)nb * 5);
  TRVG_HIP(hipMemsetAsync(cnt.get(), 0, nb * 8, ctx.stream));
  {
    ProfScope ps(ctx, KC_TWOPT, 2. * double(ctx.g.nmesh) * 16.);
    auto kern = ctx.tab.interlace ? k_twopt_fourier<double, 5, true>
                                  : k_twopt_fourier<double, 5, false>;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(RED_THREADS),
                       twopt_smem(nb, 5, ctx.g), ctx.stream, fa, fb, ctx.g, ctx.tab,
                       dmk(sbar.real(), sbar.imag()), ell, m, map->view(), nb, cnt.get(),
                       part.get());
    TRVG_LAUNCH_CHECK();
  }
  reduce_cols(ctx, part.get(), int(grid), nb * 5, sum.get());
  const unsigned long long* hc = ctx.download(cnt.get(), nb);
  const double* hs = ctx.download(sum.get(), (size_t)nb * 5);
  ctx.keep_until_flush(std::move(cnt));
  ctx.keep_until_flush(std::move(part));
  ctx.keep_until_flush(std::move(sum));
  ctx.flush();
  BinnedStat s;
  s.count.resize(nb);
  s.eff.resize(nb);
  s.val.resize(nb);
  s.sn.resize(nb);
  for (int i = 0; i < nb; i++) {
    s.count[i] = int(hc[i]);
    if (hc[i] != 0) {
      const double n = double(hc[i]);
      s.val[i] = cplx(hs[5 * i], hs[5 * i + 1]) / n;
      s.sn[i] = cplx(hs[5 * i + 2], hs[5 * i + 3]) / n;
      s.eff[i] = hs[5 * i + 4] / n;
    } else {
      s.val[i] = s.sn[i] = cplx(0., 0.);
      s.eff[i] = b.bin_centres[i];
    }
  }
  return s;
}

// FieldStats::compute_ylm_wgtd_2pt_stats_in_config (field.cpp:2705-2945):
// the shot-noise-subtracted (fa fb^* / W - S C1 / W) / V transformed back
// into t, then per bin n_pairs, mean r and mean t y_lm(r^) over fine bins of
// width 0.1.
BinnedStat stats_config(Context& ctx, double2* t, const double2* fa, const double2* fb,
                        cplx sbar, int ell, int m, const trvg_binning& b) {
  const int nb = b.num_bins;
  LoadTwopt<double, double> lt{fa, fb, ctx.tab, dmk(sbar.real(), sbar.imag()),
                               1. / ctx.g.vol};
  fft3d<double>(ctx, t, lt, +1, 2. * double(ctx.g.nmesh) * 16.);
  DevArray<double> edges(nb + 1);
  ctx.upload(edges.get(), b.bin_edges, nb + 1);
  const unsigned grid = grid_for(ctx, ctx.g.nmesh, 4);
  DevArray<unsigned long long> cnt(nb);
  DevArray<double> part((size_t)grid * nb * 3), sum((size_t)nb * 3);
  TRVG_HIP(hipMemsetAsync(cnt.get(), 0, nb * 8, ctx.stream));
  {
    ProfScope ps(ctx, KC_TWOPT, double(ctx.g.nmesh) * 16.);
    hipLaunchKernelGGL((k_3pcf_shot_bins<double, true>), dim3(grid), dim3(RED_THREADS),
                       size_t(nb) * (3 * 8 + 4), ctx.stream, t, ctx.g, ell, m, 0, 0,
                       edges.get(), nb, cnt.get(), part.get());
    TRVG_LAUNCH_CHECK();
  }
  reduce_cols(ctx, part.get(), int(grid), nb * 3, sum.get());
  const unsigned long long* hc = ctx.download(cnt.get(), nb);
  const double* hs = ctx.download(sum.get(), (size_t)nb * 3);
  ctx.keep_until_flush(std::move(edges));
  ctx.keep_until_flush(std::move(cnt));
  ctx.keep_until_flush(std::move(part));
  ctx.keep_until_flush(std::move(sum));
  ctx.flush();
  BinnedStat s;
  s.count.resize(nb);
  s.eff.resize(nb);
  s.val.resize(nb);
  s.sn.assign(nb, cplx(0., 0.));
  for (int i = 0; i < nb; i++) {
    s.count[i] = int(hc[i]);
    if (hc[i] != 0) {
      const double n = double(hc[i]);
      s.eff[i] = hs[3 * i] / n;
      s.val[i] = cplx(hs[3 * i + 1], hs[3 * i + 2]) / n;
    } else {
      s.eff[i] = b.bin_centres[i];
      s.val[i] = cplx(0., 0.);
    }
  }
  return s;
}

// (2 ell + 1)(2 L + 1) (ell 0 L; 0 0 0)(ell 0 L; m 0 M) (twopt.cpp:45-49).
double coupling_2pt(int ell, int L, int m, int M) {
  return (2 * ell + 1) * (2 * L + 1) * wigner_3j(ell, 0, L, 0, 0, 0) *
         wigner_3j(ell, 0, L, m, 0, M);
}

TwoPtResult finish(const trvg_binning& b, std::vector<int>&& count,
                   std::vector<double>&& eff, std::vector<cplx>&& raw,
                   std::vector<cplx>&& shot, double norm) {
  TwoPtResult out;
  out.bin.assign(b.bin_centres, b.bin_centres + b.num_bins);
  out.count = std::move(count);
  out.eff = std::move(eff);
  out.raw = std::move(raw);
  out.shot = std::move(shot);
  for (auto& v : out.raw) v *= norm;
  for (auto& v : out.shot) v *= norm;
  return out;
}

}  // namespace

TwoPtResult run_twopt_survey(Context& ctx, const DeviceCatalogue* d,
                             const DeviceCatalogue& r, bool fourier,
                             const trvg_binning& b, double alpha, double norm) {
  const int L = ctx.p.ELL, ell1 = ctx.p.ELL, nb = b.num_bins;
  if (L < 0) throw Error(TRVG_ERR_PARAM, "Multipole degree must be non-negative.");
  const FieldKind kind = d ? FK_PAIR : FK_SINGLE;
  const long long nm = ctx.g.nmesh;
  DevArray<double2> dn00(nm), dnLM(nm), work(nm);
  twopt_field(ctx, dn00.get(), work.get(), kind, d, r, alpha, 0, 0);
  // Shot-noise amplitudes sum_d y w^2 + alpha^2 sum_r y w^2 (twopt.cpp:298-
  // 381), for every M.
  std::vector<cplx> snr = sn_sums(ctx, r, L), snd;
  if (d) snd = sn_sums(ctx, *d, L);
  std::vector<int> count(nb, 0);
  std::vector<double> eff(nb, 0.);
  std::vector<cplx> raw(nb, cplx(0., 0.)), shot(nb, cplx(0., 0.));
  for (int M = -L; M <= L; M++) {
    twopt_field(ctx, dnLM.get(), work.get(), kind, d, r, alpha, L, M);
    const cplx sn_amp = (d ? snd[M + L] : cplx(0., 0.)) + std::pow(alpha, 2) * snr[M + L];
    for (int m1 = -ell1; m1 <= ell1; m1++) {
      const double coupling = coupling_2pt(ell1, L, m1, M);
      if (std::fabs(coupling) < 1.e-10) continue;
      BinnedStat s = fourier ? stats_fourier(ctx, dnLM.get(), dn00.get(), sn_amp, ell1, m1, b)
                             : stats_config(ctx, work.get(), dnLM.get(), dn00.get(), sn_amp,
                                            ell1, m1, b);
      for (int i = 0; i < nb; i++) {
        raw[i] += coupling * s.val[i];
        shot[i] += coupling * s.sn[i];
      }
      if (M == 0 && m1 == 0) {
        count = s.count;
        eff = s.eff;
      }
    }
  }
  ctx.flush();
  return finish(b, std::move(count), std::move(eff), std::move(raw), std::move(shot), norm);
}

TwoPtResult run_twopt_box(Context& ctx, const DeviceCatalogue& d, bool fourier,
                          const trvg_binning& b, double norm) {
  const int L = ctx.p.ELL, nb = b.num_bins;
  if (L < 0) throw Error(TRVG_ERR_PARAM, "Multipole degree must be non-negative.");
  const long long nm = ctx.g.nmesh;
  DevArray<double2> dn(nm), work(nm);
  twopt_field(ctx, dn.get(), work.get(), FK_BOX, nullptr, d, 1., 0, 0);
  const cplx sn_amp(double(d.n), 0.);
  BinnedStat s = fourier ? stats_fourier(ctx, dn.get(), dn.get(), sn_amp, L, 0, b)
                         : stats_config(ctx, work.get(), dn.get(), dn.get(), sn_amp, L, 0, b);
  const double f = double(2 * L + 1);
  std::vector<cplx> raw(nb), shot(nb);
  for (int i = 0; i < nb; i++) {
    raw[i] = f * s.val[i];
    shot[i] = f * s.sn[i];
  }
  ctx.flush();
  return finish(b, std::move(s.count), std::move(s.eff), std::move(raw), std::move(shot),
                norm);
}

double mesh_square_integral(Context& ctx, const DeviceCatalogue& c) {
  DevArray<double2> f(ctx.g.nmesh);
  zero_mesh(ctx, f.get());
  assign_add<double>(ctx, f.get(), c, WM_W, 0, 0, 1.);
  const unsigned grid = grid_for(ctx, ctx.g.nmesh, 4);
  DevArray<double> part(grid), sum(1);
  hipLaunchKernelGGL((k_pow_sum<2>), dim3(grid), dim3(RED_THREADS), 0, ctx.stream, f.get(),
                     f.get(), ctx.g.nmesh, part.get());
  TRVG_LAUNCH_CHECK();
  reduce_cols(ctx, part.get(), int(grid), 1, sum.get());
  double h = 0.;
  TRVG_HIP(hipMemcpyAsync(&h, sum.get(), 8, hipMemcpyDeviceToHost, ctx.stream));
  TRVG_HIP(hipStreamSynchronize(ctx.stream));
  return h * ctx.g.vol_cell;
}

double mesh_cross_integral(Context& ctx, const DeviceCatalogue& d, const DeviceCatalogue& r) {
  DevArray<double2> fd(ctx.g.nmesh), fr(ctx.g.nmesh);
  zero_mesh(ctx, fd.get());
  zero_mesh(ctx, fr.get());
  assign_add<double>(ctx, fd.get(), d, WM_W, 0, 0, 1.);
  assign_add<double>(ctx, fr.get(), r, WM_W, 0, 0, 1.);
  const unsigned grid = grid_for(ctx, ctx.g.nmesh, 4);
  DevArray<double> part(grid), sum(1);
  hipLaunchKernelGGL((k_pow_sum<1>), dim3(grid), dim3(RED_THREADS), 0, ctx.stream, fd.get(),
                     fr.get(), ctx.g.nmesh, part.get());
  TRVG_LAUNCH_CHECK();
  reduce_cols(ctx, part.get(), int(grid), 1, sum.get());
  double h = 0.;
  TRVG_HIP(hipMemcpyAsync(&h, sum.get(), 8, hipMemcpyDeviceToHost, ctx.stream));
  TRVG_HIP(hipStreamSynchronize(ctx.stream));
  return h * ctx.g.vol_cell;
}

}  // namespace trvg
