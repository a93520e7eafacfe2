// engine_drivers.hip -- estimator drivers of the gfx950 engine.
//
// The drivers compute what src/triumvirate/src/threept.cpp computes, term by
// term in the same (m1, m2, M) enumeration, but each distinct quantity once:
//   * one shell-filtered IFFT per distinct (bin, side) instead of two per
//     bin pair (threept.cpp:664-669, 725-730);
//   * one shot-noise IFFT per term instead of one per bin pair
//     (field.cpp:3269-3345 called at threept.cpp:950-964), followed by a
//     radial histogram so each pair costs O(#radii) spline evaluations;
//   * G_LM and the shot-noise field dn_LM share one assignment + FFT
//     (threept.cpp:452-457, 777-784), and for (L, M) = (0, 0) both are the
//     already-transformed dn_00.
// Results differ from the reference only by floating-point summation order.
#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>

#include "engine.hpp"

namespace trvg {

// ---------------------------------------------------------------------------
// FFT
// ---------------------------------------------------------------------------

template <typename T>
static const c2<T>* twiddles(int n, int sign) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, c2<T>*> cache;
  int dev = 0;
  TRVG_HIP(hipGetDevice(&dev));
  auto key = std::make_tuple(dev, n, sign);
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  std::vector<c2<T>> h(n);
  for (int m = 0; m < n; m++) {
    double a = 2. * M_PI * double(m) / double(n);
    h[m] = mk<T>(T(std::cos(a)), T(sign * std::sin(a)));
  }
  c2<T>* d = nullptr;
  TRVG_HIP(hipMalloc(&d, n * sizeof(c2<T>)));
  TRVG_HIP(hipMemcpy(d, h.data(), n * sizeof(c2<T>), hipMemcpyHostToDevice));
  cache[key] = d;
  return d;
}

template <typename T> struct FftEpt { static constexpr int v = 8; };
template <> struct FftEpt<float> { static constexpr int v = 16; };

template <typename T, int LOGN, class Loader>
static void launch_z(Context& ctx, c2<T>* data, const Loader& load, int sign) {
  constexpr int EPT = FftEpt<T>::v;
  constexpr int N = 1 << LOGN;
  constexpr int E = FFT_THREADS * EPT;
  constexpr int NL = E / N;
  size_t smem = size_t(NL * (N + 1) + N) * sizeof(c2<T>);
  long long nblocks = ctx.g.nmesh / E;
  hipLaunchKernelGGL((fft_pass_z<T, LOGN, EPT, Loader>), dim3(unsigned(nblocks)),
                     dim3(FFT_THREADS), smem, ctx.stream, data, load,
                     twiddles<T>(N, sign), ctx.g.n[0], ctx.g.n[1], sign);
  TRVG_LAUNCH_CHECK();
}

template <typename T, int LOGN>
static void launch_strided(Context& ctx, c2<T>* data, long long S,
                           long long outer_stride, long long inner, int sign) {
  constexpr int EPT = FftEpt<T>::v;
  constexpr int N = 1 << LOGN;
  constexpr int E = FFT_THREADS * EPT;
  constexpr int NL = E / N;
  size_t smem = size_t(NL * (N + 1) + N) * sizeof(c2<T>);
  long long nblocks = ctx.g.nmesh / E;
  hipLaunchKernelGGL((fft_pass_strided<T, LOGN, EPT>), dim3(unsigned(nblocks)),
                     dim3(FFT_THREADS), smem, ctx.stream, data,
                     twiddles<T>(N, sign), S, outer_stride, inner, sign);
  TRVG_LAUNCH_CHECK();
}

static int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) l++;
  return l;
}

#define TRVG_LOGN_SWITCH(lg, CALL)                               \
  switch (lg) {                                                  \
    case 3: { constexpr int LG = 3; CALL; } break;               \
    case 4: { constexpr int LG = 4; CALL; } break;               \
    case 5: { constexpr int LG = 5; CALL; } break;               \
    case 6: { constexpr int LG = 6; CALL; } break;               \
    case 7: { constexpr int LG = 7; CALL; } break;               \
    case 8: { constexpr int LG = 8; CALL; } break;               \
    case 9: { constexpr int LG = 9; CALL; } break;               \
    case 10: { constexpr int LG = 10; CALL; } break;             \
    default: throw Error(TRVG_ERR_PARAM, "Unsupported FFT size."); \
  }

template <typename T, class Loader>
void fft3d(Context& ctx, c2<T>* data, const Loader& load, int sign) {
  const Geom& g = ctx.g;
  TRVG_LOGN_SWITCH(ilog2(g.n[2]), (launch_z<T, LG, Loader>(ctx, data, load, sign)));
  TRVG_LOGN_SWITCH(ilog2(g.n[1]), (launch_strided<T, LG>(
      ctx, data, g.n[2], (long long)g.n[1] * g.n[2], g.n[2], sign)));
  TRVG_LOGN_SWITCH(ilog2(g.n[0]), (launch_strided<T, LG>(
      ctx, data, (long long)g.n[1] * g.n[2], 0, (long long)g.n[1] * g.n[2], sign)));
}

// ---------------------------------------------------------------------------
// Small launch helpers
// ---------------------------------------------------------------------------

static unsigned grid_for(const Context& ctx, long long work, int per_cu = 8) {
  long long b = (work + RED_THREADS - 1) / RED_THREADS;
  long long cap = (long long)ctx.sm_count * per_cu;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return unsigned(b);
}

template <typename T>
static void zero_field(Context& ctx, c2<T>* f) {
  hipLaunchKernelGGL((k_zero<T>), dim3(grid_for(ctx, ctx.g.nmesh)),
                     dim3(RED_THREADS), 0, ctx.stream, f, ctx.g.nmesh);
  TRVG_LAUNCH_CHECK();
}

template <typename T>
static void assign(Context& ctx, c2<T>* mesh, const DeviceCatalogue& c,
                   int mode, int ell, int m, double scale) {
  int cplx_w = ((mode == WM_YLM_W || mode == WM_YLMC_W2) && !(ell == 0 && m == 0));
  unsigned grid = grid_for(ctx, c.n, 16);
  DevParticles pv = c.view();
  if ((mode == WM_YLM_W || mode == WM_YLMC_W2) && pv.los == nullptr)
    throw Error(TRVG_ERR_DATA, "Lines of sight are required for survey-type fields.");
  switch (ctx.p.assignment_order) {
    case 1: hipLaunchKernelGGL((k_assign<T, 1>), dim3(grid), dim3(256), 0, ctx.stream,
                               mesh, pv, ctx.g, mode, ell, m, scale, cplx_w, 0); break;
    case 2: hipLaunchKernelGGL((k_assign<T, 2>), dim3(grid), dim3(256), 0, ctx.stream,
                               mesh, pv, ctx.g, mode, ell, m, scale, cplx_w, 0); break;
    case 3: hipLaunchKernelGGL((k_assign<T, 3>), dim3(grid), dim3(256), 0, ctx.stream,
                               mesh, pv, ctx.g, mode, ell, m, scale, cplx_w, 0); break;
    default: hipLaunchKernelGGL((k_assign<T, 4>), dim3(grid), dim3(256), 0, ctx.stream,
                                mesh, pv, ctx.g, mode, ell, m, scale, cplx_w, 0); break;
  }
  TRVG_LAUNCH_CHECK();
}

// Survey field  sum_d y w - alpha sum_r y w  (quad: conj(y) w^2 with
// +alpha^2), or box field (unit weights, optional mean subtraction), then
// the forward transform with the V_cell factor.
template <typename T>
static void build_field_k(Context& ctx, c2<T>* f, const DeviceCatalogue& d,
                          const DeviceCatalogue* r, bool survey, double alpha,
                          int ell, int m, bool quad, bool fluctuation) {
  zero_field<T>(ctx, f);
  T shift = T(0);
  if (survey) {
    int mode = quad ? WM_YLMC_W2 : WM_YLM_W;
    assign<T>(ctx, f, d, mode, ell, m, 1.);
    assign<T>(ctx, f, *r, mode, ell, m, quad ? alpha * alpha : -alpha);
  } else {
    assign<T>(ctx, f, d, WM_UNIT, 0, 0, 1.);
    if (fluctuation) shift = T(-(double(d.n) / ctx.g.vol));
  }
  LoadScaleShift<T> ld{f, shift, T(ctx.g.vol_cell)};
  fft3d<T>(ctx, f, ld, -1);
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------

struct ShellStats {
  std::vector<int> nmodes;
  std::vector<double> keff;
};

static ShellStats shell_stats(Context& ctx, const trvg_binning& kb) {
  const int nb = kb.num_bins;
  DevArray<double> edges(nb + 1);
  TRVG_HIP(hipMemcpyAsync(edges.get(), kb.bin_edges, (nb + 1) * 8,
                          hipMemcpyHostToDevice, ctx.stream));
  unsigned grid = grid_for(ctx, ctx.g.nmesh, 4);
  DevArray<unsigned long long> cnt(nb);
  DevArray<double> part((size_t)grid * nb), ks(nb);
  TRVG_HIP(hipMemsetAsync(cnt.get(), 0, nb * 8, ctx.stream));
  size_t smem = nb * (8 + 4);
  hipLaunchKernelGGL(k_shell_stats, dim3(grid), dim3(RED_THREADS), smem,
                     ctx.stream, ctx.g, edges.get(), nb, cnt.get(), part.get());
  TRVG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_partials, dim3((nb + 255) / 256), dim3(256), 0,
                     ctx.stream, part.get(), int(grid), nb, ks.get());
  TRVG_LAUNCH_CHECK();
  std::vector<unsigned long long> hc(nb);
  ShellStats s;
  s.keff.resize(nb);
  TRVG_HIP(hipMemcpyAsync(hc.data(), cnt.get(), nb * 8, hipMemcpyDeviceToHost, ctx.stream));
  TRVG_HIP(hipMemcpyAsync(s.keff.data(), ks.get(), nb * 8, hipMemcpyDeviceToHost, ctx.stream));
  TRVG_HIP(hipStreamSynchronize(ctx.stream));
  s.nmodes.resize(nb);
  for (int b = 0; b < nb; b++) {
    s.nmodes[b] = int(hc[b]);
    s.keff[b] /= double(s.nmodes[b]);   // NaN for an empty shell, as k_eff /= 0
  }
  return s;
}

template <typename T>
static std::vector<cplx> pair_reduce(Context& ctx, const c2<T>* FA, int na,
                                     const c2<T>* FB, int nbf, const c2<T>* G,
                                     const std::vector<int2>& pairs) {
  const int P = int(pairs.size());
  std::vector<cplx> out(P);
  const int group = RED_THREADS * PR_MAXP;
  size_t smem = size_t(PR_CELLS) * (na + nbf) * sizeof(double2);
  if (smem > 65536)
    throw Error(TRVG_ERR_PARAM, "Too many distinct shells for one pair reduction.");
  long long nchunks = (ctx.g.nmesh + PR_CELLS - 1) / PR_CELLS;
  unsigned grid = unsigned(std::min<long long>(nchunks, (long long)ctx.sm_count * 4));
  for (int p0 = 0; p0 < P; p0 += group) {
    int np = std::min(group, P - p0);
    DevArray<int2> dp(np);
    TRVG_HIP(hipMemcpyAsync(dp.get(), pairs.data() + p0, np * sizeof(int2),
                            hipMemcpyHostToDevice, ctx.stream));
    DevArray<double2> part((size_t)grid * np), sum(np);
    hipLaunchKernelGGL((k_pair_reduce<T>), dim3(grid), dim3(RED_THREADS), smem,
                       ctx.stream, FA, na, FB, nbf, G, ctx.g.nmesh, dp.get(),
                       np, part.get());
    TRVG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials2, dim3((np + 255) / 256), dim3(256), 0,
                       ctx.stream, part.get(), int(grid), np, sum.get());
    TRVG_LAUNCH_CHECK();
    std::vector<double2> h(np);
    TRVG_HIP(hipMemcpyAsync(h.data(), sum.get(), np * sizeof(double2),
                            hipMemcpyDeviceToHost, ctx.stream));
    TRVG_HIP(hipStreamSynchronize(ctx.stream));
    for (int i = 0; i < np; i++) out[p0 + i] = cplx(h[i].x, h[i].y);
  }
  return out;
}

struct TwoPt {
  std::vector<int> nmodes;
  std::vector<cplx> pk, sn;
};

// FieldStats::compute_ylm_wgtd_2pt_stats_in_fourier (field.cpp:2511-2703).
template <typename T>
static TwoPt twopt_fourier(Context& ctx, const c2<T>* fa, const c2<T>* fb,
                           cplx sbar, int ell, int m, const trvg_binning& kb) {
  const int nb = kb.num_bins;
  DevArray<double> edges(nb + 1);
  TRVG_HIP(hipMemcpyAsync(edges.get(), kb.bin_edges, (nb + 1) * 8,
                          hipMemcpyHostToDevice, ctx.stream));
  unsigned grid = grid_for(ctx, ctx.g.nmesh, 4);
  DevArray<unsigned long long> cnt(nb);
  DevArray<double> part((size_t)grid * nb * 4), sum(nb * 4);
  TRVG_HIP(hipMemsetAsync(cnt.get(), 0, nb * 8, ctx.stream));
  size_t smem = nb * (4 * 8 + 4);
  hipLaunchKernelGGL((k_twopt_fourier<T>), dim3(grid), dim3(RED_THREADS), smem,
                     ctx.stream, fa, fb, ctx.g, ctx.tab,
                     dmk(sbar.real(), sbar.imag()), ell, m, edges.get(), nb,
                     cnt.get(), part.get());
  TRVG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_partials, dim3((nb * 4 + 255) / 256), dim3(256), 0,
                     ctx.stream, part.get(), int(grid), nb * 4, sum.get());
  TRVG_LAUNCH_CHECK();
  std::vector<unsigned long long> hc(nb);
  std::vector<double> hs(nb * 4);
  TRVG_HIP(hipMemcpyAsync(hc.data(), cnt.get(), nb * 8, hipMemcpyDeviceToHost, ctx.stream));
  TRVG_HIP(hipMemcpyAsync(hs.data(), sum.get(), nb * 32, hipMemcpyDeviceToHost, ctx.stream));
  TRVG_HIP(hipStreamSynchronize(ctx.stream));
  TwoPt r;
  r.nmodes.resize(nb);
  r.pk.resize(nb);
  r.sn.resize(nb);
  for (int b = 0; b < nb; b++) {
    r.nmodes[b] = int(hc[b]);
    if (hc[b] != 0) {
      double n = double(hc[b]);
      r.pk[b] = cplx(hs[4 * b], hs[4 * b + 1]) / n;
      r.sn[b] = cplx(hs[4 * b + 2], hs[4 * b + 3]) / n;
    } else {
      r.pk[b] = 0.;
      r.sn[b] = 0.;
    }
  }
  return r;
}

struct ShotBins3pcf {
  std::vector<int> npairs;
  std::vector<double> r;
  std::vector<cplx> xi;
};

// FieldStats::compute_uncoupled_shotnoise_for_3pcf binning stage
// (field.cpp:3097-3190).
template <typename T>
static ShotBins3pcf shot_bins_3pcf(Context& ctx, const c2<T>* t, int la, int ma,
                                   int lb, int mb, const trvg_binning& rb) {
  const int nb = rb.num_bins;
  DevArray<double> edges(nb + 1);
  TRVG_HIP(hipMemcpyAsync(edges.get(), rb.bin_edges, (nb + 1) * 8,
                          hipMemcpyHostToDevice, ctx.stream));
  unsigned grid = grid_for(ctx, ctx.g.nmesh, 4);
  DevArray<unsigned long long> cnt(nb);
  DevArray<double> part((size_t)grid * nb * 3), sum(nb * 3);
  TRVG_HIP(hipMemsetAsync(cnt.get(), 0, nb * 8, ctx.stream));
  size_t smem = nb * (3 * 8 + 4);
  hipLaunchKernelGGL((k_3pcf_shot_bins<T>), dim3(grid), dim3(RED_THREADS), smem,
                     ctx.stream, t, ctx.g, la, ma, lb, mb, edges.get(), nb,
                     cnt.get(), part.get());
  TRVG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_partials, dim3((nb * 3 + 255) / 256), dim3(256), 0,
                     ctx.stream, part.get(), int(grid), nb * 3, sum.get());
  TRVG_LAUNCH_CHECK();
  std::vector<unsigned long long> hc(nb);
  std::vector<double> hs(nb * 3);
  TRVG_HIP(hipMemcpyAsync(hc.data(), cnt.get(), nb * 8, hipMemcpyDeviceToHost, ctx.stream));
  TRVG_HIP(hipMemcpyAsync(hs.data(), sum.get(), nb * 24, hipMemcpyDeviceToHost, ctx.stream));
  TRVG_HIP(hipStreamSynchronize(ctx.stream));
  ShotBins3pcf s;
  s.npairs.resize(nb);
  s.r.resize(nb);
  s.xi.resize(nb);
  const double nf = 1 / ctx.g.vol_cell * (((ctx.p.ell1 + ctx.p.ell2) % 2) ? -1. : 1.);
  for (int b = 0; b < nb; b++) {
    s.npairs[b] = int(hc[b]);
    if (hc[b] != 0) {
      double n = double(hc[b]);
      s.r[b] = hs[3 * b] / n;
      s.xi[b] = cplx(hs[3 * b + 1], hs[3 * b + 2]) / n;
      s.xi[b] *= nf / n;
    } else {
      s.r[b] = rb.bin_centres[b];
      s.xi[b] = 0.;
    }
  }
  return s;
}

// S_ij_k for each (k_a, k_b) (field.cpp:3198-3396) from the shot-noise
// field t(x).
template <typename T>
static std::vector<cplx> shot_pairs_bispec(Context& ctx, const c2<T>* t, int la,
                                           int ma, int lb, int mb,
                                           const Spline& sa, const Spline& sb,
                                           const std::vector<double>& ka,
                                           const std::vector<double>& kb) {
  const int P = int(ka.size());
  const Geom& g = ctx.g;
  std::vector<cplx> out(P);
  DevArray<double> dka(P), dkb(P);
  TRVG_HIP(hipMemcpyAsync(dka.get(), ka.data(), P * 8, hipMemcpyHostToDevice, ctx.stream));
  TRVG_HIP(hipMemcpyAsync(dkb.get(), kb.data(), P * 8, hipMemcpyHostToDevice, ctx.stream));
  const bool cubic = (g.dr[0] == g.dr[1] && g.dr[1] == g.dr[2]);
  if (cubic) {
    long long nq = (long long)(g.n[0] / 2) * (g.n[0] / 2)
                   + (long long)(g.n[1] / 2) * (g.n[1] / 2)
                   + (long long)(g.n[2] / 2) * (g.n[2] / 2) + 1;
    DevArray<double> Q(2 * nq);
    hipLaunchKernelGGL(k_zero_d, dim3(grid_for(ctx, 2 * nq)), dim3(RED_THREADS), 0,
                       ctx.stream, Q.get(), 2 * nq);
    TRVG_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_radial_hist<T>), dim3(grid_for(ctx, g.nmesh, 8)),
                       dim3(RED_THREADS), 0, ctx.stream, t, g, la, ma, lb, mb,
                       Q.get());
    TRVG_LAUNCH_CHECK();
    DevArray<double2> res(P);
    hipLaunchKernelGGL(k_shot_pairs_hist, dim3(P), dim3(RED_THREADS), 0,
                       ctx.stream, Q.get(), nq, g.dr[0], sa.dev(), sb.dev(),
                       dka.get(), dkb.get(), g.vol_cell, res.get());
    TRVG_LAUNCH_CHECK();
    std::vector<double2> h(P);
    TRVG_HIP(hipMemcpyAsync(h.data(), res.get(), P * sizeof(double2),
                            hipMemcpyDeviceToHost, ctx.stream));
    TRVG_HIP(hipStreamSynchronize(ctx.stream));
    for (int p = 0; p < P; p++) out[p] = cplx(h[p].x, h[p].y);
    return out;
  }
  const int GP = 8;
  unsigned grid = grid_for(ctx, g.nmesh, 4);
  for (int p0 = 0; p0 < P; p0 += GP) {
    int np = std::min(GP, P - p0);
    DevArray<double2> part((size_t)grid * np), sum(np);
    hipLaunchKernelGGL((k_shot_pairs_direct<T>), dim3(grid), dim3(RED_THREADS), 0,
                       ctx.stream, t, g, la, ma, lb, mb, sa.dev(), sb.dev(),
                       dka.get() + p0, dkb.get() + p0, np, part.get());
    TRVG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials2, dim3(1), dim3(256), 0, ctx.stream,
                       part.get(), int(grid), np, sum.get());
    TRVG_LAUNCH_CHECK();
    std::vector<double2> h(np);
    TRVG_HIP(hipMemcpyAsync(h.data(), sum.get(), np * sizeof(double2),
                            hipMemcpyDeviceToHost, ctx.stream));
    TRVG_HIP(hipStreamSynchronize(ctx.stream));
    for (int q = 0; q < np; q++) out[p0 + q] = cplx(h[q].x, h[q].y) * g.vol_cell;
  }
  return out;
}

static cplx sbar_catalogue(Context& ctx, const DeviceCatalogue& c, int ell, int m) {
  unsigned grid = grid_for(ctx, c.n, 4);
  DevArray<double2> part(grid);
  hipLaunchKernelGGL(k_sbar, dim3(grid), dim3(RED_THREADS), 0, ctx.stream,
                     c.view(), ell, m, part.get());
  TRVG_LAUNCH_CHECK();
  std::vector<double2> h(grid);
  TRVG_HIP(hipMemcpyAsync(h.data(), part.get(), grid * sizeof(double2),
                          hipMemcpyDeviceToHost, ctx.stream));
  TRVG_HIP(hipStreamSynchronize(ctx.stream));
  double sx = 0., sy = 0.;
  for (unsigned b = 0; b < grid; b++) { sx += h[b].x; sy += h[b].y; }
  return cplx(sx, sy);
}

// ---------------------------------------------------------------------------
// Term enumeration shared by the bispectrum and 3PCF drivers
// (threept.cpp:354-419, 1580-1640).
// ---------------------------------------------------------------------------

struct Term {
  int m1, m2, M;
  double coupling, f_mirror, f_mirror_w3j;
};

static std::vector<Term> enumerate_terms(const trvg_params& p, bool survey) {
  std::vector<Term> terms;
  auto inverse = [](const Term& a, int m1, int m2, int M) {
    return a.m1 + m1 == 0 && a.m2 + m2 == 0 && a.M + M == 0;
  };
  for (int m1 = -p.ell1; m1 <= p.ell1; m1++) {
    for (int m2 = -p.ell2; m2 <= p.ell2; m2++) {
      if (survey) {
        bool vanish = true;
        for (int M = -p.ELL; M <= p.ELL; M++)
          if (std::fabs(coupling_3pt(p.ell1, p.ell2, p.ELL, m1, m2, M)) > 1.e-9) {
            vanish = false;
            break;
          }
        if (vanish) continue;
      }
      int Mlo = survey ? -p.ELL : 0, Mhi = survey ? p.ELL : 0;
      for (int M = Mlo; M <= Mhi; M++) {
        bool redundant = false;
        for (auto& t : terms) if (inverse(t, m1, m2, M)) redundant = true;
        if (redundant) continue;
        double c = coupling_3pt(p.ell1, p.ell2, p.ELL, m1, m2, M);
        if (std::fabs(c) < 1.e-9) continue;
        bool zeros = (m1 == 0 && m2 == 0 && M == 0);
        Term t{m1, m2, M, c, zeros ? 0. : ((p.ELL % 2) ? -1. : 1.), zeros ? 0. : 1.};
        terms.push_back(t);
      }
    }
  }
  return terms;
}

static cplx ipow(int n) {
  static const cplx I(0., 1.);
  return std::pow(I, n);  // threept.cpp:271-272
}

// Distinct bins of a side and the slab index of each bin.
static std::vector<int> distinct(const std::vector<int>& v) {
  std::vector<int> u(v);
  std::sort(u.begin(), u.end());
  u.erase(std::unique(u.begin(), u.end()), u.end());
  return u;
}

// ---------------------------------------------------------------------------
// Bispectrum
// ---------------------------------------------------------------------------

template <typename T>
Result run_bispec(Context& ctx, const DeviceCatalogue& d,
                  const DeviceCatalogue* r, bool survey,
                  const trvg_binning& kb, double norm_factor) {
  const trvg_params& p = ctx.p;
  const Geom& g = ctx.g;
  const long long nm = g.nmesh;
  const double alpha = survey ? d.wstotal / r->wstotal : 1.;
  const cplx factor_phase = ipow(p.ell1 + p.ell2);
  const int nb = kb.num_bins;
  std::vector<int> ra, cb;
  pair_bins(p.shape, nb, p.idx_bin, ra, cb);
  const int dim = int(ra.size());

  Result out;
  out.dim = dim;
  out.b1.resize(dim); out.b2.resize(dim); out.e1.resize(dim); out.e2.resize(dim);
  out.n1.resize(dim); out.n2.resize(dim);
  out.raw.assign(dim, 0.); out.shot.assign(dim, 0.);

  // Common fields.
  DevArray<c2<T>> dn00(nm), N00(nm);
  build_field_k<T>(ctx, dn00.get(), d, r, survey, alpha, 0, 0, false, true);
  build_field_k<T>(ctx, N00.get(), d, r, survey, alpha, 0, 0, true, false);

  ShellStats ss = shell_stats(ctx, kb);
  for (int dv = 0; dv < dim; dv++) {
    out.b1[dv] = kb.bin_centres[ra[dv]];
    out.b2[dv] = kb.bin_centres[cb[dv]];
    out.e1[dv] = ss.keff[ra[dv]];
    out.e2[dv] = ss.keff[cb[dv]];
    out.n1[dv] = ss.nmodes[ra[dv]];
    out.n2[dv] = ss.nmodes[cb[dv]];
  }

  Spline sja, sjb;
  build_spline(p.ell1, sja, ctx.stream);
  build_spline(p.ell2, sjb, ctx.stream);

  std::vector<int> rows = distinct(ra), cols = distinct(cb);
  DevArray<c2<T>> G(nm), tfield(nm), dnLM, NLM;

  for (const Term& tm : enumerate_terms(p, survey)) {
    const bool lm_zero = !survey || (p.ELL == 0 && tm.M == 0);
    const c2<T>* dnLMp = dn00.get();
    const c2<T>* NLMp = N00.get();
    if (!lm_zero) {
      if (!dnLM.get()) { dnLM.reset(nm); NLM.reset(nm); }
      build_field_k<T>(ctx, dnLM.get(), d, r, true, alpha, p.ELL, tm.M, false, false);
      build_field_k<T>(ctx, NLM.get(), d, r, true, alpha, p.ELL, tm.M, true, false);
      dnLMp = dnLM.get();
      NLMp = NLM.get();
    }

    // G_LM(x): compensated, inverse-transformed (threept.cpp:452-459).
    LoadDeconv<T> ldg{dnLMp, ctx.tab, 1. / g.vol};
    fft3d<T>(ctx, G.get(), ldg, +1);

    // Shell fields F_a (l1, m1) for the row bins, F_b (l2, m2) for columns.
    const bool same_side = (p.ell1 == p.ell2 && tm.m1 == tm.m2);
    DevArray<c2<T>> FA(nm * rows.size());
    std::vector<int> ia(nb, -1), ib(nb, -1);
    for (size_t q = 0; q < rows.size(); q++) {
      int b = rows[q];
      LoadShell<T> ls{dn00.get(), ctx.tab, g, kb.bin_edges[b], kb.bin_edges[b + 1],
                      p.ell1, tm.m1};
      fft3d<T>(ctx, FA.get() + nm * q, ls, +1);
      ia[b] = int(q);
    }
    DevArray<c2<T>> FBown;
    const c2<T>* FBp = FA.get();
    int nfb = int(rows.size());
    bool cols_in_rows = std::all_of(cols.begin(), cols.end(),
                                    [&](int b) { return ia[b] >= 0; });
    if (same_side && cols_in_rows) {
      for (int b : cols) ib[b] = ia[b];
    } else {
      FBown.reset(nm * cols.size());
      for (size_t q = 0; q < cols.size(); q++) {
        int b = cols[q];
        LoadShell<T> ls{dn00.get(), ctx.tab, g, kb.bin_edges[b],
                        kb.bin_edges[b + 1], p.ell2, tm.m2};
        fft3d<T>(ctx, FBown.get() + nm * q, ls, +1);
        ib[b] = int(q);
      }
      FBp = FBown.get();
      nfb = int(cols.size());
    }
    std::vector<int2> pairs(dim);
    for (int dv = 0; dv < dim; dv++) pairs[dv] = make_int2(ia[ra[dv]], ib[cb[dv]]);
    std::vector<cplx> S = pair_reduce<T>(ctx, FA.get(), int(rows.size()), FBp, nfb,
                                         G.get(), pairs);
    for (int dv = 0; dv < dim; dv++) {
      cplx s = S[dv] / (double(ss.nmodes[ra[dv]]) * double(ss.nmodes[cb[dv]]));
      out.raw[dv] += tm.coupling * g.vol_cell * (s + tm.f_mirror * std::conj(s));
    }

    // Shot noise (threept.cpp:771-966).
    cplx Sbar;
    if (survey) {
      Sbar = sbar_catalogue(ctx, d, p.ELL, tm.M)
             + std::pow(alpha, 3) * sbar_catalogue(ctx, *r, p.ELL, tm.M);
    } else {
      Sbar = double(d.n);
    }
    if (p.ell1 == 0 && p.ell2 == 0) {
      cplx sijk = tm.coupling * Sbar;
      for (int dv = 0; dv < dim; dv++) out.shot[dv] += sijk;
    }
    if (p.ell2 == 0) {
      TwoPt st = twopt_fourier<T>(ctx, dn00.get(), NLMp, Sbar, p.ell1, tm.m1, kb);
      for (int dv = 0; dv < dim; dv++) {
        int b = ra[dv];
        cplx s = tm.coupling * (st.pk[b] - st.sn[b]);
        out.shot[dv] += s + tm.f_mirror * std::conj(s);
      }
    }
    if (p.ell1 == 0) {
      TwoPt st = twopt_fourier<T>(ctx, dn00.get(), NLMp, Sbar, p.ell2, tm.m2, kb);
      for (int dv = 0; dv < dim; dv++) {
        int b = cb[dv];
        cplx s = tm.coupling * (st.pk[b] - st.sn[b]);
        out.shot[dv] += s + tm.f_mirror * std::conj(s);
      }
    }
    LoadTwopt<T> lt{dnLMp, N00.get(), ctx.tab, dmk(Sbar.real(), Sbar.imag()),
                    1. / g.vol};
    fft3d<T>(ctx, tfield.get(), lt, +1);
    std::vector<cplx> sij = shot_pairs_bispec<T>(ctx, tfield.get(), p.ell1, tm.m1,
                                                 p.ell2, tm.m2, sja, sjb,
                                                 out.e1, out.e2);
    for (int dv = 0; dv < dim; dv++) {
      cplx s = tm.coupling * sij[dv];
      out.shot[dv] += factor_phase * (s + tm.f_mirror_w3j * std::conj(s));
    }
  }
  for (int dv = 0; dv < dim; dv++) {
    out.raw[dv] *= norm_factor;
    out.shot[dv] *= norm_factor;
  }
  return out;
}

// ---------------------------------------------------------------------------
// Three-point correlation function
// ---------------------------------------------------------------------------

template <typename T>
Result run_3pcf(Context& ctx, const DeviceCatalogue& d,
                const DeviceCatalogue* r, bool survey,
                const trvg_binning& rb, double norm_factor) {
  const trvg_params& p = ctx.p;
  const Geom& g = ctx.g;
  const long long nm = g.nmesh;
  const double alpha = survey ? d.wstotal / r->wstotal : 1.;
  const cplx factor_phase = ipow(p.ell1 + p.ell2);
  const double factor_parity = ((p.ell1 + p.ell2) % 2) ? -1. : 1.;
  const int nb = rb.num_bins;
  std::vector<int> ra, cb;
  pair_bins(p.shape, nb, p.idx_bin, ra, cb);
  const int dim = int(ra.size());

  Result out;
  out.dim = dim;
  out.b1.resize(dim); out.b2.resize(dim); out.e1.resize(dim); out.e2.resize(dim);
  out.n1.resize(dim); out.n2.resize(dim);
  out.raw.assign(dim, 0.); out.shot.assign(dim, 0.);

  DevArray<c2<T>> dn00(nm), N00(nm);
  build_field_k<T>(ctx, dn00.get(), d, r, survey, alpha, 0, 0, false, true);
  build_field_k<T>(ctx, N00.get(), d, r, survey, alpha, 0, 0, true, false);

  Spline sja, sjb;
  build_spline(p.ell1, sja, ctx.stream);
  build_spline(p.ell2, sjb, ctx.stream);

  std::vector<int> rows = distinct(ra), cols = distinct(cb);
  DevArray<c2<T>> G(nm), tfield(nm), dnLM;
  std::vector<double> reff(nb, 0.);
  bool first = true;

  for (const Term& tm : enumerate_terms(p, survey)) {
    const bool lm_zero = !survey || (p.ELL == 0 && tm.M == 0);
    const c2<T>* dnLMp = dn00.get();
    if (!lm_zero) {
      if (!dnLM.get()) dnLM.reset(nm);
      build_field_k<T>(ctx, dnLM.get(), d, r, true, alpha, p.ELL, tm.M, false, false);
      dnLMp = dnLM.get();
    }
    cplx Sbar;
    if (survey) {
      Sbar = sbar_catalogue(ctx, d, p.ELL, tm.M)
             + std::pow(alpha, 3) * sbar_catalogue(ctx, *r, p.ELL, tm.M);
    } else {
      Sbar = double(d.n);
    }

    // Shot noise first: defines r_eff and n_pairs (threept.cpp:1223-1367).
    LoadTwopt<T> lt{dnLMp, N00.get(), ctx.tab, dmk(Sbar.real(), Sbar.imag()),
                    1. / g.vol};
    fft3d<T>(ctx, tfield.get(), lt, +1);
    ShotBins3pcf sb = shot_bins_3pcf<T>(ctx, tfield.get(), p.ell1, tm.m1, p.ell2,
                                        tm.m2, rb);
    for (int dv = 0; dv < dim; dv++) {
      if (ra[dv] != cb[dv]) continue;
      if (p.shape == TRVG_SHAPE_OFFDIAG && p.idx_bin != 0) continue;
      int b = ra[dv];
      out.shot[dv] += tm.coupling * factor_parity
                      * (sb.xi[b] + tm.f_mirror_w3j * std::conj(sb.xi[b]));
    }
    if (first) {
      for (int dv = 0; dv < dim; dv++) {
        out.b1[dv] = rb.bin_centres[ra[dv]];
        out.b2[dv] = rb.bin_centres[cb[dv]];
        out.e1[dv] = sb.r[ra[dv]];
        out.e2[dv] = sb.r[cb[dv]];
        out.n1[dv] = sb.npairs[ra[dv]];
        out.n2[dv] = sb.npairs[cb[dv]];
      }
      reff = sb.r;
      first = false;
    }

    LoadDeconv<T> ldg{dnLMp, ctx.tab, 1. / g.vol};
    fft3d<T>(ctx, G.get(), ldg, +1);

    const bool same_side = (p.ell1 == p.ell2 && tm.m1 == tm.m2);
    DevArray<c2<T>> FA(nm * rows.size());
    std::vector<int> ia(nb, -1), ib(nb, -1);
    for (size_t q = 0; q < rows.size(); q++) {
      int b = rows[q];
      LoadSjl<T> ls{dn00.get(), ctx.tab, g, sja.dev(), reff[b], p.ell1, tm.m1};
      fft3d<T>(ctx, FA.get() + nm * q, ls, +1);
      ia[b] = int(q);
    }
    DevArray<c2<T>> FBown;
    const c2<T>* FBp = FA.get();
    int nfb = int(rows.size());
    bool cols_in_rows = std::all_of(cols.begin(), cols.end(),
                                    [&](int b) { return ia[b] >= 0; });
    if (same_side && cols_in_rows) {
      for (int b : cols) ib[b] = ia[b];
    } else {
      FBown.reset(nm * cols.size());
      for (size_t q = 0; q < cols.size(); q++) {
        int b = cols[q];
        LoadSjl<T> ls{dn00.get(), ctx.tab, g, sjb.dev(), reff[b], p.ell2, tm.m2};
        fft3d<T>(ctx, FBown.get() + nm * q, ls, +1);
        ib[b] = int(q);
      }
      FBp = FBown.get();
      nfb = int(cols.size());
    }
    std::vector<int2> pairs(dim);
    for (int dv = 0; dv < dim; dv++) pairs[dv] = make_int2(ia[ra[dv]], ib[cb[dv]]);
    std::vector<cplx> S = pair_reduce<T>(ctx, FA.get(), int(rows.size()), FBp, nfb,
                                         G.get(), pairs);
    for (int dv = 0; dv < dim; dv++) {
      cplx s = S[dv];
      out.raw[dv] += tm.coupling * g.vol_cell * factor_phase
                     * (s + tm.f_mirror * std::conj(s));
    }
  }
  for (int dv = 0; dv < dim; dv++) {
    out.raw[dv] *= norm_factor;
    out.shot[dv] *= norm_factor;
  }
  return out;
}

// calc_bispec_normalisation_from_mesh (threept.cpp:138-149 ->
// field.cpp:2017-2065) before the 1/alpha^3 factor.
template <typename T>
double run_mesh_norm(Context& ctx, const DeviceCatalogue& c) {
  DevArray<c2<T>> f(ctx.g.nmesh);
  zero_field<T>(ctx, f.get());
  assign<T>(ctx, f.get(), c, WM_W, 0, 0, 1.);
  unsigned grid = grid_for(ctx, ctx.g.nmesh, 4);
  DevArray<double> part(grid), sum(1);
  hipLaunchKernelGGL((k_cube_sum<T>), dim3(grid), dim3(RED_THREADS), 0,
                     ctx.stream, f.get(), ctx.g.nmesh, part.get());
  TRVG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, ctx.stream,
                     part.get(), int(grid), 1, sum.get());
  TRVG_LAUNCH_CHECK();
  double h = 0.;
  TRVG_HIP(hipMemcpyAsync(&h, sum.get(), 8, hipMemcpyDeviceToHost, ctx.stream));
  TRVG_HIP(hipStreamSynchronize(ctx.stream));
  return 1. / (h * ctx.g.vol_cell);
}

template Result run_bispec<double>(Context&, const DeviceCatalogue&,
                                   const DeviceCatalogue*, bool,
                                   const trvg_binning&, double);
template Result run_bispec<float>(Context&, const DeviceCatalogue&,
                                  const DeviceCatalogue*, bool,
                                  const trvg_binning&, double);
template Result run_3pcf<double>(Context&, const DeviceCatalogue&,
                                 const DeviceCatalogue*, bool,
                                 const trvg_binning&, double);
template Result run_3pcf<float>(Context&, const DeviceCatalogue&,
                                const DeviceCatalogue*, bool,
                                const trvg_binning&, double);
template double run_mesh_norm<double>(Context&, const DeviceCatalogue&);
template double run_mesh_norm<float>(Context&, const DeviceCatalogue&);
template void fft3d<double, LoadSame<double>>(Context&, double2*,
                                              const LoadSame<double>&, int);
template void fft3d<float, LoadSame<float>>(Context&, float2*,
                                            const LoadSame<float>&, int);

}  // namespace trvg
