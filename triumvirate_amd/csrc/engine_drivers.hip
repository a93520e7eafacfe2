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
#include <set>
#include <tuple>
#include <type_traits>

#include "batch_kernels.hpp"
#include "diff_kernels.hpp"
#include "engine.hpp"

namespace trvg {

// ---------------------------------------------------------------------------
// FFT
// ---------------------------------------------------------------------------

template <typename T>
static const c2<T>* twiddles(int n, int sign);

// Twiddles of a transform in context ctx: borrowed tables where the context
// has them, else the engine's cache.
template <typename T>
static const c2<T>* ctx_twiddles(const Context& ctx, int n, int sign) {
  int lg = 0;
  while ((1 << lg) < n) lg++;
  const void* b = (lg < 11 && (1 << lg) == n) ? ctx.tw[sign > 0 ? 1 : 0][lg] : nullptr;
  return b ? static_cast<const c2<T>*>(b) : twiddles<T>(n, sign);
}

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


template <typename T, int LOGN>
static size_t fft_smem(bool row_flags) {
  constexpr int N = 1 << LOGN;
  constexpr int NL = FFT_THREADS * FftEpt<T>::v / N;
  // Lines, twiddles, then (strided passes) one byte per row of pruning flags.
  return size_t(NL * (N + 1) + N) * sizeof(c2<T>) + (row_flags ? N : 0);
}

// The register-resident z pass for FP32 meshes of 256 and 512 with loaders
// that only read (the masked rows of shell transforms, plain transforms):
// measured faster there.  FP64 passes and loaders with per-cell arithmetic
// keep the LDS-staged pass, whose lower register use keeps more waves
// resident.
template <class Loader> struct ZRegLoader : std::false_type {};
template <> struct ZRegLoader<LoadMaskedZBatch<float>> : std::true_type {};
template <> struct ZRegLoader<LoadSame<float>> : std::true_type {};

template <typename T, int LOGN, class Loader>
constexpr bool use_zreg() {
  return ZRegShape<LOGN>::ok && sizeof(T) == 4 && ZRegLoader<Loader>::value;
}

template <typename T, int LOGN, class Loader>
static void launch_z(Context& ctx, c2<T>* data, const Loader& load, int sign,
                     double read_bytes, double write_bytes, const Prune& pr) {
  constexpr int EPT = FftEpt<T>::v;
  constexpr int N = 1 << LOGN;
  constexpr int E = FFT_THREADS * EPT;
  const long long nblocks = ctx.g.nmesh / E;
  const size_t smem = fft_smem<T, LOGN>(false);
  const c2<T>* tw = ctx_twiddles<T>(ctx, N, sign);
  ProfScope ps(ctx, sizeof(T) == 8 ? KC_FFT_Z64 : KC_FFT_Z, read_bytes + write_bytes);
  if constexpr (use_zreg<T, LOGN, Loader>()) {
    static_assert(E == zreg_block<T>(), "z pass block size");
    hipLaunchKernelGGL((fft_pass_z_reg<T, LOGN, Loader>), dim3(unsigned(nblocks)),
                       dim3(zreg_threads<T, LOGN>()), (zreg_smem<T, LOGN>()), ctx.stream, data,
                       load, tw, ctx.g.n[1], sign, pr);
  } else if (ctx.g.n[1] % (E / N) == 0)
    hipLaunchKernelGGL((fft_pass_z<T, LOGN, EPT, true, Loader>), dim3(unsigned(nblocks)),
                       dim3(FFT_THREADS), smem, ctx.stream, data, load, tw,
                       ctx.g.n[1], sign, pr);
  else
    hipLaunchKernelGGL((fft_pass_z<T, LOGN, EPT, false, Loader>), dim3(unsigned(nblocks)),
                       dim3(FFT_THREADS), smem, ctx.stream, data, load, tw,
                       ctx.g.n[1], sign, pr);
  TRVG_LAUNCH_CHECK();
}

template <typename T, int LOGN>
static void launch_strided(Context& ctx, c2<T>* data, long long S,
                           long long outer_stride, long long inner, int sign,
                           const Prune& pr, int prune_mode, double bytes) {
  constexpr int EPT = FftEpt<T>::v;
  constexpr int N = 1 << LOGN;
  constexpr int E = FFT_THREADS * EPT;
  const long long nblocks = ctx.g.nmesh / E;
  const size_t smem = fft_smem<T, LOGN>(true);
  const c2<T>* tw = ctx_twiddles<T>(ctx, N, sign);
  ProfScope ps(ctx, sizeof(T) == 8 ? KC_FFT_STRIDED64 : KC_FFT_STRIDED, bytes);
  if (inner % (E / N) == 0)
    hipLaunchKernelGGL((fft_pass_strided<T, LOGN, EPT, true>), dim3(unsigned(nblocks)),
                       dim3(FFT_THREADS), smem, ctx.stream, data, tw, S, outer_stride,
                       inner, sign, pr, prune_mode);
  else
    hipLaunchKernelGGL((fft_pass_strided<T, LOGN, EPT, false>), dim3(unsigned(nblocks)),
                       dim3(FFT_THREADS), smem, ctx.stream, data, tw, S, outer_stride,
                       inner, sign, pr, prune_mode);
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

// Axes of power-of-two length up to 1024 take the compile-time passes of
// fft.hpp; other (even) lengths the mixed-radix passes of fft_gen.hpp.  Both
// give a workgroup lines_per_block(n) lines (2048 / n for the former).
static bool pow2_axis(int n) { return n >= 8 && n <= 1024 && (n & (n - 1)) == 0; }
// The compile-time passes take whole blocks of 2048 elements: a power-of-two
// axis of a mesh whose cell count is not a multiple of 2048 (other axes not
// powers of two) goes through the mixed-radix passes too.
static bool fast_axis(const Geom& g, int n) { return pow2_axis(n) && g.nmesh % 2048 == 0; }
static int lines_per_block(int n) { return gen_lines_per_block(n); }
static long long blocks_of(long long nlines, int nl) { return (nlines + nl - 1) / nl; }

// Dynamic LDS above the default 64 KiB needs the kernel's attribute raised.
template <class K>
static void gen_lds(K kernel, size_t bytes) {
  if (bytes > 64 * 1024)
    TRVG_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
}

template <typename T, class Loader>
static void launch_z_gen(Context& ctx, c2<T>* data, const Loader& load, int sign,
                         double bytes, const Prune& pr) {
  const int N = ctx.g.n[2];
  const GenPlan P = make_gen_plan(N);
  const long long nlines = (long long)ctx.g.n[0] * ctx.g.n[1];
  const size_t smem = gen_smem<T>(N, false);
  auto kern = fft_pass_z_gen<T, Loader>;
  gen_lds(kern, smem);
  ProfScope ps(ctx, sizeof(T) == 8 ? KC_FFT_Z64 : KC_FFT_Z, bytes);
  hipLaunchKernelGGL(kern, dim3(unsigned(blocks_of(nlines, P.nl))), dim3(FFT_THREADS), smem,
                     ctx.stream, data, load, ctx_twiddles<T>(ctx, N, sign), ctx.g.n[1],
                     nlines, sign, pr, P);
  TRVG_LAUNCH_CHECK();
}

template <typename T>
static void launch_strided_gen(Context& ctx, c2<T>* data, int N, long long S,
                               long long outer_stride, long long inner, int sign,
                               const Prune& pr, int prune_mode, double bytes) {
  const GenPlan P = make_gen_plan(N);
  const long long nlines = ctx.g.nmesh / N;
  const size_t smem = gen_smem<T>(N, true);
  auto kern = fft_pass_strided_gen<T>;
  gen_lds(kern, smem);
  ProfScope ps(ctx, sizeof(T) == 8 ? KC_FFT_STRIDED64 : KC_FFT_STRIDED, bytes);
  hipLaunchKernelGGL(kern, dim3(unsigned(blocks_of(nlines, P.nl))), dim3(FFT_THREADS), smem,
                     ctx.stream, data, ctx_twiddles<T>(ctx, N, sign), S, outer_stride, inner,
                     nlines, pr, prune_mode, P);
  TRVG_LAUNCH_CHECK();
}

template <typename T, class BatchLoader>
static void launch_z_batch_gen(Context& ctx, c2<T>* base, const BatchLoader& load, int sign,
                               const Prune& pr, const double* kcut2, const int2* work,
                               int nwork, double bytes) {
  const int N = ctx.g.n[2];
  const GenPlan P = make_gen_plan(N);
  const long long nlines = (long long)ctx.g.n[0] * ctx.g.n[1];
  const size_t smem = gen_smem<T>(N, false);
  auto kern = fft_pass_z_batch_gen<T, BatchLoader>;
  gen_lds(kern, smem);
  ProfScope ps(ctx, sizeof(T) == 8 ? KC_FFT_Z64 : KC_FFT_Z, bytes);
  hipLaunchKernelGGL(kern, dim3(unsigned(nwork)), dim3(FFT_THREADS), smem, ctx.stream, base,
                     ctx.g.nmesh, load, ctx_twiddles<T>(ctx, N, sign), ctx.g.n[1], nlines,
                     sign, pr, kcut2, work, P);
  TRVG_LAUNCH_CHECK();
}

template <typename T>
static void launch_strided_batch_gen(Context& ctx, c2<T>* base, int N, long long S,
                                     long long outer_stride, long long inner, int sign,
                                     const Prune& pr, const double* kcut2, const int2* work,
                                     int nwork, int prune_mode, double bytes) {
  const GenPlan P = make_gen_plan(N);
  const long long nlines = ctx.g.nmesh / N;
  const size_t smem = gen_smem<T>(N, true);
  auto kern = fft_pass_strided_batch_gen<T>;
  gen_lds(kern, smem);
  ProfScope ps(ctx, sizeof(T) == 8 ? KC_FFT_STRIDED64 : KC_FFT_STRIDED, bytes);
  hipLaunchKernelGGL(kern, dim3(unsigned(nwork)), dim3(FFT_THREADS), smem, ctx.stream, base,
                     ctx.g.nmesh, ctx_twiddles<T>(ctx, N, sign), S, outer_stride, inner,
                     nlines, pr, kcut2, work, prune_mode, P);
  TRVG_LAUNCH_CHECK();
}

Prune make_prune(const Geom& g, double k_hi) {
  Prune pr{};
  pr.on = 0;
  const double kcut2 = k_hi * k_hi * (1. + 1.e-9);
  // Worth it only if the band leaves part of the (kx, ky) plane empty.
  const double kx_max = (g.n[0] / 2) * g.dk[0];
  const double ky_max = (g.n[1] / 2) * g.dk[1];
  if (kx_max * kx_max + ky_max * ky_max < kcut2) return pr;
  pr.on = 1;
  pr.kcut2 = kcut2;
  pr.dk0 = g.dk[0];
  pr.dk1 = g.dk[1];
  pr.n0 = g.n[0];
  pr.n1 = g.n[1];
  return pr;
}

// ---------------------------------------------------------------------------
// Batched shell transforms: one launch per pass for a whole slab of bins,
// over only the blocks whose lines lie in each bin's band.
// ---------------------------------------------------------------------------

// Blocks of the z and y passes holding a line inside a pruning radius, for
// one geometry (computed once per radius).  The host predicate carries a
// relative margin over the kernels' exactly rounded one, so every block the
// kernels would use is listed (a listed block with no needed line only
// writes zeros that no later pass reads).
struct PruneLists {
  std::vector<int> zblocks, yblocks;
  double d_lines = 0., slabs = 0.;   // needed z lines and y slabs
};

static const PruneLists& prune_lists(const Geom& g, double kcut2, int nl_z,
                                     int nl_y) {
  static std::mutex mu;
  static std::map<std::vector<double>, std::unique_ptr<PruneLists>> cache;
  std::vector<double> key = {double(g.n[0]), double(g.n[1]), double(g.n[2]),
                             g.dk[0], g.dk[1], kcut2, double(nl_z), double(nl_y)};
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return *it->second;
  auto pl = std::make_unique<PruneLists>();
  const double kc = kcut2 * (1. + 1.e-12);
  std::vector<char> slab_in(g.n[0]), line_in((size_t)g.n[0] * g.n[1]);
  for (int i = 0; i < g.n[0]; i++) {
    const double kx = shifted(i, g.n[0]) * g.dk[0];
    slab_in[i] = kx * kx < kc;
    pl->slabs += (kx * kx < kcut2);
    for (int j = 0; j < g.n[1]; j++) {
      const double ky = shifted(j, g.n[1]) * g.dk[1];
      line_in[(size_t)i * g.n[1] + j] = kx * kx + ky * ky < kc;
      pl->d_lines += (kx * kx + ky * ky < kcut2);
    }
  }
  const long long nlines = (long long)g.n[0] * g.n[1];
  for (long long b = 0; b * nl_z < nlines; b++) {
    bool any = false;
    for (int l = 0; l < nl_z && !any; l++) any = line_in[b * nl_z + l];
    if (any) pl->zblocks.push_back(int(b));
  }
  const long long ylines = (long long)g.n[0] * g.n[2];   // (i, k) lines
  for (long long b = 0; b * nl_y < ylines; b++) {
    bool any = false;
    for (int l = 0; l < nl_y && !any; l++) any = slab_in[(b * nl_y + l) / g.n[2]];
    if (any) pl->yblocks.push_back(int(b));
  }
  return *(cache[key] = std::move(pl));
}

template <typename T, int LOGN, class BatchLoader>
static void launch_z_batch(Context& ctx, c2<T>* base, const BatchLoader& load,
                           int sign, const Prune& pr, const double* kcut2,
                           const int2* work, int nwork, double bytes) {
  constexpr int EPT = FftEpt<T>::v;
  constexpr int N = 1 << LOGN;
  constexpr int E = FFT_THREADS * EPT;
  const size_t smem = fft_smem<T, LOGN>(false);
  const c2<T>* tw = ctx_twiddles<T>(ctx, N, sign);
  ProfScope ps(ctx, sizeof(T) == 8 ? KC_FFT_Z64 : KC_FFT_Z, bytes);
  bool reg = false;
  if constexpr (use_zreg<T, LOGN, BatchLoader>()) {
    static_assert(E == zreg_block<T>(), "z pass block size");
    reg = !pr.on;   // the register pass takes every listed block
    if (reg)
      hipLaunchKernelGGL((fft_pass_z_batch_reg<T, LOGN, BatchLoader>),
                         dim3(unsigned((nwork + 1) / 2)), dim3(zreg_threads<T, LOGN>()),
                         (zreg_smem<T, LOGN>()), ctx.stream, base, ctx.g.nmesh, load, tw,
                         ctx.g.n[1], sign, work, nwork);
  }
  if (!reg) {
    if (ctx.g.n[1] % (E / N) == 0)
      hipLaunchKernelGGL((fft_pass_z_batch<T, LOGN, EPT, true, BatchLoader>),
                         dim3(unsigned(nwork)), dim3(FFT_THREADS), smem, ctx.stream,
                         base, ctx.g.nmesh, load, tw, ctx.g.n[1], sign, pr, kcut2, work);
    else
      hipLaunchKernelGGL((fft_pass_z_batch<T, LOGN, EPT, false, BatchLoader>),
                         dim3(unsigned(nwork)), dim3(FFT_THREADS), smem, ctx.stream,
                         base, ctx.g.nmesh, load, tw, ctx.g.n[1], sign, pr, kcut2, work);
  }
  TRVG_LAUNCH_CHECK();
}

template <typename T, int LOGN>
static void launch_strided_batch(Context& ctx, c2<T>* base, long long S,
                                 long long outer_stride, long long inner,
                                 int sign, const Prune& pr, const double* kcut2,
                                 const int2* work, int nwork, int prune_mode,
                                 double bytes) {
  constexpr int EPT = FftEpt<T>::v;
  constexpr int N = 1 << LOGN;
  constexpr int E = FFT_THREADS * EPT;
  const size_t smem = fft_smem<T, LOGN>(true);
  const c2<T>* tw = ctx_twiddles<T>(ctx, N, sign);
  ProfScope ps(ctx, sizeof(T) == 8 ? KC_FFT_STRIDED64 : KC_FFT_STRIDED, bytes);
  if (inner % (E / N) == 0)
    hipLaunchKernelGGL((fft_pass_strided_batch<T, LOGN, EPT, true>), dim3(unsigned(nwork)),
                       dim3(FFT_THREADS), smem, ctx.stream, base, ctx.g.nmesh, tw, S,
                       outer_stride, inner, sign, pr, kcut2, work, prune_mode);
  else
    hipLaunchKernelGGL((fft_pass_strided_batch<T, LOGN, EPT, false>), dim3(unsigned(nwork)),
                       dim3(FFT_THREADS), smem, ctx.stream, base, ctx.g.nmesh, tw, S,
                       outer_stride, inner, sign, pr, kcut2, work, prune_mode);
  TRVG_LAUNCH_CHECK();
}

// Transform of several fields in one launch per pass: field f of the work
// lists lives at base + f * nmesh and loads through load.field(f); with
// `prune`, kcut2[f] is field f's pruning radius and the lists hold only the
// blocks that radius needs (z and y passes).
template <typename T, class BatchLoader>
void fft3d_batch(Context& ctx, c2<T>* base, const BatchLoader& load, int sign,
                 bool prune, const double* kcut2, const int2* wz, int nz,
                 const int2* wy, int ny, const int2* wx, int nx) {
  const Geom& g = ctx.g;
  Prune pr{};
  pr.on = prune ? 1 : 0;
  pr.dk0 = g.dk[0];
  pr.dk1 = g.dk[1];
  pr.n0 = g.n[0];
  pr.n1 = g.n[1];
  const double blk = double(FFT_THREADS * FftEpt<T>::v) * sizeof(c2<T>);
  if (nz > 0) {
    if (fast_axis(g, g.n[2])) {
      TRVG_LOGN_SWITCH(ilog2(g.n[2]), (launch_z_batch<T, LG>(
          ctx, base, load, sign, pr, kcut2, wz, nz, 2. * blk * nz)));
    } else {
      launch_z_batch_gen<T>(ctx, base, load, sign, pr, kcut2, wz, nz, 2. * blk * nz);
    }
  }
  if (ny > 0) {
    if (fast_axis(g, g.n[1])) {
      TRVG_LOGN_SWITCH(ilog2(g.n[1]), (launch_strided_batch<T, LG>(
          ctx, base, g.n[2], (long long)g.n[1] * g.n[2], g.n[2], sign, pr, kcut2, wy,
          ny, 1, 2. * blk * ny)));
    } else {
      launch_strided_batch_gen<T>(ctx, base, g.n[1], g.n[2], (long long)g.n[1] * g.n[2],
                                  g.n[2], sign, pr, kcut2, wy, ny, 1, 2. * blk * ny);
    }
  }
  if (nx > 0) {
    if (fast_axis(g, g.n[0])) {
      TRVG_LOGN_SWITCH(ilog2(g.n[0]), (launch_strided_batch<T, LG>(
          ctx, base, (long long)g.n[1] * g.n[2], 0, (long long)g.n[1] * g.n[2], sign,
          pr, kcut2, wx, nx, 2, 2. * blk * nx)));
    } else {
      launch_strided_batch_gen<T>(ctx, base, g.n[0], (long long)g.n[1] * g.n[2], 0,
                                  (long long)g.n[1] * g.n[2], sign, pr, kcut2, wx, nx, 2,
                                  2. * blk * nx);
    }
  }
}

// Work lists of a shell slab transform over `ngroups` groups of all
// `nbins` bins: per group nt transforms (nbins, or (nbins + 1) / 2 when
// packed, two real bins per transform), transform f = group * nt + t.
// Returns the bins' bands (two per transform when packed, padded with an
// empty band), each transform's pruning radius (+inf when the band covers
// the whole plane), and the z / y / x blocks.
template <typename T>
void shell_slab_work(const Geom& g, const trvg_binning& kb, int nbins,
                     int ngroups, bool packed, std::vector<ShellBand>& bands,
                     std::vector<double>& kcut2, std::vector<int2>& wz,
                     std::vector<int2>& wy, std::vector<int2>& wx) {
  const int nl_z = lines_per_block(g.n[2]), nl_y = lines_per_block(g.n[1]);
  const long long nblk = blocks_of((long long)g.n[1] * g.n[2], lines_per_block(g.n[0]));
  const int per = packed ? 2 : 1;
  const int nt = (nbins + per - 1) / per;
  bands.assign((size_t)nt * per, ShellBand{HUGE_VAL, HUGE_VAL});
  std::vector<double> kc(nt, 0.);
  for (int q = 0; q < nbins; q++) {
    bands[q] = shell_band(kb.bin_edges[q], kb.bin_edges[q + 1]);
    Prune pq = make_prune(g, kb.bin_edges[q + 1]);
    kc[q / per] = std::max(kc[q / per], pq.on ? pq.kcut2 : HUGE_VAL);
  }
  std::vector<const PruneLists*> pls(nt);
  for (int t = 0; t < nt; t++) pls[t] = &prune_lists(g, kc[t], nl_z, nl_y);
  kcut2.clear(); wz.clear(); wy.clear(); wx.clear();
  for (int gi = 0; gi < ngroups; gi++)
    for (int t = 0; t < nt; t++) {
      const int f = gi * nt + t;
      kcut2.push_back(kc[t]);
      for (int w : pls[t]->zblocks) wz.push_back(make_int2(f, w));
      for (int w : pls[t]->yblocks) wy.push_back(make_int2(f, w));
      for (long long w = 0; w < nblk; w++) wx.push_back(make_int2(f, int(w)));
    }
}

template void shell_slab_work<double>(const Geom&, const trvg_binning&, int, int,
                                      bool, std::vector<ShellBand>&,
                                      std::vector<double>&, std::vector<int2>&,
                                      std::vector<int2>&, std::vector<int2>&);
template void shell_slab_work<float>(const Geom&, const trvg_binning&, int, int,
                                     bool, std::vector<ShellBand>&,
                                     std::vector<double>&, std::vector<int2>&,
                                     std::vector<int2>&, std::vector<int2>&);
#define TRVG_FFT3D_BATCH_INST(T, L)                                             \
  template void fft3d_batch<T, L>(Context&, c2<T>*, const L&, int, bool,        \
                                  const double*, const int2*, int, const int2*, \
                                  int, const int2*, int);
TRVG_FFT3D_BATCH_INST(double, LoadScaleShiftMocks)
TRVG_FFT3D_BATCH_INST(double, LoadTwoptMocks)
TRVG_FFT3D_BATCH_INST(double, LoadShellMocks<double>)
TRVG_FFT3D_BATCH_INST(float, LoadShellMocks<float>)
TRVG_FFT3D_BATCH_INST(double, LoadShellPairMocks<double>)
TRVG_FFT3D_BATCH_INST(float, LoadShellPairMocks<float>)
TRVG_FFT3D_BATCH_INST(double, LoadDeconvMocks<double>)
TRVG_FFT3D_BATCH_INST(float, LoadDeconvMocks<float>)
#undef TRVG_FFT3D_BATCH_INST

// Shell fields of a real source field are real for m = 0 and even l (the
// spectrum y_l0(k^) f(k) / W(k) is Hermitian).
static bool real_shells(int ell, int m) { return m == 0 && ell % 2 == 0; }

// ---------------------------------------------------------------------------
// Shell transforms, x-first: the x pass carries the shell loader (lines
// along x at fixed (k_y, k_z), pruned on k_y^2 + k_z^2), the y pass is
// pruned on k_z, and the last pass runs along the contiguous z lines, so
// the full-size output of every field is written in whole 16 KiB blocks
// rather than as short strided segments.
// ---------------------------------------------------------------------------

// Prune over the (z, y) axes: P2(k) tests k_z, P1(k, j) tests (k_z, k_y).
static Prune prune_zy(const Geom& g) {
  Prune pr{};
  pr.on = 1;
  pr.dk0 = g.dk[2];
  pr.dk1 = g.dk[1];
  pr.n0 = g.n[2];
  pr.n1 = g.n[1];
  return pr;
}

struct PruneListsXF {
  std::vector<int> xblocks, yblocks;
  double x_lines = 0., y_lines = 0.;   // needed x lines, written y lines
};

static const PruneListsXF& prune_lists_xfirst(const Geom& g, double kcut2, int nl_x,
                                              int nl_y) {
  static std::mutex mu;
  static std::map<std::vector<double>, std::unique_ptr<PruneListsXF>> cache;
  std::vector<double> key = {double(g.n[0]), double(g.n[1]), double(g.n[2]),
                             g.dk[1], g.dk[2], kcut2, double(nl_x), double(nl_y)};
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return *it->second;
  auto pl = std::make_unique<PruneListsXF>();
  const double kc = kcut2 * (1. + 1.e-12);
  const int n1 = g.n[1], n2 = g.n[2];
  std::vector<char> kz_in(n2), line_in((size_t)n1 * n2);
  double nkz = 0.;
  for (int k = 0; k < n2; k++) {
    const double kz = shifted(k, n2) * g.dk[2];
    kz_in[k] = kz * kz < kc;
    nkz += (kz * kz < kcut2);
    for (int j = 0; j < n1; j++) {
      const double ky = shifted(j, n1) * g.dk[1];
      line_in[(size_t)j * n2 + k] = kz * kz + ky * ky < kc;
      pl->x_lines += (kz * kz + ky * ky < kcut2);
    }
  }
  pl->y_lines = nkz * g.n[0];
  const long long xl = (long long)n1 * n2;
  for (long long b = 0; b * nl_x < xl; b++) {
    bool any = false;
    for (int l = 0; l < nl_x && !any; l++) any = line_in[b * nl_x + l];
    if (any) pl->xblocks.push_back(int(b));
  }
  const long long yl = (long long)g.n[0] * n2;
  for (long long b = 0; b * nl_y < yl; b++) {
    bool any = false;
    for (int l = 0; l < nl_y && !any; l++) any = kz_in[(b * nl_y + l) % n2];
    if (any) pl->yblocks.push_back(int(b));
  }
  return *(cache[key] = std::move(pl));
}

template <typename T, int LOGN, class BatchLoader>
static void launch_x_load_batch(Context& ctx, c2<T>* base, const BatchLoader& load,
                                const Prune& pr, const double* kcut2, const int2* work,
                                int nwork, double bytes) {
  constexpr int EPT = FftEpt<T>::v;
  constexpr int N = 1 << LOGN;
  constexpr int E = FFT_THREADS * EPT;
  const size_t smem = fft_smem<T, LOGN>(false);
  const c2<T>* tw = ctx_twiddles<T>(ctx, N, +1);
  ProfScope ps(ctx, sizeof(T) == 8 ? KC_FFT_STRIDED64 : KC_FFT_STRIDED, bytes);
  if (ctx.g.n[2] % (E / N) == 0)
    hipLaunchKernelGGL((fft_pass_x_load_batch<T, LOGN, EPT, true, BatchLoader>),
                       dim3(unsigned(nwork)), dim3(FFT_THREADS), smem, ctx.stream, base,
                       ctx.g.nmesh, load, tw, ctx.g.n[1], ctx.g.n[2], +1, pr, kcut2, work);
  else
    hipLaunchKernelGGL((fft_pass_x_load_batch<T, LOGN, EPT, false, BatchLoader>),
                       dim3(unsigned(nwork)), dim3(FFT_THREADS), smem, ctx.stream, base,
                       ctx.g.nmesh, load, tw, ctx.g.n[1], ctx.g.n[2], +1, pr, kcut2, work);
  TRVG_LAUNCH_CHECK();
}

template <typename T, class BatchLoader>
static void launch_x_load_batch_gen(Context& ctx, c2<T>* base, const BatchLoader& load,
                                    const Prune& pr, const double* kcut2, const int2* work,
                                    int nwork, double bytes) {
  const int N = ctx.g.n[0];
  const GenPlan P = make_gen_plan(N);
  const size_t smem = gen_smem<T>(N, false);
  auto kern = fft_pass_x_load_batch_gen<T, BatchLoader>;
  gen_lds(kern, smem);
  ProfScope ps(ctx, sizeof(T) == 8 ? KC_FFT_STRIDED64 : KC_FFT_STRIDED, bytes);
  hipLaunchKernelGGL(kern, dim3(unsigned(nwork)), dim3(FFT_THREADS), smem, ctx.stream, base,
                     ctx.g.nmesh, load, ctx_twiddles<T>(ctx, N, +1), ctx.g.n[1], ctx.g.n[2],
                     pr, kcut2, work, P);
  TRVG_LAUNCH_CHECK();
}

template <typename T>
static void xfirst_yz(Context& ctx, c2<T>* base, const double* kcut2, const int2* wy, int ny,
                      const int2* wz, int nz, double by, double bz);

// The three passes of a batched x-first transform: field f of the work
// lists lives at base + f * nmesh, loads through load.field(f) and is pruned
// with radius kcut2[f] (+inf: no pruning); wz lists every block.
template <typename T, class BatchLoader>
static void xfirst_passes(Context& ctx, c2<T>* base, const BatchLoader& load,
                          const double* kcut2, const int2* wx, int nx, const int2* wy,
                          int ny, const int2* wz, int nz, double bx, double by, double bz) {
  const Geom& g = ctx.g;
  if (nx > 0) {
    if (fast_axis(g, g.n[0])) {
      TRVG_LOGN_SWITCH(ilog2(g.n[0]), (launch_x_load_batch<T, LG>(
          ctx, base, load, prune_zy(g), kcut2, wx, nx, bx)));
    } else {
      launch_x_load_batch_gen<T>(ctx, base, load, prune_zy(g), kcut2, wx, nx, bx);
    }
  }
  xfirst_yz<T>(ctx, base, kcut2, wy, ny, wz, nz, by, bz);
}

// The y and z passes of a batched x-first transform.
template <typename T>
static void xfirst_yz(Context& ctx, c2<T>* base, const double* kcut2, const int2* wy, int ny,
                      const int2* wz, int nz, double by, double bz) {
  const Geom& g = ctx.g;
  const Prune pr = prune_zy(g);
  if (ny > 0) {
    if (fast_axis(g, g.n[1])) {
      TRVG_LOGN_SWITCH(ilog2(g.n[1]), (launch_strided_batch<T, LG>(
          ctx, base, g.n[2], (long long)g.n[1] * g.n[2], g.n[2], +1, pr, kcut2, wy, ny, 3,
          by)));
    } else {
      launch_strided_batch_gen<T>(ctx, base, g.n[1], g.n[2], (long long)g.n[1] * g.n[2],
                                  g.n[2], +1, pr, kcut2, wy, ny, 3, by);
    }
  }
  // z pass over every block, reading only the k_z rows the y pass wrote.
  Prune off{};
  off.on = 0;
  LoadMaskedZBatch<T> lz{base, g.nmesh, kcut2, g.dk[2], g.n[2]};
  if (nz > 0) {
    if (fast_axis(g, g.n[2])) {
      TRVG_LOGN_SWITCH(ilog2(g.n[2]), (launch_z_batch<T, LG>(
          ctx, base, lz, +1, off, kcut2, wz, nz, bz)));
    } else {
      launch_z_batch_gen<T>(ctx, base, lz, +1, off, kcut2, wz, nz, bz);
    }
  }
}

// Shell fields of (ell, m) for `bins`: the backward transform of
// y_lm(k^) src(k) / W(k) on each bin's shell, all bins per pass in one
// launch.  Unpacked: field q at base + q * nmesh.  Packed (real shells:
// real source, even l, m = 0): bins q = 2p, 2p + 1 are the real and
// imaginary parts of base + p * nmesh.
// Launch plan of a shell slab transform (bands, pruning radii, work lists
// and profiler byte counts), a function of the geometry and the bins only:
// built once per device and reused by every later call.
struct ShellPlan {
  DevArray<ShellBand> bands;
  DevArray<double> kc;
  DevArray<int2> w;
  int nx = 0, ny = 0, nz = 0;
  double bx = 0., by = 0., bz = 0.;
};

template <typename T>
static std::shared_ptr<const ShellPlan> shell_plan(Context& ctx, const trvg_binning& kb,
                                                   const std::vector<int>& bins,
                                                   const std::vector<int>& nmodes,
                                                   bool packed) {
  const Geom& g = ctx.g;
  const int nf = int(bins.size());
  std::vector<double> key = {double(ctx.dev), double(sizeof(T)), double(packed)};
  for (int a = 0; a < 3; a++) { key.push_back(g.n[a]); key.push_back(g.dk[a]); }
  for (int b : bins) {
    key.push_back(kb.bin_edges[b]);
    key.push_back(kb.bin_edges[b + 1]);
    key.push_back(nmodes[b]);
  }
  static std::mutex mu;
  static std::map<std::vector<double>, std::shared_ptr<const ShellPlan>> cache;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  const int nt = packed ? (nf + 1) / 2 : nf;
  const int per = packed ? 2 : 1;
  const int nl_x = lines_per_block(g.n[0]), nl_y = lines_per_block(g.n[1]);
  const ShellBand none{HUGE_VAL, HUGE_VAL};
  std::vector<ShellBand> bands((size_t)nt * per, none);
  std::vector<double> kc(nt, 0.);
  // z blocks: every (i, j) line of the last pass.
  const long long nblk = blocks_of((long long)g.n[0] * g.n[1], lines_per_block(g.n[2]));
  std::vector<int2> wx, wy, wz;
  const double cb = sizeof(c2<T>);
  auto pl = std::make_shared<ShellPlan>();
  for (int q = 0; q < nf; q++) {
    const int b = bins[q];
    bands[q] = shell_band(kb.bin_edges[b], kb.bin_edges[b + 1]);
    Prune pq = make_prune(g, kb.bin_edges[b + 1]);
    kc[q / per] = std::max(kc[q / per], pq.on ? pq.kcut2 : HUGE_VAL);
    pl->bx += double(nmodes[b]) * sizeof(double2);
  }
  for (int f = 0; f < nt; f++) {
    const PruneListsXF& px = prune_lists_xfirst(g, kc[f], nl_x, nl_y);
    for (int w : px.xblocks) wx.push_back(make_int2(f, w));
    for (int w : px.yblocks) wy.push_back(make_int2(f, w));
    for (long long w = 0; w < nblk; w++) wz.push_back(make_int2(f, int(w)));
    const double xw = px.x_lines * g.n[0] * cb;
    const double yw = px.y_lines * g.n[1] * cb;
    pl->bx += xw;
    pl->by += xw + yw;
    pl->bz += yw + double(g.nmesh) * cb;
  }
  pl->nx = int(wx.size());
  pl->ny = int(wy.size());
  pl->nz = int(wz.size());
  pl->bands.reset(bands.size());
  pl->kc.reset(nt);
  pl->w.reset(wx.size() + wy.size() + wz.size());
  wx.insert(wx.end(), wy.begin(), wy.end());
  wx.insert(wx.end(), wz.begin(), wz.end());
  TRVG_HIP(hipMemcpyAsync(pl->bands.get(), bands.data(), bands.size() * sizeof(ShellBand),
                          hipMemcpyHostToDevice, ctx.stream));
  TRVG_HIP(hipMemcpyAsync(pl->kc.get(), kc.data(), nt * sizeof(double),
                          hipMemcpyHostToDevice, ctx.stream));
  TRVG_HIP(hipMemcpyAsync(pl->w.get(), wx.data(), wx.size() * sizeof(int2),
                          hipMemcpyHostToDevice, ctx.stream));
  TRVG_HIP(hipStreamSynchronize(ctx.stream));
  std::lock_guard<std::mutex> lk(mu);
  if (cache.size() >= 64) cache.clear();   // bounded; plans in use stay alive
  cache[key] = pl;
  return pl;
}

template <typename T>
void shell_slab_ifft(Context& ctx, c2<T>* base, const double2* src,
                     const trvg_binning& kb, const std::vector<int>& bins,
                     const std::vector<int>& nmodes, int ell, int m, bool packed) {
  const Geom& g = ctx.g;
  if (bins.empty()) return;
  const std::shared_ptr<const ShellPlan> pl = shell_plan<T>(ctx, kb, bins, nmodes, packed);
  // The plan outlives the queued launches even if the cache drops it.
  ctx.keep.push_back(std::const_pointer_cast<ShellPlan>(pl));
  const int2* px = pl->w.get();
  const int2* py = px + pl->nx;
  const int2* pz = py + pl->ny;
  if (packed) {
    LoadShellPairBatch<T, double> ld{src, ctx.tab, g, pl->bands.get(), ell, m};
    xfirst_passes<T>(ctx, base, ld, pl->kc.get(), px, pl->nx, py, pl->ny, pz, pl->nz,
                     pl->bx, pl->by, pl->bz);
  } else {
    LoadShellBatch<T, double> ld{src, ctx.tab, g, pl->bands.get(), ell, m};
    xfirst_passes<T>(ctx, base, ld, pl->kc.get(), px, pl->nx, py, pl->ny, pz, pl->nz,
                     pl->bx, pl->by, pl->bz);
  }
}

template <typename T, int LOGN>
static void launch_x_sjl(Context& ctx, c2<T>* base, int nt, const double2* src,
                         const Spline& sp, const double* r, int ell, int m, bool packed,
                         double bytes) {
  constexpr int EPT = FftEpt<T>::v;
  constexpr int N = 1 << LOGN;
  constexpr int E = FFT_THREADS * EPT;
  const Geom& g = ctx.g;
  const size_t smem = fft_smem<T, LOGN>(false);
  const unsigned grid = unsigned((long long)g.n[1] * g.n[2] / (E / N));
  ProfScope ps(ctx, sizeof(T) == 8 ? KC_FFT_STRIDED64 : KC_FFT_STRIDED, bytes);
  if (packed)
    hipLaunchKernelGGL((fft_x_sjl_multi<T, LOGN, EPT, true>), dim3(grid), dim3(FFT_THREADS),
                       smem, ctx.stream, base, g.nmesh, nt, src, ctx.tab, g, sp.dev(), r, ell,
                       m, ctx_twiddles<T>(ctx, N, +1));
  else
    hipLaunchKernelGGL((fft_x_sjl_multi<T, LOGN, EPT, false>), dim3(grid), dim3(FFT_THREADS),
                       smem, ctx.stream, base, g.nmesh, nt, src, ctx.tab, g, sp.dev(), r, ell,
                       m, ctx_twiddles<T>(ctx, N, +1));
  TRVG_LAUNCH_CHECK();
}

template <typename T>
static void launch_x_sjl_gen(Context& ctx, c2<T>* base, int nt, const double2* src,
                             const Spline& sp, const double* r, int ell, int m, bool packed,
                             double bytes) {
  const Geom& g = ctx.g;
  const int N = g.n[0];
  const GenPlan P = make_gen_plan(N);
  const size_t smem = gen_smem<T>(N, false);
  const unsigned grid = unsigned(blocks_of((long long)g.n[1] * g.n[2], P.nl));
  ProfScope ps(ctx, sizeof(T) == 8 ? KC_FFT_STRIDED64 : KC_FFT_STRIDED, bytes);
  if (packed) {
    auto kern = fft_x_sjl_multi_gen<T, true>;
    gen_lds(kern, smem);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(FFT_THREADS), smem, ctx.stream, base, g.nmesh,
                       nt, src, ctx.tab, g, sp.dev(), r, ell, m,
                       ctx_twiddles<T>(ctx, N, +1), P);
  } else {
    auto kern = fft_x_sjl_multi_gen<T, false>;
    gen_lds(kern, smem);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(FFT_THREADS), smem, ctx.stream, base, g.nmesh,
                       nt, src, ctx.tab, g, sp.dev(), r, ell, m,
                       ctx_twiddles<T>(ctx, N, +1), P);
  }
  TRVG_LAUNCH_CHECK();
}

// j_l-weighted shell fields of (ell, m) for `bins` (3PCF,
// inv_fourier_transform_sjl_ylm_wgtd_field, field.cpp:1908-2010), r the
// bins' effective radii: x-first batched transforms over every block (no
// band limit); layout and packing as shell_slab_ifft.
template <typename T>
void sjl_slab_ifft(Context& ctx, c2<T>* base, const double2* src, const Spline& sp,
                   const std::vector<double>& r, const std::vector<int>& bins, int ell,
                   int m, bool packed) {
  const Geom& g = ctx.g;
  const int nf = int(bins.size());
  if (nf == 0) return;
  const int nt = packed ? (nf + 1) / 2 : nf;
  const int per = packed ? 2 : 1;
  std::vector<double> rr((size_t)nt * per, -1.);
  for (int q = 0; q < nf; q++) rr[q] = r[bins[q]];
  DevArray<double> dr(rr.size());
  ctx.upload(dr.get(), rr.data(), rr.size());
  // Work lists (y blocks, then every z block, per transform) and the
  // unpruned radii depend on the geometry and the transform count only:
  // built once per device.
  std::vector<double> key = {double(ctx.dev), double(sizeof(T)), double(nt)};
  for (int a = 0; a < 3; a++) key.push_back(g.n[a]);
  static std::mutex mu;
  static std::map<std::vector<double>, std::shared_ptr<ShellPlan>> cache;
  std::shared_ptr<ShellPlan> plan;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) plan = it->second;
  }
  if (!plan) {
    const PruneListsXF& pl = prune_lists_xfirst(g, HUGE_VAL, lines_per_block(g.n[0]),
                                                lines_per_block(g.n[1]));
    const long long nblk = blocks_of((long long)g.n[0] * g.n[1], lines_per_block(g.n[2]));
    std::vector<double> kc(nt, HUGE_VAL);
    std::vector<int2> w;   // y blocks, then z blocks
    for (int f = 0; f < nt; f++)
      for (int b : pl.yblocks) w.push_back(make_int2(f, b));
    const size_t ny = w.size();
    for (int f = 0; f < nt; f++)
      for (long long b = 0; b < nblk; b++) w.push_back(make_int2(f, int(b)));
    plan = std::make_shared<ShellPlan>();
    plan->ny = int(ny);
    plan->nz = int(w.size() - ny);
    plan->kc.reset(nt);
    plan->w.reset(w.size());
    TRVG_HIP(hipMemcpyAsync(plan->kc.get(), kc.data(), nt * sizeof(double),
                            hipMemcpyHostToDevice, ctx.stream));
    TRVG_HIP(hipMemcpyAsync(plan->w.get(), w.data(), w.size() * sizeof(int2),
                            hipMemcpyHostToDevice, ctx.stream));
    TRVG_HIP(hipStreamSynchronize(ctx.stream));
    std::lock_guard<std::mutex> lk(mu);
    if (cache.size() >= 16) cache.clear();   // bounded; plans in use stay alive
    cache[key] = plan;
  }
  ctx.keep.push_back(plan);
  const size_t ny = size_t(plan->ny), nz = size_t(plan->nz);
  const int2* dwp = plan->w.get();
  // Algorithmic bytes: the source read once, then each pass writes (and the
  // later ones read) every field.
  const double fb = double(g.nmesh) * sizeof(c2<T>) * nt;
  const double bx = double(g.nmesh) * sizeof(double2) + fb, by = 2. * fb, bz = 2. * fb;
  if (fast_axis(g, g.n[0])) {
    TRVG_LOGN_SWITCH(ilog2(g.n[0]), (launch_x_sjl<T, LG>(ctx, base, nt, src, sp, dr.get(),
                                                          ell, m, packed, bx)));
  } else {
    launch_x_sjl_gen<T>(ctx, base, nt, src, sp, dr.get(), ell, m, packed, bx);
  }
  // The radii stay owned until the stream is synchronised: returned to the
  // shared pool now, another call could take and overwrite them while the
  // x pass still reads them.
  ctx.keep_until_flush(std::move(dr));
  xfirst_yz<T>(ctx, base, plan->kc.get(), dwp, int(ny), dwp + ny, int(nz), by, bz);
}

template <typename T, class Loader>
void fft3d(Context& ctx, c2<T>* data, const Loader& load, int sign,
           double read_bytes, const Prune* prune) {
  const Geom& g = ctx.g;
  Prune pr{};
  pr.on = 0;
  if (prune) pr = *prune;
  const double cb = sizeof(c2<T>);
  // Lines / slabs each pass touches (for the profiler's algorithmic bytes).
  double d_lines = double(g.n[0]) * g.n[1], slabs = g.n[0];
  if (pr.on) {
    d_lines = 0.;
    slabs = 0.;
    for (int i = 0; i < g.n[0]; i++) {
      double kx = shifted(i, g.n[0]) * g.dk[0];
      if (kx * kx < pr.kcut2) slabs += 1.;
      for (int j = 0; j < g.n[1]; j++) {
        double ky = shifted(j, g.n[1]) * g.dk[1];
        if (kx * kx + ky * ky < pr.kcut2) d_lines += 1.;
      }
    }
  }
  if (read_bytes < 0.) read_bytes = double(g.nmesh) * cb;
  const double zw = d_lines * g.n[2] * cb;
  const double slab_vals = slabs * g.n[1] * g.n[2];
  if (fast_axis(g, g.n[2])) {
    TRVG_LOGN_SWITCH(ilog2(g.n[2]), (launch_z<T, LG, Loader>(
        ctx, data, load, sign, read_bytes, zw, pr)));
  } else {
    launch_z_gen<T>(ctx, data, load, sign, read_bytes + zw, pr);
  }
  if (fast_axis(g, g.n[1])) {
    TRVG_LOGN_SWITCH(ilog2(g.n[1]), (launch_strided<T, LG>(
        ctx, data, g.n[2], (long long)g.n[1] * g.n[2], g.n[2], sign, pr, 1,
        zw + slab_vals * cb)));
  } else {
    launch_strided_gen<T>(ctx, data, g.n[1], g.n[2], (long long)g.n[1] * g.n[2], g.n[2],
                          sign, pr, 1, zw + slab_vals * cb);
  }
  if (fast_axis(g, g.n[0])) {
    TRVG_LOGN_SWITCH(ilog2(g.n[0]), (launch_strided<T, LG>(
        ctx, data, (long long)g.n[1] * g.n[2], 0, (long long)g.n[1] * g.n[2], sign,
        pr, 2, slab_vals * cb + double(g.nmesh) * cb)));
  } else {
    launch_strided_gen<T>(ctx, data, g.n[0], (long long)g.n[1] * g.n[2], 0,
                          (long long)g.n[1] * g.n[2], sign, pr, 2,
                          slab_vals * cb + double(g.nmesh) * cb);
  }
}

// ---------------------------------------------------------------------------
// Small launch helpers
// ---------------------------------------------------------------------------

template <typename T>
static void zero_field(Context& ctx, c2<T>* f) {
  ProfScope ps(ctx, KC_MISC, double(ctx.g.nmesh) * sizeof(c2<T>));
  hipLaunchKernelGGL((k_zero<T>), dim3(grid_for(ctx, ctx.g.nmesh)),
                     dim3(RED_THREADS), 0, ctx.stream, f, ctx.g.nmesh);
  TRVG_LAUNCH_CHECK();
}

TileSorter::TileSorter(Context& ctx, const DeviceCatalogue& c)
    : ctx_(ctx), c_(c), tv_(std::make_unique<TiledView>()) {
  const Geom& g = ctx.g;
  for (int a = 0; a < 3; a++) {
    tv_->n[a] = g.n[a];
    tv_->L[a] = g.L[a];
    tv_->tg.nt[a] = (g.n[a] + TILE - 1) / TILE;
  }
  tv_->tg.ntiles = (long long)tv_->tg.nt[0] * tv_->tg.nt[1] * tv_->tg.nt[2];
  const long long nt = tv_->tg.ntiles;
  tile_of_.reset(c.n);
  key2_.reset(c.n);
  ccount_.reset((nt + 1) * TILE_CLASSES);
  TRVG_HIP(hipMemsetAsync(ccount_.get(), 0, (nt + 1) * TILE_CLASSES * 4, ctx.stream));
}

void TileSorter::count(long long begin, long long end) {
  if (end <= begin) return;
  ProfScope ps(ctx_, KC_ASSIGN, double(end - begin) * 32.);
  hipLaunchKernelGGL(k_tile_count, dim3(grid_for(ctx_, end - begin, 16)), dim3(256), 0,
                     ctx_.stream, c_.view(), begin, end, ctx_.g, tv_->tg, tile_of_.get(),
                     ccount_.get(), key2_.get());
  TRVG_LAUNCH_CHECK();
}

void TileSorter::finish(bool move_w, bool move_los) {
  Context& ctx = ctx_;
  const DeviceCatalogue& c = c_;
  move_los = move_los && c.los.get() != nullptr;
  const long long nt = tv_->tg.ntiles;
  DevArray<unsigned int> counts(nt + 1), offs(nt + 2), nch(nt), woff(nt + 1);
  hipLaunchKernelGGL(k_tile_sum, dim3(grid_for(ctx, nt + 1)), dim3(256), 0, ctx.stream,
                     ccount_.get(), nt + 1, counts.get());
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, ctx.stream,
                     counts.get(), nt + 1, offs.get());
  TRVG_LAUNCH_CHECK();
  DevArray<double> sx(c.n), sy(c.n), sz(c.n), sw(move_w ? c.n : 0),
      slos(move_los ? 3 * c.n : 0);
  {
    ProfScope ps(ctx, KC_ASSIGN, c.n * (20. + 48. + (move_w ? 16. : 0.) + (move_los ? 48. : 0.)));
    DevArray<double2> rec(2 * size_t(c.n));
    hipLaunchKernelGGL(k_tile_scatter, dim3(grid_for(ctx, c.n, 16)), dim3(256), 0,
                       ctx.stream, c.view(), tile_of_.get(), offs.get(), ccount_.get(),
                       key2_.get(), rec.get(), move_w, slos.get());
    hipLaunchKernelGGL(k_tile_columns, dim3(grid_for(ctx, c.n, 16)), dim3(256), 0,
                       ctx.stream, rec.get(), c.n, sx.get(), sy.get(), sz.get(), sw.get());
    TRVG_LAUNCH_CHECK();
    ctx.keep_until_flush(std::move(rec));
  }
  // Work list: chunks of at most TILE_CHUNK particles per tile.
  tv_->work_cap = nt + (c.n + TILE_CHUNK - 1) / TILE_CHUNK;
  tv_->work.reset(tv_->work_cap);
  tv_->meta.reset(2);
  const unsigned tgrid = grid_for(ctx, nt);
  hipLaunchKernelGGL(k_tile_chunks, dim3(tgrid), dim3(256), 0, ctx.stream, offs.get(),
                     nt, nch.get());
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, ctx.stream, nch.get(), nt,
                     woff.get());
  hipLaunchKernelGGL(k_tile_work, dim3(tgrid), dim3(256), 0, ctx.stream, offs.get(),
                     woff.get(), nt, tv_->work.get(), tv_->meta.get());
  TRVG_LAUNCH_CHECK();
  // The permuted columns replace the caller-order ones; the old columns and
  // the sort's temporaries stay allocated until the stream has passed them.
  std::swap(c.x, sx); std::swap(c.y, sy); std::swap(c.z, sz);
  if (move_w) std::swap(c.w, sw);
  if (move_los) std::swap(c.los, slos);
  ctx.keep_until_flush(std::move(sx)); ctx.keep_until_flush(std::move(sy));
  ctx.keep_until_flush(std::move(sz)); ctx.keep_until_flush(std::move(sw));
  ctx.keep_until_flush(std::move(slos));
  ctx.keep_until_flush(std::move(tile_of_)); ctx.keep_until_flush(std::move(counts));
  ctx.keep_until_flush(std::move(offs)); ctx.keep_until_flush(std::move(ccount_));
  ctx.keep_until_flush(std::move(key2_)); ctx.keep_until_flush(std::move(nch));
  ctx.keep_until_flush(std::move(woff));
  // Exact grids for later calls once the host has the counts.
  const unsigned int* hm = ctx.download(tv_->meta.get(), 2);
  TiledView* tvp = tv_.get();
  const long long n = c.n;
  ctx.defer([tvp, hm, n] {
    tvp->nwork = hm[0];
    tvp->nstraggle = n - (long long)hm[1];
  });
  c.tiled = std::move(tv_);
}

// Tile order of `c` for the context's mesh (built once per geometry and
// cached on the catalogue): a counting sort by tile with the within-tile
// bank-class order, the columns permuted in place, and the work list built
// on the device.  Nothing here waits for the device.
static const TiledView& tiled_view(Context& ctx, const DeviceCatalogue& c) {
  if (c.tiled && c.tiled->matches(ctx.g)) return *c.tiled;
  TileSorter ts(ctx, c);
  ts.count(0, c.n);
  ts.finish();
  return *c.tiled;
}

// Grid of the tiled kernels: the exact work count once known, else the bound.
static unsigned tile_grid(const TiledView& tv) {
  return unsigned(tv.nwork >= 0 ? tv.nwork : tv.work_cap);
}

template <typename T, int ORDER>
static void assign_order(Context& ctx, c2<T>* mesh, const DeviceCatalogue& c,
                         const TiledView& tv, int mode, int ell, int m, double scale,
                         int cplx_w) {
  DevParticles sv = c.view();
  if (tile_grid(tv) > 0) {
    hipLaunchKernelGGL((k_assign_tiled<T, ORDER>), dim3(tile_grid(tv)), dim3(256), 0,
                       ctx.stream, mesh, sv, ctx.g, tv.tg, tv.work.get(), tv.meta.get(),
                       mode, ell, m, scale, cplx_w);
    TRVG_LAUNCH_CHECK();
  }
  // Stragglers [meta[1], n) keep the reference's flattened-index rule.
  if (tv.nstraggle != 0) {
    hipLaunchKernelGGL((k_assign<T, ORDER>), dim3(tv.nstraggle > 0
                                                      ? grid_for(ctx, tv.nstraggle, 16)
                                                      : 64u),
                       dim3(256), 0, ctx.stream, mesh, sv, ctx.g, mode, ell, m, scale,
                       cplx_w, 0, tv.meta.get() + 1);
    TRVG_LAUNCH_CHECK();
  }
}

// imag: a real weight (m = 0 or a unit/scalar mode) goes into the
// imaginary parts of the mesh instead of the real parts.
template <typename T>
static void assign(Context& ctx, c2<T>* mesh, const DeviceCatalogue& c,
                   int mode, int ell, int m, double scale, bool imag = false) {
  // y_l0 is real (its imaginary part is a signed zero): m = 0 fields take
  // real parts only.
  int cplx_w = ((mode == WM_YLM_W || mode == WM_YLMC_W2) && m != 0);
  if (imag) {
    if (cplx_w) throw Error(TRVG_ERR_INTERNAL, "Complex weights have no imaginary slot.");
    cplx_w = 2;
  }
  if ((mode == WM_YLM_W || mode == WM_YLMC_W2) && c.los.get() == nullptr)
    throw Error(TRVG_ERR_DATA, "Lines of sight are required for survey-type fields.");
  const TiledView& tv = tiled_view(ctx, c);
  // Algorithmic bytes: particle columns read once (position, weight,
  // line of sight when used) plus one write of the mesh.
  double pbytes = 24. + (mode == WM_UNIT ? 0. : 8.) +
                  ((mode == WM_YLM_W || mode == WM_YLMC_W2) ? 24. : 0.);
  ProfScope ps(ctx, KC_ASSIGN, c.n * pbytes + double(ctx.g.nmesh) * sizeof(c2<T>));
  switch (ctx.p.assignment_order) {
    case 1: assign_order<T, 1>(ctx, mesh, c, tv, mode, ell, m, scale, cplx_w); break;
    case 2: assign_order<T, 2>(ctx, mesh, c, tv, mode, ell, m, scale, cplx_w); break;
    case 3: assign_order<T, 3>(ctx, mesh, c, tv, mode, ell, m, scale, cplx_w); break;
    default: assign_order<T, 4>(ctx, mesh, c, tv, mode, ell, m, scale, cplx_w); break;
  }
}

template <typename T, int ORDER>
static void assign_dual_order(Context& ctx, T* ra, T* rb, const DeviceCatalogue& c,
                              const TiledView& tv, int ell, int m, double scale_a,
                              double scale_b) {
  DevParticles sv = c.view();
  const bool cplx = m != 0;
  if (tile_grid(tv) > 0) {
    const size_t smem = size_t(cplx ? 4 : 2) * TILE_CELLS * sizeof(double);
    if (cplx)
      hipLaunchKernelGGL((k_assign_tiled_dual<T, ORDER, true>), dim3(tile_grid(tv)),
                         dim3(256), smem, ctx.stream, ra, rb, sv, ctx.g, tv.tg,
                         tv.work.get(), tv.meta.get(), ell, m, scale_a, scale_b);
    else
      hipLaunchKernelGGL((k_assign_tiled_dual<T, ORDER, false>), dim3(tile_grid(tv)),
                         dim3(256), smem, ctx.stream, ra, rb, sv, ctx.g, tv.tg,
                         tv.work.get(), tv.meta.get(), ell, m, scale_a, scale_b);
    TRVG_LAUNCH_CHECK();
  }
  if (tv.nstraggle != 0) {
    // Real outputs sharing one mesh (rb = ra + 1): B goes to imaginary parts.
    const bool shared = rb == ra + 1;
    c2<T>* ma = reinterpret_cast<c2<T>*>(ra);
    c2<T>* mb = shared ? ma : reinterpret_cast<c2<T>*>(rb);
    const int cwa = cplx ? 1 : 0, cwb = cplx ? 1 : (shared ? 2 : 0);
    const unsigned sg = tv.nstraggle > 0 ? grid_for(ctx, tv.nstraggle, 16) : 64u;
    hipLaunchKernelGGL((k_assign<T, ORDER>), dim3(sg), dim3(256), 0, ctx.stream, ma, sv,
                       ctx.g, WM_YLM_W, ell, m, scale_a, cwa, 0, tv.meta.get() + 1);
    hipLaunchKernelGGL((k_assign<T, ORDER>), dim3(sg), dim3(256), 0, ctx.stream, mb, sv,
                       ctx.g, WM_YLMC_W2, ell, m, scale_b, cwb, 0, tv.meta.get() + 1);
    TRVG_LAUNCH_CHECK();
  }
}

// A += scale_a sum y_lm w W / V_cell and B += scale_b sum y_lm^* w^2 W /
// V_cell over catalogue c in one pass (survey catalogues: lines of sight
// needed); ra, rb point at the real part of cell 0 of each output.
template <typename T>
static void assign_dual(Context& ctx, T* ra, T* rb, const DeviceCatalogue& c, int ell,
                        int m, double scale_a, double scale_b) {
  if (c.los.get() == nullptr)
    throw Error(TRVG_ERR_DATA, "Lines of sight are required for survey-type fields.");
  const TiledView& tv = tiled_view(ctx, c);
  // Algorithmic bytes: position, weight and line of sight read once, the
  // output values written once.
  const double outb = (m != 0 ? 2. : 1.) * double(ctx.g.nmesh) * sizeof(c2<T>);
  ProfScope ps(ctx, KC_ASSIGN, c.n * 56. + outb);
  switch (ctx.p.assignment_order) {
    case 1: assign_dual_order<T, 1>(ctx, ra, rb, c, tv, ell, m, scale_a, scale_b); break;
    case 2: assign_dual_order<T, 2>(ctx, ra, rb, c, tv, ell, m, scale_a, scale_b); break;
    case 3: assign_dual_order<T, 3>(ctx, ra, rb, c, tv, ell, m, scale_a, scale_b); break;
    default: assign_dual_order<T, 4>(ctx, ra, rb, c, tv, ell, m, scale_a, scale_b); break;
  }
}

template <typename T>
void assign_add(Context& ctx, c2<T>* mesh, const DeviceCatalogue& c, int mode,
                int ell, int m, double scale) {
  assign<T>(ctx, mesh, c, mode, ell, m, scale);
}
template void assign_add<double>(Context&, double2*, const DeviceCatalogue&, int, int,
                                 int, double);

template <typename T>
void assign_field(Context& ctx, c2<T>* mesh, const DeviceCatalogue& c, int mode,
                  int ell, int m, double scale) {
  zero_field<T>(ctx, mesh);
  assign<T>(ctx, mesh, c, mode, ell, m, scale);
}
template void assign_field<double>(Context&, double2*, const DeviceCatalogue&,
                                   int, int, int, double);
template void assign_field<float>(Context&, float2*, const DeviceCatalogue&,
                                  int, int, int, double);

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

// One catalogue's field alpha sum_c y w (quad: alpha^2 sum_c y^* w^2),
// forward transformed with the V_cell factor (field.cpp:1325-1362,
// 1449-1489).
static void build_single_field_k(Context& ctx, double2* f, const DeviceCatalogue& c,
                                 double alpha, int ell, int m, bool quad) {
  zero_field<double>(ctx, f);
  assign<double>(ctx, f, c, quad ? WM_YLMC_W2 : WM_YLM_W, ell, m,
                 quad ? alpha * alpha : alpha);
  LoadScaleShift<double> ld{f, 0., ctx.g.vol_cell};
  fft3d<double>(ctx, f, ld, -1);
}

// Survey pair of one (l, m): the density contrast
//   dn = sum_d y w - alpha sum_r y w
// and its shot-noise field
//   N = sum_d y^* w^2 + alpha^2 sum_r y^* w^2,
// both assigned in one pass per catalogue and forward transformed.  For
// m = 0 both fields are real: they are assigned as dn + i N into one mesh,
// transformed once and separated by Hermitian symmetry.  `data_assigned`
// (catalogues of uniform weights only): dn already holds S_d, as
// prefill_survey_data leaves it.
static void build_survey_pair_k(Context& ctx, double2* dn, double2* N,
                                const DeviceCatalogue& d, const DeviceCatalogue& r,
                                double alpha, int ell, int m,
                                bool data_assigned = false) {
  double* rdn = reinterpret_cast<double*>(dn);
  // Catalogues of uniform weights: one accumulation per particle, S_c =
  // sum_c y w W, and the pair from the transforms by linearity (for a
  // uniform w, y^* w^2 = w (y w)^*).
  const bool uniform = std::isfinite(d.wuni) && std::isfinite(r.wuni);
  if (data_assigned && !uniform)
    throw Error(TRVG_ERR_PARAM, "Prefilled data fields need uniform weights.");
  if (m == 0) {
    if (uniform) {
      if (!data_assigned) {
        zero_field<double>(ctx, dn);
        assign<double>(ctx, dn, d, WM_YLM_W, ell, 0, 1.);
      }
      assign<double>(ctx, dn, r, WM_YLM_W, ell, 0, 1., true);
    } else {
      zero_field<double>(ctx, dn);
      assign_dual<double>(ctx, rdn, rdn + 1, d, ell, m, 1., 1.);
      assign_dual<double>(ctx, rdn, rdn + 1, r, ell, m, -alpha, alpha * alpha);
    }
    LoadScaleShift<double> la{dn, 0., ctx.g.vol_cell};
    fft3d<double>(ctx, dn, la, -1);
    ProfScope ps(ctx, KC_MISC, 3. * double(ctx.g.nmesh) * sizeof(double2));
    // Plain separation (dn + i N), or dn = S_d - alpha S_r and
    // N = w_d S_d + alpha^2 w_r S_r from S_d + i S_r.
    const double pa = 1., pb = uniform ? -alpha : 0.;
    const double qa = uniform ? d.wuni : 0., qb = uniform ? alpha * alpha * r.wuni : 1.;
    hipLaunchKernelGGL(k_unpack_hermitian, dim3(grid_for(ctx, ctx.g.nmesh)), dim3(256), 0,
                       ctx.stream, dn, N, ctx.g, pa, pb, qa, qb);
    TRVG_LAUNCH_CHECK();
    return;
  }
  if (uniform) {
    // S_d in dn, S_r in a scratch mesh; N (out of place) first, then dn in
    // place over S_d, each formed by its transform's first-pass loader.
    DevArray<double2> sr(ctx.g.nmesh);
    if (!data_assigned) zero_field<double>(ctx, dn);
    zero_field<double>(ctx, sr.get());
    if (!data_assigned) assign<double>(ctx, dn, d, WM_YLM_W, ell, m, 1.);
    assign<double>(ctx, sr.get(), r, WM_YLM_W, ell, m, 1.);
    const double vc = ctx.g.vol_cell, rb2 = 2. * double(ctx.g.nmesh) * sizeof(double2);
    LoadLinear2<double> lb{dn, sr.get(), d.wuni * vc, alpha * alpha * r.wuni * vc, 1};
    fft3d<double>(ctx, N, lb, -1, rb2);
    LoadLinear2<double> la{dn, sr.get(), vc, -alpha * vc, 0};
    fft3d<double>(ctx, dn, la, -1, rb2);
    ctx.keep_until_flush(std::move(sr));
    return;
  }
  zero_field<double>(ctx, dn);
  zero_field<double>(ctx, N);
  assign_dual<double>(ctx, rdn, reinterpret_cast<double*>(N), d, ell, m, 1., 1.);
  assign_dual<double>(ctx, rdn, reinterpret_cast<double*>(N), r, ell, m, -alpha,
                      alpha * alpha);
  LoadScaleShift<double> la{dn, 0., ctx.g.vol_cell};
  fft3d<double>(ctx, dn, la, -1);
  LoadScaleShift<double> lb{N, 0., ctx.g.vol_cell};
  fft3d<double>(ctx, N, lb, -1);
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------

// n_modes and k_eff per bin depend on the geometry and the edges only:
// computed once per device and reused.
ShellStats shell_stats(Context& ctx, const trvg_binning& kb) {
  const int nb = kb.num_bins;
  std::vector<double> key = {double(ctx.dev)};
  for (int a = 0; a < 3; a++) { key.push_back(ctx.g.n[a]); key.push_back(ctx.g.dk[a]); }
  key.insert(key.end(), kb.bin_edges, kb.bin_edges + nb + 1);
  static std::mutex mu;
  static std::map<std::vector<double>, ShellStats> cache;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  DevArray<double> edges(nb + 1);
  TRVG_HIP(hipMemcpyAsync(edges.get(), kb.bin_edges, (nb + 1) * 8,
                          hipMemcpyHostToDevice, ctx.stream));
  unsigned grid = grid_for(ctx, ctx.g.nmesh, 4);
  DevArray<unsigned long long> cnt(nb);
  DevArray<double> part((size_t)grid * nb), ks(nb);
  TRVG_HIP(hipMemsetAsync(cnt.get(), 0, nb * 8, ctx.stream));
  size_t smem = nb * (8 + 4);
  ProfScope ps(ctx, KC_SHELL_STATS, 0.);
  hipLaunchKernelGGL(k_shell_stats, dim3(grid), dim3(RED_THREADS), smem,
                     ctx.stream, ctx.g, edges.get(), nb, cnt.get(), part.get());
  TRVG_LAUNCH_CHECK();
  reduce_cols(ctx, part.get(), int(grid), nb, ks.get());
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
  std::lock_guard<std::mutex> lk(mu);
  if (cache.size() >= 64) cache.clear();
  cache[key] = s;
  return s;
}

// The two-point binning map: z-line windows and offsets on the host (the
// window only has to contain every cell with |k| <= twopt_kmax), the exact
// bin rule per cell on the device.
std::shared_ptr<const TwoptMapData> twopt_map(Context& ctx, const double* edges, int nb) {
  if (nb < 1 || nb >= int(TWOPT_NONE)) throw std::invalid_argument("twopt_map: bin count");
  std::vector<double> key = {double(ctx.dev)};
  for (int a = 0; a < 3; a++) { key.push_back(ctx.g.n[a]); key.push_back(ctx.g.dk[a]); }
  key.insert(key.end(), edges, edges + nb + 1);
  static std::mutex mu;
  static std::map<std::vector<double>, std::shared_ptr<const TwoptMapData>> cache;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  const Geom& g = ctx.g;
  const long long nlines = (long long)g.n[0] * g.n[1];
  const int half = g.n[2] / 2;
  const double kmax = twopt_kmax(edges[nb]);
  std::vector<int> K(nlines);
  std::vector<long long> off(nlines);
  long long tot = 0;
  for (long long l = 0; l < nlines; l++) {
    const double kx = shifted(int(l / g.n[1]), g.n[0]) * g.dk[0];
    const double ky = shifted(int(l % g.n[1]), g.n[1]) * g.dk[1];
    const double rem = kmax * kmax - (kx * kx + ky * ky);
    off[l] = tot;
    if (rem < 0.) { K[l] = -1; continue; }
    K[l] = int(std::sqrt(rem) / g.dk[2]) + 1;
    tot += (K[l] >= half ? half : K[l]) + 1;
  }
  auto d = std::make_shared<TwoptMapData>();
  d->K.reset(nlines);
  d->off.reset(nlines);
  d->bin.reset(tot > 0 ? tot : 1);
  DevArray<double> de(nb + 1);
  TRVG_HIP(hipMemcpyAsync(d->K.get(), K.data(), nlines * sizeof(int), hipMemcpyHostToDevice,
                          ctx.stream));
  TRVG_HIP(hipMemcpyAsync(d->off.get(), off.data(), nlines * sizeof(long long),
                          hipMemcpyHostToDevice, ctx.stream));
  TRVG_HIP(hipMemcpyAsync(de.get(), edges, (nb + 1) * 8, hipMemcpyHostToDevice, ctx.stream));
  hipLaunchKernelGGL(k_twopt_binmap, dim3(grid_for(ctx, nlines * 64, 4)), dim3(RED_THREADS), 0,
                     ctx.stream, g, d->view(), de.get(), nb, d->bin.get());
  TRVG_LAUNCH_CHECK();
  // The host vectors and `de` must outlive the copies and the build.
  TRVG_HIP(hipStreamSynchronize(ctx.stream));
  std::lock_guard<std::mutex> lk(mu);
  if (cache.size() >= 16) cache.clear();   // bounded; maps in use stay alive
  cache[key] = d;
  return d;
}

// Launch of the direct-to-LDS pair kernel (opting in to more than 64 KiB of
// dynamic LDS where a launch needs it).
template <typename T, int MODE>
static void launch_pair_lds(Context& ctx, unsigned grid, size_t smem, const c2<T>* fa,
                            int nas, const c2<T>* fb, int nbs, const c2<T>* G,
                            long long nm, const int2* tiles, int nt, double2* part) {
  if (smem > 64 * 1024)
    TRVG_HIP(hipFuncSetAttribute((const void*)k_pair_lds<T, MODE>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, int(smem)));
  hipLaunchKernelGGL((k_pair_lds<T, MODE>), dim3(grid), dim3(256), smem, ctx.stream, fa,
                     nas, fb, nbs, G, nm, tiles, nt, part);
  TRVG_LAUNCH_CHECK();
}

// S_p = sum_x F_a(x) F_b(x) G(x) for pairs p = (a, b) of logical fields,
// handed to `sink` at the next ctx.flush() (the launches are only queued).
// A side / B side packed: real fields stored two per complex mesh (logical
// field 2q + c is component c of mesh q); g_real: G is real.
template <typename T>
static void pair_reduce(Context& ctx, const c2<T>* FA, int na, bool a_packed,
                        const c2<T>* FB, int nbf, bool b_packed, const c2<T>* G,
                        bool g_real, const std::vector<int2>& pairs,
                        std::function<void(std::vector<cplx>&&)> sink) {
  if (a_packed && !b_packed) {
    // The product is symmetric in (a, b): put the packed side second.
    std::vector<int2> sw(pairs.size());
    for (size_t d = 0; d < pairs.size(); d++) sw[d] = make_int2(pairs[d].y, pairs[d].x);
    pair_reduce<T>(ctx, FB, nbf, false, FA, na, true, G, g_real, sw, std::move(sink));
    return;
  }
  if (a_packed && !g_real)
    throw Error(TRVG_ERR_INTERNAL, "Real shell pairs need a real G field.");
  const int mode = (a_packed && b_packed) ? 2 : (b_packed ? 1 : 0);
  const long long nm = ctx.g.nmesh;
  const int rbn = (na + 3) / 4, cbn = (nbf + 3) / 4;
  // Active 4x4 tiles of the (ia, ib) grid.
  std::vector<char> active((size_t)rbn * cbn, 0);
  for (const int2& pr : pairs) active[(size_t)(pr.x / 4) * cbn + pr.y / 4] = 1;
  // Per launch group: its tiles, origin and the host copy of its sums.
  struct Group {
    std::vector<int2> tl;
    int ra0, ca0;
    const double2* h;
  };
  std::vector<Group> groups;
  // Column blocks per launch and row blocks to match: at most 32 fields a
  // side, PT_MAXT tiles and PT_MAXROWS staged rows (4 per block a side).
  const int CB = std::min(cbn, 8);
  const int RB = std::max(1, std::min(std::min((PT_MAXROWS - 1 - 4 * CB) / 4, PT_MAXT / CB), 8));
  // Whole chunks; the kernel reads the cells past the last one directly.
  const long long nchunks = nm / PairLds<T>::LDC;
  for (int c0b = 0; c0b < cbn; c0b += CB) {
    const int c1b = std::min(cbn, c0b + CB);
    for (int r0b = 0; r0b < rbn; r0b += RB) {
      const int r1b = std::min(rbn, r0b + RB);
      std::vector<int2> tl;
      for (int rb = r0b; rb < r1b; rb++)
        for (int cb = c0b; cb < c1b; cb++)
          if (active[(size_t)rb * cbn + cb]) tl.push_back(make_int2(4 * (rb - r0b), 4 * (cb - c0b)));
      if (tl.empty()) continue;
      const int ra0 = 4 * r0b, ca0 = 4 * c0b;
      const int nas = std::min(na, 4 * r1b) - ra0;
      const int nbs = std::min(nbf, 4 * c1b) - ca0;
      const int nt = int(tl.size());
      DevArray<int2> dt(nt);
      ctx.upload(dt.get(), tl.data(), nt);
      // Resident blocks only (each loops over chunks): as many per CU as
      // the double-buffered staging allows.
      const bool same_side = mode == 2 && FA == FB && ra0 == ca0 && nas == nbs;
      const int kmode = same_side ? 3 : mode;
      const size_t smem = pair_lds_smem<T>(kmode, nas, nbs);
      const int per_cu = std::max(1, std::min(8, int((160 * 1024) / smem)));
      const unsigned grid = unsigned(std::min<long long>(nchunks, (long long)ctx.sm_count * per_cu));
      DevArray<double2> part((size_t)grid * nt * 16), sum((size_t)nt * 16);
      {
        // Packed sides start at mesh ra0 / 2 (group starts are multiples of 4).
        const c2<T>* fa = FA + (long long)(mode == 2 ? ra0 / 2 : ra0) * nm;
        const c2<T>* fb = FB + (long long)(mode >= 1 ? ca0 / 2 : ca0) * nm;
        const double ma = pair_rows_a(kmode, nas), mb = pair_rows_b(kmode, nbs);
        ProfScope ps(ctx, KC_PAIR_REDUCE, (ma + mb + 1.) * nm * sizeof(c2<T>));
        switch (kmode) {
          case 3: launch_pair_lds<T, 3>(ctx, grid, smem, fa, nas, fb, nbs, G, nm, dt.get(), nt, part.get()); break;
          case 2: launch_pair_lds<T, 2>(ctx, grid, smem, fa, nas, fb, nbs, G, nm, dt.get(), nt, part.get()); break;
          case 1: launch_pair_lds<T, 1>(ctx, grid, smem, fa, nas, fb, nbs, G, nm, dt.get(), nt, part.get()); break;
          default: launch_pair_lds<T, 0>(ctx, grid, smem, fa, nas, fb, nbs, G, nm, dt.get(), nt, part.get()); break;
        }
        TRVG_LAUNCH_CHECK();
      }
      reduce_cols(ctx, part.get(), int(grid), nt * 16, sum.get());
      groups.push_back(Group{tl, ra0, ca0, ctx.download(sum.get(), (size_t)nt * 16)});
      ctx.keep_until_flush(std::move(dt));
      ctx.keep_until_flush(std::move(part));
      ctx.keep_until_flush(std::move(sum));
    }
  }
  ctx.defer([groups = std::move(groups), pairs, na, nbf, sink = std::move(sink)]() {
    // Sums for every pair of every active tile, indexed [ia * nbf + ib].
    std::vector<cplx> full((size_t)na * nbf, cplx(0., 0.));
    for (const Group& gr : groups)
      for (size_t t = 0; t < gr.tl.size(); t++)
        for (int i = 0; i < 4; i++)
          for (int j = 0; j < 4; j++) {
            const int a = gr.ra0 + gr.tl[t].x + i, b = gr.ca0 + gr.tl[t].y + j;
            if (a < na && b < nbf) {
              const double2 v = gr.h[t * 16 + i * 4 + j];
              full[(size_t)a * nbf + b] = cplx(v.x, v.y);
            }
          }
    std::vector<cplx> out(pairs.size());
    for (size_t d = 0; d < pairs.size(); d++)
      out[d] = full[(size_t)pairs[d].x * nbf + pairs[d].y];
    sink(std::move(out));
  });
}

struct TwoPt {
  std::vector<int> nmodes;
  std::vector<cplx> pk, sn;
};

// FieldStats::compute_ylm_wgtd_2pt_stats_in_fourier (field.cpp:2511-2703).
template <typename T>
static void twopt_fourier(Context& ctx, const c2<T>* fa, const c2<T>* fb, cplx sbar,
                          int ell, int m, const trvg_binning& kb, double modes_in_range,
                          std::function<void(TwoPt&&)> sink) {
  const int nb = kb.num_bins;
  std::shared_ptr<const TwoptMapData> map = twopt_map(ctx, kb.bin_edges, nb);
  // Latency-bound (dependent field loads per cell): eight workgroups per CU.
  unsigned grid = grid_for(ctx, ctx.g.nmesh, 8);
  DevArray<unsigned long long> cnt(nb);
  DevArray<double> part((size_t)grid * nb * 4), sum(nb * 4);
  TRVG_HIP(hipMemsetAsync(cnt.get(), 0, nb * 8, ctx.stream));
  // Compulsory reads: both fields at the modes inside the binning range.
  ProfScope ps(ctx, KC_TWOPT, 2. * modes_in_range * sizeof(c2<T>));
  auto kern = ctx.tab.interlace ? k_twopt_fourier<T, 4, true> : k_twopt_fourier<T, 4, false>;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(RED_THREADS),
                     twopt_smem(nb, 4, ctx.g), ctx.stream, fa, fb, ctx.g, ctx.tab,
                     dmk(sbar.real(), sbar.imag()), ell, m, map->view(), nb,
                     cnt.get(), part.get());
  TRVG_LAUNCH_CHECK();
  reduce_cols(ctx, part.get(), int(grid), nb * 4, sum.get());
  const unsigned long long* hc = ctx.download(cnt.get(), nb);
  const double* hs = ctx.download(sum.get(), (size_t)nb * 4);
  ctx.defer([map]() {});   // keeps the map alive until the launch has run
  ctx.keep_until_flush(std::move(cnt));
  ctx.keep_until_flush(std::move(part));
  ctx.keep_until_flush(std::move(sum));
  ctx.defer([hc, hs, nb, sink = std::move(sink)]() {
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
    sink(std::move(r));
  });
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
  ProfScope ps(ctx, KC_TWOPT, double(ctx.g.nmesh) * sizeof(c2<T>));
  hipLaunchKernelGGL((k_3pcf_shot_bins<T>), dim3(grid), dim3(RED_THREADS), smem,
                     ctx.stream, t, ctx.g, la, ma, lb, mb, edges.get(), nb,
                     cnt.get(), part.get());
  TRVG_LAUNCH_CHECK();
  reduce_cols(ctx, part.get(), int(grid), nb * 3, sum.get());
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
// Sorted distinct values (bins of a side).
static std::vector<int> distinct(const std::vector<int>& v) {
  std::vector<int> u(v);
  std::sort(u.begin(), u.end());
  u.erase(std::unique(u.begin(), u.end()), u.end());
  return u;
}

// Distinct bins of each side of a pair list and the pairs re-indexed into
// them: wavenumber lists (k_eff of the bins) for the shot-noise tables.
struct PairBinLists {
  std::vector<double> ka, kb;
  std::vector<int2> pidx;
};
static PairBinLists pair_bin_lists(const std::vector<double>& keff,
                                   const std::vector<int>& ra,
                                   const std::vector<int>& cb) {
  PairBinLists L;
  std::vector<int> ua = distinct(ra), ub = distinct(cb);
  for (int b : ua) L.ka.push_back(keff[b]);
  for (int b : ub) L.kb.push_back(keff[b]);
  L.pidx.resize(ra.size());
  for (size_t d = 0; d < ra.size(); d++)
    L.pidx[d] = make_int2(int(std::lower_bound(ua.begin(), ua.end(), ra[d]) - ua.begin()),
                          int(std::lower_bound(ub.begin(), ub.end(), cb[d]) - ub.begin()));
  return L;
}

// Index of the radial histogram's staged entries by n = a^2 + b^2 + c^2
// (see k_radial_stage / k_radial_gather), built once per device and mesh.
struct RadialIndex {
  DevArray<int> ofs, idx;
  long long nq = 0, nentries = 0;
};

static std::shared_ptr<const RadialIndex> radial_index(Context& ctx) {
  const Geom& g = ctx.g;
  const std::vector<long long> key = {ctx.dev, g.n[0], g.n[1], g.n[2]};
  static std::mutex mu;
  static std::map<std::vector<long long>, std::shared_ptr<const RadialIndex>> cache;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  const int h0 = g.n[0] / 2 + 1, h1 = g.n[1] / 2 + 1, h2 = g.n[2] / 2 + 1;
  const bool swap = g.n[0] == g.n[1];
  const long long nrows = swap ? (long long)h0 * (h0 + 1) / 2 : (long long)h0 * h1;
  const long long ne = nrows * h2;
  if (ne >= (1LL << 31))
    throw Error(TRVG_ERR_PARAM, "Mesh too large for the radial histogram index.");
  auto ri = std::make_shared<RadialIndex>();
  ri->nq = (long long)(g.n[0] / 2) * (g.n[0] / 2) + (long long)(g.n[1] / 2) * (g.n[1] / 2) +
           (long long)(g.n[2] / 2) * (g.n[2] / 2) + 1;
  ri->nentries = ne;
  std::vector<long long> rowsq(nrows);
  for (int a = 0; a < h0; a++)
    for (int b = 0; b < (swap ? a + 1 : h1); b++)
      rowsq[swap ? (long long)a * (a + 1) / 2 + b : (long long)a * h1 + b] =
          (long long)a * a + (long long)b * b;
  std::vector<int> ofs(ri->nq + 1, 0), idx(ne);
  for (long long r = 0; r < nrows; r++)
    for (int c = 0; c < h2; c++) ofs[rowsq[r] + (long long)c * c + 1]++;
  for (long long n = 0; n < ri->nq; n++) ofs[n + 1] += ofs[n];
  std::vector<int> cur(ofs.begin(), ofs.end() - 1);
  for (long long r = 0; r < nrows; r++)
    for (int c = 0; c < h2; c++) idx[cur[rowsq[r] + (long long)c * c]++] = int(r * h2 + c);
  ri->ofs.reset(ri->nq + 1);
  ri->idx.reset(ne);
  TRVG_HIP(hipMemcpyAsync(ri->ofs.get(), ofs.data(), (ri->nq + 1) * sizeof(int),
                          hipMemcpyHostToDevice, ctx.stream));
  TRVG_HIP(hipMemcpyAsync(ri->idx.get(), idx.data(), ne * sizeof(int), hipMemcpyHostToDevice,
                          ctx.stream));
  TRVG_HIP(hipStreamSynchronize(ctx.stream));   // the host vectors go out of scope
  std::lock_guard<std::mutex> lk(mu);
  if (cache.size() >= 8) cache.clear();   // bounded; indexes in use stay alive
  cache[key] = ri;
  return ri;
}

template <typename T>
static void shot_pairs_bispec(Context& ctx, const c2<T>* t, int la, int ma, int lb,
                              int mb, const Spline& sa, const Spline& sb,
                              const std::vector<double>& keff, const std::vector<int>& ra,
                              const std::vector<int>& cb,
                              std::function<void(std::vector<cplx>&&)> sink) {
  const int P = int(ra.size());
  const Geom& g = ctx.g;
  const bool cubic = (g.dr[0] == g.dr[1] && g.dr[1] == g.dr[2]);
  if (cubic) {
    long long nq = (long long)(g.n[0] / 2) * (g.n[0] / 2)
                   + (long long)(g.n[1] / 2) * (g.n[1] / 2)
                   + (long long)(g.n[2] / 2) * (g.n[2] / 2) + 1;
    const PairBinLists L = pair_bin_lists(keff, ra, cb);
    const int nka = int(L.ka.size()), nkb = int(L.kb.size());
    const int nch = shot_tab_chunk(nka, nkb);
    const size_t smem = shot_tab_smem(nka, nkb, nch);
    if (smem > 160 * 1024)
      throw Error(TRVG_ERR_PARAM, "Too many wavenumber bins for the shot-noise tables.");
    DevArray<double> dka(nka), dkb(nkb);
    DevArray<int2> dp(P);
    ctx.upload(dka.get(), L.ka.data(), nka);
    ctx.upload(dkb.get(), L.kb.data(), nkb);
    ctx.upload(dp.get(), L.pidx.data(), P);
    const long long nchunks = (nq + nch - 1) / nch;
    const std::shared_ptr<const RadialIndex> ri = radial_index(ctx);
    DevArray<double> Q(2 * nq);
    DevArray<double2> S(ri->nentries);
    DevArray<double2> part((size_t)nchunks * P), res(P);
    {
      // Algorithmic traffic: t read once (the staging and the index are
      // the price of summing without atomics).
      ProfScope ps(ctx, KC_RADIAL, double(g.nmesh) * sizeof(c2<T>));
      hipLaunchKernelGGL((k_radial_stage<T>), dim3(grid_for(ctx, ri->nentries, 8)),
                         dim3(RED_THREADS), 0, ctx.stream, t, g, la, ma, lb, mb, S.get());
      TRVG_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_radial_gather, dim3(grid_for(ctx, nq * RG_LANES, 8)),
                         dim3(RED_THREADS), 0,
                         ctx.stream, S.get(), ri->ofs.get(), ri->idx.get(), nq, Q.get());
      TRVG_LAUNCH_CHECK();
    }
    {
      // The histogram is read once; partials written and reduced once.
      ProfScope ps(ctx, KC_SHOT_PAIRS, double(nq) * 16. + 2. * nchunks * P * 16.);
      hipLaunchKernelGGL(k_shot_pairs_tab, dim3(unsigned(nchunks)), dim3(256), smem,
                         ctx.stream, Q.get(), nq, g.dr[0], sa.dev(), sb.dev(), dka.get(),
                         nka, dkb.get(), nkb, dp.get(), P, nch, g.vol_cell, part.get());
      TRVG_LAUNCH_CHECK();
      reduce_cols(ctx, part.get(), int(nchunks), P, res.get());
    }
    const double2* h = ctx.download(res.get(), P);
    ctx.keep_until_flush(std::move(dka));
    ctx.keep_until_flush(std::move(dkb));
    ctx.keep_until_flush(std::move(dp));
    ctx.keep_until_flush(std::move(Q));
    ctx.keep_until_flush(std::move(S));
    ctx.keep_until_flush(std::move(part));
    ctx.keep_until_flush(std::move(res));
    ctx.defer([h, P, ri, sink = std::move(sink)]() {
      std::vector<cplx> out(P);
      for (int p = 0; p < P; p++) out[p] = cplx(h[p].x, h[p].y);
      sink(std::move(out));
    });
    return;
  }
  std::vector<double> ka(P), kb(P);
  for (int d = 0; d < P; d++) { ka[d] = keff[ra[d]]; kb[d] = keff[cb[d]]; }
  DevArray<double> dka(P), dkb(P);
  ctx.upload(dka.get(), ka.data(), P);
  ctx.upload(dkb.get(), kb.data(), P);
  const int GP = 8;
  unsigned grid = grid_for(ctx, g.nmesh, 4);
  std::vector<const double2*> hs;
  for (int p0 = 0; p0 < P; p0 += GP) {
    int np = std::min(GP, P - p0);
    DevArray<double2> part((size_t)grid * np), sum(np);
    ProfScope ps(ctx, KC_SHOT_PAIRS, double(g.nmesh) * sizeof(c2<T>));
    hipLaunchKernelGGL((k_shot_pairs_direct<T>), dim3(grid), dim3(RED_THREADS), 0,
                       ctx.stream, t, g, la, ma, lb, mb, sa.dev(), sb.dev(),
                       dka.get() + p0, dkb.get() + p0, np, part.get());
    TRVG_LAUNCH_CHECK();
    reduce_cols(ctx, part.get(), int(grid), np, sum.get());
    hs.push_back(ctx.download(sum.get(), np));
    ctx.keep_until_flush(std::move(part));
    ctx.keep_until_flush(std::move(sum));
  }
  ctx.keep_until_flush(std::move(dka));
  ctx.keep_until_flush(std::move(dkb));
  ctx.defer([hs, P, GP, vol = g.vol_cell, sink = std::move(sink)]() {
    std::vector<cplx> out(P);
    for (int p0 = 0, gi = 0; p0 < P; p0 += GP, gi++) {
      const int np = std::min(GP, P - p0);
      for (int q = 0; q < np; q++)
        out[p0 + q] = cplx(hs[gi][q].x, hs[gi][q].y) * vol;
    }
    sink(std::move(out));
  });
}

// S-bar_LM = S-bar_LM[data] + alpha^3 S-bar_LM[randoms] for every (L, M)
// in `keys` (threept.cpp:771-800; with no data catalogue, alpha^3
// S-bar_LM[randoms], threept.cpp:209-237): one pass per catalogue for up to
// SBAR_MAX_KEYS keys, partials downloaded, one synchronisation.  Call with
// no deferred work pending on the context.
static std::map<std::pair<int, int>, cplx> sbar_table(
    Context& ctx, const DeviceCatalogue* d, const DeviceCatalogue& r, double alpha,
    const std::set<std::pair<int, int>>& keys) {
  const std::vector<std::pair<int, int>> kv(keys.begin(), keys.end());
  struct Chunk {
    size_t k0;
    int nk;
    const double2* hd;
    const double2* hr;
    unsigned gd, gr;
  };
  std::vector<Chunk> chunks;
  for (size_t k0 = 0; k0 < kv.size(); k0 += SBAR_MAX_KEYS) {
    SbarKeys sk{};
    sk.n = int(std::min<size_t>(SBAR_MAX_KEYS, kv.size() - k0));
    for (int q = 0; q < sk.n; q++) {
      sk.ell[q] = kv[k0 + q].first;
      sk.m[q] = kv[k0 + q].second;
    }
    Chunk c{k0, sk.n, nullptr, nullptr, 0u, 0u};
    for (int side = d ? 0 : 1; side < 2; side++) {
      const DeviceCatalogue& cat = side == 0 ? *d : r;
      const unsigned grid = grid_for(ctx, cat.n, 4);
      DevArray<double2> part((size_t)grid * sk.n);
      hipLaunchKernelGGL(k_sbar_multi, dim3(grid), dim3(RED_THREADS), 0, ctx.stream,
                         cat.view(), sk, part.get());
      TRVG_LAUNCH_CHECK();
      const double2* h = ctx.download(part.get(), (size_t)grid * sk.n);
      ctx.keep_until_flush(std::move(part));
      if (side == 0) { c.hd = h; c.gd = grid; } else { c.hr = h; c.gr = grid; }
    }
    chunks.push_back(c);
  }
  ctx.flush();
  std::map<std::pair<int, int>, cplx> out;
  for (const Chunk& c : chunks)
    for (int q = 0; q < c.nk; q++) {
      double dx = 0., dy = 0., rx = 0., ry = 0.;
      for (unsigned b = 0; b < c.gd && c.hd; b++) { dx += c.hd[b * c.nk + q].x; dy += c.hd[b * c.nk + q].y; }
      for (unsigned b = 0; b < c.gr; b++) { rx += c.hr[b * c.nk + q].x; ry += c.hr[b * c.nk + q].y; }
      out[kv[c.k0 + q]] = cplx(dx, dy) + std::pow(alpha, 3) * cplx(rx, ry);
    }
  return out;
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

// ---------------------------------------------------------------------------
// Shell-field slabs, reused across terms: a slab holds the inverse-
// transformed shell fields of one (l, m) for a set of bins.  B_202 needs the
// (0, 0) side in every term, B_000 / zeta_000 share one slab for both sides.
// ---------------------------------------------------------------------------

template <typename T>
struct ShellSlab {
  int ell = -1, m = 0;
  std::vector<int> index;   // bin -> field index in `data`, -1 if absent
  int nfields = 0;
  bool packed = false;      // real fields, two per complex mesh
  DevArray<c2<T>> data;
  bool covers(int l, int mm, const std::vector<int>& bins) const {
    if (ell != l || m != mm) return false;
    for (int b : bins)
      if (b >= int(index.size()) || index[b] < 0) return false;
    return true;
  }
};

template <typename T>
class SlabCache {
 public:
  explicit SlabCache(size_t capacity = 2) : cap_(capacity < 2 ? 2 : capacity) {}

  // Returns a slab of (l, m) covering `bins`, building it with
  // make(slab_base, bins) if needed (field q of the slab at
  // slab_base + q * nmesh, or component q % 2 of mesh q / 2 when packed);
  // evicts the least recently used slab other than `keep` when full.
  template <class Make>
  ShellSlab<T>* get(Context& ctx, int l, int m, const std::vector<int>& bins,
                    int nb, const ShellSlab<T>* keep, bool packed, Make&& make) {
    for (size_t i = 0; i < entries_.size(); i++)
      if (entries_[i]->covers(l, m, bins)) {
        auto e = std::move(entries_[i]);      // move to the back (MRU)
        entries_.erase(entries_.begin() + i);
        entries_.push_back(std::move(e));
        return entries_.back().get();
      }
    if (entries_.size() >= cap_) {
      for (size_t i = 0; i < entries_.size(); i++)
        if (entries_[i].get() != keep) {
          entries_.erase(entries_.begin() + i);
          break;
        }
    }
    auto e = std::make_unique<ShellSlab<T>>();
    e->ell = l;
    e->m = m;
    e->index.assign(nb, -1);
    e->nfields = int(bins.size());
    e->packed = packed;
    const long long nm = ctx.g.nmesh;
    e->data.reset((size_t)nm * (packed ? (bins.size() + 1) / 2 : bins.size()));
    make(e->data.get(), bins);
    for (size_t q = 0; q < bins.size(); q++) e->index[bins[q]] = int(q);
    entries_.push_back(std::move(e));
    return entries_.back().get();
  }

 private:
  size_t cap_;
  std::vector<std::unique_ptr<ShellSlab<T>>> entries_;
};

// ---------------------------------------------------------------------------
// Bispectrum
// ---------------------------------------------------------------------------

SurveyPrefill prefill_survey_data(Context& ctx, const DeviceCatalogue& d,
                                  const std::vector<Multipole>& mps) {
  SurveyPrefill pre;
  if (!std::isfinite(d.wuni)) return pre;
  // Keys in the order the estimator uses them: the common (0, 0) pair, then
  // the (L, M) of each term that is not (0, 0).
  std::vector<std::pair<int, int>> keys{{0, 0}};
  for (const Multipole& mp : mps) {
    trvg_params p = ctx.p;
    p.ell1 = mp.ell1; p.ell2 = mp.ell2; p.ELL = mp.ELL;
    for (const Term& tm : enumerate_terms(p, true))
      if (!(p.ELL == 0 && tm.M == 0)) keys.push_back({p.ELL, tm.M});
  }
  constexpr size_t cap = size_t(4) << 30;
  const size_t bytes = size_t(ctx.g.nmesh) * sizeof(double2);
  size_t held = 0;
  for (const auto& k : keys) {
    if (pre.sd.count(k)) continue;
    if (held + bytes > cap) break;
    held += bytes;
    // As build_survey_pair_k assigns the data side of a uniform pair.
    DevArray<double2> f(ctx.g.nmesh);
    zero_field<double>(ctx, f.get());
    assign<double>(ctx, f.get(), d, WM_YLM_W, k.first, k.second, 1.);
    pre.sd.emplace(k, std::move(f));
  }
  return pre;
}

template <typename T>
std::vector<Result> run_bispec_multi(Context& ctx, const DeviceCatalogue& d,
                                     const DeviceCatalogue* r, bool survey,
                                     const trvg_binning& kb,
                                     const std::vector<Multipole>& mps,
                                     SurveyPrefill* pre) {
  const Geom& g = ctx.g;
  const long long nm = g.nmesh;
  const double alpha = survey ? d.wstotal / r->wstotal : 1.;
  const int nb = kb.num_bins;

  // S_d of (L, M) assigned ahead of the random catalogue (uniform weights
  // on both sides only), taken once; an empty array when there is none.
  const bool uniform = survey && std::isfinite(d.wuni) && std::isfinite(r->wuni);
  auto take_prefill = [&](int L, int M) {
    DevArray<double2> f;
    if (pre && uniform) {
      auto it = pre->sd.find({L, M});
      if (it != pre->sd.end()) {
        f = std::move(it->second);
        pre->sd.erase(it);
      }
    }
    return f;
  };

  // Common fields in FP64 whatever the shell precision T: the shot-noise
  // path (two-point statistics, S_ij_k) subtracts nearly equal quantities.
  DevArray<double2> dn00 = take_prefill(0, 0), N00(nm);
  const bool dn00_pre = dn00.get() != nullptr;
  if (!dn00_pre) dn00.reset(nm);
  if (survey) {
    build_survey_pair_k(ctx, dn00.get(), N00.get(), d, *r, alpha, 0, 0, dn00_pre);
  } else {
    build_field_k<double>(ctx, dn00.get(), d, r, false, alpha, 0, 0, false, true);
    build_field_k<double>(ctx, N00.get(), d, r, false, alpha, 0, 0, true, false);
  }
  ShellStats ss = shell_stats(ctx, kb);
  double modes_in_range = 0.;
  for (int n : ss.nmodes) modes_in_range += n;

  auto spline = [&](int ell) -> const Spline& {
    return cached_spline(ctx.dev, ell, ctx.stream);
  };

  DevArray<c2<T>> G(nm);
  // t(x) in the working precision (an FP32 transform in the FP32 path; its
  // radial histogram and the pair sums over it accumulate in FP64).
  DevArray<c2<T>> tfield(nm);
  DevArray<double2> dnLM, NLM;
  // Inputs of the shot-noise field held in tfield: (L, M) of its dn field
  // (-1 for dn00) and Sbar; a periodic box uses one t(x) for every term and
  // multipole.
  int t_ell = -2, t_m = 0;
  cplx t_sbar(0., 0.);
  SlabCache<T> slabs(mps.size() > 1 ? 3 : 2);
  auto make_shell = [&](int l, int mm) {
    return [&, l, mm](c2<T>* base, const std::vector<int>& bins) {
      shell_slab_ifft<T>(ctx, base, dn00.get(), kb, bins, ss.nmodes, l, mm,
                         real_shells(l, mm));
    };
  };

  // S-bar_LM of every term of every multipole, computed up front.
  std::map<std::pair<int, int>, cplx> sbar;
  if (survey) {
    std::set<std::pair<int, int>> keys;
    for (const Multipole& mp : mps) {
      trvg_params p = ctx.p;
      p.ell1 = mp.ell1; p.ell2 = mp.ell2; p.ELL = mp.ELL;
      for (const Term& tm : enumerate_terms(p, survey)) keys.insert({p.ELL, tm.M});
    }
    sbar = sbar_table(ctx, &d, *r, alpha, keys);
  }

  std::vector<Result> results;
  for (const Multipole& mp : mps) {
    trvg_params p = ctx.p;
    p.ell1 = mp.ell1; p.ell2 = mp.ell2; p.ELL = mp.ELL;
    p.shape = mp.shape; p.idx_bin = mp.idx_bin;
    const cplx factor_phase = ipow(p.ell1 + p.ell2);
    std::vector<int> ra, cb;
    pair_bins(p.shape, nb, p.idx_bin, ra, cb);
    const int dim = int(ra.size());
    std::vector<int> rows = distinct(ra), cols = distinct(cb);

    Result out;
    out.dim = dim;
    out.b1.resize(dim); out.b2.resize(dim); out.e1.resize(dim); out.e2.resize(dim);
    out.n1.resize(dim); out.n2.resize(dim);
    out.raw.assign(dim, 0.); out.shot.assign(dim, 0.);
    for (int dv = 0; dv < dim; dv++) {
      out.b1[dv] = kb.bin_centres[ra[dv]];
      out.b2[dv] = kb.bin_centres[cb[dv]];
      out.e1[dv] = ss.keff[ra[dv]];
      out.e2[dv] = ss.keff[cb[dv]];
      out.n1[dv] = ss.nmodes[ra[dv]];
      out.n2[dv] = ss.nmodes[cb[dv]];
    }
    const Spline& sja = spline(p.ell1);
    const Spline& sjb = spline(p.ell2);

    for (const Term& tm : enumerate_terms(p, survey)) {
      const bool lm_zero = !survey || (p.ELL == 0 && tm.M == 0);
      const double2* dnLMp = dn00.get();
      const double2* NLMp = N00.get();
      if (!lm_zero) {
        if (!NLM.get()) NLM.reset(nm);
        DevArray<double2> own = take_prefill(p.ELL, tm.M);
        if (own.get()) {
          build_survey_pair_k(ctx, own.get(), NLM.get(), d, *r, alpha, p.ELL, tm.M, true);
          dnLMp = own.get();
          ctx.keep_until_flush(std::move(own));   // read by this term's kernels
        } else {
          if (!dnLM.get()) dnLM.reset(nm);
          build_survey_pair_k(ctx, dnLM.get(), NLM.get(), d, *r, alpha, p.ELL, tm.M);
          dnLMp = dnLM.get();
        }
        NLMp = NLM.get();
      }

      // G_LM(x): compensated, inverse-transformed (threept.cpp:452-459).
      LoadDeconv<T, double> ldg{dnLMp, ctx.tab, 1. / g.vol};
      fft3d<T>(ctx, G.get(), ldg, +1, double(nm) * sizeof(double2));

      // G_LM is real for M = 0 (real weights Y_L0) and in a periodic box.
      const bool g_real = !survey || tm.M == 0;
      // Shell fields F_a (l1, m1) for the row bins, F_b (l2, m2) for columns.
      ShellSlab<T>* A = slabs.get(ctx, p.ell1, tm.m1, rows, nb, nullptr,
                                  real_shells(p.ell1, tm.m1), make_shell(p.ell1, tm.m1));
      ShellSlab<T>* B = slabs.get(ctx, p.ell2, tm.m2, cols, nb, A,
                                  real_shells(p.ell2, tm.m2), make_shell(p.ell2, tm.m2));
      std::vector<int2> pairs(dim);
      for (int dv = 0; dv < dim; dv++)
        pairs[dv] = make_int2(A->index[ra[dv]], B->index[cb[dv]]);
      // Results arrive at the flush after the last term (the sinks add
      // into `out`, which outlives them).
      const double coupling = tm.coupling, f_mirror = tm.f_mirror;
      pair_reduce<T>(ctx, A->data.get(), A->nfields, A->packed, B->data.get(),
                     B->nfields, B->packed, G.get(), g_real, pairs,
                     [&out, &ra, &cb, &ss, dim, coupling, f_mirror,
                      vc = g.vol_cell](std::vector<cplx>&& S) {
                       for (int dv = 0; dv < dim; dv++) {
                         cplx sv = S[dv] / (double(ss.nmodes[ra[dv]]) *
                                            double(ss.nmodes[cb[dv]]));
                         out.raw[dv] += coupling * vc * (sv + f_mirror * std::conj(sv));
                       }
                     });

      // Shot noise (threept.cpp:771-966).
      const cplx Sbar = survey ? sbar.at({p.ELL, tm.M}) : cplx(double(d.n), 0.);
      if (p.ell1 == 0 && p.ell2 == 0) {
        cplx sijk = tm.coupling * Sbar;
        for (int dv = 0; dv < dim; dv++) out.shot[dv] += sijk;
      }
      // Two-point terms of the row side (ell2 = 0) and/or the column side
      // (ell1 = 0); with both sides monopoles of one m the statistics are
      // shared.
      auto twopt_sink = [&out, &ra, &cb, dim, coupling, f_mirror](bool rows, bool cols) {
        return [&out, &ra, &cb, dim, coupling, f_mirror, rows, cols](TwoPt&& st) {
          for (int side = 0; side < 2; side++) {
            if (!(side == 0 ? rows : cols)) continue;
            const std::vector<int>& bins = side == 0 ? ra : cb;
            for (int dv = 0; dv < dim; dv++) {
              int bb = bins[dv];
              cplx sv = coupling * (st.pk[bb] - st.sn[bb]);
              out.shot[dv] += sv + f_mirror * std::conj(sv);
            }
          }
        };
      };
      const bool shared = p.ell1 == 0 && p.ell2 == 0 && tm.m1 == tm.m2;
      if (p.ell2 == 0)
        twopt_fourier<double>(ctx, dn00.get(), NLMp, Sbar, p.ell1, tm.m1, kb,
                              modes_in_range, twopt_sink(true, shared));
      if (p.ell1 == 0 && !shared)
        twopt_fourier<double>(ctx, dn00.get(), NLMp, Sbar, p.ell2, tm.m2, kb,
                              modes_in_range, twopt_sink(false, true));
      const int key_ell = lm_zero ? -1 : p.ELL, key_m = lm_zero ? 0 : tm.M;
      if (key_ell != t_ell || key_m != t_m || Sbar != t_sbar) {
        LoadTwopt<T, double> lt{dnLMp, N00.get(), ctx.tab,
                             dmk(Sbar.real(), Sbar.imag()), 1. / g.vol};
        fft3d<T>(ctx, tfield.get(), lt, +1, 2. * double(nm) * sizeof(double2));
        t_ell = key_ell;
        t_m = key_m;
        t_sbar = Sbar;
      }
      shot_pairs_bispec<T>(
          ctx, tfield.get(), p.ell1, tm.m1, p.ell2, tm.m2, sja, sjb, ss.keff, ra, cb,
          [&out, dim, coupling, factor_phase, fmw = tm.f_mirror_w3j](std::vector<cplx>&& sij) {
            for (int dv = 0; dv < dim; dv++) {
              cplx sv = coupling * sij[dv];
              out.shot[dv] += factor_phase * (sv + fmw * std::conj(sv));
            }
          });
    }
    ctx.flush();
    for (int dv = 0; dv < dim; dv++) {
      out.raw[dv] *= mp.norm;
      out.shot[dv] *= mp.norm;
    }
    results.push_back(std::move(out));
  }
  return results;
}

template <typename T>
Result run_bispec(Context& ctx, const DeviceCatalogue& d,
                  const DeviceCatalogue* r, bool survey,
                  const trvg_binning& kb, double norm_factor, SurveyPrefill* pre) {
  Multipole mp{ctx.p.ell1, ctx.p.ell2, ctx.p.ELL, ctx.p.shape, ctx.p.idx_bin,
               norm_factor};
  return std::move(run_bispec_multi<T>(ctx, d, r, survey, kb, {mp}, pre)[0]);
}

// ---------------------------------------------------------------------------
// Three-point correlation function
// ---------------------------------------------------------------------------

template <typename T>
Result run_3pcf(Context& ctx, const DeviceCatalogue& d,
                const DeviceCatalogue* r, bool survey,
                const trvg_binning& rb, double norm_factor,
                const Window3pcf* win) {
  const trvg_params& p = ctx.p;
  const Geom& g = ctx.g;
  const long long nm = g.nmesh;
  const double alpha = win ? win->alpha : (survey ? d.wstotal / r->wstotal : 1.);
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

  DevArray<double2> dn00(nm), N00(nm);
  if (win) {
    build_single_field_k(ctx, dn00.get(), d, alpha, 0, 0, false);
    build_single_field_k(ctx, N00.get(), d, alpha, 0, 0, true);
  } else if (survey) {
    build_survey_pair_k(ctx, dn00.get(), N00.get(), d, *r, alpha, 0, 0);
  } else {
    build_field_k<double>(ctx, dn00.get(), d, r, false, alpha, 0, 0, false, true);
    build_field_k<double>(ctx, N00.get(), d, r, false, alpha, 0, 0, true, false);
  }

  const Spline& sja = cached_spline(ctx.dev, p.ell1, ctx.stream);
  const Spline& sjb = cached_spline(ctx.dev, p.ell2, ctx.stream);

  std::vector<int> rows = distinct(ra), cols = distinct(cb);
  DevArray<c2<T>> G(nm);
  DevArray<c2<T>> tfield(nm);
  DevArray<double2> dnLM;
  SlabCache<T> slabs;
  std::vector<double> reff(nb, 0.);
  bool first = true;
  std::map<std::pair<int, int>, cplx> sbar;
  if (survey) {
    std::set<std::pair<int, int>> keys;
    for (const Term& tm : enumerate_terms(p, survey)) keys.insert({p.ELL, tm.M});
    sbar = win ? sbar_table(ctx, nullptr, d, alpha, keys)
               : sbar_table(ctx, &d, *r, alpha, keys);
  }

  for (const Term& tm : enumerate_terms(p, survey)) {
    const bool lm_zero = !survey || (p.ELL == 0 && tm.M == 0);
    const double2* dnLMp = dn00.get();
    if (!lm_zero) {
      if (!dnLM.get()) dnLM.reset(nm);
      if (win)
        build_single_field_k(ctx, dnLM.get(), d, alpha, p.ELL, tm.M, false);
      else
        build_field_k<double>(ctx, dnLM.get(), d, r, true, alpha, p.ELL, tm.M, false, false);
      dnLMp = dnLM.get();
    }
    const cplx Sbar = survey ? sbar.at({p.ELL, tm.M}) : cplx(double(d.n), 0.);

    // Shot noise first: defines r_eff and n_pairs (threept.cpp:1223-1367).
    LoadTwopt<T, double> lt{dnLMp, N00.get(), ctx.tab, dmk(Sbar.real(), Sbar.imag()),
                         1. / g.vol};
    fft3d<T>(ctx, tfield.get(), lt, +1, 2. * double(nm) * sizeof(double2));
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

    LoadDeconv<T, double> ldg{dnLMp, ctx.tab, 1. / g.vol};
    fft3d<T>(ctx, G.get(), ldg, +1, double(nm) * sizeof(double2));
    if (win && win->wa_power >= 0) {
      // Wide-angle power-law kernel (field.cpp:1727-1762).
      ProfScope ps(ctx, KC_MISC, 2. * double(nm) * sizeof(c2<T>));
      hipLaunchKernelGGL((k_wide_angle<T>), dim3(grid_for(ctx, nm)), dim3(RED_THREADS), 0,
                         ctx.stream, G.get(), g, win->wa_power);
      TRVG_LAUNCH_CHECK();
    }

    auto make_sjl = [&](int l, int mm, const Spline& sp) {
      return [&, l, mm](c2<T>* base, const std::vector<int>& bins) {
        sjl_slab_ifft<T>(ctx, base, dn00.get(), sp, reff, bins, l, mm, real_shells(l, mm));
      };
    };
    ShellSlab<T>* A = slabs.get(ctx, p.ell1, tm.m1, rows, nb, nullptr,
                                real_shells(p.ell1, tm.m1), make_sjl(p.ell1, tm.m1, sja));
    ShellSlab<T>* B = slabs.get(ctx, p.ell2, tm.m2, cols, nb, A,
                                real_shells(p.ell2, tm.m2), make_sjl(p.ell2, tm.m2, sjb));
    std::vector<int2> pairs(dim);
    for (int dv = 0; dv < dim; dv++)
      pairs[dv] = make_int2(A->index[ra[dv]], B->index[cb[dv]]);
    // G_LM is real for M = 0 (real weights Y_L0) and in a periodic box.
    const bool g_real = !survey || tm.M == 0;
    pair_reduce<T>(ctx, A->data.get(), A->nfields, A->packed, B->data.get(), B->nfields,
                   B->packed, G.get(), g_real, pairs,
                   [&out, dim, c = tm.coupling * g.vol_cell * factor_phase,
                    fm = tm.f_mirror](std::vector<cplx>&& S) {
                     for (int dv = 0; dv < dim; dv++)
                       out.raw[dv] += c * (S[dv] + fm * std::conj(S[dv]));
                   });
  }
  ctx.flush();
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
  // A single assignment and reduction: always FP64.
  DevArray<double2> f(ctx.g.nmesh);
  zero_field<double>(ctx, f.get());
  assign<double>(ctx, f.get(), c, WM_W, 0, 0, 1.);
  unsigned grid = grid_for(ctx, ctx.g.nmesh, 4);
  DevArray<double> part(grid), sum(1);
  hipLaunchKernelGGL((k_cube_sum<double>), dim3(grid), dim3(RED_THREADS), 0,
                     ctx.stream, f.get(), ctx.g.nmesh, part.get());
  TRVG_LAUNCH_CHECK();
  reduce_cols(ctx, part.get(), int(grid), 1, sum.get());
  double h = 0.;
  TRVG_HIP(hipMemcpyAsync(&h, sum.get(), 8, hipMemcpyDeviceToHost, ctx.stream));
  TRVG_HIP(hipStreamSynchronize(ctx.stream));
  return 1. / (h * ctx.g.vol_cell);
}

template std::vector<Result> run_bispec_multi<double>(
    Context&, const DeviceCatalogue&, const DeviceCatalogue*, bool,
    const trvg_binning&, const std::vector<Multipole>&, SurveyPrefill*);
template std::vector<Result> run_bispec_multi<float>(
    Context&, const DeviceCatalogue&, const DeviceCatalogue*, bool,
    const trvg_binning&, const std::vector<Multipole>&, SurveyPrefill*);
template Result run_bispec<double>(Context&, const DeviceCatalogue&,
                                   const DeviceCatalogue*, bool,
                                   const trvg_binning&, double, SurveyPrefill*);
template Result run_bispec<float>(Context&, const DeviceCatalogue&,
                                  const DeviceCatalogue*, bool,
                                  const trvg_binning&, double, SurveyPrefill*);
template Result run_3pcf<double>(Context&, const DeviceCatalogue&,
                                 const DeviceCatalogue*, bool,
                                 const trvg_binning&, double, const Window3pcf*);
template Result run_3pcf<float>(Context&, const DeviceCatalogue&,
                                const DeviceCatalogue*, bool,
                                const trvg_binning&, double, const Window3pcf*);
template double run_mesh_norm<double>(Context&, const DeviceCatalogue&);
template double run_mesh_norm<float>(Context&, const DeviceCatalogue&);
template void fft3d<double, LoadSame<double>>(Context&, double2*,
                                              const LoadSame<double>&, int,
                                              double, const Prune*);
template void fft3d<float, LoadSame<float>>(Context&, float2*,
                                            const LoadSame<float>&, int, double,
                                            const Prune*);

// Loaders of the two-point path (engine_twopt.hip).
template void fft3d<double, LoadScaleShift<double>>(Context&, double2*,
                                                    const LoadScaleShift<double>&, int,
                                                    double, const Prune*);
template void fft3d<double, LoadTwopt<double, double>>(Context&, double2*,
                                                       const LoadTwopt<double, double>&,
                                                       int, double, const Prune*);

// Loaders of the differentiable path (engine_diff.hip).
#define TRVG_FFT3D_INST(T, L)                                                   \
  template void fft3d<T, L>(Context&, c2<T>*, const L&, int, double, const Prune*);
TRVG_FFT3D_INST(double, LoadReal<double>)
TRVG_FFT3D_INST(float, LoadReal<float>)
TRVG_FFT3D_INST(double, LoadShellCotangent<double>)
TRVG_FFT3D_INST(float, LoadShellCotangent<float>)
#define TRVG_COMMA ,
TRVG_FFT3D_INST(double, LoadDeconv<double TRVG_COMMA double>)
TRVG_FFT3D_INST(float, LoadDeconv<float TRVG_COMMA float>)
TRVG_FFT3D_INST(double, LoadShell<double TRVG_COMMA double>)
TRVG_FFT3D_INST(float, LoadShell<float TRVG_COMMA float>)
#undef TRVG_COMMA
#undef TRVG_FFT3D_INST

}  // namespace trvg
