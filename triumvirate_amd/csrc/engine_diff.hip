// engine_diff.hip -- differentiable periodic-box bispectrum B_000 (diag
// form) and mesh assignment: forward, hand-written adjoints, C ABI.
//
// The forward is the chain of compute_bispec_in_gpp_box (threept.cpp:
// 1473-2188) for l1 = l2 = L = 0 with the diag form, starting from a given
// real mesh delta(x) (the unweighted field fluctuation of
// field.cpp:1235-1243 when delta comes from trvg_diff_assign minus the
// mean).  Shell fields are recomputed in the backward pass instead of being
// stored.  Device pointers in, host vectors of per-bin values out.
#include <cmath>
#include <vector>

#include "diff_kernels.hpp"
#include "engine.hpp"

namespace trvg {

template <typename T>
static void diff_common_fields(Context& ctx, const T* mesh, c2<T>* D, c2<T>* G) {
  const Geom& g = ctx.g;
  fft3d<T>(ctx, D, LoadReal<T>{mesh, T(g.vol_cell)}, -1,
           double(g.nmesh) * sizeof(T));
  fft3d<T>(ctx, G, LoadDeconv<T, T>{D, ctx.tab, 1. / g.vol}, +1);
}

template <typename T>
static void diff_shell(Context& ctx, const c2<T>* D, c2<T>* F,
                       const trvg_binning& kb, int b, int nmodes) {
  LoadShell<T, T> ls{D, ctx.tab, ctx.g,
                     shell_band(kb.bin_edges[b], kb.bin_edges[b + 1]), 0, 0};
  Prune pr = make_prune(ctx.g, kb.bin_edges[b + 1]);
  fft3d<T>(ctx, F, ls, +1, double(nmodes) * sizeof(c2<T>), &pr);
}

template <typename T>
std::vector<double> diff_bispec_forward(Context& ctx, const trvg_binning& kb,
                                        double norm, const T* mesh,
                                        std::vector<int>& nmodes) {
  const Geom& g = ctx.g;
  const long long nm = g.nmesh;
  const int nb = kb.num_bins;
  ShellStats ss = shell_stats(ctx, kb);
  nmodes = ss.nmodes;
  DevArray<c2<T>> D(nm), G(nm), F(nm);
  diff_common_fields<T>(ctx, mesh, D.get(), G.get());
  const unsigned grid = grid_for(ctx, nm, 4);
  DevArray<double> part((size_t)grid * nb), sums(nb);   // [grid][nb]
  for (int b = 0; b < nb; b++) {
    diff_shell<T>(ctx, D.get(), F.get(), kb, b, ss.nmodes[b]);
    ProfScope ps(ctx, KC_PAIR_REDUCE, 2. * double(nm) * sizeof(c2<T>));
    hipLaunchKernelGGL((k_f2g_partials<T>), dim3(grid), dim3(RED_THREADS), 0,
                       ctx.stream, F.get(), G.get(), nm, part.get() + b, nb);
    TRVG_LAUNCH_CHECK();
  }
  reduce_cols(ctx, part.get(), int(grid), nb, sums.get());
  std::vector<double> h(nb);
  TRVG_HIP(hipMemcpyAsync(h.data(), sums.get(), nb * 8, hipMemcpyDeviceToHost, ctx.stream));
  TRVG_HIP(hipStreamSynchronize(ctx.stream));
  std::vector<double> out(nb);
  for (int b = 0; b < nb; b++) {
    const double n = double(ss.nmodes[b]);
    out[b] = norm * g.vol_cell * h[b] / (n * n);   // NaN for an empty bin
  }
  return out;
}

template <typename T>
void diff_bispec_backward(Context& ctx, const trvg_binning& kb, double norm,
                          const T* mesh, const double* grad_bk, T* grad_mesh) {
  const Geom& g = ctx.g;
  const long long nm = g.nmesh;
  const int nb = kb.num_bins;
  ShellStats ss = shell_stats(ctx, kb);
  DevArray<c2<T>> D(nm), G(nm), F(nm), acc(nm);
  DevArray<T> gG(nm);
  diff_common_fields<T>(ctx, mesh, D.get(), G.get());
  const unsigned grid = grid_for(ctx, nm);
  hipLaunchKernelGGL((k_zero<T>), dim3(grid), dim3(RED_THREADS), 0, ctx.stream,
                     acc.get(), nm);
  TRVG_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_zero_real<T>), dim3(grid), dim3(RED_THREADS), 0,
                     ctx.stream, gG.get(), nm);
  TRVG_LAUNCH_CHECK();
  for (int b = 0; b < nb; b++) {
    const double n = double(ss.nmodes[b]);
    if (grad_bk[b] == 0. || ss.nmodes[b] == 0) continue;
    const double a = grad_bk[b] * norm * g.vol_cell / (n * n);
    diff_shell<T>(ctx, D.get(), F.get(), kb, b, ss.nmodes[b]);
    // In place: F <- FFT[2 a F G]; gG += a F^2.
    fft3d<T>(ctx, F.get(), LoadShellCotangent<T>{F.get(), G.get(), gG.get(), a}, -1,
             double(nm) * (2. * sizeof(c2<T>) + 2. * sizeof(T)));
    ProfScope ps(ctx, KC_MISC, 3. * double(nm) * sizeof(c2<T>));
    hipLaunchKernelGGL((k_adj_shell_acc<T>), dim3(grid), dim3(RED_THREADS), 0,
                       ctx.stream, acc.get(), F.get(), g, ctx.tab,
                       shell_band(kb.bin_edges[b], kb.bin_edges[b + 1]));
    TRVG_LAUNCH_CHECK();
  }
  fft3d<T>(ctx, F.get(), LoadReal<T>{gG.get(), T(1)}, -1, double(nm) * sizeof(T));
  hipLaunchKernelGGL((k_adj_all_acc<T>), dim3(grid), dim3(RED_THREADS), 0,
                     ctx.stream, acc.get(), F.get(), g, ctx.tab, 1. / g.vol);
  TRVG_LAUNCH_CHECK();
  fft3d<T>(ctx, acc.get(), LoadSame<T>{acc.get()}, +1);
  hipLaunchKernelGGL((k_real_part<T>), dim3(grid), dim3(RED_THREADS), 0,
                     ctx.stream, grad_mesh, acc.get(), nm, g.vol_cell);
  TRVG_LAUNCH_CHECK();
  TRVG_HIP(hipStreamSynchronize(ctx.stream));
}

template <typename T, int ORDER>
static void diff_assign_order(Context& ctx, T* mesh, const double* pos,
                              const double* w, long long n, bool accumulate) {
  const long long nm = ctx.g.nmesh;
  if (!accumulate) {
    hipLaunchKernelGGL((k_zero_real<T>), dim3(grid_for(ctx, nm)), dim3(RED_THREADS),
                       0, ctx.stream, mesh, nm);
    TRVG_LAUNCH_CHECK();
  }
  if (n > 0) {
    ProfScope ps(ctx, KC_ASSIGN, double(n) * 32. + double(nm) * sizeof(T));
    hipLaunchKernelGGL((k_assign_real<T, ORDER>), dim3(grid_for(ctx, n, 16)),
                       dim3(256), 0, ctx.stream, mesh, pos, w, n, ctx.g);
    TRVG_LAUNCH_CHECK();
  }
  TRVG_HIP(hipStreamSynchronize(ctx.stream));
}

template <typename T>
void diff_assign(Context& ctx, T* mesh, const double* pos, const double* w,
                 long long n, bool accumulate) {
  switch (ctx.p.assignment_order) {
    case 1: diff_assign_order<T, 1>(ctx, mesh, pos, w, n, accumulate); break;
    case 2: diff_assign_order<T, 2>(ctx, mesh, pos, w, n, accumulate); break;
    case 3: diff_assign_order<T, 3>(ctx, mesh, pos, w, n, accumulate); break;
    default: diff_assign_order<T, 4>(ctx, mesh, pos, w, n, accumulate); break;
  }
}

template <typename T, int ORDER>
static void diff_assign_adjoint_order(Context& ctx, const T* gmesh,
                                      const double* pos, const double* w,
                                      long long n, double* gpos, double* gw) {
  if (n > 0) {
    ProfScope ps(ctx, KC_ASSIGN, double(n) * (32. + 32.));
    hipLaunchKernelGGL((k_assign_real_adjoint<T, ORDER>), dim3(grid_for(ctx, n, 16)),
                       dim3(256), 0, ctx.stream, gmesh, pos, w, n, ctx.g, gpos, gw);
    TRVG_LAUNCH_CHECK();
  }
  TRVG_HIP(hipStreamSynchronize(ctx.stream));
}

template <typename T>
void diff_assign_adjoint(Context& ctx, const T* gmesh, const double* pos,
                         const double* w, long long n, double* gpos, double* gw) {
  switch (ctx.p.assignment_order) {
    case 1: diff_assign_adjoint_order<T, 1>(ctx, gmesh, pos, w, n, gpos, gw); break;
    case 2: diff_assign_adjoint_order<T, 2>(ctx, gmesh, pos, w, n, gpos, gw); break;
    case 3: diff_assign_adjoint_order<T, 3>(ctx, gmesh, pos, w, n, gpos, gw); break;
    default: diff_assign_adjoint_order<T, 4>(ctx, gmesh, pos, w, n, gpos, gw); break;
  }
}

template std::vector<double> diff_bispec_forward<double>(
    Context&, const trvg_binning&, double, const double*, std::vector<int>&);
template std::vector<double> diff_bispec_forward<float>(
    Context&, const trvg_binning&, double, const float*, std::vector<int>&);
template void diff_bispec_backward<double>(Context&, const trvg_binning&, double,
                                           const double*, const double*, double*);
template void diff_bispec_backward<float>(Context&, const trvg_binning&, double,
                                          const float*, const double*, float*);
template void diff_assign<double>(Context&, double*, const double*, const double*,
                                  long long, bool);
template void diff_assign<float>(Context&, float*, const double*, const double*,
                                 long long, bool);
template void diff_assign_adjoint<double>(Context&, const double*, const double*,
                                          const double*, long long, double*, double*);
template void diff_assign_adjoint<float>(Context&, const float*, const double*,
                                         const double*, long long, double*, double*);

}  // namespace trvg
