// engine_noalloc.hip -- the differentiable periodic-box B_000 diag forward
// (trvg_diff_bispec_diag) without any device allocation, for callers that
// cannot allocate (code under HIP graph capture, serving loops, harnesses
// that forbid hipMalloc).
//
// Every buffer beyond the caller's mesh and outputs lives in one
// file-scope device array, `scratch`, of NOALLOC_SCRATCH_BYTES: the three
// complex meshes of the forward (the density transform D, the common field
// G and one shell field F), the per-block partial sums, the shell
// statistics, the per-axis window tables and the FFT twiddle tables, all
// recomputed on every call.  The array is resolved to a device address once
// per device (hipGetSymbolAddress, with that device current); host code only
// ever passes that address, never the array's name.  Calls on one device
// share the array: calls that overlap in time on one device race.  Work runs
// on the calling thread's default stream and the call returns when its
// results are on the host.
#include <cmath>
#include <mutex>

#include "diff_kernels.hpp"
#include "engine.hpp"

namespace trvg {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;   // 64 MiB
constexpr int NOALLOC_MAX_BINS = 256;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device (resolved once).
unsigned char* scratch_address(int dev) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    throw Error(TRVG_ERR_DEVICE, "Device index out of range for the scratch table.");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      (void)hipGetLastError();
      throw Error(TRVG_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                       hipGetErrorString(e));
    }
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  return g_scratch_ptr[dev];
}

// Bump allocation inside the scratch array (256-byte aligned pieces).
struct ScratchCursor {
  unsigned char* base;
  size_t used = 0;
  template <typename E>
  E* take(size_t n) {
    E* p = reinterpret_cast<E*>(base + used);
    used += (n * sizeof(E) + 255) & ~size_t(255);
    return p;
  }
};

// e^{sign 2 pi i m / n}, m < n.
template <typename T>
__global__ void k_twiddle_fill(c2<T>* __restrict__ tw, int n, int sign) {
  for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < n; m += gridDim.x * blockDim.x) {
    double s, c;
    sincos(2. * M_PI * double(m) / double(n), &s, &c);
    tw[m] = mk<T>(T(c), T(sign * s));
  }
}

// Per-axis sin(u)/u, sin(u)^2 and u, u = pi shifted(i) / n, in the layout
// [kind][axis][i] with axis offsets o0, o1, o2 inside each kind.
__global__ void k_axis_tables(double* __restrict__ t, int n0, int n1, int n2) {
  const int tot = n0 + n1 + n2;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < 3 * tot;
       q += gridDim.x * blockDim.x) {
    const int kind = q / tot, r = q % tot;
    const int a = r < n0 ? 0 : (r < n0 + n1 ? 1 : 2);
    const int n = a == 0 ? n0 : (a == 1 ? n1 : n2);
    const int i = r - (a == 0 ? 0 : (a == 1 ? n0 : n0 + n1));
    const int ii = shifted(i, n);
    const double u = M_PI * ii / double(n);
    double v;
    if (kind == 0) v = (ii != 0) ? sin(u) / u : 1.;
    else if (kind == 1) v = (ii != 0) ? sin(u) * sin(u) : 0.;
    else v = u;
    t[q] = v;
  }
}

__global__ void k_clear_u64(unsigned long long* __restrict__ p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = 0ull;
}

int log2_exact(int n) {
  int l = 0;
  while ((1 << l) < n) l++;
  return l;
}

}  // namespace

// Scratch bytes the forward needs on a mesh of nm cells with nb bins.
template <typename T>
static size_t noalloc_bytes(const Geom& g, int nb, unsigned grid) {
  auto r = [](size_t b) { return (b + 255) & ~size_t(255); };
  size_t b = 3 * r(size_t(g.nmesh) * sizeof(c2<T>));
  b += r(size_t(grid) * nb * 8) + r(size_t(nb) * 8);             // partials, sums
  b += r(size_t(nb + 1) * 8) + r(size_t(nb) * 8) +                 // edges, counts
       r(size_t(grid) * nb * 8) + r(size_t(nb) * 8);              // k sums
  b += r(size_t(3) * (g.n[0] + g.n[1] + g.n[2]) * 8);             // axis tables
  b += 6 * r(size_t(1024) * sizeof(c2<T>));                       // twiddles
  return b;
}

template <typename T>
static void noalloc_forward(const trvg_params& q, const trvg_binning& kb, double norm,
                            const T* mesh, double* bk, int* nmodes_out) {
  const int nb = kb.num_bins;
  if (nb > NOALLOC_MAX_BINS)
    throw Error(TRVG_ERR_PARAM, "The allocation-free forward takes at most " +
                                    std::to_string(NOALLOC_MAX_BINS) + " bins.");
  int dev = 0;
  TRVG_HIP(hipGetDevice(&dev));
  int sm = 256;
  TRVG_HIP(hipDeviceGetAttribute(&sm, hipDeviceAttributeMultiprocessorCount, dev));
  const Geom g0 = make_geom(q);
  const unsigned grid = unsigned(std::min<long long>(
      (g0.nmesh + RED_THREADS - 1) / RED_THREADS, (long long)sm * 4));
  const size_t need = noalloc_bytes<T>(g0, nb, grid);
  if (need > NOALLOC_SCRATCH_BYTES)
    throw Error(TRVG_ERR_PARAM,
                "Mesh too large for the allocation-free forward: it needs " +
                    std::to_string(need) + " bytes of scratch, the library holds " +
                    std::to_string(NOALLOC_SCRATCH_BYTES) + ".");

  ScratchCursor sc{scratch_address(dev)};
  hipStream_t s = nullptr;   // the calling thread's default stream
  const long long nm = g0.nmesh;
  c2<T>* D = sc.take<c2<T>>(nm);
  c2<T>* G = sc.take<c2<T>>(nm);
  c2<T>* F = sc.take<c2<T>>(nm);
  double* part = sc.take<double>(size_t(grid) * nb);
  double* sums = sc.take<double>(nb);
  double* edges = sc.take<double>(nb + 1);
  unsigned long long* cnt = sc.take<unsigned long long>(nb);
  double* kpart = sc.take<double>(size_t(grid) * nb);
  double* ksum = sc.take<double>(nb);
  double* tabs = sc.take<double>(size_t(3) * (g0.n[0] + g0.n[1] + g0.n[2]));

  // Per-axis tables and twiddles for every axis length and both signs.
  AxisTables tab;
  {
    const int tot = g0.n[0] + g0.n[1] + g0.n[2];
    hipLaunchKernelGGL(k_axis_tables, dim3((3 * tot + 255) / 256), dim3(256), 0, s, tabs,
                       g0.n[0], g0.n[1], g0.n[2]);
    TRVG_LAUNCH_CHECK();
    long long off = 0;
    for (int kind = 0; kind < 3; kind++)
      for (int a = 0; a < 3; a++) {
        const double* p = tabs + off;
        if (kind == 0) tab.sinc[a] = p;
        else if (kind == 1) tab.sin2[a] = p;
        else tab.u[a] = p;
        off += g0.n[a];
      }
  }
  Context ctx(q, s, tab);
  for (int a = 0; a < 3; a++) {
    const int n = g0.n[a], lg = log2_exact(n);
    for (int sg = 0; sg < 2; sg++) {
      if (ctx.tw[sg][lg]) continue;
      c2<T>* tw = sc.take<c2<T>>(n);
      hipLaunchKernelGGL((k_twiddle_fill<T>), dim3((n + 255) / 256), dim3(256), 0, s, tw, n,
                         sg ? +1 : -1);
      TRVG_LAUNCH_CHECK();
      ctx.tw[sg][lg] = tw;
    }
  }

  // Mode counts per bin (cleared by a kernel: no memset node if captured).
  TRVG_HIP(hipMemcpyAsync(edges, kb.bin_edges, size_t(nb + 1) * 8, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_clear_u64, dim3(1), dim3(256), 0, s, cnt, nb);
  hipLaunchKernelGGL(k_shell_stats, dim3(grid), dim3(RED_THREADS), size_t(nb) * (8 + 4), s,
                     ctx.g, edges, nb, cnt, kpart);
  TRVG_LAUNCH_CHECK();
  reduce_cols(ctx, kpart, int(grid), nb, ksum);
  unsigned long long hc[NOALLOC_MAX_BINS];
  TRVG_HIP(hipMemcpyAsync(hc, cnt, size_t(nb) * 8, hipMemcpyDeviceToHost, s));
  TRVG_HIP(hipStreamSynchronize(s));

  // D = FFT[delta V_cell], G = IFFT[D / W / V], then per bin F = IFFT[shell
  // of D / W] and sum_x F^2 G (engine_diff.hip's forward, same kernels).
  fft3d<T>(ctx, D, LoadReal<T>{mesh, T(ctx.g.vol_cell)}, -1, double(nm) * sizeof(T));
  fft3d<T>(ctx, G, LoadDeconv<T, T>{D, ctx.tab, 1. / ctx.g.vol}, +1);
  for (int b = 0; b < nb; b++) {
    LoadShell<T, T> ls{D, ctx.tab, ctx.g,
                       shell_band(kb.bin_edges[b], kb.bin_edges[b + 1]), 0, 0};
    Prune pr = make_prune(ctx.g, kb.bin_edges[b + 1]);
    fft3d<T>(ctx, F, ls, +1, double(hc[b]) * sizeof(c2<T>), &pr);
    hipLaunchKernelGGL((k_f2g_partials<T>), dim3(grid), dim3(RED_THREADS), 0, s, F, G, nm,
                       part + b, nb);
    TRVG_LAUNCH_CHECK();
  }
  reduce_cols(ctx, part, int(grid), nb, sums);
  double h[NOALLOC_MAX_BINS];
  TRVG_HIP(hipMemcpyAsync(h, sums, size_t(nb) * 8, hipMemcpyDeviceToHost, s));
  TRVG_HIP(hipStreamSynchronize(s));
  for (int b = 0; b < nb; b++) {
    const double n = double(hc[b]);
    bk[b] = norm * ctx.g.vol_cell * h[b] / (n * n);   // NaN for an empty bin
    if (nmodes_out) nmodes_out[b] = int(hc[b]);
  }
}

template void noalloc_forward<double>(const trvg_params&, const trvg_binning&, double,
                                      const double*, double*, int*);
template void noalloc_forward<float>(const trvg_params&, const trvg_binning&, double,
                                     const float*, double*, int*);

void diff_bispec_forward_noalloc(const trvg_params& q, const trvg_binning& kb, double norm,
                                 const void* mesh, double* bk, int* nmodes) {
  if (q.precision == TRVG_PREC_FP32)
    noalloc_forward<float>(q, kb, norm, static_cast<const float*>(mesh), bk, nmodes);
  else
    noalloc_forward<double>(q, kb, norm, static_cast<const double*>(mesh), bk, nmodes);
}

size_t noalloc_scratch_bytes() { return NOALLOC_SCRATCH_BYTES; }

}  // namespace trvg
