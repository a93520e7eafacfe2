// fft_gen.hpp -- FFT passes for mesh axes that are not powers of two.
//
// An axis of length N (even, N <= 2048) is transformed in LDS by a
// mixed-radix Stockham sequence N = r_0 r_1 ... (radix 4 while it divides,
// then 2, then the odd primes in ascending order).  Stage s with span
// Ns = r_0 ... r_{s-1} and radix p forms every output directly,
//   out[(j / Ns) Ns p + k + r Ns] = sum_q in[j + q N/p] w^{q (k N/(Ns p) + r N/p)},
// j < N/p, k = j mod Ns, r < p, w = e^{sign 2 pi i / N}, reading one LDS
// buffer and writing the other (no register arrays, so any prime radix
// works; a radix-p stage costs p products per output).  Sign convention
// and normalisation as the power-of-two passes (FFTW's, unnormalised).
//
// A workgroup takes NL = max(1, 2048 / N) lines of one axis (the last block
// of an axis may hold fewer); the passes mirror the general (non-FAST)
// paths of fft.hpp: the z pass with a loader, the strided passes with the
// same pruning modes, the x-first loader pass and the j_l multi-field x
// pass, each also over (field, block) work lists.
#pragma once

#include "fft.hpp"

namespace trvg {

constexpr int GEN_ELEMS = FFT_THREADS * 8;   // target elements per workgroup
constexpr int GEN_MAX_N = 2048;
constexpr int GEN_MAX_STAGES = 16;

struct GenPlan {
  int n;      // line length
  int nl;     // lines per workgroup
  int nst;    // stages
  int radix[GEN_MAX_STAGES];
};

inline int gen_lines_per_block(int n) { return n <= GEN_ELEMS ? GEN_ELEMS / n : 1; }

inline GenPlan make_gen_plan(int n) {
  GenPlan p{};
  p.n = n;
  p.nl = gen_lines_per_block(n);
  int m = n;
  auto push = [&](int r) { p.radix[p.nst++] = r; };
  while (m % 4 == 0) { push(4); m /= 4; }
  while (m % 2 == 0) { push(2); m /= 2; }
  for (int q = 3; m > 1; q += 2)
    while (m % q == 0) { push(q); m /= q; }
  return p;
}

// Dynamic LDS of a pass: two line buffers (padded stride N + 1), the N
// twiddles and, for strided passes, one byte per row of pruning flags.
template <typename T>
inline size_t gen_smem(int n, bool row_flags) {
  const int nl = gen_lines_per_block(n);
  return size_t(2 * nl * (n + 1) + n) * sizeof(c2<T>) + (row_flags ? size_t(n) : 0);
}

// The nl lines of a (line stride N + 1) are transformed; returns the buffer
// (a or b) holding the result.  Ends with a barrier.
template <typename T>
__device__ c2<T>* lds_fft_gen(c2<T>* a, c2<T>* b, const c2<T>* W, const GenPlan& P,
                              int nl) {
  const int N = P.n, LS = N + 1;
  const int total = nl * N;
  int Ns = 1;
  for (int s = 0; s < P.nst; s++) {
    const int p = P.radix[s];
    const int Np = N / p;
    const int tstep = N / (Ns * p);
    for (int o = threadIdx.x; o < total; o += FFT_THREADS) {
      const int line = o / N, rem = o % N;
      const int r = rem / Np, j = rem % Np;
      const int k = j % Ns;
      const int be = (k * tstep + r * Np) % N;
      const c2<T>* src = a + line * LS + j;
      c2<T> acc = mk<T>(T(0), T(0));
      int e = 0;
      for (int q = 0; q < p; q++) {
        acc = cadd<T>(acc, cmul<T>(src[q * Np], W[e]));
        e += be;
        if (e >= N) e -= N;
      }
      b[line * LS + (j / Ns) * Ns * p + k + r * Ns] = acc;
    }
    __syncthreads();
    c2<T>* t = a;
    a = b;
    b = t;
    Ns *= p;
  }
  return a;
}

template <typename T>
struct GenSmem {
  c2<T>* a;
  c2<T>* b;
  c2<T>* W;
  unsigned char* flags;
  __device__ explicit GenSmem(const GenPlan& P) {
    extern __shared__ char smem_raw[];
    const int LS = P.n + 1;
    a = reinterpret_cast<c2<T>*>(smem_raw);
    b = a + P.nl * LS;
    W = b + P.nl * LS;
    flags = reinterpret_cast<unsigned char*>(W + P.n);
  }
};

// z pass (contiguous lines, line index = i n1 + j) with a loader.
template <typename T, class Loader>
__device__ void fft_z_block_gen(c2<T>* __restrict__ data, const Loader& load,
                                const c2<T>* __restrict__ tw, int n1, long long nlines,
                                int sign, const Prune& pr, const GenPlan& P,
                                long long blk) {
  (void)sign;
  const int N = P.n, LS = N + 1;
  GenSmem<T> sm(P);
  const long long line0 = blk * P.nl;
  if (line0 >= nlines) return;
  const int nl = int(min((long long)P.nl, nlines - line0));
  if (pr.on) {
    bool any = false;
    for (int l = 0; l < nl && !any; l++) {
      const long long line = line0 + l;
      any = pr.p1(int(line / n1), int(line % n1));
    }
    if (!any) return;
  }
  for (int m = threadIdx.x; m < N; m += FFT_THREADS) sm.W[m] = tw[m];
  for (int idx = threadIdx.x; idx < nl * N; idx += FFT_THREADS) {
    const int l = idx / N, k = idx % N;
    const long long line = line0 + l;
    sm.a[l * LS + k] = load(line * N + k, int(line / n1), int(line % n1), k);
  }
  __syncthreads();
  const c2<T>* R = lds_fft_gen<T>(sm.a, sm.b, sm.W, P, nl);
  for (int idx = threadIdx.x; idx < nl * N; idx += FFT_THREADS) {
    const int l = idx / N, k = idx % N;
    st_mesh<T>(data + (line0 + l) * N + k, R[l * LS + k]);
  }
}

template <typename T, class Loader>
__global__ void __launch_bounds__(FFT_THREADS)
fft_pass_z_gen(c2<T>* __restrict__ data, Loader load, const c2<T>* __restrict__ tw,
               int n1, long long nlines, int sign, Prune pr, GenPlan P) {
  fft_z_block_gen<T>(data, load, tw, n1, nlines, sign, pr, P, blockIdx.x);
}

template <typename T, class BatchLoader>
__global__ void __launch_bounds__(FFT_THREADS)
fft_pass_z_batch_gen(c2<T>* __restrict__ base, long long fstride, BatchLoader load,
                     const c2<T>* __restrict__ tw, int n1, long long nlines, int sign,
                     Prune pr, const double* __restrict__ kcut2,
                     const int2* __restrict__ work, GenPlan P) {
  const int2 wk = work[blockIdx.x];
  Prune p = pr;
  p.kcut2 = kcut2[wk.x];
  fft_z_block_gen<T>(base + wk.x * fstride, load.field(wk.x), tw, n1, nlines, sign, p,
                     P, wk.y);
}

// Strided pass (line `id` starts at (id / inner) outer_stride + id % inner,
// element stride S) with the pruning modes of fft_strided_block.
template <typename T>
__device__ void fft_strided_block_gen(c2<T>* __restrict__ data, const c2<T>* __restrict__ tw,
                                      long long S, long long outer_stride, long long inner,
                                      long long nlines, const Prune& pr, int prune_mode,
                                      const GenPlan& P, long long blk) {
  const int N = P.n, LS = N + 1;
  GenSmem<T> sm(P);
  const long long id0 = blk * P.nl;
  if (id0 >= nlines) return;
  const int nl = int(min((long long)P.nl, nlines - id0));
  if (pr.on && (prune_mode == 1 || prune_mode == 3)) {
    bool any = false;
    for (int l = 0; l < nl && !any; l++) {
      const long long id = id0 + l;
      any = pr.p2(int(prune_mode == 1 ? id / inner : id % inner));
    }
    if (!any) return;
  }
  for (int m = threadIdx.x; m < N; m += FFT_THREADS) sm.W[m] = tw[m];
  // Element (n, l): consecutive threads take consecutive lines.
  for (int idx = threadIdx.x; idx < nl * N; idx += FFT_THREADS) {
    const int n = idx / nl, l = idx % nl;
    const long long id = id0 + l;
    const long long base = (id / inner) * outer_stride + id % inner;
    bool need = true;
    if (pr.on) {
      if (prune_mode == 1) {
        const int i = int(id / inner);
        need = pr.p2(i) && pr.p1(i, n);
      } else if (prune_mode == 2) {
        need = pr.p2(n);
      } else if (prune_mode == 3) {
        need = pr.p1(int(id % inner), n);
      }
    }
    sm.a[l * LS + n] = need ? ld_mesh<T>(data + base + (long long)n * S) : mk<T>(T(0), T(0));
  }
  __syncthreads();
  const c2<T>* R = lds_fft_gen<T>(sm.a, sm.b, sm.W, P, nl);
  for (int idx = threadIdx.x; idx < nl * N; idx += FFT_THREADS) {
    const int n = idx / nl, l = idx % nl;
    const long long id = id0 + l;
    if (pr.on && prune_mode == 1 && !pr.p2(int(id / inner))) continue;
    if (pr.on && prune_mode == 3 && !pr.p2(int(id % inner))) continue;
    const long long base = (id / inner) * outer_stride + id % inner;
    st_mesh<T>(data + base + (long long)n * S, R[l * LS + n]);
  }
}

template <typename T>
__global__ void __launch_bounds__(FFT_THREADS)
fft_pass_strided_gen(c2<T>* __restrict__ data, const c2<T>* __restrict__ tw, long long S,
                     long long outer_stride, long long inner, long long nlines, Prune pr,
                     int prune_mode, GenPlan P) {
  fft_strided_block_gen<T>(data, tw, S, outer_stride, inner, nlines, pr, prune_mode, P,
                           blockIdx.x);
}

template <typename T>
__global__ void __launch_bounds__(FFT_THREADS)
fft_pass_strided_batch_gen(c2<T>* __restrict__ base, long long fstride,
                           const c2<T>* __restrict__ tw, long long S, long long outer_stride,
                           long long inner, long long nlines, Prune pr,
                           const double* __restrict__ kcut2, const int2* __restrict__ work,
                           int prune_mode, GenPlan P) {
  const int2 wk = work[blockIdx.x];
  Prune p = pr;
  p.kcut2 = kcut2[wk.x];
  fft_strided_block_gen<T>(base + wk.x * fstride, tw, S, outer_stride, inner, nlines, p,
                           prune_mode, P, wk.y);
}

// First pass of the x-first order: lines along x (element stride n1 n2,
// line id = j n2 + k) with a loader load(g, i, j, k); blocks with no line
// where P1(k, j) are skipped.
template <typename T, class Loader>
__device__ void fft_x_load_block_gen(c2<T>* __restrict__ data, const Loader& load,
                                     const c2<T>* __restrict__ tw, int n1, int n2,
                                     const Prune& pr, const GenPlan& P, long long blk) {
  const int N = P.n, LS = N + 1;
  GenSmem<T> sm(P);
  const long long S = (long long)n1 * n2;
  const long long id0 = blk * P.nl;
  if (id0 >= S) return;
  const int nl = int(min((long long)P.nl, S - id0));
  if (pr.on) {
    bool any = false;
    for (int l = 0; l < nl && !any; l++) {
      const long long id = id0 + l;
      any = pr.p1(int(id % n2), int(id / n2));
    }
    if (!any) return;
  }
  for (int m = threadIdx.x; m < N; m += FFT_THREADS) sm.W[m] = tw[m];
  for (int idx = threadIdx.x; idx < nl * N; idx += FFT_THREADS) {
    const int n = idx / nl, l = idx % nl;
    const long long id = id0 + l;
    sm.a[l * LS + n] = load(id + (long long)n * S, n, int(id / n2), int(id % n2));
  }
  __syncthreads();
  const c2<T>* R = lds_fft_gen<T>(sm.a, sm.b, sm.W, P, nl);
  for (int idx = threadIdx.x; idx < nl * N; idx += FFT_THREADS) {
    const int n = idx / nl, l = idx % nl;
    st_mesh<T>(data + id0 + l + (long long)n * S, R[l * LS + n]);
  }
}

template <typename T, class BatchLoader>
__global__ void __launch_bounds__(FFT_THREADS)
fft_pass_x_load_batch_gen(c2<T>* __restrict__ base, long long fstride, BatchLoader load,
                          const c2<T>* __restrict__ tw, int n1, int n2, Prune pr,
                          const double* __restrict__ kcut2, const int2* __restrict__ work,
                          GenPlan P) {
  const int2 wk = work[blockIdx.x];
  Prune p = pr;
  p.kcut2 = kcut2[wk.x];
  fft_x_load_block_gen<T>(base + wk.x * fstride, load.field(wk.x), tw, n1, n2, p, P,
                          wk.y);
}

}  // namespace trvg
