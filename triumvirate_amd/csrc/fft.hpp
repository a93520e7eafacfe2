// fft.hpp -- 3-D complex FFT for gfx950: three batched 1-D passes, each a
// radix-4/radix-2 Stockham autosort transform held in LDS (the
// formulation of Govindaraju et al., "High performance discrete Fourier
// transforms on graphics processors", SC 2008).
//
// Sizes: powers of two from 8 to 1024 per axis.  Sign convention as FFTW
// (forward e^{-ikx}, backward e^{+ikx}), unnormalised.
//
// Each pass is in place: a workgroup loads E = 256*EPT complex elements
// (L = E/n lines of one axis), transforms them in LDS and writes them back
// to the same addresses, so no second buffer is needed.  The first pass
// (along z, contiguous lines) takes a "loader" functor that produces each
// input element from one or more source fields; shell filters, window
// deconvolution, scalings and two-point products are fused there instead
// of costing a separate read+write sweep of the mesh.
#pragma once

#include "common.hpp"

namespace trvg {

constexpr int FFT_THREADS = 256;

// Global stores and loads of the passes' mesh data (one place for their
// cache policy; plain accesses measured best: streaming hints gained on
// some passes and lost on others).
template <typename T>
__device__ __forceinline__ void st_mesh(c2<T>* p, c2<T> v) { *p = v; }

template <typename T>
__device__ __forceinline__ c2<T> ld_mesh(const c2<T>* p) { return *p; }

// Elements per thread of a pass (a workgroup holds 256 * EPT elements).
// Eight for both precisions: the FP32 passes are latency-bound, and the
// smaller LDS footprint admits twice the workgroups per CU.
template <typename T> struct FftEpt { static constexpr int v = 8; };

// Workgroups are dispatched to the 8 XCDs round-robin by index; mapping
// index b to logical block (b % 8) * (n / 8) + b / 8 gives each XCD (and
// its L2) a contiguous range of blocks, so neighbouring blocks that share
// cache lines run under one L2.  Identity when n is not a multiple of 8.
__device__ __forceinline__ long long xcd_block(long long b, long long n) {
  if (n % 8 != 0) return b;
  return (b % 8) * (n / 8) + b / 8;
}

// Support pruning for band-limited inverse transforms: input cells exist only
// where |k| < k_hi, so (with kcut2 = k_hi^2 (1 + 1e-9))
//   line (i, j) of the z pass is needed only if kx^2 + ky^2 < kcut2 (P1);
//   slab i of the y pass only if kx^2 < kcut2 (P2).
// Passes skip what is not needed, and later passes read only rows that
// earlier passes wrote, treating the rest as zero.
struct Prune {
  int on;
  double kcut2;
  double dk0, dk1;
  int n0, n1;
  // Explicitly rounded operations: every pass must evaluate the predicates
  // identically (no contraction differences between kernels).
  __device__ __forceinline__ bool p2(int i) const {
    double kx = __dmul_rn(double(shifted(i, n0)), dk0);
    return __dmul_rn(kx, kx) < kcut2;
  }
  __device__ __forceinline__ bool p1(int i, int j) const {
    double kx = __dmul_rn(double(shifted(i, n0)), dk0);
    double ky = __dmul_rn(double(shifted(j, n1)), dk1);
    return __dadd_rn(__dmul_rn(kx, kx), __dmul_rn(ky, ky)) < kcut2;
  }
};

// In-LDS Stockham transform of L lines of length N (N = 2^LOGN) stored as
// buf[line * (N + 1) + n] (one pad element per line against bank
// conflicts).  W[m] = e^{sign 2 pi i m / N} (m < N) in LDS.
template <typename T, int LOGN, int EPT>
__device__ __forceinline__ void lds_fft_lines(c2<T>* buf, const c2<T>* W,
                                              int sign) {
  constexpr int N = 1 << LOGN;
  constexpr int LS = N + 1;  // padded LDS line stride
  constexpr int E = FFT_THREADS * EPT;
  constexpr int NLINES = E / N;
  const int t = threadIdx.x;
  int Ns = 1;
  int rem = LOGN;
#pragma unroll
  for (int stage = 0; stage < (LOGN + 1) / 2; stage++) {
    if (rem >= 2) {
      // radix-4
      constexpr int NB = E / 4;               // butterflies per workgroup
      constexpr int PER = NB / FFT_THREADS;   // = EPT / 4
      c2<T> v[PER > 0 ? PER : 1][4];
      int jj[PER > 0 ? PER : 1], ll[PER > 0 ? PER : 1];
#pragma unroll
      for (int u = 0; u < PER; u++) {
        int b = t + u * FFT_THREADS;
        int line = b / (N / 4);
        int j = b % (N / 4);
        jj[u] = j;
        ll[u] = line;
        int k = j & (Ns - 1);
        const c2<T>* src = buf + line * LS;
#pragma unroll
        for (int q = 0; q < 4; q++) v[u][q] = src[j + q * (N / 4)];
        int tstep = k * (N / (Ns * 4));
        v[u][1] = cmul<T>(v[u][1], W[tstep]);
        v[u][2] = cmul<T>(v[u][2], W[2 * tstep]);
        v[u][3] = cmul<T>(v[u][3], W[3 * tstep]);
        c2<T> t0 = cadd<T>(v[u][0], v[u][2]);
        c2<T> t1 = csub<T>(v[u][0], v[u][2]);
        c2<T> t2 = cadd<T>(v[u][1], v[u][3]);
        c2<T> d = csub<T>(v[u][1], v[u][3]);
        // t3 = sign * i * d
        c2<T> t3 = (sign < 0) ? mk<T>(d.y, -d.x) : mk<T>(-d.y, d.x);
        v[u][0] = cadd<T>(t0, t2);
        v[u][2] = csub<T>(t0, t2);
        v[u][1] = cadd<T>(t1, t3);
        v[u][3] = csub<T>(t1, t3);
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < PER; u++) {
        int j = jj[u];
        int k = j & (Ns - 1);
        int idxD = (j / Ns) * Ns * 4 + k;
        c2<T>* dst = buf + ll[u] * LS;
#pragma unroll
        for (int q = 0; q < 4; q++) dst[idxD + q * Ns] = v[u][q];
      }
      __syncthreads();
      Ns *= 4;
      rem -= 2;
    } else if (rem == 1) {
      constexpr int NB = E / 2;
      constexpr int PER = NB / FFT_THREADS;   // = EPT / 2
      c2<T> v[PER][2];
      int jj[PER], ll[PER];
#pragma unroll
      for (int u = 0; u < PER; u++) {
        int b = t + u * FFT_THREADS;
        int line = b / (N / 2);
        int j = b % (N / 2);
        jj[u] = j;
        ll[u] = line;
        int k = j & (Ns - 1);
        const c2<T>* src = buf + line * LS;
        v[u][0] = src[j];
        v[u][1] = cmul<T>(src[j + N / 2], W[k * (N / (Ns * 2))]);
        c2<T> a = v[u][0], bb = v[u][1];
        v[u][0] = cadd<T>(a, bb);
        v[u][1] = csub<T>(a, bb);
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < PER; u++) {
        int j = jj[u];
        int k = j & (Ns - 1);
        int idxD = (j / Ns) * Ns * 2 + k;
        c2<T>* dst = buf + ll[u] * LS;
        dst[idxD] = v[u][0];
        dst[idxD + Ns] = v[u][1];
      }
      __syncthreads();
      Ns *= 2;
      rem -= 1;
    }
  }
  (void)NLINES;
}

// ---------------------------------------------------------------------------
// Register-resident DFTs for the z pass.
// ---------------------------------------------------------------------------

// cos(pi k / 16) for k in [0, 16] (k > 8 by symmetry).
__host__ __device__ constexpr double cos_pi16(int k) {
  return k > 8   ? -cos_pi16(16 - k)
         : k == 0 ? 1.0
         : k == 1 ? 0.9807852804032304
         : k == 2 ? 0.9238795325112867
         : k == 3 ? 0.8314696123025452
         : k == 4 ? 0.7071067811865476
         : k == 5 ? 0.5555702330196023
         : k == 6 ? 0.38268343236508984
         : k == 7 ? 0.19509032201612833
                  : 0.0;
}

// d * e^{sgn i pi k / 16} for k in [0, 16); k is a constant after
// unrolling, so the trivial factors (1, +-i) cost no multiplications.
template <typename T>
__device__ __forceinline__ c2<T> rot_pi16(c2<T> d, int k, T sgn) {
  if (k == 0) return d;
  if (k == 8) return mk<T>(-sgn * d.y, sgn * d.x);
  const T c = T(cos_pi16(k));
  const T sn = sgn * T(cos_pi16(8 - k > 0 ? 8 - k : k - 8));
  return mk<T>(d.x * c - d.y * sn, d.x * sn + d.y * c);
}

// One decimation-in-frequency radix-2 stage of length LEN over v[0, R),
// then the stages of length LEN / 2 ... 2.
template <typename T, int R, int LEN>
struct DifStages {
  static __device__ __forceinline__ void run(c2<T>* v, T sgn) {
    constexpr int H = LEN / 2;
#pragma unroll
    for (int s = 0; s < R; s += LEN) {
#pragma unroll
      for (int j = 0; j < H; j++) {
        const c2<T> a = v[s + j], b = v[s + j + H];
        v[s + j] = cadd<T>(a, b);
        v[s + j + H] = rot_pi16<T>(csub<T>(a, b), (32 / LEN) * j, sgn);
      }
    }
    DifStages<T, R, H>::run(v, sgn);
  }
};
template <typename T, int R>
struct DifStages<T, R, 1> {
  static __device__ __forceinline__ void run(c2<T>*, T) {}
};

__host__ __device__ constexpr int bitrev(int k, int bits) {
  return bits == 0 ? 0 : ((k & 1) << (bits - 1)) | bitrev(k >> 1, bits - 1);
}
__host__ __device__ constexpr int ilog2c(int r) { return r <= 1 ? 0 : 1 + ilog2c(r / 2); }

// v <- DFT_R(v): X[k] = sum_n v[n] e^{sgn 2 pi i n k / R}, natural order in
// and out (R a power of two <= 32, held in registers).
template <typename T, int R>
__device__ __forceinline__ void dft_reg(c2<T>* v, T sgn) {
  static_assert(R >= 1 && R <= 32 && (R & (R - 1)) == 0, "register DFT size");
  DifStages<T, R, R>::run(v, sgn);
  c2<T> o[R];
#pragma unroll
  for (int k = 0; k < R; k++) o[k] = v[bitrev(k, ilog2c(R))];
#pragma unroll
  for (int k = 0; k < R; k++) v[k] = o[k];
}

// Shape of the register z pass for N = 2^LOGN = A * B: A threads per line,
// B elements per thread (B a multiple of A); a workgroup takes the block of
// every pass, FFT_THREADS * FftEpt<T> elements.
template <int LOGN> struct ZRegShape { static constexpr bool ok = false; };
template <> struct ZRegShape<8> {
  static constexpr bool ok = true;
  static constexpr int A = 16, B = 16;
};
template <> struct ZRegShape<9> {
  static constexpr bool ok = true;
  static constexpr int A = 16, B = 32;
};
template <typename T>
constexpr int zreg_block() { return FFT_THREADS * FftEpt<T>::v; }
template <typename T, int LOGN>
constexpr int zreg_threads() { return zreg_block<T>() / ZRegShape<LOGN>::B; }
template <typename T, int LOGN>
constexpr size_t zreg_smem() {
  return size_t((zreg_block<T>() >> LOGN) * ZRegShape<LOGN>::B *
                    (ZRegShape<LOGN>::A + 1) +
                (1 << LOGN)) * sizeof(c2<T>);
}

// z pass of block `blk` (L = zreg_block / N whole lines) as
// N = A * B: with n = n1 + A n2 and k = k2 + B k1,
//   X[k2 + B k1] = sum_n1 w_A^{n1 k1} w_N^{n1 k2} sum_n2 x[n1 + A n2] w_B^{n2 k2}.
// Thread (line, n1) loads its B elements (lanes of a line read A adjacent
// elements per load), takes the inner DFT in registers, applies w_N^{n1 k2}
// and writes to LDS; after one exchange each thread takes B / A outer DFTs
// of length A in registers and stores A adjacent outputs per lane group.
// One LDS round trip per element instead of one per radix-4 stage.
// Loads of block `blk`: thread (l, q) takes elements k = q + A n2 of line
// l of the block (lanes of a line read A adjacent elements per load).
template <typename T, int LOGN, class Loader>
__device__ __forceinline__ void zreg_load(c2<T>* v, const Loader& load, int n1,
                                          long long blk) {
  constexpr int N = 1 << LOGN;
  constexpr int A = ZRegShape<LOGN>::A, B = ZRegShape<LOGN>::B;
  constexpr int L = zreg_block<T>() / N;
  const int l = threadIdx.x / A, q = threadIdx.x % A;
  const long long line = blk * L + l;
  const int i = int(line / n1), j = int(line % n1);
  const long long g0 = line * N;
#pragma unroll
  for (int n2 = 0; n2 < B; n2++) {
    const int k = q + A * n2;
    v[n2] = load(g0 + k, i, j, k);
  }
}

// Transform of the loaded block and its store: the inner DFT in registers,
// w_N^{n1 k2} and one LDS exchange, the outer DFTs in registers, A adjacent
// outputs per lane group.  W (the N twiddles) is in LDS; the LDS line
// buffer may be reused by the next block of the workgroup.
template <typename T, int LOGN>
__device__ __forceinline__ void zreg_transform_store(c2<T>* v, c2<T>* __restrict__ data,
                                                     const c2<T>* W, c2<T>* buf,
                                                     int sign, long long blk) {
  constexpr int N = 1 << LOGN;
  constexpr int A = ZRegShape<LOGN>::A, B = ZRegShape<LOGN>::B;
  constexpr int L = zreg_block<T>() / N;
  constexpr int AP = A + 1;        // padded stride between k2 rows in LDS
  constexpr int LSTR = B * AP;     // LDS elements per line
  const int l = threadIdx.x / A, q = threadIdx.x % A;
  const T sgn = T(sign);
  dft_reg<T, B>(v, sgn);
  __syncthreads();
  c2<T>* lb = buf + l * LSTR;
  lb[q] = v[0];
#pragma unroll
  for (int k2 = 1; k2 < B; k2++) lb[k2 * AP + q] = cmul<T>(v[k2], W[q * k2]);
  __syncthreads();
  c2<T>* out = data + (blk * L + l) * N;
#pragma unroll
  for (int r = 0; r < B / A; r++) {
    const int k2 = q + A * r;
    c2<T> u[A];
#pragma unroll
    for (int m = 0; m < A; m++) u[m] = lb[k2 * AP + m];
    dft_reg<T, A>(u, sgn);
#pragma unroll
    for (int k1 = 0; k1 < A; k1++) st_mesh<T>(out + k2 + B * k1, u[k1]);
  }
}

// z pass of block `blk` (L = zreg_block / N whole lines) as
// N = A * B: with n = n1 + A n2 and k = k2 + B k1,
//   X[k2 + B k1] = sum_n1 w_A^{n1 k1} w_N^{n1 k2} sum_n2 x[n1 + A n2] w_B^{n2 k2}.
// One LDS round trip per element instead of one per radix-4 stage.
template <typename T, int LOGN, class Loader>
__device__ __forceinline__ void fft_z_block_reg(c2<T>* __restrict__ data,
                                                const Loader& load,
                                                const c2<T>* __restrict__ tw, int n1,
                                                int sign, const Prune& pr, long long blk) {
  constexpr int N = 1 << LOGN;
  constexpr int A = ZRegShape<LOGN>::A, B = ZRegShape<LOGN>::B;
  constexpr int L = zreg_block<T>() / N;
  constexpr int NT = L * A;
  extern __shared__ char smem_raw[];
  c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
  c2<T>* W = buf + L * B * (A + 1);
  const int t = threadIdx.x;
  if (pr.on) {
    // Skip the workgroup if none of its lines is needed (uniform exit).
    bool any = false;
    for (int ll = 0; ll < L && !any; ll++) {
      const long long ln = blk * L + ll;
      any = pr.p1(int(ln / n1), int(ln % n1));
    }
    if (!any) return;
  }
  for (int m = t; m < N; m += NT) W[m] = tw[m];
  c2<T> v[B];
  zreg_load<T, LOGN>(v, load, n1, blk);
  zreg_transform_store<T, LOGN>(v, data, W, buf, sign, blk);
}

template <typename T, int LOGN, class Loader>
__global__ void __launch_bounds__((zreg_threads<T, LOGN>()))
fft_pass_z_reg(c2<T>* __restrict__ data, Loader load, const c2<T>* __restrict__ tw,
               int n1, int sign, Prune pr) {
  fft_z_block_reg<T, LOGN>(data, load, tw, n1, sign, pr, xcd_block(blockIdx.x, gridDim.x));
}

// Batched unpruned register z pass, two work items (field, block) per
// workgroup: the loads of the second are issued before the first is
// transformed, so each wave keeps twice the bytes in flight.
template <typename T, int LOGN, class BatchLoader>
__global__ void __launch_bounds__((zreg_threads<T, LOGN>()))
fft_pass_z_batch_reg(c2<T>* __restrict__ base, long long fstride, BatchLoader load,
                     const c2<T>* __restrict__ tw, int n1, int sign,
                     const int2* __restrict__ work, int nwork) {
  constexpr int N = 1 << LOGN;
  constexpr int A = ZRegShape<LOGN>::A, B = ZRegShape<LOGN>::B;
  constexpr int L = zreg_block<T>() / N;
  constexpr int NT = L * A;
  extern __shared__ char smem_raw[];
  c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
  c2<T>* W = buf + L * B * (A + 1);
  for (int m = threadIdx.x; m < N; m += NT) W[m] = tw[m];
  const long long b0 = 2LL * xcd_block(blockIdx.x, gridDim.x);
  const bool two = b0 + 1 < nwork;   // uniform over the workgroup
  const int2 w0 = work[b0];
  c2<T> v[B], vn[B];
  zreg_load<T, LOGN>(v, load.field(w0.x), n1, w0.y);
  int2 w1 = w0;
  if (two) {
    w1 = work[b0 + 1];
    zreg_load<T, LOGN>(vn, load.field(w1.x), n1, w1.y);
  }
  zreg_transform_store<T, LOGN>(v, base + w0.x * fstride, W, buf, sign, w0.y);
  if (two) zreg_transform_store<T, LOGN>(vn, base + w1.x * fstride, W, buf, sign, w1.y);
}

// First pass along z (contiguous lines), with a fused loader, for block
// `blk` of the mesh at `data`.
// Loader: c2<T> operator()(long long g, int i, int j, int k) const.
// FAST (n1 a multiple of the lines per block, as for every mesh whose z
// size fits NLINES lines per row): the block's lines share i, so indices
// need no division.
template <typename T, int LOGN, int EPT, bool FAST, class Loader>
__device__ __forceinline__ void fft_z_block(c2<T>* __restrict__ data,
                                            const Loader& load,
                                            const c2<T>* __restrict__ tw,
                                            int n1, int sign, const Prune& pr,
                                            long long blk) {
  constexpr int N = 1 << LOGN;
  constexpr int E = FFT_THREADS * EPT;
  constexpr int NLINES = E / N;
  extern __shared__ char smem_raw[];
  constexpr int LS = N + 1;
  c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
  c2<T>* W = buf + NLINES * LS;
  const int t = threadIdx.x;
  const long long line0 = blk * NLINES;
  const long long base = line0 * N;
  const int i0 = int(line0 / n1), j0 = int(line0 % n1);
  if (pr.on) {
    // Skip the workgroup if none of its lines is needed (uniform exit).
    bool any = false;
    for (int l = 0; l < NLINES && !any; l++) {
      if (FAST) {
        any = pr.p1(i0, j0 + l);
      } else {
        const long long line = line0 + l;
        any = pr.p1(int(line / n1), int(line % n1));
      }
    }
    if (!any) return;
  }
  for (int m = t; m < N; m += FFT_THREADS) W[m] = tw[m];
#pragma unroll
  for (int e = 0; e < EPT; e++) {
    const int idx = t + e * FFT_THREADS;
    const int k = idx % N;
    int i, j;
    if (FAST) {
      i = i0;
      j = j0 + idx / N;
    } else {
      const long long line = line0 + idx / N;
      i = int(line / n1);
      j = int(line % n1);
    }
    buf[(idx / N) * LS + k] = load(base + idx, i, j, k);
  }
  __syncthreads();
  lds_fft_lines<T, LOGN, EPT>(buf, W, sign);
  c2<T>* out = data + base;
#pragma unroll
  for (int e = 0; e < EPT; e++) {
    const int idx = t + e * FFT_THREADS;
    st_mesh<T>(out + idx, buf[(idx / N) * LS + idx % N]);
  }
}

template <typename T, int LOGN, int EPT, bool FAST, class Loader>
__global__ void __launch_bounds__(FFT_THREADS)
fft_pass_z(c2<T>* __restrict__ data, Loader load, const c2<T>* __restrict__ tw,
           int n1, int sign, Prune pr) {
  fft_z_block<T, LOGN, EPT, FAST>(data, load, tw, n1, sign, pr,
                                  xcd_block(blockIdx.x, gridDim.x));
}

// Batched first pass over several fields (a slab of meshes `fstride` apart):
// work[b] = (field, block) lists only the blocks with needed lines, so
// blocks outside every band cost nothing.  BatchLoader::field(q) returns
// the loader of field q; kcut2[q] is that field's pruning radius.
template <typename T, int LOGN, int EPT, bool FAST, class BatchLoader>
__global__ void __launch_bounds__(FFT_THREADS)
fft_pass_z_batch(c2<T>* __restrict__ base, long long fstride, BatchLoader load,
                 const c2<T>* __restrict__ tw, int n1, int sign, Prune pr,
                 const double* __restrict__ kcut2, const int2* __restrict__ work) {
  const int2 wk = work[xcd_block(blockIdx.x, gridDim.x)];
  Prune p = pr;
  p.kcut2 = kcut2[wk.x];
  fft_z_block<T, LOGN, EPT, FAST>(base + wk.x * fstride, load.field(wk.x), tw,
                                  n1, sign, p, wk.y);
}

// Strided pass: lines of length N with element stride S.  Line `id` starts
// at (id / inner) * outer_stride + id % inner, so consecutive lines are
// adjacent in memory within a slice:
//   y pass: S = n2, inner = n2, outer_stride = n1*n2;
//   x pass: S = n1*n2, inner = n1*n2, outer_stride = 0.
// Each workgroup takes L = E/N consecutive lines.
//
// prune_mode (with pr.on): 1 = y pass (slab i = id / inner must satisfy P2,
// row j read only where P1(i, j)); 2 = x pass (row i read only where P2(i));
// 3 = y pass of the x-first order (Prune over the (z, y) axes: line k =
// id % inner written only where P2(k), row j read only where P1(k, j)).
//
// FAST (inner a multiple of L): the block's lines lie in one slice, so the
// slice's pruning flags go to LDS once per block and each thread steps a
// pointer by a constant per element (thread t always owns line t % L).
template <typename T, int LOGN, int EPT, bool FAST>
__device__ __forceinline__ void fft_strided_block(
    c2<T>* __restrict__ data, const c2<T>* __restrict__ tw, long long S,
    long long outer_stride, long long inner, int sign, const Prune& pr,
    int prune_mode, long long blk) {
  constexpr int N = 1 << LOGN;
  constexpr int E = FFT_THREADS * EPT;
  constexpr int NLINES = E / N;
  constexpr int NSTEP = FFT_THREADS / NLINES;   // rows advanced per element
  extern __shared__ char smem_raw[];
  constexpr int LS = N + 1;
  c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
  c2<T>* W = buf + NLINES * LS;
  unsigned char* row_in = reinterpret_cast<unsigned char*>(W + N);   // [N]
  const int t = threadIdx.x;
  const long long id0 = blk * NLINES;
  if (FAST) {
    const long long slice0 = id0 / inner;
    const bool m3 = pr.on && prune_mode == 3;
    if (pr.on && prune_mode == 1 && !pr.p2(int(slice0))) return;
    if (m3) {
      bool any = false;
      for (int l = 0; l < NLINES && !any; l++) any = pr.p2(int(id0 % inner) + l);
      if (!any) return;
    }
    for (int m = t; m < N; m += FFT_THREADS) {
      W[m] = tw[m];
      bool need = true;
      if (pr.on) {
        if (prune_mode == 1) need = pr.p1(int(slice0), m);
        else if (prune_mode == 2) need = pr.p2(m);
      }
      row_in[m] = need;
    }
    __syncthreads();
    const int l = t % NLINES, n0 = t / NLINES;
    const int kl = int(id0 % inner) + l;   // this thread's line (prune mode 3)
    c2<T>* line = data + (slice0 * outer_stride + id0 % inner + l);
    const long long step = (long long)NSTEP * S;
    c2<T>* p = line + (long long)n0 * S;
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      const int n = n0 + e * NSTEP;
      const bool need = m3 ? pr.p1(kl, n) : bool(row_in[n]);
      buf[l * LS + n] = need ? ld_mesh<T>(p) : mk<T>(T(0), T(0));
      p += step;
    }
    __syncthreads();
    lds_fft_lines<T, LOGN, EPT>(buf, W, sign);
    if (m3 && !pr.p2(kl)) return;
    p = line + (long long)n0 * S;
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      st_mesh<T>(p, buf[l * LS + n0 + e * NSTEP]);
      p += step;
    }
    return;
  }
  if (pr.on && prune_mode == 1) {
    bool any = false;
    for (int l = 0; l < NLINES && !any; l++) any = pr.p2(int((id0 + l) / inner));
    if (!any) return;
  }
  if (pr.on && prune_mode == 3) {
    bool any = false;
    for (int l = 0; l < NLINES && !any; l++) any = pr.p2(int((id0 + l) % inner));
    if (!any) return;
  }
  for (int m = t; m < N; m += FFT_THREADS) W[m] = tw[m];
  // Global element (n, l): line id0 + l, position n; LDS line-major [l][n].
#pragma unroll
  for (int e = 0; e < EPT; e++) {
    const int idx = t + e * FFT_THREADS;
    const int n = idx / NLINES, l = idx % NLINES;
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
    buf[l * LS + n] = need ? ld_mesh<T>(data + base + (long long)n * S) : mk<T>(T(0), T(0));
  }
  __syncthreads();
  lds_fft_lines<T, LOGN, EPT>(buf, W, sign);
#pragma unroll
  for (int e = 0; e < EPT; e++) {
    const int idx = t + e * FFT_THREADS;
    const int n = idx / NLINES, l = idx % NLINES;
    const long long id = id0 + l;
    const long long base = (id / inner) * outer_stride + id % inner;
    if (pr.on && prune_mode == 1 && !pr.p2(int(id / inner))) continue;
    if (pr.on && prune_mode == 3 && !pr.p2(int(id % inner))) continue;
    st_mesh<T>(data + base + (long long)n * S, buf[l * LS + n]);
  }
}

template <typename T, int LOGN, int EPT, bool FAST>
__global__ void __launch_bounds__(FFT_THREADS)
fft_pass_strided(c2<T>* __restrict__ data, const c2<T>* __restrict__ tw,
                 long long S, long long outer_stride, long long inner,
                 int sign, Prune pr, int prune_mode) {
  fft_strided_block<T, LOGN, EPT, FAST>(data, tw, S, outer_stride, inner, sign, pr,
                                        prune_mode, xcd_block(blockIdx.x, gridDim.x));
}

// Batched strided pass over a slab of fields, work[b] = (field, block).
template <typename T, int LOGN, int EPT, bool FAST>
__global__ void __launch_bounds__(FFT_THREADS)
fft_pass_strided_batch(c2<T>* __restrict__ base, long long fstride,
                       const c2<T>* __restrict__ tw, long long S,
                       long long outer_stride, long long inner, int sign,
                       Prune pr, const double* __restrict__ kcut2,
                       const int2* __restrict__ work, int prune_mode) {
  const int2 wk = work[xcd_block(blockIdx.x, gridDim.x)];
  Prune p = pr;
  p.kcut2 = kcut2[wk.x];
  fft_strided_block<T, LOGN, EPT, FAST>(base + wk.x * fstride, tw, S, outer_stride,
                                        inner, sign, p, prune_mode, wk.y);
}

// First pass of the x-first order: lines along x (length N = n0, element
// stride S = n1 n2), line id = j n2 + k, with a fused loader
// load(g, i, j, k).  Prune over the (z, y) axes: a block runs only if one of
// its lines has P1(k, j); it writes all its lines (lines outside the band
// load as zeros).  FAST (n2 a multiple of the lines per block): the block's
// lines share j.
template <typename T, int LOGN, int EPT, bool FAST, class Loader>
__device__ __forceinline__ void fft_x_load_block(c2<T>* __restrict__ data,
                                                 const Loader& load,
                                                 const c2<T>* __restrict__ tw,
                                                 int n1, int n2, int sign,
                                                 const Prune& pr, long long blk) {
  constexpr int N = 1 << LOGN;
  constexpr int E = FFT_THREADS * EPT;
  constexpr int NLINES = E / N;
  constexpr int NSTEP = FFT_THREADS / NLINES;
  extern __shared__ char smem_raw[];
  constexpr int LS = N + 1;
  c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
  c2<T>* W = buf + NLINES * LS;
  const int t = threadIdx.x;
  const long long S = (long long)n1 * n2;
  const long long id0 = blk * NLINES;
  if (pr.on) {
    bool any = false;
    for (int l = 0; l < NLINES && !any; l++) {
      const long long id = id0 + l;
      any = pr.p1(int(id % n2), int(id / n2));
    }
    if (!any) return;
  }
  for (int m = t; m < N; m += FFT_THREADS) W[m] = tw[m];
  if (FAST) {
    const int l = t % NLINES, r0 = t / NLINES;
    const int j = int(id0 / n2), k = int(id0 % n2) + l;
    const long long g0 = id0 + l;
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      const int n = r0 + e * NSTEP;
      buf[l * LS + n] = load(g0 + (long long)n * S, n, j, k);
    }
  } else {
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      const int idx = t + e * FFT_THREADS;
      const int n = idx / NLINES, l = idx % NLINES;
      const long long id = id0 + l;
      buf[l * LS + n] = load(id + (long long)n * S, n, int(id / n2), int(id % n2));
    }
  }
  __syncthreads();
  lds_fft_lines<T, LOGN, EPT>(buf, W, sign);
  if (FAST) {
    const int l = t % NLINES, r0 = t / NLINES;
    c2<T>* p = data + id0 + l + (long long)r0 * S;
    const long long step = (long long)NSTEP * S;
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      st_mesh<T>(p, buf[l * LS + r0 + e * NSTEP]);
      p += step;
    }
  } else {
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      const int idx = t + e * FFT_THREADS;
      const int n = idx / NLINES, l = idx % NLINES;
      st_mesh<T>(data + id0 + l + (long long)n * S, buf[l * LS + n]);
    }
  }
}

// Batched x-first loader pass, work[b] = (field, block).
template <typename T, int LOGN, int EPT, bool FAST, class BatchLoader>
__global__ void __launch_bounds__(FFT_THREADS)
fft_pass_x_load_batch(c2<T>* __restrict__ base, long long fstride, BatchLoader load,
                      const c2<T>* __restrict__ tw, int n1, int n2, int sign, Prune pr,
                      const double* __restrict__ kcut2, const int2* __restrict__ work) {
  const int2 wk = work[xcd_block(blockIdx.x, gridDim.x)];
  Prune p = pr;
  p.kcut2 = kcut2[wk.x];
  fft_x_load_block<T, LOGN, EPT, FAST>(base + wk.x * fstride, load.field(wk.x), tw, n1,
                                       n2, sign, p, wk.y);
}

// Identity loader (plain in-place transform).
template <typename T>
struct LoadSame {
  const c2<T>* src;
  __device__ c2<T> operator()(long long g, int, int, int) const {
    return ld_mesh<T>(src + g);
  }
};

}  // namespace trvg
