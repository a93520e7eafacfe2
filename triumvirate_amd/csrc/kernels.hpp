// kernels.hpp -- mesh kernels of the three-point engine (gfx950).
//
// All bin sums, normalisation sums and shot-noise amplitudes accumulate in
// FP64; T (double or float) is only the storage/arithmetic type of mesh
// values.  Integer counts (mode / pair counts) are exact.
#pragma once

#include "common.hpp"
#include "fft.hpp"

namespace trvg {

constexpr int RED_THREADS = 256;

// ---------------------------------------------------------------------------
// Particles
// ---------------------------------------------------------------------------

// Device catalogue: SoA FP64 columns.  w = ws * wc.  los may be null.
struct DevParticles {
  const double* x;
  const double* y;
  const double* z;
  const double* w;
  const double* los;   // 3 per particle (observer-frame unit vectors)
  long long n;
};

enum WeightMode {
  WM_UNIT = 0,       // 1
  WM_W = 1,          // w
  WM_YLM_W = 2,      // w * Y(los)                    (field.cpp:1246-1323)
  WM_YLMC_W2 = 3,    // w^2 * conj(Y(los))            (field.cpp:1364-1447)
};

__device__ __forceinline__ double2 particle_weight(const DevParticles& p,
                                                   long long i, int mode,
                                                   int ell, int m) {
  if (mode == WM_UNIT) return dmk(1., 0.);
  double w = p.w[i];
  if (mode == WM_W) return dmk(w, 0.);
  double2 y = ylm_reduced_dev(ell, m, p.los[3 * i], p.los[3 * i + 1],
                              p.los[3 * i + 2]);
  if (mode == WM_YLM_W) return dmk(y.x * w, y.y * w);
  double w2 = w * w;
  return dmk(y.x * w2, -y.y * w2);
}

// Mesh assignment (field.cpp:569-1112): window weights and periodic index
// rules per scheme; cell value += w * W_x W_y W_z / V_cell (times `scale`,
// e.g. -alpha for randoms).  `shadow` selects the half-cell-shifted pass
// used for interlacing.  Hardware FP64/FP32 atomics on the coarse-grained
// mesh.
template <typename T, int ORDER>
__global__ void __launch_bounds__(256)
k_assign(c2<T>* __restrict__ mesh, DevParticles p, Geom g, int mode, int ell,
         int m, double scale, int cplx_w, int shadow) {
  const double inv_vol_cell = 1 / g.vol_cell;
  for (long long pid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       pid < p.n; pid += (long long)gridDim.x * blockDim.x) {
    const double pos[3] = {p.x[pid], p.y[pid], p.z[pid]};
    int ijk[ORDER][3];
    double win[ORDER][3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const int ng = g.n[a];
      double loc = __ddiv_rn(__dmul_rn(double(ng), pos[a]), g.L[a]);
      if (shadow) {
        loc = __dadd_rn(loc, 0.5);
        if (loc > ng) loc = __dadd_rn(loc, -double(ng));
      }
      int idx = int(loc);
      double s = __dadd_rn(loc, -double(idx));
      if (ORDER == 1) {
        if (s >= 0.5) idx = (idx == ng - 1) ? 0 : idx + 1;
        ijk[0][a] = idx;
        win[0][a] = 1.;
      } else if (ORDER == 2) {
        ijk[0][a] = idx;
        ijk[1][a] = (idx == ng - 1) ? 0 : idx + 1;
        win[0][a] = 1. - s;
        win[1][a] = s;
      } else if (ORDER == 3) {
        if (s < 0.5) {
          ijk[0][a] = (idx == 0) ? ng - 1 : idx - 1;
          ijk[1][a] = idx;
          ijk[2][a] = (idx == ng - 1) ? 0 : idx + 1;
          win[0][a] = 0.5 * (0.5 - s) * (0.5 - s);
          win[1][a] = 0.75 - s * s;
          win[2][a] = 0.5 * (0.5 + s) * (0.5 + s);
        } else {
          ijk[0][a] = idx;
          ijk[1][a] = (idx == ng - 1) ? 0 : idx + 1;
          ijk[2][a] = (ijk[1][a] == ng - 1) ? 0 : ijk[1][a] + 1;
          double sp = 1. - s;
          win[0][a] = 0.5 * (0.5 + sp) * (0.5 + sp);
          win[1][a] = 0.75 - sp * sp;
          win[2][a] = 0.5 * (0.5 - sp) * (0.5 - sp);
        }
      } else {
        ijk[0][a] = (idx == 0) ? ng - 1 : idx - 1;
        ijk[1][a] = idx;
        ijk[2][a] = (idx == ng - 1) ? 0 : idx + 1;
        ijk[3][a] = (ijk[2][a] == ng - 1) ? 0 : ijk[2][a] + 1;
        double sm = 1. - s;
        win[0][a] = 1. / 6 * sm * sm * sm;
        win[1][a] = 1. / 6 * (4. - 6. * s * s + 3. * s * s * s);
        win[2][a] = 1. / 6 * (4. - 6. * sm * sm + 3. * sm * sm * sm);
        win[3][a] = 1. / 6 * s * s * s;
      }
    }
    double2 wt = particle_weight(p, pid, mode, ell, m);
    double wre = inv_vol_cell * wt.x * scale;
    double wim = inv_vol_cell * wt.y * scale;
#pragma unroll
    for (int il = 0; il < ORDER; il++) {
#pragma unroll
      for (int jl = 0; jl < ORDER; jl++) {
        const long long row =
            ((long long)ijk[il][0] * g.n[1] + ijk[jl][1]) * g.n[2];
        const double wij = win[il][0] * win[jl][1];
#pragma unroll
        for (int kl = 0; kl < ORDER; kl++) {
          const long long gid = row + ijk[kl][2];
          // Same guard as the reference: only the flattened index is
          // checked (a particle at x = L lands in a neighbouring row).
          if (gid < 0 || gid >= g.nmesh) continue;
          const double wv = wij * win[kl][2];
          T* cell = reinterpret_cast<T*>(mesh + gid);
          unsafeAtomicAdd(cell, T(wre * wv));
          if (cplx_w) unsafeAtomicAdd(cell + 1, T(wim * wv));
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Fused first-pass loaders (see fft.hpp)
// ---------------------------------------------------------------------------

// (src + shift) * scale, shift real (MeshField::fourier_transform with the
// vol_cell factor, field.cpp:1496-1509; mean subtraction of
// compute_unweighted_field_fluctuations_insitu, field.cpp:1225-1244).
template <typename T>
struct LoadScaleShift {
  const c2<T>* src;
  T shift, scale;
  __device__ c2<T> operator()(long long g, int, int, int) const {
    c2<T> v = src[g];
    return mk<T>((v.x + shift) * scale, v.y * scale);
  }
};

// src / W(k) * scale (apply_assignment_compensation then the 1/V of
// inv_fourier_transform, field.cpp:1657-1720, 1764-1785).
template <typename T, typename S = T>
struct LoadDeconv {
  const c2<S>* src;
  AxisTables tab;
  double scale;
  __device__ c2<T> operator()(long long g, int i, int j, int k) const {
    c2<S> v = src[g];
    double f = scale / window_at(tab, i, j, k);
    return mk<T>(T(double(v.x) * f), T(double(v.y) * f));
  }
};

// Shell filter: y_lm(k^) src(k) / W(k) for k_lo <= |k| < k_hi, else 0
// (inv_fourier_transform_ylm_wgtd_field_band_limited, field.cpp:1792-1906;
// the 1/n_modes factor is applied to the pair sums instead).  The shell
// test runs on |k|^2 against ShellBand thresholds, which selects exactly
// the modes whose correctly rounded |k| passes the reference's test.
template <typename T, typename S = T>
struct LoadShell {
  const c2<S>* src;
  AxisTables tab;
  Geom geo;
  ShellBand band;
  int ell, m;
  __device__ c2<T> operator()(long long g, int i, int j, int k) const {
    double kx = grid_coord(i, geo.n[0], geo.dk[0]);
    double ky = grid_coord(j, geo.n[1], geo.dk[1]);
    double kz = grid_coord(k, geo.n[2], geo.dk[2]);
    if (!band.contains(mag2_exact(kx, ky, kz))) return mk<T>(T(0), T(0));
    c2<S> v = src[g];
    double iw = 1. / window_at(tab, i, j, k);
    double2 fk = dmk(double(v.x) * iw, double(v.y) * iw);
    double2 y = ylm_reduced_dev(ell, m, kx, ky, kz);
    double2 r = dmul(y, fk);
    return mk<T>(T(r.x), T(r.y));
  }
};

// Two real shells in one transform: for a real source field, even l and
// m = 0 the shell field of each bin is real, so the transform of
// S_0(k) + i S_1(k) returns F_0(x) + i F_1(x).  The bins are disjoint, so
// each mode belongs to at most one of the two bands.
template <typename T, typename S = T>
struct LoadShellPair {
  const c2<S>* src;
  AxisTables tab;
  Geom geo;
  ShellBand band0, band1;
  int ell, m;
  __device__ c2<T> operator()(long long g, int i, int j, int k) const {
    double kx = grid_coord(i, geo.n[0], geo.dk[0]);
    double ky = grid_coord(j, geo.n[1], geo.dk[1]);
    double kz = grid_coord(k, geo.n[2], geo.dk[2]);
    const double k2 = mag2_exact(kx, ky, kz);
    const bool in0 = band0.contains(k2);
    if (!in0 && !band1.contains(k2)) return mk<T>(T(0), T(0));
    c2<S> v = src[g];
    double iw = 1. / window_at(tab, i, j, k);
    double2 fk = dmk(double(v.x) * iw, double(v.y) * iw);
    double2 r = dmul(ylm_reduced_dev(ell, m, kx, ky, kz), fk);
    return in0 ? mk<T>(T(r.x), T(r.y)) : mk<T>(T(-r.y), T(r.x));
  }
};

template <typename T, typename S = T>
struct LoadShellPairBatch {
  const c2<S>* src;
  AxisTables tab;
  Geom geo;
  const ShellBand* bands;   // two per packed field
  int ell, m;
  __device__ LoadShellPair<T, S> field(int p) const {
    return LoadShellPair<T, S>{src, tab, geo, bands[2 * p], bands[2 * p + 1], ell, m};
  }
};

// LoadShell for field q of a batched slab transform (band of bin q).
template <typename T, typename S = T>
struct LoadShellBatch {
  const c2<S>* src;
  AxisTables tab;
  Geom geo;
  const ShellBand* bands;
  int ell, m;
  __device__ LoadShell<T, S> field(int q) const {
    return LoadShell<T, S>{src, tab, geo, bands[q], ell, m};
  }
};

// j_l(|k| r) y_lm(k^) src(k) / W(k) / V
// (inv_fourier_transform_sjl_ylm_wgtd_field, field.cpp:1908-2010).
template <typename T, typename S = T>
struct LoadSjl {
  const c2<S>* src;
  AxisTables tab;
  Geom geo;
  SplineDev sj;
  double r;
  int ell, m;
  __device__ c2<T> operator()(long long g, int i, int j, int k) const {
    double kx = grid_coord(i, geo.n[0], geo.dk[0]);
    double ky = grid_coord(j, geo.n[1], geo.dk[1]);
    double kz = grid_coord(k, geo.n[2], geo.dk[2]);
    double km = mag3_exact(kx, ky, kz);
    c2<S> v = src[g];
    double iw = 1. / window_at(tab, i, j, k);
    double2 fk = dmk(double(v.x) * iw, double(v.y) * iw);
    double2 y = ylm_reduced_dev(ell, m, kx, ky, kz);
    double2 yf = dmul(y, fk);
    double jv = sj.eval(km * r);
    return mk<T>(T(jv * yf.x / geo.vol), T(jv * yf.y / geo.vol));
  }
};

// (f_a f_b^* / C1 - S C1 / C1) / V  (field.cpp:3007-3025, 3271-3289).
template <typename T>
struct LoadTwopt {
  const c2<T>* fa;
  const c2<T>* fb;
  AxisTables tab;
  double2 sbar;
  double inv_vol;
  __device__ c2<T> operator()(long long g, int i, int j, int k) const {
    c2<T> a = fa[g], b = fb[g];
    double2 pk = dmul(dmk(a.x, a.y), dmk(b.x, -double(b.y)));
    double c1 = alias_at(tab, i, j, k);
    double w = tab.interlace ? window_at(tab, i, j, k) * window_at(tab, i, j, k)
                             : c1;
    double2 sn = dmk(sbar.x * c1 / w, sbar.y * c1 / w);
    pk = dmk(pk.x / w - sn.x, pk.y / w - sn.y);
    return mk<T>(T(pk.x * inv_vol), T(pk.y * inv_vol));
  }
};

// ---------------------------------------------------------------------------
// Bin location
// ---------------------------------------------------------------------------

// Bin b with edges[b] <= v < edges[b+1], or -1.  Edges ascending.
__device__ __forceinline__ int find_bin(const double* edges, int nb, double v) {
  if (!(edges[0] <= v && v < edges[nb])) return -1;
  int lo = 0, hi = nb;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (edges[mid] <= v) lo = mid; else hi = mid;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// Shell statistics: n_modes and sum |k| per bin under the exact rule
// k_lo <= |k| < k_hi (field.cpp:1815-1830).  Per-block partial sums.
// ---------------------------------------------------------------------------

static __global__ void __launch_bounds__(RED_THREADS)
k_shell_stats(Geom g, const double* __restrict__ edges, int nb,
              unsigned long long* __restrict__ counts,
              double* __restrict__ ksum_part) {
  extern __shared__ double sh_ks[];
  unsigned int* sh_cnt = reinterpret_cast<unsigned int*>(sh_ks + nb);
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    sh_ks[b] = 0.;
    sh_cnt[b] = 0u;
  }
  __syncthreads();
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       c < g.nmesh; c += (long long)gridDim.x * blockDim.x) {
    int k = int(c % g.n[2]);
    long long ij = c / g.n[2];
    int j = int(ij % g.n[1]), i = int(ij / g.n[1]);
    double km = mag3_exact(grid_coord(i, g.n[0], g.dk[0]),
                           grid_coord(j, g.n[1], g.dk[1]),
                           grid_coord(k, g.n[2], g.dk[2]));
    int b = find_bin(edges, nb, km);
    if (b >= 0) {
      atomicAdd(&sh_cnt[b], 1u);
      atomicAdd(&sh_ks[b], km);
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    atomicAdd(&counts[b], (unsigned long long)sh_cnt[b]);
    ksum_part[(long long)blockIdx.x * nb + b] = sh_ks[b];
  }
}

// Sum per-block partials in a fixed order (deterministic).
// Column sums of partials [nrows][width]: a block takes RC_COLS adjacent
// columns (contiguous loads) with RC_ROWS row lanes, then a fixed-shape LDS
// tree; the summation order depends only on (nrows, width).
constexpr int RC_COLS = 4;
constexpr int RC_ROWS = 64;

__device__ __forceinline__ double vzero(double) { return 0.; }
__device__ __forceinline__ double2 vzero(double2) { return make_double2(0., 0.); }
__device__ __forceinline__ double vadd(double a, double b) { return a + b; }
__device__ __forceinline__ double2 vadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}

template <typename V>
__device__ __forceinline__ void k_reduce_cols_body(const V* __restrict__ part, int nrows, int width, V* __restrict__ out) {
  __shared__ V red[RC_ROWS][RC_COLS];
  const int c = threadIdx.x % RC_COLS, r = threadIdx.x / RC_COLS;
  const int col = blockIdx.x * RC_COLS + c;
  V s = vzero(V());
  if (col < width)
    for (int b = r; b < nrows; b += RC_ROWS) s = vadd(s, part[(long long)b * width + col]);
  red[r][c] = s;
  __syncthreads();
  for (int h = RC_ROWS / 2; h > 0; h >>= 1) {
    if (r < h) red[r][c] = vadd(red[r][c], red[r + h][c]);
    __syncthreads();
  }
  if (r == 0 && col < width) out[col] = red[0][c];
}

template <typename V>
static __global__ void __launch_bounds__(RC_COLS * RC_ROWS)
k_reduce_cols(const V* __restrict__ part, int nrows, int width, V* __restrict__ out) {
  k_reduce_cols_body<V>(part, nrows, width, out);
}

// ---------------------------------------------------------------------------
// Two-point statistics in Fourier space with the reference's two-level
// binning (field.cpp:2511-2703): fine bin ik = int(|k| / 1e-5) (< 1e6);
// the fine bin enters coarse bin b when edges[b] <= ik * 1e-5 < edges[b+1].
// Accumulates n_modes, sum(f_a f_b^*/C1 * y) and sum(S C1/C1 * y) per bin.
// Output per block: [nb] counts (global atomics) and [nb][4] doubles.
// ---------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ void k_twopt_fourier_body(const c2<T>* __restrict__ fa, const c2<T>* __restrict__ fb,
                Geom g, AxisTables tab, double2 sbar, int ell, int m,
                const double* __restrict__ edges, int nb,
                unsigned long long* __restrict__ counts,
                double* __restrict__ part) {
  extern __shared__ double sh2[];   // [nb][4] then counts
  unsigned int* sh_cnt = reinterpret_cast<unsigned int*>(sh2 + 4 * nb);
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    sh2[4 * b] = sh2[4 * b + 1] = sh2[4 * b + 2] = sh2[4 * b + 3] = 0.;
    sh_cnt[b] = 0u;
  }
  __syncthreads();
  const int n_sample = 1000000;
  const double dk_sample = 1.e-5;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       c < g.nmesh; c += (long long)gridDim.x * blockDim.x) {
    int k = int(c % g.n[2]);
    long long ij = c / g.n[2];
    int j = int(ij % g.n[1]), i = int(ij / g.n[1]);
    double kx = grid_coord(i, g.n[0], g.dk[0]);
    double ky = grid_coord(j, g.n[1], g.dk[1]);
    double kz = grid_coord(k, g.n[2], g.dk[2]);
    double km = mag3_exact(kx, ky, kz);
    int ik = int(__ddiv_rn(km, dk_sample));
    if (ik < 0 || ik >= n_sample) continue;
    int b = find_bin(edges, nb, __dmul_rn(double(ik), dk_sample));
    if (b < 0) continue;
    c2<T> a = fa[c], bb = fb[c];
    double2 pk = dmul(dmk(a.x, a.y), dmk(bb.x, -double(bb.y)));
    double c1 = alias_at(tab, i, j, k);
    double w = tab.interlace ? window_at(tab, i, j, k) * window_at(tab, i, j, k)
                             : c1;
    double2 sn = dmk(sbar.x * c1 / w, sbar.y * c1 / w);
    pk = dmk(pk.x / w, pk.y / w);
    double2 y = ylm_reduced_dev(ell, m, kx, ky, kz);
    pk = dmul(pk, y);
    sn = dmul(sn, y);
    atomicAdd(&sh_cnt[b], 1u);
    atomicAdd(&sh2[4 * b], pk.x);
    atomicAdd(&sh2[4 * b + 1], pk.y);
    atomicAdd(&sh2[4 * b + 2], sn.x);
    atomicAdd(&sh2[4 * b + 3], sn.y);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    atomicAdd(&counts[b], (unsigned long long)sh_cnt[b]);
    for (int q = 0; q < 4; q++)
      part[((long long)blockIdx.x * nb + b) * 4 + q] = sh2[4 * b + q];
  }
}

template <typename T>
static __global__ void __launch_bounds__(RED_THREADS)
k_twopt_fourier(const c2<T>* __restrict__ fa, const c2<T>* __restrict__ fb,
                Geom g, AxisTables tab, double2 sbar, int ell, int m,
                const double* __restrict__ edges, int nb,
                unsigned long long* __restrict__ counts,
                double* __restrict__ part) {
  k_twopt_fourier_body<T>(fa, fb, g, tab, sbar, ell, m, edges, nb, counts, part);
}

// ---------------------------------------------------------------------------
// Three-point correlation shot noise: fine bins of width 1 (< 1e5)
// (field.cpp:3097-3170): n_pairs, sum r, sum t(x) y_a(x^) y_b(x^).
// ---------------------------------------------------------------------------

template <typename T>
static __global__ void __launch_bounds__(RED_THREADS)
k_3pcf_shot_bins(const c2<T>* __restrict__ t, Geom g, int la, int ma, int lb,
                 int mb, const double* __restrict__ edges, int nb,
                 unsigned long long* __restrict__ counts,
                 double* __restrict__ part) {
  extern __shared__ double sh3[];   // [nb][3] then counts
  unsigned int* sh_cnt = reinterpret_cast<unsigned int*>(sh3 + 3 * nb);
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    sh3[3 * b] = sh3[3 * b + 1] = sh3[3 * b + 2] = 0.;
    sh_cnt[b] = 0u;
  }
  __syncthreads();
  const int n_sample = 100000;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       c < g.nmesh; c += (long long)gridDim.x * blockDim.x) {
    int k = int(c % g.n[2]);
    long long ij = c / g.n[2];
    int j = int(ij % g.n[1]), i = int(ij / g.n[1]);
    double rx = grid_coord(i, g.n[0], g.dr[0]);
    double ry = grid_coord(j, g.n[1], g.dr[1]);
    double rz = grid_coord(k, g.n[2], g.dr[2]);
    double rm = mag3_exact(rx, ry, rz);
    int ir = int(rm);   // r / dr_sample with dr_sample = 1
    if (ir < 0 || ir >= n_sample) continue;
    int b = find_bin(edges, nb, double(ir));
    if (b < 0) continue;
    c2<T> tv = t[c];
    double2 ya = ylm_reduced_dev(la, ma, rx, ry, rz);
    double2 yb = ylm_reduced_dev(lb, mb, rx, ry, rz);
    double2 v = dmul(dmk(tv.x, tv.y), dmul(ya, yb));
    atomicAdd(&sh_cnt[b], 1u);
    atomicAdd(&sh3[3 * b], rm);
    atomicAdd(&sh3[3 * b + 1], v.x);
    atomicAdd(&sh3[3 * b + 2], v.y);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    atomicAdd(&counts[b], (unsigned long long)sh_cnt[b]);
    for (int q = 0; q < 3; q++)
      part[((long long)blockIdx.x * nb + b) * 3 + q] = sh3[3 * b + q];
  }
}

// ---------------------------------------------------------------------------
// Bispectrum shot noise S_ij_k (field.cpp:3198-3396)
//   S(k_a, k_b) = V_cell sum_x j_a(k_a |x|) j_b(k_b |x|) y_a(x^) y_b(x^) t(x).
// With equal cell sizes on the three axes |x| depends on the cell only
// through n = i^2 + j^2 + k^2, so the angular-weighted field is first
// histogrammed into Q[n] and each pair then costs one pass over n.
// ---------------------------------------------------------------------------

// S_p = V_cell sum_n j_a(ka[p.x] r_n) j_b(kb[p.y] r_n) Q[n], r_n = sqrt(n) dr,
// over entries with Q[n] != 0.  The pairs index two short lists of distinct
// wavenumbers (one per bin of each side), so a block evaluates the Bessel
// splines once per (wavenumber, n) of its chunk of `nch` radii into LDS and
// each thread then reduces pairs from the tables.  Output: per-block
// partials [gridDim.x][P] for a column reduction.
__device__ __forceinline__ void k_shot_pairs_tab_body(
    const double* __restrict__ Q, long long nq, double dr, SplineDev sa,
    SplineDev sb, const double* __restrict__ ka, int nka,
    const double* __restrict__ kb, int nkb, const int2* __restrict__ pidx, int P,
    int nch, double vol_cell, double2* __restrict__ part) {
  extern __shared__ double sp_lds[];
  const int ld = nch + 1;                   // odd row stride: rows on distinct banks
  double* ja = sp_lds;                      // [nka][ld]
  double* jb = ja + (size_t)nka * ld;       // [nkb][ld]
  double* qs = jb + (size_t)nkb * ld;       // [nch][2]
  const long long n0 = (long long)blockIdx.x * nch;
  const int cnt = int(min((long long)nch, nq - n0));
  for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
    qs[2 * i] = Q[2 * (n0 + i)];
    qs[2 * i + 1] = Q[2 * (n0 + i) + 1];
  }
  const int nev = (nka + nkb) * cnt;
  for (int e = threadIdx.x; e < nev; e += blockDim.x) {
    const int row = e / cnt, i = e - row * cnt;
    const double r = sqrt(double(n0 + i)) * dr;
    if (row < nka) ja[row * ld + i] = sa.eval(ka[row] * r);
    else jb[(row - nka) * ld + i] = sb.eval(kb[row - nka] * r);
  }
  __syncthreads();
  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    const int2 q = pidx[p];
    const double* a = ja + q.x * ld;
    const double* c = jb + q.y * ld;
    double sx = 0., sy = 0.;
    for (int i = 0; i < cnt; i++) {
      const double qx = qs[2 * i], qy = qs[2 * i + 1];
      if (qx == 0. && qy == 0.) continue;
      const double jj = a[i] * c[i];
      sx += jj * qx;
      sy += jj * qy;
    }
    part[(long long)blockIdx.x * P + p] = dmk(sx * vol_cell, sy * vol_cell);
  }
}

static __global__ void __launch_bounds__(256)
k_shot_pairs_tab(const double* __restrict__ Q, long long nq, double dr,
                 SplineDev sa, SplineDev sb, const double* __restrict__ ka, int nka,
                 const double* __restrict__ kb, int nkb,
                 const int2* __restrict__ pidx, int P, int nch, double vol_cell,
                 double2* __restrict__ part) {
  k_shot_pairs_tab_body(Q, nq, dr, sa, sb, ka, nka, kb, nkb, pidx, P, nch, vol_cell,
                        part);
}

// Radii per block of k_shot_pairs_tab (at most 128, LDS within 64 KiB) and
// its dynamic LDS bytes.
inline int shot_tab_chunk(int nka, int nkb) {
  const int rows = nka + nkb;
  const int fit = (8192 - rows) / (rows + 2);
  return fit < 1 ? 1 : (fit > 128 ? 128 : fit);
}
inline size_t shot_tab_smem(int nka, int nkb, int nch) {
  return ((size_t)(nka + nkb) * (nch + 1) + 2 * (size_t)nch) * sizeof(double);
}

// General cells: direct sum over the mesh for a group of pairs.
template <typename T>
static __global__ void __launch_bounds__(RED_THREADS)
k_shot_pairs_direct(const c2<T>* __restrict__ t, Geom g, int la, int ma,
                    int lb, int mb, SplineDev sa, SplineDev sb,
                    const double* __restrict__ ka, const double* __restrict__ kb,
                    int P, double2* __restrict__ part) {
  constexpr int GP = 8;
  double ax[GP], ay[GP];
  for (int q = 0; q < GP; q++) ax[q] = ay[q] = 0.;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       c < g.nmesh; c += (long long)gridDim.x * blockDim.x) {
    int k = int(c % g.n[2]);
    long long ij = c / g.n[2];
    int j = int(ij % g.n[1]), i = int(ij / g.n[1]);
    double rx = grid_coord(i, g.n[0], g.dr[0]);
    double ry = grid_coord(j, g.n[1], g.dr[1]);
    double rz = grid_coord(k, g.n[2], g.dr[2]);
    double rm = mag3_exact(rx, ry, rz);
    c2<T> tv = t[c];
    double2 ya = ylm_reduced_dev(la, ma, rx, ry, rz);
    double2 yb = ylm_reduced_dev(lb, mb, rx, ry, rz);
    double2 v = dmul(dmk(tv.x, tv.y), dmul(ya, yb));
    for (int q = 0; q < GP && q < P; q++) {
      double jj = sa.eval(ka[q] * rm) * sb.eval(kb[q] * rm);
      ax[q] += jj * v.x;
      ay[q] += jj * v.y;
    }
  }
  __shared__ double red[RED_THREADS];
  for (int q = 0; q < GP && q < P; q++) {
    for (int comp = 0; comp < 2; comp++) {
      red[threadIdx.x] = comp ? ay[q] : ax[q];
      __syncthreads();
      for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
      }
      if (threadIdx.x == 0) {
        if (comp) part[(long long)blockIdx.x * P + q].y = red[0];
        else part[(long long)blockIdx.x * P + q].x = red[0];
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// Scalar reductions
// ---------------------------------------------------------------------------

// Per-block partial of sum_x Re(f)^3 (calc_grid_based_powlaw_norm,
// field.cpp:2017-2065).
template <typename T>
static __global__ void __launch_bounds__(RED_THREADS)
k_cube_sum(const c2<T>* __restrict__ f, long long n, double* __restrict__ part) {
  __shared__ double red[RED_THREADS];
  double s = 0.;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x) {
    double v = f[c].x;
    s += v * v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int q = RED_THREADS / 2; q > 0; q >>= 1) {
    if (threadIdx.x < q) red[threadIdx.x] += red[threadIdx.x + q];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Per-block partial of sum_p y_lm(los_p) w_p^3 (threept.cpp:156-236).
static __global__ void __launch_bounds__(RED_THREADS)
k_sbar(DevParticles p, int ell, int m, double2* __restrict__ part) {
  __shared__ double rr[RED_THREADS], ri[RED_THREADS];
  double sx = 0., sy = 0.;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < p.n;
       i += (long long)gridDim.x * blockDim.x) {
    double w = p.w[i];
    double w3 = w * w * w;
    double2 y = (ell == 0 && m == 0)
                    ? dmk(1., 0.)
                    : ylm_reduced_dev(ell, m, p.los[3 * i], p.los[3 * i + 1],
                                      p.los[3 * i + 2]);
    sx += y.x * w3;
    sy += y.y * w3;
  }
  rr[threadIdx.x] = sx;
  ri[threadIdx.x] = sy;
  __syncthreads();
  for (int q = RED_THREADS / 2; q > 0; q >>= 1) {
    if (threadIdx.x < q) {
      rr[threadIdx.x] += rr[threadIdx.x + q];
      ri[threadIdx.x] += ri[threadIdx.x + q];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = dmk(rr[0], ri[0]);
}

// Elementwise helpers.
template <typename T>
__global__ void k_zero(c2<T>* __restrict__ f, long long n) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x)
    f[c] = mk<T>(T(0), T(0));
}

static __global__ void k_zero_d(double* __restrict__ f, long long n) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x)
    f[c] = 0.;
}

template <typename T>
__global__ void k_copy(c2<T>* __restrict__ dst, const c2<T>* __restrict__ src,
                       long long n) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x)
    dst[c] = src[c];
}

}  // namespace trvg

namespace trvg {

// ---------------------------------------------------------------------------
// Tile-binned assignment.
//
// Particles are counting-sorted by the 8^3-cell tile that holds their base
// cell index (the int(loc) of every scheme).  A workgroup accumulates one
// chunk of a tile's particles into an 11^3 LDS block (the tile plus the
// [-1, +2] window halo) in FP64 and flushes it to the mesh with one global
// atomic per touched cell.  For base indices inside the mesh every scheme's
// wrap rule (field.cpp:655-1033) is cell = (idx + offset) mod n, which is
// what the flush applies.  Particles whose base index falls outside the mesh
// (positions outside [0, L)) are "stragglers" and keep the reference's
// flattened-index rule through k_assign.
// ---------------------------------------------------------------------------

constexpr int TILE = 8;
constexpr int TILE_H = TILE + 3;                    // halo -1 .. +2
constexpr int TILE_CELLS = TILE_H * TILE_H * TILE_H;  // 1331

struct TileGeom {
  int nt[3];          // tiles per axis
  long long ntiles;
};

// Base cell index per axis of a particle (the int(loc) of field.cpp).
__device__ __forceinline__ void base_cell(const Geom& g, double x, double y,
                                          double z, int idx[3], double s[3]) {
  const double pos[3] = {x, y, z};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    double loc = __ddiv_rn(__dmul_rn(double(g.n[a]), pos[a]), g.L[a]);
    idx[a] = int(loc);
    s[a] = __dadd_rn(loc, -double(idx[a]));
  }
}

// Tile id per particle (ntiles for stragglers) and per-tile counts.
static __global__ void __launch_bounds__(256)
k_tile_count(DevParticles p, Geom g, TileGeom tg, int* __restrict__ tile_of,
             unsigned int* __restrict__ counts) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < p.n;
       i += (long long)gridDim.x * blockDim.x) {
    int idx[3];
    double s[3];
    base_cell(g, p.x[i], p.y[i], p.z[i], idx, s);
    long long t = tg.ntiles;
    if (idx[0] >= 0 && idx[0] < g.n[0] && idx[1] >= 0 && idx[1] < g.n[1] &&
        idx[2] >= 0 && idx[2] < g.n[2]) {
      t = ((long long)(idx[0] / TILE) * tg.nt[1] + idx[1] / TILE) * tg.nt[2] +
          idx[2] / TILE;
    }
    tile_of[i] = int(t);
    atomicAdd(&counts[t], 1u);
  }
}

// Exclusive scan of n counts by one workgroup (n up to a few 1e6).
static __global__ void __launch_bounds__(1024)
k_scan_counts(const unsigned int* __restrict__ counts, long long n,
              unsigned int* __restrict__ offsets) {
  __shared__ unsigned int sh[1024];
  __shared__ unsigned int carry;
  if (threadIdx.x == 0) carry = 0u;
  __syncthreads();
  for (long long base = 0; base < n; base += 1024) {
    long long i = base + threadIdx.x;
    unsigned int v = i < n ? counts[i] : 0u;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      unsigned int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < n) offsets[i] = carry + sh[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += sh[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) offsets[n] = carry;
}

// Scatter particle columns into tile order.
static __global__ void __launch_bounds__(256)
k_tile_scatter(DevParticles p, const int* __restrict__ tile_of,
               unsigned int* __restrict__ cursor, double* __restrict__ sx,
               double* __restrict__ sy, double* __restrict__ sz,
               double* __restrict__ sw, double* __restrict__ slos) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < p.n;
       i += (long long)gridDim.x * blockDim.x) {
    unsigned int d = atomicAdd(&cursor[tile_of[i]], 1u);
    sx[d] = p.x[i];
    sy[d] = p.y[i];
    sz[d] = p.z[i];
    sw[d] = p.w[i];
    if (slos) {
      slos[3 * (long long)d] = p.los[3 * i];
      slos[3 * (long long)d + 1] = p.los[3 * i + 1];
      slos[3 * (long long)d + 2] = p.los[3 * i + 2];
    }
  }
}

// One work item: a chunk [start, start + count) of one tile's particles.
struct TileWork {
  int tile;
  int count;
  long long start;
};

template <typename T, int ORDER>
__global__ void __launch_bounds__(256)
k_assign_tiled(c2<T>* __restrict__ mesh, DevParticles p, Geom g, TileGeom tg,
               const TileWork* __restrict__ work, int mode, int ell, int m,
               double scale, int cplx_w) {
  __shared__ double acc_re[TILE_CELLS];
  __shared__ double acc_im[TILE_CELLS];
  const TileWork wk = work[blockIdx.x];
  for (int c = threadIdx.x; c < TILE_CELLS; c += blockDim.x) {
    acc_re[c] = 0.;
    acc_im[c] = 0.;
  }
  const int t = wk.tile;
  const int tz = t % tg.nt[2];
  const int ty = (t / tg.nt[2]) % tg.nt[1];
  const int tx = t / (tg.nt[2] * tg.nt[1]);
  const int o0[3] = {tx * TILE - 1, ty * TILE - 1, tz * TILE - 1};
  const double inv_vol_cell = 1 / g.vol_cell;
  __syncthreads();
  for (int q = threadIdx.x; q < wk.count; q += blockDim.x) {
    const long long pid = wk.start + q;
    int idx[3];
    double s[3];
    base_cell(g, p.x[pid], p.y[pid], p.z[pid], idx, s);
    int lo[3];            // local LDS index of the first covered cell
    double win[ORDER][3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const double sa = s[a];
      int first = idx[a];
      if (ORDER == 1) {
        if (sa >= 0.5) first = idx[a] + 1;
        win[0][a] = 1.;
      } else if (ORDER == 2) {
        win[0][a] = 1. - sa;
        win[1][a] = sa;
      } else if (ORDER == 3) {
        if (sa < 0.5) {
          first = idx[a] - 1;
          win[0][a] = 0.5 * (0.5 - sa) * (0.5 - sa);
          win[1][a] = 0.75 - sa * sa;
          win[2][a] = 0.5 * (0.5 + sa) * (0.5 + sa);
        } else {
          double sp = 1. - sa;
          win[0][a] = 0.5 * (0.5 + sp) * (0.5 + sp);
          win[1][a] = 0.75 - sp * sp;
          win[2][a] = 0.5 * (0.5 - sp) * (0.5 - sp);
        }
      } else {
        first = idx[a] - 1;
        double sm = 1. - sa;
        win[0][a] = 1. / 6 * sm * sm * sm;
        win[1][a] = 1. / 6 * (4. - 6. * sa * sa + 3. * sa * sa * sa);
        win[2][a] = 1. / 6 * (4. - 6. * sm * sm + 3. * sm * sm * sm);
        win[3][a] = 1. / 6 * sa * sa * sa;
      }
      lo[a] = first - o0[a];
    }
    double2 wt = particle_weight(p, pid, mode, ell, m);
    const double wre = inv_vol_cell * wt.x * scale;
    const double wim = inv_vol_cell * wt.y * scale;
#pragma unroll
    for (int il = 0; il < ORDER; il++) {
#pragma unroll
      for (int jl = 0; jl < ORDER; jl++) {
        const int row = ((lo[0] + il) * TILE_H + (lo[1] + jl)) * TILE_H + lo[2];
        const double wij = win[il][0] * win[jl][1];
#pragma unroll
        for (int kl = 0; kl < ORDER; kl++) {
          const double wv = wij * win[kl][2];
          atomicAdd(&acc_re[row + kl], wre * wv);
          if (cplx_w) atomicAdd(&acc_im[row + kl], wim * wv);
        }
      }
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < TILE_CELLS; c += blockDim.x) {
    const double vr = acc_re[c], vi = acc_im[c];
    if (vr == 0. && vi == 0.) continue;
    const int lz = c % TILE_H, ly = (c / TILE_H) % TILE_H, lx = c / (TILE_H * TILE_H);
    int gx = o0[0] + lx, gy = o0[1] + ly, gz = o0[2] + lz;
    gx = (gx % g.n[0] + g.n[0]) % g.n[0];
    gy = (gy % g.n[1] + g.n[1]) % g.n[1];
    gz = (gz % g.n[2] + g.n[2]) % g.n[2];
    T* cell = reinterpret_cast<T*>(mesh + ((long long)gx * g.n[1] + gy) * g.n[2] + gz);
    unsafeAtomicAdd(cell, T(vr));
    if (cplx_w) unsafeAtomicAdd(cell + 1, T(vi));
  }
}

}  // namespace trvg

namespace trvg {

// ---------------------------------------------------------------------------
// Register-tiled pair product-reduction.
//   S[ia][ib] = sum_x FA[ia](x) FB[ib](x) G(x)
// A launch covers a list of 4x4 tiles of the (ia, ib) grid (up to PT_MAXT
// tiles).  Each block stages PT_CELLS cells of every field in LDS (one HBM
// read per field and cell), forms H = FB * G once, and thread (tile, group)
// accumulates its 16 pair sums in FP64 registers over its cells of every
// chunk the block visits.  Output: per-block partials [gridDim.x][ntiles*16].
// ---------------------------------------------------------------------------

constexpr int PT_CELLS = 64;
constexpr int PT_MAXT = 32;

// Radial histogram over octant representatives: the cells (+-a, +-b, +-c)
// share n = a^2 + b^2 + c^2, so each thread sums its (up to 8) mirror cells
// and issues one atomic pair per representative.
template <typename T>
__device__ __forceinline__ void k_radial_hist_oct_body(const c2<T>* __restrict__ t, Geom g, int la, int ma, int lb,
                  int mb, double* __restrict__ Q) {
  const int h0 = g.n[0] / 2 + 1, h1 = g.n[1] / 2 + 1, h2 = g.n[2] / 2 + 1;
  const long long nrep = (long long)h0 * h1 * h2;
  const bool trivial = (la == 0 && ma == 0 && lb == 0 && mb == 0);
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < nrep;
       r += (long long)gridDim.x * blockDim.x) {
    const int c = int(r % h2);
    const int b = int((r / h2) % h1);
    const int a = int(r / ((long long)h2 * h1));
    int is[2], js[2], ks[2];
    const int ni = (a == 0 || 2 * a == g.n[0]) ? 1 : 2;
    const int nj = (b == 0 || 2 * b == g.n[1]) ? 1 : 2;
    const int nk = (c == 0 || 2 * c == g.n[2]) ? 1 : 2;
    is[0] = a; is[1] = g.n[0] - a;
    js[0] = b; js[1] = g.n[1] - b;
    ks[0] = c; ks[1] = g.n[2] - c;
    double sx = 0., sy = 0.;
    for (int u = 0; u < ni; u++)
      for (int v = 0; v < nj; v++)
        for (int w = 0; w < nk; w++) {
          const int i = is[u], j = js[v], k = ks[w];
          const long long cell = ((long long)i * g.n[1] + j) * g.n[2] + k;
          c2<T> tv = t[cell];
          double2 val = dmk(tv.x, tv.y);
          if (!trivial) {
            double rx = shifted(i, g.n[0]) * g.dr[0];
            double ry = shifted(j, g.n[1]) * g.dr[1];
            double rz = shifted(k, g.n[2]) * g.dr[2];
            double2 ya = ylm_reduced_dev(la, ma, rx, ry, rz);
            double2 yb = ylm_reduced_dev(lb, mb, rx, ry, rz);
            val = dmul(val, dmul(ya, yb));
          }
          sx += val.x;
          sy += val.y;
        }
    const long long nsq = (long long)a * a + (long long)b * b + (long long)c * c;
    unsafeAtomicAdd(&Q[2 * nsq], sx);
    unsafeAtomicAdd(&Q[2 * nsq + 1], sy);
  }
}

template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_radial_hist_oct(const c2<T>* __restrict__ t, Geom g, int la, int ma, int lb,
                  int mb, double* __restrict__ Q) {
  k_radial_hist_oct_body<T>(t, g, la, ma, lb, mb, Q);
}

}  // namespace trvg

namespace trvg {

}  // namespace trvg

namespace trvg {

// Register-tiled pair product-reduction
//   S[ia][ib] = sum_x FA[ia](x) FB[ib](x) G(x).
// A launch covers a list of 4x4 tiles of the (ia, ib) grid (up to PT_MAXT
// tiles).  Each block stages PT_CELLS cells of every field in LDS (one HBM
// read per field and cell), forms H = FB * G once, and thread (tile, group)
// accumulates its 16 pair sums over its cells of every chunk the block
// visits: in T within a chunk, promoted to FP64 accumulators once per
// chunk.  The next chunk's global loads are issued into registers before
// the current chunk is reduced.  Staging map: thread t owns cell
// c = t % PT_CELLS of the chunk and fields f = t / PT_CELLS + k * (256 /
// PT_CELLS), so each thread loads G[c] once.  Output: per-block partials
// [gridDim.x][ntiles * 16].
constexpr int PT3_SLOTS = 8;   // per-thread staged values per side (<= 32 fields)

// MODE 0: complex fields.  MODE 1: the B fields are real and packed two per
// complex mesh (logical field 2p + c is component c of mesh p).
template <typename T, int MODE>
__device__ __forceinline__ void k_pair_tiles3_body(const c2<T>* __restrict__ FA, int na, const c2<T>* __restrict__ FB,
              int nbf, const c2<T>* __restrict__ G, long long nmesh,
              const int2* __restrict__ tiles, int ntiles,
              double2* __restrict__ part) {
  extern __shared__ char sm3_raw[];
  c2<T>* sa = reinterpret_cast<c2<T>*>(sm3_raw);           // [na][PT_CELLS]
  c2<T>* sh = sa + (size_t)na * PT_CELLS;                   // [nbf][PT_CELLS]
  constexpr int FSTEP = 256 / PT_CELLS;                     // 4
  const int ngroups = 256 / ntiles;
  const int group = threadIdx.x % ngroups;
  const int tile = threadIdx.x / ngroups;
  const bool active = tile < ntiles;
  const int2 t0 = tiles[active ? tile : 0];
  double ar[16], ai[16];
#pragma unroll
  for (int q = 0; q < 16; q++) ar[q] = ai[q] = 0.;
  int ia[4], ib[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    ia[q] = min(t0.x + q, na - 1) * PT_CELLS;
    ib[q] = min(t0.y + q, nbf - 1) * PT_CELLS;
  }
  const int sc = threadIdx.x % PT_CELLS;   // staged cell
  const int sf = threadIdx.x / PT_CELLS;   // first staged field
  c2<T> ra[PT3_SLOTS], rb[PT3_SLOTS], rg;
  const long long nchunks = (nmesh + PT_CELLS - 1) / PT_CELLS;
  constexpr bool A_PACKED = false, B_PACKED = (MODE == 1);
  // Staged meshes per side: packed sides hold two logical fields per mesh.
  const int ma = A_PACKED ? (na + 1) / 2 : na;
  const int mb = B_PACKED ? (nbf + 1) / 2 : nbf;
  auto fetch = [&](long long ch) {
    const long long gi = ch * PT_CELLS + sc;
    const bool ok = (ch < nchunks) && (gi < nmesh);
#pragma unroll
    for (int k = 0; k < PT3_SLOTS; k++) {
      const int f = sf + k * FSTEP;
      ra[k] = (ok && f < ma) ? FA[(long long)f * nmesh + gi] : mk<T>(T(0), T(0));
      rb[k] = (ok && f < mb) ? FB[(long long)f * nmesh + gi] : mk<T>(T(0), T(0));
    }
    rg = ok ? G[gi] : mk<T>(T(0), T(0));
  };
  long long ch = blockIdx.x;
  fetch(ch);
  for (; ch < nchunks; ch += gridDim.x) {
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PT3_SLOTS; k++) {
      const int f = sf + k * FSTEP;
      if (A_PACKED) {
        if (2 * f < na) sa[2 * f * PT_CELLS + sc] = mk<T>(ra[k].x, T(0));
        if (2 * f + 1 < na) sa[(2 * f + 1) * PT_CELLS + sc] = mk<T>(ra[k].y, T(0));
      } else if (f < na) {
        sa[f * PT_CELLS + sc] = ra[k];
      }
      if (MODE == 1) {
        if (2 * f < nbf) sh[2 * f * PT_CELLS + sc] = mk<T>(rb[k].x * rg.x, rb[k].x * rg.y);
        if (2 * f + 1 < nbf)
          sh[(2 * f + 1) * PT_CELLS + sc] = mk<T>(rb[k].y * rg.x, rb[k].y * rg.y);
      } else if (f < nbf) {
        sh[f * PT_CELLS + sc] = cmul<T>(rb[k], rg);
      }
    }
    __syncthreads();
    fetch(ch + gridDim.x);
    if (active) {
      T pr[16], pi[16];
#pragma unroll
      for (int q = 0; q < 16; q++) pr[q] = pi[q] = T(0);
      for (int c = group; c < PT_CELLS; c += ngroups) {
        c2<T> x[4], y[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          x[q] = sa[ia[q] + c];
          y[q] = sh[ib[q] + c];
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
#pragma unroll
          for (int j = 0; j < 4; j++) {
            pr[i * 4 + j] += x[i].x * y[j].x - x[i].y * y[j].y;
            pi[i * 4 + j] += x[i].x * y[j].y + x[i].y * y[j].x;
          }
        }
      }
#pragma unroll
      for (int q = 0; q < 16; q++) {
        ar[q] += double(pr[q]);
        ai[q] += double(pi[q]);
      }
    }
  }
  __syncthreads();
  double2* red = reinterpret_cast<double2*>(sm3_raw);   // [256]
  for (int q = 0; q < 16; q++) {
    red[threadIdx.x] = active ? dmk(ar[q], ai[q]) : dmk(0., 0.);
    __syncthreads();
    if (threadIdx.x < ntiles) {
      double sx = 0., sy = 0.;
      for (int gr = 0; gr < ngroups; gr++) {
        double2 v = red[threadIdx.x * ngroups + gr];
        sx += v.x;
        sy += v.y;
      }
      part[(long long)blockIdx.x * ntiles * 16 + threadIdx.x * 16 + q] = dmk(sx, sy);
    }
    __syncthreads();
  }
}

// All-real pair reduction: A and B real and packed two per complex mesh, G
// real (shells of even l, m = 0 with an M = 0 or periodic-box G).  Same
// tiles and output as the complex body with imaginary parts zero; values
// are staged as scalars and chunks are PR2_CELLS cells, since each pair
// costs one FMA per cell here.
constexpr int PR2_CELLS = 128;

template <typename T>
__device__ __forceinline__ void k_pair_real_body(
    const c2<T>* __restrict__ FA, int na, const c2<T>* __restrict__ FB, int nbf,
    const c2<T>* __restrict__ G, long long nmesh, const int2* __restrict__ tiles,
    int ntiles, double2* __restrict__ part) {
  extern __shared__ char smr_raw[];
  T* sa = reinterpret_cast<T*>(smr_raw);                 // [na][PR2_CELLS]
  T* sh = sa + (size_t)na * PR2_CELLS;                    // [nbf][PR2_CELLS]
  constexpr int FSTEP = 256 / PR2_CELLS;                  // packed meshes per pass
  constexpr int SLOTS = 8;                                // 16 meshes = 32 fields
  const int ngroups = 256 / ntiles;
  const int group = threadIdx.x % ngroups;
  const int tile = threadIdx.x / ngroups;
  const bool active = tile < ntiles;
  const int2 t0 = tiles[active ? tile : 0];
  double ar[16];
#pragma unroll
  for (int q = 0; q < 16; q++) ar[q] = 0.;
  int ia[4], ib[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    ia[q] = min(t0.x + q, na - 1) * PR2_CELLS;
    ib[q] = min(t0.y + q, nbf - 1) * PR2_CELLS;
  }
  const int sc = threadIdx.x % PR2_CELLS;
  const int sf = threadIdx.x / PR2_CELLS;
  const int ma = (na + 1) / 2, mb = (nbf + 1) / 2;
  c2<T> ra[SLOTS], rb[SLOTS];
  T rg;
  const long long nchunks = (nmesh + PR2_CELLS - 1) / PR2_CELLS;
  auto fetch = [&](long long ch) {
    const long long gi = ch * PR2_CELLS + sc;
    const bool ok = (ch < nchunks) && (gi < nmesh);
#pragma unroll
    for (int k = 0; k < SLOTS; k++) {
      const int f = sf + k * FSTEP;
      ra[k] = (ok && f < ma) ? FA[(long long)f * nmesh + gi] : mk<T>(T(0), T(0));
      rb[k] = (ok && f < mb) ? FB[(long long)f * nmesh + gi] : mk<T>(T(0), T(0));
    }
    rg = ok ? G[gi].x : T(0);
  };
  long long ch = blockIdx.x;
  fetch(ch);
  for (; ch < nchunks; ch += gridDim.x) {
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SLOTS; k++) {
      const int f = sf + k * FSTEP;
      if (2 * f < na) sa[2 * f * PR2_CELLS + sc] = ra[k].x;
      if (2 * f + 1 < na) sa[(2 * f + 1) * PR2_CELLS + sc] = ra[k].y;
      if (2 * f < nbf) sh[2 * f * PR2_CELLS + sc] = rb[k].x * rg;
      if (2 * f + 1 < nbf) sh[(2 * f + 1) * PR2_CELLS + sc] = rb[k].y * rg;
    }
    __syncthreads();
    fetch(ch + gridDim.x);
    if (active) {
      T pr[16];
#pragma unroll
      for (int q = 0; q < 16; q++) pr[q] = T(0);
      for (int c = group; c < PR2_CELLS; c += ngroups) {
        T x[4], y[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          x[q] = sa[ia[q] + c];
          y[q] = sh[ib[q] + c];
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) pr[i * 4 + j] += x[i] * y[j];
      }
#pragma unroll
      for (int q = 0; q < 16; q++) ar[q] += double(pr[q]);
    }
  }
  __syncthreads();
  double* red = reinterpret_cast<double*>(smr_raw);   // [256]
  for (int q = 0; q < 16; q++) {
    red[threadIdx.x] = active ? ar[q] : 0.;
    __syncthreads();
    if (threadIdx.x < ntiles) {
      double sx = 0.;
      for (int gr = 0; gr < ngroups; gr++) sx += red[threadIdx.x * ngroups + gr];
      part[(long long)blockIdx.x * ntiles * 16 + threadIdx.x * 16 + q] = dmk(sx, 0.);
    }
    __syncthreads();
  }
}

// Bytes of dynamic LDS of a pair-reduction launch of a given MODE.
template <typename T>
inline size_t pair_smem(int mode, int nas, int nbs) {
  const size_t staged = mode == 2 ? (size_t)(nas + nbs) * PR2_CELLS * sizeof(T)
                                  : (size_t)(nas + nbs) * PT_CELLS * sizeof(c2<T>);
  return staged > 256 * sizeof(double2) ? staged : 256 * sizeof(double2);
}

template <typename T, int MODE>
__global__ void __launch_bounds__(256)
k_pair_tiles3(const c2<T>* __restrict__ FA, int na, const c2<T>* __restrict__ FB,
              int nbf, const c2<T>* __restrict__ G, long long nmesh,
              const int2* __restrict__ tiles, int ntiles,
              double2* __restrict__ part) {
  if constexpr (MODE == 2)
    k_pair_real_body<T>(FA, na, FB, nbf, G, nmesh, tiles, ntiles, part);
  else
    k_pair_tiles3_body<T, MODE>(FA, na, FB, nbf, G, nmesh, tiles, ntiles, part);
}

}  // namespace trvg
