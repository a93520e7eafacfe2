// kernels.hpp -- mesh kernels of the three-point engine (gfx950).
//
// All bin sums, normalisation sums and shot-noise amplitudes accumulate in
// FP64; T (double or float) is only the storage/arithmetic type of mesh
// values.  Integer counts (mode / pair counts) are exact.
#pragma once

#include "common.hpp"
#include "fft.hpp"
#include "fft_gen.hpp"

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
// used for interlacing.  Particles [*first, n) (all of them when `first` is
// null) are assigned.  Hardware FP64/FP32 atomics on the coarse-grained
// mesh.
template <typename T, int ORDER>
__global__ void __launch_bounds__(256)
k_assign(c2<T>* __restrict__ mesh, DevParticles p, Geom g, int mode, int ell,
         int m, double scale, int cplx_w, int shadow,
         const unsigned int* __restrict__ first) {
  const double inv_vol_cell = 1 / g.vol_cell;
  const long long p0 = first ? (long long)*first : 0;
  for (long long pid = p0 + blockIdx.x * (long long)blockDim.x + threadIdx.x;
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
        } else if (!shadow) {
          ijk[0][a] = idx;
          ijk[1][a] = (idx == ng - 1) ? 0 : idx + 1;
          ijk[2][a] = (ijk[1][a] == ng - 1) ? 0 : ijk[1][a] + 1;
          double sp = 1. - s;
          win[0][a] = 0.5 * (0.5 + sp) * (0.5 + sp);
          win[1][a] = 0.75 - sp * sp;
          win[2][a] = 0.5 * (0.5 - sp) * (0.5 - sp);
        } else {
          // The reference's shadow pass tests the base index for the last
          // cell too (field.cpp:944-950): kept for identical fields.
          ijk[0][a] = idx;
          ijk[1][a] = (idx == ng - 1) ? 0 : idx + 1;
          ijk[2][a] = (idx == ng - 1) ? 0 : ijk[1][a] + 1;
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
          if (cplx_w == 2) {           // real weight into the imaginary part
            unsafeAtomicAdd(cell + 1, T(wre * wv));
            continue;
          }
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

// Final z pass of the x-first shell transform (in place): rows with
// k_z^2 >= kcut2 were never written by the y pass and read as zero (the
// same predicate as Prune::p2 over the z axis).
template <typename T>
struct LoadMaskedZ {
  const c2<T>* src;
  double kcut2, dkz;
  int nz;
  __device__ c2<T> operator()(long long g, int, int, int k) const {
    const double kz = __dmul_rn(double(shifted(k, nz)), dkz);
    return __dmul_rn(kz, kz) < kcut2 ? ld_mesh<T>(src + g) : mk<T>(T(0), T(0));
  }
};

template <typename T>
struct LoadMaskedZBatch {
  const c2<T>* base;
  long long fstride;
  const double* kcut2;
  double dkz;
  int nz;
  __device__ LoadMaskedZ<T> field(int f) const {
    return LoadMaskedZ<T>{base + (long long)f * fstride, kcut2[f], dkz, nz};
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

// x pass (x-first order) of every j_l shell field of one (l, m) at once,
// field f = j_l(|k| r_f) y_lm(k^) src(k) / W(k) / V
// (inv_fourier_transform_sjl_ylm_wgtd_field, field.cpp:1908-2010):
// a workgroup takes NLINES x lines, forms y_lm(k^) src(k) / W(k) and |k|
// of its elements once, then for each transform fills LDS with
// j_l(|k| r_f) times that factor (PACKED: two radii in the real and
// imaginary parts, r < 0 absent), transforms and writes field f.
template <typename T, int LOGN, int EPT, bool PACKED>
__global__ void __launch_bounds__(FFT_THREADS)
fft_x_sjl_multi(c2<T>* __restrict__ base, long long fstride, int nt,
                const double2* __restrict__ src, AxisTables tab, Geom geo, SplineDev sj,
                const double* __restrict__ r, int ell, int m,
                const c2<T>* __restrict__ tw) {
  constexpr int N = 1 << LOGN;
  constexpr int E = FFT_THREADS * EPT;
  constexpr int NLINES = E / N;
  constexpr int LS = N + 1;
  extern __shared__ char smem_raw[];
  c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
  c2<T>* W = buf + NLINES * LS;
  const int t = threadIdx.x;
  const int n1 = geo.n[1], n2 = geo.n[2];
  const long long S = (long long)n1 * n2;
  const long long id0 = xcd_block(blockIdx.x, gridDim.x) * NLINES;
  for (int q = t; q < N; q += FFT_THREADS) W[q] = tw[q];
  // Per element: y_lm src / W / V in the working precision (FP32 fields
  // take j_l and the product in FP32) and |k|.
  c2<T> fac[EPT];
  double km[EPT];
  const double inv_vol = 1. / geo.vol;
#pragma unroll
  for (int e = 0; e < EPT; e++) {
    const int idx = t + e * FFT_THREADS;
    const int n = idx / NLINES, l = idx % NLINES;
    const long long id = id0 + l;
    const int j = int(id / n2), k = int(id % n2);
    const double kx = grid_coord(n, geo.n[0], geo.dk[0]);
    const double ky = grid_coord(j, n1, geo.dk[1]);
    const double kz = grid_coord(k, n2, geo.dk[2]);
    km[e] = mag3_exact(kx, ky, kz);
    const double2 v = src[id + (long long)n * S];
    const double iw = inv_vol / window_at(tab, n, j, k);
    const double2 fe = dmul(ylm_reduced_dev(ell, m, kx, ky, kz), dmk(v.x * iw, v.y * iw));
    fac[e] = mk<T>(T(fe.x), T(fe.y));
  }
  for (int f = 0; f < nt; f++) {
    const double r0 = PACKED ? r[2 * f] : r[f];
    const double r1 = PACKED ? r[2 * f + 1] : -1.;
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      const int idx = t + e * FFT_THREADS;
      T j0, j1 = T(0);
      if (sizeof(T) == 4) {
        j0 = T(sj.eval32(km[e] * r0));
        if (PACKED && r1 >= 0.) j1 = T(sj.eval32(km[e] * r1));
      } else {
        j0 = T(sj.eval(km[e] * r0));
        if (PACKED && r1 >= 0.) j1 = T(sj.eval(km[e] * r1));
      }
      c2<T> val;
      if (PACKED) {
        val = mk<T>(j0 * fac[e].x - j1 * fac[e].y, j0 * fac[e].y + j1 * fac[e].x);
      } else {
        val = mk<T>(j0 * fac[e].x, j0 * fac[e].y);
      }
      buf[(idx % NLINES) * LS + idx / NLINES] = val;
    }
    __syncthreads();
    lds_fft_lines<T, LOGN, EPT>(buf, W, +1);
    c2<T>* out = base + (long long)f * fstride;
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      const int idx = t + e * FFT_THREADS;
      const int n = idx / NLINES, l = idx % NLINES;
      out[id0 + l + (long long)n * S] = buf[l * LS + n];
    }
    __syncthreads();
  }
}

// fft_x_sjl_multi for an x axis that is not a power of two (mixed-radix
// transform of fft_gen.hpp; nl <= NL lines per workgroup, at most eight
// elements per thread).
template <typename T, bool PACKED>
__global__ void __launch_bounds__(FFT_THREADS)
fft_x_sjl_multi_gen(c2<T>* __restrict__ base, long long fstride, int nt,
                    const double2* __restrict__ src, AxisTables tab, Geom geo,
                    SplineDev sj, const double* __restrict__ r, int ell, int m,
                    const c2<T>* __restrict__ tw, GenPlan P) {
  constexpr int MAXE = GEN_ELEMS / FFT_THREADS;
  const int N = P.n, LS = N + 1;
  GenSmem<T> sm(P);
  const int n1 = geo.n[1], n2 = geo.n[2];
  const long long S = (long long)n1 * n2;
  const long long id0 = (long long)blockIdx.x * P.nl;
  if (id0 >= S) return;
  const int nl = int(min((long long)P.nl, S - id0));
  const int total = nl * N;
  for (int q = threadIdx.x; q < N; q += FFT_THREADS) sm.W[q] = tw[q];
  c2<T> fac[MAXE];
  double km[MAXE];
  const double inv_vol = 1. / geo.vol;
#pragma unroll
  for (int e = 0; e < MAXE; e++) {
    const int idx = threadIdx.x + e * FFT_THREADS;
    fac[e] = mk<T>(T(0), T(0));
    km[e] = 0.;
    if (idx >= total) continue;
    const int n = idx / nl, l = idx % nl;
    const long long id = id0 + l;
    const int j = int(id / n2), k = int(id % n2);
    const double kx = grid_coord(n, geo.n[0], geo.dk[0]);
    const double ky = grid_coord(j, n1, geo.dk[1]);
    const double kz = grid_coord(k, n2, geo.dk[2]);
    km[e] = mag3_exact(kx, ky, kz);
    const double2 v = src[id + (long long)n * S];
    const double iw = inv_vol / window_at(tab, n, j, k);
    const double2 fe = dmul(ylm_reduced_dev(ell, m, kx, ky, kz), dmk(v.x * iw, v.y * iw));
    fac[e] = mk<T>(T(fe.x), T(fe.y));
  }
  for (int f = 0; f < nt; f++) {
    const double r0 = PACKED ? r[2 * f] : r[f];
    const double r1 = PACKED ? r[2 * f + 1] : -1.;
#pragma unroll
    for (int e = 0; e < MAXE; e++) {
      const int idx = threadIdx.x + e * FFT_THREADS;
      if (idx >= total) continue;
      T j0, j1 = T(0);
      if (sizeof(T) == 4) {
        j0 = T(sj.eval32(km[e] * r0));
        if (PACKED && r1 >= 0.) j1 = T(sj.eval32(km[e] * r1));
      } else {
        j0 = T(sj.eval(km[e] * r0));
        if (PACKED && r1 >= 0.) j1 = T(sj.eval(km[e] * r1));
      }
      const c2<T> val = PACKED ? mk<T>(j0 * fac[e].x - j1 * fac[e].y, j0 * fac[e].y + j1 * fac[e].x)
                               : mk<T>(j0 * fac[e].x, j0 * fac[e].y);
      sm.a[(idx % nl) * LS + idx / nl] = val;
    }
    __syncthreads();
    const c2<T>* R = lds_fft_gen<T>(sm.a, sm.b, sm.W, P, nl);
    c2<T>* out = base + (long long)f * fstride;
    for (int idx = threadIdx.x; idx < total; idx += FFT_THREADS) {
      const int n = idx / nl, l = idx % nl;
      st_mesh<T>(out + id0 + l + (long long)n * S, R[l * LS + n]);
    }
    __syncthreads();
  }
}

// (f_a f_b^* / C1 - S C1 / C1) / V  (field.cpp:3007-3025, 3271-3289).
template <typename T, typename S = T>
struct LoadTwopt {
  const c2<S>* fa;
  const c2<S>* fb;
  AxisTables tab;
  double2 sbar;
  double inv_vol;
  double fb_dc = 0.;   // added to fb at k = 0 (fb = an unsubtracted field's
                       // fluctuation spectrum)
  __device__ c2<T> operator()(long long g, int i, int j, int k) const {
    c2<S> a = fa[g], b = fb[g];
    if (g == 0) b.x += S(fb_dc);
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
// Accumulates n_modes, sum(f_a f_b^*/W * y), sum(S C1/W * y) (and sum |k|)
// per bin.
//
// The coarse bin of every cell depends on the geometry and the edges only,
// and is the same at k_z and -k_z.  A TwoptMap holds it per z line (i, j)
// over the half window q = 0..K of k_z indices (all q <= n/2 when the window
// covers the line), built once by k_twopt_binmap; the statistics kernel
// visits each half-window cell q with its mirror n - q, sharing the bin,
// the window factors and y_lm (y(-k_z) = (-1)^(l + |m|) y(k_z)).
// ---------------------------------------------------------------------------

constexpr unsigned short TWOPT_NONE = 0xFFFF;

struct TwoptMap {
  const int* K;               // per line: half-window end (-1: nothing in range)
  const long long* off;       // per line: offset of its half window in `bin`
  const unsigned short* bin;  // coarse bin per half-window cell, TWOPT_NONE outside
};

// Largest |k| that can enter a bin (top edge plus two fine-bin widths).
__host__ __device__ inline double twopt_kmax(double top_edge) {
  return (top_edge + 2. * 1.e-5) * (1. + 1.e-12);
}

// One wave per z line; the reference's exact rule per half-window cell.
static __global__ void __launch_bounds__(RED_THREADS)
k_twopt_binmap(Geom g, TwoptMap map, const double* __restrict__ edges, int nb,
               unsigned short* __restrict__ bin) {
  const int n_sample = 1000000;
  const double dk_sample = 1.e-5;
  const long long nlines = (long long)g.n[0] * g.n[1];
  const int half = g.n[2] / 2;
  const int nw = blockDim.x / 64, wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  for (long long l = (long long)blockIdx.x * nw + wave; l < nlines;
       l += (long long)gridDim.x * nw) {
    const int K = map.K[l];
    if (K < 0) continue;
    const int nq = (K >= half ? half : K) + 1;
    const int i = int(l / g.n[1]), j = int(l % g.n[1]);
    const double kx = grid_coord(i, g.n[0], g.dk[0]);
    const double ky = grid_coord(j, g.n[1], g.dk[1]);
    unsigned short* out = bin + map.off[l];
    for (int q = lane; q < nq; q += 64) {
      const double kz = grid_coord(q, g.n[2], g.dk[2]);
      const double km = mag3_exact(kx, ky, kz);
      const int ik = int(__ddiv_rn(km, dk_sample));
      int b = -1;
      if (ik >= 0 && ik < n_sample) b = find_bin(edges, nb, __dmul_rn(double(ik), dk_sample));
      out[q] = b < 0 ? TWOPT_NONE : (unsigned short)b;
    }
  }
}

// Dynamic LDS of the statistics kernel.
__host__ __device__ inline size_t twopt_smem(int nb, int nv, const Geom& g) {
  return size_t(nb) * (nv * 8 + 4) + size_t(2) * (g.n[2] / 2 + 1) * 8;
}

// ILV: interlaced meshes (W and C1 of the interlaced average), a template
// parameter so that the plain kernel does not carry the isotropic C1 code
// (which costs it occupancy).
template <typename T, int NV = 4, bool ILV = false>
__device__ __forceinline__ void k_twopt_fourier_body(const c2<T>* __restrict__ fa, const c2<T>* __restrict__ fb,
                Geom g, AxisTables tab, double2 sbar, int ell, int m, TwoptMap map, int nb,
                unsigned long long* __restrict__ counts,
                double* __restrict__ part, double fb_dc = 0.) {
  // NV = 4: sums of pk (re, im) and sn (re, im) per bin; NV = 5 adds the
  // sum of |k| (the power spectrum's effective wavenumber).
  extern __shared__ double sh2[];   // [nb][NV], z tables [2][n/2 + 1], counts
  const int half = g.n[2] / 2;
  double* zc1 = sh2 + NV * nb;       // C1 factor of k_z (plain windows)
  double* zrw = zc1 + half + 1;      // 1/W factor of k_z
  unsigned int* sh_cnt = reinterpret_cast<unsigned int*>(zrw + half + 1);
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    for (int q = 0; q < NV; q++) sh2[NV * b + q] = 0.;
    sh_cnt[b] = 0u;
  }
  // W = C1 for plain meshes; W = (sinc_x sinc_y sinc_z)^(2p) with
  // interlacing (C1 then from the isotropic form, alias_at).
  for (int q = threadIdx.x; q <= half; q += blockDim.x) {
    if (!ILV) {
      const double c = c1_axis(tab.sin2[2][q], tab.order);
      zc1[q] = c;
      zrw[q] = 1. / c;
    } else {
      double p = tab.sinc[2][q];
      const double s = p;
      for (int o = 1; o < tab.order; o++) p *= s;
      zc1[q] = 0.;
      zrw[q] = 1. / (p * p);
    }
  }
  __syncthreads();
  const int am = m < 0 ? -m : m;
  const double ysgn = ((ell + am) & 1) ? -1. : 1.;
  const long long nlines = (long long)g.n[0] * g.n[1];
  // Each lane takes a contiguous run of the line's half window, so
  // consecutive cells of a lane mostly share a bin: sums stay in registers
  // and go to LDS only when the bin changes and at the end of the run.
  const int nw = blockDim.x / 64, wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  for (long long l = (long long)blockIdx.x * nw + wave; l < nlines;
       l += (long long)gridDim.x * nw) {
   const int K = map.K[l];
   if (K < 0) continue;
   const bool whole = K >= half;
   const int nq = (whole ? half : K) + 1;
   const int i = int(l / g.n[1]), j = int(l % g.n[1]);
   const double kx = grid_coord(i, g.n[0], g.dk[0]);
   const double ky = grid_coord(j, g.n[1], g.dk[1]);
   double cxy = 0., rxy;
   if (!ILV) {
     cxy = c1_axis(tab.sin2[0][i], tab.order) * c1_axis(tab.sin2[1][j], tab.order);
     rxy = 1. / cxy;
   } else {
     double p = tab.sinc[0][i] * tab.sinc[1][j];
     const double s = p;
     for (int o = 1; o < tab.order; o++) p *= s;
     rxy = 1. / (p * p);
   }
   const unsigned short* lb = map.bin + map.off[l];
   const long long base = l * g.n[2];
   const int run = (nq + 63) / 64;
   const int lo = lane * run, hi = lo + run < nq ? lo + run : nq;
   int cur = -1;
   unsigned int acc_n = 0u;
   double a0 = 0., a1 = 0., a2 = 0., a3 = 0., a4 = 0.;
   // Two cells (and their mirrors) per step, all loads issued before either
   // is used; the loads do not wait for the bin (cells outside the binning
   // are rare inside the window).
   constexpr int U = 2;
   for (int q0 = lo; q0 < hi; q0 += U) {
    int bq[U];
    c2<T> A[U], B[U], AM[U], BM[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int qq = q0 + u < hi ? q0 + u : q0;
      const bool pair = qq != 0 && !(whole && qq == half);
      const long long c = base + qq, cm = pair ? base + g.n[2] - qq : c;
      bq[u] = q0 + u < hi ? int(lb[qq]) : int(TWOPT_NONE);
      A[u] = fa[c];
      B[u] = fb[c];
      AM[u] = fa[cm];
      BM[u] = fb[cm];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
     const int b = bq[u];
     if (b == TWOPT_NONE) continue;
     const int q = q0 + u;
     const bool pair = q != 0 && !(whole && q == half);
     const double kz = grid_coord(q, g.n[2], g.dk[2]);
     c2<T> bb = B[u];
     if (base + q == 0) bb.x += T(fb_dc);   // k = 0 offset (as LoadTwopt::fb_dc)
     double2 pk = dmul(dmk(A[u].x, A[u].y), dmk(bb.x, -double(bb.y)));
     if (pair) {
       const double2 pm = dmul(dmk(AM[u].x, AM[u].y), dmk(BM[u].x, -double(BM[u].y)));
       pk.x += ysgn * pm.x;
       pk.y += ysgn * pm.y;
     }
     const double c1 = ILV ? alias_at(tab, i, j, q) : cxy * zc1[q];
     const double iw = rxy * zrw[q];
     const double cw = c1 * iw * (pair ? 1. + ysgn : 1.);
     double2 sn = dmk(sbar.x * cw, sbar.y * cw);
     pk = dmk(pk.x * iw, pk.y * iw);
     const double2 y = ylm_reduced_dev(ell, m, kx, ky, kz);
     pk = dmul(pk, y);
     sn = dmul(sn, y);
     if (b != cur) {
       if (acc_n) {
         atomicAdd(&sh_cnt[cur], acc_n);
         atomicAdd(&sh2[NV * cur], a0);
         atomicAdd(&sh2[NV * cur + 1], a1);
         atomicAdd(&sh2[NV * cur + 2], a2);
         atomicAdd(&sh2[NV * cur + 3], a3);
         if (NV == 5) atomicAdd(&sh2[NV * cur + 4], a4);
       }
       cur = b;
       acc_n = 0u;
       a0 = a1 = a2 = a3 = a4 = 0.;
     }
     acc_n += pair ? 2u : 1u;
     a0 += pk.x;
     a1 += pk.y;
     a2 += sn.x;
     a3 += sn.y;
     if (NV == 5) {
       const double km = mag3_exact(kx, ky, kz);
       a4 += pair ? 2. * km : km;
     }
    }
   }
   if (acc_n) {
     atomicAdd(&sh_cnt[cur], acc_n);
     atomicAdd(&sh2[NV * cur], a0);
     atomicAdd(&sh2[NV * cur + 1], a1);
     atomicAdd(&sh2[NV * cur + 2], a2);
     atomicAdd(&sh2[NV * cur + 3], a3);
     if (NV == 5) atomicAdd(&sh2[NV * cur + 4], a4);
   }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    atomicAdd(&counts[b], (unsigned long long)sh_cnt[b]);
    for (int q = 0; q < NV; q++)
      part[((long long)blockIdx.x * nb + b) * NV + q] = sh2[NV * b + q];
  }
}

template <typename T, int NV = 4, bool ILV = false>
static __global__ void __launch_bounds__(RED_THREADS)
k_twopt_fourier(const c2<T>* __restrict__ fa, const c2<T>* __restrict__ fb,
                Geom g, AxisTables tab, double2 sbar, int ell, int m, TwoptMap map,
                int nb, unsigned long long* __restrict__ counts,
                double* __restrict__ part) {
  k_twopt_fourier_body<T, NV, ILV>(fa, fb, g, tab, sbar, ell, m, map, nb, counts, part);
}

// ---------------------------------------------------------------------------
// Three-point correlation shot noise: fine bins of width 1 (< 1e5)
// (field.cpp:3097-3170): n_pairs, sum r, sum t(x) y_a(x^) y_b(x^).
// ---------------------------------------------------------------------------

template <typename T, bool TENTHS = false>
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
  // Fine bins of width 1 (1e5 of them) for the three-point shot noise, of
  // width 0.1 (1e6) for the two-point correlation function
  // (field.cpp:2875-2877).
  const int n_sample = TENTHS ? 1000000 : 100000;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       c < g.nmesh; c += (long long)gridDim.x * blockDim.x) {
    int k = int(c % g.n[2]);
    long long ij = c / g.n[2];
    int j = int(ij % g.n[1]), i = int(ij / g.n[1]);
    double rx = grid_coord(i, g.n[0], g.dr[0]);
    double ry = grid_coord(j, g.n[1], g.dr[1]);
    double rz = grid_coord(k, g.n[2], g.dr[2]);
    double rm = mag3_exact(rx, ry, rz);
    // r / dr_sample, truncated; the fine bin enters a coarse bin by its
    // lower edge ir * dr_sample.
    int ir = TENTHS ? int(__ddiv_rn(rm, 0.1)) : int(rm);
    if (ir < 0 || ir >= n_sample) continue;
    int b = find_bin(edges, nb, TENTHS ? __dmul_rn(double(ir), 0.1) : double(ir));
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


// G(x) |x|^-power for |x| >= 1e-6 (wide-angle power-law kernel,
// field.cpp:1727-1762).
template <typename T>
__global__ void k_wide_angle(c2<T>* __restrict__ G, Geom g, int power) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < g.nmesh;
       c += (long long)gridDim.x * blockDim.x) {
    const int k = int(c % g.n[2]);
    const long long ij = c / g.n[2];
    const int j = int(ij % g.n[1]), i = int(ij / g.n[1]);
    const double r = mag3_exact(grid_coord(i, g.n[0], g.dr[0]),
                                grid_coord(j, g.n[1], g.dr[1]),
                                grid_coord(k, g.n[2], g.dr[2]));
    if (r < 1.e-6) continue;
    const double f = pow(r, -double(power));
    const c2<T> v = G[c];
    G[c] = mk<T>(T(double(v.x) * f), T(double(v.y) * f));
  }
}

// Per-block partials of sum_p y_lm(los_p) w_p^3 (threept.cpp:156-236), or
// of sum_p y_lm(los_p) w_p^2 (twopt.cpp:298-381), for several (l, m) at
// once, part[block * n + q]; one pass over the particles.
constexpr int SBAR_MAX_KEYS = 8;
struct SbarKeys {
  int n;
  int ell[SBAR_MAX_KEYS], m[SBAR_MAX_KEYS];
  int wpow;   // power of the weight: 2 (power spectrum), else 3
};

static __global__ void __launch_bounds__(RED_THREADS)
k_sbar_multi(DevParticles p, SbarKeys keys, double2* __restrict__ part) {
  __shared__ double rr[RED_THREADS], ri[RED_THREADS];
  double sx[SBAR_MAX_KEYS], sy[SBAR_MAX_KEYS];
#pragma unroll
  for (int q = 0; q < SBAR_MAX_KEYS; q++) sx[q] = sy[q] = 0.;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < p.n;
       i += (long long)gridDim.x * blockDim.x) {
    const double w = p.w[i];
    const double w3 = keys.wpow == 2 ? w * w : w * w * w;
#pragma unroll
    for (int q = 0; q < SBAR_MAX_KEYS; q++) {
      if (q >= keys.n) break;
      const double2 y = (keys.ell[q] == 0 && keys.m[q] == 0)
                            ? dmk(1., 0.)
                            : ylm_reduced_dev(keys.ell[q], keys.m[q], p.los[3 * i],
                                              p.los[3 * i + 1], p.los[3 * i + 2]);
      sx[q] += y.x * w3;
      sy[q] += y.y * w3;
    }
  }
#pragma unroll
  for (int q = 0; q < SBAR_MAX_KEYS; q++) {
    if (q >= keys.n) break;
    rr[threadIdx.x] = sx[q];
    ri[threadIdx.x] = sy[q];
    __syncthreads();
    for (int h = RED_THREADS / 2; h > 0; h >>= 1) {
      if (threadIdx.x < h) {
        rr[threadIdx.x] += rr[threadIdx.x + h];
        ri[threadIdx.x] += ri[threadIdx.x + h];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(long long)blockIdx.x * keys.n + q] = dmk(rr[0], ri[0]);
    __syncthreads();
  }
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

// Bank classes of the tile-local cells (FP64 accumulators: 32 classes of
// the LDS banks).  Within a tile, particles are ordered by (rank in class,
// class), so the lanes of one LDS atomic instruction start at distinct
// classes; every stencil offset shifts all lanes alike, keeping them
// distinct.
constexpr int TILE_CLASSES = 32;

// Tile id per particle of [begin, end) (ntiles for stragglers), its bank
// class, and its rank within (tile, class) as key2 = rank * 32 + class, from
// one atomic on the (tile, class) counter.  Lanes of a wave that hold the
// same key (spatially ordered input) take their ranks from one atomic.
static __global__ void __launch_bounds__(256)
k_tile_count(DevParticles p, long long begin, long long end, Geom g, TileGeom tg,
             int* __restrict__ tile_of, unsigned int* __restrict__ ccount,
             unsigned int* __restrict__ key2) {
  const int lane = __lane_id();
  for (long long i0 = begin + blockIdx.x * (long long)blockDim.x; i0 < end;
       i0 += (long long)gridDim.x * blockDim.x) {
    const long long i = i0 + threadIdx.x;
    const bool valid = i < end;
    long long t = tg.ntiles;
    int cls = 0;
    if (valid) {
      int idx[3];
      double s[3];
      base_cell(g, p.x[i], p.y[i], p.z[i], idx, s);
      if (idx[0] >= 0 && idx[0] < g.n[0] && idx[1] >= 0 && idx[1] < g.n[1] &&
          idx[2] >= 0 && idx[2] < g.n[2]) {
        t = ((long long)(idx[0] / TILE) * tg.nt[1] + idx[1] / TILE) * tg.nt[2] +
            idx[2] / TILE;
        const int lx = idx[0] % TILE + 1, ly = idx[1] % TILE + 1, lz = idx[2] % TILE + 1;
        cls = ((lx * TILE_H + ly) * TILE_H + lz) % TILE_CLASSES;
      }
    }
    const unsigned int key = unsigned(t) * TILE_CLASSES + unsigned(cls);
    unsigned long long pending = __ballot(valid);
    unsigned int rank = 0u;
    bool have = false;
    for (int round = 0; round < 4 && pending != 0ull; round++) {
      const int leader = __ffsll((long long)pending) - 1;
      const unsigned int k0 = __shfl(key, leader);
      const unsigned long long m = __ballot(valid && key == k0) & pending;
      const unsigned int nm = unsigned(__popcll(m));
      if (nm < 4u) break;                 // mostly distinct keys: direct atomics
      unsigned int base = 0u;
      if (lane == leader) base = atomicAdd(&ccount[k0], nm);
      base = __shfl(base, leader);
      if ((m >> lane) & 1ull) {
        rank = base + unsigned(__popcll(m & ((1ull << lane) - 1ull)));
        have = true;
      }
      pending &= ~m;
    }
    if (valid && !have) rank = atomicAdd(&ccount[key], 1u);
    if (valid) {
      tile_of[i] = int(t);
      key2[i] = rank * TILE_CLASSES + unsigned(cls);
    }
  }
}

// Particles per tile (t < ntiles) and stragglers (t = ntiles) from the
// (tile, class) counters.
static __global__ void __launch_bounds__(256)
k_tile_sum(const unsigned int* __restrict__ ccount, long long nkeys,
           unsigned int* __restrict__ counts) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nkeys;
       t += (long long)gridDim.x * blockDim.x) {
    unsigned int s = 0u;
#pragma unroll 8
    for (int c = 0; c < TILE_CLASSES; c++) s += ccount[t * TILE_CLASSES + c];
    counts[t] = s;
  }
}

// Exclusive scan of n counts by one workgroup (offsets[n] = total): each
// thread sums a contiguous segment (a multiple of four words, read and
// written as 16-byte vectors; `counts` and `offsets` 16-byte aligned), the
// 1024 segment sums are scanned in LDS, and each thread writes its
// segment's offsets.
static __global__ void __launch_bounds__(1024)
k_scan_counts(const unsigned int* __restrict__ counts, long long n,
              unsigned int* __restrict__ offsets) {
  __shared__ unsigned int sh[1024];
  const long long per = ((n + 1023) / 1024 + 3) & ~3LL;
  const long long lo = threadIdx.x * per < n ? threadIdx.x * per : n;
  const long long hi = lo + per < n ? lo + per : n;
  unsigned int own = 0u;
  long long i = lo;
  for (; i + 4 <= hi; i += 4) {
    const uint4 v = *reinterpret_cast<const uint4*>(counts + i);
    own += v.x + v.y + v.z + v.w;
  }
  for (; i < hi; i++) own += counts[i];
  sh[threadIdx.x] = own;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const unsigned int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  unsigned int run = sh[threadIdx.x] - own;
  for (i = lo; i + 4 <= hi; i += 4) {
    const uint4 v = *reinterpret_cast<const uint4*>(counts + i);
    uint4 o;
    o.x = run;
    o.y = o.x + v.x;
    o.z = o.y + v.y;
    o.w = o.z + v.z;
    run = o.w + v.w;
    *reinterpret_cast<uint4*>(offsets + i) = o;
  }
  for (; i < hi; i++) {
    const unsigned int v = counts[i];
    offsets[i] = run;
    run += v;
  }
  if (threadIdx.x == 1023) offsets[n] = sh[1023];
}

// Scatter particles into tile order, ordered within a tile by (rank in
// class, class): position = sum over classes c of min(n_c, rank) plus the
// classes before the particle's that have more than rank members.  Each
// particle lands as one 32-byte record {x, y, z, w} (w = 0 unless `with_w`)
// rather than one word per column: for input in random order every store
// is a partial line, and one record is one partial line instead of three or
// four.  k_tile_columns turns the records back into columns.
static __global__ void __launch_bounds__(256)
k_tile_scatter(DevParticles p, const int* __restrict__ tile_of,
               const unsigned int* __restrict__ offs, const unsigned int* __restrict__ ccount,
               const unsigned int* __restrict__ key2, double2* __restrict__ rec,
               bool with_w, double* __restrict__ slos) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < p.n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long t = tile_of[i];
    const unsigned int r = key2[i] / TILE_CLASSES, cls = key2[i] % TILE_CLASSES;
    // The tile's 32 class counts (one 128-byte line) as 16-byte vectors.
    const uint4* cc = reinterpret_cast<const uint4*>(ccount + t * TILE_CLASSES);
    unsigned int pos = 0;
#pragma unroll
    for (int c4 = 0; c4 < TILE_CLASSES / 4; c4++) {
      const uint4 v = cc[c4];
      const unsigned int nn[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const unsigned int c = unsigned(4 * c4 + k), n = nn[k];
        pos += n < r ? n : r;
        if (c < cls && n > r) pos++;
      }
    }
    const unsigned int d = offs[t] + pos;
    rec[2 * (long long)d] = make_double2(p.x[i], p.y[i]);
    rec[2 * (long long)d + 1] = make_double2(p.z[i], with_w ? p.w[i] : 0.);
    if (slos) {
      slos[3 * (long long)d] = p.los[3 * i];
      slos[3 * (long long)d + 1] = p.los[3 * i + 1];
      slos[3 * (long long)d + 2] = p.los[3 * i + 2];
    }
  }
}

// Tile-ordered records {x, y, z, w} of k_tile_scatter back to columns
// (sw null: the weights stay where they are).
static __global__ void __launch_bounds__(256)
k_tile_columns(const double2* __restrict__ rec, long long n, double* __restrict__ sx,
               double* __restrict__ sy, double* __restrict__ sz, double* __restrict__ sw) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double2 a = rec[2 * i], b = rec[2 * i + 1];
    sx[i] = a.x;
    sy[i] = a.y;
    sz[i] = b.x;
    if (sw) sw[i] = b.y;
  }
}

// One work item: a chunk [start, start + count) of one tile's particles.
struct TileWork {
  int tile;
  int count;
  long long start;
};

// Particles per work item at most (one workgroup each).
constexpr int TILE_CHUNK = 2048;

// Work items of each tile: ceil(count / TILE_CHUNK) (offs: exclusive scan
// of the per-tile counts).
static __global__ void __launch_bounds__(256)
k_tile_chunks(const unsigned int* __restrict__ offs, long long nt,
              unsigned int* __restrict__ nch) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nt;
       t += (long long)gridDim.x * blockDim.x)
    nch[t] = (offs[t + 1] - offs[t] + TILE_CHUNK - 1) / TILE_CHUNK;
}

// The work list (woff: exclusive scan of the chunk counts) and
// meta = {work items, particles inside the mesh}.
static __global__ void __launch_bounds__(256)
k_tile_work(const unsigned int* __restrict__ offs, const unsigned int* __restrict__ woff,
            long long nt, TileWork* __restrict__ work, unsigned int* __restrict__ meta) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nt;
       t += (long long)gridDim.x * blockDim.x) {
    const unsigned int b = offs[t], e = offs[t + 1];
    unsigned int w = woff[t];
    for (unsigned int s0 = b; s0 < e; s0 += TILE_CHUNK, w++) {
      const unsigned int c = e - s0 < unsigned(TILE_CHUNK) ? e - s0 : unsigned(TILE_CHUNK);
      work[w] = TileWork{int(t), int(c), (long long)s0};
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    meta[0] = woff[nt];
    meta[1] = offs[nt];
  }
}

// Window weights of a particle along the three axes and the tile-local
// index (halo origin o0) of the first covered cell per axis.
template <int ORDER>
__device__ __forceinline__ void tile_stencil(const Geom& g, const int o0[3],
                                             double x, double y, double z,
                                             int lo[3], double win[ORDER][3]) {
  int idx[3];
  double s[3];
  base_cell(g, x, y, z, idx, s);
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
}

// Tile halo origin of work item wk.
__device__ __forceinline__ void tile_origin(const TileGeom& tg, int t, int o0[3]) {
  const int tz = t % tg.nt[2];
  const int ty = (t / tg.nt[2]) % tg.nt[1];
  const int tx = t / (tg.nt[2] * tg.nt[1]);
  o0[0] = tx * TILE - 1;
  o0[1] = ty * TILE - 1;
  o0[2] = tz * TILE - 1;
}

// Periodic mesh cell of tile-local cell c.
__device__ __forceinline__ long long tile_cell(const Geom& g, const int o0[3], int c) {
  const int lz = c % TILE_H, ly = (c / TILE_H) % TILE_H, lx = c / (TILE_H * TILE_H);
  int gx = o0[0] + lx, gy = o0[1] + ly, gz = o0[2] + lz;
  gx = (gx % g.n[0] + g.n[0]) % g.n[0];
  gy = (gy % g.n[1] + g.n[1]) % g.n[1];
  gz = (gz % g.n[2] + g.n[2]) % g.n[2];
  return ((long long)gx * g.n[1] + gy) * g.n[2] + gz;
}

template <typename T, int ORDER>
__global__ void __launch_bounds__(256)
k_assign_tiled(c2<T>* __restrict__ mesh, DevParticles p, Geom g, TileGeom tg,
               const TileWork* __restrict__ work, const unsigned int* __restrict__ meta,
               int mode, int ell, int m, double scale, int cplx_w) {
  __shared__ double acc_re[TILE_CELLS];
  __shared__ double acc_im[TILE_CELLS];
  if (blockIdx.x >= meta[0]) return;   // grid sized by a bound
  const TileWork wk = work[blockIdx.x];
  for (int c = threadIdx.x; c < TILE_CELLS; c += blockDim.x) {
    acc_re[c] = 0.;
    acc_im[c] = 0.;
  }
  int o0[3];
  tile_origin(tg, wk.tile, o0);
  const double inv_vol_cell = 1 / g.vol_cell;
  __syncthreads();
  for (int q = threadIdx.x; q < wk.count; q += blockDim.x) {
    const long long pid = wk.start + q;
    int lo[3];
    double win[ORDER][3];
    tile_stencil<ORDER>(g, o0, p.x[pid], p.y[pid], p.z[pid], lo, win);
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
          if (cplx_w == 1) atomicAdd(&acc_im[row + kl], wim * wv);
        }
      }
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < TILE_CELLS; c += blockDim.x) {
    const double vr = acc_re[c], vi = acc_im[c];
    if (vr == 0. && vi == 0.) continue;
    T* cell = reinterpret_cast<T*>(mesh + tile_cell(g, o0, c));
    if (cplx_w == 2) {             // real weight into the imaginary part
      unsafeAtomicAdd(cell + 1, T(vr));
      continue;
    }
    unsafeAtomicAdd(cell, T(vr));
    if (cplx_w) unsafeAtomicAdd(cell + 1, T(vi));
  }
}

// Both survey fields of one (l, m) in one pass over the particles:
//   A += scale_a  y_lm w   W / V_cell   (density contrast)
//   B += scale_b  y_lm^* w^2 W / V_cell (its shot-noise field)
// ra, rb: the real part of cell 0 of each output (cells two T apart, the
// imaginary part following the real part).  CPLX = false for m = 0 (real
// weights: real parts only, so both fields may share one complex mesh,
// rb = ra + 1).  Dynamic LDS: (CPLX ? 4 : 2) * TILE_CELLS doubles.
template <typename T, int ORDER, bool CPLX>
__global__ void __launch_bounds__(256)
k_assign_tiled_dual(T* __restrict__ ra, T* __restrict__ rb,
                    DevParticles p, Geom g, TileGeom tg,
                    const TileWork* __restrict__ work,
                    const unsigned int* __restrict__ meta, int ell, int m,
                    double scale_a, double scale_b) {
  constexpr int NACC = CPLX ? 4 : 2;
  extern __shared__ double acc_dual[];   // [NACC][TILE_CELLS]
  if (blockIdx.x >= meta[0]) return;   // grid sized by a bound
  const TileWork wk = work[blockIdx.x];
  for (int c = threadIdx.x; c < NACC * TILE_CELLS; c += blockDim.x) acc_dual[c] = 0.;
  int o0[3];
  tile_origin(tg, wk.tile, o0);
  const double inv_vol_cell = 1 / g.vol_cell;
  double* a_re = acc_dual;
  double* b_re = acc_dual + TILE_CELLS;
  double* a_im = acc_dual + 2 * TILE_CELLS;
  double* b_im = acc_dual + 3 * TILE_CELLS;
  __syncthreads();
  for (int q = threadIdx.x; q < wk.count; q += blockDim.x) {
    const long long pid = wk.start + q;
    int lo[3];
    double win[ORDER][3];
    tile_stencil<ORDER>(g, o0, p.x[pid], p.y[pid], p.z[pid], lo, win);
    const double2 y = ylm_reduced_dev(ell, m, p.los[3 * pid], p.los[3 * pid + 1],
                                      p.los[3 * pid + 2]);
    const double w = p.w[pid];
    const double fa = inv_vol_cell * w * scale_a;
    const double fb = inv_vol_cell * (w * w) * scale_b;
    const double are = y.x * fa, aim = y.y * fa;
    const double bre = y.x * fb, bim = -y.y * fb;
#pragma unroll
    for (int il = 0; il < ORDER; il++) {
#pragma unroll
      for (int jl = 0; jl < ORDER; jl++) {
        const int row = ((lo[0] + il) * TILE_H + (lo[1] + jl)) * TILE_H + lo[2];
        const double wij = win[il][0] * win[jl][1];
#pragma unroll
        for (int kl = 0; kl < ORDER; kl++) {
          const double wv = wij * win[kl][2];
          atomicAdd(&a_re[row + kl], are * wv);
          atomicAdd(&b_re[row + kl], bre * wv);
          if (CPLX) {
            atomicAdd(&a_im[row + kl], aim * wv);
            atomicAdd(&b_im[row + kl], bim * wv);
          }
        }
      }
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < TILE_CELLS; c += blockDim.x) {
    const double va = a_re[c], vb = b_re[c];
    const double wa = CPLX ? a_im[c] : 0., wb = CPLX ? b_im[c] : 0.;
    if (va == 0. && vb == 0. && wa == 0. && wb == 0.) continue;
    const long long gc = tile_cell(g, o0, c);
    T* ca = ra + 2 * gc;
    T* cb = rb + 2 * gc;
    unsafeAtomicAdd(ca, T(va));
    unsafeAtomicAdd(cb, T(vb));
    if (CPLX) {
      unsafeAtomicAdd(ca + 1, T(wa));
      unsafeAtomicAdd(cb + 1, T(wb));
    }
  }
}

// Spectra of two real fields transformed together as P = FFT[a + i b]:
//   a(k) = (P(k) + P(-k)^*) / 2,  b(k) = (P(k) - P(-k)^*) / (2i),
// written as P <- pa a + pb b and B <- qa a + qb b (plain separation:
// (1, 0, 0, 1)).  Each thread owns the pair (k, -k), so the update is in
// place without a race.
static __global__ void __launch_bounds__(256)
k_unpack_hermitian(double2* __restrict__ P, double2* __restrict__ B, Geom g,
                   double pa, double pb, double qa, double qb) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < g.nmesh;
       c += (long long)gridDim.x * blockDim.x) {
    const int k = int(c % g.n[2]);
    const long long ij = c / g.n[2];
    const int j = int(ij % g.n[1]), i = int(ij / g.n[1]);
    const int im = (g.n[0] - i) % g.n[0], jm = (g.n[1] - j) % g.n[1],
              km = (g.n[2] - k) % g.n[2];
    const long long cm = ((long long)im * g.n[1] + jm) * g.n[2] + km;
    if (cm < c) continue;
    const double2 p = P[c], q = P[cm];
    // cell c: uses P(-k) = q; cell cm: uses P(k) = p.
    const double2 a_c = dmk(0.5 * (p.x + q.x), 0.5 * (p.y - q.y));
    const double2 b_c = dmk(0.5 * (p.y + q.y), -0.5 * (p.x - q.x));
    const double2 a_m = dmk(0.5 * (q.x + p.x), 0.5 * (q.y - p.y));
    const double2 b_m = dmk(0.5 * (q.y + p.y), -0.5 * (q.x - p.x));
    P[c] = dmk(pa * a_c.x + pb * b_c.x, pa * a_c.y + pb * b_c.y);
    B[c] = dmk(qa * a_c.x + qb * b_c.x, qa * a_c.y + qb * b_c.y);
    if (cm != c) {
      P[cm] = dmk(pa * a_m.x + pb * b_m.x, pa * a_m.y + pb * b_m.y);
      B[cm] = dmk(qa * a_m.x + qb * b_m.x, qa * a_m.y + qb * b_m.y);
    }
  }
}

// ca a + cb b, or its conjugate, per cell (real coefficients): the survey
// pair of one (l, m != 0) from the per-catalogue sums S_d = sum_d y w W,
// S_r = sum_r y w W of catalogues with uniform weights w_d, w_r:
//   dn = S_d - alpha S_r,   N = w_d S_d^* + alpha^2 w_r S_r^*
// (N = sum y^* w^2 W = w conj(S) per catalogue), formed as the first pass
// of each transform reads them.
template <typename T>
struct LoadLinear2 {
  const c2<T>* a;
  const c2<T>* b;
  double ca, cb;
  int conj;
  __device__ c2<T> operator()(long long g, int, int, int) const {
    const c2<T> u = a[g], v = b[g];
    const double re = ca * double(u.x) + cb * double(v.x);
    const double im = ca * double(u.y) + cb * double(v.y);
    return mk<T>(T(re), T(conj ? -im : im));
  }
};

}  // namespace trvg

namespace trvg {

// Most 4x4 pair tiles per pair-reduction launch (four thread groups each),
// and most staged mesh rows (both sides and G), which keeps the double-
// buffered staging of the pair kernel within the 160 KiB of LDS.
constexpr int PT_MAXT = 64;
constexpr int PT_MAXROWS = 76;

// U consecutive LDS elements read as one (16-byte aligned) access.
template <typename E, int U>
struct alignas(sizeof(E) * U >= 16 ? 16 : sizeof(E) * U) LdsPack {
  E v[U];
};

// Sum over the (up to 8) mirror cells (+-a, +-b, +-c) of t y_a y_b, with
// y_lm evaluated once at (a, b, c) dr and mirrored by sign flips.
template <typename T>
__device__ __forceinline__ double2 oct_mirror_sum(const c2<T>* __restrict__ t, const Geom& g,
                                                  int a, int b, int c, int la, int ma,
                                                  int lb, int mb, bool trivial) {
  int is[2], js[2], ks[2];
  const int ni = (a == 0 || 2 * a == g.n[0]) ? 1 : 2;
  const int nj = (b == 0 || 2 * b == g.n[1]) ? 1 : 2;
  const int nk = (c == 0 || 2 * c == g.n[2]) ? 1 : 2;
  is[0] = a; is[1] = g.n[0] - a;
  js[0] = b; js[1] = g.n[1] - b;
  ks[0] = c; ks[1] = g.n[2] - c;
  double2 ya0 = dmk(1., 0.), yb0 = dmk(1., 0.);
  if (!trivial) {
    const double ax = fabs(shifted(is[0], g.n[0]) * g.dr[0]);
    const double ay = fabs(shifted(js[0], g.n[1]) * g.dr[1]);
    const double az = fabs(shifted(ks[0], g.n[2]) * g.dr[2]);
    ya0 = ylm_reduced_dev(la, ma, ax, ay, az);
    yb0 = ylm_reduced_dev(lb, mb, ax, ay, az);
  }
  // All mirror cells are loaded before any is used (independent loads in
  // flight together), then summed in the order (u, v, w).
  // (Absent mirrors index the representative itself, so every address is
  // inside the mesh.)
  c2<T> tv8[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int u = q >> 2, v = (q >> 1) & 1, w = q & 1;
    const int iu = u < ni ? u : 0, jv = v < nj ? v : 0, kw = w < nk ? w : 0;
    tv8[q] = t[((long long)is[iu] * g.n[1] + js[jv]) * g.n[2] + ks[kw]];
  }
  double sx = 0., sy = 0.;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int u = q >> 2, v = (q >> 1) & 1, w = q & 1;
    if (!(u < ni && v < nj && w < nk)) continue;
    const int i = is[u], j = js[v], k = ks[w];
    const c2<T> tv = tv8[q];
    double2 val = dmk(tv.x, tv.y);
    if (!trivial) {
      const bool fx = shifted(i, g.n[0]) < 0, fy = shifted(j, g.n[1]) < 0,
                 fz = shifted(k, g.n[2]) < 0;
      val = dmul(val, dmul(ylm_mirror(ya0, fx, fy, fz, la, ma),
                           ylm_mirror(yb0, fx, fy, fz, lb, mb)));
    }
    sx += val.x;
    sy += val.y;
  }
  return dmk(sx, sy);
}

// Radial histogram Q[n] = sum over cells with i^2 + j^2 + k^2 = n (shifted
// indices) of t y_a y_b.  Octant representatives (a, b, c) >= 0 carry their
// mirror cells; with n_x = n_y the rows (a, b) and (b, a) share every n, so
// rows are taken with a >= b and summed together.  One wave per row, lanes
// over c: one atomic pair per (row, c).
template <typename T>
__device__ __forceinline__ void k_radial_hist_oct_body(const c2<T>* __restrict__ t, Geom g, int la, int ma, int lb,
                  int mb, double* __restrict__ Q) {
  const int h0 = g.n[0] / 2 + 1, h1 = g.n[1] / 2 + 1, h2 = g.n[2] / 2 + 1;
  const bool trivial = (la == 0 && ma == 0 && lb == 0 && mb == 0);
  const bool swap = g.n[0] == g.n[1];
  const int nrows = swap ? h0 * (h0 + 1) / 2 : h0 * h1;
  const int nw = blockDim.x / 64, lane = threadIdx.x % 64;
  for (int r = blockIdx.x * nw + threadIdx.x / 64; r < nrows; r += gridDim.x * nw) {
    int a, b;
    if (swap) {
      a = int((sqrt(8. * r + 1.) - 1.) / 2.);
      while (a * (a + 1) / 2 > r) a--;
      while ((a + 1) * (a + 2) / 2 <= r) a++;
      b = r - a * (a + 1) / 2;
    } else {
      a = r / h1;
      b = r % h1;
    }
    for (int c = lane; c < h2; c += 64) {
      double2 s = oct_mirror_sum<T>(t, g, a, b, c, la, ma, lb, mb, trivial);
      if (swap && a != b) {
        const double2 s2 = oct_mirror_sum<T>(t, g, b, a, c, la, ma, lb, mb, trivial);
        s.x += s2.x;
        s.y += s2.y;
      }
      const long long nsq = (long long)a * a + (long long)b * b + (long long)c * c;
      unsafeAtomicAdd(&Q[2 * nsq], s.x);
      unsafeAtomicAdd(&Q[2 * nsq + 1], s.y);
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_radial_hist_oct(const c2<T>* __restrict__ t, Geom g, int la, int ma, int lb,
                  int mb, double* __restrict__ Q) {
  k_radial_hist_oct_body<T>(t, g, la, ma, lb, mb, Q);
}

// Radial histogram without atomics, in two passes.  k_radial_stage writes,
// per octant row (a, b) (rows with a >= b and the (b, a) row folded in when
// n_x = n_y, as k_radial_hist_oct) and c, the mirror sum of t y_a y_b to
// S[row * h_z + c] (lanes over c: coalesced reads of t and writes of S).
// k_radial_gather then sums, for each n, the entries with
// a^2 + b^2 + c^2 = n listed by a geometry-only index (ofs, idx) in a fixed
// order, so the histogram is the same on every run.
template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_radial_stage(const c2<T>* __restrict__ t, Geom g, int la, int ma, int lb, int mb,
               double2* __restrict__ S) {
  const int h0 = g.n[0] / 2 + 1, h1 = g.n[1] / 2 + 1, h2 = g.n[2] / 2 + 1;
  const bool trivial = (la == 0 && ma == 0 && lb == 0 && mb == 0);
  const bool swap = g.n[0] == g.n[1];
  const int nrows = swap ? h0 * (h0 + 1) / 2 : h0 * h1;
  const int nw = blockDim.x / 64, lane = threadIdx.x % 64;
  for (int r = blockIdx.x * nw + threadIdx.x / 64; r < nrows; r += gridDim.x * nw) {
    int a, b;
    if (swap) {
      a = int((sqrt(8. * r + 1.) - 1.) / 2.);
      while (a * (a + 1) / 2 > r) a--;
      while ((a + 1) * (a + 2) / 2 <= r) a++;
      b = r - a * (a + 1) / 2;
    } else {
      a = r / h1;
      b = r % h1;
    }
    for (int c = lane; c < h2; c += 64) {
      double2 s = oct_mirror_sum<T>(t, g, a, b, c, la, ma, lb, mb, trivial);
      if (swap && a != b) {
        const double2 s2 = oct_mirror_sum<T>(t, g, b, a, c, la, ma, lb, mb, trivial);
        s.x += s2.x;
        s.y += s2.y;
      }
      S[(long long)r * h2 + c] = s;
    }
  }
}

// Q[2n], Q[2n + 1] = sum of S[idx[e]] over e in [ofs[n], ofs[n + 1]): eight
// lanes per n, each over every eighth entry, then a fixed butterfly.
constexpr int RG_LANES = 8;

static __global__ void __launch_bounds__(RED_THREADS)
k_radial_gather(const double2* __restrict__ S, const int* __restrict__ ofs,
                const int* __restrict__ idx, long long nq, double* __restrict__ Q) {
  const int sub = threadIdx.x % RG_LANES;
  const long long per_block = blockDim.x / RG_LANES;
  // Every lane of a group runs the same trip count, so the butterfly below
  // always has all eight lanes.
  for (long long n0 = (long long)blockIdx.x * per_block; n0 < nq;
       n0 += (long long)gridDim.x * per_block) {
    const long long n = n0 + threadIdx.x / RG_LANES;
    double sx = 0., sy = 0.;
    if (n < nq) {
      const int e1 = ofs[n + 1];
      for (int e = ofs[n] + sub; e < e1; e += RG_LANES) {
        const double2 v = S[idx[e]];
        sx += v.x;
        sy += v.y;
      }
    }
#pragma unroll
    for (int o = 1; o < RG_LANES; o <<= 1) {
      sx += __shfl_xor(sx, o);
      sy += __shfl_xor(sy, o);
    }
    if (n < nq && sub == 0) {
      Q[2 * n] = sx;
      Q[2 * n + 1] = sy;
    }
  }
}

}  // namespace trvg

namespace trvg {

// ---------------------------------------------------------------------------
// Pair product-reduction with direct-to-LDS staging (gfx950
// global_load_lds_dwordx4), double buffered:
//   S[ia][ib] = sum_x FA[ia](x) FB[ib](x) G(x)
// A chunk is LDC cells (1 KiB of c2<T> per mesh row: one wave-wide
// 16-byte-per-lane load).  Rows: the MA meshes of the A side, the MB meshes
// of the B side, then G.  While a block reduces chunk c from one buffer, the
// loads of its next chunk land in the other; no register staging.  Thread
// (tile, group) folds G into one side per cell (H = F_b G for its tile's
// four B fields, or F_a G for its four A rows when B is real, so each pair
// costs two real FMAs) and accumulates 16 pair sums in T within a chunk,
// promoted to FP64 once per chunk.  MODE 0: complex sides; 1: B side real, two per mesh; 2: both sides
// real, two per mesh, G real; 3: as 2 with both sides the same fields (read
// once).  Cells past the last whole chunk are read from global memory.
// Output: per-block partials [gridDim.x][ntiles * 16].
// ---------------------------------------------------------------------------

template <typename T>
struct PairLds {
  static constexpr int LDC = 1024 / int(sizeof(c2<T>));   // cells per chunk
  static constexpr int ROWB = 1024 + 16;                  // row stride (bytes), bank skew
  static constexpr int PER_LANE = 16 / int(sizeof(c2<T>)); // cells per lane load
};

// Mesh rows of each side for a launch.
__host__ __device__ inline int pair_rows_a(int mode, int na) {
  return mode >= 2 ? (na + 1) / 2 : na;
}
__host__ __device__ inline int pair_rows_b(int mode, int nbf) {
  return mode == 3 ? 0 : (mode >= 1 ? (nbf + 1) / 2 : nbf);
}

template <typename T>
inline size_t pair_lds_smem(int mode, int nas, int nbs) {
  const size_t rows = size_t(pair_rows_a(mode, nas) + pair_rows_b(mode, nbs) + 1);
  const size_t buf = 2 * rows * PairLds<T>::ROWB;
  return buf > 256 * sizeof(double2) ? buf : 256 * sizeof(double2);
}

// Two consecutive cells of one LDS row.
template <typename T>
__device__ __forceinline__ LdsPack<c2<T>, 2> lds_cells2(const char* row, int c) {
  return *reinterpret_cast<const LdsPack<c2<T>, 2>*>(row + (size_t)c * sizeof(c2<T>));
}

template <typename T, int MODE>
__device__ __forceinline__ void k_pair_lds_body(
    const c2<T>* __restrict__ FA, int na, const c2<T>* __restrict__ FB, int nbf,
    const c2<T>* __restrict__ G, long long nmesh, const int2* __restrict__ tiles,
    int ntiles, double2* __restrict__ part) {
  using L = PairLds<T>;
  constexpr bool CPLX = MODE <= 1;
  extern __shared__ __attribute__((aligned(16))) char smp_raw[];
  const int MA = pair_rows_a(MODE, na), MB = pair_rows_b(MODE, nbf);
  const int R = MA + MB + 1;
  const int ng = 256 / ntiles;
  const int group = threadIdx.x % ng;
  const int tile = threadIdx.x / ng;
  const bool active = tile < ntiles;
  const int2 t0 = tiles[active ? tile : 0];
  // LDS rows read by this thread's tile (clamped to existing meshes; the
  // host discards pairs past na / nbf).
  int xr[4], yr[4];
  if (MODE >= 2) {
    const int a0 = t0.x / 2, b0 = t0.y / 2;
#pragma unroll
    for (int q = 0; q < 2; q++) {
      xr[q] = min(a0 + q, MA - 1);
      yr[q] = (MODE == 3) ? min(b0 + q, MA - 1) : MA + min(b0 + q, MB - 1);
    }
    xr[2] = xr[3] = yr[2] = yr[3] = 0;
  } else {
#pragma unroll
    for (int q = 0; q < 4; q++) xr[q] = min(t0.x + q, na - 1);
    if (MODE == 1) {
      const int b0 = t0.y / 2;
      yr[0] = MA + min(b0, MB - 1);
      yr[1] = MA + min(b0 + 1, MB - 1);
      yr[2] = yr[3] = 0;
    } else {
#pragma unroll
      for (int q = 0; q < 4; q++) yr[q] = MA + min(t0.y + q, nbf - 1);
    }
  }
  const int grow = MA + MB;
  double ar[16], ai[16];
#pragma unroll
  for (int q = 0; q < 16; q++) ar[q] = ai[q] = 0.;

  const long long nchunks = nmesh / L::LDC;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  // Wave w loads rows w, w + 4, ... of chunk ch into buffer b.
  auto issue = [&](long long ch, int b) {
    char* base = smp_raw + (size_t)b * R * L::ROWB;
    const long long off = ch * L::LDC + (long long)lane * L::PER_LANE;
    for (int r = wave; r < R; r += 4) {
      const c2<T>* src = r < MA ? FA + (long long)r * nmesh
                                : (r < MA + MB ? FB + (long long)(r - MA) * nmesh : G);
      __builtin_amdgcn_global_load_lds(
          (const void*)(src + off),
          (__attribute__((address_space(3))) void*)(
              (__attribute__((address_space(3))) char*)(base + (size_t)r * L::ROWB)),
          16, 0, 0);
    }
  };
  long long ch = blockIdx.x;
  int cur = 0;
  if (ch < nchunks) issue(ch, 0);
  for (; ch < nchunks; ch += gridDim.x) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (ch + gridDim.x < nchunks) issue(ch + gridDim.x, cur ^ 1);
    if (active) {
      const char* buf = smp_raw + (size_t)cur * R * L::ROWB;
      T pr[16], pi[16];
#pragma unroll
      for (int q = 0; q < 16; q++) pr[q] = pi[q] = T(0);
      for (int cb = group; cb < L::LDC / 2; cb += ng) {
        const int c = 2 * cb;
        const LdsPack<c2<T>, 2> gg = lds_cells2<T>(buf + (size_t)grow * L::ROWB, c);
        if (MODE >= 2) {
          LdsPack<c2<T>, 2> xa[2], yb[2];
#pragma unroll
          for (int q = 0; q < 2; q++) {
            xa[q] = lds_cells2<T>(buf + (size_t)xr[q] * L::ROWB, c);
            yb[q] = lds_cells2<T>(buf + (size_t)yr[q] * L::ROWB, c);
          }
#pragma unroll
          for (int u = 0; u < 2; u++) {
            const T g = gg.v[u].x;
            T x[4], h[4];
#pragma unroll
            for (int q = 0; q < 2; q++) {
              x[2 * q] = xa[q].v[u].x;
              x[2 * q + 1] = xa[q].v[u].y;
              h[2 * q] = yb[q].v[u].x * g;
              h[2 * q + 1] = yb[q].v[u].y * g;
            }
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
              for (int j = 0; j < 4; j++) pr[i * 4 + j] += x[i] * h[j];
          }
        } else {
          LdsPack<c2<T>, 2> xa[4], yb[4];
#pragma unroll
          for (int q = 0; q < 4; q++) xa[q] = lds_cells2<T>(buf + (size_t)xr[q] * L::ROWB, c);
#pragma unroll
          for (int q = 0; q < (MODE == 1 ? 2 : 4); q++)
            yb[q] = lds_cells2<T>(buf + (size_t)yr[q] * L::ROWB, c);
#pragma unroll
          for (int u = 0; u < 2; u++) {
            const c2<T> g = gg.v[u];
            if (MODE == 1) {
              // Real B: form A_a G once per row, then 2 FMAs per pair.
              c2<T> ag[4];
              T b[4];
#pragma unroll
              for (int i = 0; i < 4; i++) ag[i] = cmul<T>(xa[i].v[u], g);
#pragma unroll
              for (int q = 0; q < 2; q++) {
                b[2 * q] = yb[q].v[u].x;
                b[2 * q + 1] = yb[q].v[u].y;
              }
#pragma unroll
              for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                  pr[i * 4 + j] += ag[i].x * b[j];
                  pi[i * 4 + j] += ag[i].y * b[j];
                }
            } else {
              c2<T> h[4];
#pragma unroll
              for (int q = 0; q < 4; q++) h[q] = cmul<T>(yb[q].v[u], g);
#pragma unroll
              for (int i = 0; i < 4; i++) {
                const c2<T> a = xa[i].v[u];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                  pr[i * 4 + j] += a.x * h[j].x - a.y * h[j].y;
                  pi[i * 4 + j] += a.x * h[j].y + a.y * h[j].x;
                }
              }
            }
          }
        }
      }
#pragma unroll
      for (int q = 0; q < 16; q++) {
        ar[q] += double(pr[q]);
        if (CPLX) ai[q] += double(pi[q]);
      }
    }
    cur ^= 1;
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  // Cells past the last whole chunk (a mesh whose cell count is not a
  // multiple of LDC): read from global memory by the block next in turn.
  const long long tail0 = nchunks * L::LDC;
  if (tail0 < nmesh && blockIdx.x == nchunks % gridDim.x && active) {
    auto row = [&](int r) -> const c2<T>* {
      return r < MA ? FA + (long long)r * nmesh
                    : (r < MA + MB ? FB + (long long)(r - MA) * nmesh : G);
    };
    T pr[16], pi[16];
#pragma unroll
    for (int q = 0; q < 16; q++) pr[q] = pi[q] = T(0);
    for (long long c = tail0 + group; c < nmesh; c += ng) {
      const c2<T> g = row(grow)[c];
      if (MODE >= 2) {
        T x[4], h[4];
#pragma unroll
        for (int q = 0; q < 2; q++) {
          const c2<T> xa = row(xr[q])[c], yb = row(yr[q])[c];
          x[2 * q] = xa.x;
          x[2 * q + 1] = xa.y;
          h[2 * q] = yb.x * g.x;
          h[2 * q + 1] = yb.y * g.x;
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) pr[i * 4 + j] += x[i] * h[j];
      } else if (MODE == 1) {
        c2<T> ag[4];
        T b[4];
#pragma unroll
        for (int i = 0; i < 4; i++) ag[i] = cmul<T>(row(xr[i])[c], g);
#pragma unroll
        for (int q = 0; q < 2; q++) {
          const c2<T> yb = row(yr[q])[c];
          b[2 * q] = yb.x;
          b[2 * q + 1] = yb.y;
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) {
            pr[i * 4 + j] += ag[i].x * b[j];
            pi[i * 4 + j] += ag[i].y * b[j];
          }
      } else {
        c2<T> h[4];
#pragma unroll
        for (int q = 0; q < 4; q++) h[q] = cmul<T>(row(yr[q])[c], g);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const c2<T> a = row(xr[i])[c];
#pragma unroll
          for (int j = 0; j < 4; j++) {
            pr[i * 4 + j] += a.x * h[j].x - a.y * h[j].y;
            pi[i * 4 + j] += a.x * h[j].y + a.y * h[j].x;
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
      ar[q] += double(pr[q]);
      if (CPLX) ai[q] += double(pi[q]);
    }
  }
  double2* red = reinterpret_cast<double2*>(smp_raw);   // [256]
  for (int q = 0; q < 16; q++) {
    red[threadIdx.x] = active ? dmk(ar[q], CPLX ? ai[q] : 0.) : dmk(0., 0.);
    __syncthreads();
    if (threadIdx.x < ntiles) {
      double sx = 0., sy = 0.;
      for (int gr = 0; gr < ng; gr++) {
        const double2 v = red[threadIdx.x * ng + gr];
        sx += v.x;
        sy += v.y;
      }
      part[(long long)blockIdx.x * ntiles * 16 + threadIdx.x * 16 + q] = dmk(sx, sy);
    }
    __syncthreads();
  }
}

template <typename T, int MODE>
__global__ void __launch_bounds__(256)
k_pair_lds(const c2<T>* __restrict__ FA, int na, const c2<T>* __restrict__ FB, int nbf,
           const c2<T>* __restrict__ G, long long nmesh, const int2* __restrict__ tiles,
           int ntiles, double2* __restrict__ part) {
  k_pair_lds_body<T, MODE>(FA, na, FB, nbf, G, nmesh, tiles, ntiles, part);
}

}  // namespace trvg
