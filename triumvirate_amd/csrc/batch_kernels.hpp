// batch_kernels.hpp -- kernels of the batched periodic-box bispectrum: many
// catalogues ("mocks") of one geometry per launch.  Mock b of a batch is
// blockIdx.y (or, in the FFT work lists, part of the field index), and every
// per-mock quantity that varies between batches (particle offsets, counts,
// normalisations) is read from device memory, so the whole batch is one
// fixed launch sequence that a HIP graph can capture and replay.
#pragma once

#include "diff_kernels.hpp"
#include "kernels.hpp"

namespace trvg {

// ---- tiled assignment of a batch (all on the device, graph-capturable) ----
// A counting sort by (mock, tile) key b (ntiles + 1) + t, t = ntiles for
// particles whose base cell lies outside the mesh (they keep the direct
// index rule of k_assign), then one workgroup per key accumulates its
// tile's halo block in LDS and adds it to the mesh (k_assign_tiled's
// scheme, in FP64).

// Key of every particle and its rank within the key.
static __global__ void __launch_bounds__(256)
k_tile_key_batch(const double* __restrict__ x, const double* __restrict__ y,
                 const double* __restrict__ z, const long long* __restrict__ offs,
                 Geom g, TileGeom tg, unsigned int* __restrict__ counts,
                 unsigned int* __restrict__ key, unsigned int* __restrict__ rank) {
  const int b = blockIdx.y;
  const long long nk = tg.ntiles + 1;
  const long long p0 = offs[b], p1 = offs[b + 1];
  for (long long p = p0 + blockIdx.x * (long long)blockDim.x + threadIdx.x; p < p1;
       p += (long long)gridDim.x * blockDim.x) {
    int idx[3];
    double s[3];
    base_cell(g, x[p], y[p], z[p], idx, s);
    long long t = tg.ntiles;
    if (idx[0] >= 0 && idx[0] < g.n[0] && idx[1] >= 0 && idx[1] < g.n[1] &&
        idx[2] >= 0 && idx[2] < g.n[2])
      t = ((long long)(idx[0] / TILE) * tg.nt[1] + idx[1] / TILE) * tg.nt[2] + idx[2] / TILE;
    const unsigned int k = unsigned(b * nk + t);
    key[p] = k;
    rank[p] = atomicAdd(&counts[k], 1u);
  }
}

// Exclusive scan of n values in segments of 1024 (segment totals to sums);
// k_scan_counts then scans the totals and k_scan_add adds them back.
static __global__ void __launch_bounds__(1024)
k_scan_seg(const unsigned int* __restrict__ in, long long n, unsigned int* __restrict__ out,
           unsigned int* __restrict__ sums) {
  __shared__ unsigned int sh[1024];
  const long long i = (long long)blockIdx.x * 1024 + threadIdx.x;
  const unsigned int v = i < n ? in[i] : 0u;
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const unsigned int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = sh[threadIdx.x] - v;
  if (threadIdx.x == 1023) sums[blockIdx.x] = sh[1023];
}

static __global__ void __launch_bounds__(1024)
k_scan_add(unsigned int* __restrict__ out, long long n, const unsigned int* __restrict__ soff,
           int nseg) {
  const long long i = (long long)blockIdx.x * 1024 + threadIdx.x;
  if (i < n) out[i] += soff[blockIdx.x];
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = soff[nseg];
}

// Positions in key order.
static __global__ void __launch_bounds__(256)
k_tile_scatter_batch(const double* __restrict__ x, const double* __restrict__ y,
                     const double* __restrict__ z, const long long* __restrict__ offs,
                     const unsigned int* __restrict__ key, const unsigned int* __restrict__ rank,
                     const unsigned int* __restrict__ koff, double* __restrict__ sx,
                     double* __restrict__ sy, double* __restrict__ sz) {
  const int b = blockIdx.y;
  const long long p0 = offs[b], p1 = offs[b + 1];
  for (long long p = p0 + blockIdx.x * (long long)blockDim.x + threadIdx.x; p < p1;
       p += (long long)gridDim.x * blockDim.x) {
    const unsigned int d = koff[key[p]] + rank[p];
    sx[d] = x[p];
    sy[d] = y[p];
    sz[d] = z[p];
  }
}

// One workgroup per key (mock b, tile t): unit weights, real FP64 field,
// cell += W / V_cell (field.cpp:569-1112 rules, as k_assign with WM_UNIT).
template <int ORDER>
static __global__ void __launch_bounds__(256)
k_assign_tiled_batch(double2* __restrict__ mesh, long long nm, const double* __restrict__ sx,
                     const double* __restrict__ sy, const double* __restrict__ sz,
                     const unsigned int* __restrict__ koff, Geom g, TileGeom tg) {
  __shared__ double acc[TILE_CELLS];
  const long long nk = tg.ntiles + 1;
  const long long key = blockIdx.x;
  const int b = int(key / nk);
  const long long t = key % nk;
  const unsigned int p0 = koff[key], p1 = koff[key + 1];
  if (p0 == p1) return;
  double2* out = mesh + (long long)b * nm;
  const double inv_vol_cell = 1 / g.vol_cell;
  if (t == tg.ntiles) {
    // Base cell outside the mesh: the direct rule (flattened-index guard).
    for (unsigned int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
      const double pos[3] = {sx[p], sy[p], sz[p]};
      int ix[3][ORDER];
      double wn[3][ORDER], dw[3][ORDER];
#pragma unroll
      for (int a = 0; a < 3; a++)
        stencil_1d<ORDER>(pos[a], g.n[a], g.L[a], ix[a], wn[a], dw[a]);
#pragma unroll
      for (int il = 0; il < ORDER; il++)
#pragma unroll
        for (int jl = 0; jl < ORDER; jl++) {
          const long long row = ((long long)ix[0][il] * g.n[1] + ix[1][jl]) * g.n[2];
          const double wij = inv_vol_cell * wn[0][il] * wn[1][jl];
#pragma unroll
          for (int kl = 0; kl < ORDER; kl++) {
            const long long gid = row + ix[2][kl];
            if (gid < 0 || gid >= g.nmesh) continue;
            unsafeAtomicAdd(reinterpret_cast<double*>(out + gid), wij * wn[2][kl]);
          }
        }
    }
    return;
  }
  for (int c = threadIdx.x; c < TILE_CELLS; c += blockDim.x) acc[c] = 0.;
  int o0[3];
  tile_origin(tg, int(t), o0);
  __syncthreads();
  for (unsigned int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    int lo[3];
    double win[ORDER][3];
    tile_stencil<ORDER>(g, o0, sx[p], sy[p], sz[p], lo, win);
#pragma unroll
    for (int il = 0; il < ORDER; il++)
#pragma unroll
      for (int jl = 0; jl < ORDER; jl++) {
        const int row = ((lo[0] + il) * TILE_H + (lo[1] + jl)) * TILE_H + lo[2];
        const double wij = win[il][0] * win[jl][1];
#pragma unroll
        for (int kl = 0; kl < ORDER; kl++)
          atomicAdd(&acc[row + kl], inv_vol_cell * (wij * win[kl][2]));
      }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < TILE_CELLS; c += blockDim.x) {
    const double v = acc[c];
    if (v == 0.) continue;
    unsafeAtomicAdd(reinterpret_cast<double*>(out + tile_cell(g, o0, c)), v);
  }
}

static __global__ void __launch_bounds__(256)
k_fill_zero_u32(unsigned int* __restrict__ a, long long n) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x)
    a[c] = 0u;
}

// Zero a FP64 array (graph-safe replacement of a memset).
static __global__ void __launch_bounds__(256)
k_fill_zero_d(double* __restrict__ a, long long n) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x)
    a[c] = 0.;
}

static __global__ void __launch_bounds__(256)
k_fill_zero_u64(unsigned long long* __restrict__ a, long long n) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x)
    a[c] = 0ull;
}

// ---- batched loaders (field(q) gives the per-field loader) ---------------

// (v - N_b / V) V_cell for mock b.
struct LoadScaleShiftMocks {
  const double2* src;
  long long stride;
  const double* nmock;
  double vol, vol_cell;
  __device__ LoadScaleShift<double> field(int b) const {
    return LoadScaleShift<double>{src + (long long)b * stride,
                                  -(nmock[b] / vol), vol_cell};
  }
};

// Shell field (mock f / nbins, bin f % nbins).
template <typename T>
struct LoadShellMocks {
  const double2* src;
  long long stride;
  AxisTables tab;
  Geom geo;
  const ShellBand* bands;
  int nbins, ell, m;
  __device__ LoadShell<T, double> field(int f) const {
    return LoadShell<T, double>{src + (long long)(f / nbins) * stride, tab, geo,
                                bands[f % nbins], ell, m};
  }
};

// Packed real shell pair (mock f / nt, bins 2 (f % nt) and 2 (f % nt) + 1).
template <typename T>
struct LoadShellPairMocks {
  const double2* src;
  long long stride;
  AxisTables tab;
  Geom geo;
  const ShellBand* bands;
  int nt, ell, m;
  __device__ LoadShellPair<T, double> field(int f) const {
    const int t = f % nt;
    return LoadShellPair<T, double>{src + (long long)(f / nt) * stride, tab, geo,
                                    bands[2 * t], bands[2 * t + 1], ell, m};
  }
};

// Compensated, 1/V-scaled field of mock b.
template <typename T>
struct LoadDeconvMocks {
  const double2* src;
  long long stride;
  AxisTables tab;
  double scale;
  __device__ LoadDeconv<T, double> field(int b) const {
    return LoadDeconv<T, double>{src + (long long)b * stride, tab, scale};
  }
};

// Shot-noise field of mock b with Sbar = N_b (periodic box).
// fb: dn00 of each mock; N00 differs from it only by N_b at k = 0.
struct LoadTwoptMocks {
  const double2* fa;
  const double2* fb;
  long long stride;
  AxisTables tab;
  const double* nmock;
  double inv_vol;
  __device__ LoadTwopt<double> field(int b) const {
    const long long o = (long long)b * stride;
    return LoadTwopt<double>{fa + o, fb + o, tab, dmk(nmock[b], 0.), inv_vol, nmock[b]};
  }
};

// ---- batched reductions ----------------------------------------------------

template <typename T, int MODE>
static __global__ void __launch_bounds__(256)
k_pair_lds_batch(const c2<T>* __restrict__ FA, int na, long long sa,
                 const c2<T>* __restrict__ FB, int nbf, long long sb,
                 const c2<T>* __restrict__ G, long long sg, long long nmesh,
                 const int2* __restrict__ tiles, int ntiles,
                 double2* __restrict__ part, long long sp) {
  const long long b = blockIdx.y;
  k_pair_lds_body<T, MODE>(FA + b * sa, na, FB + b * sb, nbf, G + b * sg, nmesh, tiles,
                           ntiles, part + b * sp);
}

template <typename V>
static __global__ void __launch_bounds__(RC_COLS * RC_ROWS)
k_reduce_cols_batch(const V* __restrict__ part, int nrows, int width,
                    V* __restrict__ out) {
  const long long b = blockIdx.y;
  k_reduce_cols_body<V>(part + b * nrows * width, nrows, width, out + b * width);
}

// Two-point statistics of mock b (fa = dn00, fb = N00, Sbar = N_b); with
// fb_dc, fb is dn00 and N_b is added at k = 0 to form N00.  Plans are
// three-point and so never interlaced.
static __global__ void __launch_bounds__(RED_THREADS)
k_twopt_fourier_batch(const double2* __restrict__ fa,
                      const double2* __restrict__ fb, long long stride, Geom g,
                      AxisTables tab, const double* __restrict__ nmock, int ell,
                      int m, TwoptMap map, int nb,
                      unsigned long long* __restrict__ counts,
                      double* __restrict__ part, int fb_dc) {
  const long long b = blockIdx.y;
  k_twopt_fourier_body<double>(fa + b * stride, fb + b * stride, g, tab,
                               dmk(nmock[b], 0.), ell, m, map, nb,
                               counts + b * nb, part + b * gridDim.x * nb * 4,
                               fb_dc ? nmock[b] : 0.);
}

static __global__ void __launch_bounds__(RED_THREADS)
k_radial_hist_oct_batch(const double2* __restrict__ t, long long stride, Geom g,
                        int la, int ma, int lb, int mb, double* __restrict__ Q,
                        long long nq) {
  const long long b = blockIdx.y;
  k_radial_hist_oct_body<double>(t + b * stride, g, la, ma, lb, mb, Q + b * 2 * nq);
}

static __global__ void __launch_bounds__(256)
k_shot_pairs_tab_batch(const double* __restrict__ Q, long long nq, double dr,
                       SplineDev sa, SplineDev sb, const double* __restrict__ ka,
                       int nka, const double* __restrict__ kb, int nkb,
                       const int2* __restrict__ pidx, int P, int nch, double vol_cell,
                       double2* __restrict__ part) {
  const long long b = blockIdx.y;
  k_shot_pairs_tab_body(Q + b * 2 * nq, nq, dr, sa, sb, ka, nka, kb, nkb, pidx, P, nch,
                        vol_cell, part + b * gridDim.x * P);
}

// ---- epilogues: per-mock accumulation of the estimator terms ------------

// raw[b][dv] += c V_cell (s + f conj s), s = S[b][idx[dv]] / (n1 n2), for
// the entries dv of one tile group (ent[e] = (dv, idx)).
static __global__ void __launch_bounds__(256)
k_epi_pairs(double2* __restrict__ raw, int dim, const double2* __restrict__ S,
            int width, const int2* __restrict__ ent, int nent,
            const double* __restrict__ n1, const double* __restrict__ n2,
            double coupling, double f_mirror, double vol_cell) {
  const int b = blockIdx.y;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nent;
       e += gridDim.x * blockDim.x) {
    const int dv = ent[e].x;
    const double2 v = S[(long long)b * width + ent[e].y];
    const double d = n1[dv] * n2[dv];
    const double sx = v.x / d, sy = v.y / d;
    double2& r = raw[(long long)b * dim + dv];
    r.x += coupling * vol_cell * (sx + f_mirror * sx);
    r.y += coupling * vol_cell * (sy - f_mirror * sy);
  }
}

// shot[b][dv] += c N_b (the S_ijk term of l1 = l2 = 0, Sbar = N).
static __global__ void __launch_bounds__(256)
k_epi_const(double2* __restrict__ shot, int dim, const double* __restrict__ nmock,
            double coupling) {
  const int b = blockIdx.y;
  for (int dv = blockIdx.x * blockDim.x + threadIdx.x; dv < dim;
       dv += gridDim.x * blockDim.x)
    shot[(long long)b * dim + dv].x += coupling * nmock[b];
}

// Two-point term: shot[b][dv] += s + f conj s, s = c (pk[bb] - sn[bb]) with
// per-bin means of the summed [b][nb][4] statistics (0 for an empty bin).
static __global__ void __launch_bounds__(256)
k_epi_twopt(double2* __restrict__ shot, int dim, const double* __restrict__ sums,
            const unsigned long long* __restrict__ counts, int nb,
            const int* __restrict__ bin_of, double coupling, double f_mirror) {
  const int b = blockIdx.y;
  for (int dv = blockIdx.x * blockDim.x + threadIdx.x; dv < dim;
       dv += gridDim.x * blockDim.x) {
    const int bb = bin_of[dv];
    const unsigned long long c = counts[(long long)b * nb + bb];
    double px = 0., py = 0., sx = 0., sy = 0.;
    if (c != 0ull) {
      const double n = double(c);
      const double* s = sums + ((long long)b * nb + bb) * 4;
      px = s[0] / n; py = s[1] / n; sx = s[2] / n; sy = s[3] / n;
    }
    const double vx = coupling * (px - sx), vy = coupling * (py - sy);
    double2& r = shot[(long long)b * dim + dv];
    r.x += vx + f_mirror * vx;
    r.y += vy - f_mirror * vy;
  }
}

// S_ij_k term: shot[b][dv] += phase (s + f conj s), s = c sij[b][dv].
static __global__ void __launch_bounds__(256)
k_epi_sij(double2* __restrict__ shot, int dim, const double2* __restrict__ sij,
          double coupling, double f_w3j, double2 phase) {
  const int b = blockIdx.y;
  for (int dv = blockIdx.x * blockDim.x + threadIdx.x; dv < dim;
       dv += gridDim.x * blockDim.x) {
    const double2 v = sij[(long long)b * dim + dv];
    const double sx = coupling * v.x, sy = coupling * v.y;
    const double2 w = dmk(sx + f_w3j * sx, sy - f_w3j * sy);
    double2& r = shot[(long long)b * dim + dv];
    r.x += phase.x * w.x - phase.y * w.y;
    r.y += phase.x * w.y + phase.y * w.x;
  }
}

// raw, shot *= norm[b]; entries zeroed first by the caller's fill.
static __global__ void __launch_bounds__(256)
k_epi_norm(double2* __restrict__ raw, double2* __restrict__ shot, int dim,
           const double* __restrict__ norm) {
  const int b = blockIdx.y;
  for (int dv = blockIdx.x * blockDim.x + threadIdx.x; dv < dim;
       dv += gridDim.x * blockDim.x) {
    const double f = norm[b];
    double2& r = raw[(long long)b * dim + dv];
    double2& s = shot[(long long)b * dim + dv];
    r.x *= f; r.y *= f;
    s.x *= f; s.y *= f;
  }
}

}  // namespace trvg
