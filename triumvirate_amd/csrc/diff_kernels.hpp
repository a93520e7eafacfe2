// diff_kernels.hpp -- kernels of the differentiable periodic-box bispectrum
// (B_000 diag form): real-mesh assignment and its adjoint, the fused
// loaders of the adjoint transforms, and the shell accumulations.
//
// Forward (the reference chain of compute_bispec_in_gpp_box,
// threept.cpp:1473-2188, restricted to l1 = l2 = L = 0):
//   D(k)   = V_cell FFT[delta](k)
//   G(x)   = IFFT[D / (W V)](x)
//   F_b(x) = IFFT[shell_b D / W](x)
//   B_b    = norm V_cell / n_b^2  sum_x F_b(x)^2 G(x)
// Adjoint, with a_b = dL/dB_b norm V_cell / n_b^2 and both k-space masks
// even in k:
//   dL/ddelta = V_cell IFFT[ sum_b shell_b / W  FFT[2 a_b F_b G]
//                            + 1 / (W V)        FFT[sum_b a_b F_b^2] ].
#pragma once

#include "kernels.hpp"

namespace trvg {

// Assignment stencil of one axis (the index and weight rules of k_assign,
// field.cpp:569-1112) with dW/ds, s the in-cell offset in cell units.
template <int ORDER>
__device__ __forceinline__ void stencil_1d(double pos, int ng, double L,
                                           int idx_out[ORDER],
                                           double win[ORDER],
                                           double dwin[ORDER]) {
  const double loc = __ddiv_rn(__dmul_rn(double(ng), pos), L);
  int idx = int(loc);
  const double s = __dadd_rn(loc, -double(idx));
  auto nxt = [ng](int i) { return (i == ng - 1) ? 0 : i + 1; };
  auto prv = [ng](int i) { return (i == 0) ? ng - 1 : i - 1; };
  if (ORDER == 1) {
    if (s >= 0.5) idx = nxt(idx);
    idx_out[0] = idx;
    win[0] = 1.;
    dwin[0] = 0.;
  } else if (ORDER == 2) {
    idx_out[0] = idx;
    idx_out[1] = nxt(idx);
    win[0] = 1. - s;
    win[1] = s;
    dwin[0] = -1.;
    dwin[1] = 1.;
  } else if (ORDER == 3) {
    if (s < 0.5) {
      idx_out[0] = prv(idx);
      idx_out[1] = idx;
      idx_out[2] = nxt(idx);
      win[0] = 0.5 * (0.5 - s) * (0.5 - s);
      win[1] = 0.75 - s * s;
      win[2] = 0.5 * (0.5 + s) * (0.5 + s);
      dwin[0] = -(0.5 - s);
      dwin[1] = -2. * s;
      dwin[2] = 0.5 + s;
    } else {
      idx_out[0] = idx;
      idx_out[1] = nxt(idx);
      idx_out[2] = nxt(idx_out[1]);
      const double sp = 1. - s;
      win[0] = 0.5 * (0.5 + sp) * (0.5 + sp);
      win[1] = 0.75 - sp * sp;
      win[2] = 0.5 * (0.5 - sp) * (0.5 - sp);
      dwin[0] = -(0.5 + sp);
      dwin[1] = 2. * sp;
      dwin[2] = 0.5 - sp;
    }
  } else {
    idx_out[0] = prv(idx);
    idx_out[1] = idx;
    idx_out[2] = nxt(idx);
    idx_out[3] = nxt(idx_out[2]);
    const double sm = 1. - s;
    win[0] = 1. / 6 * sm * sm * sm;
    win[1] = 1. / 6 * (4. - 6. * s * s + 3. * s * s * s);
    win[2] = 1. / 6 * (4. - 6. * sm * sm + 3. * sm * sm * sm);
    win[3] = 1. / 6 * s * s * s;
    dwin[0] = -0.5 * sm * sm;
    dwin[1] = -2. * s + 1.5 * s * s;
    dwin[2] = 2. * sm - 1.5 * sm * sm;
    dwin[3] = 0.5 * s * s;
  }
}

// n(x) += w_p W(x - x_p) / V_cell on a real mesh (positions n x 3,
// row-major; weights may be null for unit weights).
template <typename T, int ORDER>
__global__ void __launch_bounds__(256)
k_assign_real(T* __restrict__ mesh, const double* __restrict__ pos,
              const double* __restrict__ w, long long n, Geom g) {
  const double inv_vol_cell = 1 / g.vol_cell;
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n;
       p += (long long)gridDim.x * blockDim.x) {
    int ix[3][ORDER];
    double wn[3][ORDER], dw[3][ORDER];
#pragma unroll
    for (int a = 0; a < 3; a++)
      stencil_1d<ORDER>(pos[3 * p + a], g.n[a], g.L[a], ix[a], wn[a], dw[a]);
    const double wp = (w ? w[p] : 1.) * inv_vol_cell;
#pragma unroll
    for (int il = 0; il < ORDER; il++)
#pragma unroll
      for (int jl = 0; jl < ORDER; jl++) {
        const long long row = ((long long)ix[0][il] * g.n[1] + ix[1][jl]) * g.n[2];
        const double wij = wp * wn[0][il] * wn[1][jl];
#pragma unroll
        for (int kl = 0; kl < ORDER; kl++) {
          const long long gid = row + ix[2][kl];
          if (gid < 0 || gid >= g.nmesh) continue;
          unsafeAtomicAdd(mesh + gid, T(wij * wn[2][kl]));
        }
      }
  }
}

// Adjoint of k_assign_real: for each particle, gather the mesh cotangent
// through the stencil (grad_w = sum gm W / V_cell) and its derivative
// (grad_pos = w / V_cell sum gm dW/dx).  Either output may be null.
template <typename T, int ORDER>
__global__ void __launch_bounds__(256)
k_assign_real_adjoint(const T* __restrict__ gmesh, const double* __restrict__ pos,
                      const double* __restrict__ w, long long n, Geom g,
                      double* __restrict__ gpos, double* __restrict__ gw) {
  const double inv_vol_cell = 1 / g.vol_cell;
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n;
       p += (long long)gridDim.x * blockDim.x) {
    int ix[3][ORDER];
    double wn[3][ORDER], dw[3][ORDER];
#pragma unroll
    for (int a = 0; a < 3; a++)
      stencil_1d<ORDER>(pos[3 * p + a], g.n[a], g.L[a], ix[a], wn[a], dw[a]);
    double val = 0., dx = 0., dy = 0., dz = 0.;
#pragma unroll
    for (int il = 0; il < ORDER; il++)
#pragma unroll
      for (int jl = 0; jl < ORDER; jl++) {
        const long long row = ((long long)ix[0][il] * g.n[1] + ix[1][jl]) * g.n[2];
#pragma unroll
        for (int kl = 0; kl < ORDER; kl++) {
          const long long gid = row + ix[2][kl];
          if (gid < 0 || gid >= g.nmesh) continue;
          const double v = double(gmesh[gid]);
          val += v * wn[0][il] * wn[1][jl] * wn[2][kl];
          dx += v * dw[0][il] * wn[1][jl] * wn[2][kl];
          dy += v * wn[0][il] * dw[1][jl] * wn[2][kl];
          dz += v * wn[0][il] * wn[1][jl] * dw[2][kl];
        }
      }
    const double wp = (w ? w[p] : 1.) * inv_vol_cell;
    if (gw) gw[p] = val * inv_vol_cell;
    if (gpos) {
      gpos[3 * p] = wp * dx * (double(g.n[0]) / g.L[0]);
      gpos[3 * p + 1] = wp * dy * (double(g.n[1]) / g.L[1]);
      gpos[3 * p + 2] = wp * dz * (double(g.n[2]) / g.L[2]);
    }
  }
}

// Real mesh -> complex transform input (v scale, 0).
template <typename T>
struct LoadReal {
  const T* src;
  T scale;
  __device__ c2<T> operator()(long long g, int, int, int) const {
    return mk<T>(src[g] * scale, T(0));
  }
};

// First-pass loader of FFT[2 a F G]: reads the shell field F and G(x)
// (real parts), returns (2 a F G, 0) and adds a F^2 to gG (each cell is
// loaded exactly once per transform, so the update is race-free).
template <typename T>
struct LoadShellCotangent {
  const c2<T>* F;
  const c2<T>* G;
  T* gG;
  double a;
  __device__ c2<T> operator()(long long g, int, int, int) const {
    const double f = double(F[g].x), gg = double(G[g].x);
    gG[g] = T(double(gG[g]) + a * f * f);
    return mk<T>(T(2. * a * f * gg), T(0));
  }
};

// sum_x Re F(x)^2 Re G(x), per-block FP64 partials.
template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_f2g_partials(const c2<T>* __restrict__ F, const c2<T>* __restrict__ G,
               long long n, double* __restrict__ part, int stride) {
  __shared__ double red[RED_THREADS];
  double s = 0.;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x) {
    const double f = double(F[c].x);
    s += f * f * double(G[c].x);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = RED_THREADS / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long long)blockIdx.x * stride] = red[0];
}

// acc(k) += t(k) / W(k) on the shell of `band` (the predicate of LoadShell).
template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_adj_shell_acc(c2<T>* __restrict__ acc, const c2<T>* __restrict__ t, Geom g,
                AxisTables tab, ShellBand band) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       c < g.nmesh; c += (long long)gridDim.x * blockDim.x) {
    const int k = int(c % g.n[2]);
    const long long ij = c / g.n[2];
    const int j = int(ij % g.n[1]), i = int(ij / g.n[1]);
    const double kx = grid_coord(i, g.n[0], g.dk[0]);
    const double ky = grid_coord(j, g.n[1], g.dk[1]);
    const double kz = grid_coord(k, g.n[2], g.dk[2]);
    if (!band.contains(mag2_exact(kx, ky, kz))) continue;
    const double iw = 1. / window_at(tab, i, j, k);
    c2<T> a = acc[c], v = t[c];
    acc[c] = mk<T>(T(double(a.x) + double(v.x) * iw),
                   T(double(a.y) + double(v.y) * iw));
  }
}

// acc(k) += t(k) s / W(k) on every mode.
template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_adj_all_acc(c2<T>* __restrict__ acc, const c2<T>* __restrict__ t, Geom g,
              AxisTables tab, double s) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       c < g.nmesh; c += (long long)gridDim.x * blockDim.x) {
    const int k = int(c % g.n[2]);
    const long long ij = c / g.n[2];
    const int j = int(ij % g.n[1]), i = int(ij / g.n[1]);
    const double f = s / window_at(tab, i, j, k);
    c2<T> a = acc[c], v = t[c];
    acc[c] = mk<T>(T(double(a.x) + double(v.x) * f),
                   T(double(a.y) + double(v.y) * f));
  }
}

// Zero a real mesh.
template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_zero_real(T* __restrict__ f, long long n) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x)
    f[c] = T(0);
}

// out(x) = scale Re src(x).
template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_real_part(T* __restrict__ out, const c2<T>* __restrict__ src, long long n,
            double scale) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x)
    out[c] = T(double(src[c].x) * scale);
}

}  // namespace trvg
