// common.hpp -- shared host/device definitions of the gfx950 engine.
//
// Layout: every mesh field is an interleaved complex array of
// ngrid[0]*ngrid[1]*ngrid[2] cells, row-major with the z index fastest
// (the index order of MeshField::ret_grid_index,
// src/triumvirate/src/field.cpp:533-538).  T2 is double2 (FP64 path) or
// float2 (FP32 path).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/trv_gpu.h"

namespace trvg {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define TRVG_HIP(expr)                                                      \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess) {                                                 \
      throw ::trvg::Error(TRVG_ERR_DEVICE, std::string(#expr) + " failed: " \
                                               + hipGetErrorString(e_));    \
    }                                                                       \
  } while (0)

#define TRVG_LAUNCH_CHECK() TRVG_HIP(hipGetLastError())

template <typename T> struct C2;
template <> struct C2<double> { using type = double2; };
template <> struct C2<float> { using type = float2; };
template <typename T> using c2 = typename C2<T>::type;

template <typename T>
__host__ __device__ __forceinline__ c2<T> mk(T a, T b) {
  c2<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__host__ __device__ __forceinline__ c2<T> cadd(c2<T> a, c2<T> b) {
  return mk<T>(a.x + b.x, a.y + b.y);
}
template <typename T>
__host__ __device__ __forceinline__ c2<T> csub(c2<T> a, c2<T> b) {
  return mk<T>(a.x - b.x, a.y - b.y);
}
template <typename T>
__host__ __device__ __forceinline__ c2<T> cmul(c2<T> a, c2<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__host__ __device__ __forceinline__ c2<T> cscale(c2<T> a, T s) {
  return mk<T>(a.x * s, a.y * s);
}

__host__ __device__ __forceinline__ double2 dmk(double a, double b) {
  return make_double2(a, b);
}
__host__ __device__ __forceinline__ double2 dmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ double2 dconj(double2 a) {
  return make_double2(a.x, -a.y);
}

// ---------------------------------------------------------------------------
// Mesh geometry (uniform over every kernel)
// ---------------------------------------------------------------------------

struct Geom {
  int n[3];
  long long nmesh;
  double L[3];
  double dk[3];       // 2 pi / L
  double dr[3];       // L / n
  double vol, vol_cell;
};

inline Geom make_geom(const trvg_params& p) {
  Geom g;
  for (int a = 0; a < 3; a++) {
    g.n[a] = p.ngrid[a];
    g.L[a] = p.boxsize[a];
    g.dk[a] = 2. * M_PI / p.boxsize[a];
    g.dr[a] = p.boxsize[a] / p.ngrid[a];
  }
  g.nmesh = (long long)g.n[0] * g.n[1] * g.n[2];
  g.vol = g.L[0] * g.L[1] * g.L[2];
  g.vol_cell = g.vol / double(g.nmesh);
  return g;
}

// Signed FFT frequency / position index (field.cpp:540-562).
__host__ __device__ __forceinline__ int shifted(int i, int n) {
  return (i < n / 2) ? i : i - n;
}

// |v| = sqrt(x*x + y*y + z*z) with every operation IEEE round-to-nearest
// and no contraction, so that bin membership tests agree bit-for-bit with
// the reference's plain FP64 arithmetic (maths.cpp:57-59).
// k_lo <= |k| < k_hi with |k| = sqrt_rn(|k|^2), evaluated on |k|^2:
// k2_lo (k2_hi) is the least double s with sqrt_rn(s) >= k_lo (k_hi), so
// the comparison of |k|^2 selects exactly the same modes (shell_band()).
struct ShellBand {
  double k2_lo, k2_hi;
  __device__ __forceinline__ bool contains(double k2) const {
    return k2 >= k2_lo && k2 < k2_hi;
  }
};

// |v|^2 with the rounding sequence of mag3_exact.
__device__ __forceinline__ double mag2_exact(double x, double y, double z) {
  return __dadd_rn(__dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y)), __dmul_rn(z, z));
}

__device__ __forceinline__ double mag3_exact(double x, double y, double z) {
  double s = __dadd_rn(__dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y)),
                       __dmul_rn(z, z));
  return __dsqrt_rn(s);
}

// Wavevector / separation component: (i or i - n) * step, exactly rounded.
__device__ __forceinline__ double grid_coord(int i, int n, double step) {
  return __dmul_rn(double(shifted(i, n)), step);
}

// ---------------------------------------------------------------------------
// Reduced spherical harmonic  sqrt(4 pi/(2l+1)) conj(Y_lm(v^))
// (maths.cpp:171-220).  Special cases as the reference: 1 for (0,0),
// 0 for |v| < 1e-9, azimuth 0 for r_xy < 1e-9.  The azimuthal factor
// e^{i m phi} is formed as ((x + i y)/r_xy)^m instead of through acos.
// ---------------------------------------------------------------------------

__host__ __device__ inline double sph_plm_dev(int l, int m, double x) {
  double pmm = 1.;
  double somx2 = sqrt((1. - x) * (1. + x));
  for (int i = 1; i <= m; i++) pmm *= -somx2 * sqrt((2. * i - 1.) / (2. * i));
  pmm *= sqrt((2. * m + 1.) / (4. * M_PI));
  if (l == m) return pmm;
  double pmmp1 = x * sqrt(2. * m + 3.) * pmm;
  if (l == m + 1) return pmmp1;
  double pll = 0.;
  for (int ll = m + 2; ll <= l; ll++) {
    double a = sqrt((4. * ll * ll - 1.) / (double(ll * ll) - double(m * m)));
    double b = sqrt(((2. * ll + 1.) * (double((ll - 1) * (ll - 1)) - double(m * m)))
                    / ((2. * ll - 3.) * (double(ll * ll) - double(m * m))));
    pll = a * x * pmmp1 - b * pmm;
    pmm = pmmp1;
    pmmp1 = pll;
  }
  return pll;
}

__host__ __device__ inline double2 ylm_reduced_generic(int ell, int m, double x,
                                                       double y, double z) {
  if (ell == 0 && m == 0) return dmk(1., 0.);
  double r = sqrt(x * x + y * y + z * z);
  if (fabs(r) < 1.e-9) return dmk(0., 0.);
  double mu = z / r;
  double rxy = sqrt(x * x + y * y);
  double c = 1., s = 0.;
  if (fabs(rxy) >= 1.e-9) {
    c = x / rxy;
    s = y / rxy;
  }
  int am = m < 0 ? -m : m;
  double er = 1., ei = 0.;  // (c + i s)^|m|
  for (int q = 0; q < am; q++) {
    double t = er * c - ei * s;
    ei = er * s + ei * c;
    er = t;
  }
  if (m < 0) ei = -ei;  // e^{i m phi}
  double p = sph_plm_dev(ell, am, mu);
  double parity = (m < 0 && (am & 1)) ? -1. : 1.;
  double norm = sqrt(4. * M_PI / (2. * ell + 1.));
  // parity * conj(e^{i m phi} P) * norm
  return dmk(parity * er * p * norm, -parity * ei * p * norm);
}

// Constants of the associated Legendre recurrence for l <= YLM_LMAX, as
// sph_plm_dev forms them (sqrt evaluated at compile time by Newton's
// iteration; within one unit in the last place of the run-time values).
constexpr int YLM_LMAX = 8;

constexpr double cx_sqrt(double x) {
  double r = x > 1. ? x : 1.;
  for (int i = 0; i < 200; i++) {
    const double n = 0.5 * (r + x / r);
    if (n == r) break;
    r = n;
  }
  return r;
}

struct YlmCoef {
  double pm0[YLM_LMAX + 1];                    // prod_i -sqrt((2i-1)/(2i)) sqrt((2m+1)/(4 pi))
  double c1[YLM_LMAX + 1];                     // sqrt(2m + 3)
  double a[YLM_LMAX + 1][YLM_LMAX + 1];        // [l][m] recurrence coefficients
  double b[YLM_LMAX + 1][YLM_LMAX + 1];
  double norm[YLM_LMAX + 1];                   // sqrt(4 pi / (2l + 1))
};

constexpr YlmCoef make_ylm_coef() {
  YlmCoef c{};
  const double pi = 3.14159265358979323846;
  for (int m = 0; m <= YLM_LMAX; m++) {
    double p = 1.;
    for (int i = 1; i <= m; i++) p *= -cx_sqrt((2. * i - 1.) / (2. * i));
    c.pm0[m] = p * cx_sqrt((2. * m + 1.) / (4. * pi));
    c.c1[m] = cx_sqrt(2. * m + 3.);
  }
  for (int l = 0; l <= YLM_LMAX; l++) {
    c.norm[l] = cx_sqrt(4. * pi / (2. * l + 1.));
    for (int m = 0; m + 2 <= l; m++) {
      const double d = double(l * l) - double(m * m);
      c.a[l][m] = cx_sqrt((4. * l * l - 1.) / d);
      c.b[l][m] = cx_sqrt(((2. * l + 1.) * (double((l - 1) * (l - 1)) - double(m * m))) /
                          ((2. * l - 3.) * d));
    }
  }
  return c;
}

constexpr YlmCoef YLM_COEF = make_ylm_coef();

// The same function with one square root and one division: with
// u = (x + i y) / r, sin^|m|(theta) e^{i |m| phi} = u^|m|, so the
// P_l^|m| recurrence runs on P_l^|m| / sin^|m|(theta) in mu = z / r alone
// and the azimuthal factor needs no r_xy (near the z axis, r_xy < 1e-9,
// u = (sin(theta), 0) with the reference's azimuth 0).
__host__ __device__ inline double2 ylm_reduced_dev(int ell, int m, double x,
                                                   double y, double z) {
  if (ell == 0 && m == 0) return dmk(1., 0.);
  if (ell > YLM_LMAX) return ylm_reduced_generic(ell, m, x, y, z);
  const double r = sqrt(x * x + y * y + z * z);
  if (fabs(r) < 1.e-9) return dmk(0., 0.);
  const double ir = 1. / r;
  const double mu = z * ir;
  const int am = m < 0 ? -m : m;
  double er = 1., ei = 0.;
  if (am > 0) {
    const double rxy2 = x * x + y * y;
    double ur = x * ir, ui = y * ir;
    if (rxy2 < 1.e-18) {
      // Azimuth 0 and sin(theta) from mu, as the reference forms them.
      const double mr = z / r;
      ur = sqrt((1. - mr) * (1. + mr));
      ui = 0.;
    }
    er = ur;
    ei = ui;
    for (int q = 1; q < am; q++) {
      const double t = er * ur - ei * ui;
      ei = er * ui + ei * ur;
      er = t;
    }
    if (m < 0) ei = -ei;
  }
  double p = YLM_COEF.pm0[am];
  if (ell > am) {
    double pmm = p, pmmp1 = mu * YLM_COEF.c1[am] * pmm;
    for (int ll = am + 2; ll <= ell; ll++) {
      const double pll = YLM_COEF.a[ll][am] * mu * pmmp1 - YLM_COEF.b[ll][am] * pmm;
      pmm = pmmp1;
      pmmp1 = pll;
    }
    p = pmmp1;
  }
  const double parity = (m < 0 && (am & 1)) ? -1. : 1.;
  const double f = parity * p * YLM_COEF.norm[ell];
  return dmk(f * er, -f * ei);
}

// y_lm (as ylm_reduced_dev) at (+-x, +-y, +-z) from its value y at
// (x, y, z) >= 0: x -> -x maps e^{i m phi} to (-1)^m conj(e^{i m phi}),
// y -> -y to its conjugate, z -> -z multiplies P_l^|m| by (-1)^(l + |m|).
// Negation is exact, so the result equals a direct evaluation.
__host__ __device__ inline double2 ylm_mirror(double2 y, bool fx, bool fy, bool fz,
                                              int ell, int m) {
  const int am = m < 0 ? -m : m;
  double sgn = 1.;
  if (fx && (am & 1)) sgn = -sgn;
  if (fz && ((ell + am) & 1)) sgn = -sgn;
  const bool cj = fx != fy;
  return dmk(sgn * y.x, cj ? -(sgn * y.y) : sgn * y.y);
}

// ---------------------------------------------------------------------------
// Spherical Bessel j_l (FP64), same algorithm as the host node generator:
// ascending series for x < 1 + l, upward recurrence otherwise.
// ---------------------------------------------------------------------------

__host__ __device__ inline double sph_bessel_dev(int l, double x) {
  if (x == 0.) return (l == 0) ? 1. : 0.;
  if (x < 1. + l) {
    double pref = 1.;
    for (int i = 1; i <= l; i++) pref *= x / (2. * i + 1.);
    double term = 1., sum = 1., h = -0.5 * x * x;
    for (int k = 1; k < 200; k++) {
      term *= h / (k * (2. * l + 2. * k + 1.));
      sum += term;
      if (fabs(term) < 1.e-18 * fabs(sum)) break;
    }
    return pref * sum;
  }
  double sx = sin(x), cx = cos(x);
  double j0 = sx / x;
  if (l == 0) return j0;
  double j1 = sx / (x * x) - cx / x;
  for (int n = 1; n < l; n++) {
    double j2 = (2. * n + 1.) / x * j1 - j0;
    j0 = j1;
    j1 = j2;
  }
  return j1;
}

// Natural cubic spline of j_l on x_i = 0.05 i (maths.cpp:309-375).
// y: node values; c: second derivative / 2 at the nodes.
struct SplineDev {
  // Per interval i: (y_i, b_i, c_i, d_i) of the cubic on [x_i, x_i+1)
  // (b, d as GSL's cspline evaluation forms them), and an FP32 copy.
  const double4* cf;
  const float4* cf32;
  int n;          // number of nodes
  int ell;
  double step;    // 0.05
  double inv_step;
  double split;   // max(1000, l^2)

  // Interval i with x_i <= x < x_{i+1}, x_i = 0 + step * i (exact: the
  // estimate from x / step is corrected against the node abscissae).
  __device__ __forceinline__ int interval(double x) const {
    int i = int(x * inv_step);
    if (i > n - 2) i = n - 2;
    if (i < 0) i = 0;
    if (x < __dmul_rn(step, double(i)) && i > 0) i--;
    else if (i + 1 < n - 1 && x >= __dmul_rn(step, double(i + 1))) i++;
    return i;
  }

  __device__ double eval(double x) const {
    if (x >= split) return sph_bessel_dev(ell, x);
    const int i = interval(x);
    const double delx = x - __dmul_rn(step, double(i));
    const double4 q = cf[i];
    return q.x + delx * (q.y + delx * (q.z + delx * q.w));
  }

  // FP32 evaluation (FP32 fields): the interval from the FP64 argument, the
  // cubic in FP32.  At an interval edge the neighbouring cubic is used,
  // which the spline's continuity makes equivalent to FP32 rounding.
  __device__ float eval32(double x) const {
    if (x >= split) return float(sph_bessel_dev(ell, x));
    int i = int(x * inv_step);
    if (i > n - 2) i = n - 2;
    if (i < 0) i = 0;
    const float dl = float(x - __dmul_rn(step, double(i)));
    const float4 q = cf32[i];
    return q.x + dl * (q.y + dl * (q.z + dl * q.w));
  }
};

// ---------------------------------------------------------------------------
// Per-axis tables for the assignment window W(k) and the aliasing C1(k)
// (field.cpp:1114-1201, 3403-3564), filled on the host in FP64.
// ---------------------------------------------------------------------------

struct AxisTables {
  const double* sinc[3];   // sin(u)/u (1 at u = 0), u = pi * shifted(i) / n
  const double* sin2[3];   // sin(u)^2 (0 at u = 0)
  const double* u[3];      // u itself (for the interlaced iso-approximation)
  int order;
  int interlace;
};

__device__ __forceinline__ double window_at(const AxisTables& t, int i, int j,
                                            int k) {
  double wk = t.sinc[0][i] * t.sinc[1][j] * t.sinc[2][k];
  double w = wk;
  for (int q = 1; q < t.order; q++) w *= wk;
  return w;
}

__device__ __forceinline__ double c1_axis(double s2, int order) {
  if (order == 1) return 1.;
  if (order == 2) return 1. - 2. / 3. * s2;
  if (order == 3) return 1. - s2 + 2. / 15. * s2 * s2;
  return 1. - 4. / 3. * s2 + 2. / 5. * s2 * s2 - 4. / 315. * s2 * s2 * s2;
}

__device__ __forceinline__ double alias_at(const AxisTables& t, int i, int j,
                                           int k) {
  if (!t.interlace) {
    return c1_axis(t.sin2[0][i], t.order) * c1_axis(t.sin2[1][j], t.order)
           * c1_axis(t.sin2[2][k], t.order);
  }
  double ux = t.u[0][i], uy = t.u[1][j], uz = t.u[2][k];
  double uh = sqrt(ux * ux + uy * uy + uz * uz) / 2.;
  double s2h = sin(uh) * sin(uh);
  double c2h = cos(uh) * cos(uh);
  double cc = c2h;
  for (int q = 1; q < t.order; q++) cc *= c2h;
  return cc * (t.order == 1 ? 1. : c1_axis(s2h, t.order));
}

}  // namespace trvg
