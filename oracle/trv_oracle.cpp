// trv_oracle.cpp -- CPU FP64 restatement of Triumvirate's three-point
// estimator path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle against which the HIP
// engine is checked: only the test suite, the smoke check and the benchmark's
// CPU-baseline leg load it.  It is never linked into, or called by, the
// product library.
//
// It restates the reference algorithm from its source (file:line cited per
// function, relative to the reference repository root) in plain C++17 +
// OpenMP.  No reference source is copied; FFTW and GSL are replaced by
// routines written here from their mathematical definitions:
//   - complex FFT: iterative radix-2 Cooley-Tukey (power-of-two sizes) with a
//     direct O(n^2) DFT for other sizes; FFTW sign convention
//     (forward e^{-i k x}, backward e^{+i k x}, both unnormalised);
//   - normalised associated Legendre functions: the standard three-term
//     recurrence in l at fixed m (Condon-Shortley phase included);
//   - spherical Bessel j_l: ascending power series for small argument,
//     upward recurrence from j_0, j_1 otherwise;
//   - natural cubic spline: tridiagonal solve for the second-derivative
//     coefficients, piecewise-cubic evaluation on the located interval;
//   - Wigner 3-j symbols: Racah's closed-form sum.
//
// Two schedules are provided.  Schedule 0 ("reference") follows the loop
// structure of src/triumvirate/src/threept.cpp exactly, including its
// redundant recomputation (two shell IFFTs and one shot-noise IFFT per bin
// pair); it is the timed CPU baseline.  Schedule 1 ("dedup") computes each
// distinct shell field once; the results differ only by floating-point
// summation order.
//
// Parity pins: see tests/test_oracle_*.py (reference fixture files, closed
// forms, library cross-checks).

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdarg>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/trv_gpu.h"

namespace {

using cplx = std::complex<double>;
const cplx I_UNIT(0., 1.);

// --------------------------------------------------------------------------
// Errors (mirror trv::sys::InvalidParameterError / InvalidDataError)
// --------------------------------------------------------------------------

struct OracleError : std::runtime_error {
  int code;
  OracleError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw OracleError(code, buf);
}

thread_local std::string g_last_error;
int g_schedule = 0;  // 0 = reference loop structure, 1 = deduplicated

// --------------------------------------------------------------------------
// Maths
// --------------------------------------------------------------------------

// |v| with the operation order of trv::maths::get_vec3d_magnitude
// (src/triumvirate/src/maths.cpp:57-59): sqrt(x*x + y*y + z*z).
inline double mag3(const double v[3]) {
  return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

// Normalised associated Legendre function
//   sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) P_l^m(x),  m >= 0,
// with the Condon-Shortley phase (-1)^m, as returned by GSL's
// gsl_sf_legendre_sphPlm (called at src/triumvirate/src/maths.cpp:211).
double sph_plm(int l, int m, double x) {
  // P~_m^m = (-1)^m sqrt((2m+1)/(4 pi)) sqrt(prod_{i=1..m} (2i-1)/(2i)) (1-x^2)^{m/2}
  double pmm = 1.;
  double somx2 = std::sqrt((1. - x) * (1. + x));
  for (int i = 1; i <= m; i++) {
    pmm *= -somx2 * std::sqrt((2. * i - 1.) / (2. * i));
  }
  pmm *= std::sqrt((2. * m + 1.) / (4. * M_PI));
  if (l == m) return pmm;
  double pmmp1 = x * std::sqrt(2. * m + 3.) * pmm;
  if (l == m + 1) return pmmp1;
  double pll = 0.;
  for (int ll = m + 2; ll <= l; ll++) {
    double a = std::sqrt((4. * ll * ll - 1.) / (double(ll * ll) - double(m * m)));
    double b = std::sqrt(
      ((2. * ll + 1.) * (double((ll - 1) * (ll - 1)) - double(m * m)))
      / ((2. * ll - 3.) * (double(ll * ll) - double(m * m))));
    pll = a * x * pmmp1 - b * pmm;
    pmm = pmmp1;
    pmmp1 = pll;
  }
  return pll;
}

// Reduced spherical harmonic sqrt(4 pi/(2l+1)) conj(Y_lm(r^)), following
// SphericalHarmonicCalculator::calc_reduced_spherical_harmonic
// (src/triumvirate/src/maths.cpp:171-220): 1 for (0,0); 0 if |r| < 1e-9;
// phi = acos(x / r_xy), reflected to 2 pi - phi for y < 0; phi = 0 if
// r_xy < 1e-9.
cplx ylm_reduced(int ell, int m, const double pos[3]) {
  const double eps = 1.e-9;
  if (ell == 0 && m == 0) return cplx(1., 0.);
  double r2 = 0.;
  for (int a = 0; a < 3; a++) r2 += pos[a] * pos[a];
  double r = std::sqrt(r2);
  if (std::fabs(r) < eps) return cplx(0., 0.);
  double mu = pos[2] / r;
  double rxy = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
  double phi = 0.;
  if (std::fabs(rxy) >= eps) {
    phi = std::acos(pos[0] / rxy);
    if (pos[1] < 0.) phi = -phi + 2. * M_PI;
  }
  cplx y = std::exp(I_UNIT * double(m) * phi) * sph_plm(ell, std::abs(m), mu);
  double parity = ((m - std::abs(m)) / 2) % 2 == 0 ? 1. : -1.;
  y = parity * std::conj(y);
  y *= std::sqrt(4. * M_PI / (2. * ell + 1.));
  return y;
}

// Wigner 3-j symbol (j1 j2 j3; m1 m2 m3) for integer arguments, by Racah's
// formula.  Replaces gsl_sf_coupling_3j (src/triumvirate/src/maths.cpp:167-169).
double lfact(int n) { return std::lgamma(double(n) + 1.); }

double wigner_3j(int j1, int j2, int j3, int m1, int m2, int m3) {
  if (m1 + m2 + m3 != 0) return 0.;
  if (j3 < std::abs(j1 - j2) || j3 > j1 + j2) return 0.;
  if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.;
  double pre = 0.5 * (lfact(j1 + j2 - j3) + lfact(j1 - j2 + j3)
                      + lfact(-j1 + j2 + j3) - lfact(j1 + j2 + j3 + 1)
                      + lfact(j1 + m1) + lfact(j1 - m1) + lfact(j2 + m2)
                      + lfact(j2 - m2) + lfact(j3 + m3) + lfact(j3 - m3));
  int kmin = std::max(0, std::max(j2 - j3 - m1, j1 - j3 + m2));
  int kmax = std::min(j1 + j2 - j3, std::min(j1 - m1, j2 + m2));
  double sum = 0.;
  for (int k = kmin; k <= kmax; k++) {
    double t = -(lfact(k) + lfact(j1 + j2 - j3 - k) + lfact(j1 - m1 - k)
                 + lfact(j2 + m2 - k) + lfact(j3 - j2 + m1 + k)
                 + lfact(j3 - j1 - m2 + k));
    double s = (k % 2 == 0) ? 1. : -1.;
    sum += s * std::exp(pre + t);
  }
  int ph = j1 - j2 - m3;
  double sign = (((ph % 2) + 2) % 2 == 0) ? 1. : -1.;
  return sign * sum;
}

const double EPS_COUPLING = 1.e-9;  // src/triumvirate/src/maths.cpp:165

// (2l1+1)(2l2+1)(2L+1) 3j(l1 l2 L; 0 0 0) 3j(l1 l2 L; m1 m2 M),
// trv::calc_coupling_coeff_3pt (src/triumvirate/src/threept.cpp:65-71).
double coupling_3pt(int l1, int l2, int L, int m1, int m2, int M) {
  return double(2 * l1 + 1) * double(2 * l2 + 1) * double(2 * L + 1)
         * wigner_3j(l1, l2, L, 0, 0, 0) * wigner_3j(l1, l2, L, m1, m2, M);
}

// Spherical Bessel function j_l(x), x >= 0, FP64.  Replaces
// gsl_sf_bessel_jl (src/triumvirate/src/maths.cpp:331, :370).
double sph_bessel_j(int l, double x) {
  if (x == 0.) return (l == 0) ? 1. : 0.;
  if (x < 1. + l) {
    // x^l / (2l+1)!! * sum_k (-x^2/2)^k / (k! (2l+3)(2l+5)...(2l+2k+1))
    double pref = 1.;
    for (int i = 1; i <= l; i++) pref *= x / (2. * i + 1.);
    double term = 1., sum = 1., h = -0.5 * x * x;
    for (int k = 1; k < 200; k++) {
      term *= h / (k * (2. * l + 2. * k + 1.));
      sum += term;
      if (std::fabs(term) < 1.e-18 * std::fabs(sum)) break;
    }
    return pref * sum;
  }
  double s = std::sin(x), c = std::cos(x);
  double j0 = s / x;
  if (l == 0) return j0;
  double j1 = s / (x * x) - c / x;
  for (int n = 1; n < l; n++) {
    double j2 = (2. * n + 1.) / x * j1 - j0;
    j0 = j1;
    j1 = j2;
  }
  return j1;
}

// Natural cubic spline of j_l on x_i = 0.05 i, i = 0 .. split/0.05, with
// split = max(1000, l^2); j_l evaluated directly beyond the split.
// trv::maths::SphericalBesselCalculator (src/triumvirate/src/maths.cpp:309-375,
// constants include/maths.hpp:305-306).
struct BesselSpline {
  int order = 0;
  double split = 1000.;
  double step = 0.05;
  std::vector<double> x, y, c;  // c = second derivative / 2 at nodes

  explicit BesselSpline(int ell) : order(ell) {
    split = (split >= double(order * order)) ? split : double(order * order);
    int n = int((split - 0.) / step) + 1;
    x.resize(n);
    y.resize(n);
    for (int i = 0; i < n; i++) {
      x[i] = 0. + step * i;
      y[i] = sph_bessel_j(order, x[i]);
    }
    // Natural boundary: c_0 = c_{n-1} = 0; interior rows
    //   h_{i-1} c_{i-1} + 2 (h_{i-1} + h_i) c_i + h_i c_{i+1}
    //     = 3 ((y_{i+1} - y_i)/h_i - (y_i - y_{i-1})/h_{i-1}).
    c.assign(n, 0.);
    int m = n - 2;
    std::vector<double> diag(m), off(m), rhs(m);
    for (int i = 1; i <= m; i++) {
      double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
      diag[i - 1] = 2. * (h0 + h1);
      off[i - 1] = h1;
      rhs[i - 1] = 3. * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
    }
    // Thomas algorithm (symmetric, diagonally dominant).
    for (int i = 1; i < m; i++) {
      double f = off[i - 1] / diag[i - 1];
      diag[i] -= f * off[i - 1];
      rhs[i] -= f * rhs[i - 1];
    }
    std::vector<double> sol(m);
    sol[m - 1] = rhs[m - 1] / diag[m - 1];
    for (int i = m - 2; i >= 0; i--) {
      sol[i] = (rhs[i] - off[i] * sol[i + 1]) / diag[i];
    }
    for (int i = 0; i < m; i++) c[i + 1] = sol[i];
  }

  double eval(double xv) const {
    if (xv >= split) return sph_bessel_j(order, xv);
    // Interval with x[i] <= xv < x[i+1] (binary search).
    int lo = 0, hi = int(x.size()) - 1;
    while (hi > lo + 1) {
      int mid = (lo + hi) / 2;
      if (x[mid] > xv) hi = mid; else lo = mid;
    }
    double dx = x[lo + 1] - x[lo];
    double dy = y[lo + 1] - y[lo];
    double delx = xv - x[lo];
    double b = dy / dx - dx * (c[lo + 1] + 2. * c[lo]) / 3.;
    double d = (c[lo + 1] - c[lo]) / (3. * dx);
    return y[lo] + delx * (b + delx * (c[lo] + delx * d));
  }
};

// --------------------------------------------------------------------------
// FFT (FP64, in place, unnormalised; sign = -1 forward, +1 backward)
// --------------------------------------------------------------------------

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

struct Fft1d {
  int n;
  int sign;
  std::vector<cplx> tw;       // e^{sign 2 pi i k / n}, k < n
  std::vector<int> bitrev;

  Fft1d(int n_, int sign_) : n(n_), sign(sign_), tw(n_) {
    for (int k = 0; k < n; k++) {
      double a = 2. * M_PI * double(k) / double(n);
      tw[k] = cplx(std::cos(a), sign * std::sin(a));
    }
    if (is_pow2(n)) {
      bitrev.resize(n);
      int lg = 0;
      while ((1 << lg) < n) lg++;
      for (int i = 0; i < n; i++) {
        int r = 0;
        for (int b = 0; b < lg; b++) if (i & (1 << b)) r |= 1 << (lg - 1 - b);
        bitrev[i] = r;
      }
    }
  }

  void run(cplx* a, cplx* scratch) const {
    if (!is_pow2(n)) {
      for (int k = 0; k < n; k++) {
        cplx s = 0.;
        for (int j = 0; j < n; j++) s += a[j] * tw[(long long)j * k % n];
        scratch[k] = s;
      }
      std::copy(scratch, scratch + n, a);
      return;
    }
    for (int i = 0; i < n; i++) {
      int r = bitrev[i];
      if (r > i) std::swap(a[i], a[r]);
    }
    for (int len = 2; len <= n; len <<= 1) {
      int half = len >> 1;
      int step = n / len;
      for (int i = 0; i < n; i += len) {
        for (int j = 0; j < half; j++) {
          cplx t = a[i + j + half] * tw[j * step];
          cplx u = a[i + j];
          a[i + j] = u + t;
          a[i + j + half] = u - t;
        }
      }
    }
  }
};

void fft3d(cplx* data, const int ng[3], int sign) {
  const long long n0 = ng[0], n1 = ng[1], n2 = ng[2];
  // z (contiguous lines)
  {
    Fft1d f(ng[2], sign);
#pragma omp parallel
    {
      std::vector<cplx> scr(n2);
#pragma omp for schedule(static)
      for (long long l = 0; l < n0 * n1; l++) f.run(data + l * n2, scr.data());
    }
  }
  // y and x: gather blocks of contiguous columns into line buffers.
  const int B = 16;
  for (int axis = 1; axis >= 0; axis--) {
    long long n = (axis == 1) ? n1 : n0;
    long long stride = (axis == 1) ? n2 : n1 * n2;
    long long inner = stride;                       // contiguous run
    long long outer = (axis == 1) ? n0 : 1;
    Fft1d f(int(n), sign);
    long long nblk = (inner + B - 1) / B;
#pragma omp parallel
    {
      std::vector<cplx> buf(n * B), scr(n);
#pragma omp for schedule(static)
      for (long long ob = 0; ob < outer * nblk; ob++) {
        long long o = ob / nblk, b = ob % nblk;
        long long base = o * n * stride + b * B;
        int w = int(std::min<long long>(B, inner - b * B));
        for (long long i = 0; i < n; i++)
          for (int t = 0; t < w; t++) buf[t * n + i] = data[base + i * stride + t];
        for (int t = 0; t < w; t++) f.run(buf.data() + t * n, scr.data());
        for (long long i = 0; i < n; i++)
          for (int t = 0; t < w; t++) data[base + i * stride + t] = buf[t * n + i];
      }
    }
  }
}

// --------------------------------------------------------------------------
// Parameters, binning, catalogues
// --------------------------------------------------------------------------

struct Params {
  double boxsize[3];
  int ngrid[3];
  int order;
  bool interlace;
  int ell1, ell2, ELL;
  int shape;
  int idx_bin;
  long long nmesh;
  double volume;
};

Params make_params(const trvg_params* p) {
  Params q;
  for (int a = 0; a < 3; a++) {
    q.boxsize[a] = p->boxsize[a];
    q.ngrid[a] = p->ngrid[a];
    if (q.ngrid[a] <= 0 || !(q.boxsize[a] > 0.))
      fail(TRVG_ERR_PARAM, "Invalid mesh/box: ngrid[%d]=%d boxsize[%d]=%g.",
           a, q.ngrid[a], a, q.boxsize[a]);
  }
  q.order = p->assignment_order;
  if (q.order < 1 || q.order > 4)
    fail(TRVG_ERR_PARAM, "Mesh assignment order must be 1-4: %d.", q.order);
  q.interlace = p->interlace != 0;
  q.ell1 = p->ell1; q.ell2 = p->ell2; q.ELL = p->ELL;
  q.shape = p->shape;
  q.idx_bin = p->idx_bin;
  q.volume = q.boxsize[0] * q.boxsize[1] * q.boxsize[2];
  q.nmesh = (long long)q.ngrid[0] * q.ngrid[1] * q.ngrid[2];
  return q;
}

struct Catalogue {
  long long n = 0;
  std::vector<double> x, y, z, nz, ws, wc, w;
  double wstotal = 0., wtotal = 0.;
};

Catalogue load_catalogue(const trvg_catalogue* c) {
  if (c == nullptr || c->x == nullptr || c->y == nullptr || c->z == nullptr)
    fail(TRVG_ERR_DATA, "Particle data are uninitialised.");
  Catalogue out;
  out.n = c->ntotal;
  long long s = c->stride > 0 ? c->stride : 1;
  out.x.resize(out.n); out.y.resize(out.n); out.z.resize(out.n);
  out.nz.resize(out.n); out.ws.resize(out.n); out.wc.resize(out.n);
  out.w.resize(out.n);
  for (long long i = 0; i < out.n; i++) {
    out.x[i] = c->x[i * s];
    out.y[i] = c->y[i * s];
    out.z[i] = c->z[i * s];
    out.nz[i] = c->nz ? c->nz[i * s] : 0.;
    out.ws[i] = c->ws ? c->ws[i * s] : 1.;
    out.wc[i] = c->wc ? c->wc[i * s] : 1.;
    out.w[i] = out.ws[i] * out.wc[i];  // src/triumvirate/src/particles.cpp:553
  }
  for (long long i = 0; i < out.n; i++) {
    out.wtotal += out.w[i];
    out.wstotal += out.ws[i];
  }
  return out;
}

// --------------------------------------------------------------------------
// Mesh field (src/triumvirate/src/field.cpp MeshField)
// --------------------------------------------------------------------------

struct Mesh {
  const Params& p;
  std::vector<cplx> f, fs;  // field and interlacing shadow
  double dr[3], dk[3], vol, vol_cell;
  int window_order = -1;
  std::vector<double> window;

  explicit Mesh(const Params& p_) : p(p_), f(p_.nmesh) {
    if (p.interlace) fs.assign(p.nmesh, 0.);
    for (int a = 0; a < 3; a++) {
      dr[a] = p.boxsize[a] / p.ngrid[a];
      dk[a] = 2. * M_PI / p.boxsize[a];
    }
    vol = p.volume;
    vol_cell = vol / double(p.nmesh);
  }

  long long gidx(int i, int j, int k) const {
    return ((long long)i * p.ngrid[1] + j) * p.ngrid[2] + k;
  }
  // field.cpp:555-562 (wavevector), :546-553 (position vector)
  void kvec(int i, int j, int k, double kv[3]) const {
    kv[0] = (i < p.ngrid[0] / 2) ? i * dk[0] : (i - p.ngrid[0]) * dk[0];
    kv[1] = (j < p.ngrid[1] / 2) ? j * dk[1] : (j - p.ngrid[1]) * dk[1];
    kv[2] = (k < p.ngrid[2] / 2) ? k * dk[2] : (k - p.ngrid[2]) * dk[2];
  }
  void rvec(int i, int j, int k, double rv[3]) const {
    rv[0] = (i < p.ngrid[0] / 2) ? i * dr[0] : (i - p.ngrid[0]) * dr[0];
    rv[1] = (j < p.ngrid[1] / 2) ? j * dr[1] : (j - p.ngrid[1]) * dr[1];
    rv[2] = (k < p.ngrid[2] / 2) ? k * dr[2] : (k - p.ngrid[2]) * dr[2];
  }

  void reset() {
    std::fill(f.begin(), f.end(), cplx(0., 0.));
    if (p.interlace) std::fill(fs.begin(), fs.end(), cplx(0., 0.));
  }

  // Mesh assignment, field.cpp:569-1112.  The window rules and wrap-around
  // index rules are restated per scheme; the TSC shadow pass keeps the
  // reference's index rule for s >= 0.5 (field.cpp:944-950).
  void assign_pass(const Catalogue& c, const std::vector<cplx>& wgt,
                   std::vector<cplx>& out, bool shadow) {
    const int order = p.order;
    const double inv_vol_cell = 1 / vol_cell;
    const long long nmesh = p.nmesh;
#pragma omp parallel for schedule(static)
    for (long long pid = 0; pid < c.n; pid++) {
      int ijk[4][3];
      double win[4][3];
      const double pos[3] = {c.x[pid], c.y[pid], c.z[pid]};
      for (int a = 0; a < 3; a++) {
        const int ng = p.ngrid[a];
        double loc = ng * pos[a] / p.boxsize[a];
        if (shadow) {
          loc += 0.5;
          if (loc > ng) loc -= ng;
        }
        int idx = int(loc);
        if (order == 1) {
          if (loc - idx >= 0.5) idx = (idx == ng - 1) ? 0 : idx + 1;
          ijk[0][a] = idx;
          win[0][a] = 1.;
        } else if (order == 2) {
          ijk[0][a] = idx;
          ijk[1][a] = (idx == ng - 1) ? 0 : idx + 1;
          double s = loc - idx;
          win[0][a] = 1. - s;
          win[1][a] = s;
        } else if (order == 3) {
          if (loc - idx < 0.5) {
            ijk[0][a] = (idx == 0) ? ng - 1 : idx - 1;
            ijk[1][a] = idx;
            ijk[2][a] = (idx == ng - 1) ? 0 : idx + 1;
          } else if (!shadow) {
            ijk[0][a] = idx;
            ijk[1][a] = (idx == ng - 1) ? 0 : idx + 1;
            ijk[2][a] = (ijk[1][a] == ng - 1) ? 0 : ijk[1][a] + 1;
          } else {
            ijk[0][a] = idx;
            ijk[1][a] = (idx == ng - 1) ? 0 : ijk[0][a] + 1;
            ijk[2][a] = (idx == ng - 1) ? 0 : ijk[1][a] + 1;
          }
          double s = loc - idx;
          if (s < 0.5) {
            win[0][a] = 1. / 2 * (1. / 2 - s) * (1. / 2 - s);
            win[1][a] = 3. / 4 - s * s;
            win[2][a] = 1. / 2 * (1. / 2 + s) * (1. / 2 + s);
          } else {
            s = 1 - s;
            win[0][a] = 1. / 2 * (1. / 2 + s) * (1. / 2 + s);
            win[1][a] = 3. / 4 - s * s;
            win[2][a] = 1. / 2 * (1. / 2 - s) * (1. / 2 - s);
          }
        } else {
          ijk[0][a] = (idx == 0) ? ng - 1 : idx - 1;
          ijk[1][a] = idx;
          ijk[2][a] = (idx == ng - 1) ? 0 : idx + 1;
          ijk[3][a] = (ijk[2][a] == ng - 1) ? 0 : ijk[2][a] + 1;
          double s = loc - idx;
          win[0][a] = 1. / 6 * (1. - s) * (1. - s) * (1. - s);
          win[1][a] = 1. / 6 * (4. - 6. * s * s + 3. * s * s * s);
          win[2][a] = 1. / 6 * (4. - 6. * (1. - s) * (1. - s)
                                + 3. * (1. - s) * (1. - s) * (1. - s));
          win[3][a] = 1. / 6 * s * s * s;
        }
      }
      for (int il = 0; il < order; il++)
        for (int jl = 0; jl < order; jl++)
          for (int kl = 0; kl < order; kl++) {
            long long g = gidx(ijk[il][0], ijk[jl][1], ijk[kl][2]);
            if (0 <= g && g < nmesh) {
              double re = inv_vol_cell * wgt[pid].real() * win[il][0]
                          * win[jl][1] * win[kl][2];
              double im = inv_vol_cell * wgt[pid].imag() * win[il][0]
                          * win[jl][1] * win[kl][2];
              double* cell = reinterpret_cast<double*>(&out[g]);
#pragma omp atomic
              cell[0] += re;
#pragma omp atomic
              cell[1] += im;
            }
          }
    }
  }

  void assign(const Catalogue& c, const std::vector<cplx>& wgt) {
    reset();
    assign_pass(c, wgt, f, false);
    if (p.interlace) assign_pass(c, wgt, fs, true);
  }

  // field.cpp:1208-1223
  void unweighted_field(const Catalogue& c) {
    std::vector<cplx> w(c.n, cplx(1., 0.));
    assign(c, w);
  }
  // field.cpp:1225-1244 (only the real part of the main field is shifted)
  void unweighted_fluctuations(const Catalogue& c) {
    unweighted_field(c);
    double nbar = double(c.n) / vol;
    for (long long g = 0; g < p.nmesh; g++) f[g] -= nbar;
  }

  static std::vector<cplx> ylm_weights(const Catalogue& c, const double* los,
                                       int ell, int m, int wpow, bool conj) {
    std::vector<cplx> w(c.n);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < c.n; i++) {
      double l3[3] = {los[3 * i], los[3 * i + 1], los[3 * i + 2]};
      cplx y = ylm_reduced(ell, m, l3);
      if (conj) y = std::conj(y);
      double wp = (wpow == 1) ? c.w[i] : std::pow(c.w[i], wpow);
      w[i] = cplx(y.real() * wp, y.imag() * wp);
    }
    return w;
  }

  // field.cpp:1246-1323: data minus alpha times randoms.
  void ylm_wgtd_field(const Catalogue& d, const Catalogue& r,
                      const double* los_d, const double* los_r, double alpha,
                      int ell, int m) {
    assign(d, ylm_weights(d, los_d, ell, m, 1, false));
    Mesh fr(p);
    fr.assign(r, ylm_weights(r, los_r, ell, m, 1, false));
    for (long long g = 0; g < p.nmesh; g++) {
      f[g] = cplx(f[g].real() - alpha * fr.f[g].real(),
                  f[g].imag() - alpha * fr.f[g].imag());
    }
    if (p.interlace) {
      for (long long g = 0; g < p.nmesh; g++) {
        fs[g] = cplx(fs[g].real() - alpha * fr.fs[g].real(),
                     fs[g].imag() - alpha * fr.fs[g].imag());
      }
    }
  }

  // field.cpp:1364-1447: conjugated Y_lm, squared weights, plus alpha^2.
  void ylm_wgtd_quad_field(const Catalogue& d, const Catalogue& r,
                           const double* los_d, const double* los_r,
                           double alpha, int ell, int m) {
    assign(d, ylm_weights(d, los_d, ell, m, 2, true));
    Mesh fr(p);
    fr.assign(r, ylm_weights(r, los_r, ell, m, 2, true));
    double a2 = std::pow(alpha, 2);
    for (long long g = 0; g < p.nmesh; g++) {
      f[g] = cplx(f[g].real() + a2 * fr.f[g].real(),
                  f[g].imag() + a2 * fr.f[g].imag());
    }
    if (p.interlace) {
      for (long long g = 0; g < p.nmesh; g++) {
        fs[g] = cplx(fs[g].real() + a2 * fr.fs[g].real(),
                     fs[g].imag() + a2 * fr.fs[g].imag());
      }
    }
  }

  // field.cpp:1325-1362: one catalogue, Y_lm-weighted, times alpha (the
  // main field only: the interlacing shadow keeps the unscaled sum).
  void ylm_wgtd_field_single(const Catalogue& c, const double* los,
                             double alpha, int ell, int m) {
    assign(c, ylm_weights(c, los, ell, m, 1, false));
    for (long long g = 0; g < p.nmesh; g++)
      f[g] = cplx(f[g].real() * alpha, f[g].imag() * alpha);
  }

  // field.cpp:1449-1489: one catalogue, conj(Y_lm) w^2, times alpha^2.
  void ylm_wgtd_quad_field_single(const Catalogue& c, const double* los,
                                  double alpha, int ell, int m) {
    assign(c, ylm_weights(c, los, ell, m, 2, true));
    const double a2 = std::pow(alpha, 2);
    for (long long g = 0; g < p.nmesh; g++)
      f[g] = cplx(f[g].real() * a2, f[g].imag() * a2);
  }

  // field.cpp:1727-1762: f(x) |x|^-(i + j) where |x| >= 1e-6.
  void wide_angle_kernel(int power) {
    for (int i = 0; i < p.ngrid[0]; i++)
      for (int j = 0; j < p.ngrid[1]; j++)
        for (int k = 0; k < p.ngrid[2]; k++) {
          double rv[3];
          rvec(i, j, k, rv);
          double r = mag3(rv);
          if (r < 1.e-6) continue;
          f[gidx(i, j, k)] *= std::pow(r, -power);
        }
  }

  // field.cpp:1496-1655
  void fourier_transform() {
    for (long long g = 0; g < p.nmesh; g++) f[g] *= vol_cell;
    fft3d(f.data(), p.ngrid, -1);
    if (p.interlace) {
      for (long long g = 0; g < p.nmesh; g++) fs[g] *= vol_cell;
      fft3d(fs.data(), p.ngrid, -1);
#pragma omp parallel for collapse(3)
      for (int i = 0; i < p.ngrid[0]; i++)
        for (int j = 0; j < p.ngrid[1]; j++)
          for (int k = 0; k < p.ngrid[2]; k++) {
            long long g = gidx(i, j, k);
            double m0 = (i < p.ngrid[0] / 2) ? double(i) / p.ngrid[0]
                                             : double(i) / p.ngrid[0] - 1;
            double m1 = (j < p.ngrid[1] / 2) ? double(j) / p.ngrid[1]
                                             : double(j) / p.ngrid[1] - 1;
            double m2 = (k < p.ngrid[2] / 2) ? double(k) / p.ngrid[2]
                                             : double(k) / p.ngrid[2] - 1;
            double arg = M_PI * (m0 + m1 + m2);
            double re = f[g].real() + std::cos(arg) * fs[g].real()
                        - std::sin(arg) * fs[g].imag();
            double im = f[g].imag() + std::sin(arg) * fs[g].real()
                        + std::cos(arg) * fs[g].imag();
            f[g] = cplx(re / 2., im / 2.);
          }
    }
  }

  // field.cpp:1657-1720
  void inv_fourier_transform() {
    for (long long g = 0; g < p.nmesh; g++) f[g] /= vol;
    fft3d(f.data(), p.ngrid, +1);
  }

  // field.cpp:1114-1201: pow(prod sinc(pi i/N), order), shifted indices.
  void compute_window(int order) {
    if (window_order == order) return;
    window.resize(p.nmesh);
#pragma omp parallel for collapse(3)
    for (int i = 0; i < p.ngrid[0]; i++)
      for (int j = 0; j < p.ngrid[1]; j++)
        for (int k = 0; k < p.ngrid[2]; k++) {
          int ii = (i < p.ngrid[0] / 2) ? i : i - p.ngrid[0];
          int jj = (j < p.ngrid[1] / 2) ? j : j - p.ngrid[1];
          int kk = (k < p.ngrid[2] / 2) ? k : k - p.ngrid[2];
          double ux = M_PI * ii / double(p.ngrid[0]);
          double uy = M_PI * jj / double(p.ngrid[1]);
          double uz = M_PI * kk / double(p.ngrid[2]);
          double wx = (ii != 0) ? std::sin(ux) / ux : 1.;
          double wy = (jj != 0) ? std::sin(uy) / uy : 1.;
          double wz = (kk != 0) ? std::sin(uz) / uz : 1.;
          window[gidx(i, j, k)] = std::pow(wx * wy * wz, order);
        }
    window_order = order;
  }

  // field.cpp:1764-1785
  void apply_compensation() {
    compute_window(p.order);
    for (long long g = 0; g < p.nmesh; g++) f[g] /= window[g];
  }

  // field.cpp:1792-1906: shell k_lo <= |k| < k_hi, y(k^) f(k)/W(k), backward
  // FFT, divided by the mode count.
  void ifft_ylm_band_limited(const Mesh& src, const std::vector<cplx>& ylm,
                             double k_lo, double k_hi, double& k_eff,
                             int& nmodes) {
    compute_window(p.order);
    double keff = 0.;
    int nm = 0;
#pragma omp parallel for collapse(3) reduction(+ : keff, nm)
    for (int i = 0; i < p.ngrid[0]; i++)
      for (int j = 0; j < p.ngrid[1]; j++)
        for (int k = 0; k < p.ngrid[2]; k++) {
          long long g = gidx(i, j, k);
          double kv[3];
          kvec(i, j, k, kv);
          double km = mag3(kv);
          if (k_lo <= km && km < k_hi) {
            cplx fk = src.f[g];
            fk /= window[g];
            f[g] = ylm[g] * fk;
            keff += km;
            nm++;
          } else {
            f[g] = 0.;
          }
        }
    fft3d(f.data(), p.ngrid, +1);
    for (long long g = 0; g < p.nmesh; g++) f[g] /= double(nm);
    k_eff = keff / double(nm);
    nmodes = nm;
  }

  // field.cpp:1908-2010: j_l(|k| r) y(k^) f(k)/W(k)/V, backward FFT.
  void ifft_sjl_ylm(const Mesh& src, const std::vector<cplx>& ylm,
                    const BesselSpline& sj, double r) {
    compute_window(p.order);
#pragma omp parallel for collapse(3)
    for (int i = 0; i < p.ngrid[0]; i++)
      for (int j = 0; j < p.ngrid[1]; j++)
        for (int k = 0; k < p.ngrid[2]; k++) {
          long long g = gidx(i, j, k);
          double kv[3];
          kvec(i, j, k, kv);
          double km = mag3(kv);
          cplx fk = src.f[g];
          fk /= window[g];
          double jv = sj.eval(km * r);
          cplx yf = ylm[g] * fk;
          f[g] = cplx(jv * yf.real() / vol, jv * yf.imag() / vol);
        }
    fft3d(f.data(), p.ngrid, +1);
  }

  // field.cpp:2017-2065
  double grid_powlaw_norm(const Catalogue& c, int order) {
    std::vector<cplx> w(c.n);
    for (long long i = 0; i < c.n; i++) w[i] = cplx(c.w[i], 0.);
    assign(c, w);
    double vi = 0.;
#pragma omp parallel for reduction(+ : vi)
    for (long long g = 0; g < p.nmesh; g++) vi += std::pow(f[g].real(), order);
    vi *= vol_cell;
    return 1. / vi;
  }
};

// Reduced spherical harmonics on the k-grid and r-grid,
// maths.cpp:222-302.
std::vector<cplx> ylm_table(const Mesh& m, int ell, int mm, bool fourier) {
  std::vector<cplx> out(m.p.nmesh);
#pragma omp parallel for collapse(3)
  for (int i = 0; i < m.p.ngrid[0]; i++)
    for (int j = 0; j < m.p.ngrid[1]; j++)
      for (int k = 0; k < m.p.ngrid[2]; k++) {
        double v[3];
        if (fourier) m.kvec(i, j, k, v); else m.rvec(i, j, k, v);
        out[m.gidx(i, j, k)] = ylm_reduced(ell, mm, v);
      }
  return out;
}

// --------------------------------------------------------------------------
// Field statistics (field.cpp FieldStats)
// --------------------------------------------------------------------------

struct Stats {
  const Params& p;
  Mesh geom;
  std::vector<double> alias;
  bool alias_ini = false;
  std::vector<int> nmodes, npairs;
  std::vector<double> k, r;
  std::vector<cplx> pk, sn, xi;

  explicit Stats(const Params& p_) : p(p_), geom(p_) {}

  // field.cpp:3403-3564 (C1 aliasing; interlaced iso-approximation)
  void compute_alias() {
    if (alias_ini) return;
    alias.resize(p.nmesh);
#pragma omp parallel for collapse(3)
    for (int i = 0; i < p.ngrid[0]; i++)
      for (int j = 0; j < p.ngrid[1]; j++)
        for (int k = 0; k < p.ngrid[2]; k++) {
          int ii = (i < p.ngrid[0] / 2) ? i : i - p.ngrid[0];
          int jj = (j < p.ngrid[1] / 2) ? j : j - p.ngrid[1];
          int kk = (k < p.ngrid[2] / 2) ? k : k - p.ngrid[2];
          double ux = M_PI * ii / double(p.ngrid[0]);
          double uy = M_PI * jj / double(p.ngrid[1]);
          double uz = M_PI * kk / double(p.ngrid[2]);
          double val;
          if (p.interlace) {
            double uh = std::sqrt(ux * ux + uy * uy + uz * uz) / 2.;
            double s2h = std::sin(uh) * std::sin(uh);
            double c2h = std::cos(uh) * std::cos(uh);
            double cc = std::pow(c2h, p.order), cs = 1.;
            if (p.order == 2) cs = 1. - 2. / 3. * s2h;
            else if (p.order == 3) cs = 1. - s2h + 2. / 15. * s2h * s2h;
            else if (p.order == 4)
              cs = 1. - 4. / 3. * s2h + 2. / 5. * s2h * s2h
                   - 4. / 315. * s2h * s2h * s2h;
            val = cc * cs;
          } else {
            double sx = (ii != 0) ? std::sin(ux) * std::sin(ux) : 0.;
            double sy = (jj != 0) ? std::sin(uy) * std::sin(uy) : 0.;
            double sz = (kk != 0) ? std::sin(uz) * std::sin(uz) : 0.;
            if (p.order == 1) {
              val = 1.;
            } else if (p.order == 2) {
              val = (1. - 2. / 3. * sx) * (1. - 2. / 3. * sy) * (1. - 2. / 3. * sz);
            } else if (p.order == 3) {
              val = (1. - sx + 2. / 15. * sx * sx) * (1. - sy + 2. / 15. * sy * sy)
                    * (1. - sz + 2. / 15. * sz * sz);
            } else {
              val = (1. - 4. / 3. * sx + 2. / 5. * sx * sx - 4. / 315. * sx * sx * sx)
                    * (1. - 4. / 3. * sy + 2. / 5. * sy * sy - 4. / 315. * sy * sy * sy)
                    * (1. - 4. / 3. * sz + 2. / 5. * sz * sz - 4. / 315. * sz * sz * sz);
            }
          }
          alias[geom.gidx(i, j, k)] = val;
        }
    alias_ini = true;
  }

  // Window used for f_a f_b^* (pk) and for the shot-noise amplitude (sn):
  // C1 without interlacing, W_a W_b with it (field.cpp:2554-2557).
  double win_pk(long long g, int i, int j, int k) {
    if (!p.interlace) return alias[g];
    (void)i; (void)j; (void)k;
    return geom.window[g] * geom.window[g];
  }

  // field.cpp:2511-2703: two-level fine binning (1e6 bins of 1e-5).
  void twopt_fourier(const Mesh& fa, const Mesh& fb, cplx sn_amp, int ell,
                     int m, const trvg_binning& kb) {
    compute_alias();
    if (p.interlace) geom.compute_window(p.order);
    const int n_sample = 1000000;
    const double dk_sample = 1.e-5;
    std::vector<int> nm_s(n_sample, 0);
    std::vector<double> k_s(n_sample, 0.), pr(n_sample, 0.), pi(n_sample, 0.),
        sr(n_sample, 0.), si(n_sample, 0.);
#pragma omp parallel for collapse(3)
    for (int i = 0; i < p.ngrid[0]; i++)
      for (int j = 0; j < p.ngrid[1]; j++)
        for (int k2 = 0; k2 < p.ngrid[2]; k2++) {
          long long g = geom.gidx(i, j, k2);
          double kv[3];
          geom.kvec(i, j, k2, kv);
          double km = mag3(kv);
          int ik = int(km / dk_sample);
          if (0 <= ik && ik < n_sample) {
            cplx pkm = fa.f[g] * std::conj(fb.f[g]);
            cplx snm = sn_amp * alias[g];
            double w = win_pk(g, i, j, k2);
            pkm /= w;
            snm /= w;
            cplx y = ylm_reduced(ell, m, kv);
            pkm *= y;
            snm *= y;
#pragma omp atomic
            nm_s[ik]++;
#pragma omp atomic
            k_s[ik] += km;
#pragma omp atomic
            pr[ik] += pkm.real();
#pragma omp atomic
            pi[ik] += pkm.imag();
#pragma omp atomic
            sr[ik] += snm.real();
#pragma omp atomic
            si[ik] += snm.imag();
          }
        }
    int nb = kb.num_bins;
    nmodes.assign(nb, 0); k.assign(nb, 0.); pk.assign(nb, 0.); sn.assign(nb, 0.);
    for (int b = 0; b < nb; b++) {
      double lo = kb.bin_edges[b], hi = kb.bin_edges[b + 1];
      for (int i = 0; i < n_sample; i++) {
        double kk = i * dk_sample;
        if (lo <= kk && kk < hi) {
          nmodes[b] += nm_s[i];
          k[b] += k_s[i];
          pk[b] += cplx(pr[i], pi[i]);
          sn[b] += cplx(sr[i], si[i]);
        }
      }
      if (nmodes[b] != 0) {
        k[b] /= double(nmodes[b]);
        pk[b] /= double(nmodes[b]);
        sn[b] /= double(nmodes[b]);
      } else {
        k[b] = kb.bin_centres[b];
        pk[b] = 0.;
        sn[b] = 0.;
      }
    }
  }

  // (f_a f_b^* / C1 - S C1 / C1) / V, then backward FFT
  // (field.cpp:3007-3025, 3271-3289).
  std::vector<cplx> twopt_config_field(const Mesh& fa, const Mesh& fb,
                                       cplx sn_amp) {
    compute_alias();
    if (p.interlace) geom.compute_window(p.order);
    std::vector<cplx> t(p.nmesh);
#pragma omp parallel for collapse(3)
    for (int i = 0; i < p.ngrid[0]; i++)
      for (int j = 0; j < p.ngrid[1]; j++)
        for (int k2 = 0; k2 < p.ngrid[2]; k2++) {
          long long g = geom.gidx(i, j, k2);
          cplx pkm = fa.f[g] * std::conj(fb.f[g]);
          cplx snm = sn_amp * alias[g];
          double w = win_pk(g, i, j, k2);
          pkm /= w;
          snm /= w;
          pkm -= snm;
          t[g] = cplx(pkm.real() / p.volume, pkm.imag() / p.volume);
        }
    fft3d(t.data(), p.ngrid, +1);
    return t;
  }

  // field.cpp:2705-2945: the shot-noise-subtracted product transformed
  // back, then 1e6 fine bins of width 0.1 by int(r / 0.1), Y_lm(r^)
  // weighted, assigned to coarse bins by the fine lower edge, mean per bin
  // (empty bins: centre and 0).
  void twopt_config(const Mesh& fa, const Mesh& fb, cplx sn_amp, int ell,
                    int m, const trvg_binning& rb) {
    std::vector<cplx> t = twopt_config_field(fa, fb, sn_amp);
    const int n_sample = 1000000;
    const double dr_sample = 1.e-1;
    std::vector<int> np_s(n_sample, 0);
    std::vector<double> r_s(n_sample, 0.), xr(n_sample, 0.), xim(n_sample, 0.);
#pragma omp parallel for collapse(3)
    for (int i = 0; i < p.ngrid[0]; i++)
      for (int j = 0; j < p.ngrid[1]; j++)
        for (int k2 = 0; k2 < p.ngrid[2]; k2++) {
          long long g = geom.gidx(i, j, k2);
          double rv[3];
          geom.rvec(i, j, k2, rv);
          double rm = mag3(rv);
          int ir = int(rm / dr_sample);
          if (0 <= ir && ir < n_sample) {
            cplx xp = t[g] * ylm_reduced(ell, m, rv);
#pragma omp atomic
            np_s[ir]++;
#pragma omp atomic
            r_s[ir] += rm;
#pragma omp atomic
            xr[ir] += xp.real();
#pragma omp atomic
            xim[ir] += xp.imag();
          }
        }
    int nb = rb.num_bins;
    npairs.assign(nb, 0); r.assign(nb, 0.); xi.assign(nb, 0.);
    for (int b = 0; b < nb; b++) {
      double lo = rb.bin_edges[b], hi = rb.bin_edges[b + 1];
      for (int i = 0; i < n_sample; i++) {
        double rr = i * dr_sample;
        if (lo <= rr && rr < hi) {
          npairs[b] += np_s[i];
          r[b] += r_s[i];
          xi[b] += cplx(xr[i], xim[i]);
        }
      }
      if (npairs[b] != 0) {
        r[b] /= double(npairs[b]);
        xi[b] /= double(npairs[b]);
      } else {
        r[b] = rb.bin_centres[b];
        xi[b] = 0.;
      }
    }
  }

  // field.cpp:2947-3196: fine bins of width 1 (1e5 bins), mean, then the
  // extra (-1)^(l1+l2) / (V_cell npairs) factor.
  void shotnoise_3pcf(const Mesh& fa, const Mesh& fb,
                      const std::vector<cplx>& ya, const std::vector<cplx>& yb,
                      cplx sn_amp, const trvg_binning& rb) {
    std::vector<cplx> t = twopt_config_field(fa, fb, sn_amp);
    const int n_sample = 100000;
    const double dr_sample = 1.;
    std::vector<int> np_s(n_sample, 0);
    std::vector<double> r_s(n_sample, 0.), xr(n_sample, 0.), xim(n_sample, 0.);
#pragma omp parallel for collapse(3)
    for (int i = 0; i < p.ngrid[0]; i++)
      for (int j = 0; j < p.ngrid[1]; j++)
        for (int k2 = 0; k2 < p.ngrid[2]; k2++) {
          long long g = geom.gidx(i, j, k2);
          double rv[3];
          geom.rvec(i, j, k2, rv);
          double rm = mag3(rv);
          int ir = int(rm / dr_sample);
          if (0 <= ir && ir < n_sample) {
            cplx xp = t[g];
            xp *= ya[g] * yb[g];
#pragma omp atomic
            np_s[ir]++;
#pragma omp atomic
            r_s[ir] += rm;
#pragma omp atomic
            xr[ir] += xp.real();
#pragma omp atomic
            xim[ir] += xp.imag();
          }
        }
    int nb = rb.num_bins;
    npairs.assign(nb, 0); r.assign(nb, 0.); xi.assign(nb, 0.);
    for (int b = 0; b < nb; b++) {
      double lo = rb.bin_edges[b], hi = rb.bin_edges[b + 1];
      for (int i = 0; i < n_sample; i++) {
        double rr = i * dr_sample;
        if (lo <= rr && rr < hi) {
          npairs[b] += np_s[i];
          r[b] += r_s[i];
          xi[b] += cplx(xr[i], xim[i]);
        }
      }
      if (npairs[b] != 0) {
        r[b] /= double(npairs[b]);
        xi[b] /= double(npairs[b]);
      } else {
        r[b] = rb.bin_centres[b];
        xi[b] = 0.;
      }
    }
    double nf = 1 / geom.vol_cell * std::pow(-1, p.ell1 + p.ell2);
    for (int b = 0; b < nb; b++) {
      if (npairs[b] != 0) xi[b] *= nf / double(npairs[b]);
    }
  }

  // field.cpp:3198-3396: V_cell sum_x j_a(k_a r) j_b(k_b r) y_a y_b t(x).
  cplx shotnoise_bispec_from_t(const std::vector<cplx>& t,
                               const std::vector<cplx>& ya,
                               const std::vector<cplx>& yb,
                               const BesselSpline& sa, const BesselSpline& sb,
                               double ka, double kb) {
    double sr = 0., si = 0.;
#pragma omp parallel for collapse(3) reduction(+ : sr, si)
    for (int i = 0; i < p.ngrid[0]; i++)
      for (int j = 0; j < p.ngrid[1]; j++)
        for (int k2 = 0; k2 < p.ngrid[2]; k2++) {
          long long g = geom.gidx(i, j, k2);
          double rv[3];
          geom.rvec(i, j, k2, rv);
          double rm = mag3(rv);
          double ja = sa.eval(ka * rm);
          double jb = sb.eval(kb * rm);
          cplx s = t[g];
          s *= ja * jb * ya[g] * yb[g];
          sr += s.real();
          si += s.imag();
        }
    return cplx(sr, si) * geom.vol_cell;
  }
};

// --------------------------------------------------------------------------
// Estimator drivers (src/triumvirate/src/threept.cpp)
// --------------------------------------------------------------------------

struct Triplet { int m1, m2, M; };

bool is_inverse(const Triplet& a, const Triplet& b) {
  return a.m1 + b.m1 == 0 && a.m2 + b.m2 == 0 && a.M + b.M == 0;
}

int result_dim(int shape, int nb, int idx_bin) {
  switch (shape) {
    case TRVG_SHAPE_DIAG: case TRVG_SHAPE_ROW: return nb;
    case TRVG_SHAPE_OFFDIAG: return nb - std::abs(idx_bin);
    case TRVG_SHAPE_FULL: return nb * nb;
    case TRVG_SHAPE_TRIU: return nb * (nb + 1) / 2;
    default: fail(TRVG_ERR_PARAM, "Three-point statistic form is not recognised.");
  }
}

// (row bin, column bin) of each data-vector entry, threept.cpp:465-770.
void pair_bins(int shape, int nb, int idx_bin, std::vector<int>& ra,
               std::vector<int>& cb) {
  int dim = result_dim(shape, nb, idx_bin);
  ra.resize(dim);
  cb.resize(dim);
  if (shape == TRVG_SHAPE_DIAG) {
    for (int d = 0; d < dim; d++) ra[d] = cb[d] = d;
  } else if (shape == TRVG_SHAPE_OFFDIAG) {
    for (int d = 0; d < dim; d++) {
      if (idx_bin >= 0) { ra[d] = d; cb[d] = d + std::abs(idx_bin); }
      else { ra[d] = d + std::abs(idx_bin); cb[d] = d; }
    }
  } else if (shape == TRVG_SHAPE_ROW) {
    for (int d = 0; d < dim; d++) { ra[d] = idx_bin; cb[d] = d; }
  } else if (shape == TRVG_SHAPE_FULL) {
    for (int r = 0; r < nb; r++)
      for (int c = 0; c < nb; c++) { ra[r * nb + c] = r; cb[r * nb + c] = c; }
  } else {
    for (int r = 0; r < nb; r++)
      for (int c = r; c < nb; c++) {
        int d = (2 * nb - r + 1) * r / 2 + (c - r);
        ra[d] = r;
        cb[d] = c;
      }
  }
}

void validate_coupling(const Params& p) {
  if (std::fabs(wigner_3j(p.ell1, p.ell2, p.ELL, 0, 0, 0)) < EPS_COUPLING)
    fail(TRVG_ERR_PARAM,
         "Specified three-point correlator multipole vanishes identically "
         "owing to zero-valued Wigner 3-j symbol.");
}

cplx ipow(int n) { return std::pow(I_UNIT, n); }  // threept.cpp:271-272

struct Output {
  int dim;
  std::vector<double> b1, b2, e1, e2;
  std::vector<int> n1, n2;
  std::vector<cplx> raw, shot;
  explicit Output(int d)
      : dim(d), b1(d, 0.), b2(d, 0.), e1(d, 0.), e2(d, 0.), n1(d, 0), n2(d, 0),
        raw(d, 0.), shot(d, 0.) {}
};

// Sum over the mesh of F_a F_b G (threept.cpp:495-511).
cplx triple_sum(const Mesh& a, const Mesh& b, const Mesh& g) {
  double sr = 0., si = 0.;
#pragma omp parallel for reduction(+ : sr, si)
  for (long long i = 0; i < a.p.nmesh; i++) {
    cplx v = a.f[i] * b.f[i] * g.f[i];
    sr += v.real();
    si += v.imag();
  }
  return cplx(sr, si);
}

// Particle shot-noise amplitude, threept.cpp:156-236.
cplx sbar_lm(const Catalogue& c, const double* los, int ell, int m) {
  double sr = 0., si = 0.;
#pragma omp parallel for reduction(+ : sr, si)
  for (long long i = 0; i < c.n; i++) {
    double l3[3] = {los[3 * i], los[3 * i + 1], los[3 * i + 2]};
    cplx y = ylm_reduced(ell, m, l3) * std::pow(c.w[i], 3);
    sr += y.real();
    si += y.imag();
  }
  return cplx(sr, si);
}

// The common survey / box bispectrum body.  `survey` selects the
// survey-type field builders (threept.cpp:248-1012) or the periodic-box ones
// (threept.cpp:1473-2188).
Output bispec_impl(bool survey, const Catalogue& d, const Catalogue* r,
                   const double* los_d, const double* los_r, const Params& p,
                   const trvg_binning& kb, double norm_factor) {
  validate_coupling(p);
  double alpha = survey ? d.wstotal / r->wstotal : 1.;
  cplx factor_phase = ipow(p.ell1 + p.ell2);
  int nb = kb.num_bins;
  int dim = result_dim(p.shape, nb, p.idx_bin);
  std::vector<int> ra, cb;
  pair_bins(p.shape, nb, p.idx_bin, ra, cb);
  Output out(dim);

  Mesh dn00(p);
  if (survey) dn00.ylm_wgtd_field(d, *r, los_d, los_r, alpha, 0, 0);
  else dn00.unweighted_fluctuations(d);
  dn00.fourier_transform();
  double vol_cell = dn00.vol_cell;

  Mesh N00(p);
  if (survey) N00.ylm_wgtd_quad_field(d, *r, los_d, los_r, alpha, 0, 0);
  else N00.unweighted_field(d);
  N00.fourier_transform();

  BesselSpline sja(p.ell1), sjb(p.ell2);
  Stats stats(p);

  std::vector<Triplet> done;
  int count_terms = 0;
  for (int m1 = -p.ell1; m1 <= p.ell1; m1++) {
    for (int m2 = -p.ell2; m2 <= p.ell2; m2++) {
      int Mlo = survey ? -p.ELL : 0, Mhi = survey ? p.ELL : 0;
      if (survey) {
        bool vanish = true;
        for (int M = -p.ELL; M <= p.ELL; M++)
          if (std::fabs(coupling_3pt(p.ell1, p.ell2, p.ELL, m1, m2, M)) > EPS_COUPLING) {
            vanish = false;
            break;
          }
        if (vanish) continue;
      }
      std::vector<cplx> yka, ykb, yra, yrb;
      bool tables = false;
      for (int M = Mlo; M <= Mhi; M++) {
        Triplet cur{m1, m2, M};
        bool redundant = false;
        for (auto& t : done) if (is_inverse(cur, t)) redundant = true;
        if (redundant) continue;
        double coupling = coupling_3pt(p.ell1, p.ell2, p.ELL, m1, m2, M);
        if (std::fabs(coupling) < EPS_COUPLING) continue;
        bool zeros = (m1 == 0 && m2 == 0 && M == 0);
        double f_mirror = zeros ? 0. : std::pow(-1., p.ELL);
        double f_mirror_w3j = zeros ? 0. : 1.;
        if (!tables) {
          yka = ylm_table(dn00, p.ell1, m1, true);
          ykb = ylm_table(dn00, p.ell2, m2, true);
          yra = ylm_table(dn00, p.ell1, m1, false);
          yrb = ylm_table(dn00, p.ell2, m2, false);
          tables = true;
        }

        // Raw bispectrum.
        Mesh G(p);
        if (survey) G.ylm_wgtd_field(d, *r, los_d, los_r, alpha, p.ELL, M);
        else G.unweighted_fluctuations(d);
        G.fourier_transform();
        G.apply_compensation();
        G.inv_fourier_transform();

        if (g_schedule == 0) {
          Mesh Fa(p), Fb(p);
          // The row form transforms the fixed first shell once per term
          // (threept.cpp:604-612); every other form transforms both shells
          // for every entry.
          double ea_row = 0.;
          int na_row = 0;
          if (p.shape == TRVG_SHAPE_ROW) {
            Fa.ifft_ylm_band_limited(dn00, yka, kb.bin_edges[p.idx_bin],
                                     kb.bin_edges[p.idx_bin + 1], ea_row,
                                     na_row);
          }
          for (int dv = 0; dv < dim; dv++) {
            double ea = ea_row, eb;
            int na = na_row, nbm;
            if (p.shape != TRVG_SHAPE_ROW) {
              Fa.ifft_ylm_band_limited(dn00, yka, kb.bin_edges[ra[dv]],
                                       kb.bin_edges[ra[dv] + 1], ea, na);
            }
            Fb.ifft_ylm_band_limited(dn00, ykb, kb.bin_edges[cb[dv]],
                                     kb.bin_edges[cb[dv] + 1], eb, nbm);
            if (count_terms == 0) {
              out.b1[dv] = kb.bin_centres[ra[dv]];
              out.b2[dv] = kb.bin_centres[cb[dv]];
              out.e1[dv] = ea;
              out.e2[dv] = eb;
              out.n1[dv] = na;
              out.n2[dv] = nbm;
            }
            cplx s = triple_sum(Fa, Fb, G);
            out.raw[dv] += coupling * vol_cell * (s + f_mirror * std::conj(s));
          }
        } else {
          // Deduplicated: one shell field per distinct (bin, side).
          std::vector<int> rows(ra), cols(cb);
          std::sort(rows.begin(), rows.end());
          rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
          std::sort(cols.begin(), cols.end());
          cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
          std::vector<Mesh> FA, FB;
          std::vector<double> EA(nb), EB(nb);
          std::vector<int> NA(nb), NB(nb), ia(nb, -1), ib(nb, -1);
          for (int b : rows) {
            FA.emplace_back(p);
            FA.back().ifft_ylm_band_limited(dn00, yka, kb.bin_edges[b],
                                            kb.bin_edges[b + 1], EA[b], NA[b]);
            ia[b] = int(FA.size()) - 1;
          }
          for (int b : cols) {
            FB.emplace_back(p);
            FB.back().ifft_ylm_band_limited(dn00, ykb, kb.bin_edges[b],
                                            kb.bin_edges[b + 1], EB[b], NB[b]);
            ib[b] = int(FB.size()) - 1;
          }
          for (int dv = 0; dv < dim; dv++) {
            if (count_terms == 0) {
              out.b1[dv] = kb.bin_centres[ra[dv]];
              out.b2[dv] = kb.bin_centres[cb[dv]];
              out.e1[dv] = EA[ra[dv]];
              out.e2[dv] = EB[cb[dv]];
              out.n1[dv] = NA[ra[dv]];
              out.n2[dv] = NB[cb[dv]];
            }
            cplx s = triple_sum(FA[ia[ra[dv]]], FB[ib[cb[dv]]], G);
            out.raw[dv] += coupling * vol_cell * (s + f_mirror * std::conj(s));
          }
        }
        // Shot noise.
        cplx Sbar;
        Mesh dnLM(p);
        Mesh* dn_for_sn = &dn00;
        Mesh* N_for_2pt = &N00;
        Mesh NLM(p);
        if (survey) {
          dnLM.ylm_wgtd_field(d, *r, los_d, los_r, alpha, p.ELL, M);
          dnLM.fourier_transform();
          NLM.ylm_wgtd_quad_field(d, *r, los_d, los_r, alpha, p.ELL, M);
          NLM.fourier_transform();
          Sbar = sbar_lm(d, los_d, p.ELL, M)
                 + std::pow(alpha, 3) * sbar_lm(*r, los_r, p.ELL, M);
          dn_for_sn = &dnLM;
          N_for_2pt = &NLM;
        } else {
          Sbar = double(d.n);
        }

        if (p.ell1 == 0 && p.ell2 == 0) {
          cplx sijk = coupling * Sbar;
          for (int dv = 0; dv < dim; dv++) out.shot[dv] += sijk;
        }
        if (p.ell2 == 0) {
          stats.twopt_fourier(dn00, *N_for_2pt, Sbar, p.ell1, m1, kb);
          for (int dv = 0; dv < dim; dv++) {
            int b = ra[dv];
            cplx s = coupling * (stats.pk[b] - stats.sn[b]);
            out.shot[dv] += s + f_mirror * std::conj(s);
          }
        }
        if (p.ell1 == 0) {
          stats.twopt_fourier(dn00, *N_for_2pt, Sbar, p.ell2, m2, kb);
          for (int dv = 0; dv < dim; dv++) {
            int b = cb[dv];
            cplx s = coupling * (stats.pk[b] - stats.sn[b]);
            out.shot[dv] += s + f_mirror * std::conj(s);
          }
        }
        if (g_schedule == 0) {
          for (int dv = 0; dv < dim; dv++) {
            std::vector<cplx> t = stats.twopt_config_field(*dn_for_sn, N00, Sbar);
            cplx s = coupling * stats.shotnoise_bispec_from_t(
                                    t, yra, yrb, sja, sjb, out.e1[dv], out.e2[dv]);
            out.shot[dv] += factor_phase * (s + f_mirror_w3j * std::conj(s));
          }
        } else {
          std::vector<cplx> t = stats.twopt_config_field(*dn_for_sn, N00, Sbar);
          for (int dv = 0; dv < dim; dv++) {
            cplx s = coupling * stats.shotnoise_bispec_from_t(
                                    t, yra, yrb, sja, sjb, out.e1[dv], out.e2[dv]);
            out.shot[dv] += factor_phase * (s + f_mirror_w3j * std::conj(s));
          }
        }
        done.push_back(cur);
        count_terms++;
      }
    }
  }
  for (int dv = 0; dv < dim; dv++) {
    out.raw[dv] *= norm_factor;
    out.shot[dv] *= norm_factor;
  }
  return out;
}

// Three-point correlation function body: survey (threept.cpp:1014-1471) or
// periodic box (threept.cpp:2190-2619).
// Window of a random catalogue (threept.cpp:2621-3077): with `window` the
// catalogue is *r alone (single-catalogue fields and shot-noise amplitude
// scaled by alpha), `alpha_win` its contrast and `wa_power` = i + j of the
// wide-angle kernel on G (negative: none).
struct Window3 {
  bool on = false;
  double alpha = 1.;
  int wa_power = -1;
};

Output threepcf_impl(bool survey, const Catalogue& d, const Catalogue* r,
                     const double* los_d, const double* los_r, const Params& p,
                     const trvg_binning& rb, double norm_factor,
                     const Window3& win = Window3()) {
  validate_coupling(p);
  double alpha = win.on ? win.alpha : (survey ? d.wstotal / r->wstotal : 1.);
  cplx factor_phase = ipow(p.ell1 + p.ell2);
  double factor_parity = std::pow(-1., p.ell1 + p.ell2);
  int nb = rb.num_bins;
  int dim = result_dim(p.shape, nb, p.idx_bin);
  std::vector<int> ra, cb;
  pair_bins(p.shape, nb, p.idx_bin, ra, cb);
  Output out(dim);

  Mesh dn00(p);
  if (win.on) dn00.ylm_wgtd_field_single(*r, los_r, alpha, 0, 0);
  else if (survey) dn00.ylm_wgtd_field(d, *r, los_d, los_r, alpha, 0, 0);
  else dn00.unweighted_fluctuations(d);
  dn00.fourier_transform();
  double vol_cell = dn00.vol_cell;

  Mesh N00(p);
  if (win.on) N00.ylm_wgtd_quad_field_single(*r, los_r, alpha, 0, 0);
  else if (survey) N00.ylm_wgtd_quad_field(d, *r, los_d, los_r, alpha, 0, 0);
  else N00.unweighted_field(d);
  N00.fourier_transform();

  BesselSpline sja(p.ell1), sjb(p.ell2);
  Stats stats(p);
  std::vector<Triplet> done;
  int count_terms = 0;
  for (int m1 = -p.ell1; m1 <= p.ell1; m1++) {
    for (int m2 = -p.ell2; m2 <= p.ell2; m2++) {
      int Mlo = survey ? -p.ELL : 0, Mhi = survey ? p.ELL : 0;
      if (survey) {
        bool vanish = true;
        for (int M = -p.ELL; M <= p.ELL; M++)
          if (std::fabs(coupling_3pt(p.ell1, p.ell2, p.ELL, m1, m2, M)) > EPS_COUPLING) {
            vanish = false;
            break;
          }
        if (vanish) continue;
      }
      std::vector<cplx> yka, ykb, yra, yrb;
      bool tables = false;
      for (int M = Mlo; M <= Mhi; M++) {
        Triplet cur{m1, m2, M};
        bool redundant = false;
        for (auto& t : done) if (is_inverse(cur, t)) redundant = true;
        if (redundant) continue;
        double coupling = coupling_3pt(p.ell1, p.ell2, p.ELL, m1, m2, M);
        if (std::fabs(coupling) < EPS_COUPLING) continue;
        bool zeros = (m1 == 0 && m2 == 0 && M == 0);
        double f_mirror = zeros ? 0. : std::pow(-1., p.ELL);
        double f_mirror_w3j = zeros ? 0. : 1.;
        if (!tables) {
          yra = ylm_table(dn00, p.ell1, m1, false);
          yrb = ylm_table(dn00, p.ell2, m2, false);
          yka = ylm_table(dn00, p.ell1, m1, true);
          ykb = ylm_table(dn00, p.ell2, m2, true);
          tables = true;
        }

        // Shot noise first (it defines r_eff and npairs).
        Mesh dnLM(p);
        Mesh* dn_for_sn = &dn00;
        cplx Sbar;
        if (win.on) {
          dnLM.ylm_wgtd_field_single(*r, los_r, alpha, p.ELL, M);
          dnLM.fourier_transform();
          Sbar = std::pow(alpha, 3) * sbar_lm(*r, los_r, p.ELL, M);
          dn_for_sn = &dnLM;
        } else if (survey) {
          dnLM.ylm_wgtd_field(d, *r, los_d, los_r, alpha, p.ELL, M);
          dnLM.fourier_transform();
          Sbar = sbar_lm(d, los_d, p.ELL, M)
                 + std::pow(alpha, 3) * sbar_lm(*r, los_r, p.ELL, M);
          dn_for_sn = &dnLM;
        } else {
          Sbar = double(d.n);
        }
        stats.shotnoise_3pcf(*dn_for_sn, N00, yra, yrb, Sbar, rb);
        for (int dv = 0; dv < dim; dv++) {
          if (ra[dv] != cb[dv]) continue;  // Kronecker delta, threept.cpp:1236-1285
          if (p.shape == TRVG_SHAPE_OFFDIAG && p.idx_bin != 0) continue;
          int b = ra[dv];
          out.shot[dv] += coupling * factor_parity
                          * (stats.xi[b] + f_mirror_w3j * std::conj(stats.xi[b]));
        }
        if (count_terms == 0) {
          for (int dv = 0; dv < dim; dv++) {
            out.b1[dv] = rb.bin_centres[ra[dv]];
            out.b2[dv] = rb.bin_centres[cb[dv]];
            out.e1[dv] = stats.r[ra[dv]];
            out.e2[dv] = stats.r[cb[dv]];
            out.n1[dv] = stats.npairs[ra[dv]];
            out.n2[dv] = stats.npairs[cb[dv]];
          }
        }

        // Raw 3PCF.
        Mesh G(p);
        if (win.on) G.ylm_wgtd_field_single(*r, los_r, alpha, p.ELL, M);
        else if (survey) G.ylm_wgtd_field(d, *r, los_d, los_r, alpha, p.ELL, M);
        else G.unweighted_fluctuations(d);
        G.fourier_transform();
        G.apply_compensation();
        G.inv_fourier_transform();
        if (win.on && win.wa_power >= 0) G.wide_angle_kernel(win.wa_power);

        if (g_schedule == 0) {
          Mesh Fa(p), Fb(p);
          for (int dv = 0; dv < dim; dv++) {
            Fa.ifft_sjl_ylm(dn00, yka, sja, out.e1[dv]);
            Fb.ifft_sjl_ylm(dn00, ykb, sjb, out.e2[dv]);
            cplx s = triple_sum(Fa, Fb, G);
            out.raw[dv] += coupling * vol_cell * factor_phase
                           * (s + f_mirror * std::conj(s));
          }
        } else {
          std::vector<int> rows(ra), cols(cb);
          std::sort(rows.begin(), rows.end());
          rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
          std::sort(cols.begin(), cols.end());
          cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
          // r_eff per bin, recorded from the first term.
          std::vector<double> reff_a(nb, 0.), reff_b(nb, 0.);
          for (int dv = 0; dv < dim; dv++) {
            reff_a[ra[dv]] = out.e1[dv];
            reff_b[cb[dv]] = out.e2[dv];
          }
          std::vector<Mesh> FA, FB;
          std::vector<int> ia(nb, -1), ib(nb, -1);
          for (int b : rows) {
            FA.emplace_back(p);
            FA.back().ifft_sjl_ylm(dn00, yka, sja, reff_a[b]);
            ia[b] = int(FA.size()) - 1;
          }
          for (int b : cols) {
            FB.emplace_back(p);
            FB.back().ifft_sjl_ylm(dn00, ykb, sjb, reff_b[b]);
            ib[b] = int(FB.size()) - 1;
          }
          for (int dv = 0; dv < dim; dv++) {
            cplx s = triple_sum(FA[ia[ra[dv]]], FB[ib[cb[dv]]], G);
            out.raw[dv] += coupling * vol_cell * factor_phase
                           * (s + f_mirror * std::conj(s));
          }
        }
        done.push_back(cur);
        count_terms++;
      }
    }
  }
  for (int dv = 0; dv < dim; dv++) {
    out.raw[dv] *= norm_factor;
    out.shot[dv] *= norm_factor;
  }
  return out;
}

// --------------------------------------------------------------------------
// Two-point statistics (src/twopt.cpp)
// --------------------------------------------------------------------------

// twopt.cpp:45-49
double coupling_2pt(int ell, int ELL, int m, int M) {
  return (2 * ell + 1) * (2 * ELL + 1) * wigner_3j(ell, 0, ELL, 0, 0, 0) *
         wigner_3j(ell, 0, ELL, m, 0, M);
}

// twopt.cpp:298-381: sum_p y_lm(los_p) w_p^2 over one catalogue.
cplx sn_amp_powspec(const Catalogue& c, const double* los, int ell, int m) {
  double sr = 0., si = 0.;
#pragma omp parallel for reduction(+ : sr, si)
  for (long long i = 0; i < c.n; i++) {
    double l3[3] = {los[3 * i], los[3 * i + 1], los[3 * i + 2]};
    cplx y = ylm_reduced(ell, m, l3) * std::pow(c.w[i], 2);
    sr += y.real();
    si += y.imag();
  }
  return cplx(sr, si);
}

struct Output2 {
  std::vector<double> bin, eff;
  std::vector<int> cnt;
  std::vector<cplx> raw, shot;
  explicit Output2(int nb) : bin(nb, 0.), eff(nb, 0.), cnt(nb, 0), raw(nb, 0.), shot(nb, 0.) {}
};

// The survey-type two-point body (twopt.cpp:388-496 in Fourier space,
// 498-607 in configuration space, 795-901 for the window of one random
// catalogue): dn_00, then for every M the field dn_LM, its shot-noise
// amplitude and the coupled statistics of dn_LM against dn_00.
Output2 twopt_survey_impl(bool fourier, const Catalogue* d, const Catalogue& r,
                          const double* los_d, const double* los_r,
                          const Params& p, const trvg_binning& b,
                          double alpha, double norm_factor) {
  const int ELL = p.ELL, ell1 = p.ELL, nb = b.num_bins;
  Output2 out(nb);
  Mesh dn00(p);
  if (d) dn00.ylm_wgtd_field(*d, r, los_d, los_r, alpha, 0, 0);
  else dn00.ylm_wgtd_field_single(r, los_r, alpha, 0, 0);
  dn00.fourier_transform();
  Stats stats(p);
  for (int M = -ELL; M <= ELL; M++) {
    Mesh dnLM(p);
    if (d) dnLM.ylm_wgtd_field(*d, r, los_d, los_r, alpha, ELL, M);
    else dnLM.ylm_wgtd_field_single(r, los_r, alpha, ELL, M);
    dnLM.fourier_transform();
    cplx sn_amp = d ? sn_amp_powspec(*d, los_d, ELL, M) +
                          std::pow(alpha, 2) * sn_amp_powspec(r, los_r, ELL, M)
                    : std::pow(alpha, 2) * sn_amp_powspec(r, los_r, ELL, M);
    for (int m1 = -ell1; m1 <= ell1; m1++) {
      double coupling = coupling_2pt(ell1, ELL, m1, M);
      if (std::fabs(coupling) < 1.e-10) continue;
      if (fourier) {
        stats.twopt_fourier(dnLM, dn00, sn_amp, ell1, m1, b);
        for (int i = 0; i < nb; i++) {
          out.raw[i] += coupling * stats.pk[i];
          out.shot[i] += coupling * stats.sn[i];
        }
      } else {
        stats.twopt_config(dnLM, dn00, sn_amp, ell1, m1, b);
        for (int i = 0; i < nb; i++) out.raw[i] += coupling * stats.xi[i];
      }
      if (M == 0 && m1 == 0) {
        for (int i = 0; i < nb; i++) {
          out.cnt[i] = fourier ? stats.nmodes[i] : stats.npairs[i];
          out.eff[i] = fourier ? stats.k[i] : stats.r[i];
        }
      }
    }
  }
  for (int i = 0; i < nb; i++) {
    out.bin[i] = b.bin_centres[i];
    out.raw[i] *= norm_factor;
    out.shot[i] *= norm_factor;
  }
  return out;
}

// twopt.cpp:609-701 (Fourier) and 703-793 (configuration): the unweighted
// field fluctuation against itself, (2 ELL + 1) times the M = 0 statistic.
Output2 twopt_box_impl(bool fourier, const Catalogue& d, const Params& p,
                       const trvg_binning& b, double norm_factor) {
  const int nb = b.num_bins;
  Output2 out(nb);
  Mesh dn(p);
  dn.unweighted_fluctuations(d);
  dn.fourier_transform();
  cplx sn_amp = double(d.n);
  Stats stats(p);
  const double f = double(2 * p.ELL + 1);
  if (fourier) stats.twopt_fourier(dn, dn, sn_amp, p.ELL, 0, b);
  else stats.twopt_config(dn, dn, sn_amp, p.ELL, 0, b);
  for (int i = 0; i < nb; i++) {
    out.bin[i] = b.bin_centres[i];
    out.cnt[i] = fourier ? stats.nmodes[i] : stats.npairs[i];
    out.eff[i] = fourier ? stats.k[i] : stats.r[i];
    out.raw[i] = norm_factor * f * (fourier ? stats.pk[i] : stats.xi[i]);
    out.shot[i] = fourier ? norm_factor * f * stats.sn[i] : cplx(0., 0.);
  }
  return out;
}

void write_powspec(const Output2& o, trvg_powspec_result* res) {
  const int nb = int(o.bin.size());
  if (res->dim < nb)
    fail(TRVG_ERR_PARAM, "Result capacity %d < data-vector dimension %d.", res->dim, nb);
  for (int i = 0; i < nb; i++) {
    res->kbin[i] = o.bin[i]; res->keff[i] = o.eff[i]; res->nmodes[i] = o.cnt[i];
    res->pk_raw[2 * i] = o.raw[i].real(); res->pk_raw[2 * i + 1] = o.raw[i].imag();
    res->pk_shot[2 * i] = o.shot[i].real(); res->pk_shot[2 * i + 1] = o.shot[i].imag();
  }
  res->dim = nb;
}

void write_2pcf(const Output2& o, trvg_2pcf_result* res) {
  const int nb = int(o.bin.size());
  if (res->dim < nb)
    fail(TRVG_ERR_PARAM, "Result capacity %d < data-vector dimension %d.", res->dim, nb);
  for (int i = 0; i < nb; i++) {
    res->rbin[i] = o.bin[i]; res->reff[i] = o.eff[i]; res->npairs[i] = o.cnt[i];
    res->xi[2 * i] = o.raw[i].real(); res->xi[2 * i + 1] = o.raw[i].imag();
  }
  res->dim = nb;
}

void write_bispec(const Output& o, trvg_bispec_result* res) {
  if (res->dim < o.dim)
    fail(TRVG_ERR_PARAM, "Result capacity %d < data-vector dimension %d.",
         res->dim, o.dim);
  for (int i = 0; i < o.dim; i++) {
    res->k1_bin[i] = o.b1[i]; res->k2_bin[i] = o.b2[i];
    res->k1_eff[i] = o.e1[i]; res->k2_eff[i] = o.e2[i];
    res->nmodes_1[i] = o.n1[i]; res->nmodes_2[i] = o.n2[i];
    res->bk_raw[2 * i] = o.raw[i].real(); res->bk_raw[2 * i + 1] = o.raw[i].imag();
    res->bk_shot[2 * i] = o.shot[i].real(); res->bk_shot[2 * i + 1] = o.shot[i].imag();
  }
  res->dim = o.dim;
}

void write_3pcf(const Output& o, trvg_3pcf_result* res) {
  if (res->dim < o.dim)
    fail(TRVG_ERR_PARAM, "Result capacity %d < data-vector dimension %d.",
         res->dim, o.dim);
  for (int i = 0; i < o.dim; i++) {
    res->r1_bin[i] = o.b1[i]; res->r2_bin[i] = o.b2[i];
    res->r1_eff[i] = o.e1[i]; res->r2_eff[i] = o.e2[i];
    res->npairs_1[i] = o.n1[i]; res->npairs_2[i] = o.n2[i];
    res->zeta_raw[2 * i] = o.raw[i].real(); res->zeta_raw[2 * i + 1] = o.raw[i].imag();
    res->zeta_shot[2 * i] = o.shot[i].real(); res->zeta_shot[2 * i + 1] = o.shot[i].imag();
  }
  res->dim = o.dim;
}

template <class F>
int guarded(F&& fn) {
  try {
    g_last_error.clear();
    fn();
    return TRVG_OK;
  } catch (const OracleError& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "Host allocation failed.";
    return TRVG_ERR_ALLOC;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return TRVG_ERR_INTERNAL;
  }
}

}  // namespace

// ==========================================================================
// C API (prefix trvo_: same call shapes as the trvg_ product ABI)
// ==========================================================================

extern "C" {

const char* trvo_last_error(void) { return g_last_error.c_str(); }

void trvo_set_schedule(int s) { g_schedule = s; }

void trvo_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int trvo_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int trvo_result_dim(int shape, int nb, int idx_bin, int* dim) {
  return guarded([&] { *dim = result_dim(shape, nb, idx_bin); });
}

void trvo_ylm(int ell, int m, const double* pos, double* out) {
  cplx y = ylm_reduced(ell, m, pos);
  out[0] = y.real();
  out[1] = y.imag();
}

double trvo_wigner_3j(int j1, int j2, int j3, int m1, int m2, int m3) {
  return wigner_3j(j1, j2, j3, m1, m2, m3);
}

double trvo_sph_bessel(int ell, double x) { return sph_bessel_j(ell, x); }

// Evaluates the spline for ell at n points.
void trvo_bessel_spline(int ell, const double* x, double* out, long long n) {
  BesselSpline s(ell);
  for (long long i = 0; i < n; i++) out[i] = s.eval(x[i]);
}

// FFT of an interleaved complex array of ng[0]*ng[1]*ng[2] cells.
void trvo_fft3d(double* data, const int* ng, int sign) {
  fft3d(reinterpret_cast<cplx*>(data), ng, sign);
}

// Weighted assignment of a catalogue with complex weights (interleaved) to
// `out` (interleaved complex, nmesh cells); `shadow` selects the
// interlacing half-cell-shifted pass.
int trvo_assign(const trvg_catalogue* cat, const double* weights,
                const trvg_params* params, int shadow, double* out) {
  return guarded([&] {
    Params p = make_params(params);
    p.interlace = false;
    Catalogue c = load_catalogue(cat);
    std::vector<cplx> w(c.n);
    for (long long i = 0; i < c.n; i++) w[i] = cplx(weights[2 * i], weights[2 * i + 1]);
    Mesh m(p);
    std::vector<cplx> f(p.nmesh, 0.);
    m.assign_pass(c, w, f, shadow != 0);
    std::memcpy(out, static_cast<const void*>(f.data()), sizeof(cplx) * p.nmesh);
  });
}

// Fourier-space assignment window W(k) on the mesh of `params` (order from
// params->assignment_order), nmesh doubles (field.cpp:1114-1201).
int trvo_window(const trvg_params* params, double* out) {
  return guarded([&] {
    Params p = make_params(params);
    p.interlace = false;
    Mesh m(p);
    m.compute_window(p.order);
    std::memcpy(out, m.window.data(), sizeof(double) * p.nmesh);
  });
}

// Aliasing function C1(k) of FieldStats (field.cpp:3403-3564) on the mesh
// of `params`, with or without the interlaced approximation, nmesh doubles.
int trvo_alias(const trvg_params* params, double* out) {
  return guarded([&] {
    Params p = make_params(params);
    Stats st(p);
    st.compute_alias();
    std::memcpy(out, st.alias.data(), sizeof(double) * p.nmesh);
  });
}

// Forward transform of an assigned field as MeshField::fourier_transform,
// with an optional interlacing shadow (both interleaved complex, in place).
int trvo_fourier_transform(const trvg_params* params, double* field,
                           double* shadow) {
  return guarded([&] {
    Params p = make_params(params);
    p.interlace = (shadow != nullptr);
    Mesh m(p);
    std::memcpy(static_cast<void*>(m.f.data()), field, sizeof(cplx) * p.nmesh);
    if (shadow)
      std::memcpy(static_cast<void*>(m.fs.data()), shadow, sizeof(cplx) * p.nmesh);
    m.fourier_transform();
    std::memcpy(field, static_cast<const void*>(m.f.data()), sizeof(cplx) * p.nmesh);
  });
}

int trvo_calc_bispec_normalisation_from_particles(const trvg_catalogue* cat,
                                                  double alpha, double* norm) {
  return guarded([&] {
    Catalogue c = load_catalogue(cat);
    double s = 0.;
    for (long long i = 0; i < c.n; i++)
      s += c.ws[i] * std::pow(c.nz[i], 2) * std::pow(c.wc[i], 3);
    if (s == 0.)
      fail(TRVG_ERR_DATA,
           "Particle 'nz' values appear to be all zeros. "
           "Check the input catalogue contains valid 'nz' field.");
    *norm = 1. / (alpha * s);
  });
}

int trvo_calc_bispec_normalisation_from_mesh(const trvg_catalogue* cat,
                                             const trvg_params* params,
                                             double alpha, double* norm) {
  return guarded([&] {
    Params p = make_params(params);
    Catalogue c = load_catalogue(cat);
    Mesh m(p);
    double nf = m.grid_powlaw_norm(c, 3);
    *norm = nf / std::pow(alpha, 3);
  });
}

// Bounded CPU-baseline sample: the per-pair loop body of the reference
// schedule (threept.cpp:465-517 and 950-964 with field.cpp:1792-1906,
// 3198-3396) run `npairs` times on the mesh of `params` (random dense
// fields; the cost does not depend on field values), cycling over the
// bins of `kb`.  Returns the wall time of the loop in seconds.
int trvo_bench_pairs(const trvg_params* params, const trvg_binning* kb,
                     int npairs, double* seconds) {
  return guarded([&] {
    Params p = make_params(params);
    p.interlace = false;
    std::vector<cplx> rnd(p.nmesh);
    unsigned long long st = 88172645463325252ULL;
    for (long long g = 0; g < p.nmesh; g++) {
      st ^= st << 13; st ^= st >> 7; st ^= st << 17;
      double a = double(st % 1000003) / 1000003. - 0.5;
      st ^= st << 13; st ^= st >> 7; st ^= st << 17;
      double b = double(st % 1000003) / 1000003. - 0.5;
      rnd[g] = cplx(a, b);
    }
    Mesh dn00(p), N00(p), G(p), Fa(p), Fb(p);
    dn00.f = rnd;
    N00.f = rnd;
    G.f = rnd;
    Stats stats(p);
    std::vector<cplx> yka = ylm_table(dn00, p.ell1, 0, true);
    std::vector<cplx> ykb = ylm_table(dn00, p.ell2, 0, true);
    std::vector<cplx> yra = ylm_table(dn00, p.ell1, 0, false);
    std::vector<cplx> yrb = ylm_table(dn00, p.ell2, 0, false);
    BesselSpline sja(p.ell1), sjb(p.ell2);
    stats.compute_alias();
    dn00.compute_window(p.order);
    Fa.compute_window(p.order);
    Fb.compute_window(p.order);
    double t0 = omp_get_wtime();
    double sink = 0.;
    for (int q = 0; q < npairs; q++) {
      int b = q % kb->num_bins;
      double ea, eb;
      int na, nbm;
      Fa.ifft_ylm_band_limited(dn00, yka, kb->bin_edges[b], kb->bin_edges[b + 1], ea, na);
      Fb.ifft_ylm_band_limited(dn00, ykb, kb->bin_edges[b], kb->bin_edges[b + 1], eb, nbm);
      cplx s = triple_sum(Fa, Fb, G);
      std::vector<cplx> t = stats.twopt_config_field(dn00, N00, cplx(1., 0.));
      cplx sn = stats.shotnoise_bispec_from_t(t, yra, yrb, sja, sjb, ea, eb);
      sink += s.real() + sn.real();
    }
    *seconds = omp_get_wtime() - t0;
    if (sink == 12345.6789) *seconds += 0.;  // keep the work observable
  });
}

int trvo_compute_bispec(const trvg_catalogue* data, const trvg_catalogue* rand,
                        const double* los_data, const double* los_rand,
                        const trvg_params* params, const trvg_binning* kb,
                        double norm_factor, trvg_bispec_result* out) {
  return guarded([&] {
    Params p = make_params(params);
    p.interlace = false;  // parameters.cpp:1240-1249
    Catalogue d = load_catalogue(data), r = load_catalogue(rand);
    write_bispec(bispec_impl(true, d, &r, los_data, los_rand, p, *kb, norm_factor), out);
  });
}

int trvo_compute_bispec_in_gpp_box(const trvg_catalogue* data,
                                   const trvg_params* params,
                                   const trvg_binning* kb, double norm_factor,
                                   trvg_bispec_result* out) {
  return guarded([&] {
    Params p = make_params(params);
    p.interlace = false;
    Catalogue d = load_catalogue(data);
    write_bispec(bispec_impl(false, d, nullptr, nullptr, nullptr, p, *kb, norm_factor), out);
  });
}

int trvo_compute_3pcf(const trvg_catalogue* data, const trvg_catalogue* rand,
                      const double* los_data, const double* los_rand,
                      const trvg_params* params, const trvg_binning* rb,
                      double norm_factor, trvg_3pcf_result* out) {
  return guarded([&] {
    Params p = make_params(params);
    p.interlace = false;
    Catalogue d = load_catalogue(data), r = load_catalogue(rand);
    write_3pcf(threepcf_impl(true, d, &r, los_data, los_rand, p, *rb, norm_factor), out);
  });
}

int trvo_compute_3pcf_in_gpp_box(const trvg_catalogue* data,
                                 const trvg_params* params,
                                 const trvg_binning* rb, double norm_factor,
                                 trvg_3pcf_result* out) {
  return guarded([&] {
    Params p = make_params(params);
    p.interlace = false;
    Catalogue d = load_catalogue(data);
    write_3pcf(threepcf_impl(false, d, nullptr, nullptr, nullptr, p, *rb, norm_factor), out);
  });
}

int trvo_compute_3pcf_window(const trvg_catalogue* rand, const double* los_rand,
                             const trvg_params* params, const trvg_binning* rb,
                             double alpha, double norm_factor, int wide_angle,
                             int i_wa, int j_wa, trvg_3pcf_result* out) {
  return guarded([&] {
    Params p = make_params(params);
    p.interlace = false;
    Catalogue r = load_catalogue(rand);
    Window3 w;
    w.on = true;
    w.alpha = alpha;
    w.wa_power = wide_angle ? i_wa + j_wa : -1;
    write_3pcf(threepcf_impl(true, r, &r, los_rand, los_rand, p, *rb, norm_factor, w), out);
  });
}

int trvo_calc_powspec_normalisation_from_particles(const trvg_catalogue* cat,
                                                   double alpha, double* norm) {
  return guarded([&] {
    Catalogue c = load_catalogue(cat);
    double s = 0.;
    for (long long i = 0; i < c.n; i++) s += c.ws[i] * c.nz[i] * std::pow(c.wc[i], 2);
    if (s == 0.)
      fail(TRVG_ERR_DATA,
           "Particle 'nz' values appear to be all zeros. "
           "Check the input catalogue contains valid 'nz' field.");
    *norm = 1. / (alpha * s);
  });
}

int trvo_calc_powspec_normalisation_from_mesh(const trvg_catalogue* cat,
                                              const trvg_params* params,
                                              double alpha, double* norm) {
  return guarded([&] {
    Params p = make_params(params);
    Catalogue c = load_catalogue(cat);
    Mesh m(p);
    *norm = m.grid_powlaw_norm(c, 2) / std::pow(alpha, 2);
  });
}

// twopt.cpp:111-228 after the re-alignment: both catalogues assigned with
// their weights w, 1 / (alpha V_cell sum_x n_d n_r).
int trvo_calc_powspec_normalisation_from_meshes(const trvg_catalogue* data,
                                                const trvg_catalogue* rand,
                                                const trvg_params* params,
                                                double alpha, double* norm) {
  return guarded([&] {
    Params p = make_params(params);
    Catalogue d = load_catalogue(data), r = load_catalogue(rand);
    Mesh md(p), mr(p);
    std::vector<cplx> wd(d.n), wr(r.n);
    for (long long i = 0; i < d.n; i++) wd[i] = cplx(d.w[i], 0.);
    for (long long i = 0; i < r.n; i++) wr[i] = cplx(r.w[i], 0.);
    md.assign(d, wd);
    mr.assign(r, wr);
    double s = 0.;
    for (long long g = 0; g < p.nmesh; g++) s += md.f[g].real() * mr.f[g].real();
    *norm = 1. / (alpha * md.vol_cell * s);
  });
}

int trvo_calc_powspec_shotnoise_from_particles(const trvg_catalogue* cat,
                                               double alpha, double* sn) {
  return guarded([&] {
    (void)alpha;
    Catalogue c = load_catalogue(cat);
    double s = 0.;
    for (long long i = 0; i < c.n; i++) s += std::pow(c.ws[i], 2) * std::pow(c.wc[i], 2);
    *sn = s;
  });
}

int trvo_compute_powspec(const trvg_catalogue* data, const trvg_catalogue* rand,
                         const double* los_data, const double* los_rand,
                         const trvg_params* params, const trvg_binning* kb,
                         double norm_factor, trvg_powspec_result* out) {
  return guarded([&] {
    Params p = make_params(params);
    Catalogue d = load_catalogue(data), r = load_catalogue(rand);
    double alpha = d.wstotal / r.wstotal;
    write_powspec(twopt_survey_impl(true, &d, r, los_data, los_rand, p, *kb, alpha,
                                    norm_factor), out);
  });
}

int trvo_compute_corrfunc(const trvg_catalogue* data, const trvg_catalogue* rand,
                          const double* los_data, const double* los_rand,
                          const trvg_params* params, const trvg_binning* rb,
                          double norm_factor, trvg_2pcf_result* out) {
  return guarded([&] {
    Params p = make_params(params);
    Catalogue d = load_catalogue(data), r = load_catalogue(rand);
    double alpha = d.wstotal / r.wstotal;
    write_2pcf(twopt_survey_impl(false, &d, r, los_data, los_rand, p, *rb, alpha,
                                 norm_factor), out);
  });
}

int trvo_compute_powspec_in_gpp_box(const trvg_catalogue* data,
                                    const trvg_params* params,
                                    const trvg_binning* kb, double norm_factor,
                                    trvg_powspec_result* out) {
  return guarded([&] {
    Params p = make_params(params);
    Catalogue d = load_catalogue(data);
    write_powspec(twopt_box_impl(true, d, p, *kb, norm_factor), out);
  });
}

int trvo_compute_corrfunc_in_gpp_box(const trvg_catalogue* data,
                                     const trvg_params* params,
                                     const trvg_binning* rb, double norm_factor,
                                     trvg_2pcf_result* out) {
  return guarded([&] {
    Params p = make_params(params);
    Catalogue d = load_catalogue(data);
    write_2pcf(twopt_box_impl(false, d, p, *rb, norm_factor), out);
  });
}

int trvo_compute_corrfunc_window(const trvg_catalogue* rand, const double* los_rand,
                                 const trvg_params* params, const trvg_binning* rb,
                                 double alpha, double norm_factor,
                                 trvg_2pcf_result* out) {
  return guarded([&] {
    Params p = make_params(params);
    Catalogue r = load_catalogue(rand);
    write_2pcf(twopt_survey_impl(false, nullptr, r, nullptr, los_rand, p, *rb, alpha,
                                 norm_factor), out);
  });
}

}  // extern "C"
