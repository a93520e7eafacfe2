// engine_core.hip -- host runtime pieces of the engine: memory pool,
// catalogue upload, spline construction, window tables, Wigner 3-j and the
// binning-shape index maps.
#include <algorithm>
#include <cmath>
#include <cstring>

#include <map>
#include <tuple>

#include "engine.hpp"

namespace trvg {

// ---------------------------------------------------------------------------
// Pool
// ---------------------------------------------------------------------------

Pool& pool() {
  static Pool* p = new Pool();  // process lifetime; device memory is
                                // reclaimed with the context at exit
  return *p;
}

void* Pool::get(size_t bytes) {
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = free_.find(bytes);
    if (it != free_.end()) {
      void* p = it->second;
      free_.erase(it);
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    release_all();
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error(TRVG_ERR_ALLOC,
                  "Device allocation of " + std::to_string(bytes) +
                      " bytes failed: " + hipGetErrorString(e));
    }
  }
  return p;
}

void Pool::put(void* p, size_t bytes) {
  std::lock_guard<std::mutex> lk(mu_);
  free_.emplace(bytes, p);
}

void Pool::release_all() {
  std::lock_guard<std::mutex> lk(mu_);
  for (auto& kv : free_) (void)hipFree(kv.second);
  free_.clear();
}

// ---------------------------------------------------------------------------
// Spherical Bessel spline (maths.cpp:309-375): nodes x_i = 0.05 i up to
// split = max(1000, l^2); natural boundary; second-derivative/2
// coefficients from the tridiagonal system (Thomas algorithm).
// ---------------------------------------------------------------------------

void spline_nodes(int ell, std::vector<double>& y, std::vector<double>& c,
                  double& split) {
  const double step = 0.05;
  split = 1000.;
  split = (split >= double(ell * ell)) ? split : double(ell * ell);
  int n = int((split - 0.) / step) + 1;
  std::vector<double> x(n);
  y.resize(n);
  for (int i = 0; i < n; i++) {
    x[i] = 0. + step * i;
    y[i] = sph_bessel_dev(ell, x[i]);
  }
  c.assign(n, 0.);
  int m = n - 2;
  std::vector<double> diag(m), off(m), rhs(m), sol(m);
  for (int i = 1; i <= m; i++) {
    double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
    diag[i - 1] = 2. * (h0 + h1);
    off[i - 1] = h1;
    rhs[i - 1] = 3. * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
  }
  for (int i = 1; i < m; i++) {
    double f = off[i - 1] / diag[i - 1];
    diag[i] -= f * off[i - 1];
    rhs[i] -= f * rhs[i - 1];
  }
  sol[m - 1] = rhs[m - 1] / diag[m - 1];
  for (int i = m - 2; i >= 0; i--) sol[i] = (rhs[i] - off[i] * sol[i + 1]) / diag[i];
  for (int i = 0; i < m; i++) c[i + 1] = sol[i];
}

double sq_threshold(double t) {
  if (!(t > 0.)) return -1.;
  if (std::isinf(t)) return t;
  double s = t * t;
  while (s > 0. && std::sqrt(s) >= t) s = std::nextafter(s, 0.);
  while (std::sqrt(s) < t) s = std::nextafter(s, HUGE_VAL);
  return s;
}

static std::mutex g_spline_mu;
static std::map<std::pair<int, int>, std::unique_ptr<Spline>> g_splines;

const Spline& cached_spline(int dev, int ell, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_spline_mu);
  auto key = std::make_pair(dev, ell);
  auto it = g_splines.find(key);
  if (it != g_splines.end()) return *it->second;
  auto sp = std::make_unique<Spline>();
  build_spline(ell, *sp, s);
  return *(g_splines[key] = std::move(sp));
}

void build_spline(int ell, Spline& sp, hipStream_t s) {
  std::vector<double> y, c;
  double split;
  spline_nodes(ell, y, c, split);
  sp.ell = ell;
  sp.split = split;
  sp.n = int(y.size());
  // Interval coefficients as GSL's cspline evaluation forms them:
  // b = dy/dx - dx (c_i+1 + 2 c_i) / 3, d = (c_i+1 - c_i) / (3 dx).
  std::vector<double4> cf(sp.n - 1);
  std::vector<float4> cf32(sp.n - 1);
  for (int i = 0; i + 1 < sp.n; i++) {
    const double xlo = sp.step * double(i), xhi = sp.step * double(i + 1);
    const double dx = xhi - xlo;
    const double dy = y[i + 1] - y[i];
    const double t = dx * (c[i + 1] + 2. * c[i]);
    const double b = dy / dx - t / 3.;
    const double d = (c[i + 1] - c[i]) / (3. * dx);
    cf[i] = make_double4(y[i], b, c[i], d);
    cf32[i] = make_float4(float(y[i]), float(b), float(c[i]), float(d));
  }
  sp.cf.reset(sp.n - 1);
  sp.cf32.reset(sp.n - 1);
  TRVG_HIP(hipMemcpyAsync(sp.cf.get(), cf.data(), (sp.n - 1) * sizeof(double4),
                          hipMemcpyHostToDevice, s));
  TRVG_HIP(hipMemcpyAsync(sp.cf32.get(), cf32.data(), (sp.n - 1) * sizeof(float4),
                          hipMemcpyHostToDevice, s));
  TRVG_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

// Streams are reused across calls: a free list per device (all at the
// default priority).
static std::mutex g_stream_mu;
static std::map<int, std::vector<hipStream_t>> g_free_streams;

hipStream_t acquire_stream(int dev) {
  {
    std::lock_guard<std::mutex> lk(g_stream_mu);
    auto& v = g_free_streams[dev];
    if (!v.empty()) {
      hipStream_t s = v.back();
      v.pop_back();
      return s;
    }
  }
  hipStream_t s = nullptr;
  TRVG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  return s;
}

void release_stream(int dev, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_stream_mu);
  g_free_streams[dev].push_back(s);
}

// Per-axis tables depend only on the mesh sizes: built once per (device,
// sizes) and kept for the process lifetime.
struct AxisTableData {
  DevArray<double> data;
  long long offs[3][3];
};
static std::mutex g_tab_mu;
static std::map<std::tuple<int, int, int, int>, std::unique_ptr<AxisTableData>> g_tabs;

static const AxisTableData& axis_tables(int dev, const Geom& g, hipStream_t stream) {
  std::lock_guard<std::mutex> lk(g_tab_mu);
  auto key = std::make_tuple(dev, g.n[0], g.n[1], g.n[2]);
  auto it = g_tabs.find(key);
  if (it != g_tabs.end()) return *it->second;
  // sinc, sin^2, u for each axis (field.cpp:1138-1149, 3441-3465).
  auto t = std::make_unique<AxisTableData>();
  long long tot = 3LL * (g.n[0] + g.n[1] + g.n[2]);
  std::vector<double> h(tot);
  long long off = 0;
  for (int kind = 0; kind < 3; kind++) {
    for (int a = 0; a < 3; a++) {
      t->offs[kind][a] = off;
      for (int i = 0; i < g.n[a]; i++) {
        int ii = shifted(i, g.n[a]);
        double u = M_PI * ii / double(g.n[a]);
        double v;
        if (kind == 0) v = (ii != 0) ? std::sin(u) / u : 1.;
        else if (kind == 1) v = (ii != 0) ? std::sin(u) * std::sin(u) : 0.;
        else v = u;
        h[off + i] = v;
      }
      off += g.n[a];
    }
  }
  t->data.reset(tot);
  TRVG_HIP(hipMemcpyAsync(t->data.get(), h.data(), tot * 8, hipMemcpyHostToDevice, stream));
  TRVG_HIP(hipStreamSynchronize(stream));
  return *(g_tabs[key] = std::move(t));
}

void* PinnedArena::alloc(size_t bytes) {
  bytes = (bytes + 255) & ~size_t(255);
  while (chunk_ < chunks_.size()) {
    if (used_ + bytes <= chunks_[chunk_].size) {
      void* p = chunks_[chunk_].p + used_;
      used_ += bytes;
      return p;
    }
    chunk_++;
    used_ = 0;
  }
  Chunk c;
  c.size = std::max(bytes, size_t(16) << 20);
  TRVG_HIP(hipHostMalloc((void**)&c.p, c.size));
  chunks_.push_back(c);
  chunk_ = chunks_.size() - 1;
  used_ = bytes;
  return c.p;
}

PinnedArena::~PinnedArena() {
  for (auto& c : chunks_) (void)hipHostFree(c.p);
}

// One arena per pooled stream (a stream serves one Context at a time).
static PinnedArena& pinned_arena(hipStream_t s) {
  static std::mutex mu;
  static std::map<hipStream_t, std::unique_ptr<PinnedArena>>* arenas =
      new std::map<hipStream_t, std::unique_ptr<PinnedArena>>();
  std::lock_guard<std::mutex> lk(mu);
  auto& a = (*arenas)[s];
  if (!a) a = std::make_unique<PinnedArena>();
  return *a;
}

void Context::flush() {
  TRVG_HIP(hipStreamSynchronize(stream));
  keep.clear();
  std::vector<std::function<void()>> d;
  d.swap(deferred);
  for (auto& f : d) f();
}

Context::Context(const trvg_params& params) : p(params) {
  if (p.device >= 0) TRVG_HIP(hipSetDevice(p.device));
  g = make_geom(p);
  TRVG_HIP(hipGetDevice(&dev));
  stream = acquire_stream(dev);
  pin = &pinned_arena(stream);
  pin->reset();
  TRVG_HIP(hipDeviceGetAttribute(&sm_count, hipDeviceAttributeMultiprocessorCount, dev));
  const AxisTableData& t = axis_tables(dev, g, stream);
  for (int a = 0; a < 3; a++) {
    tab.sinc[a] = t.data.get() + t.offs[0][a];
    tab.sin2[a] = t.data.get() + t.offs[1][a];
    tab.u[a] = t.data.get() + t.offs[2][a];
  }
  tab.order = p.assignment_order;
  tab.interlace = p.interlace;
}

Context::Context(const trvg_params& params, hipStream_t s, const AxisTables& t)
    : p(params), stream(s), tab(t), pooled(false) {
  g = make_geom(p);
  TRVG_HIP(hipGetDevice(&dev));
  TRVG_HIP(hipDeviceGetAttribute(&sm_count, hipDeviceAttributeMultiprocessorCount, dev));
  tab.order = p.assignment_order;
  tab.interlace = p.interlace;
}

Context::~Context() {
  if (!pooled) {
    (void)hipStreamSynchronize(stream);
    return;
  }
  if (stream) {
    (void)hipStreamSynchronize(stream);
    keep.clear();
    try {
      flush_profile();
    } catch (...) {
    }
    release_stream(dev, stream);
  }
}

void Context::flush_profile() {
  if (prof.empty()) return;
  TRVG_HIP(hipStreamSynchronize(stream));
  for (auto& r : prof) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess)
      Profiler::get().add(r.kc, ms, r.bytes);
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  prof.clear();
}

// ---------------------------------------------------------------------------
// Profiler
// ---------------------------------------------------------------------------

const char* kernel_class_name(int kc) {
  static const char* names[KC_COUNT] = {
      "fft_pass_z_f32", "fft_pass_strided_f32", "assign", "pair_reduce",
      "twopt_bins", "shell_stats", "shot_radial_hist", "shot_pairs", "misc",
      "fft_pass_z_f64", "fft_pass_strided_f64"};
  return (kc >= 0 && kc < KC_COUNT) ? names[kc] : "";
}

Profiler& Profiler::get() {
  static Profiler* p = new Profiler();
  return *p;
}

void Profiler::enable(bool on) {
  std::lock_guard<std::mutex> lk(mu_);
  on_ = on;
  for (int i = 0; i < KC_COUNT; i++) {
    ms_[i] = 0.;
    n_[i] = 0;
    bytes_[i] = 0.;
  }
}

void Profiler::add(int kc, double ms, double bytes) {
  std::lock_guard<std::mutex> lk(mu_);
  ms_[kc] += ms;
  n_[kc] += 1;
  bytes_[kc] += bytes;
}

void Profiler::read(int kc, double* ms, long long* n, double* bytes) {
  std::lock_guard<std::mutex> lk(mu_);
  *ms = ms_[kc];
  *n = n_[kc];
  *bytes = bytes_[kc];
}

ProfScope::ProfScope(Context& ctx, int kc, double bytes)
    : ctx_(ctx), kc_(kc), bytes_(bytes) {
  if (!ctx.pooled || !Profiler::get().enabled_for(kc)) return;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(ctx_.stream, &cs) != hipSuccess ||
      cs != hipStreamCaptureStatusNone)
    return;   // no events inside a graph being captured
  TRVG_HIP(hipEventCreate(&a_));
  TRVG_HIP(hipEventRecord(a_, ctx_.stream));
}

ProfScope::~ProfScope() {
  if (a_ == nullptr) return;
  hipEvent_t b = nullptr;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a_);
    return;
  }
  (void)hipEventRecord(b, ctx_.stream);
  ctx_.prof.push_back(ProfRecord{kc_, bytes_, a_, b});
  if (ctx_.prof.size() > 4096) {
    try {
      ctx_.flush_profile();
    } catch (...) {
    }
  }
}

// ---------------------------------------------------------------------------
// Wigner 3-j (Racah's formula) and coupling (threept.cpp:65-71).
// ---------------------------------------------------------------------------

static double lfact(int n) { return std::lgamma(double(n) + 1.); }

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
    sum += ((k % 2 == 0) ? 1. : -1.) * std::exp(pre + t);
  }
  int ph = j1 - j2 - m3;
  return ((((ph % 2) + 2) % 2 == 0) ? 1. : -1.) * sum;
}

double coupling_3pt(int l1, int l2, int L, int m1, int m2, int M) {
  return double(2 * l1 + 1) * double(2 * l2 + 1) * double(2 * L + 1)
         * wigner_3j(l1, l2, L, 0, 0, 0) * wigner_3j(l1, l2, L, m1, m2, M);
}

int result_dim(int shape, int nb, int idx_bin) {
  switch (shape) {
    case TRVG_SHAPE_DIAG: case TRVG_SHAPE_ROW: return nb;
    case TRVG_SHAPE_OFFDIAG: return nb - std::abs(idx_bin);
    case TRVG_SHAPE_FULL: return nb * nb;
    case TRVG_SHAPE_TRIU: return nb * (nb + 1) / 2;
    default:
      throw Error(TRVG_ERR_PARAM, "Three-point statistic form is not recognised.");
  }
}

// (row bin, column bin) per data-vector entry (threept.cpp:465-770).
void pair_bins(int shape, int nb, int idx_bin, std::vector<int>& ra,
               std::vector<int>& cb) {
  int dim = result_dim(shape, nb, idx_bin);
  ra.assign(dim, 0);
  cb.assign(dim, 0);
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

// Mesh axes: even lengths in [8, 2048].  Powers of two up to 1024 take the
// compile-time FFT passes, other lengths the mixed-radix ones.  Odd lengths
// are refused: the octant histogram of the shot-noise terms and the
// Hermitian unpacking of packed transforms pair k with -k through a
// Nyquist plane (the reference's cutoff-derived meshes are always even,
// parameters.cpp:1434-1463).
static bool mesh_axis_ok(int n) { return n >= 8 && n <= 2048 && n % 2 == 0; }

void require_pow2_mesh(const trvg_params& p, const char* what) {
  for (int a = 0; a < 3; a++) {
    const int n = p.ngrid[a];
    if (n < 8 || n > 1024 || (n & (n - 1)) != 0)
      throw Error(TRVG_ERR_PARAM, std::string(what) +
                                      " needs mesh numbers that are powers of two in "
                                      "[8, 1024]: ngrid[" + std::to_string(a) + "] = " +
                                      std::to_string(n) + ".");
  }
}

void validate_params(const trvg_params& p, int num_bins) {
  for (int a = 0; a < 3; a++) {
    if (!(p.boxsize[a] > 0.))
      throw Error(TRVG_ERR_PARAM, "Box size must be positive.");
    if (!mesh_axis_ok(p.ngrid[a]))
      throw Error(TRVG_ERR_PARAM,
                  "Mesh number per dimension must be even and in [8, 2048]: "
                  "ngrid[" + std::to_string(a) + "] = " + std::to_string(p.ngrid[a]) +
                      ".");
  }
  long long nmesh = (long long)p.ngrid[0] * p.ngrid[1] * p.ngrid[2];
  if (nmesh < 2048)
    throw Error(TRVG_ERR_PARAM, "Mesh must have at least 2048 cells.");
  if (p.assignment_order < 1 || p.assignment_order > 4)
    throw Error(TRVG_ERR_PARAM,
                "Mesh assignment scheme must be 'ngp', 'cic', 'tsc' or 'pcs'.");
  if (p.precision != TRVG_PREC_FP64 && p.precision != TRVG_PREC_FP32)
    throw Error(TRVG_ERR_PARAM, "Precision must be 64 or 32.");
  if (p.ell1 < 0 || p.ell2 < 0 || p.ELL < 0)
    throw Error(TRVG_ERR_PARAM, "Multipole degrees must be non-negative.");
  if (num_bins <= 0)
    throw Error(TRVG_ERR_PARAM, "Number of bins must be positive.");
  if ((p.shape == TRVG_SHAPE_ROW || p.shape == TRVG_SHAPE_OFFDIAG) &&
      std::abs(p.idx_bin) >= num_bins)
    throw Error(TRVG_ERR_PARAM,
                "Bin index `idx_bin` must be < `num_bins` in absolute value.");
  if (p.shape == TRVG_SHAPE_ROW && p.idx_bin < 0)
    throw Error(TRVG_ERR_PARAM, "Row index `idx_bin` must be non-negative.");
  (void)result_dim(p.shape, num_bins, p.idx_bin);
  if (std::fabs(wigner_3j(p.ell1, p.ell2, p.ELL, 0, 0, 0)) < 1.e-9)
    throw Error(TRVG_ERR_PARAM,
                "Specified three-point correlator multipole vanishes "
                "identically owing to zero-valued Wigner 3-j symbol.");
}

}  // namespace trvg
