// engine.hpp -- host-side runtime of the gfx950 three-point engine:
// device memory pool, FFT plans (twiddle tables), per-axis window tables,
// spherical-Bessel splines and the templated estimator drivers.
#pragma once

#include <complex>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "fft.hpp"
#include "kernels.hpp"

namespace trvg {

using cplx = std::complex<double>;

// ---------------------------------------------------------------------------
// Device memory: a size-keyed cache of hipMalloc'ed blocks, reused across
// calls so that steady-state calls do not allocate.
// ---------------------------------------------------------------------------

class Pool {
 public:
  void* get(size_t bytes);
  void put(void* p, size_t bytes);
  void release_all();
  ~Pool() { release_all(); }

 private:
  std::mutex mu_;
  std::multimap<size_t, void*> free_;
};

Pool& pool();

template <typename E>
class DevArray {
 public:
  DevArray() = default;
  explicit DevArray(size_t n) { reset(n); }
  ~DevArray() { free_mem(); }
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
  DevArray(DevArray&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevArray& operator=(DevArray&& o) noexcept {
    if (this != &o) { free_mem(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  void reset(size_t n) {
    free_mem();
    n_ = n;
    p_ = n ? static_cast<E*>(pool().get(n * sizeof(E))) : nullptr;
  }
  E* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  void free_mem() {
    if (p_) pool().put(p_, n_ * sizeof(E));
    p_ = nullptr;
    n_ = 0;
  }
  E* p_ = nullptr;
  size_t n_ = 0;
};

// ---------------------------------------------------------------------------
// Device catalogue (owned copy of the columns the kernels read).
// ---------------------------------------------------------------------------

// Tile order of a catalogue for one mesh geometry (see k_assign_tiled):
// the catalogue's columns are permuted in place into tile order, particles
// inside the mesh first, then the stragglers (positions outside [0, L)).
// The chunked work list and its length are built on the device, so the
// sort needs no host round trip; `meta` holds {work items, particles
// inside the mesh}.  The host learns both after the first call's
// synchronisation and then launches exact grids.
struct TiledView {
  int n[3] = {0, 0, 0};
  double L[3] = {0., 0., 0.};
  TileGeom tg;
  DevArray<TileWork> work;
  DevArray<unsigned int> meta;
  long long work_cap = 0;   // bound on the work items (grid of the first calls)
  long long nwork = -1;     // exact count once known on the host, else -1
  long long nstraggle = -1; // likewise
  bool matches(const Geom& g) const {
    return n[0] == g.n[0] && n[1] == g.n[1] && n[2] == g.n[2] &&
           L[0] == g.L[0] && L[1] == g.L[1] && L[2] == g.L[2];
  }
};

// Device copy of the columns the kernels read.  The column order is the
// engine's (tile order of the last mesh geometry used), not the caller's:
// every use is a sum over particles.  A catalogue is not to be used by two
// calls at once.
struct DeviceCatalogue {
  long long n = 0;
  mutable DevArray<double> x, y, z, w, los;
  double wstotal = 0., wtotal = 0.;
  double wuni = NAN;   // the weight ws wc of every particle, NaN if they differ
  mutable std::unique_ptr<TiledView> tiled;  // built on first use
  DevParticles view() const {
    DevParticles p;
    p.x = x.get(); p.y = y.get(); p.z = z.get(); p.w = w.get();
    p.los = los.get();
    p.n = n;
    return p;
  }
};

struct Context;

// Tile sort of a device catalogue for the context's mesh in two phases, so
// that counting can follow an upload chunk by chunk: count() for each range
// of particles once its columns are on the device (stream order), then
// finish() (scan, in-place scatter into tile order, work list), which
// caches the result on the catalogue (c.tiled).  Nothing waits for the
// device.
class TileSorter {
 public:
  TileSorter(Context& ctx, const DeviceCatalogue& c);
  void count(long long begin, long long end);
  // Permutes the columns into tile order.  A weight column whose values
  // are all equal (move_w false) and lines of sight that the caller forms
  // afterwards from the sorted positions (move_los false) are not moved.
  void finish(bool move_w = true, bool move_los = true);

 private:
  Context& ctx_;
  const DeviceCatalogue& c_;
  std::unique_ptr<TiledView> tv_;
  DevArray<int> tile_of_;
  DevArray<unsigned int> ccount_, key2_;
};

// Streams of the engine (pooled per device, default priority).
hipStream_t acquire_stream(int dev);
void release_stream(int dev, hipStream_t s);

// Upload a (possibly strided) host catalogue; `los` optional (n x 3).  The
// device columns are valid in the order of stream s (the host buffers may
// be reused on return).  With a context `sort_for` (s is then its stream),
// the copies run on a separate stream, chunk by chunk, while the context's
// stream counts each arrived chunk for the tile sort of the context's mesh,
// which the upload completes; with los32 the lines of sight cross as FP32
// (FP32-field calls) and are widened on the device.  Without `los` and
// with `observer` (3 coordinates in the frame of the positions), the lines
// of sight are formed on the device from the positions relative to the
// observer, as the chunks arrive.  With `fixed_pos` (FP32-field calls
// with a context) the positions cross as 32-bit fixed-point offsets in
// [-L, 2L) per axis (step 3L / 2^32, 2.2e-6 Mpc/h at L = 3200 Mpc/h) and are
// decoded on the device to the nearest step centre; a chunk with a position
// outside that range crosses in FP64.
void upload_catalogue(const trvg_catalogue* c, const double* los,
                      DeviceCatalogue& out, hipStream_t s,
                      Context* sort_for = nullptr, bool los32 = false,
                      const double* observer = nullptr, bool fixed_pos = false);

// ---------------------------------------------------------------------------
// Spherical Bessel spline (host build, device tables).
// ---------------------------------------------------------------------------

struct Spline {
  int ell = 0;
  double split = 1000.;
  double step = 0.05;
  int n = 0;
  DevArray<double4> cf;
  DevArray<float4> cf32;
  SplineDev dev() const {
    SplineDev s;
    s.cf = cf.get(); s.cf32 = cf32.get();
    s.n = n; s.ell = ell; s.step = step; s.inv_step = 1. / step;
    s.split = split;
    return s;
  }
};

void build_spline(int ell, Spline& sp, hipStream_t s);

// Process-wide spline for degree ell on device dev (built on first use).
const Spline& cached_spline(int dev, int ell, hipStream_t s);

// Host natural-spline node values and coefficients (for tests).
void spline_nodes(int ell, std::vector<double>& y, std::vector<double>& c,
                  double& split);

// ---------------------------------------------------------------------------
// Kernel-class profiler: when enabled, every launch is bracketed by HIP
// events on the engine stream; elapsed times, launch counts and the
// algorithmic bytes of each launch are accumulated per class.
// ---------------------------------------------------------------------------

enum KernelClass {
  KC_FFT_Z = 0,        // first FFT pass (with fused loader), FP32 fields
  KC_FFT_STRIDED,      // y / x FFT passes, FP32 fields
  KC_ASSIGN,           // particle -> mesh assignment
  KC_PAIR_REDUCE,      // sum_x F_a F_b G over bin pairs
  KC_TWOPT,            // two-point binned statistics
  KC_SHELL_STATS,      // mode counts / k_eff
  KC_RADIAL,           // shot-noise radial histogram
  KC_SHOT_PAIRS,       // shot-noise per-pair sums
  KC_MISC,             // zero fills, partial sums, particle sums
  KC_FFT_Z64,          // first FFT pass, FP64 fields
  KC_FFT_STRIDED64,    // y / x FFT passes, FP64 fields
  KC_COUNT
};

const char* kernel_class_name(int kc);

struct ProfRecord {
  int kc;
  double bytes;
  hipEvent_t a, b;
};

class Profiler {
 public:
  static Profiler& get();
  bool enabled() const { return on_; }
  bool enabled_for(int kc) const { return on_ && (only_ < 0 || only_ == kc); }
  void enable(bool on);
  void only(int kc) { only_ = kc; }
  void add(int kc, double ms, double bytes);
  void read(int kc, double* ms, long long* n, double* bytes);

 private:
  std::mutex mu_;
  bool on_ = false;
  int only_ = -1;
  double ms_[KC_COUNT] = {};
  long long n_[KC_COUNT] = {};
  double bytes_[KC_COUNT] = {};
};

// ---------------------------------------------------------------------------
// Per-call context.
// ---------------------------------------------------------------------------

// Pinned host memory for the asynchronous transfers of one engine call: a
// bump allocator over page-locked chunks, one arena per pooled stream
// (reset when a call takes the stream, after the previous call has
// synchronised it).
class PinnedArena {
 public:
  void* alloc(size_t bytes);
  void reset() { chunk_ = 0; used_ = 0; }
  ~PinnedArena();

 private:
  struct Chunk {
    char* p;
    size_t size;
  };
  std::vector<Chunk> chunks_;
  size_t chunk_ = 0, used_ = 0;
};

struct Context {
  trvg_params p;
  Geom g;
  hipStream_t stream = nullptr;
  int dev = 0;
  AxisTables tab;               // per-axis sinc, sin^2, u (cached tables)
  int sm_count = 256;
  std::vector<ProfRecord> prof;
  PinnedArena* pin = nullptr;
  // Host work waiting for device results, and device buffers those results
  // are still being copied out of: both released by flush().
  std::vector<std::function<void()>> deferred;
  std::vector<std::shared_ptr<void>> keep;
  // False for a context on borrowed resources (stream, axis tables,
  // twiddles): it takes nothing from the engine's pools and records no
  // profiler events.
  bool pooled = true;
  // Borrowed twiddle tables by sign (-1: [0], +1: [1]) and log2 length;
  // null entries use the engine's cached tables.
  const void* tw[2][11] = {};

  explicit Context(const trvg_params& params);
  // A context on borrowed resources; allocates nothing.
  Context(const trvg_params& params, hipStream_t s, const AxisTables& t);
  ~Context();
  // Resolve pending profiler records (synchronises the stream).
  void flush_profile();

  // Host -> device without blocking the host (staged in pinned memory).
  template <typename E>
  void upload(E* dst, const E* src, size_t n) {
    if (n == 0) return;
    void* h = pin->alloc(n * sizeof(E));
    std::memcpy(h, src, n * sizeof(E));
    TRVG_HIP(hipMemcpyAsync(dst, h, n * sizeof(E), hipMemcpyHostToDevice, stream));
  }
  // Device -> host into pinned memory; the values are valid after flush().
  template <typename E>
  const E* download(const E* src, size_t n) {
    E* h = static_cast<E*>(pin->alloc((n ? n : 1) * sizeof(E)));
    if (n) TRVG_HIP(hipMemcpyAsync(h, src, n * sizeof(E), hipMemcpyDeviceToHost, stream));
    return h;
  }
  template <typename E>
  void keep_until_flush(DevArray<E>&& a) {
    keep.push_back(std::make_shared<DevArray<E>>(std::move(a)));
  }
  void defer(std::function<void()> f) { deferred.push_back(std::move(f)); }
  // Synchronise the stream, then run the deferred host work in order.
  void flush();
};

// Brackets one kernel launch with profiler events (no-op when disabled).
class ProfScope {
 public:
  ProfScope(Context& ctx, int kc, double bytes);
  ~ProfScope();

 private:
  Context& ctx_;
  int kc_;
  double bytes_;
  hipEvent_t a_ = nullptr;
};

// Least double s >= 0 with sqrt_rn(s) >= t (-1 when every s passes), and
// the band of a bin [k_lo, k_hi) in |k|^2.
double sq_threshold(double t);
inline ShellBand shell_band(double k_lo, double k_hi) {
  return ShellBand{sq_threshold(k_lo), sq_threshold(k_hi)};
}

// Launch helpers shared by the drivers.
inline unsigned grid_for(const Context& ctx, long long work, int per_cu = 8) {
  long long b = (work + RED_THREADS - 1) / RED_THREADS;
  long long cap = (long long)ctx.sm_count * per_cu;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return unsigned(b);
}

// Column sums of a [nrows][width] partials array (fixed order, so results
// do not depend on scheduling).
template <typename V>
inline void reduce_cols(Context& ctx, const V* part, int nrows, int width, V* out) {
  hipLaunchKernelGGL((k_reduce_cols<V>), dim3((width + RC_COLS - 1) / RC_COLS),
                     dim3(RC_COLS * RC_ROWS), 0, ctx.stream, part, nrows, width, out);
  TRVG_LAUNCH_CHECK();
}

// Mode counts and mean |k| of each k-bin (exactly rounded |k|; k_eff is
// NaN for an empty bin, as the reference divides by zero).
struct ShellStats {
  std::vector<int> nmodes;
  std::vector<double> keff;
};
ShellStats shell_stats(Context& ctx, const trvg_binning& kb);

// Coarse bin of each half-window cell of the z lines for the two-point
// statistics (see TwoptMap); built once per device, geometry and edges.
struct TwoptMapData {
  DevArray<int> K;
  DevArray<long long> off;
  DevArray<unsigned short> bin;
  TwoptMap view() const { return {K.get(), off.get(), bin.get()}; }
};
std::shared_ptr<const TwoptMapData> twopt_map(Context& ctx, const double* edges, int nb);

// Maths shared with the drivers.
double wigner_3j(int j1, int j2, int j3, int m1, int m2, int m3);
double coupling_3pt(int l1, int l2, int L, int m1, int m2, int M);
int result_dim(int shape, int nb, int idx_bin);
void pair_bins(int shape, int nb, int idx_bin, std::vector<int>& ra,
               std::vector<int>& cb);
void validate_params(const trvg_params& p, int num_bins);
// Batched plans, the differentiable path and the allocation-free entry
// point keep power-of-two meshes only.
void require_pow2_mesh(const trvg_params& p, const char* what);

// FFT of a mesh: first pass (z) loads through `Loader`, then y and x passes.
// read_bytes: compulsory bytes the loader reads (profiler accounting; a
// negative value means one in-place read of the mesh).
template <typename T, class Loader>
void fft3d(Context& ctx, c2<T>* data, const Loader& load, int sign,
           double read_bytes = -1., const Prune* prune = nullptr);

// Pruning descriptor for an inverse transform whose input vanishes for
// |k| >= k_hi (disabled when the band covers the whole kx range).
Prune make_prune(const Geom& g, double k_hi);

// Zero `mesh` and assign catalogue `c` with weight mode (WeightMode) and
// scale through the production (tile-binned) path.
template <typename T>
void assign_field(Context& ctx, c2<T>* mesh, const DeviceCatalogue& c, int mode,
                  int ell, int m, double scale);

// Add catalogue `c` with weight mode and scale into `mesh` (no zeroing).
template <typename T>
void assign_add(Context& ctx, c2<T>* mesh, const DeviceCatalogue& c, int mode,
                int ell, int m, double scale);

// Two-point statistics (engine_twopt.hip).  Binned values of one
// power-spectrum or correlation-function multipole.
struct TwoPtResult {
  std::vector<double> bin, eff;
  std::vector<int> count;
  std::vector<cplx> raw, shot;
};
// Survey pair (d, r) or, with d null, the window of one random catalogue;
// fourier selects the power spectrum, else the correlation function.
TwoPtResult run_twopt_survey(Context& ctx, const DeviceCatalogue* d,
                             const DeviceCatalogue& r, bool fourier,
                             const trvg_binning& b, double alpha, double norm);
// Periodic box (global plane-parallel).
TwoPtResult run_twopt_box(Context& ctx, const DeviceCatalogue& d, bool fourier,
                          const trvg_binning& b, double norm);
// sum_x n(x)^2 V_cell of the w-weighted mesh, and sum_x n_d(x) n_r(x)
// V_cell of two catalogues (the mesh normalisations of the power spectrum).
double mesh_square_integral(Context& ctx, const DeviceCatalogue& c);
double mesh_cross_integral(Context& ctx, const DeviceCatalogue& d,
                           const DeviceCatalogue& r);

// Estimators.
struct Result {
  int dim = 0;
  std::vector<double> b1, b2, e1, e2;
  std::vector<int> n1, n2;
  std::vector<cplx> raw, shot;
};

// One multipole of a multi-multipole call.
struct Multipole {
  int ell1, ell2, ELL, shape, idx_bin;
  double norm;
};

// Data-catalogue fields of a survey bispectrum call, assigned while the
// random catalogue is still crossing from the host.  For catalogues of
// uniform weights the survey pair of each (L, M) starts from
// S_d = sum_d y_LM w W, which needs no random particle; the estimator takes
// each mesh (moves it out) at its first use and, when the random catalogue
// turns out not to have uniform weights, uses none of them.
struct SurveyPrefill {
  std::map<std::pair<int, int>, DevArray<double2>> sd;
};
// S_d of every (L, M) that `mps` use, in order of use, while their meshes
// stay within 4 GiB; empty unless `d` has uniform weights.
SurveyPrefill prefill_survey_data(Context& ctx, const DeviceCatalogue& d,
                                  const std::vector<Multipole>& mps);

// Several bispectrum multipoles on one catalogue pair, sharing the common
// fields, shell statistics and shell slabs (`pre`: data fields assigned
// ahead, or null).
template <typename T>
std::vector<Result> run_bispec_multi(Context& ctx, const DeviceCatalogue& d,
                                     const DeviceCatalogue* r, bool survey,
                                     const trvg_binning& kb,
                                     const std::vector<Multipole>& mps,
                                     SurveyPrefill* pre = nullptr);

// Batched periodic-box bispectrum plans (engine_batch.hip).
struct BatchPlanBase {
  virtual ~BatchPlanBase() = default;
  // Catalogues b < nmocks with counts[b] particles, positions concatenated;
  // outs[b * nmultipoles + q] receives multipole q of catalogue b.
  virtual void run(int nmocks, const long long* counts, const double* x,
                   const double* y, const double* z, trvg_bispec_result* outs) = 0;
  // As run(), with x / y / z device pointers.
  virtual void run_device(int nmocks, const long long* counts, const double* x,
                          const double* y, const double* z,
                          trvg_bispec_result* outs) = 0;
  // tuned[0..4]: assignment / pair / statistics blocks per catalogue, pair
  // row-block cap, graph captured (0/1).
  virtual void info(int* tuned) const = 0;
};
BatchPlanBase* make_batch_plan(const trvg_params& p, const std::vector<Multipole>& mps,
                               const trvg_binning& kb, int batch, long long capacity,
                               int autotune, int use_graph);

// Differentiable periodic-box B_000 diag form (engine_diff.hip).
template <typename T>
std::vector<double> diff_bispec_forward(Context& ctx, const trvg_binning& kb,
                                        double norm, const T* mesh,
                                        std::vector<int>& nmodes);
template <typename T>
void diff_bispec_backward(Context& ctx, const trvg_binning& kb, double norm,
                          const T* mesh, const double* grad_bk, T* grad_mesh);
template <typename T>
void diff_assign(Context& ctx, T* mesh, const double* pos, const double* w,
                 long long n, bool accumulate);
template <typename T>
void diff_assign_adjoint(Context& ctx, const T* gmesh, const double* pos,
                         const double* w, long long n, double* gpos, double* gw);

// The forward above without any device allocation: every buffer in the
// library's fixed scratch array (engine_noalloc.hip); throws
// TRVG_ERR_PARAM when the mesh or the binning exceeds it.
void diff_bispec_forward_noalloc(const trvg_params& q, const trvg_binning& kb,
                                 double norm, const void* mesh, double* bk,
                                 int* nmodes);
size_t noalloc_scratch_bytes();

template <typename T>
Result run_bispec(Context& ctx, const DeviceCatalogue& d,
                  const DeviceCatalogue* r, bool survey,
                  const trvg_binning& kb, double norm_factor,
                  SurveyPrefill* pre = nullptr);

// The window of one random catalogue (threept.cpp:2621-3077): fields of
// the catalogue alone scaled by alpha, and the wide-angle kernel
// |x|^-wa_power on G when wa_power >= 0.
struct Window3pcf {
  double alpha = 1.;
  int wa_power = -1;
};

template <typename T>
Result run_3pcf(Context& ctx, const DeviceCatalogue& d,
                const DeviceCatalogue* r, bool survey,
                const trvg_binning& rb, double norm_factor,
                const Window3pcf* win = nullptr);

template <typename T>
double run_mesh_norm(Context& ctx, const DeviceCatalogue& c);

}  // namespace trvg
