// upload.hip -- host -> HBM transfer of particle catalogues.
//
// The caller's catalogue is host memory of any kind (numpy columns, the
// 56-byte records of trv::ParticleData viewed with a stride, pageable).  A
// direct copy from pageable memory runs through the runtime's own staging
// at a few GB/s, and a single host thread gathering the columns is slower
// still.  Here a pool of host threads gathers fixed-size chunks of the
// columns (x, y, z, w = ws * wc and the lines of sight) into a ring of
// page-locked buffers, and each chunk is copied to the device
// asynchronously on the call's stream while the next chunk is gathered.
// The weight column of a catalogue whose particles all carry the same
// ws * wc is not transferred: the device fills it.  The host returns once
// every chunk is gathered; the copies complete in stream order before the
// kernels that read them.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdlib>
#include <map>
#include <thread>

#include "engine.hpp"

namespace trvg {

// ---------------------------------------------------------------------------
// Host worker pool
// ---------------------------------------------------------------------------

namespace {

int host_thread_count() {
  for (const char* var : {"TRVG_HOST_THREADS", "OMP_NUM_THREADS"}) {
    const char* v = std::getenv(var);
    if (v && std::atoi(v) > 0) return std::min(std::atoi(v), 64);
  }
  unsigned hw = std::thread::hardware_concurrency();
  return int(std::max(1u, std::min(hw, 16u)));
}

// Persistent threads running one job at a time: run(fn) calls fn(t) for
// t = 0 .. size()-1 (t = 0 on the calling thread) and returns when every
// call has returned.  Jobs from different host threads are serialised.
class HostWorkers {
 public:
  static HostWorkers& get() {
    static HostWorkers* w = new HostWorkers(host_thread_count());
    return *w;
  }
  int size() const { return nthreads_; }

  void run(const std::function<void(int)>& fn) {
    std::lock_guard<std::mutex> serial(run_mu_);
    if (nthreads_ == 1) {
      fn(0);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &fn;
      pending_ = nthreads_ - 1;
      gen_++;
    }
    cv_.notify_all();
    fn(0);
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return pending_ == 0; });
    job_ = nullptr;
  }

 private:
  explicit HostWorkers(int n) : nthreads_(n) {
    for (int t = 1; t < n; t++) {
      std::thread th([this, t] { loop(t); });
      th.detach();   // lives for the process
    }
  }
  void loop(int t) {
    long seen = 0;
    for (;;) {
      const std::function<void(int)>* job;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        job = job_;
      }
      (*job)(t);
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (--pending_ == 0) done_cv_.notify_one();
      }
    }
  }

  int nthreads_;
  std::mutex run_mu_, mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int)>* job_ = nullptr;
  long gen_ = 0;
  int pending_ = 0;
};

// ---------------------------------------------------------------------------
// Page-locked staging ring (one per stream: a stream serves one call at a
// time)
// ---------------------------------------------------------------------------

constexpr int RING_SLOTS = 3;
constexpr size_t RING_BYTES = size_t(32) << 20;   // per slot

struct StagingRing {
  char* buf[RING_SLOTS] = {};
  hipEvent_t ev[RING_SLOTS] = {};
  bool pending[RING_SLOTS] = {};
  int dev = -1;
  ~StagingRing() {
    for (int i = 0; i < RING_SLOTS; i++) {
      if (buf[i]) (void)hipHostFree(buf[i]);
      if (ev[i]) (void)hipEventDestroy(ev[i]);
    }
    if (tail) (void)hipEventDestroy(tail);
  }
  // Slot k, once the copies that last read it have completed.
  char* acquire(int k) {
    if (buf[k] == nullptr) {
      TRVG_HIP(hipHostMalloc((void**)&buf[k], RING_BYTES));
      TRVG_HIP(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
    }
    if (pending[k]) TRVG_HIP(hipEventSynchronize(ev[k]));
    pending[k] = false;
    return buf[k];
  }
  void release(int k, hipStream_t s) {
    TRVG_HIP(hipEventRecord(ev[k], s));
    pending[k] = true;
  }
  // Event for the end of an upload's copies (waited on by other streams).
  hipEvent_t tail_event() {
    if (tail == nullptr) TRVG_HIP(hipEventCreateWithFlags(&tail, hipEventDisableTiming));
    return tail;
  }
  hipEvent_t tail = nullptr;
};

StagingRing& staging_ring(hipStream_t s) {
  static std::mutex mu;
  static auto* rings = new std::map<hipStream_t, std::unique_ptr<StagingRing>>();
  std::lock_guard<std::mutex> lk(mu);
  auto& r = (*rings)[s];
  if (!r) r = std::make_unique<StagingRing>();
  return *r;
}

__global__ void k_fill_d(double* __restrict__ f, long long n, double v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    f[i] = v;
}

__global__ void k_widen(double* __restrict__ d, const float* __restrict__ f, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    d[i] = double(f[i]);
}

void widen_device(double* d, const float* f, long long n, hipStream_t s) {
  if (n <= 0) return;
  long long blocks = std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_widen, dim3(unsigned(blocks)), dim3(256), 0, s, d, f, n);
  TRVG_LAUNCH_CHECK();
}

void fill_device(double* f, long long n, double v, hipStream_t s) {
  if (n <= 0) return;
  long long blocks = std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_d, dim3(unsigned(blocks)), dim3(256), 0, s, f, n, v);
  TRVG_LAUNCH_CHECK();
}

// Fixed-point positions of one chunk ([x block, y block, z block] of m
// words each) to FP64: x = lo + (q + 1/2) step per axis.
struct FixedAxes {
  double lo[3], step[3], inv[3];
};

__global__ void k_unfix(const unsigned int* __restrict__ q, long long m, FixedAxes f,
                        double* __restrict__ x, double* __restrict__ y,
                        double* __restrict__ z) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < m;
       i += (long long)gridDim.x * blockDim.x) {
    x[i] = f.lo[0] + (double(q[i]) + 0.5) * f.step[0];
    y[i] = f.lo[1] + (double(q[m + i]) + 0.5) * f.step[1];
    z[i] = f.lo[2] + (double(q[2 * m + i]) + 0.5) * f.step[2];
  }
}

// Lines of sight of particles [lo, hi) from their positions relative to the
// observer o: (r - o) / |r - o|, a zero norm taken as 1, in the operation
// order of ParticleCatalogue.compute_los (catalogue.py:437-458:
// sqrt(x^2 + y^2 + z^2), then each component divided by it), every
// operation rounded as on the host.
__global__ void k_los_observer(const double* __restrict__ x, const double* __restrict__ y,
                               const double* __restrict__ z, double* __restrict__ los,
                               long long lo, long long hi, double ox, double oy, double oz) {
  for (long long i = lo + blockIdx.x * (long long)blockDim.x + threadIdx.x; i < hi;
       i += (long long)gridDim.x * blockDim.x) {
    const double a = __dsub_rn(x[i], ox), b = __dsub_rn(y[i], oy), c = __dsub_rn(z[i], oz);
    double r = __dsqrt_rn(__dadd_rn(__dadd_rn(__dmul_rn(a, a), __dmul_rn(b, b)),
                                    __dmul_rn(c, c)));
    if (r == 0.) r = 1.;
    los[3 * i] = __ddiv_rn(a, r);
    los[3 * i + 1] = __ddiv_rn(b, r);
    los[3 * i + 2] = __ddiv_rn(c, r);
  }
}

void los_from_observer(DeviceCatalogue& d, const double* o, long long lo, long long hi,
                       hipStream_t s) {
  if (hi <= lo) return;
  long long blocks = std::min<long long>((hi - lo + 255) / 256, 4096);
  hipLaunchKernelGGL(k_los_observer, dim3(unsigned(blocks)), dim3(256), 0, s, d.x.get(),
                     d.y.get(), d.z.get(), d.los.get(), lo, hi, o[0], o[1], o[2]);
  TRVG_LAUNCH_CHECK();
}

// Per-thread partial results of one chunk (padded to a cache line).
struct alignas(64) Partial {
  double wst, wt;
  int uniform;
};

// Staging buffers of one chunk of m particles: FP64 columns x, y, z, w, or
// the three fixed-point blocks q (x, y, z words) and w.
struct SlotCols {
  double *x, *y, *z, *w;
  unsigned int* q;
  long long m;
};

// Gather of particles base + [lo, hi) into a slot (positions only; the
// weights are gathered by gather_weights).  FIXED: positions as 32-bit
// fixed point, returning whether a position lay outside [-L, 2L) (its word
// is then 0 and the chunk must cross in FP64).  CONTIG: unit stride, so
// the loop reads the columns directly and vectorises.  Stores are
// streaming: the host never reads the staging buffer back.
template <bool FIXED, bool CONTIG>
int gather_positions(const trvg_catalogue* c, long long st, long long base, long long lo,
                     long long hi, const FixedAxes& fx, const SlotCols& s) {
  const double* __restrict__ X = c->x;
  const double* __restrict__ Y = c->y;
  const double* __restrict__ Z = c->z;
  if (CONTIG) {
    X += base;
    Y += base;
    Z += base;
  }
  const long long m = s.m;
  if constexpr (FIXED) {
    unsigned int* __restrict__ qx = s.q;
    unsigned int* __restrict__ qy = s.q + m;
    unsigned int* __restrict__ qz = s.q + 2 * m;
    const double l0 = fx.lo[0], l1 = fx.lo[1], l2 = fx.lo[2];
    const double i0 = fx.inv[0], i1 = fx.inv[1], i2 = fx.inv[2];
    int bad = 0;
    for (long long i = lo; i < hi; i++) {
      const long long g = CONTIG ? i : (base + i) * st;
      const double u0 = (X[g] - l0) * i0, u1 = (Y[g] - l1) * i1, u2 = (Z[g] - l2) * i2;
      const bool k0 = u0 >= 0. && u0 < 4294967296.;
      const bool k1 = u1 >= 0. && u1 < 4294967296.;
      const bool k2 = u2 >= 0. && u2 < 4294967296.;
      bad |= int(!(k0 && k1 && k2));
      __builtin_nontemporal_store(k0 ? (unsigned int)u0 : 0u, &qx[i]);
      __builtin_nontemporal_store(k1 ? (unsigned int)u1 : 0u, &qy[i]);
      __builtin_nontemporal_store(k2 ? (unsigned int)u2 : 0u, &qz[i]);
    }
    return bad;
  } else {
    double* __restrict__ bx = s.x;
    double* __restrict__ by = s.y;
    double* __restrict__ bz = s.z;
    for (long long i = lo; i < hi; i++) {
      const long long g = CONTIG ? i : (base + i) * st;
      __builtin_nontemporal_store(X[g], &bx[i]);
      __builtin_nontemporal_store(Y[g], &by[i]);
      __builtin_nontemporal_store(Z[g], &bz[i]);
    }
    return 0;
  }
}

// Weights w = ws * wc (particles.cpp:553) of particles base + [lo, hi):
// the sums of ws and w, and whether every w equals w0.  The weight column
// is written only once a weight differs from w0 (a uniform chunk is not
// copied); the range's uniform prefix is filled then.  Absent ws and wc
// columns are unit weights: nothing to read or write.
Partial gather_weights(const trvg_catalogue* c, long long st, long long base, long long lo,
                       long long hi, double w0, double* __restrict__ bw) {
  if (c->ws == nullptr && c->wc == nullptr)
    return Partial{double(hi - lo), double(hi - lo), w0 == 1.};
  double ps = 0., pw = 0.;
  int uni = 1;
  for (long long i = lo; i < hi; i++) {
    const long long g = (base + i) * st;
    const double ws = c->ws ? c->ws[g] : 1.;
    const double wc = c->wc ? c->wc[g] : 1.;
    const double w = ws * wc;
    if (uni && w != w0) {
      for (long long k = lo; k < i; k++) bw[k] = w0;
      uni = 0;
    }
    if (!uni) __builtin_nontemporal_store(w, &bw[i]);
    ps += ws;
    pw += w;
  }
  return Partial{ps, pw, uni};
}

}  // namespace

// ---------------------------------------------------------------------------
// Catalogue upload
// ---------------------------------------------------------------------------

void upload_catalogue(const trvg_catalogue* c, const double* los,
                      DeviceCatalogue& out, hipStream_t s, Context* sort_for,
                      bool los32, const double* observer, bool fixed_pos) {
  if (c == nullptr || c->x == nullptr || c->y == nullptr || c->z == nullptr)
    throw Error(TRVG_ERR_DATA, "Particle data are uninitialised.");
  if (c->ntotal <= 0)
    throw Error(TRVG_ERR_DATA, "Particle catalogue is empty.");
  const long long n = c->ntotal;
  const long long st = c->stride > 0 ? c->stride : 1;
  out.tiled.reset();
  out.n = n;
  out.x.reset(n); out.y.reset(n); out.z.reset(n); out.w.reset(n);
  const bool derive = los == nullptr && observer != nullptr;
  if (derive && !(std::isfinite(observer[0]) && std::isfinite(observer[1]) &&
                  std::isfinite(observer[2])))
    throw Error(TRVG_ERR_PARAM, "Observer position must be finite.");
  if (los || derive) out.los.reset(3 * n);
  else out.los.reset(0);
  los32 = los32 && los != nullptr && sort_for != nullptr;
  fixed_pos = fixed_pos && sort_for != nullptr;
  FixedAxes fx{};
  if (fixed_pos) {
    for (int a = 0; a < 3; a++) {
      const double L = sort_for->g.L[a];
      fx.lo[a] = -L;
      fx.step[a] = 3. * L / 4294967296.;
      fx.inv[a] = 4294967296. / (3. * L);
    }
  }

  // Copies on their own stream when the upload feeds a tile sort, so that
  // the counting of chunk k overlaps the copy of chunk k + 1.
  hipStream_t cs = s;
  int dev = 0;
  if (sort_for) {
    dev = sort_for->dev;
    cs = acquire_stream(dev);
  }
  std::unique_ptr<TileSorter> sorter;
  if (sort_for) sorter = std::make_unique<TileSorter>(*sort_for, out);
  DevArray<float> los_f(los32 ? 3 * n : 0);   // FP32 lines of sight (widened)
  DevArray<unsigned int> pos_q(fixed_pos ? 3 * n : 0);   // fixed-point positions

  // w = ws * wc (particles.cpp:553); the reference weight of the first
  // particle decides uniformity.
  const double w0 = (c->ws ? c->ws[0] : 1.) * (c->wc ? c->wc[0] : 1.);
  const size_t pbytes = 32 + (los ? (los32 ? 12 : 24) : 0);
  const long long chunk = (long long)(RING_BYTES / pbytes);
  StagingRing& ring = staging_ring(cs);
  HostWorkers& pool = HostWorkers::get();
  const int nt = pool.size();
  std::vector<Partial> part(nt);
  double wst = 0., wt = 0.;
  bool uniform = true;   // every chunk so far has w == w0 throughout
  int slot = 0;
  try {
    for (long long base = 0; base < n; base += chunk, slot = (slot + 1) % RING_SLOTS) {
      const long long m = std::min(chunk, n - base);
      // Slot layout: FP64 x, y, z, w, then the lines of sight; with fixed-
      // point positions w, then x, y, z as words, then the lines of sight.
      char* slot_buf = ring.acquire(slot);
      double* bx = reinterpret_cast<double*>(slot_buf);
      double* by = bx + m;
      double* bz = by + m;
      double* bw = bz + m;
      double* bl = bw + m;
      unsigned int* bq = nullptr;
      bool fixed_chunk = fixed_pos;
      if (fixed_chunk) {
        bw = reinterpret_cast<double*>(slot_buf);
        bq = reinterpret_cast<unsigned int*>(bw + m);
        bl = reinterpret_cast<double*>(slot_buf + size_t(m) * 20 + 7 - (size_t(m) * 20 + 7) % 8);
      }
      float* bl32 = reinterpret_cast<float*>(bl);
      std::vector<int> out_of_range(nt, 0);
      const SlotCols cols{bx, by, bz, bw, bq, m};
      pool.run([&](int t) {
        const long long lo = m * t / nt, hi = m * (t + 1) / nt;
        int bad;
        if (fixed_chunk)
          bad = st == 1 ? gather_positions<true, true>(c, st, base, lo, hi, fx, cols)
                        : gather_positions<true, false>(c, st, base, lo, hi, fx, cols);
        else
          bad = st == 1 ? gather_positions<false, true>(c, st, base, lo, hi, fx, cols)
                        : gather_positions<false, false>(c, st, base, lo, hi, fx, cols);
        out_of_range[t] = bad;
        part[t] = gather_weights(c, st, base, lo, hi, w0, bw);
        if (los && hi > lo) {
          if (los32) {
            const double* src = los + 3 * (base + lo);
            for (long long q = 0; q < 3 * (hi - lo); q++)
              bl32[3 * lo + q] = float(src[q]);
          } else {
            std::memcpy(bl + 3 * lo, los + 3 * (base + lo), size_t(hi - lo) * 24);
          }
        }
        // Orders the streaming stores before the copy reads the buffer.
        __builtin_ia32_sfence();
      });
      bool refetch = false;
      for (int t = 0; t < nt; t++) refetch = refetch || out_of_range[t];
      if (refetch) {
        // A position outside [-L, 2L): this chunk crosses in FP64 (the
        // slot holds 32 bytes per particle plus the lines of sight).
        fixed_chunk = false;
        bx = reinterpret_cast<double*>(slot_buf);
        by = bx + m;
        bz = by + m;
        bw = bz + m;
        double* bl64 = bw + m;
        pool.run([&](int t) {
          const long long lo = m * t / nt, hi = m * (t + 1) / nt;
          for (long long i = lo; i < hi; i++) {
            const long long g = (base + i) * st;
            bx[i] = c->x[g];
            by[i] = c->y[g];
            bz[i] = c->z[g];
            bw[i] = (c->ws ? c->ws[g] : 1.) * (c->wc ? c->wc[g] : 1.);
          }
          if (los && hi > lo) {
            if (los32) {
              const double* src = los + 3 * (base + lo);
              float* d32 = reinterpret_cast<float*>(bl64);
              for (long long q = 0; q < 3 * (hi - lo); q++) d32[3 * lo + q] = float(src[q]);
            } else {
              std::memcpy(bl64 + 3 * lo, los + 3 * (base + lo), size_t(hi - lo) * 24);
            }
          }
        });
        bl = bl64;
        bl32 = reinterpret_cast<float*>(bl);
      }
      bool chunk_uniform = true;
      for (int t = 0; t < nt; t++) {   // fixed order: deterministic sums
        wst += part[t].wst;
        wt += part[t].wt;
        chunk_uniform = chunk_uniform && part[t].uniform;
      }
      if (!chunk_uniform && !refetch) {
        // Ranges that stayed uniform left their weights unwritten.
        pool.run([&](int t) {
          if (!part[t].uniform) return;
          const long long lo = m * t / nt, hi = m * (t + 1) / nt;
          for (long long i = lo; i < hi; i++) bw[i] = w0;
        });
      }
      const size_t cb = size_t(m) * 8;
      if (fixed_chunk) {
        TRVG_HIP(hipMemcpyAsync(pos_q.get() + 3 * base, bq, 3 * size_t(m) * 4,
                                hipMemcpyHostToDevice, cs));
        const long long blocks = std::min<long long>((m + 255) / 256, 4096);
        hipLaunchKernelGGL(k_unfix, dim3(unsigned(blocks)), dim3(256), 0, cs,
                           pos_q.get() + 3 * base, m, fx, out.x.get() + base,
                           out.y.get() + base, out.z.get() + base);
        TRVG_LAUNCH_CHECK();
      } else {
        TRVG_HIP(hipMemcpyAsync(out.x.get() + base, bx, cb, hipMemcpyHostToDevice, cs));
        TRVG_HIP(hipMemcpyAsync(out.y.get() + base, by, cb, hipMemcpyHostToDevice, cs));
        TRVG_HIP(hipMemcpyAsync(out.z.get() + base, bz, cb, hipMemcpyHostToDevice, cs));
      }
      if (!(uniform && chunk_uniform)) {
        if (uniform) fill_device(out.w.get(), base, w0, cs);   // the uniform prefix
        uniform = false;
        TRVG_HIP(hipMemcpyAsync(out.w.get() + base, bw, cb, hipMemcpyHostToDevice, cs));
      }
      if (los32) {
        TRVG_HIP(hipMemcpyAsync(los_f.get() + 3 * base, bl32, 3 * size_t(m) * 4,
                                hipMemcpyHostToDevice, cs));
        widen_device(out.los.get() + 3 * base, los_f.get() + 3 * base, 3 * m, cs);
      } else if (los) {
        TRVG_HIP(hipMemcpyAsync(out.los.get() + 3 * base, bl, 3 * cb,
                                hipMemcpyHostToDevice, cs));
      }
      // With a tile sort, lines of sight are formed once from the sorted
      // positions instead of being permuted with them.
      if (derive && !sorter) los_from_observer(out, observer, base, base + m, cs);
      ring.release(slot, cs);
      if (sorter) {
        // The positions of this chunk are on the device once the slot's
        // event has passed (w is not read by the count).
        TRVG_HIP(hipStreamWaitEvent(s, ring.ev[slot], 0));
        sorter->count(base, base + m);
      }
    }
    if (uniform) fill_device(out.w.get(), n, w0, cs);
    if (sorter) {
      TRVG_HIP(hipEventRecord(ring.tail_event(), cs));
      TRVG_HIP(hipStreamWaitEvent(s, ring.tail_event(), 0));
      sorter->finish(/*move_w=*/!uniform, /*move_los=*/los != nullptr);
      if (derive) los_from_observer(out, observer, 0, n, s);
      sort_for->keep_until_flush(std::move(los_f));
      sort_for->keep_until_flush(std::move(pos_q));
      release_stream(dev, cs);
    }
  } catch (...) {
    if (sort_for) {
      (void)hipStreamSynchronize(cs);
      release_stream(dev, cs);
    }
    throw;
  }
  out.wstotal = wst;
  out.wtotal = wt;
  out.wuni = uniform ? w0 : NAN;
}

}  // namespace trvg
