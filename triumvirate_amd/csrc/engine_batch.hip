// engine_batch.hip -- batched periodic-box bispectrum over many catalogues
// ("mocks") of one geometry: a plan holds every buffer and table for B
// mocks, the launch parameters picked by an autotuner, and a HIP graph of
// the whole batch, captured once and replayed for every batch.
//
// Per batch the sequence is that of compute_bispec_in_gpp_box
// (threept.cpp:1473-2188) for each requested multipole, with every kernel
// spanning all mocks: assignment, dn00(k) and N00(k), the shell slabs of
// every (l, m) side (shared across multipoles), G(x), the pair reductions,
// the two-point and S_ij_k shot-noise terms, and the normalisation.  Host
// work per batch: pack positions, one copy in, one graph launch, one copy
// out.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "batch_kernels.hpp"
#include "engine.hpp"

namespace trvg {

// Declared in engine_drivers.hip.
template <typename T>
void shell_slab_work(const Geom& g, const trvg_binning& kb, int nbins,
                     int nfield_groups, bool packed, std::vector<ShellBand>& bands,
                     std::vector<double>& kcut2, std::vector<int2>& wz,
                     std::vector<int2>& wy, std::vector<int2>& wx);
template <typename T, class BatchLoader>
void fft3d_batch(Context& ctx, c2<T>* base, const BatchLoader& load, int sign,
                 bool prune, const double* kcut2, const int2* wz, int nz,
                 const int2* wy, int ny, const int2* wx, int nx);

namespace {

struct Term {
  int m1, m2, M;
  double coupling, f_mirror, f_mirror_w3j;
};

// Terms of a periodic-box multipole (threept.cpp:1580-1640 with M = 0).
std::vector<Term> gpp_terms(int l1, int l2, int L) {
  std::vector<Term> terms;
  for (int m1 = -l1; m1 <= l1; m1++)
    for (int m2 = -l2; m2 <= l2; m2++) {
      const int M = 0;
      bool redundant = false;
      for (auto& t : terms)
        if (t.m1 + m1 == 0 && t.m2 + m2 == 0 && t.M + M == 0) redundant = true;
      if (redundant) continue;
      const double c = coupling_3pt(l1, l2, L, m1, m2, M);
      if (std::fabs(c) < 1.e-9) continue;
      const bool zeros = (m1 == 0 && m2 == 0);
      terms.push_back(Term{m1, m2, M, c, zeros ? 0. : ((L % 2) ? -1. : 1.),
                           zeros ? 0. : 1.});
    }
  return terms;
}

struct PairGroup {
  int ra0, ca0, nas, nbs, nt;
  DevArray<int2> tiles;
  DevArray<int2> ent;   // (dv, index into the group's [nt * 16] sums)
  int nent = 0;
};

struct TermPlan {
  Term t;
  int slab_a, slab_b;
  // Pair kernel: 0 both sides complex, 1 second side packed real, 2 both
  // sides packed real; swap: the first side is slab_b (the packed side of a
  // mixed pair goes second).
  int mode = 0;
  bool swap = false;
  std::vector<std::unique_ptr<PairGroup>> groups;
};

// Shells of (l, m) are real for even l and m = 0 (real Y_l0, Hermitian
// dn00): such slabs hold two bins per complex mesh.
bool real_slab(int l, int m) { return m == 0 && l % 2 == 0; }

struct MultipolePlan {
  int l1, l2, L, shape, idx_bin, dim;
  std::vector<int> ra, cb;
  std::vector<TermPlan> terms;
  DevArray<double> n1, n2;            // per entry: mode counts
  // Shot-noise tables: k_eff of the distinct bins of each side, the entries
  // as (side-a index, side-b index), radii per table block.
  DevArray<double> ka, kb;
  DevArray<int2> pidx;
  int nka = 0, nkb = 0, nch = 0, nchunks = 0;
  size_t tab_smem = 0;
  DevArray<int> bin_a, bin_b;       // per entry: row / column bin
  DevArray<double2> raw, shot;      // [B][dim]
  double2 phase;
};

// Launch parameters chosen by the autotuner (blocks per mock).
struct Tuning {
  int assign_blocks = 64;
  int pair_blocks = 32;
  int stat_blocks = 32;
  int rb_cap = 8;
};

}  // namespace

template <typename T>
class BatchPlan : public BatchPlanBase {
 public:
  BatchPlan(const trvg_params& p, const std::vector<Multipole>& mps,
            const trvg_binning& kb, int batch, long long capacity, int autotune,
            int use_graph)
      : ctx_(p), B_(batch), cap_(capacity), autotune_(autotune != 0),
        use_graph_(use_graph != 0) {
    const Geom& g = ctx_.g;
    if (!(g.dr[0] == g.dr[1] && g.dr[1] == g.dr[2]))
      throw Error(TRVG_ERR_PARAM, "Batched plans need cubic mesh cells.");
    nb_ = kb.num_bins;
    edges_.assign(kb.bin_edges, kb.bin_edges + nb_ + 1);
    centres_.assign(kb.bin_centres, kb.bin_centres + nb_);
    kb_.num_bins = nb_;
    kb_.bin_edges = edges_.data();
    kb_.bin_centres = centres_.data();
    ss_ = shell_stats(ctx_, kb_);
    nm_ = g.nmesh;
    nq_ = (long long)(g.n[0] / 2) * (g.n[0] / 2) + (long long)(g.n[1] / 2) * (g.n[1] / 2) +
          (long long)(g.n[2] / 2) * (g.n[2] / 2) + 1;

    // Host staging (pinned) and device inputs.
    TRVG_HIP(hipHostMalloc((void**)&h_pos_, 3 * cap_ * sizeof(double)));
    TRVG_HIP(hipHostMalloc((void**)&h_meta_, (3 * B_ + 1) * sizeof(double)));
    d_x_.reset(cap_); d_y_.reset(cap_); d_z_.reset(cap_);
    d_offs_.reset(B_ + 1);
    // Tiled assignment: keys (mock, tile), their counts and offsets, the
    // particles in key order.
    for (int a = 0; a < 3; a++) tg_.nt[a] = (g.n[a] + TILE - 1) / TILE;
    tg_.ntiles = (long long)tg_.nt[0] * tg_.nt[1] * tg_.nt[2];
    nkeys_ = (long long)B_ * (tg_.ntiles + 1);
    nseg_ = int((nkeys_ + 1023) / 1024);
    d_kcount_.reset(nkeys_);
    d_koff_.reset(nkeys_ + 1);
    d_ksum_.reset(nseg_);
    d_ksoff_.reset(nseg_ + 1);
    d_key_.reset(cap_); d_rank_.reset(cap_);
    d_sx_.reset(cap_); d_sy_.reset(cap_); d_sz_.reset(cap_);
    d_nmock_.reset(B_);
    d_norm_.reset(B_);
    TRVG_HIP(hipHostMalloc((void**)&h_offs_, (B_ + 1) * sizeof(long long)));

    // Common fields.
    D_.reset((size_t)B_ * nm_);
    tfield_.reset((size_t)B_ * nm_);
    G_.reset((size_t)B_ * nm_);
    Q_.reset((size_t)B_ * 2 * nq_);
    tmap_ = twopt_map(ctx_, edges_.data(), nb_);

    // Work lists: unpruned passes over B fields, and the shell slabs
    // (complex and packed real).
    full_lists(B_, wfull_);
    for (int pk = 0; pk < 2; pk++) {
      SlabWork& sw = swork_[pk];
      std::vector<ShellBand> bands;
      std::vector<double> kc;
      std::vector<int2> wz, wy, wx;
      shell_slab_work<T>(g, kb_, nb_, B_, pk == 1, bands, kc, wz, wy, wx);
      sw.nz = int(wz.size()); sw.ny = int(wy.size()); sw.nx = int(wx.size());
      upload(sw.bands, bands);
      upload(sw.kcut2, kc);
      std::vector<int2> all(wz);
      all.insert(all.end(), wy.begin(), wy.end());
      all.insert(all.end(), wx.begin(), wx.end());
      upload(sw.work, all);
    }

    // Multipoles, terms, slabs, pair groups.
    for (const Multipole& mp : mps) {
      auto P = std::make_unique<MultipolePlan>();
      P->l1 = mp.ell1; P->l2 = mp.ell2; P->L = mp.ELL;
      P->shape = mp.shape; P->idx_bin = mp.idx_bin;
      pair_bins(mp.shape, nb_, mp.idx_bin, P->ra, P->cb);
      P->dim = int(P->ra.size());
      std::vector<double> n1(P->dim), n2(P->dim);
      for (int dv = 0; dv < P->dim; dv++) {
        n1[dv] = ss_.nmodes[P->ra[dv]];
        n2[dv] = ss_.nmodes[P->cb[dv]];
      }
      upload(P->n1, n1); upload(P->n2, n2);
      {
        std::vector<int> ua(P->ra), ub(P->cb);
        for (auto* u : {&ua, &ub}) {
          std::sort(u->begin(), u->end());
          u->erase(std::unique(u->begin(), u->end()), u->end());
        }
        std::vector<double> ka, kbv;
        for (int b : ua) ka.push_back(ss_.keff[b]);
        for (int b : ub) kbv.push_back(ss_.keff[b]);
        std::vector<int2> pidx(P->dim);
        for (int dv = 0; dv < P->dim; dv++)
          pidx[dv] = make_int2(
              int(std::lower_bound(ua.begin(), ua.end(), P->ra[dv]) - ua.begin()),
              int(std::lower_bound(ub.begin(), ub.end(), P->cb[dv]) - ub.begin()));
        upload(P->ka, ka); upload(P->kb, kbv); upload(P->pidx, pidx);
        P->nka = int(ka.size()); P->nkb = int(kbv.size());
        P->nch = shot_tab_chunk(P->nka, P->nkb);
        P->tab_smem = shot_tab_smem(P->nka, P->nkb, P->nch);
        if (P->tab_smem > 160 * 1024)
          throw Error(TRVG_ERR_PARAM, "Too many wavenumber bins for the shot-noise tables.");
        P->nchunks = int((nq_ + P->nch - 1) / P->nch);
      }
      upload(P->bin_a, P->ra); upload(P->bin_b, P->cb);
      P->raw.reset((size_t)B_ * P->dim);
      P->shot.reset((size_t)B_ * P->dim);
      const int ph = (mp.ell1 + mp.ell2) % 4;   // i^(l1 + l2), threept.cpp:271-272
      P->phase = ph == 0 ? dmk(1., 0.) : ph == 1 ? dmk(0., 1.) : ph == 2 ? dmk(-1., 0.)
                                                              : dmk(0., -1.);
      for (const Term& t : gpp_terms(mp.ell1, mp.ell2, mp.ELL)) {
        TermPlan tp;
        tp.t = t;
        tp.slab_a = slab_index(mp.ell1, t.m1);
        tp.slab_b = slab_index(mp.ell2, t.m2);
        // G is real for every periodic-box term (M = 0).
        const bool pa = slab_packed_[tp.slab_a], pb = slab_packed_[tp.slab_b];
        tp.mode = (pa && pb) ? 2 : (pa || pb) ? 1 : 0;
        tp.swap = pa && !pb;
        P->terms.push_back(std::move(tp));
      }
      (void)cached_spline(ctx_.dev, mp.ell1, ctx_.stream);
      (void)cached_spline(ctx_.dev, mp.ell2, ctx_.stream);
      mps_.push_back(std::move(P));
    }
    int maxdim = 0;
    size_t maxtab = 0;
    for (auto& P : mps_) {
      maxdim = std::max(maxdim, P->dim);
      maxtab = std::max(maxtab, (size_t)P->nchunks * P->dim);
    }
    sij_.reset((size_t)B_ * maxdim);
    tab_part_.reset((size_t)B_ * maxtab);
    counts_.reset((size_t)B_ * nb_);
    tsum_.reset((size_t)B_ * nb_ * 4);
    build_pair_groups();
    TRVG_HIP(hipStreamSynchronize(ctx_.stream));
  }

  ~BatchPlan() override {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (h_pos_) (void)hipHostFree(h_pos_);
    if (h_meta_) (void)hipHostFree(h_meta_);
    if (h_offs_) (void)hipHostFree(h_offs_);
  }

  void info(int* tuned) const override {
    tuned[0] = tune_.assign_blocks;
    tuned[1] = tune_.pair_blocks;
    tuned[2] = tune_.stat_blocks;
    tuned[3] = tune_.rb_cap;
    tuned[4] = exec_ ? 1 : 0;
  }

  void run(int nmocks, const long long* counts, const double* x, const double* y,
           const double* z, trvg_bispec_result* outs) override {
    run_impl(nmocks, counts, x, y, z, false, outs);
  }

  void run_device(int nmocks, const long long* counts, const double* x,
                  const double* y, const double* z, trvg_bispec_result* outs) override {
    run_impl(nmocks, counts, x, y, z, true, outs);
  }

 private:
  void run_impl(int nmocks, const long long* counts, const double* x,
                const double* y, const double* z, bool on_device,
                trvg_bispec_result* outs) {
    if (nmocks < 1 || nmocks > B_)
      throw Error(TRVG_ERR_PARAM, "Batch holds 1 to " + std::to_string(B_) + " catalogues.");
    long long total = 0;
    for (int b = 0; b < nmocks; b++) {
      if (counts[b] < 0) throw Error(TRVG_ERR_PARAM, "Negative particle count.");
      total += counts[b];
    }
    if (total > cap_)
      throw Error(TRVG_ERR_PARAM, "Batch of " + std::to_string(total) +
                                      " particles exceeds the plan capacity " +
                                      std::to_string(cap_) + ".");
    for (int b = 0; b < nmocks; b++)
      if (counts[b] == 0) throw Error(TRVG_ERR_DATA, "Empty catalogue in batch.");
    // Stage positions and per-mock metadata (unused slots: no particles).
    if (!on_device) {
      std::memcpy(h_pos_, x, total * 8);
      std::memcpy(h_pos_ + cap_, y, total * 8);
      std::memcpy(h_pos_ + 2 * cap_, z, total * 8);
    }
    h_offs_[0] = 0;
    const double V = ctx_.g.vol;
    for (int b = 0; b < B_; b++) {
      const long long n = b < nmocks ? counts[b] : 0;
      h_offs_[b + 1] = h_offs_[b] + n;
      const double nz = double(n) / V;
      h_meta_[b] = double(n);
      // Particle normalisation with nz = N / V (threept.cpp:96-136).
      h_meta_[B_ + b] = n ? 1. / (double(n) * nz * nz) : 0.;
    }
    hipStream_t s = ctx_.stream;
    if (on_device) {
      // Inputs already in device memory (written by other streams before
      // this call returns to them: synchronise the device first).
      TRVG_HIP(hipDeviceSynchronize());
      TRVG_HIP(hipMemcpyAsync(d_x_.get(), x, total * 8, hipMemcpyDeviceToDevice, s));
      TRVG_HIP(hipMemcpyAsync(d_y_.get(), y, total * 8, hipMemcpyDeviceToDevice, s));
      TRVG_HIP(hipMemcpyAsync(d_z_.get(), z, total * 8, hipMemcpyDeviceToDevice, s));
    } else {
      TRVG_HIP(hipMemcpyAsync(d_x_.get(), h_pos_, total * 8, hipMemcpyHostToDevice, s));
      TRVG_HIP(hipMemcpyAsync(d_y_.get(), h_pos_ + cap_, total * 8, hipMemcpyHostToDevice, s));
      TRVG_HIP(hipMemcpyAsync(d_z_.get(), h_pos_ + 2 * cap_, total * 8, hipMemcpyHostToDevice, s));
    }
    TRVG_HIP(hipMemcpyAsync(d_offs_.get(), h_offs_, (B_ + 1) * 8, hipMemcpyHostToDevice, s));
    TRVG_HIP(hipMemcpyAsync(d_nmock_.get(), h_meta_, B_ * 8, hipMemcpyHostToDevice, s));
    TRVG_HIP(hipMemcpyAsync(d_norm_.get(), h_meta_ + B_, B_ * 8, hipMemcpyHostToDevice, s));

    if (!warmed_) {
      enqueue();                       // first use: twiddles, code objects
      if (autotune_) tune();
      warmed_ = true;
      if (use_graph_) capture();
    }
    if (exec_) TRVG_HIP(hipGraphLaunch(exec_, s));
    else enqueue();

    // Results.
    for (size_t q = 0; q < mps_.size(); q++) {
      MultipolePlan& P = *mps_[q];
      std::vector<double2> raw((size_t)nmocks * P.dim), shot((size_t)nmocks * P.dim);
      TRVG_HIP(hipMemcpyAsync(raw.data(), P.raw.get(), raw.size() * 16, hipMemcpyDeviceToHost, s));
      TRVG_HIP(hipMemcpyAsync(shot.data(), P.shot.get(), shot.size() * 16, hipMemcpyDeviceToHost, s));
      TRVG_HIP(hipStreamSynchronize(s));
      for (int b = 0; b < nmocks; b++) {
        trvg_bispec_result* o = &outs[(size_t)b * mps_.size() + q];
        if (o->dim < P.dim)
          throw Error(TRVG_ERR_PARAM, "Result capacity below the data-vector dimension.");
        for (int dv = 0; dv < P.dim; dv++) {
          const int a = P.ra[dv], c = P.cb[dv];
          o->k1_bin[dv] = centres_[a]; o->k2_bin[dv] = centres_[c];
          o->k1_eff[dv] = ss_.keff[a]; o->k2_eff[dv] = ss_.keff[c];
          o->nmodes_1[dv] = ss_.nmodes[a]; o->nmodes_2[dv] = ss_.nmodes[c];
          const double2 r = raw[(size_t)b * P.dim + dv], sh = shot[(size_t)b * P.dim + dv];
          o->bk_raw[2 * dv] = r.x; o->bk_raw[2 * dv + 1] = r.y;
          o->bk_shot[2 * dv] = sh.x; o->bk_shot[2 * dv + 1] = sh.y;
        }
        o->dim = P.dim;
      }
    }
  }

  template <typename E>
  void upload(DevArray<E>& d, const std::vector<E>& h) {
    d.reset(h.size());
    if (!h.empty())
      TRVG_HIP(hipMemcpyAsync(d.get(), h.data(), h.size() * sizeof(E),
                              hipMemcpyHostToDevice, ctx_.stream));
  }

  template <typename U>
  void full_lists(int nfields, std::vector<int2>& w) {
    constexpr int E = FFT_THREADS * FftEpt<U>::v;
    const long long nblk = nm_ / E;
    w.clear();
    for (int f = 0; f < nfields; f++)
      for (long long k = 0; k < nblk; k++) w.push_back(make_int2(f, int(k)));
  }
  void full_lists(int nfields, DevArray<int2>& d64) {
    std::vector<int2> w;
    full_lists<double>(nfields, w);
    upload(d64, w);
    nfull64_ = int(w.size());
    std::vector<int2> w32;
    full_lists<T>(nfields, w32);
    upload(wfullT_, w32);
    nfullT_ = int(w32.size());
  }

  int slab_index(int l, int m) {
    for (size_t i = 0; i < slabs_.size(); i++)
      if (slabs_[i].first == l && slabs_[i].second == m) return int(i);
    slabs_.push_back({l, m});
    slab_packed_.push_back(real_slab(l, m));
    slab_data_.emplace_back();
    slab_data_.back().reset((size_t)B_ * slab_meshes(int(slabs_.size() - 1)) * nm_);
    return int(slabs_.size() - 1);
  }

  // Meshes per mock of slab i.
  int slab_meshes(int i) const { return slab_packed_[i] ? (nb_ + 1) / 2 : nb_;
  }

  void build_pair_groups() {
    for (auto& P : mps_) {
      for (TermPlan& tp : P->terms) {
        tp.groups.clear();
        const int na = nb_, nbf = nb_;
        const int rbn = (na + 3) / 4, cbn = (nbf + 3) / 4;
        // Tile rows index the first side's bins (slab_b's when swapped).
        const std::vector<int>& rows = tp.swap ? P->cb : P->ra;
        const std::vector<int>& cols = tp.swap ? P->ra : P->cb;
        std::vector<char> active((size_t)rbn * cbn, 0);
        for (int dv = 0; dv < P->dim; dv++)
          active[(size_t)(rows[dv] / 4) * cbn + cols[dv] / 4] = 1;
        const int CB = std::min(cbn, 8);
        const int RB = std::max(1, std::min(std::min((PT_MAXROWS - 1 - 4 * CB) / 4,
                                                     PT_MAXT / CB),
                                            tune_.rb_cap));
        for (int c0b = 0; c0b < cbn; c0b += CB) {
          const int c1b = std::min(cbn, c0b + CB);
          for (int r0b = 0; r0b < rbn; r0b += RB) {
            const int r1b = std::min(rbn, r0b + RB);
            std::vector<int2> tl;
            for (int rb = r0b; rb < r1b; rb++)
              for (int cb = c0b; cb < c1b; cb++)
                if (active[(size_t)rb * cbn + cb])
                  tl.push_back(make_int2(4 * (rb - r0b), 4 * (cb - c0b)));
            if (tl.empty()) continue;
            auto gr = std::make_unique<PairGroup>();
            gr->ra0 = 4 * r0b;
            gr->ca0 = 4 * c0b;
            gr->nas = std::min(na, 4 * r1b) - gr->ra0;
            gr->nbs = std::min(nbf, 4 * c1b) - gr->ca0;
            gr->nt = int(tl.size());
            std::vector<int2> ent;
            for (int dv = 0; dv < P->dim; dv++) {
              const int a = rows[dv] - gr->ra0, c = cols[dv] - gr->ca0;
              if (a < 0 || c < 0 || a >= 4 * (r1b - r0b) || c >= 4 * (c1b - c0b)) continue;
              for (int t = 0; t < gr->nt; t++)
                if (tl[t].x == 4 * (a / 4) && tl[t].y == 4 * (c / 4))
                  ent.push_back(make_int2(dv, t * 16 + (a % 4) * 4 + c % 4));
            }
            upload(gr->tiles, tl);
            upload(gr->ent, ent);
            gr->nent = int(ent.size());
            tp.groups.push_back(std::move(gr));
          }
        }
      }
    }
    size_t need = 0;
    for (auto& P : mps_)
      for (TermPlan& tp : P->terms)
        for (auto& gr : tp.groups)
          need = std::max(need, (size_t)gr->nt * 16);
    pair_width_ = need;
    pair_part_.reset((size_t)B_ * kMaxPairBlocks * need);
    pair_sum_.reset((size_t)B_ * need);
    tpart_.reset((size_t)B_ * kMaxStatBlocks * nb_ * 4);
  }

  // ---- the batch sequence ------------------------------------------------

  void enqueue() {
    const Geom& g = ctx_.g;
    hipStream_t s = ctx_.stream;
    const long long total = (long long)B_ * nm_;
    const unsigned zg = grid_for(ctx_, 2 * total);
    // dn00(k) of every mock (FP64); N00(k) equals it except for N_b at
    // k = 0, which its two consumers add themselves.
    hipLaunchKernelGGL(k_fill_zero_d, dim3(zg), dim3(256), 0, s,
                       reinterpret_cast<double*>(D_.get()), 2 * total);
    TRVG_LAUNCH_CHECK();
    launch_assign();
    LoadScaleShiftMocks ls{D_.get(), nm_, d_nmock_.get(), g.vol, g.vol_cell};
    fft3d_batch<double>(ctx_, D_.get(), ls, -1, false, swork_[0].kcut2.get(),
                        wfull_.get(), nfull64_, wfull_.get(), nfull64_,
                        wfull_.get(), nfull64_);
    // Shell slabs (all bins of every mock per pass).
    for (size_t i = 0; i < slabs_.size(); i++) {
      const SlabWork& sw = swork_[slab_packed_[i] ? 1 : 0];
      const int2* wz = sw.work.get();
      if (slab_packed_[i]) {
        LoadShellPairMocks<T> lsh{D_.get(), nm_, ctx_.tab, g, sw.bands.get(),
                                  (nb_ + 1) / 2, slabs_[i].first, slabs_[i].second};
        fft3d_batch<T>(ctx_, slab_data_[i].get(), lsh, +1, true, sw.kcut2.get(),
                       wz, sw.nz, wz + sw.nz, sw.ny, wz + sw.nz + sw.ny, sw.nx);
      } else {
        LoadShellMocks<T> lsh{D_.get(), nm_, ctx_.tab, g, sw.bands.get(), nb_,
                              slabs_[i].first, slabs_[i].second};
        fft3d_batch<T>(ctx_, slab_data_[i].get(), lsh, +1, true, sw.kcut2.get(),
                       wz, sw.nz, wz + sw.nz, sw.ny, wz + sw.nz + sw.ny, sw.nx);
      }
    }
    // G(x) = IFFT[dn00 / (W V)] (every periodic-box term uses dn00).
    LoadDeconvMocks<T> ld{D_.get(), nm_, ctx_.tab, 1. / g.vol};
    fft3d_batch<T>(ctx_, G_.get(), ld, +1, false, swork_[0].kcut2.get(), wfullT_.get(),
                   nfullT_, wfullT_.get(), nfullT_, wfullT_.get(), nfullT_);
    // Shot-noise field t(x) = IFFT[(dn00 N00^* - N) / C1] / V: the same for
    // every term and multipole of a periodic box (threept.cpp:950-964).
    LoadTwoptMocks lt{D_.get(), D_.get(), nm_, ctx_.tab, d_nmock_.get(), 1. / g.vol};
    fft3d_batch<double>(ctx_, tfield_.get(), lt, +1, false, swork_[0].kcut2.get(),
                        wfull_.get(), nfull64_, wfull_.get(), nfull64_,
                        wfull_.get(), nfull64_);
    for (auto& P : mps_) run_multipole(*P);
  }

  void launch_assign() {
    const Geom& g = ctx_.g;
    hipStream_t s = ctx_.stream;
    const dim3 pgrid(tune_.assign_blocks, B_);
    hipLaunchKernelGGL(k_fill_zero_u32, dim3(grid_for(ctx_, nkeys_)), dim3(256), 0, s,
                       d_kcount_.get(), nkeys_);
    hipLaunchKernelGGL(k_tile_key_batch, pgrid, dim3(256), 0, s, d_x_.get(), d_y_.get(),
                       d_z_.get(), d_offs_.get(), g, tg_, d_kcount_.get(), d_key_.get(),
                       d_rank_.get());
    hipLaunchKernelGGL(k_scan_seg, dim3(nseg_), dim3(1024), 0, s, d_kcount_.get(), nkeys_,
                       d_koff_.get(), d_ksum_.get());
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, d_ksum_.get(),
                       (long long)nseg_, d_ksoff_.get());
    hipLaunchKernelGGL(k_scan_add, dim3(nseg_), dim3(1024), 0, s, d_koff_.get(), nkeys_,
                       d_ksoff_.get(), nseg_);
    hipLaunchKernelGGL(k_tile_scatter_batch, pgrid, dim3(256), 0, s, d_x_.get(), d_y_.get(),
                       d_z_.get(), d_offs_.get(), d_key_.get(), d_rank_.get(), d_koff_.get(),
                       d_sx_.get(), d_sy_.get(), d_sz_.get());
    TRVG_LAUNCH_CHECK();
    const dim3 kgrid(static_cast<unsigned>(nkeys_), 1, 1);
    switch (ctx_.p.assignment_order) {
      case 1: hipLaunchKernelGGL((k_assign_tiled_batch<1>), kgrid, dim3(256), 0, s, D_.get(),
                                 nm_, d_sx_.get(), d_sy_.get(), d_sz_.get(), d_koff_.get(), g,
                                 tg_); break;
      case 2: hipLaunchKernelGGL((k_assign_tiled_batch<2>), kgrid, dim3(256), 0, s, D_.get(),
                                 nm_, d_sx_.get(), d_sy_.get(), d_sz_.get(), d_koff_.get(), g,
                                 tg_); break;
      case 3: hipLaunchKernelGGL((k_assign_tiled_batch<3>), kgrid, dim3(256), 0, s, D_.get(),
                                 nm_, d_sx_.get(), d_sy_.get(), d_sz_.get(), d_koff_.get(), g,
                                 tg_); break;
      default: hipLaunchKernelGGL((k_assign_tiled_batch<4>), kgrid, dim3(256), 0, s, D_.get(),
                                  nm_, d_sx_.get(), d_sy_.get(), d_sz_.get(), d_koff_.get(), g,
                                  tg_); break;
    }
    TRVG_LAUNCH_CHECK();
  }


  template <int MODE>
  void launch_pairs_mode(const c2<T>* A, int nas, long long sa, const c2<T>* Bs, int nbs,
                         long long sb, PairGroup& gr, size_t smem) {
    if (smem > 64 * 1024)
      TRVG_HIP(hipFuncSetAttribute((const void*)k_pair_lds_batch<T, MODE>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(smem)));
    hipLaunchKernelGGL((k_pair_lds_batch<T, MODE>), dim3(tune_.pair_blocks, B_), dim3(256),
                       smem, ctx_.stream, A, nas, sa, Bs, nbs, sb, G_.get(), nm_, nm_,
                       gr.tiles.get(), gr.nt, pair_part_.get(),
                       (long long)tune_.pair_blocks * gr.nt * 16);
    TRVG_LAUNCH_CHECK();
  }

  void launch_pairs(TermPlan& tp, PairGroup& gr) {
    const int ia = tp.swap ? tp.slab_b : tp.slab_a;
    const int ib = tp.swap ? tp.slab_a : tp.slab_b;
    // Packed sides start at mesh ra0 / 2 (group starts are multiples of 4).
    const int a0 = slab_packed_[ia] ? gr.ra0 / 2 : gr.ra0;
    const int b0 = slab_packed_[ib] ? gr.ca0 / 2 : gr.ca0;
    const c2<T>* A = slab_data_[ia].get() + (long long)a0 * nm_;
    const c2<T>* Bs = slab_data_[ib].get() + (long long)b0 * nm_;
    const long long sa = (long long)slab_meshes(ia) * nm_;
    const long long sb = (long long)slab_meshes(ib) * nm_;
    const int mode = (tp.mode == 2 && A == Bs && gr.nas == gr.nbs && sa == sb) ? 3 : tp.mode;
    const size_t smem = pair_lds_smem<T>(mode, gr.nas, gr.nbs);
    switch (mode) {
      case 3: launch_pairs_mode<3>(A, gr.nas, sa, Bs, gr.nbs, sb, gr, smem); break;
      case 2: launch_pairs_mode<2>(A, gr.nas, sa, Bs, gr.nbs, sb, gr, smem); break;
      case 1: launch_pairs_mode<1>(A, gr.nas, sa, Bs, gr.nbs, sb, gr, smem); break;
      default: launch_pairs_mode<0>(A, gr.nas, sa, Bs, gr.nbs, sb, gr, smem); break;
    }
  }


  void run_multipole(MultipolePlan& P) {
    const Geom& g = ctx_.g;
    hipStream_t s = ctx_.stream;
    const long long nd = 2LL * B_ * P.dim;
    hipLaunchKernelGGL(k_fill_zero_d, dim3(grid_for(ctx_, nd)), dim3(256), 0, s,
                       reinterpret_cast<double*>(P.raw.get()), nd);
    hipLaunchKernelGGL(k_fill_zero_d, dim3(grid_for(ctx_, nd)), dim3(256), 0, s,
                       reinterpret_cast<double*>(P.shot.get()), nd);
    TRVG_LAUNCH_CHECK();
    const unsigned eg = unsigned(std::max(1, std::min(64, (P.dim + 255) / 256)));
    const Spline& sja = cached_spline(ctx_.dev, P.l1, s);
    const Spline& sjb = cached_spline(ctx_.dev, P.l2, s);
    for (TermPlan& tp : P.terms) {
      const Term& t = tp.t;
      // Raw bispectrum: pair sums of the shell slabs with G.
      for (auto& gr : tp.groups) {
        launch_pairs(tp, *gr);
        const int w = gr->nt * 16;
        hipLaunchKernelGGL((k_reduce_cols_batch<double2>),
                           dim3((w + RC_COLS - 1) / RC_COLS, B_), dim3(RC_COLS * RC_ROWS),
                           0, s, pair_part_.get(), tune_.pair_blocks, w, pair_sum_.get());
        hipLaunchKernelGGL(k_epi_pairs, dim3(eg, B_), dim3(256), 0, s, P.raw.get(),
                           P.dim, pair_sum_.get(), w, gr->ent.get(), gr->nent,
                           P.n1.get(), P.n2.get(), t.coupling, t.f_mirror, g.vol_cell);
        TRVG_LAUNCH_CHECK();
      }
      // Shot noise (threept.cpp:1760-1960): Sbar = N.
      if (P.l1 == 0 && P.l2 == 0) {
        hipLaunchKernelGGL(k_epi_const, dim3(eg, B_), dim3(256), 0, s, P.shot.get(),
                           P.dim, d_nmock_.get(), t.coupling);
        TRVG_LAUNCH_CHECK();
      }
      if (P.l2 == 0) twopt(P, t, P.l1, t.m1, P.bin_a.get(), true);
      // Both sides monopoles: reuse the first side's statistics.
      if (P.l1 == 0)
        twopt(P, t, P.l2, t.m2, P.bin_b.get(), !(P.l2 == 0 && t.m1 == t.m2));
      const long long nQ = (long long)B_ * 2 * nq_;
      hipLaunchKernelGGL(k_fill_zero_d, dim3(grid_for(ctx_, nQ)), dim3(256), 0, s,
                         Q_.get(), nQ);
      const long long nrep = (long long)(g.n[0] / 2 + 1) * (g.n[1] / 2 + 1) * (g.n[2] / 2 + 1);
      const unsigned rg = unsigned(std::min<long long>(tune_.stat_blocks * 4,
                                                       (nrep + 255) / 256));
      hipLaunchKernelGGL(k_radial_hist_oct_batch, dim3(rg, B_), dim3(RED_THREADS), 0, s,
                         tfield_.get(), nm_, g, P.l1, t.m1, P.l2, t.m2, Q_.get(), nq_);
      hipLaunchKernelGGL(k_shot_pairs_tab_batch, dim3(P.nchunks, B_), dim3(256),
                         P.tab_smem, s, Q_.get(), nq_, g.dr[0], sja.dev(), sjb.dev(),
                         P.ka.get(), P.nka, P.kb.get(), P.nkb, P.pidx.get(), P.dim,
                         P.nch, g.vol_cell, tab_part_.get());
      hipLaunchKernelGGL((k_reduce_cols_batch<double2>),
                         dim3((P.dim + RC_COLS - 1) / RC_COLS, B_), dim3(RC_COLS * RC_ROWS),
                         0, s, tab_part_.get(), P.nchunks, P.dim, sij_.get());
      hipLaunchKernelGGL(k_epi_sij, dim3(eg, B_), dim3(256), 0, s, P.shot.get(), P.dim,
                         sij_.get(), t.coupling, t.f_mirror_w3j, P.phase);
      TRVG_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_epi_norm, dim3(eg, B_), dim3(256), 0, s, P.raw.get(),
                       P.shot.get(), P.dim, d_norm_.get());
    TRVG_LAUNCH_CHECK();
  }

  // Two-point shot-noise term of one side; compute = false reuses the
  // statistics left in counts_ / tsum_ by the previous call.
  void twopt(MultipolePlan& P, const Term& t, int ell, int m, const int* bin_of,
             bool compute) {
    hipStream_t s = ctx_.stream;
    if (compute) twopt_stats(ell, m);
    const unsigned eg = unsigned(std::max(1, std::min(64, (P.dim + 255) / 256)));
    hipLaunchKernelGGL(k_epi_twopt, dim3(eg, B_), dim3(256), 0, s, P.shot.get(), P.dim,
                       tsum_.get(), counts_.get(), nb_, bin_of, t.coupling, t.f_mirror);
    TRVG_LAUNCH_CHECK();
  }

  void twopt_stats(int ell, int m) {
    const Geom& g = ctx_.g;
    hipStream_t s = ctx_.stream;
    hipLaunchKernelGGL(k_fill_zero_u64, dim3(1), dim3(256), 0, s, counts_.get(),
                       (long long)B_ * nb_);
    hipLaunchKernelGGL(k_twopt_fourier_batch, dim3(tune_.stat_blocks, B_),
                       dim3(RED_THREADS), twopt_smem(nb_, 4, g), s, D_.get(), D_.get(),
                       nm_, g, ctx_.tab, d_nmock_.get(), ell, m, tmap_->view(), nb_,
                       counts_.get(), tpart_.get(), 1);
    const int w = nb_ * 4;
    hipLaunchKernelGGL((k_reduce_cols_batch<double>), dim3((w + RC_COLS - 1) / RC_COLS, B_),
                       dim3(RC_COLS * RC_ROWS), 0, s, tpart_.get(), tune_.stat_blocks, w,
                       tsum_.get());
    TRVG_LAUNCH_CHECK();
  }

  // ---- autotuner ---------------------------------------------------------

  template <class F>
  float time_it(F&& f) {
    hipEvent_t a, b;
    TRVG_HIP(hipEventCreate(&a));
    TRVG_HIP(hipEventCreate(&b));
    f();   // warm
    TRVG_HIP(hipEventRecord(a, ctx_.stream));
    f();
    TRVG_HIP(hipEventRecord(b, ctx_.stream));
    TRVG_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    TRVG_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
  }

  // Picks blocks per mock for the assignment, the pair reduction (and its
  // row-block cap) and the binned statistics by timing each candidate on
  // the current batch; every candidate computes the same sums.
  void tune() {
    auto pick = [&](int* slot, std::initializer_list<int> cands, auto&& body) {
      float best = 1e30f;
      int arg = *slot;
      for (int c : cands) {
        *slot = c;
        const float ms = time_it(body);
        if (ms < best) { best = ms; arg = c; }
      }
      *slot = arg;
    };
    const long long total = (long long)B_ * nm_;
    pick(&tune_.assign_blocks, {16, 32, 64, 128, 256}, [&] {
      hipLaunchKernelGGL(k_fill_zero_d, dim3(grid_for(ctx_, 2 * total)), dim3(256), 0,
                         ctx_.stream, reinterpret_cast<double*>(D_.get()), 2 * total);
      launch_assign();
    });
    MultipolePlan& P0 = *mps_[0];
    TermPlan& tp0 = P0.terms[0];
    pick(&tune_.pair_blocks, {8, 16, 32, 64, kMaxPairBlocks}, [&] {
      for (auto& gr : tp0.groups) launch_pairs(tp0, *gr);
    });
    float best = 1e30f;
    int arg = tune_.rb_cap;
    for (int cap : {4, 8}) {
      tune_.rb_cap = cap;
      build_pair_groups();
      const float ms = time_it([&] {
        for (auto& gr : mps_[0]->terms[0].groups) launch_pairs(mps_[0]->terms[0], *gr);
      });
      if (ms < best) { best = ms; arg = cap; }
    }
    tune_.rb_cap = arg;
    build_pair_groups();
    pick(&tune_.stat_blocks, {8, 16, 32, kMaxStatBlocks}, [&] {
      hipLaunchKernelGGL(k_fill_zero_u64, dim3(1), dim3(256), 0, ctx_.stream,
                         counts_.get(), (long long)B_ * nb_);
      hipLaunchKernelGGL(k_twopt_fourier_batch, dim3(tune_.stat_blocks, B_),
                         dim3(RED_THREADS), twopt_smem(nb_, 4, ctx_.g), ctx_.stream,
                         D_.get(), D_.get(), nm_, ctx_.g, ctx_.tab, d_nmock_.get(), 0, 0,
                         tmap_->view(), nb_, counts_.get(), tpart_.get(), 1);
    });
    TRVG_HIP(hipStreamSynchronize(ctx_.stream));
  }

  void capture() {
    TRVG_HIP(hipStreamSynchronize(ctx_.stream));
    TRVG_HIP(hipStreamBeginCapture(ctx_.stream, hipStreamCaptureModeThreadLocal));
    try {
      enqueue();
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(ctx_.stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    TRVG_HIP(hipStreamEndCapture(ctx_.stream, &graph_));
    TRVG_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
  }

  static constexpr int kMaxPairBlocks = 128;
  static constexpr int kMaxStatBlocks = 64;

  Context ctx_;
  int B_;
  long long cap_;
  bool autotune_, use_graph_, warmed_ = false;
  Tuning tune_;
  int nb_ = 0;
  long long nm_ = 0, nq_ = 0;
  std::vector<double> edges_, centres_;
  trvg_binning kb_{};
  ShellStats ss_;
  double* h_pos_ = nullptr;
  double* h_meta_ = nullptr;
  long long* h_offs_ = nullptr;
  DevArray<double> d_x_, d_y_, d_z_, d_nmock_, d_norm_;
  std::shared_ptr<const TwoptMapData> tmap_;
  DevArray<long long> d_offs_;
  TileGeom tg_{};
  long long nkeys_ = 0;
  int nseg_ = 0;
  DevArray<unsigned int> d_kcount_, d_koff_, d_ksum_, d_ksoff_, d_key_, d_rank_;
  DevArray<double> d_sx_, d_sy_, d_sz_;
  DevArray<double2> D_, tfield_;
  DevArray<c2<T>> G_;
  DevArray<double> Q_;
  // Shell slab work lists: [0] one bin per mesh, [1] packed real pairs.
  struct SlabWork {
    DevArray<ShellBand> bands;
    DevArray<double> kcut2;
    DevArray<int2> work;   // z, then y, then x blocks
    int nz = 0, ny = 0, nx = 0;
  } swork_[2];
  DevArray<int2> wfull_, wfullT_;
  int nfull64_ = 0, nfullT_ = 0;
  std::vector<std::pair<int, int>> slabs_;
  std::vector<bool> slab_packed_;
  std::vector<DevArray<c2<T>>> slab_data_;
  std::vector<std::unique_ptr<MultipolePlan>> mps_;
  size_t pair_width_ = 0;
  DevArray<double2> pair_part_, pair_sum_, sij_, tab_part_;
  DevArray<double> tpart_, tsum_;
  DevArray<unsigned long long> counts_;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
};

BatchPlanBase* make_batch_plan(const trvg_params& p, const std::vector<Multipole>& mps,
                               const trvg_binning& kb, int batch, long long capacity,
                               int autotune, int use_graph) {
  if (p.precision == TRVG_PREC_FP32)
    return new BatchPlan<float>(p, mps, kb, batch, capacity, autotune, use_graph);
  return new BatchPlan<double>(p, mps, kb, batch, capacity, autotune, use_graph);
}

}  // namespace trvg
