// The batched form: many independent realizations driven in lockstep on one
// stream by one host thread.  Each realization is an Engine<D> (its own cell
// store, voxel lists, word stream and helper threads) built on the driver's
// stream; the driver takes over the launches: every draw round and every
// refinement step of all the realizations that need it is ONE launch
// (k_batch.hpp), so the per-launch latency that bounds a single realization's
// draw loop is paid once for the whole group.  A realization's own sequence of
// operations -- its words, draws, rounds and refinements -- is exactly the one
// proceed() runs, so its placed set is the same.
// Not a header of its own: engine.hip includes it once, after Engine<D>.

namespace rsa_amd {

template <int D>
class BatchDriver {
 public:
  using E = Engine<D>;
  using Prep = typename E::Prep;

  BatchDriver(int n, const double* side, const double* radius, const uint64_t* seed, int nbshots, int active,
              uint64_t* count, double* volume, bool extract, rsa_gpu_batch_stats* bst)
      : n_(n), side_(side), radius_(radius), seed_(seed), nbshots_(nbshots),
        active_(std::max(1, std::min(active, kMaxBatch))), count_(count), volume_(volume), extract_(extract),
        bst_(bst) {}

  ~BatchDriver() {
    stop_threads();
    live_.clear();  // engines first: they sync the stream
    pre_.clear();
    fin_q_.clear();
    if (st_) (void)hipStreamSynchronize(st_);
    PinnedPool::get().release(arena_);
    PinnedPool::get().release(hctr_);
    dtab_.release();
    rblob_.release();
    if (st_) (void)hipStreamDestroy(st_);
  }

  std::vector<std::vector<double>> pos_, rad_;

  void run() {
    HIP_OK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    StreamScope scope(st_, nullptr);  // a buffer released by an engine waits for the shared stream
    arena_ = static_cast<char*>(PinnedPool::get().alloc(kArena, hipHostMallocCoherent | hipHostMallocMapped));
    HIP_OK(hipHostGetDevicePointer((void**)&arena_dev_, arena_, 0));
    hctr_ = static_cast<int*>(
        PinnedPool::get().alloc(sizeof(int) * C_NUM * kMaxBatch, hipHostMallocCoherent | hipHostMallocMapped));
    HIP_OK(hipHostGetDevicePointer((void**)&hctr_dev_, hctr_, 0));
    dtab_.reserve(sizeof(DrawDesc<D>) * kMaxBatch + 16);
    if (extract_) {
      pos_.resize(n_);
      rad_.resize(n_);
    }
    const auto t0 = std::chrono::steady_clock::now();
    // Realizations are set up (engine, generator, word stream, lattice) by the
    // admit thread and finished (end-of-run checks, extraction) by the finish
    // thread, both on the same stream: this thread only prepares, launches and
    // reads back draws and refinements, and the device keeps working meanwhile.
    admit_th_ = std::thread([this] { admit_loop(); });
    fin_th_ = std::thread([this] { finish_loop(); });
    while (taken_ < n_ || !live_.empty()) {
      timed(t_admit_, [&] { admit(); });
      // The group is driven as two halves (by slot parity): while the device runs
      // one half's round, the host reads the other half's results and prepares and
      // launches its next round.
      bool launched = false;
      for (int h = 0; h < 2; h++) {
        if (!inflight_[h].empty()) {
          std::vector<R*> again = collect(inflight_[h]);
          inflight_[h].clear();
          if (!again.empty()) {
            launch(again, false);
            inflight_[h] = again;
            launched = true;
            continue;
          }
        }
        std::vector<R*> first;
        timed(t_plan_, [&] {
          for (auto& b : live_)
            if (b->s == R::READY && b->half == h) plan(*b, first, false);
        });
        if (!first.empty()) {
          launch(first, true);
          inflight_[h] = first;
          launched = true;
        }
      }
      std::vector<R*> ref;
      int ready = 0;
      for (auto& b : live_) {
        if (b->s == R::REFINE) ref.push_back(b.get());
        if (b->s == R::READY) ready++;
      }
      const bool busy = !inflight_[0].empty() || !inflight_[1].empty();
      // refinements wait for company: run them once half the group waits, or when
      // nothing else can move (a realization's own order of steps is unchanged)
      if (!ref.empty() && (ref.size() >= refine_frac_ * live_.size() || (ready == 0 && !busy)))
        timed(t_refine_, [&] { refine(ref); });
      else if (!launched && !busy && ready > 0) {
        // nothing could go: wait for the first realization whose draw is being prepared
        for (auto& b : live_) {
          if (b->s != R::READY) continue;
          std::vector<R*> first;
          timed(t_plan_, [&] { plan(*b, first, true); });
          if (!first.empty()) {
            launch(first, true);
            inflight_[b->half] = first;
          }
          break;
        }
      }
      timed(t_finish_, [&] { finish(); });
    }
    timed(t_finish_, [&] { drain(); });
    if (bst_) {
      bst_->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      bst_->ticks = ticks_;
      bst_->draw_launch_groups = draw_groups_;
      bst_->refine_launch_groups = refine_groups_;
      bst_->max_active = max_active_;
      bst_->draws = draws_;
      bst_->refines = refines_;
      bst_->t_admit = t_admit_;
      bst_->t_plan = t_plan_;
      bst_->t_wait = t_wait_;
      bst_->t_refine = t_refine_;
      bst_->t_finish = t_finish_;
      bst_->t_prep = t_prep_;
      bst_->t_words = t_words_;
      bst_->t_refine_wait = t_refine_wait_;
    }
  }

 private:
  // one realization in flight
  struct R {
    int k = 0;
    std::unique_ptr<E> e;
    rsa_gpu_stats st{};
    enum State { READY, ROUND, REFINE, DONE } s = READY;
    Prep cur;
    uint64_t w = 0, old_nb = 0;
    int seq = 0, r = 0;
    bool lists_kept = false, refilter = false, first = false, cont = true, lists = false;
    int half = 0;  // which half of the group it is driven with
    int64_t local_new = 0;
    std::chrono::steady_clock::time_point t0;
  };

  static constexpr size_t kArena = size_t(16) << 20;

  int n_;
  const double *side_, *radius_;
  const uint64_t* seed_;
  int nbshots_, active_;
  uint64_t* count_;
  double* volume_;
  bool extract_;
  rsa_gpu_batch_stats* bst_;
  hipStream_t st_ = nullptr;
  char* arena_ = nullptr;
  char* arena_dev_ = nullptr;
  size_t arena_off_ = 0;
  int* hctr_ = nullptr;
  int* hctr_dev_ = nullptr;
  DevBuf<uint8_t> dtab_, rblob_;
  std::vector<DrawDesc<D>> tab_;  // host copy of a launch's descriptors
  std::vector<std::unique_ptr<R>> live_;
  std::vector<R*> inflight_[2];  // each half's draw round in flight
  // refinements wait until this fraction of the group waits for one
  // (RSA_GPU_BATCH_REFINE_FRAC, a tuning setting; default one half)
  double refine_frac_ = [] {
    const char* e = std::getenv("RSA_GPU_BATCH_REFINE_FRAC");
    return e ? std::atof(e) : 0.5;
  }();
  int next_ = 0;   // realizations set up so far (admit thread)
  int taken_ = 0;  // realizations made live (this thread)
  // admit and finish threads, their queues and errors (guarded by qm_)
  static constexpr int kPre = 4;  // realizations set up ahead
  std::thread admit_th_, fin_th_;
  std::mutex qm_;
  std::condition_variable qcv_;
  std::deque<std::unique_ptr<R>> pre_, fin_q_;
  bool qstop_ = false, fin_busy_ = false;
  std::string qerr_;
  int64_t ticks_ = 0, draw_groups_ = 0, refine_groups_ = 0, max_active_ = 0, draws_ = 0, refines_ = 0;
  double t_admit_ = 0, t_plan_ = 0, t_wait_ = 0, t_refine_ = 0, t_finish_ = 0, t_prep_ = 0, t_words_ = 0,
         t_refine_wait_ = 0;
  template <class F>
  static void timed(double& acc, F&& f) {
    const auto t0 = std::chrono::steady_clock::now();
    f();
    acc += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  // host bytes -> device table (through the arena and k_stage on the stream)
  void stage(const void* src, size_t bytes, void* dst) {
    const size_t n16 = (bytes + 15) / 16;
    if (arena_off_ + n16 * 16 > kArena) {  // wrap: every earlier staging copy has run once the stream is idle
      HIP_OK(hipStreamSynchronize(st_));
      arena_off_ = 0;
    }
    std::memcpy(arena_ + arena_off_, src, bytes);
    hipLaunchKernelGGL(k_stage, dim3(nblocks((int64_t)n16)), dim3(WG), 0, st_,
                       reinterpret_cast<const int4*>(arena_dev_ + arena_off_), reinterpret_cast<int4*>(dst), (int)n16);
    arena_off_ += n16 * 16;
  }

  // uniform_generate<D, 1>(rsa_domain([0, side]^D, 1, r), r, nbshots, 10, seed): the
  // domain, the generator of one class (fraction 1, radius_generator.ixx:71-86)
  // and the start of proceed()
  std::unique_ptr<R> make(int k) {
    auto b = std::make_unique<R>();
    b->k = k;
    b->t0 = std::chrono::steady_clock::now();
    double inf[D], sup[D], v = 1.;
    for (int d = 0; d < D; d++) {
      inf[d] = 0.;
      sup[d] = side_[k];
      v *= side_[k];
    }
    b->e.reset(new E(inf, sup, 1, radius_[k], st_));
    E& e = *b->e;
    const uint64_t cnt = static_cast<uint64_t>(v * 1.0 / (volume_unit_sphere(D) * puissance<D>(radius_[k])));
    e.algo_init({RadiusClass{radius_[k], cnt, 0}}, 0.0, nbshots_);
    // proceed(seed): the word stream, the helper thread (which also copies each
    // prepared draw's words to the device), the initial lattice
    e.prep_words_ = true;
    e.reset_dev_ring();
    e.hs_.start(seed_[k]);
    e.hs_.wait_ns = 0;
    e.prep_start();
    e.lst_ = rsa_gpu_learn_stats{};
    e.level_ = 0;
    e.last_miss_ = 0.;
    e.since_refine_ = 0;
    e.miss_hist_.clear();
    e.init_voxels(e.r_max_ + e.gmin_);
    b->st.nvox_max = e.nvox_;
    return b;
  }

  void admit_loop() {
    StreamScope scope(st_, nullptr);
    try {
      while (next_ < n_) {
        {
          std::unique_lock<std::mutex> lk(qm_);
          qcv_.wait(lk, [&] { return qstop_ || (int)pre_.size() < kPre; });
          if (qstop_) return;
        }
        auto b = make(next_++);
        std::lock_guard<std::mutex> lk(qm_);
        pre_.push_back(std::move(b));
        qcv_.notify_all();
      }
    } catch (const std::exception& ex) {
      std::lock_guard<std::mutex> lk(qm_);
      if (qerr_.empty()) qerr_ = ex.what();
      qcv_.notify_all();
    }
  }

  // set-up realizations join the live group while it has room (waiting for one
  // only when nothing is live)
  void admit() {
    std::unique_lock<std::mutex> lk(qm_);
    if (live_.empty() && taken_ < n_) {  // a full start: as many as the queue holds
      const int want = std::min({active_, n_ - taken_, kPre});
      qcv_.wait(lk, [&] { return (int)pre_.size() >= want || !qerr_.empty(); });
    }
    if (!qerr_.empty()) throw Error(qerr_);
    bool took = false;
    while ((int)live_.size() < active_ && !pre_.empty()) {
      pre_.front()->half = taken_ & 1;
      live_.push_back(std::move(pre_.front()));
      pre_.pop_front();
      taken_++;
      took = true;
    }
    if (took) qcv_.notify_all();
    max_active_ = std::max<int64_t>(max_active_, (int64_t)live_.size());
  }

  // the top of proceed()'s loop for one realization: ends it, runs a host-driven
  // draw, or prepares the first round of a fast draw (into `first`)
  bool any_refine() const {
    for (auto& b : live_)
      if (b->s == R::REFINE) return true;
    return false;
  }

  // (without `wait`, a realization whose draw is not prepared yet stays READY)
  void plan(R& b, std::vector<R*>& first, bool wait) {
    E& e = *b.e;
    if (e.nvox_ == 0) {
      b.st.loop_iterations++;
      b.st.end_reason = 1;
      b.s = R::DONE;
      return;
    }
    e.hs_.set_low(b.w);
    if (wait) {
      timed(t_prep_, [&] { b.cur = e.prep_get(b.w, e.nvox_); });
    } else if (!e.prep_try_get(b.w, e.nvox_, b.cur)) {
      return;
    }
    b.st.loop_iterations++;
    b.st.loop_iterations += b.cur.zero_iters;
    const uint64_t nshots = (uint64_t)b.cur.nshots;
    const uint64_t maximal = e.current_number();
    const bool may_outreach = nshots >= maximal;
    if (may_outreach || nshots > 48000) {  // the host-driven form, alone on the stream
      const double* hrad = nullptr;
      if (b.cur.law) throw Error("batch: custom radius laws are not batched");
      int64_t rounds = 0;
      uint64_t w_next = 0;
      const uint64_t nb_new =
          (uint64_t)e.single_draw_slow(b.cur.nshots, hrad, maximal, may_outreach, b.cur.w_shots, rounds, w_next, &b.st);
      after_draw(b, nb_new, rounds, false, false);
      return;
    }
    e.ensure_cand(b.cur.nshots);
    e.init_pub();
    Prep& pr = b.cur;
    pr.id = ++e.draw_id_;
    pr.seq = ++e.seq_;
    pr.miss_base = (uint64_t)(int)e.nb_fast_ - b.old_nb;
    // the helper thread made the words readable on the device (prep_words_)
    b.seq = pr.seq;
    b.r = 0;
    b.lists_kept = false;
    b.local_new = 0;
    b.first = true;
    b.refilter = false;
    b.lists = false;
    b.s = R::ROUND;
    first.push_back(&b);
  }

  void fill(R& b, DrawDesc<D>& a) {
    E& e = *b.e;
    const Prep& pr = b.cur;
    std::memset(&a, 0, sizeof(a));
    a.ring = e.ring_.p;
    a.mask = E::kRing - 1;
    a.w0 = pr.w_shots;
    a.nshots = pr.nshots;
    a.N = (uint32_t)e.nvox_;
    a.vpos = e.hvpos_dev_[pr.buf];
    a.hrad = pr.law ? e.hrad_dev_[pr.buf] : nullptr;
    a.end_rel = pr.end_rel;
    a.vx = e.vview();
    a.cls = e.class_dev();
    a.g = e.geo_;
    a.cs = e.cs();
    a.cpos = e.cpos_.p;
    a.crad = e.crad_.p;
    a.cprio = e.cprio_.p;
    a.cphase = e.cphase_.p;
    a.ccell = e.ccell_.p;
    a.ctag = e.ctag_.p;
    a.owner = e.owner_.p;
    a.cellcc = e.cellcc_.p;
    a.cslot = e.cslot_.p;
    a.ovhead = e.ovhead_.p;
    a.next = e.next_.p;
    a.ctl = e.dctl_.p;
    a.draw_id = pr.id;
    a.forced = 1;
    a.list_thr = e.list_thr_;
    a.lab = e.lab_.p;
    a.cell2t = e.cell2t_.p;
    a.seg = e.seg_.p;
    a.len = e.len_.p;
    a.ord = e.ord_.p;
    a.slab = e.slab_.p;
    a.pool = e.pool_.p;
    a.pool_cap = (int)e.pool_.cap;
    a.gate_id = b.first ? pr.id : -1;
    a.urec = e.urec_.p;
    a.seq = b.seq;
    a.aimed = e.aimed_;
    a.miss_base = pr.miss_base;
    a.host = e.hpub_dev_;
    a.ncells = e.geo_.ncells;
    a.rl = pr.rej;
    a.refilter = b.refilter ? 1 : 0;
    a.lists = b.lists ? 1 : 0;
  }

  // one round of every draw in `group` (a first round, or the next one of draws that
  // need it): one launch per kernel for all of them
  void launch(const std::vector<R*>& group, bool first_round) {
    using LG = LaneGroup<D>;
    ticks_++;
    draw_groups_++;
    int mx = 1;
    tab_.resize(group.size());
    bool any_refilter = false, any_lists = false;
    for (size_t q = 0; q < group.size(); q++) {
      fill(*group[q], tab_[q]);
      mx = std::max(mx, group[q]->cur.nshots);
      any_refilter = any_refilter || group[q]->refilter;
      any_lists = any_lists || group[q]->lists;
    }
    stage(tab_.data(), sizeof(DrawDesc<D>) * tab_.size(), dtab_.p);
    const auto* dt = reinterpret_cast<const DrawDesc<D>*>(dtab_.p);
    const int nr = (int)group.size();
    const int bpl = (int)nblocks((int64_t)mx * LG::LPC), bps = (int)nblocks(mx);
    if (first_round) hipLaunchKernelGGL(kb_decode_filter<D>, dim3(nr * bpl), dim3(WG), 0, st_, dt, bpl);
    if (any_lists) hipLaunchKernelGGL(kb_lists<D>, dim3(nr * bps), dim3(WG), 0, st_, dt, bps);
    if (any_refilter) {
      hipLaunchKernelGGL(kb_refilter<D>, dim3(nr * bpl), dim3(WG), 0, st_, dt, bpl);
      hipLaunchKernelGGL(kb_remove_skips<D>, dim3(nr * bps), dim3(WG), 0, st_, dt, bps);
    }
    hipLaunchKernelGGL(kb_partner_lists<D>, dim3(nr * bpl), dim3(WG), 0, st_, dt, bpl);
    // the owners' workgroups and one more that publishes the round, per realization
    const int bsc = (mx + kOwnersPerBlock - 1) / kOwnersPerBlock + 1;
    hipLaunchKernelGGL(kb_sweep_commit<D>, dim3(nr * bsc), dim3(kSweepThreads), (size_t)mx, st_, dt, bsc);
    HIP_OK(hipGetLastError());
    for (R* b : group) {
      b->first = false;
      b->refilter = false;
      b->lists = false;
    }
  }

  // the results of a launched round (finish_draw's loop body, for the group): the
  // draws that need another round are returned, flagged for it
  std::vector<R*> collect(const std::vector<R*>& group) {
    std::vector<R*> again;
    for (R* bp : group) {
      R& b = *bp;
      E& e = *b.e;
      timed(t_wait_, [&] { e.wait_pub(b.seq); });
      const DrawPub h = e.hpub_[b.seq & (kPubSlots - 1)];
      if (h.label_stuck) throw Error("internal: conflict labelling made no progress");
      if (h.cell_overflow) throw Error("internal: cell capacity exceeded");
      b.st.ghost_cell_drops += h.dropped;
      b.st.image_cell_mismatch += h.img_mismatch;
      auto build_lists = [&]() {  // k_lists (forced) in the next launch, before its round
        b.lists = true;
        b.lists_kept = true;
      };
      if (h.need_lists) {  // a cell got more than CSLOT candidates: build the lists, redo the first round
        build_lists();
        b.seq = ++e.seq_;
        again.push_back(&b);
        continue;
      }
      if (h.pool_overflow) {  // long partner lists: grow the pool and redo the round
        HIP_OK(hipStreamSynchronize(st_));
        e.pool_.reserve((size_t)h.need_pool * 2);
        HIP_OK(hipMemsetAsync(&e.dctl_.p->pool_cursor, 0, 3 * sizeof(int), st_));
        if (b.r == 0 && !b.lists_kept) build_lists();
        b.seq = ++e.seq_;
        again.push_back(&b);
        continue;
      }
      b.local_new += h.acc;
      if (h.con == 0) {
        after_draw(b, (uint64_t)b.local_new, b.r + 1, true, h.gate_open != 0);
        continue;
      }
      if (++b.r >= 100000) throw Error("too many conflict rounds in one draw");
      b.lists_kept = true;
      b.refilter = true;
      b.seq = ++e.seq_;
      again.push_back(&b);
    }
    return again;
  }

  // proceed()'s bookkeeping after a draw
  void after_draw(R& b, uint64_t nb_new, int64_t rounds, bool fast, bool gate_open) {
    E& e = *b.e;
    draws_++;
    const uint64_t nshots = (uint64_t)b.cur.nshots;
    b.w = b.cur.w_end();
    b.st.draws++;
    b.st.shots += (int64_t)nshots;
    b.st.accepted += (int64_t)nb_new;
    b.st.rounds += rounds;
    b.st.rounds_max = std::max(b.st.rounds_max, rounds);
    e.update_placed(nb_new);
    b.cont = e.still();
    e.nb_fast_ += nb_new;
    const int new_nb = (int)e.nb_fast_;
    const uint64_t nb_miss = nshots - (uint64_t)(new_nb - b.old_nb);
    b.old_nb = (uint64_t)new_nb;
    const double miss_rate = nb_miss / (0.000001 + nshots);
    const bool refine_now = miss_rate > e.aimed_;
    if (fast && gate_open == refine_now) throw Error("internal: device and host refinement decisions differ");
    e.miss_hist_.push_back(miss_rate);
    e.last_miss_ = miss_rate;
    e.since_refine_ = refine_now ? 0 : e.since_refine_ + 1;
    b.cur = Prep{};
    if (refine_now) {
      b.s = R::REFINE;
    } else if (!b.cont) {
      b.st.end_reason = 2;
      b.s = R::DONE;
    } else {
      b.s = R::READY;
    }
  }

  // remove_covered + subdivide_uncovered (Engine::refine, with its thresholds for
  // the grouped order and the packed copy) for every realization of `group` at once
  void refine(const std::vector<R*>& group) {
    refine_groups_++;
    ticks_++;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<R*> g;
    for (R* b : group) {
      if (b->e->nvox_ > 0) g.push_back(b);
    }
    const int n = (int)g.size();
    if (n > 0) {
      std::vector<RefDesc<D>> tab(n);
      RefHead head{};
      head.n = n;
      // scans: 0 cell counts of grouped lists, 1 packed-copy offsets, 2 uncovered
      // voxels per block (keep), 3 children per block
      ScanHead sh[4] = {};
      std::vector<ScanDesc> sd[4];
      bool any_g = false, any_p = false, any_wide = false;
      for (int q = 0; q < n; q++) {
        E& e = *g[q]->e;
        const int64_t nv = e.nvox_;
        const unsigned nb = nblocks(nv);
        const int64_t nc = e.geo_.ncells;
        const bool grouped = nv >= e.group_min_;
        const bool packed = ipow_c(3, D) <= 32 && nv >= e.pack_min_;
        e.code_.reserve(nv);
        e.moved_.reserve(nv);
        e.blk_.reserve(nb);
        e.holes_.reserve(nv);
        e.tsum_.reserve(std::max(nblocks(nb, SCAN_T), (grouped || packed) ? nblocks(nc, SCAN_T) : 1u));
        RefDesc<D>& a = tab[q];
        std::memset(&a, 0, sizeof(a));
        a.g = e.geo_;
        a.cs = e.cs();
        a.vx = e.vview();
        any_wide = any_wide || a.vx.wide;
        a.vp = e.vparams(e.gmin_);
        a.nvox = nv;
        a.code = e.code_.p;
        a.moved = e.moved_.p;
        a.blk = e.blk_.p;
        a.holes = e.holes_.p;
        a.ctr = e.dctr_.p;
        if (grouped) {
          e.perm_.reserve(nv);
          e.vcell_.reserve(nv);
          e.vrank_.reserve(nv);
          e.ccount_.reserve(nc);
          a.perm = e.perm_.p;
          a.vcell = e.vcell_.p;
          a.vrank = e.vrank_.p;
          a.ccount = e.ccount_.p;
          sd[0].push_back(ScanDesc{e.ccount_.p, nc, e.tsum_.p, e.dctr_.p + C_GROUPED});
          any_g = true;
        }
        if (packed) {
          const int64_t ns = (int64_t)(e.nb_fast_ + e.added_);
          e.coff_.reserve((size_t)nc + 1);
          e.csph_.reserve((size_t)ns + 1);
          e.cprad_.reserve((size_t)ns + 1);
          a.coff = e.coff_.p;
          a.csph = e.csph_.p;
          a.crad = e.cprad_.p;
          a.ncs = (int)std::min<int64_t>(ns, INT32_MAX);
          a.ncap = ns;
          sd[1].push_back(ScanDesc{e.coff_.p, nc, e.tsum_.p, e.coff_.p + nc});
          any_p = true;
        }
        sd[2].push_back(ScanDesc{e.blk_.p, (int64_t)nb, e.tsum_.p, e.dctr_.p + C_KEEP});
        sd[3].push_back(ScanDesc{e.blk_.p, (int64_t)nb, e.tsum_.p, e.dctr_.p + C_NCHILD});
        const int nbw = (int)nblocks(nv, 4 * WG), ncb = (int)nblocks(nc);
        head.nb[q + 1] = head.nb[q] + (int)nb;
        head.nbw[q + 1] = head.nbw[q] + nbw;
        head.nbg[q + 1] = head.nbg[q] + (grouped ? nbw : 0);
        head.ncg[q + 1] = head.ncg[q] + (grouped ? ncb : 0);
        head.ncp[q + 1] = head.ncp[q] + (packed ? ncb : 0);
      }
      for (int k = 0; k < 4; k++) {
        sh[k].n = (int)sd[k].size();
        for (int q = 0; q < sh[k].n; q++) sh[k].nt[q + 1] = sh[k].nt[q] + (int)nblocks(sd[k][q].n, SCAN_T);
      }
      // one staging copy for the group's tables
      std::vector<char> blob;
      auto put = [&](const void* p, size_t bytes) {
        const size_t off = (blob.size() + 15) & ~size_t(15);
        blob.resize(off + bytes);
        std::memcpy(blob.data() + off, p, bytes);
        return off;
      };
      const size_t o_head = put(&head, sizeof(head));
      const size_t o_tab = put(tab.data(), sizeof(RefDesc<D>) * n);
      size_t o_sh[4], o_sd[4];
      for (int k = 0; k < 4; k++) {
        o_sh[k] = put(&sh[k], sizeof(ScanHead));
        o_sd[k] = put(sd[k].data(), sizeof(ScanDesc) * sd[k].size());
      }
      rblob_.reserve(blob.size() + 16);
      stage(blob.data(), blob.size(), rblob_.p);
      const uint8_t* base = rblob_.p;
      auto* rh = reinterpret_cast<const RefHead*>(base + o_head);
      auto* rt = reinterpret_cast<const RefDesc<D>*>(base + o_tab);
      auto scan = [&](int k) {
        if (sh[k].n == 0) return;
        auto* h = reinterpret_cast<const ScanHead*>(base + o_sh[k]);
        auto* t = reinterpret_cast<const ScanDesc*>(base + o_sd[k]);
        hipLaunchKernelGGL(kbs_tile_sums, dim3(sh[k].nt[sh[k].n]), dim3(1024), 0, st_, h, t);
        hipLaunchKernelGGL(kbs_scan_blocks, dim3(sh[k].n), dim3(1024), 0, st_, h, t);
        hipLaunchKernelGGL(kbs_tile_scan, dim3(sh[k].nt[sh[k].n]), dim3(1024), 0, st_, h, t);
      };
      if (any_g) {
        hipLaunchKernelGGL(kb_clear_ccount<D>, dim3(head.ncg[n]), dim3(WG), 0, st_, rh, rt);
        hipLaunchKernelGGL((kb_bin<D, false>), dim3(head.nbg[n]), dim3(WG), 0, st_, rh, rt);
        if (any_wide) hipLaunchKernelGGL((kb_bin<D, true>), dim3(head.nbg[n]), dim3(WG), 0, st_, rh, rt);
        scan(0);
        hipLaunchKernelGGL(kb_place<D>, dim3(head.nbg[n]), dim3(WG), 0, st_, rh, rt);
      }
      if (any_p) {
        hipLaunchKernelGGL(kb_cell_cnt<D>, dim3(head.ncp[n]), dim3(WG), 0, st_, rh, rt);
        scan(1);
        hipLaunchKernelGGL(kb_compact<D>, dim3(head.ncp[n]), dim3(WG), 0, st_, rh, rt);
      }
      hipLaunchKernelGGL((kb_cover<D, false>), dim3(head.nb[n]), dim3(WG), 0, st_, rh, rt);
      if (any_wide) hipLaunchKernelGGL((kb_cover<D, true>), dim3(head.nb[n]), dim3(WG), 0, st_, rh, rt);
      if (any_g) hipLaunchKernelGGL(kb_block_count<D>, dim3(head.nbg[n]), dim3(WG), 0, st_, rh, rt);
      scan(2);
      hipLaunchKernelGGL(kb_holes<D>, dim3(head.nbw[n]), dim3(WG), 0, st_, rh, rt);
      hipLaunchKernelGGL((kb_move<D, false>), dim3(head.nbw[n]), dim3(WG), 0, st_, rh, rt);
      if (any_wide) hipLaunchKernelGGL((kb_move<D, true>), dim3(head.nbw[n]), dim3(WG), 0, st_, rh, rt);
      hipLaunchKernelGGL(kb_child_count<D>, dim3(head.nbw[n]), dim3(WG), 0, st_, rh, rt);
      scan(3);
      hipLaunchKernelGGL(kb_gather_ctr<D>, dim3(n), dim3(64), 0, st_, rt, n, hctr_dev_);
      HIP_OK(hipGetLastError());
      timed(t_refine_wait_, [&] { HIP_OK(hipStreamSynchronize(st_)); });
      // the children of every list that keeps voxels
      std::vector<int> nxt(n, -1);
      int m = 0;
      RefHead h2{};
      std::vector<RefDesc<D>> tab2;
      for (int q = 0; q < n; q++) {
        E& e = *g[q]->e;
        std::memcpy(e.hctr_, hctr_ + (size_t)q * C_NUM, sizeof(int) * C_NUM);
        e.check_scatter();  // the previous refinement's children, this one's packed copy
        e.level_++;
        const int64_t keep = e.hctr_[C_KEEP];
        if (keep == 0) {  // subdivide_uncovered returns early on an empty list (list_of_voxels.ixx:60)
          e.nvox_ = 0;
          continue;
        }
        const int64_t nchild = e.hctr_[C_NCHILD];
        if (nchild < 0) throw Error("voxel list would exceed 2^31 entries");
        if (e.max_coord_ > ((int64_t)1 << 61)) throw Error("voxel lattice coordinates exceed 62 bits");
        const int x = 1 - e.vcur_;
        e.wide_[x] = e.wide_[e.vcur_] || 2 * e.max_coord_ + 1 > (int64_t)INT32_MAX;
        for (int d = 0; d < D; d++) {
          if (e.wide_[x])
            e.vw_[x][d].reserve(std::max<int64_t>(nchild, 1));
          else
            e.vc_[x][d].reserve(std::max<int64_t>(nchild, 1));
        }
        RefDesc<D> a = tab[q];
        a.vx = e.vview();
        a.dst = e.vview(x);
        a.keep = keep;
        a.nchild = nchild;
        tab2.push_back(a);
        nxt[q] = x;
        h2.nsc[m + 1] = h2.nsc[m] + (int)nblocks(keep);
        m++;
      }
      h2.n = m;
      if (m > 0) {
        blob.clear();
        const size_t p_head = put(&h2, sizeof(h2));
        const size_t p_tab = put(tab2.data(), sizeof(RefDesc<D>) * m);
        stage(blob.data(), blob.size(), rblob_.p);
        auto* sh2 = reinterpret_cast<const RefHead*>(base + p_head);
        auto* st2 = reinterpret_cast<const RefDesc<D>*>(base + p_tab);
        // one instance per (list width, children width) present
        bool kinds[2][2] = {};
        for (const auto& a : tab2) kinds[a.vx.wide ? 1 : 0][a.dst.wide ? 1 : 0] = true;
        if (kinds[0][0]) hipLaunchKernelGGL((kb_scatter<D, false, false>), dim3(h2.nsc[m]), dim3(WG), 0, st_, sh2, st2);
        if (kinds[0][1]) hipLaunchKernelGGL((kb_scatter<D, false, true>), dim3(h2.nsc[m]), dim3(WG), 0, st_, sh2, st2);
        if (kinds[1][0]) hipLaunchKernelGGL((kb_scatter<D, true, false>), dim3(h2.nsc[m]), dim3(WG), 0, st_, sh2, st2);
        if (kinds[1][1]) hipLaunchKernelGGL((kb_scatter<D, true, true>), dim3(h2.nsc[m]), dim3(WG), 0, st_, sh2, st2);
        HIP_OK(hipGetLastError());
      }
      for (int q = 0; q < n; q++) {
        E& e = *g[q]->e;
        // the next draw's Poisson count depends on the new list length: its
        // preparation starts now, while the device scatters the children
        if (nxt[q] < 0) {
          if (e.nvox_ > 0) e.prep_request(g[q]->w, e.nvox_);
          continue;
        }
        e.prep_request(g[q]->w, e.hctr_[C_NCHILD]);
        e.vcur_ = nxt[q];
        e.nvox_ = e.hctr_[C_NCHILD];
        e.max_coord_ = 2 * e.max_coord_;
        for (int d = 0; d < D; d++) e.vlen_[d] = 0.5 * e.vlen_[d];
      }
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (R* b : group) {
      refines_++;
      b->st.refines++;
      b->st.voxel_seconds += secs / (double)group.size();
      b->st.nvox_max = std::max<int64_t>(b->st.nvox_max, b->e->nvox_);
      if (!b->cont) {
        b->st.end_reason = 2;
        b->s = R::DONE;
      } else {
        b->s = R::READY;
      }
    }
  }

  // the end of proceed() for a finished realization (end-of-run checks on the
  // device), then its results; its engine is released (finish thread)
  void finish_one(R& b) {
    E& e = *b.e;
    const int64_t nsamp =
        std::min<int64_t>(int64_t(1) << 22, std::max<int64_t>(int64_t(1) << 16, 4 * (int64_t)e.nb_fast_));
    HIP_OK(hipMemsetAsync(e.dctr_.p + C_OVERLAP, 0, 2 * sizeof(int), st_));
    hipLaunchKernelGGL(k_check_overlap<D>, dim3(nblocks(e.geo_.ncells)), dim3(WG), 0, st_, e.geo_, e.cs(), e.dctr_.p);
    hipLaunchKernelGGL(k_check_cover<D>, dim3(nblocks(nsamp)), dim3(WG), 0, st_, e.geo_, e.cs(), nsamp,
                       0x5eed0c0ffee5eedull ^ seed_[b.k], e.gmin_, e.dctr_.p);
    HIP_OK(hipGetLastError());
    b.st.cover_samples = nsamp;
    e.read_ctr();
    b.st.overlap_pairs = e.hctr_[C_OVERLAP] / 2;
    b.st.cover_free = e.hctr_[C_UNCOVERED];
    e.check_scatter();
    e.prep_halt();
    e.hs_.halt();
    b.st.nvox_final = e.nvox_;
    b.st.words = (int64_t)b.w;
    b.st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - b.t0).count();
    const rsa_gpu_stats& s = b.st;
    if (s.end_reason != 1 || s.nvox_final != 0 || s.overlap_pairs != 0 || (D <= 3 && s.cover_free != 0))
      throw Error("batch: realization " + std::to_string(b.k) + " failed the end-of-run checks (" +
                  std::to_string(s.overlap_pairs) + " overlapping pairs, " + std::to_string(s.cover_free) +
                  " free sample points)");
    const int k = b.k;
    count_[k] = e.count();
    if (volume_) volume_[k] = e.volume_of_spheres();
    if (extract_) {
      pos_[k].resize(count_[k] * D);
      rad_[k].resize(count_[k]);
      e.extract(pos_[k].data(), rad_[k].data(), nullptr, count_[k]);
    }
  }

  void finish_loop() {
    StreamScope scope(st_, nullptr);
    try {
      for (;;) {
        std::unique_ptr<R> b;
        {
          std::unique_lock<std::mutex> lk(qm_);
          qcv_.wait(lk, [&] { return qstop_ || !fin_q_.empty(); });
          if (fin_q_.empty()) return;  // stopped with nothing left
          b = std::move(fin_q_.front());
          fin_q_.pop_front();
          fin_busy_ = true;
        }
        finish_one(*b);
        b.reset();  // the engine's buffers go back to the pools
        std::lock_guard<std::mutex> lk(qm_);
        fin_busy_ = false;
        qcv_.notify_all();
      }
    } catch (const std::exception& ex) {
      std::lock_guard<std::mutex> lk(qm_);
      if (qerr_.empty()) qerr_ = ex.what();
      fin_busy_ = false;
      qcv_.notify_all();
    }
  }

  // finished realizations go to the finish thread
  void finish() {
    std::lock_guard<std::mutex> lk(qm_);
    if (!qerr_.empty()) throw Error(qerr_);
    bool any = false;
    for (auto& b : live_)
      if (b->s == R::DONE) {
        fin_q_.push_back(std::move(b));
        any = true;
      }
    if (!any) return;
    live_.erase(std::remove_if(live_.begin(), live_.end(), [](const std::unique_ptr<R>& b) { return !b; }),
                live_.end());
    qcv_.notify_all();
  }

  // the last finished realizations, then both threads end
  void drain() {
    {
      std::unique_lock<std::mutex> lk(qm_);
      qcv_.wait(lk, [&] { return (fin_q_.empty() && !fin_busy_) || !qerr_.empty(); });
    }
    stop_threads();
    if (!qerr_.empty()) throw Error(qerr_);
  }

  void stop_threads() {
    {
      std::lock_guard<std::mutex> lk(qm_);
      qstop_ = true;
      if (!qerr_.empty()) fin_q_.clear();  // after an error nothing more is finished
    }
    qcv_.notify_all();
    if (admit_th_.joinable()) admit_th_.join();
    if (fin_th_.joinable()) fin_th_.join();
  }
};

void batch_queue_run(int dim, int n, const double* side, const double* radius, const uint64_t* seed, int nbshots,
                     int active, uint64_t* count, double* volume, double* pos, double* radius_out, uint64_t cap,
                     rsa_gpu_batch_stats* bst) {
  auto go = [&](auto tag) {
    constexpr int DD = decltype(tag)::value;
    BatchDriver<DD> drv(n, side, radius, seed, nbshots, active, count, volume, pos != nullptr, bst);
    drv.run();
    if (pos) {
      uint64_t off = 0;
      for (int k = 0; k < n; k++) {
        if (off + count[k] > cap) throw Error("batch: output capacity exceeded");
        std::memcpy(pos + off * DD, drv.pos_[k].data(), drv.pos_[k].size() * sizeof(double));
        std::memcpy(radius_out + off, drv.rad_[k].data(), drv.rad_[k].size() * sizeof(double));
        off += count[k];
      }
    }
  };
  switch (dim) {
    case 2:
      go(std::integral_constant<int, 2>{});
      break;
    case 3:
      go(std::integral_constant<int, 3>{});
      break;
    case 4:
      go(std::integral_constant<int, 4>{});
      break;
    default:
      throw Error("batch: dim must be 2, 3 or 4");
  }
}

}  // namespace rsa_amd
