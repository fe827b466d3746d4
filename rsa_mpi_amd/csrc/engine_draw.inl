// Host side of a draw (member functions of Engine<D>, included by engine.hip
// after the class): the outreach pivot, the Lemire chain and Poisson count on
// the helper thread, the publication ring, the queued draw kernels and the
// host-driven form for draws that may overshoot a class target.
// Not a header of its own: engine.hip includes it once, after Engine<D>.

namespace rsa_amd {

// median_of_medians::find_pivot_for_first_elements on one rank (median_of_medians.ixx:8-178)
template <int D>
double Engine<D>::mom_pivot(std::vector<uint64_t> values, int64_t nb_first) {
  const int64_t gsize = (int64_t)values.size();
  if (nb_first > gsize)
    throw Error("Impossible to find the pivot for N first elements if N is smaller than the size of the array");
  if (nb_first <= 0) throw Error("Impossible to find the pivot for N first elements if N  <= 0");
  if (gsize == 0) throw Error("Impossible to find the pivot for N first elements if no element");
  const double pct = nb_first / (1. * gsize);
  std::sort(values.begin(), values.end(), std::less<uint64_t>{});
  auto count_le = [&](double pivot) -> int64_t {  // upper_bound with less<uint64_t>
    if (values.empty()) return 0;
    auto it = std::upper_bound(values.begin(), values.end(), pivot, std::less<uint64_t>{});
    return (int)(it - values.begin());
  };
  double first, secnd;
  {
    int64_t piv = static_cast<int64_t>(values.size() * pct) - 1;
    int64_t mn = piv, mx = piv + 1;
    if (std::abs(values.size() * pct - 1 - piv) < 1e-8) mn -= 1;
    if (mn < 0) mn = 0;
    if (std::abs(values.size() * pct - 1 - piv - 1) < 1e-8) mx += 1;
    if (mx > int64_t(values.size()) - 1) mx = values.size() - 1;
    first = values[mn] - std::numeric_limits<double>::epsilon();
    secnd = values[mx] + std::numeric_limits<double>::epsilon();
  }
  std::vector<double> cand = {first, secnd};
  std::sort(cand.begin(), cand.end(), std::less<uint64_t>{});
  auto cmp = [&](int64_t nb, double p) { return nb <= count_le(p); };
  auto it = std::upper_bound(cand.begin(), cand.end(), nb_first, cmp);
  if (it == cand.end()) throw Error("Unexpected - 0");
  const double pmax = *it;
  double pivot = pmax;
  int64_t below = count_le(pivot);
  if (below != nb_first) {
    if (it == cand.begin()) throw Error("Unexpected - 1");
    const double pmin = *(it - 1);
    const double step = pmax - pmin;
    double add = 0.5 * step;
    for (int j = 0; j < 20; j++) {
      pivot = pmin + add;
      below = count_le(pivot);
      if (below == nb_first) break;
      if (below > nb_first) add = 0.5 * std::abs(add);
      if (below < nb_first) add = -0.5 * std::abs(add);
    }
  }
  return pivot;
}

// Voxel-index words of the shots (uniform_int_distribution's Lemire rejection,
// uniform_int_dist.h:246-268): the host holds the word stream anyway (Poisson
// draws from it), so it walks the rejections; the decode itself runs on the GPU.
template <int D>
int Engine<D>::host_chain(int nshots, uint64_t w0, uint32_t N, int b, RejList* rej) {
  const uint32_t thr = (uint32_t)(0u - N) % N;
  const uint64_t stride = 2 * D + 1;
  // offsets go to pinned buffer b when the draw takes the fast path
  uint32_t* out = nshots <= kFastShots ? hvpos_[b] : nullptr;
  uint64_t avail = w0 + (uint64_t)nshots * stride + 64;  // rejections are rare: extend when passed
  hs_.ensure(avail);
  const uint32_t* buf = hs_.buf;
  uint64_t p = w0;
  for (int i = 0; i < nshots; i++) {
    for (;;) {
      if (p >= avail) {
        avail = p + (uint64_t)(nshots - i) * stride + 64;
        hs_.ensure(avail);
      }
      if ((uint32_t)((uint64_t)buf[p & (kRing - 1)] * (uint64_t)N) >= thr) break;
      p++;
      if (rej && rej->n >= 0) {
        if (rej->n < KREJ)
          rej->idx[rej->n++] = i;
        else
          rej->n = -1;
      }
    }
    if (out) out[i] = (uint32_t)(p - w0);
    p += stride;
  }
  return (int)(p - w0);
}

// Radii of a draw's shots under a custom law: law(u) with u = generate_canonical
// over the 2 words per shot that follow the priorities (uniform_real_distribution
// on [0, 1), random.tcc:3346-3379; radius_generator.hxx:25-33).
template <int D>
void Engine<D>::law_radii(const Prep& pr, double* out) {
  thread_local std::vector<double> u;
  u.resize((size_t)pr.nshots);
  const uint64_t w = pr.w_shots + (uint64_t)pr.end_rel + (uint64_t)pr.nshots;
  hs_.ensure(w + 2 * (uint64_t)pr.nshots);
  for (int i = 0; i < pr.nshots; i++) u[i] = canonical(hs_.at(w + 2 * (uint64_t)i), hs_.at(w + 2 * (uint64_t)i + 1));
  law_(u.data(), out, pr.nshots, law_user_);
  // the cell grid is sized for max_radius and the voxel test uses min_radius
  for (int i = 0; i < pr.nshots; i++)
    if (!(out[i] >= gmin_ && out[i] <= gmax_))
      throw Error("radius law returned " + std::to_string(out[i]) + ", outside [min_radius, max_radius]");
}

// Poisson shot count(s) (compute_nb_shots_voxel, operator_algorithm.ixx:208-223)
// from position w until a draw has shots, then its rejection chain into pinned
// buffer b.  A prepared draw stays valid while the voxel count is unchanged.
template <int D>
typename Engine<D>::Prep Engine<D>::prepare(uint64_t w, int64_t nvox, int b) {
  Prep p;
  p.nv = nvox;
  p.w_start = w;
  p.buf = b;
  const int nv = (int)nvox;
  if (nv == 0) return p;
  const double intensity = 0.01 * nv;  // compute_intensity_poisson (:221-223)
  const double max_intensity = intensity;
  double mult = nbshots_ / max_intensity;
  if (mult > 1) mult = 1.;
  for (;;) {
    RingURNG g{&hs_, w};
    std::poisson_distribution<> pd(mult * intensity);
    const int n = pd(g);
    w = g.pos;
    if (n == 0) {
      p.zero_iters++;
      continue;
    }
    if ((uint64_t)n > (uint64_t)INT32_MAX / 16) throw Error("too many shots in one draw");
    p.nshots = n;
    break;
  }
  p.w_shots = w;
  p.rej.n = 0;
  p.end_rel = host_chain(p.nshots, w, (uint32_t)nvox, b, &p.rej);
  // entries past the list's end compare above every shot index
  if (p.rej.n >= 0)
    for (int t = p.rej.n; t < KREJ; t++) p.rej.idx[t] = INT_MAX;
  p.law = custom_;
  if (p.law && p.nshots <= kFastShots) law_radii(p, hrad_[b]);  // larger draws: in single_draw_slow
  return p;
}

// The helper thread: prepares the draws that follow (w, nv) while the main
// thread drives the GPU, up to kPrepAhead of them, each into the next of the
// kPrepBufs pinned offset buffers.  A request for another (w, nv) -- after a
// refinement -- discards what was prepared.
template <int D>
void Engine<D>::prep_start() {
  prep_halt();
  {
    std::lock_guard<std::mutex> lk(pm_);
    pstop_ = false;
    pgen_ = 0;
    preq_set_ = false;
    pready_.clear();
    perr_.clear();
  }
  pth_ = std::thread([this] {
    std::unique_lock<std::mutex> lk(pm_);
    uint64_t gen = 0, w = 0;
    int64_t nv = -1;
    for (;;) {
      pcv_.wait(lk, [&] { return pstop_ || (preq_set_ && (pgen_ != gen || (int)pready_.size() < kPrepAhead)); });
      if (pstop_) return;
      if (pgen_ != gen) {
        gen = pgen_;
        w = preq_w_;
        nv = preq_nv_;
      }
      const int b = pbuf_next_;
      pbuf_next_ = (pbuf_next_ + 1) % kPrepBufs;
      lk.unlock();
      Prep pr;
      std::string err;
      try {
        pr = prepare(w, nv, b);
        if (prep_words_ && pr.nshots > 0 && pr.nshots <= kFastShots)
          dev_words(pr.w_shots, pr.w_shots + (uint64_t)pr.end_rel + (uint64_t)pr.nshots);
      } catch (const std::exception& e) {
        err = e.what();
      }
      lk.lock();
      if (!err.empty()) {
        perr_ = err;
        pcv_.notify_all();
        return;
      }
      if (gen != pgen_) continue;  // superseded while it ran
      pready_.push_back(pr);
      w = pr.w_end();
      pnext_w_ = w;
      pcv_.notify_all();
    }
  });
}

template <int D>
void Engine<D>::prep_halt() {
  if (!pth_.joinable()) return;
  {
    std::lock_guard<std::mutex> lk(pm_);
    pstop_ = true;
  }
  pcv_.notify_all();
  pth_.join();
}

// The prepared draw starting at stream position w for voxel count nv: the
// queue's front, or the one the worker is preparing, or a new request.
template <int D>
void Engine<D>::prep_request(uint64_t w, int64_t nv) {
  std::lock_guard<std::mutex> lk(pm_);
  const bool front_ok = !pready_.empty() && pready_.front().w_start == w && pready_.front().nv == nv;
  const bool coming = pready_.empty() && preq_set_ && pnext_w_ == w && preq_nv_ == nv;
  if (!front_ok && !coming) {
    pready_.clear();
    pgen_++;
    preq_w_ = w;
    preq_nv_ = nv;
    pnext_w_ = w;
    preq_set_ = true;
    pcv_.notify_all();
  }
}

template <int D>
bool Engine<D>::prep_try_get(uint64_t w, int64_t nv, Prep& out) {
  prep_request(w, nv);
  std::lock_guard<std::mutex> lk(pm_);
  if (!perr_.empty()) throw Error(perr_);
  if (pready_.empty()) return false;
  out = pready_.front();
  pready_.pop_front();
  pcv_.notify_all();
  if (out.w_start != w || out.nv != nv) throw Error("internal: prepared draw out of sequence");
  return true;
}

template <int D>
typename Engine<D>::Prep Engine<D>::prep_get(uint64_t w, int64_t nv) {
  prep_request(w, nv);
  std::unique_lock<std::mutex> lk(pm_);
  pcv_.wait(lk, [&] { return !pready_.empty() || !perr_.empty(); });
  if (!perr_.empty()) throw Error(perr_);
  Prep pr = pready_.front();
  pready_.pop_front();
  pcv_.notify_all();
  if (pr.w_start != w || pr.nv != nv) throw Error("internal: prepared draw out of sequence");
  return pr;
}

template <int D>
void Engine<D>::wait_pub(int seq) {
  const auto t0 = std::chrono::steady_clock::now();
  volatile int* s = &hpub_[seq & (kPubSlots - 1)].seq;
  while (*s != seq) {
    if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(20)) {
      (void)hipStreamSynchronize(st_);
      if (*s == seq) break;
      throw Error("internal: draw results not published");
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
}

// generate_spheres + add_to_sample loop (operator_algorithm.ixx:534-582, 383-443)
// for draws that cannot overshoot the class target: three kernels per round and
// one host round trip (the round's counts reach pinned host memory from the
// publishing workgroup of k_sweep_commit; the tag update and commit run while
// the host moves on).
// enqueue_decode / enqueue_front / enqueue_settle queue a draw's first round;
// finish_draw waits for its rounds and queues the further ones.
template <int D>
void Engine<D>::init_pub() {
  if (hpub_) return;
  dctl_.reserve(1);
  HIP_OK(hipMemsetAsync(dctl_.p, 0, sizeof(DrawCtl), st_));
  hpub_ = static_cast<DrawPub*>(
      PinnedPool::get().alloc(sizeof(DrawPub) * kPubSlots, hipHostMallocCoherent | hipHostMallocMapped));
  HIP_OK(hipHostGetDevicePointer((void**)&hpub_dev_, hpub_, 0));
  std::memset(hpub_, 0, sizeof(DrawPub) * kPubSlots);
  pool_.reserve(1 << 16);
}

template <int D>
void Engine<D>::enqueue_round(const Prep& pr, int seq, int gate_id) {
  const int nshots = pr.nshots;
  using LG = LaneGroup<D>;
  const unsigned nbl = nblocks((int64_t)nshots * LG::LPC);
  const CandSet c = cset(pr.id);
  timed("k_partner_lists", (double)nshots * (8.0 * (D + 1) + 10.0), [&] {
    hipLaunchKernelGGL(k_partner_lists<D>, dim3(nbl), dim3(WG), 0, st_, geo_, nshots, c.cpos, c.crad, c.cprio,
                       c.ccell, c.ctag, lab_.p, cell2t_.p, seg_.p, len_.p, ord_.p, slab_.p, pool_.p,
                       (int)pool_.cap, dctl_.p, seq, gate_id, urec_.p, c.cellcc, c.cslot);
  });
  enqueue_settle(pr, seq, gate_id, nullptr);
}

// The next draw's decode inputs (its candidate buffers: the set of its id).
template <int D>
DecodeArgs<D> Engine<D>::decode_args(const Prep& pr, int forced) {
  const CandSet c = cset(pr.id);
  DecodeArgs<D> a{};
  a.ring = ring_.p;
  a.mask = kRing - 1;
  a.w0 = pr.w_shots;
  a.nshots = pr.nshots;
  a.N = (uint32_t)nvox_;
  a.vpos = hvpos_dev_[pr.buf];
  a.hrad = pr.law ? hrad_dev_[pr.buf] : nullptr;
  a.rl = pr.rej;
  a.end_rel = pr.end_rel;
  a.vx = vview();
  a.cls = class_dev();
  a.g = geo_;
  a.cs = cs();
  a.cpos = c.cpos;
  a.crad = c.crad;
  a.cprio = c.cprio;
  a.cphase = c.cphase;
  a.ccell = c.ccell;
  a.ctag = c.ctag;
  a.owner = c.owner;
  a.cellcc = c.cellcc;
  a.cslot = c.cslot;
  a.ovhead = c.ovhead;
  a.next = c.next;
  a.ctl = dctl_.p;
  a.draw_id = pr.id;
  a.forced = forced;
  a.list_thr = list_thr_;
  return a;
}

// A draw's decode on its own, forced (its launch opens the draw's gate): the
// first draw after the queue of draws broke off.  k_decode_filter reads the shot
// words from the device ring and the chain offsets from pinned host memory
// (written before this launch, untouched until the draw's results have been
// published).
template <int D>
void Engine<D>::enqueue_decode(Prep& pr) {
  pr.id = ++draw_id_;
  host_timed("host_dev_words", [&] { dev_words(pr.w_shots, pr.w_shots + (uint64_t)pr.end_rel + (uint64_t)pr.nshots); });
  using LG = LaneGroup<D>;
  const unsigned nbl = nblocks((int64_t)pr.nshots * LG::LPC);
  const DecodeArgs<D> a = decode_args(pr, 1);
  timed("k_decode_filter", (double)pr.nshots * (4.0 * (2 * D + 2) + 4.0 * D + 8.0 * (D + 1) + 13.0), [&] {
    auto kern = a.vx.wide ? k_decode_filter<D, true, true> : k_decode_filter<D, false, true>;
    hipLaunchKernelGGL(kern, dim3(nbl), dim3(WG), 0, st_, a.ring, a.mask, a.w0, a.nshots, a.N, a.vpos, a.hrad, a.rl,
                       a.end_rel, a.vx, a.cls, a.g, a.cs, a.cpos, a.crad, a.cprio, a.cphase, a.ccell, a.ctag, a.owner,
                       a.cellcc, a.cslot, a.ovhead, a.next, a.ctl, a.draw_id, a.forced, a.list_thr);
  });
}

// A decoded draw's first round up to its sweep: the validated filter (or, when
// its gate stays closed, the undoing of its decode's arrivals) and the partner
// lists, which also build the cell lists from the arrival slots (k_lists runs
// only when a cell got more than CSLOT candidates).
template <int D>
void Engine<D>::enqueue_front(Prep& pr, uint64_t miss_base) {
  pr.seq = ++seq_;
  pr.miss_base = miss_base;
  const int nshots = pr.nshots;
  using LG = LaneGroup<D>;
  const unsigned nbl = nblocks((int64_t)nshots * LG::LPC);
  const CandSet c = cset(pr.id);
  timed("k_filter", (double)nshots * (8.0 * (D + 1) + 5.0), [&] {
    hipLaunchKernelGGL(k_filter<D>, dim3(nbl), dim3(WG), 0, st_, geo_, cs(), nshots, c.cpos, c.crad, c.ccell, c.ctag,
                       c.cellcc, c.ovhead, dctl_.p, pr.id);
  });
  timed("k_partner_lists", (double)nshots * (8.0 * (D + 1) + 10.0), [&] {
    hipLaunchKernelGGL(k_partner_lists<D>, dim3(nbl), dim3(WG), 0, st_, geo_, nshots, c.cpos, c.crad, c.cprio,
                       c.ccell, c.ctag, lab_.p, cell2t_.p, seg_.p, len_.p, ord_.p, slab_.p, pool_.p,
                       (int)pool_.cap, dctl_.p, pr.seq, pr.id, urec_.p, c.cellcc, c.cslot);
  });
}

// A round's sweep, tag update, commit and publication; with `next`, the next
// draw's decode in the same launch (its id assigned here, its words made
// readable first).  The launch has the owners' workgroups and one more that
// publishes the round, then the decode's.
template <int D>
void Engine<D>::enqueue_settle(const Prep& pr, int seq, int gate_id, Prep* next) {
  const int nshots = pr.nshots;
  const unsigned nsc = (unsigned)((nshots + kOwnersPerBlock - 1) / kOwnersPerBlock + 1);
  const CandSet c = cset(pr.id);
  SweepArgs<D> sa{};
  sa.cs = cs();
  sa.nshots = nshots;
  sa.cpos = c.cpos;
  sa.crad = c.crad;
  sa.cphase = c.cphase;
  sa.ccell = c.ccell;
  sa.owner = c.owner;
  sa.ctag = c.ctag;
  sa.lab = lab_.p;
  sa.slab = slab_.p;
  sa.pool = pool_.p;
  sa.urec = urec_.p;
  sa.seg = seg_.p;
  sa.len = len_.p;
  sa.ord = ord_.p;
  sa.cell2t = cell2t_.p;
  sa.ctl = dctl_.p;
  sa.seq = seq;
  sa.gate_id = gate_id;
  sa.draw_id = pr.id;
  sa.aimed = aimed_;
  sa.miss_base = pr.miss_base;
  sa.host = hpub_dev_;
  sa.cellcc = c.cellcc;
  sa.cslot = c.cslot;
  sa.ncells = geo_.ncells;
  if (next == nullptr) {
    timed("k_sweep_commit", (double)nshots * (8.0 * (D + 1) + 16.0), [&] {
      hipLaunchKernelGGL(k_sweep_commit<D>, dim3(nsc), dim3(kSweepThreads), (size_t)nshots, st_, sa.cs, nshots,
                         sa.cpos, sa.crad, sa.cphase, sa.ccell, sa.owner, sa.ctag, sa.lab, sa.slab, sa.pool, sa.urec,
                         sa.seg, sa.len, sa.ord, sa.cell2t, sa.ctl, seq, gate_id, pr.id, sa.aimed, sa.miss_base,
                         sa.host, sa.cellcc, sa.cslot, sa.ncells);
    });
  } else {
    next->id = ++draw_id_;  // the publisher of this round opens its gate (id + 1) or not
    host_timed("host_dev_words",
               [&] { dev_words(next->w_shots, next->w_shots + (uint64_t)next->end_rel + (uint64_t)next->nshots); });
    using LG = LaneGroup<D>;
    const unsigned ndec = nblocks((int64_t)next->nshots * LG::LPC, kSweepThreads);
    const DecodeArgs<D> da = decode_args(*next, 2);
    timed("k_sweep_commit_decode", (double)nshots * (8.0 * (D + 1) + 16.0), [&] {
      auto kern = da.vx.wide ? k_sweep_commit_decode<D, true> : k_sweep_commit_decode<D, false>;
      hipLaunchKernelGGL(kern, dim3(nsc + ndec), dim3(kSweepThreads), (size_t)nshots, st_, sa, (int)nsc, da);
    });
  }
  HIP_OK(hipGetLastError());
}

// Waits for the rounds of draw pr (its first round queued), queues its further
// rounds; any kernel queued after the first round makes a queued next draw moot
// (its gate stays closed), which *next_live reports.
template <int D>
int64_t Engine<D>::finish_draw(const Prep& pr, int64_t& rounds, bool* next_live, bool& gate_open,
                               rsa_gpu_stats* st) {
  const int nshots = pr.nshots;
  using LG = LaneGroup<D>;
  const unsigned nbl = nblocks((int64_t)nshots * LG::LPC);
  const unsigned nbs = nblocks(nshots);
  const CandSet c = cset(pr.id);
  int seq = pr.seq;
  int64_t local_new = 0;
  int r = 0;
  // whether the cell lists of this draw are kept in seg/len/ord/cell2t: not after
  // a first round that gave up (those lists are built by its update kernel)
  bool lists_kept = false;
  auto build_lists = [&]() {  // k_lists, forced: once per draw, before the cell counts are reset
    timed("k_lists", (double)nshots * 24.0, [&] {
      hipLaunchKernelGGL(k_lists, dim3(nbs), dim3(WG), 0, st_, nshots, c.ccell, c.owner, c.ctag, c.cellcc, c.cslot,
                         c.ovhead, c.next, seg_.p, len_.p, ord_.p, cell2t_.p, dctl_.p, -1, geo_.ncells);
    });
    lists_kept = true;
  };
  for (;;) {
    host_timed("host_wait_pub", [&] { wait_pub(seq); });
    const DrawPub h = hpub_[seq & (kPubSlots - 1)];
    if (h.label_stuck) throw Error("internal: conflict labelling made no progress");
    if (h.cell_overflow) throw Error("internal: cell capacity exceeded");
    st->ghost_cell_drops += h.dropped;
    st->image_cell_mismatch += h.img_mismatch;
    if (h.need_lists) {  // a cell got more than CSLOT candidates: build the lists, redo the first round
      *next_live = false;
      HIP_OK(hipMemsetAsync(&dctl_.p->long_list[pr.id & 1], 0, sizeof(int), st_));
      build_lists();
      if (prof_) prof_add("host_list_redo", 0, 0);
      seq = ++seq_;
      enqueue_round(pr, seq, -1);
      continue;
    }
    if (h.pool_overflow) {  // long partner lists: grow the pool and redo the round
      *next_live = false;
      HIP_OK(hipStreamSynchronize(st_));  // nothing queued may hold the old pool
      pool_.reserve((size_t)h.need_pool * 2);
      HIP_OK(hipMemsetAsync(&dctl_.p->pool_cursor, 0, 3 * sizeof(int), st_));
      if (r == 0 && !lists_kept) build_lists();  // the first round's update kernel did not run
      seq = ++seq_;
      enqueue_round(pr, seq, -1);
      continue;
    }
    local_new += h.acc;
    if (h.con == 0) {
      rounds = r + 1;
      gate_open = h.gate_open != 0;
      break;
    }
    if (++r >= 100000) throw Error("too many conflict rounds in one draw");
    lists_kept = true;  // kept by the round's update kernel
    *next_live = false;
    timed("k_refilter", (double)nshots * (8.0 * (D + 1) + 5.0), [&] {
      hipLaunchKernelGGL(k_refilter<D>, dim3(nbl), dim3(WG), 0, st_, geo_, cs(), nshots, c.cpos, c.crad, c.ccell,
                         c.ctag, cell2t_.p, seg_.p, len_.p, ord_.p);
    });
    hipLaunchKernelGGL(k_remove_skips, dim3(nbs), dim3(WG), 0, st_, nshots, c.owner, seg_.p, len_.p, ord_.p, c.ctag);
    seq = ++seq_;
    enqueue_round(pr, seq, -1);
  }
  return local_new;
}

// Host-driven form of a draw; used when the draw may overshoot the class target
// (retain_only_first_elements needs the host-side pivot between sweep and update).
template <int D>
int64_t Engine<D>::single_draw_slow(int nshots, const double* hrad, uint64_t maximal, bool may_outreach, uint64_t w0,
                                    int64_t& rounds,
                               uint64_t& w_next, rsa_gpu_stats* st) {
  ensure_cand(nshots);
  const uint32_t N = (uint32_t)nvox_;
  const uint32_t thr = (uint32_t)(0u - N) % N;
  int nwin = nshots * (2 * D + 1) + std::max(1024, nshots / 4);
  const unsigned nbs = nblocks(nshots);
  for (;;) {  // word window: grows only if the rejections overflow it
    dev_words(w0, w0 + (uint64_t)nwin + (uint64_t)nshots + 64);
    reset_ctr();
    const size_t shmem = sizeof(uint32_t) * ((nwin + 31) / 32 + 1);
    if (shmem > 60000) throw Error("word window too large for one workgroup");
    timed("k_chain", (double)nwin * 4.0 + nshots * 4.0, [&] {
      hipLaunchKernelGGL(k_chain<D>, dim3(1), dim3(1024), shmem, st_, ring_.p, kRing - 1, w0, nwin, nshots, N, thr,
                         vpos_.p, dctr_.p);
    });
    read_ctr();
    if (!hctr_[C_CHAIN_OVF]) break;
    nwin *= 2;
  }
  const int end_rel = hctr_[C_END_REL];
  dev_words(w0, w0 + (uint64_t)end_rel + (uint64_t)nshots);
  const ClassDev cls = class_dev();
  hipLaunchKernelGGL(k_decode<D>, dim3(nbs), dim3(WG), 0, st_, ring_.p, kRing - 1, w0, nshots, N, vpos_.p, hrad,
                     vview(), cls, geo_, cpos_.p, crad_.p, cprio_.p, cphase_.p, ccell_.p, ctag_.p, dctr_.p);
  hipLaunchKernelGGL(k_bin_count, dim3(nbs), dim3(WG), 0, st_, nshots, ccell_.p, cellcc_.p, carr_.p, cell2t_.p,
                     touched_.p, dctr_.p);
  hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, st_, touched_.p, cellcc_.p, seg_.p, len_.p, dctr_.p);
  hipLaunchKernelGGL(k_bin_place, dim3(nbs), dim3(WG), 0, st_, nshots, ccell_.p, carr_.p, cell2t_.p, seg_.p, ord_.p);
  hipLaunchKernelGGL(k_bin_sort, dim3(nbs), dim3(WG), 0, st_, seg_.p, len_.p, ord_.p, dctr_.p);
  HIP_OK(hipGetLastError());
  w_next = w0 + (uint64_t)end_rel + (uint64_t)nshots;

  int64_t local_new = 0;
  rounds = 0;
  for (;;) {
    rounds++;
    // per-round counters
    HIP_OK(hipMemsetAsync(dctr_.p + C_CURSOR, 0, sizeof(int) * 4, st_));
    timed("k_filter", (double)nshots * (8.0 * (D + 1) + 5.0), [&] {
      hipLaunchKernelGGL(k_filter<D>, dim3(nbs), dim3(WG), 0, st_, geo_, cs(), nshots, cpos_.p, crad_.p, ccell_.p,
                         ctag_.p, cell2t_.p, seg_.p, len_.p, ord_.p);
    });
    hipLaunchKernelGGL(k_list_remove, dim3(nbs), dim3(WG), 0, st_, 0, seg_.p, len_.p, ord_.p, ctag_.p, dctr_.p);
    for (;;) {
      HIP_OK(hipMemsetAsync(dctr_.p + C_CURSOR, 0, sizeof(int), st_));
      hipLaunchKernelGGL(k_partners<D>, dim3(nbs), dim3(WG), 0, st_, geo_, nshots, cpos_.p, crad_.p, cprio_.p,
                         ccell_.p, ctag_.p, cell2t_.p, seg_.p, len_.p, ord_.p, poff_.p, pcnt_.p, plist_.p,
                         (int)plist_.cap, dctr_.p);
      read_ctr();
      if (hctr_[C_OVERFLOW] == 2) {
        plist_.reserve((size_t)hctr_[C_CURSOR] * 2);
        HIP_OK(hipMemsetAsync(dctr_.p + C_OVERFLOW, 0, sizeof(int), st_));
        continue;
      }
      break;
    }
    timed("k_label", (double)nshots * 10.0, [&] {
      hipLaunchKernelGGL(k_label, dim3(1), dim3(1024), 0, st_, nshots, ctag_.p, lab_.p, poff_.p, pcnt_.p, plist_.p,
                         dctr_.p);
    });
    read_ctr();
    if (hctr_[C_LABEL_STUCK]) throw Error("internal: conflict labelling made no progress");
    const uint64_t nb_undecided = (uint64_t)hctr_[C_ACCEPTED];
    bool outreached = false;
    uint64_t total_undecided = 0;
    if (may_outreach) {
      total_undecided = nb_undecided;
      outreached = total_undecided > maximal;
    }
    if (outreached) {  // keep the first `maximal` by priority (operator_algorithm.ixx:416-421)
      HIP_OK(hipMemsetAsync(dctr_.p + C_CURSOR, 0, sizeof(int), st_));
      hipLaunchKernelGGL(k_gather_undecided, dim3(nbs), dim3(WG), 0, st_, nshots, cprio_.p, ctag_.p, undec_.p,
                         dctr_.p);
      read_ctr();
      std::vector<uint32_t> pr(hctr_[C_CURSOR]);
      if (!pr.empty())
        HIP_OK(hipMemcpyAsync(pr.data(), undec_.p, pr.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, st_));
        HIP_OK(hipStreamSynchronize(st_));
      std::vector<uint64_t> prio(pr.begin(), pr.end());
      const double pivot = mom_pivot(prio, (int64_t)maximal);
      int64_t fin = 0;
      for (uint64_t p : prio)
        if (!((double)p > pivot)) fin++;
      hipLaunchKernelGGL(k_retain, dim3(nbs), dim3(WG), 0, st_, nshots, pivot, cprio_.p, ctag_.p);
      hipLaunchKernelGGL(k_update, dim3(nbs), dim3(WG), 0, st_, seg_.p, len_.p, ord_.p, ctag_.p, dctr_.p);
      local_new += fin;
      break;
    }
    HIP_OK(hipMemsetAsync(dctr_.p + C_REMAIN, 0, sizeof(int), st_));
    hipLaunchKernelGGL(k_update, dim3(nbs), dim3(WG), 0, st_, seg_.p, len_.p, ord_.p, ctag_.p, dctr_.p);
    local_new += (int64_t)nb_undecided;
    const uint64_t left = maximal - total_undecided;  // by-value parameter: draw-start value each round
    if (left == 0) {
      hipLaunchKernelGGL(k_list_remove, dim3(nbs), dim3(WG), 0, st_, 1, seg_.p, len_.p, ord_.p, ctag_.p, dctr_.p);
      break;
    }
    read_ctr();
    if (hctr_[C_REMAIN] == 0) break;
  }
  hipLaunchKernelGGL(k_commit<D>, dim3(nbs), dim3(WG), 0, st_, cs(), nshots, cpos_.p, crad_.p, cphase_.p, ctag_.p,
                     touched_.p, seg_.p, len_.p, ord_.p, cell2t_.p, cellcc_.p, dctr_.p);
  read_ctr();
  if (hctr_[C_OVERFLOW] == 1) throw Error("cell capacity exceeded");
  st->ghost_cell_drops += hctr_[C_DROPPED];
  st->image_cell_mismatch += hctr_[C_IMG_MISMATCH];
  return local_new;
}

// proceed<1> -> proceed_voxel (operator_algorithm.ixx:85-188, 208-236)

}  // namespace rsa_amd
