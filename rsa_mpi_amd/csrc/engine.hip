// Single-rank bias-free RSA (MarcJos/RSA_MPI, method 1 = voxel) on one MI355X.
//
// Device state: validated spheres in per-cell slots (cell grid of
// rsa_grid_traversal with its ghost layers kept as index space only: a ghost
// neighbour is read from the real cell it mirrors, filtered by the ghost-area
// band and shifted, which is what the reference's periodic self-exchange puts
// there), the voxel list as int32 SoA lattice coordinates plus one common
// length, the MT19937 word stream in an HBM ring, and per-draw candidate SoA.
//
// One draw (operator_algorithm.ixx:117-173) = decode candidates from the word
// stream, bin them into cells in shot order, run the reference's conflict
// rounds exactly (validated filter, priority sweep labels, tag update with the
// same swap-with-last list edits), commit survivors to their cells in the
// reference's in-cell order; then, when the miss rate exceeds the aimed rate,
// one fused voxel pass computes coverage (remove_covered) and the children
// masks (subdivide_uncovered), followed by the swap-with-last compaction in
// closed form and a stable children scatter.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <condition_variable>
#include <atomic>
#include <map>
#include <mutex>
#include <deque>
#include <thread>
#include <chrono>
#include <cstddef>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <random>
#include <string>
#include <tuple>
#include <vector>

#include "device_common.hpp"
#include "engine.hpp"
#include "learn.hpp"
#include "mlp.hpp"
#include "mt19937.hpp"
#include "pools.hpp"
#include "k_common.hpp"
#include "k_draw.hpp"
#include "k_refine.hpp"
#include "k_learn.hpp"
#include "k_extract.hpp"
#include "k_checks.hpp"
#include "k_batch.hpp"

namespace rsa_amd {

// ============================================================== host engine

static double volume_unit_sphere(int dim) {  // auxiliary/Geometry/Sphere.ixx:31-44
  const double pi = M_PI;
  double mult = 1;
  if (dim % 2 == 0) {
    for (int k = 0; k <= dim / 2 - 1; k++) mult *= (2 * pi) / (dim - 2 * k);
  } else {
    mult = 1. / pi;
    for (int k = 0; k <= (dim - 1) / 2; k++) mult *= (2 * pi) / (dim - 2 * k);
  }
  return mult;
}

template <int D>
static double puissance(double x) {  // AuxiFunctions.ixx:34-44
  if constexpr (D == 0) return 1.;
  if constexpr (D == 1) return x;
  if constexpr (D == 2) return x * x;
  if constexpr (D == 3) return x * (x * x);
  if constexpr (D == 4) return (x * x) * (x * x);
  return 0;
}

template <int D>
static std::vector<std::array<int, D>> tab_corner_host() {  // auxiliary/Path.ixx:85-116
  std::vector<std::array<int, D>> t(ipow_c(2, D));
  if constexpr (D == 2) {
    t = {{{0, 0}}, {{1, 1}}, {{0, 1}}, {{1, 0}}};
  } else if constexpr (D == 3) {
    t = {{{0, 0, 0}}, {{1, 1, 1}}, {{0, 0, 1}}, {{1, 1, 0}}, {{0, 1, 0}}, {{1, 0, 1}}, {{0, 1, 1}}, {{1, 0, 0}}};
  } else {
    // the reference's D >= 4 table never sets component 0 (see Path.ixx:105-115)
    for (int i = 0; i < ipow_c(2, D); i++)
      for (int d = 0; d < D; d++) t[i][d] = d == 0 ? 0 : ((i >> d) & 1);
  }
  return t;
}

template <int D>
class Engine final : public IEngine {
 public:
  // ext: a stream to run on instead of one of the engine's own (the batched form
  // drives many engines on one stream)
  Engine(const double* ginf_, const double* gsup_, int ghost_layer, double domain_rad, hipStream_t ext = nullptr) {
    if (ghost_layer < 1) throw Error("ghost_layer must be >= 1 for a periodic domain");
    gl_ = ghost_layer;
    drad_ = domain_rad;
    double dsize[D];
    for (int d = 0; d < D; d++) {  // rsa_domain.ixx:51-56 on one rank
      ginf_a[d] = ginf_[d];
      gsup_a[d] = gsup_[d];
      dsize[d] = gsup_a[d] - ginf_a[d];
      const double size = dsize[d] / 1;
      inf_[d] = ginf_a[d] + 0 * size;
      sup_[d] = ginf_a[d] + (0 + 1) * size;
      L_[d] = dsize[d];
    }
    // non-blocking: no implicit ordering with the null stream (every copy, fill
    // and launch of the engine is on its own streams, the stencil table of the
    // first define_grid included), so engines run side by side
    own_st_ = ext == nullptr;
    if (own_st_)
      HIP_OK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    else
      st_ = ext;
    HIP_OK(hipStreamCreateWithFlags(&cst_, hipStreamNonBlocking));
    // Cells of more than list_thr_ candidates send a draw's first round to the
    // host-built lists.  RSA_GPU_LIST_THRESHOLD lowers it (1 .. CSLOT) so that tests
    // run that path, which the usual workloads almost never reach.
    if (const char* e = std::getenv("RSA_GPU_LIST_THRESHOLD")) list_thr_ = std::max(1, std::min(CSLOT, std::atoi(e)));
    try {
      define_grid(domain_rad);  // the domain's own grid (rsa_domain.ixx:26); replaced by algo_init
      for (int d = 1; d < D; d++)
        if (sup_[d] - inf_[d] < 2 * domain_rad) throw Error("The domain is not sufficiently large!");
    } catch (...) {  // no destructor runs for a constructor that throws
      (void)hipStreamSynchronize(st_);
      (void)hipStreamDestroy(cst_);
      if (own_st_) (void)hipStreamDestroy(st_);
      throw;
    }
    hctr_ = static_cast<int*>(PinnedPool::get().alloc(sizeof(int) * (C_NUM + 4), hipHostMallocDefault));
    dctr_.reserve(C_NUM);
    reset_ctr();  // pool memory: counters start at zero
    ring_.reserve(kRing);
    hs_.buf = static_cast<uint32_t*>(PinnedPool::get().alloc(sizeof(uint32_t) * kRing, hipHostMallocDefault));
    for (int b = 0; b < kPrepBufs; b++) {
      hvpos_[b] = static_cast<uint32_t*>(PinnedPool::get().alloc((kFastShots + 1) * sizeof(uint32_t),
                                                                 hipHostMallocCoherent | hipHostMallocMapped));
      HIP_OK(hipHostGetDevicePointer((void**)&hvpos_dev_[b], hvpos_[b], 0));
      hrad_[b] = static_cast<double*>(PinnedPool::get().alloc((kFastShots + 1) * sizeof(double),
                                                              hipHostMallocCoherent | hipHostMallocMapped));
      HIP_OK(hipHostGetDevicePointer((void**)&hrad_dev_[b], hrad_[b], 0));
    }
    aimed_ = D <= 3 ? 0.95 : 1 - 0.05 * puissance<(D > 3 ? D - 3 : 0)>(0.1);
  }
  ~Engine() override {
    // work still queued (speculative draws) ends before any buffer is handed on
    if (st_) (void)hipStreamSynchronize(st_);
    if (cst_) (void)hipStreamSynchronize(cst_);
    rs_.arm();  // the device buffers (members) are then released without further waits
    prep_halt();
    hs_.halt();
    for (auto* h : hvpos_)
      PinnedPool::get().release(h);
    for (auto* h : hrad_)
      PinnedPool::get().release(h);
    PinnedPool::get().release(hrad_big_);
    PinnedPool::get().release(hs_.buf);
    PinnedPool::get().release(hpub_);
    for (auto& k : pend_) {
      (void)hipEventDestroy(k.a);
      (void)hipEventDestroy(k.b);
    }
    for (auto e : evpool_) (void)hipEventDestroy(e);
    if (cst_) (void)hipStreamSynchronize(cst_);
    for (auto& c : chunks_) (void)hipEventDestroy(c.ev);
    for (auto e : free_ev_) (void)hipEventDestroy(e);
    if (cst_) (void)hipStreamDestroy(cst_);
    if (st_ && own_st_) (void)hipStreamDestroy(st_);
    PinnedPool::get().release(hctr_);
  }

  int dim() const override { return D; }
  double total_volume() const override {  // rsa_domain.ixx:166-169
    double v = 1.;
    for (int d = 0; d < D; d++) v *= (gsup_a[d] - ginf_a[d]);
    return v;
  }

  // RadiusGenerator (radius_generator.ixx:9-28) then rsa_algo ctor (operator_algorithm.ixx:15-30)
  void algo_init(const std::vector<RadiusClass>& cls, double exclusion, int nbshots) override {
    StreamScope scope(st_, cst_);
    set_classes(cls, exclusion);
    nbshots_ = nbshots;
    r_max_ = gmax_;
    aimed_ = D <= 3 ? 0.95 : 1 - 0.05 * puissance<(D > 3 ? D - 3 : 0)>(0.1);
    define_grid(gmax_);
    if (gmax_ > drad_)
      throw Error("Impossible to have the maximal radius of radius generator larger than domain implicit radius");
    // cells beyond the ghost-band width would need images further than one layer
    for (int d = 0; d < D; d++)
      if (2 * drad_ > cl_[d])
        throw Error("domain radius too large for the cell grid: the ghost band must fit in one cell layer");
    alloc_cells();
    algo_ready_ = true;
  }

  // rsa_algo constructor with a custom-law RadiusGenerator (radius_generator.ixx:38-48)
  void algo_init_law(double rmin, double rmax, rsa_gpu_radius_law law, void* user, uint64_t phase, uint64_t count,
                     double exclusion, int nbshots) override {
    StreamScope scope(st_, cst_);
    (void)exclusion;  // kept by the reference's generator, not added to the law's radii
    set_law(rmin, rmax, law, user, phase, count);
    nbshots_ = nbshots;
    r_max_ = gmax_;
    aimed_ = D <= 3 ? 0.95 : 1 - 0.05 * puissance<(D > 3 ? D - 3 : 0)>(0.1);
    define_grid(gmax_);
    if (gmax_ > drad_)
      throw Error("Impossible to have the maximal radius of radius generator larger than domain implicit radius");
    for (int d = 0; d < D; d++)
      if (2 * drad_ > cl_[d])
        throw Error("domain radius too large for the cell grid: the ghost band must fit in one cell layer");
    alloc_cells();
    algo_ready_ = true;
  }

  void reset_classes(const std::vector<RadiusClass>& cls, double exclusion) override {  // :238-250
    StreamScope scope(st_, cst_);
    if (!algo_ready_) throw Error("rsa_algo not initialised");
    const auto saved = classes_;
    const double old_max = gmax_, old_min = gmin_;
    const int old_cur = cur_;
    set_classes(cls, exclusion);  // the new generator (its constructor may throw)
    if (old_max < gmax_) {
      classes_ = saved;
      gmax_ = old_max;
      gmin_ = old_min;
      cur_ = old_cur;
      throw Error("Incorrect maximal radius");
    }
    r_max_ = std::max(r_max_, gmax_);
    // the cell store keeps the spheres already placed: size it for every radius
    store_rmin_ = std::min(store_rmin_, gmin_);
    grow_cells(cell_cap(store_rmin_, r_max_ > store_rmin_ ? r_max_ : store_rmin_));
  }

  // reset_radius_generator to a RadiusGenerator with a custom law (operator_algorithm.ixx:238-250,
  // radius_generator.ixx:38-48): same checks and kept spheres as reset_classes
  void reset_law(double rmin, double rmax, rsa_gpu_radius_law law, void* user, uint64_t phase, uint64_t count,
                 double exclusion) override {
    StreamScope scope(st_, cst_);
    if (!algo_ready_) throw Error("rsa_algo not initialised");
    (void)exclusion;  // kept by the reference's generator, not added to the law's radii
    if (gmax_ < rmax) throw Error("Incorrect maximal radius");
    set_law(rmin, rmax, law, user, phase, count);  // validates its arguments before changing anything
    r_max_ = std::max(r_max_, gmax_);
    store_rmin_ = std::min(store_rmin_, gmin_);
    grow_cells(cell_cap(store_rmin_, r_max_ > store_rmin_ ? r_max_ : store_rmin_));
  }

  void set_miss_rate(double m) override { aimed_ = m; }

  void add_spheres(int64_t n, const double* pos, const double* rad, const uint64_t* phase) override {
    StreamScope scope(st_, cst_);
    if (!algo_ready_) throw Error("rsa_algo not initialised");
    if (n <= 0) return;
    std::vector<uint8_t> cnt(geo_.ncells);
    HIP_OK(hipMemcpyAsync(cnt.data(), cells_cnt_.p, geo_.ncells, hipMemcpyDeviceToHost, st_));
    HIP_OK(hipStreamSynchronize(st_));
    std::vector<int> cell(n), slot(n);
    for (int64_t i = 0; i < n; i++) {  // build_grid(..., Validated) appends in input order
      const int64_t c = cell_index<D>(geo_, pos + i * D);
      if (c < 0 || c >= geo_.ncells) throw Error("sphere centre outside the cell grid");
      if (cnt[c] >= cap_) throw Error("cell capacity exceeded");
      cell[i] = (int)c;
      slot[i] = cnt[c]++;
    }
    // the inputs in pinned, coherent host memory, read there by k_insert
    const size_t off_rad = (size_t)n * D * sizeof(double), off_cell = off_rad + (size_t)n * sizeof(double);
    const size_t off_slot = off_cell + (size_t)n * sizeof(int), off_ph = off_slot + (size_t)n * sizeof(int);
    const size_t off_last = off_ph + (size_t)n * sizeof(uint32_t), bytes = off_last + (size_t)n;
    char* h = static_cast<char*>(PinnedPool::get().alloc(bytes, hipHostMallocCoherent | hipHostMallocMapped));
    char* hd = nullptr;
    hipError_t e = hipHostGetDevicePointer((void**)&hd, h, 0);
    if (e != hipSuccess) {
      PinnedPool::get().release(h);
      HIP_OK(e);
    }
    std::memcpy(h, pos, (size_t)n * D * sizeof(double));
    std::memcpy(h + off_rad, rad, (size_t)n * sizeof(double));
    std::memcpy(h + off_cell, cell.data(), (size_t)n * sizeof(int));
    std::memcpy(h + off_slot, slot.data(), (size_t)n * sizeof(int));
    for (int64_t i = 0; i < n; i++) {
      reinterpret_cast<uint32_t*>(h + off_ph)[i] = (uint32_t)phase[i];
      reinterpret_cast<uint8_t*>(h + off_last)[i] = slot[i] + 1 == cnt[cell[i]];
    }
    hipLaunchKernelGGL(k_insert<D>, dim3(nblocks(n)), dim3(WG), 0, st_, cs(), n, (const int*)(hd + off_cell),
                       (const int*)(hd + off_slot), (const uint8_t*)(hd + off_last), (const double*)hd,
                       (const double*)(hd + off_rad), (const uint32_t*)(hd + off_ph));
    e = hipStreamSynchronize(st_);
    PinnedPool::get().release(h);
    HIP_OK(e);
    added_ += (uint64_t)n;
  }

  uint64_t count() const override { return nb_fast_; }

  // 64-bit slot after the counters in pinned memory (8-byte aligned)
  uint64_t* hsum() { return reinterpret_cast<uint64_t*>(hctr_ + ((C_NUM + 1) & ~1)); }

  // The placed spheres in extraction order.  Device to host copies go through a
  // pinned staging block (never straight into the caller's pageable memory), and
  // the received values are checked against the set checksum k_extract forms
  // on the device: a transfer that changed any value is an error, not a result.
  uint64_t extract(double* pos, double* rad, uint64_t* phase, uint64_t cap) override {
    StreamScope scope(st_, cst_);
    if (!algo_ready_) return 0;
    const unsigned nb = nblocks(geo_.ncells);
    DevBuf<int> blk;
    blk.reserve(nb);
    hipLaunchKernelGGL(k_cell_counts<D>, dim3(nb), dim3(WG), 0, st_, geo_, cells_cnt_.p, blk.p);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st_, blk.p, (int)nb, dctr_.p + C_CURSOR);
    HIP_OK(hipMemcpyAsync(hctr_ + C_CURSOR, dctr_.p + C_CURSOR, sizeof(int), hipMemcpyDeviceToHost, st_));
    HIP_OK(hipStreamSynchronize(st_));
    const int total = hctr_[C_CURSOR];
    if (pos == nullptr) return (uint64_t)total;
    if ((uint64_t)total > cap) throw Error("extract: output capacity too small");
    if (total == 0) return 0;
    DevBuf<double> dp, dr;
    DevBuf<uint32_t> dph;
    DevBuf<unsigned long long> xs;
    dp.reserve((size_t)total * D);
    dr.reserve(total);
    dph.reserve(total);
    xs.reserve(1);
    HIP_OK(hipMemsetAsync(xs.p, 0, sizeof(unsigned long long), st_));
    hipLaunchKernelGGL((k_extract<D, uint32_t>), dim3(nb), dim3(WG), 0, st_, geo_, cs(), blk.p, dp.p, dr.p,
                       dph.p, xs.p);
    HIP_OK(hipGetLastError());
    const size_t bpos = (size_t)total * D * sizeof(double), brad = (size_t)total * sizeof(double),
                 bph = (size_t)total * sizeof(uint32_t);
    char* stage = static_cast<char*>(PinnedPool::get().alloc(bpos + brad + bph, hipHostMallocDefault));
    struct Give {  // the staging block goes back to the pool on every path
      void* p;
      ~Give() { PinnedPool::get().release(p); }
    } give{stage};
    HIP_OK(hipMemcpyAsync(hsum(), xs.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, st_));
    HIP_OK(hipMemcpyAsync(stage, dp.p, bpos, hipMemcpyDeviceToHost, st_));
    HIP_OK(hipMemcpyAsync(stage + bpos, dr.p, brad, hipMemcpyDeviceToHost, st_));
    if (phase) HIP_OK(hipMemcpyAsync(stage + bpos + brad, dph.p, bph, hipMemcpyDeviceToHost, st_));
    HIP_OK(hipStreamSynchronize(st_));
    const double* sp = reinterpret_cast<const double*>(stage);
    const double* sr = reinterpret_cast<const double*>(stage + bpos);
    uint64_t h = 0;
    for (int i = 0; i < total; i++) {
      for (int d = 0; d < D; d++) h += word_mix(bits_of(sp[(size_t)i * D + d]), d);
      h += word_mix(bits_of(sr[i]), D);
    }
    if (h != *hsum())
      throw Error("extract: the spheres received on the host do not match their device checksum (" +
                  std::to_string(total) + " spheres)");
    std::memcpy(pos, sp, bpos);
    std::memcpy(rad, sr, brad);
    if (phase) {
      const uint32_t* sph = reinterpret_cast<const uint32_t*>(stage + bpos + brad);
      for (int i = 0; i < total; i++) phase[i] = sph[i];
    }
    return (uint64_t)total;
  }

  // extract() into caller-owned device memory with no allocation: the per-block
  // counts go to the static scratch array, the total through the engine's pinned
  // counters (both exist before the call).  pos (cap * D), rad (cap) and phase
  // (cap, may be null) are device pointers; with pos null only the count is returned.
  uint64_t extract_noalloc(double* pos, double* rad, uint64_t* phase, uint64_t cap) override {
    if (!algo_ready_) return 0;
    const int64_t nb = nblocks(geo_.ncells);
    if (nb > noalloc_blocks())
      throw Error("extract_noalloc: the cell grid has " + std::to_string(geo_.ncells) + " cells; the scratch holds " +
                  std::to_string((int64_t)noalloc_blocks() * WG) + " at most");
    int* s = scratch_ptr();
    hipLaunchKernelGGL(k_cell_counts<D>, dim3(nb), dim3(WG), 0, st_, geo_, cells_cnt_.p, s);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st_, s, (int)nb, s + kNoAllocBlocks);
    HIP_OK(hipMemcpyAsync(hctr_ + C_CURSOR, s + kNoAllocBlocks, sizeof(int), hipMemcpyDeviceToHost, st_));
    HIP_OK(hipStreamSynchronize(st_));
    const int total = hctr_[C_CURSOR];
    if (total < 0) throw Error("extract_noalloc: sphere count exceeds 2^31");
    if (pos == nullptr) return (uint64_t)total;
    if ((uint64_t)total > cap) throw Error("extract: output capacity too small");
    if (!rad) throw Error("extract_noalloc: null radius output");
    hipLaunchKernelGGL((k_extract<D, uint64_t>), dim3(nb), dim3(WG), 0, st_, geo_, cs(), s, pos, rad, phase,
                       (unsigned long long*)nullptr);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(st_));
    return (uint64_t)total;
  }
  // the scratch's block limit; RSA_GPU_NOALLOC_BLOCKS lowers it (a test setting, to reach the size check)
  static int64_t noalloc_blocks() {
    const char* e = std::getenv("RSA_GPU_NOALLOC_BLOCKS");
    const int64_t w = e ? std::atoll(e) : kNoAllocBlocks;
    return std::max<int64_t>(1, std::min<int64_t>(kNoAllocBlocks, w));
  }

  double volume_of_spheres() override {  // rsa_grid.ixx:61-71: per real cell, cells summed in index order
    StreamScope scope(st_, cst_);
    const uint64_t n = extract(nullptr, nullptr, nullptr, 0);
    std::vector<double> p(n * D), r(n);
    std::vector<uint64_t> ph(n);
    extract(p.data(), r.data(), ph.data(), n);
    const double mult = volume_unit_sphere(D);
    double res = 0.;
    // the reference sums per cell then adds the cell sums; redo that grouping
    std::vector<uint8_t> cnt(geo_.ncells);
    HIP_OK(hipMemcpyAsync(cnt.data(), cells_cnt_.p, geo_.ncells, hipMemcpyDeviceToHost, st_));
    HIP_OK(hipStreamSynchronize(st_));
    uint64_t k = 0;
    for (int c = 0; c < geo_.ncells; c++) {
      if (!cnt[c]) continue;
      double loc = 0.;
      for (int j = 0; j < cnt[c]; j++, k++) loc += mult * puissance<D>(r[k]);
      res += loc;
    }
    return res;
  }

  uint64_t voxels(int64_t* coords, double* len, uint64_t cap) override {
    StreamScope scope(st_, cst_);
    if (len)
      for (int d = 0; d < D; d++) len[d] = vlen_[d];
    if (coords == nullptr) return (uint64_t)nvox_;
    if ((uint64_t)nvox_ > cap) throw Error("voxels: output capacity too small");
    for (int d = 0; d < D; d++) {
      if (wide_[vcur_]) {
        std::vector<int64_t> tmp(nvox_);
        if (nvox_) {
          HIP_OK(hipMemcpyAsync(tmp.data(), vw_[vcur_][d].p, nvox_ * sizeof(int64_t), hipMemcpyDeviceToHost, st_));
          HIP_OK(hipStreamSynchronize(st_));
        }
        for (int64_t i = 0; i < nvox_; i++) coords[i * D + d] = tmp[i];
      } else {
        std::vector<int32_t> tmp(nvox_);
        if (nvox_) {
          HIP_OK(hipMemcpyAsync(tmp.data(), vc_[vcur_][d].p, nvox_ * sizeof(int32_t), hipMemcpyDeviceToHost, st_));
          HIP_OK(hipStreamSynchronize(st_));
        }
        for (int64_t i = 0; i < nvox_; i++) coords[i * D + d] = tmp[i];
      }
    }
    return (uint64_t)nvox_;
  }

  void proceed(uint64_t seed, int64_t max_draws, rsa_gpu_stats* out) override;

  void set_tuning(const rsa_gpu_tuning& t) override {
    if (t.group_min < 1 || t.pack_min < 1) throw Error("tuning: thresholds must be >= 1");
    group_min_ = t.group_min;
    pack_min_ = t.pack_min;
  }
  rsa_gpu_tuning tuning() const override { return rsa_gpu_tuning{group_min_, pack_min_}; }

  // learned predictor: trace logging, use (ordering and speculation), outcome measurement
  void learn_mode(int trace, int use, int measure) override {
    learn_trace_ = trace != 0;
    learn_use_ = use != 0;
    learn_measure_ = measure != 0;
    if (learn_trace_ || learn_use_) learn::Model::get().ensure();
  }
  void learn_stats(rsa_gpu_learn_stats* out) override { *out = lst_; }

  void profile_enable(int level) override {
    const bool on = level > 0;
    StreamScope scope(st_, cst_);
    resolve_prof();
    prof_ = on;
    prof_level_ = level;
    kprof_.clear();
  }
  bool profile_get(int k, const char** name, int64_t* launches, double* seconds, double* bytes) override {
    {
      StreamScope scope(st_, cst_);
      resolve_prof();
    }
    if (k < 0 || k >= (int)kprof_.size()) return false;
    *name = kprof_[k].name;
    *launches = kprof_[k].launches;
    *seconds = kprof_[k].seconds;
    *bytes = kprof_[k].bytes;
    return true;
  }

 private:
  // ---------------------------------------------------------------- state
  static constexpr uint64_t kRing = 1ull << 23;  // words in the device ring (32 MiB)
  // Declared first, so destroyed last: from arm() (in the destructor, once the
  // streams are idle) until every member is gone, releases do not wait.
  struct QuietRelease {
    RelStreams prev;
    bool armed = false;
    void arm() {
      prev = tl_rel;
      armed = true;
      tl_rel.set = true;
      tl_rel.s[0] = tl_rel.s[1] = nullptr;
    }
    ~QuietRelease() {
      if (armed) tl_rel = prev;
    }
  } rs_;
  double ginf_a[D], gsup_a[D], inf_[D], sup_[D], L_[D], cl_[D];
  double drad_ = 0.05;
  int gl_ = 1;
  Geo<D> geo_{};
  hipStream_t st_ = nullptr;
  hipStream_t cst_ = nullptr;  // copies of MT words into the device ring
  bool own_st_ = true;         // st_ was created by the engine (not the batch driver's)
  template <int DD>
  friend class BatchDriver;
  int* hctr_ = nullptr;
  DevBuf<int> dctr_;
  int list_thr_ = CSLOT;  // see the constructor

  // radius generator (radius_generator.hxx:131-137)
  std::vector<std::tuple<double, uint64_t, int>> classes_;
  double gmax_ = 0, gmin_ = 0;
  int cur_ = 0;
  double r_max_ = 0;
  int nbshots_ = 6000;
  double aimed_ = 0.95;
  bool algo_ready_ = false;

  // cells
  int cap_ = 8;
  DevBuf<uint8_t> cells_cnt_;
  DevBuf<int> cells_mod_;  // CellStore::mod
  DevBuf<Slot<D>> cells_sph_;
  DevBuf<uint32_t> cells_ph_;
  DevBuf<int> cell2t_, cellcc_;
  uint64_t nb_fast_ = 0;
  uint64_t added_ = 0;  // spheres given by add_spheres since the cell store was made

  // voxels
  DevBuf<int32_t> vc_[2][D];
  DevBuf<int64_t> vw_[2][D];  // the same lists once coordinates need 64 bits
  bool wide_[2] = {false, false};
  int vcur_ = 0;
  int64_t nvox_ = 0;
  double vlen_[D] = {}, vorigin_[D] = {};
  std::vector<std::array<int, D>> tab_;

  // candidates
  DevBuf<double> cpos_, crad_;
  DevBuf<uint32_t> cprio_, cphase_, vpos_, undec_;
  DevBuf<int> ccell_, carr_, touched_, seg_, len_, ord_, poff_, pcnt_, plist_;
  DevBuf<uint8_t> ctag_, lab_;

  // refine scratch
  DevBuf<uint32_t> code_, moved_, perm_;
  DevBuf<int> blk_, holes_, ccount_, vcell_, vrank_, tsum_;
  DevBuf<int> coff_;         // packed copy of the cell store: offsets per cell (+ total)
  DevBuf<Slot<D>> csph_;     // and the spheres in cell order (k_cover's listing threshold in place of the radius)
  DevBuf<double> cprad_;     // and their radii

  // word stream: generated on the host into a pinned ring (the MT19937
  // recurrence is sequential), copied asynchronously into the device ring
  struct HostRing {
    MT19937 mt;
    uint32_t* buf = nullptr;
    std::atomic<uint64_t> hi{0};   // words [0, hi) generated; the last kRing of them are in buf
    std::atomic<uint64_t> low{0};  // words below low are no longer read (host or copy engine)
    std::atomic<bool> stop{false};
    std::thread th;
    // The generator runs ahead on its own thread, at most kRing words past
    // `low`, so a slot is rewritten only after its word has been consumed.
    void start(uint64_t seed) {
      halt();
      mt.seed_seq2(0u, (uint32_t)seed);
      hi.store(0);
      low.store(0);
      stop.store(false);
      th = std::thread([this] {
        uint32_t blk[MT19937::N];
        while (!stop.load(std::memory_order_relaxed)) {
          const uint64_t h = hi.load(std::memory_order_relaxed);
          if (h + MT19937::N > low.load(std::memory_order_acquire) + kRing) {
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            continue;
          }
          const uint64_t off = h & (kRing - 1);
          if (off + MT19937::N <= kRing) {
            mt.next_block(buf + off);
          } else {  // the block wraps around the ring's end
            mt.next_block(blk);
            for (int k = 0; k < MT19937::N; k++) buf[(h + k) & (kRing - 1)] = blk[k];
          }
          hi.store(h + MT19937::N, std::memory_order_release);
        }
      });
    }
    void halt() {
      if (th.joinable()) {
        stop.store(true);
        th.join();
      }
    }
    void set_low(uint64_t v) { low.store(v, std::memory_order_release); }
    std::atomic<int64_t> wait_ns{0};  // time consumers (main and helper threads) waited for the generator
    void ensure(uint64_t upto) {
      if (upto > low.load(std::memory_order_relaxed) + kRing) throw Error("word window larger than the ring");
      if (hi.load(std::memory_order_acquire) >= upto) return;
      const auto t0 = std::chrono::steady_clock::now();
      while (hi.load(std::memory_order_acquire) < upto) {
      }
      wait_ns += std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
    }
    uint32_t at(uint64_t pos) {
      if (hi.load(std::memory_order_acquire) <= pos) ensure(pos + 1);
      return buf[pos & (kRing - 1)];
    }
    ~HostRing() { halt(); }
  };
  // URNG over the host ring with std::mt19937's result_type/min/max (see StreamURNG)
  struct RingURNG {
    using result_type = unsigned long;
    HostRing* s;
    uint64_t pos;
    static constexpr result_type min() { return 0; }
    static constexpr result_type max() { return 0xffffffffUL; }
    result_type operator()() { return s->at(pos++); }
  };
  HostRing hs_;
  DevBuf<uint32_t> ring_;
  // Words reach the device ring in chunks copied on cst_ ahead of use; st_
  // waits on a chunk's event once, before the first kernel that reads it.
  static constexpr uint64_t kChunk = 1ull << 18;  // words per copy (1 MiB)
  static constexpr uint64_t kAhead = 1ull << 21;  // copy at most this far past the current need
  struct Chunk {
    uint64_t start, end;
    hipEvent_t ev;
  };
  std::deque<Chunk> chunks_;  // issued, not yet waited on by st_, in word order
  std::deque<hipEvent_t> free_ev_;  // reused oldest first
  uint64_t dev_hi_ = 0;  // words below dev_hi_ are copied or in flight
  std::mutex dw_m_;

  CellStore<D> cs() { return CellStore<D>{cells_cnt_.p, cells_sph_.p, cells_ph_.p, cap_, cells_mod_.p}; }

  // ---------------------------------------------------------------- kernel profile
  struct KProf {
    const char* name;
    int64_t launches = 0;
    double seconds = 0, bytes = 0;
  };
  bool prof_ = false;
  std::vector<KProf> kprof_;
  // launch() runs the kernel launch; with profiling on it is bracketed by a pair
  // of events on the context's stream (the host does not wait for them), and
  // the pair is read in resolve_prof() once the stream has drained
  int prof_level_ = 0;  // 1: refine kernels only, 2: every bracketed kernel
  struct PendingK {
    const char* name;
    double bytes;
    hipEvent_t a, b;
  };
  std::vector<PendingK> pend_;
  std::vector<hipEvent_t> evpool_;
  hipEvent_t take_event() {
    if (evpool_.empty()) {
      hipEvent_t e;
      HIP_OK(hipEventCreate(&e));
      return e;
    }
    const hipEvent_t e = evpool_.back();
    evpool_.pop_back();
    return e;
  }
  template <class F>
  void timed(const char* name, double bytes, F&& launch, int level = 2) {
    if (!prof_ || level > prof_level_) {
      launch();
      return;
    }
    PendingK k{name, bytes, take_event(), take_event()};
    HIP_OK(hipEventRecord(k.a, st_));
    launch();
    HIP_OK(hipEventRecord(k.b, st_));
    pend_.push_back(k);
  }
  void resolve_prof() {
    if (pend_.empty()) return;
    HIP_OK(hipStreamSynchronize(st_));
    for (auto& k : pend_) {
      float ms = 0;
      HIP_OK(hipEventElapsedTime(&ms, k.a, k.b));
      prof_add(k.name, ms * 1e-3, k.bytes);
      evpool_.push_back(k.a);
      evpool_.push_back(k.b);
    }
    pend_.clear();
  }
  void prof_add(const char* name, double seconds, double bytes) {
    KProf* k = nullptr;
    for (auto& e : kprof_)
      if (std::strcmp(e.name, name) == 0) k = &e;
    if (!k) {
      kprof_.push_back(KProf{name});
      k = &kprof_.back();
    }
    k->launches++;
    k->seconds += seconds;
    k->bytes += bytes;
  }
  // host-side time of a step of the draw loop (profiling level >= 1)
  template <class F>
  auto host_timed(const char* name, F&& f) {
    if (!prof_) return f();
    const auto t0 = std::chrono::steady_clock::now();
    struct Acc {
      Engine* e;
      const char* n;
      std::chrono::steady_clock::time_point t0;
      ~Acc() { e->prof_add(n, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), 0); }
    } acc{this, name, t0};
    return f();
  }

  void define_grid(double rad) {  // rsa_grid_traversal.ixx:11-73
    for (int d = 0; d < D; d++) {
      const double length = rad * 2;
      const double width = sup_[d] - inf_[d];
      const int nin = static_cast<int>(width / length);
      if (nin == 0) throw Error("Too small domain wrt ghost layer");
      cl_[d] = width / nin;
      geo_.cinv[d] = 1. / cl_[d];
      geo_.cshift[d] = inf_[d] - gl_ * cl_[d];
      geo_.n[d] = nin + 2 * gl_;
      geo_.inf[d] = inf_[d];
      geo_.sup[d] = sup_[d];
      geo_.lo_band_hi[d] = inf_[d] + 2 * drad_;  // rsa_domain.ixx:117-133
      geo_.hi_band_lo[d] = sup_[d] - 2 * drad_;
      geo_.L[d] = L_[d];
    }
    geo_.gl = gl_;
    int64_t nc = 1;
    for (int d = 0; d < D; d++) nc *= geo_.n[d];
    if (nc > (int64_t)1 << 30) throw Error("cell grid too large");
    geo_.ncells = (int)nc;
    real_cells_ = 1;
    for (int d = 0; d < D; d++) real_cells_ *= geo_.n[d] - 2 * gl_;
    // stencil order (rsa_grid_traversal.ixx:30-66): libstdc++ std::sort by squared norm
    const int n3 = ipow_c(3, D);
    std::vector<int> order(n3);
    std::vector<std::array<double, D>> vec(n3);
    for (int i = 0; i < n3; i++) {
      order[i] = i;
      int tmp = i;
      for (int d = 0; d < D; d++) {
        const int r = tmp % 3;
        vec[i][d] = r - 1;
        tmp = (tmp - r) / 3;
      }
    }
    auto norm2 = [](const std::array<double, D>& v) {
      double s = 0.;
      for (int d = 0; d < D; d++) s += v[d] * v[d];
      return s;
    };
    std::sort(order.begin(), order.end(), [&](int a, int b) { return norm2(vec[a]) < norm2(vec[b]); });
    geo_.nst = n3;
    std::vector<int> packed(n3);
    for (int i = 0; i < n3; i++) {
      int v = 0;
      for (int d = 0; d < D; d++) v |= ((int)vec[order[i]][d] + 1) << (2 * d);
      packed[i] = v;
    }
    // linear index differences: cell index = c0 + n0 * (c1 + n1 * (c2 + ...))
    std::vector<int> lin(n3);
    for (int i = 0; i < n3; i++) {
      int64_t off = 0, stride = 1;
      for (int d = 0; d < D; d++) {
        off += (int64_t)vec[order[i]][d] * stride;
        stride *= geo_.n[d];
      }
      lin[i] = (int)off;
    }
    stab_.reserve(2 * n3);
    StencilTable tab{};
    tab.n = 2 * n3;
    for (int i = 0; i < n3; i++) {
      tab.v[i] = packed[i];
      tab.v[n3 + i] = lin[i];
    }
    hipLaunchKernelGGL(k_set_table, dim3(1), dim3(256), 0, st_, stab_.p, tab);
    HIP_OK(hipGetLastError());
    geo_.stab = stab_.p;
    geo_.slin = stab_.p + n3;
  }

  void set_classes(const std::vector<RadiusClass>& cls, double exclusion) {  // radius_generator.ixx:9-28
    if (cls.empty()) throw Error("Empty radius generator!");
    custom_ = false;
    classes_.clear();
    for (const auto& c : cls) classes_.emplace_back(c.radius, c.count, c.phase);
    std::sort(classes_.begin(), classes_.end(), [](auto& a, auto& b) { return std::get<0>(a) > std::get<0>(b); });
    for (size_t i = 0; i < classes_.size(); i++) std::get<0>(classes_[0]) += 0.5 * exclusion;
    gmax_ = std::get<0>(classes_[0]);
    gmin_ = std::get<0>(classes_[classes_.size() - 1]);
    cur_ = 0;
  }

  // RadiusGenerator(min, max, RandomRadiusGenerator*, n, exclusion) (radius_generator.ixx:38-48): one
  // class (0, n, 0) whose radii come from the law; the exclusion distance is kept but not added
  void set_law(double rmin, double rmax, rsa_gpu_radius_law law, void* user, uint64_t phase, uint64_t count) {
    if (!law) throw Error("null radius law");
    if (!(rmin > 0) || rmax < rmin) throw Error("radius law bounds must satisfy 0 < min_radius <= max_radius");
    classes_.clear();
    classes_.emplace_back(0., count, 0);
    gmax_ = rmax;
    gmin_ = rmin;
    cur_ = 0;
    custom_ = true;
    law_ = law;
    law_user_ = user;
    law_phase_ = (uint32_t)phase;
  }
  ClassDev class_dev() const {
    if (custom_) return ClassDev{0., law_phase_};
    return ClassDev{std::get<0>(classes_[cur_]), (uint32_t)std::get<2>(classes_[cur_])};
  }

  // Spheres per cell for radii in [rmin, rmax].  Disjoint balls of radius rmin
  // centred in the cell lie in the cell grown by rmin per side (volume bound).
  // With equal radii, if the half-cell diagonal is below 2r no two centres
  // share a half-cell, so at most 2^D (a margin keeps rounding of the
  // collision test harmless).
  int cell_cap(double rmin, double rmax) const {
    double vol = 1.;
    for (int d = 0; d < D; d++) vol *= (cl_[d] + 2 * rmin);
    const double vb = volume_unit_sphere(D) * puissance<D>(rmin);
    int cap = (int)std::ceil(vol / vb);
    if (rmax == rmin) {
      double diag2 = 0.;
      for (int d = 0; d < D; d++) diag2 += 0.25 * cl_[d] * cl_[d];
      if (std::sqrt(diag2) < 2 * rmin * (1 - 1e-9)) cap = std::min(cap, ipow_c(2, D));
    }
    cap = std::max(cap, 2);
    if (cap > 255) throw Error("radius ratio too large for the per-cell capacity");
    return cap;
  }
  // grow the per-cell slot arrays to `cap`, keeping every stored sphere
  void grow_cells(int cap) {
    if (cap <= cap_) return;
    HIP_OK(hipStreamSynchronize(st_));
    DevBuf<Slot<D>> sph;
    DevBuf<uint32_t> ph;
    sph.reserve((size_t)geo_.ncells * cap);
    ph.reserve((size_t)geo_.ncells * cap);
    fill_nan(sph.p, (int64_t)geo_.ncells * cap);  // the slots past the copied ones
    // copied by a kernel on the engine's stream (no copy-engine writes that later kernels read)
    const int64_t total = (int64_t)geo_.ncells * cap_;
    hipLaunchKernelGGL(k_grow_cells<D>, dim3(nblocks(total)), dim3(WG), 0, st_, total, cap_, cap, cells_sph_.p,
                       cells_ph_.p, sph.p, ph.p);
    HIP_OK(hipGetLastError());
    std::swap(cells_sph_.p, sph.p);
    std::swap(cells_sph_.cap, sph.cap);
    std::swap(cells_ph_.p, ph.p);
    std::swap(cells_ph_.cap, ph.cap);
    cap_ = cap;
  }
  double store_rmin_ = 0;  // smallest radius any class so far may place

  // sphere slots set to NaN (CellStore::sph past each count)
  void fill_nan(Slot<D>* sph, int64_t slots) {
    static_assert(sizeof(Slot<D>) % sizeof(double) == 0, "slots are whole doubles");
    const int64_t n = slots * (int64_t)(sizeof(Slot<D>) / sizeof(double));
    hipLaunchKernelGGL(k_fill_nan, dim3(nblocks(n)), dim3(WG), 0, st_, reinterpret_cast<double*>(sph), n);
    HIP_OK(hipGetLastError());
  }
  void alloc_cells() {
    store_rmin_ = gmin_;
    cap_ = cell_cap(gmin_, gmax_);
    cells_cnt_.reserve(geo_.ncells);
    cells_sph_.reserve((size_t)geo_.ncells * cap_);
    cells_ph_.reserve((size_t)geo_.ncells * cap_);
    fill_nan(cells_sph_.p, (int64_t)geo_.ncells * cap_);
    cells_mod_.reserve(geo_.ncells);
    HIP_OK(hipMemsetAsync(cells_mod_.p, 0xff, (size_t)geo_.ncells * sizeof(int), st_));
    cell2t_.reserve(geo_.ncells);
    cellcc_.reserve(geo_.ncells);
    ovhead_.reserve(geo_.ncells);
    HIP_OK(hipMemsetAsync(ovhead_.p, 0xff, (size_t)geo_.ncells * sizeof(int), st_));
    cslot_.reserve((size_t)geo_.ncells * CSLOT);
    cellcc1_.reserve(geo_.ncells);
    ovhead1_.reserve(geo_.ncells);
    HIP_OK(hipMemsetAsync(ovhead1_.p, 0xff, (size_t)geo_.ncells * sizeof(int), st_));
    cslot1_.reserve((size_t)geo_.ncells * CSLOT);
    HIP_OK(hipMemsetAsync(cellcc1_.p, 0, (size_t)geo_.ncells * sizeof(int), st_));
    HIP_OK(hipMemsetAsync(cells_cnt_.p, 0, geo_.ncells, st_));
    HIP_OK(hipMemsetAsync(cell2t_.p, 0xff, (size_t)geo_.ncells * sizeof(int), st_));
    HIP_OK(hipMemsetAsync(cellcc_.p, 0, (size_t)geo_.ncells * sizeof(int), st_));
    nb_fast_ = 0;
    added_ = 0;
  }

  uint64_t current_number() const { return std::get<1>(classes_[cur_]); }
  bool still() const { return (uint64_t)cur_ < classes_.size(); }
  void update_placed(uint64_t nb) {  // radius_generator.ixx:62-68
    std::get<1>(classes_[cur_]) = current_number() - nb;
    if (current_number() == 0) cur_++;
  }

  void reset_ctr() { HIP_OK(hipMemsetAsync(dctr_.p, 0, sizeof(int) * C_NUM, st_)); }
  void read_ctr() {
    HIP_OK(hipMemcpyAsync(hctr_, dctr_.p, sizeof(int) * C_NUM, hipMemcpyDeviceToHost, st_));
    HIP_OK(hipStreamSynchronize(st_));
  }
  // after read_ctr: a voxel list written inconsistently with its count is an error
  void check_scatter() const {
    if (hctr_[C_SCATTER_ERR]) throw Error("internal: voxel children disagree with their count (list not written)");
    if (hctr_[C_COMPACT_ERR]) throw Error("internal: more stored spheres than placed ones (packed copy not written)");
  }

  // one chunk [dev_hi_, end) of generated words onto cst_ (split at the ring's end)
  void issue_chunk(uint64_t end) {
    const uint64_t a = dev_hi_;
    while (dev_hi_ < end) {
      const uint64_t off = dev_hi_ & (kRing - 1);
      const uint64_t n = std::min<uint64_t>(end - dev_hi_, kRing - off);
      HIP_OK(hipMemcpyAsync(ring_.p + off, hs_.buf + off, n * sizeof(uint32_t), hipMemcpyHostToDevice, cst_));
      dev_hi_ += n;
    }
    if (dev_hi_ == a) return;
    hipEvent_t ev;
    if (free_ev_.size() > 8) {
      ev = free_ev_.front();
      free_ev_.pop_front();
    } else {
      HIP_OK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    HIP_OK(hipEventRecord(ev, cst_));
    chunks_.push_back({a, dev_hi_, ev});
  }
  // before a new stream of words: no copy in flight, nothing copied
  void reset_dev_ring() {
    HIP_OK(hipStreamSynchronize(cst_));
    for (auto& c : chunks_) free_ev_.push_back(c.ev);
    chunks_.clear();
    dev_hi_ = 0;
  }
  // Make device-ring words [from, upto) readable by the next kernel on st_.
  // Every word below hs_.low has been consumed, so a copy into the slot of word
  // p (which held word p - kRing) is safe once p < low + kRing; words at or
  // above low stay in the host ring until their copy has completed, because the
  // kernel that reads them waits for it and low only moves past them later.
  void dev_words(uint64_t from, uint64_t upto) {
    // the batched form calls it from the helper thread and, for host-driven draws,
    // from the driver's
    std::unique_lock<std::mutex> dlk(dw_m_, std::defer_lock);
    if (prep_words_) dlk.lock();
    if (dev_hi_ < from) {  // words skipped on the device (read on the host only)
      dev_hi_ = from;
    }
    while (dev_hi_ < upto) {
      const uint64_t end = std::min(upto, dev_hi_ + kChunk);
      hs_.ensure(end);
      issue_chunk(end);
    }
    // run ahead with what the producer has already generated
    const uint64_t lim = std::min(upto + kAhead, hs_.low.load(std::memory_order_relaxed) + kRing);
    for (;;) {
      const uint64_t gen = std::min(hs_.hi.load(std::memory_order_acquire), lim);
      if (gen < dev_hi_ + kChunk) break;
      issue_chunk(dev_hi_ + kChunk);
    }
    // wait for the last chunk holding words below upto (cst_ completes in order)
    hipEvent_t wait = nullptr;
    while (!chunks_.empty() && chunks_.front().start < upto) {
      if (wait) free_ev_.push_back(wait);
      wait = chunks_.front().ev;
      chunks_.pop_front();
    }
    // The host waits for the copy (issued ahead, normally done already) instead of
    // queueing a wait on st_: a stream-wait is a barrier packet that stalls its
    // whole hardware queue, and with the streams of several engines sharing the 4
    // hardware queues, two engines' barriers can each sit ahead of the marker the
    // other waits for, so neither queue moves again.
    if (wait) {
      HIP_OK(hipEventSynchronize(wait));
      free_ev_.push_back(wait);
    }
  }

  size_t cand_cap_ = 0;  // candidates every per-candidate buffer holds
  void ensure_cand(int n) {
    if ((size_t)n <= cand_cap_) return;
    HIP_OK(hipStreamSynchronize(st_));  // queued kernels may still use the old buffers
    const size_t c = (size_t)std::max(n, 1024) * 5 / 4;
    cand_cap_ = c;
    cpos_.reserve(c * D);
    crad_.reserve(c);
    cprio_.reserve(c);
    cphase_.reserve(c);
    vpos_.reserve(c);
    undec_.reserve(c);
    ccell_.reserve(c);
    carr_.reserve(c);
    touched_.reserve(c);
    seg_.reserve(c);
    len_.reserve(c);
    ord_.reserve(c * (CSLOT + 1));  // k_lists: CSLOT per owner, then the longer lists
    poff_.reserve(c);
    pcnt_.reserve(c);
    ctag_.reserve(c);
    lab_.reserve(c);
    next_.reserve(c);
    owner_.reserve(c);
    slab_.reserve(c * KMAX);
    urec_.reserve(c);
    cpos1_.reserve(c * D);
    crad1_.reserve(c);
    cprio1_.reserve(c);
    cphase1_.reserve(c);
    ccell1_.reserve(c);
    ctag1_.reserve(c);
    next1_.reserve(c);
    owner1_.reserve(c);
    // kernels of a draw whose gate stays closed read these before returning
    HIP_OK(hipMemsetAsync(owner_.p, 0, c * sizeof(uint8_t), st_));
    HIP_OK(hipMemsetAsync(ccell_.p, 0xff, c * sizeof(int), st_));
    HIP_OK(hipMemsetAsync(ctag_.p, T_DROPPED, c * sizeof(uint8_t), st_));
    HIP_OK(hipMemsetAsync(owner1_.p, 0, c * sizeof(uint8_t), st_));
    HIP_OK(hipMemsetAsync(ccell1_.p, 0xff, c * sizeof(int), st_));
    HIP_OK(hipMemsetAsync(ctag1_.p, T_DROPPED, c * sizeof(uint8_t), st_));
    if (plist_.cap < 8 * c) plist_.reserve(8 * c);
  }

  void init_voxels(double max_diag) {  // list_of_voxels.ixx:11-28, 192-211
    std::array<int, D> nv{};
    tab_ = tab_corner_host<D>();
    int64_t total = 1;
    for (int d = 0; d < D; d++) {
      const double width = sup_[d] - inf_[d];
      nv[d] = 1 + static_cast<int>(std::sqrt(D) * width / max_diag);
      vlen_[d] = width / nv[d];
      vorigin_[d] = inf_[d];
      total *= nv[d];
    }
    if (total > ((int64_t)1 << 31) - 1) throw Error("initial voxel lattice too large");
    nvox_ = total;
    vcur_ = 0;
    wide_[0] = wide_[1] = false;
    for (int d = 0; d < D; d++) vc_[0][d].reserve(total);
    int32_t* c[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int d = 0; d < D; d++) c[d] = vc_[0][d].p;
    hipLaunchKernelGGL(k_lattice<D>, dim3(nblocks(total)), dim3(WG), 0, st_, total, nv[0], D > 1 ? nv[1] : 1,
                       D > 2 ? nv[2] : 1, D > 3 ? nv[3] : 1, c[0], c[1], c[2], c[3]);
    HIP_OK(hipGetLastError());
    max_coord_ = 0;
    for (int d = 0; d < D; d++) max_coord_ = std::max<int64_t>(max_coord_, nv[d]);
  }
  int64_t max_coord_ = 0;

  VoxView<D> vview(int b = -1) const {
    if (b < 0) b = vcur_;
    VoxView<D> v;
    for (int d = 0; d < D; d++) {
      v.c[d] = vc_[b][d].p;
      v.w[d] = vw_[b][d].p;
      v.origin[d] = vorigin_[d];
      v.len[d] = vlen_[d];
    }
    v.wide = wide_[b] ? 1 : 0;
    return v;
  }

  VoxParams<D> vparams(double rmin) const {
    VoxParams<D> vp{};
    for (int d = 0; d < D; d++) {
      vp.origin[d] = vorigin_[d];
      vp.len[d] = vlen_[d];
      vp.nl[d] = 0.5 * vlen_[d];  // new_voxel_lengths (list_of_voxels.ixx:62)
    }
    // the kernels' compile-time table must be the host's, and the farthest-corner
    // test in k_cover needs its rows to form a product set
    unsigned cm[D] = {};
    for (int k = 0; k < VoxParams<D>::NCH; k++)
      for (int d = 0; d < D; d++) {
        if (tab_[k][d] != TabCorner<D>::at(k, d)) throw Error("internal: TabCorner table mismatch");
        cm[d] |= 1u << tab_[k][d];
      }
    int prod = 1;
    for (int d = 0; d < D; d++) {
      if ((cm[d] == 3) != TabCorner<D>::both(d)) throw Error("internal: TabCorner table mismatch");
      prod *= (cm[d] == 3 ? 2 : 1);
    }
    std::vector<int> seen(1 << D, 0);
    int distinct = 0;
    for (int k = 0; k < VoxParams<D>::NCH; k++) {
      int key = 0;
      for (int d = 0; d < D; d++) key |= tab_[k][d] << d;
      if (!seen[key]++) distinct++;
    }
    if (distinct != prod) throw Error("internal: TabCorner rows are not a product set");
    vp.rmin = rmin;
    // X: the largest coordinate magnitude of a voxel corner or a (shifted) sphere centre
    double X = 0.;
    for (int d = 0; d < D; d++)
      X = std::max(X, std::max(std::fabs(inf_[d]), std::fabs(sup_[d])) + std::fabs(L_[d]));
    const double eps = std::ldexp(1.0, -53);
    vp.xs = 64. * std::sqrt((double)D) * eps * X;
    vp.xs2 = 256. * D * (eps * X) * (eps * X);
    vp.h2 = 0.;
    for (int d = 0; d < D; d++) vp.h2 += (0.5 * vp.len[d]) * (0.5 * vp.len[d]);
    for (int d = 0; d < D; d++)
      if (cm[d] != 3) vp.h2 = -1.;
    return vp;
  }

  // exclusive scan of a[0, n) in place; the total (-1 above INT_MAX) to *total
  void scan_excl(int* a, int64_t n, int* total) {
    const unsigned nt = nblocks(n, SCAN_T);
    tsum_.reserve(nt);
    hipLaunchKernelGGL(k_tile_sums, dim3(nt), dim3(1024), 0, st_, a, n, tsum_.p);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st_, tsum_.p, (int)nt, total);
    hipLaunchKernelGGL(k_tile_scan, dim3(nt), dim3(1024), 0, st_, a, n, tsum_.p);
  }
  // Refinement thresholds (rsa_gpu_tuning): lists of at least group_min_ voxels
  // are walked in cell-grouped order, lists of at least pack_min_ read the
  // placed spheres from a packed copy in cell order (D <= 3).  Neither changes
  // any result; the defaults were measured on the 3D cubes of r = 0.002-0.005.
  static constexpr int64_t kGroupMin = int64_t(1) << 20;
  int64_t group_min_ = kGroupMin, pack_min_ = kGroupMin;

  // remove_covered + subdivide_uncovered (operator_algorithm.ixx:190-197)
  void refine(double rmin) {
    if (nvox_ == 0) return;
    const unsigned nb = nblocks(nvox_);
    code_.reserve(nvox_);
    moved_.reserve(nvox_);
    blk_.reserve(nb);
    holes_.reserve(nvox_);
    const VoxParams<D> vp = vparams(rmin);
    VoxView<D> vx = vview();
    // compulsory traffic: lattice coordinates in, code out, every placed sphere and cell count once
    const double cover_bytes = (double)nvox_ * ((wide_[vcur_] ? 8.0 : 4.0) * D + 4.0) +
                               (double)nb_fast_ * sizeof(Slot<D>) + geo_.ncells;
    // long lists are walked in cell-grouped order
    const bool grouped = nvox_ >= group_min_;
    // learned mode: the MLP predicts each voxel's coverage, and the walk order puts the
    // predicted-covered voxels first (within each group, the cell-grouped order)
    const bool predicted = grouped && learn_use_ && learn::Model::get().trained();
    const float density = (float)((double)(nb_fast_ + added_) / (double)std::max<int64_t>(1, real_cells_)),
                miss = (float)last_miss_, level = (float)level_;
    if (predicted) {
      pred_.reserve(nvox_);
      const learn::Model& m = learn::Model::get();
      timed("k_predict", (double)nvox_ * ((wide_[vcur_] ? 8.0 : 4.0) * D + 1.0), [&] {
        const unsigned nbp = nblocks(nvox_);
        if (wide_[vcur_])
          hipLaunchKernelGGL((k_predict<D, true>), dim3(nbp), dim3(WG), 0, st_, geo_, cs(), vx, vp, nvox_, level,
                             density, miss, m.P(), m.Wq(), pred_.p);
        else
          hipLaunchKernelGGL((k_predict<D, false>), dim3(nbp), dim3(WG), 0, st_, geo_, cs(), vx, vp, nvox_, level,
                             density, miss, m.P(), m.Wq(), pred_.p);
      }, 1);
    }
    if (grouped) {
      const int64_t nkeys = (predicted ? 2 : 1) * (int64_t)geo_.ncells;
      perm_.reserve(nvox_);
      vcell_.reserve(nvox_);
      vrank_.reserve(nvox_);
      ccount_.reserve(nkeys);
      HIP_OK(hipMemsetAsync(ccount_.p, 0, sizeof(int) * (size_t)nkeys, st_));
      const uint8_t* pr = predicted ? pred_.p : nullptr;
      timed("k_bin", (double)nvox_ * ((wide_[vcur_] ? 8.0 : 4.0) * D + 8.0), [&] {
        if (wide_[vcur_])
          hipLaunchKernelGGL((k_bin<D, true>), dim3(nblocks(nvox_, 4 * WG)), dim3(WG), 0, st_, geo_, vx, nvox_, vp,
                             ccount_.p, vcell_.p, vrank_.p, pr);
        else
          hipLaunchKernelGGL((k_bin<D, false>), dim3(nblocks(nvox_, 4 * WG)), dim3(WG), 0, st_, geo_, vx, nvox_, vp,
                             ccount_.p, vcell_.p, vrank_.p, pr);
      }, 1);
      scan_excl(ccount_.p, nkeys, dctr_.p + C_GROUPED);
      timed("k_place", (double)nvox_ * 12.0, [&] {
        hipLaunchKernelGGL(k_place, dim3(nblocks(nvox_, 4 * WG)), dim3(WG), 0, st_, nvox_, vcell_.p, vrank_.p,
                           ccount_.p, perm_.p);
      }, 1);
    }
    const uint32_t* perm = grouped ? perm_.p : nullptr;
    // the placed spheres packed in cell order (k_cover's stencil rows, D <= 3),
    // for long lists only: below pack_min_ voxels the copy costs more than it saves
    const int* coff = nullptr;
    const Slot<D>* csph = nullptr;
    const double* crad = nullptr;
    int ncs = 0;
    if (ipow_c(3, D) <= 32 && nvox_ >= pack_min_) {
      const int64_t nc = geo_.ncells;
      coff_.reserve((size_t)nc + 1);
      csph_.reserve((size_t)(nb_fast_ + added_) + 1);
      cprad_.reserve((size_t)(nb_fast_ + added_) + 1);
      timed("k_compact", (double)nc * 9.0 + (double)(nb_fast_ + added_) * 2.0 * sizeof(Slot<D>), [&] {
        hipLaunchKernelGGL(k_cell_cnt, dim3(nblocks(nc)), dim3(WG), 0, st_, nc, cells_cnt_.p, coff_.p);
        scan_excl(coff_.p, nc, coff_.p + nc);
        hipLaunchKernelGGL(k_compact<D>, dim3(nblocks(nc)), dim3(WG), 0, st_, nc, cs(), coff_.p, csph_.p, cprad_.p, vp,
                           (int64_t)(nb_fast_ + added_), dctr_.p + C_COMPACT_ERR);
      }, 1);
      coff = coff_.p;
      csph = csph_.p;
      crad = cprad_.p;
      ncs = (int)std::min<uint64_t>(nb_fast_ + added_, (uint64_t)INT32_MAX);
    }
    timed("k_cover", cover_bytes, [&] {
      if (wide_[vcur_])
        hipLaunchKernelGGL((k_cover<D, true>), dim3(nb), dim3(WG), 0, st_, geo_, cs(), vx, nvox_, vp, perm, coff,
                           csph, crad, ncs, code_.p, blk_.p);
      else
        hipLaunchKernelGGL((k_cover<D, false>), dim3(nb), dim3(WG), 0, st_, geo_, cs(), vx, nvox_, vp, perm, coff,
                           csph, crad, ncs, code_.p, blk_.p);
    }, 1);
    // the streaming passes: one wave per block of WG voxels, four blocks per workgroup
    const unsigned nbw = nblocks(nvox_, 4 * WG);
    if (grouped) hipLaunchKernelGGL(k_block_count, dim3(nbw), dim3(WG), 0, st_, nvox_, code_.p, blk_.p);
    // the predictor's outcome counters, and its traces
    const bool measured = learn_use_ || learn_measure_;
    if (measured) {
      HIP_OK(hipMemsetAsync(dctr_.p + C_L0, 0, 6 * sizeof(int), st_));
      hipLaunchKernelGGL(k_outcomes, dim3(nblocks(nvox_, 16384)), dim3(1024), 0, st_, nvox_, code_.p,
                         predicted ? pred_.p : (const uint8_t*)nullptr, grouped ? perm_.p : (const uint32_t*)nullptr,
                         dctr_.p + C_L0);
    }
    if (learn_trace_) {
      learn::Model& m = learn::Model::get();
      const int64_t stride = std::max<int64_t>(1, nvox_ / kTraceVoxels);
      const int64_t nt = (nvox_ + stride - 1) / stride;
      if (wide_[vcur_])
        hipLaunchKernelGGL((k_trace_voxels<D, true>), dim3(nblocks(nt)), dim3(WG), 0, st_, geo_, cs(), vx, vp, nvox_,
                           stride, code_.p, level, density, miss, m.tx(), m.tt(), m.cursor(), learn::kTraceCap);
      else
        hipLaunchKernelGGL((k_trace_voxels<D, false>), dim3(nblocks(nt)), dim3(WG), 0, st_, geo_, cs(), vx, vp,
                           nvox_, stride, code_.p, level, density, miss, m.tx(), m.tt(), m.cursor(),
                           learn::kTraceCap);
      lst_.traced_voxels += nt;
    }
    scan_excl(blk_.p, nb, dctr_.p + C_KEEP);
    HIP_OK(hipMemsetAsync(dctr_.p + C_MOVED, 0, sizeof(int), st_));
    // the compaction's bytes depend on the number of moved voxels (left by k_holes):
    // filled in below, once the counters are read
    const size_t holes_at = pend_.size();
    timed("k_holes", 0., [&] {
      hipLaunchKernelGGL(k_holes, dim3(nbw), dim3(WG), 0, st_, nvox_, code_.p, blk_.p, dctr_.p, holes_.p);
    }, 1);
    timed("k_move", 0., [&] {
      if (wide_[vcur_])
        hipLaunchKernelGGL((k_move<D, true>), dim3(nbw), dim3(WG), 0, st_, nvox_, code_.p, moved_.p, blk_.p, dctr_.p,
                           holes_.p, vx);
      else
        hipLaunchKernelGGL((k_move<D, false>), dim3(nbw), dim3(WG), 0, st_, nvox_, code_.p, moved_.p, blk_.p, dctr_.p,
                           holes_.p, vx);
    }, 1);
    // children per block of the kept prefix, counted against the device's keep
    // (blocks past it count 0), so that one host round trip returns both totals
    hipLaunchKernelGGL(k_child_count, dim3(nbw), dim3(WG), 0, st_, nvox_, dctr_.p, code_.p, moved_.p, blk_.p);
    scan_excl(blk_.p, nb, dctr_.p + C_NCHILD);
    read_ctr();
    check_scatter();  // the previous refinement's children
    if (measured) {
      lst_.voxels_predicted += hctr_[C_L0];
      lst_.voxels_mispredicted += hctr_[C_L0 + 1];
      lst_.voxels_covered += hctr_[C_L0 + 2];
      lst_.voxels_predicted_covered += hctr_[C_L0 + 3];
      if (predicted) lst_.voxels_covered_predicted += hctr_[C_L0 + 2];
      lst_.waves += hctr_[C_L0 + 4];
      lst_.waves_uniform += hctr_[C_L0 + 5];
    }
    level_++;
    if (pend_.size() == holes_at + 2) {
      // k_holes: codes and block offsets read, one index per hole written; k_move: the
      // same reads, then per moved voxel its coordinates read and written and its code
      const double cw = wide_[vcur_] ? 8.0 : 4.0, m = (double)hctr_[C_MOVED];
      pend_[holes_at].bytes = (double)nvox_ * 4.0 + nb * 4.0 + m * 4.0;
      pend_[holes_at + 1].bytes = (double)nvox_ * 4.0 + nb * 4.0 + m * (2.0 * cw * D + 4.0 + 4.0);
    }
    const int64_t keep = hctr_[C_KEEP];
    if (keep == 0) {  // subdivide_uncovered returns early on an empty list (list_of_voxels.ixx:60)
      nvox_ = 0;
      return;
    }
    const unsigned nb2 = nblocks(keep);
    const int64_t nchild = hctr_[C_NCHILD];
    if (nchild < 0) throw Error("voxel list would exceed 2^31 entries");
    // children coordinates 2c + t: 64-bit lists from the refinement on where they pass INT32_MAX
    if (max_coord_ > ((int64_t)1 << 61)) throw Error("voxel lattice coordinates exceed 62 bits");
    const int nxt = 1 - vcur_;
    wide_[nxt] = wide_[vcur_] || 2 * max_coord_ + 1 > (int64_t)INT32_MAX;
    for (int d = 0; d < D; d++) {
      if (wide_[nxt])
        vw_[nxt][d].reserve(std::max<int64_t>(nchild, 1));
      else
        vc_[nxt][d].reserve(std::max<int64_t>(nchild, 1));
    }
    const VoxView<D> dst = vview(nxt);
    const double cw = wide_[vcur_] ? 8.0 : 4.0, nw = wide_[nxt] ? 8.0 : 4.0;
    timed("k_scatter", (double)keep * (cw * D + 4.0) + (double)nchild * nw * D, [&] {
      auto kern = wide_[vcur_] ? (wide_[nxt] ? k_scatter<D, true, true> : k_scatter<D, true, false>)
                               : (wide_[nxt] ? k_scatter<D, false, true> : k_scatter<D, false, false>);
      hipLaunchKernelGGL(kern, dim3(nb2), dim3(WG), 0, st_, keep, code_.p, moved_.p, blk_.p, vx, dst, nchild,
                         dctr_.p + C_SCATTER_ERR);
    }, 1);
    HIP_OK(hipGetLastError());
    vcur_ = nxt;
    nvox_ = nchild;
    max_coord_ = 2 * max_coord_;
    for (int d = 0; d < D; d++) vlen_[d] = 0.5 * vlen_[d];
  }

  double mom_pivot(std::vector<uint64_t> values, int64_t nb_first);
  // A draw's host-side inputs: its Poisson shot count and the voxel words of its
  // shots (Lemire rejection chain), computed from a stream position.
  struct Prep {
    int64_t nv = -1;         // voxel count the Poisson intensity used (-1: none)
    uint64_t w_start = 0;    // stream position before the Poisson draws
    int64_t zero_iters = 0;  // Poisson draws that gave no shot
    int nshots = 0;
    uint64_t w_shots = 0;    // stream position of the first shot word
    int end_rel = 0;         // words of the shots' positions
    int buf = 0;             // pinned voxel-word buffer holding the chain
    int id = 0, seq = 0;     // draw id (gate value) and first sweep's seq, once queued
    uint64_t miss_base = 0;  // spheres placed before the draw since the miss-rate reference point
    RejList rej{};           // the chain as launch arguments (n < 0: use the pinned offsets)
    bool law = false;        // custom radius law: 2 words per shot follow the priorities
    uint64_t w_end() const { return w_shots + (uint64_t)end_rel + (uint64_t)nshots * (law ? 3 : 1); }
  };
  void law_radii(const Prep& pr, double* out);
  void init_pub();
  void enqueue_round(const Prep& pr, int seq, int gate_id);
  void enqueue_decode(Prep& pr);
  void enqueue_front(Prep& pr, uint64_t miss_base);
  void enqueue_settle(const Prep& pr, int seq, int gate_id, Prep* next);
  DecodeArgs<D> decode_args(const Prep& pr, int forced);
  int64_t finish_draw(const Prep& pr, int64_t& rounds, bool* next_live, bool& gate_open, rsa_gpu_stats* st);
  int draw_id_ = 0;
  int64_t single_draw_slow(int nshots, const double* hrad, uint64_t maximal, bool may_outreach, uint64_t w0,
                           int64_t& rounds,
                           uint64_t& w_next, rsa_gpu_stats* st);
  int host_chain(int nshots, uint64_t w0, uint32_t N, int b, RejList* rej);
  Prep prepare(uint64_t w, int64_t nvox, int b);
  // draw preparation on a helper thread (prep_start / prep_get / prep_halt)
  // kQueueDepth draws queued behind the current one (deeper queues measured no
  // faster: the draw kernels keep the device busy already, and every conflict
  // round or refinement wastes the launches of all queued draws); pinned offset
  // buffers for the current draw, the queued ones and the kPrepAhead the helper
  // prepares, with room to spare
  static constexpr int kPrepAhead = 4, kQueueDepth = 1, kPrepBufs = 12, kFastShots = 48000;
  static_assert(kPrepBufs > 1 + kQueueDepth + kPrepAhead + 1, "offset buffers in use at once");
  void prep_start();
  void prep_halt();
  Prep prep_get(uint64_t w, int64_t nv);
  // asks the helper for the draws from (w, nv) without waiting for them
  void prep_request(uint64_t w, int64_t nv);
  // the draw from (w, nv) if the helper has it ready (else false, and it is requested)
  bool prep_try_get(uint64_t w, int64_t nv, Prep& out);
  // batched form: the helper also makes each prepared draw's words readable on
  // the device (dev_words), so that the driver only launches
  bool prep_words_ = false;
  std::thread pth_;
  std::mutex pm_;
  std::condition_variable pcv_;
  bool pstop_ = false, preq_set_ = false;
  uint64_t pgen_ = 0, preq_w_ = 0, pnext_w_ = 0;  // pnext_w_: where the worker's next draw starts
  int64_t preq_nv_ = -1;
  int pbuf_next_ = 0;
  std::deque<Prep> pready_;
  std::string perr_;
  void wait_pub(int seq);
  DevBuf<DrawCtl> dctl_;
  DrawPub* hpub_ = nullptr;      // pinned, coherent: kPubSlots slots written by k_sweep_commit
  DrawPub* hpub_dev_ = nullptr;  // its device address
  int seq_ = 0;
  // pinned voxel-word offsets of prepared draws (kFastShots + 1 entries each),
  // used round-robin by the helper thread and never freed during a run
  uint32_t* hvpos_[kPrepBufs] = {};
  const uint32_t* hvpos_dev_[kPrepBufs] = {};
  // custom radius law: per-shot radii of prepared draws (pinned, read by k_decode_filter)
  double* hrad_[kPrepBufs] = {};
  const double* hrad_dev_[kPrepBufs] = {};
  double* hrad_big_ = nullptr;  // radii of host-driven draws with more shots
  const double* hrad_big_dev_ = nullptr;
  size_t hrad_big_cap_ = 0;
  bool custom_ = false;
  rsa_gpu_radius_law law_ = nullptr;
  void* law_user_ = nullptr;
  uint32_t law_phase_ = 0;
  DevBuf<int> next_, cslot_, ovhead_, slab_, pool_;
  DevBuf<URec> urec_;
  DevBuf<uint8_t> owner_;
  // What a draw's decode writes, in two sets: a draw uses the set of the parity
  // of its id, so that the next draw's decode runs while this one commits
  // (k_sweep_commit_decode).  Set 0 is the buffers above (also those of the
  // host-driven path and of the batched form).
  DevBuf<double> cpos1_, crad1_;
  DevBuf<uint32_t> cprio1_, cphase1_;
  DevBuf<int> ccell1_, next1_, cellcc1_, cslot1_, ovhead1_;
  DevBuf<uint8_t> ctag1_, owner1_;
  struct CandSet {
    double* cpos;
    double* crad;
    uint32_t* cprio;
    uint32_t* cphase;
    int* ccell;
    uint8_t* ctag;
    uint8_t* owner;
    int* next;
    int* cellcc;
    int* cslot;
    int* ovhead;
  };
  CandSet cset(int id) const {
    if (id & 1) return {cpos1_.p, crad1_.p, cprio1_.p, cphase1_.p, ccell1_.p, ctag1_.p, owner1_.p, next1_.p, cellcc1_.p,
                        cslot1_.p, ovhead1_.p};
    return {cpos_.p, crad_.p, cprio_.p, cphase_.p, ccell_.p, ctag_.p, owner_.p, next_.p, cellcc_.p, cslot_.p,
            ovhead_.p};
  }
  DevBuf<int> stab_;  // packed stencil offsets (Geo::stab), then linear ones (Geo::slin)

  // learned predictor
  static constexpr int64_t kTraceVoxels = 8192;  // voxels traced per refinement pass
  bool learn_trace_ = false, learn_use_ = false, learn_measure_ = false;
  DevBuf<uint8_t> pred_;
  rsa_gpu_learn_stats lst_{};
  int level_ = 0;            // refinements so far in this proceed
  double last_miss_ = 0.;    // the last draw's miss rate
  int since_refine_ = 0;     // draws since the last refinement
  int64_t real_cells_ = 1;
  std::vector<double> miss_hist_;
  std::vector<float> draw_x_;
  std::vector<uint8_t> draw_t_;
  // the draw head's inputs for the coming draw (its refinement is what is predicted):
  // the recent miss rates and their trend, draws since the last refinement, the shot
  // count and Poisson intensity relative to nbshots, the voxel count, the refinements so
  // far and the margin below the aimed miss rate; the last feature marks a draw sample
  void draw_features(float* x, int64_t nshots) const {
    for (int k = 0; k < mlp::NF; k++) x[k] = 0.f;
    const int n = (int)miss_hist_.size();
    for (int k = 0; k < 4; k++) x[k] = k < n ? (float)miss_hist_[n - 1 - k] : 0.f;
    x[4] = x[0] - x[1];
    x[5] = since_refine_ / 64.f;
    x[6] = (float)nshots / (float)std::max(1, nbshots_);
    x[7] = (float)(0.01 * (double)nvox_ / (double)std::max(1, nbshots_));
    x[8] = (float)(std::log2((double)nvox_ + 1.0) / 32.0);
    x[9] = level_ / 32.f;
    x[10] = (float)aimed_ - x[0];
    x[15] = 1.f;
  }
};

}  // namespace rsa_amd

// the host side of a draw: Engine<D> member functions
#include "engine_draw.inl"
// the batched form: many engines driven in lockstep on one stream
#include "batch.inl"

namespace rsa_amd {

template <int D>
void Engine<D>::proceed(uint64_t seed, int64_t max_draws, rsa_gpu_stats* st) {
  if (!algo_ready_) throw Error("rsa_algo not initialised");
  StreamScope scope(st_, cst_);
  const auto t0 = std::chrono::steady_clock::now();
  std::memset(st, 0, sizeof(*st));
  reset_dev_ring();  // no copy may read the host ring once the producer restarts
  hs_.start(seed);
  hs_.wait_ns = 0;
  prep_start();
  struct Halt {
    Engine& e;
    ~Halt() {
      e.prep_halt();
      e.hs_.halt();
    }
  } halt_at_exit{*this};
  Prep cur;               // the next draw's inputs (valid while cur.nv == nvox_ and cur.w_start == w)
  bool cur_live = false;  // its first round is already queued (behind the previous draw)
  // the draws prepared after cur, in stream order; the first `live` of them have
  // their first round queued behind cur (each gated on the one before it)
  std::deque<Prep> ahead;
  int live = 0;
  uint64_t w = 0;
  lst_ = rsa_gpu_learn_stats{};
  level_ = 0;
  last_miss_ = 0.;
  since_refine_ = 0;
  miss_hist_.clear();
  draw_x_.clear();
  draw_t_.clear();
  // the draw head predicts refinements only with a trained model
  const bool draw_head = learn_use_ && learn::Model::get().trained();
  init_voxels(r_max_ + gmin_);
  st->nvox_max = nvox_;
  uint64_t old_nb = 0;
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  double vox_ms = 0;
  for (;;) {
    st->loop_iterations++;
    if (max_draws >= 0 && st->draws >= max_draws) {
      st->end_reason = 3;
      break;
    }
    if (nvox_ == 0) {
      st->end_reason = 1;
      break;
    }
    // every kernel and copy that read ring words below w has completed (the
    // previous draw's results were waited for): the producer may reuse them
    hs_.set_low(w);
    if (cur.nv != nvox_ || cur.w_start != w) {
      if (cur_live) throw Error("internal: a queued draw lost its inputs");
      ahead.clear();  // prepared for another voxel list
      live = 0;
      cur = host_timed("host_prepare", [&] { return prep_get(w, nvox_); });
    }
    st->loop_iterations += cur.zero_iters;  // the iterations whose Poisson draw gave no shot
    const uint64_t nshots = (uint64_t)cur.nshots;
    const uint64_t maximal = current_number();
    const bool may_outreach = nshots >= maximal;
    // the draw head: will this draw trigger a refinement? (its inputs are known before the draw)
    float dx[mlp::NF];
    draw_features(dx, (int64_t)nshots);
    bool predicted_refine = false;
    if (draw_head) predicted_refine = learn::host_logit(learn::Model::get().host_params(), dx, 1) > 0.f;
    int64_t rounds = 0;
    uint64_t nb_new = 0;
    bool fast = false, gate_open = false;
    // whether the draws queued behind cur are still good (no kernel of cur came after them)
    bool spec_ok = true;
    if (may_outreach || nshots > 48000) {
      if (cur_live || live) throw Error("internal: a queued draw needs the host-driven form");
      ahead.clear();
      uint64_t w_next = 0;
      const double* hrad = nullptr;
      if (cur.law) {
        if (cur.nshots <= kFastShots) {
          hrad = hrad_dev_[cur.buf];
        } else {  // radii of a draw too large for the prepared buffers
          if (hrad_big_cap_ < (size_t)cur.nshots) {
            PinnedPool::get().release(hrad_big_);
            hrad_big_cap_ = (size_t)cur.nshots;
            hrad_big_ = static_cast<double*>(PinnedPool::get().alloc(hrad_big_cap_ * sizeof(double),
                                                                     hipHostMallocCoherent | hipHostMallocMapped));
            HIP_OK(hipHostGetDevicePointer((void**)&hrad_big_dev_, hrad_big_, 0));
          }
          law_radii(cur, hrad_big_);
          hrad = hrad_big_dev_;
        }
      }
      nb_new = (uint64_t)single_draw_slow(cur.nshots, hrad, maximal, may_outreach, cur.w_shots, rounds, w_next, st);
    } else {
      fast = true;
      init_pub();
      const auto tq = std::chrono::steady_clock::now();
      if (!cur_live) {
        ensure_cand(cur.nshots);
        enqueue_decode(cur);
        enqueue_front(cur, (uint64_t)(int)nb_fast_ - old_nb);
      }
      // The next draw is queued behind this one (kQueueDepth = 1), gated on it:
      // its decode runs in this draw's sweep-and-commit launch, then come its
      // filter and partner lists; its own sweep-and-commit launch comes with the
      // decode of the draw after it.  Its inputs hold unless this draw triggers a
      // refinement (the device then keeps the gate closed and the queued draw's
      // kernels undo its decode or return at once); it cannot change the class or
      // overshoot its target: the shots of cur and of the queued draw sum below
      // maximal.  With the draw head predicting a refinement, nothing is queued
      // (the gate would stay closed); the device still decides, as always.
      if (predicted_refine) lst_.speculation_skipped++;
      Prep* nx = nullptr;
      if (live < kQueueDepth && !predicted_refine && (max_draws < 0 || st->draws + 2 <= max_draws)) {
        if (ahead.empty()) {
          const uint64_t wn = cur.w_end();
          ahead.push_back(host_timed("host_prepare_next", [&] { return prep_get(wn, nvox_); }));
        }
        Prep& n = ahead.front();
        if (n.nshots <= 48000 && (size_t)n.nshots <= cand_cap_ && nshots + (uint64_t)n.nshots < maximal) nx = &n;
      }
      enqueue_settle(cur, cur.seq, cur.id, nx);
      if (nx != nullptr) {
        enqueue_front(*nx, 0);  // the miss count restarts after each draw
        live = 1;
      }
      if (prof_)
        prof_add("host_enqueue_draw", std::chrono::duration<double>(std::chrono::steady_clock::now() - tq).count(), 0);
      nb_new = (uint64_t)finish_draw(cur, rounds, &spec_ok, gate_open, st);
      if (!spec_ok) live = 0;  // their gates stayed closed: they are queued again, forced
    }
    w = cur.w_end();
    st->draws++;
    st->shots += (int64_t)nshots;
    st->accepted += (int64_t)nb_new;
    st->rounds += rounds;
    st->rounds_max = std::max(st->rounds_max, rounds);
    update_placed(nb_new);
    const bool cont = still();
    nb_fast_ += nb_new;
    const int new_nb = (int)nb_fast_;
    const uint64_t nb_miss = nshots - (uint64_t)(new_nb - old_nb);
    old_nb = (uint64_t)new_nb;
    const double miss_rate = nb_miss / (0.000001 + nshots);
    const bool refine_now = miss_rate > aimed_;
    if (fast && gate_open == refine_now) throw Error("internal: device and host refinement decisions differ");
    if (draw_head) {
      lst_.draws_predicted++;
      lst_.draws_mispredicted += predicted_refine != refine_now;
      lst_.refines += refine_now;
      lst_.refines_predicted += predicted_refine;
    }
    if (learn_trace_) {
      draw_x_.insert(draw_x_.end(), dx, dx + mlp::NF);
      const uint8_t t[4] = {0, 0, (uint8_t)(refine_now ? 1 : 0), 1};
      draw_t_.insert(draw_t_.end(), t, t + 4);
      lst_.traced_draws++;
    }
    miss_hist_.push_back(miss_rate);
    last_miss_ = miss_rate;
    since_refine_ = refine_now ? 0 : since_refine_ + 1;
    if (refine_now) {
      ahead.clear();  // their gates stayed closed, and the voxel list changes
      live = 0;
      HIP_OK(hipEventRecord(e0, st_));
      refine(gmin_);
      HIP_OK(hipEventRecord(e1, st_));
      HIP_OK(hipEventSynchronize(e1));
      float ms = 0;
      HIP_OK(hipEventElapsedTime(&ms, e0, e1));
      vox_ms += ms;
      st->refines++;
      st->nvox_max = std::max<int64_t>(st->nvox_max, nvox_);
    }
    if (!cont) {
      if (live) throw Error("internal: a draw was queued past the last class");
      st->end_reason = 2;
      break;
    }
    if (ahead.empty()) {
      cur = Prep{};  // prepared at the top of the loop
      cur_live = false;
    } else {
      cur = ahead.front();
      ahead.pop_front();
      cur_live = live > 0;
      if (live) live--;
    }
  }
  if (prof_) prof_add("host_wait_mt_words", 1e-9 * (double)hs_.wait_ns.load(), 0);
  if (learn_trace_ && !draw_t_.empty()) {
    HIP_OK(hipStreamSynchronize(st_));  // the voxel traces of this run are in before the draw samples
    learn::Model::get().add_samples((int64_t)draw_t_.size() / 4, draw_x_.data(), draw_t_.data());
  }
  // end-of-run checks on the device (every run): no overlap, and no free point
  // left for a sphere of the smallest radius
  const int64_t nsamp = std::min<int64_t>(int64_t(1) << 22, std::max<int64_t>(int64_t(1) << 16, 4 * (int64_t)nb_fast_));
  HIP_OK(hipMemsetAsync(dctr_.p + C_OVERLAP, 0, 2 * sizeof(int), st_));
  timed("k_check_overlap", (double)geo_.ncells + (double)nb_fast_ * sizeof(Slot<D>), [&] {
    hipLaunchKernelGGL(k_check_overlap<D>, dim3(nblocks(geo_.ncells)), dim3(WG), 0, st_, geo_, cs(), dctr_.p);
  }, 1);
  timed("k_check_cover", (double)nsamp * 8.0 * D, [&] {
    hipLaunchKernelGGL(k_check_cover<D>, dim3(nblocks(nsamp)), dim3(WG), 0, st_, geo_, cs(), nsamp,
                       0x5eed0c0ffee5eedull ^ seed, gmin_, dctr_.p);
  }, 1);
  HIP_OK(hipGetLastError());
  read_ctr();  // the stream is idle after this
  st->overlap_pairs = hctr_[C_OVERLAP] / 2;
  st->cover_samples = nsamp;
  st->cover_free = hctr_[C_UNCOVERED];
  check_scatter();
  resolve_prof();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  st->nvox_final = nvox_;
  st->words = (int64_t)w;
  st->voxel_seconds = vox_ms * 1e-3;
  st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

IEngine* make_engine(int dim, const double* inf, const double* sup, int gl, double rad) {
  switch (dim) {
    case 2:
      return new Engine<2>(inf, sup, gl, rad);
    case 3:
      return new Engine<3>(inf, sup, gl, rad);
    case 4:
      return new Engine<4>(inf, sup, gl, rad);
    default:
      throw Error("dimension must be 2, 3 or 4");
  }
}

}  // namespace rsa_amd
