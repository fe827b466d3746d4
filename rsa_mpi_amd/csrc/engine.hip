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

namespace rsa_amd {

#define HIP_OK(expr)                                                                             \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      throw Error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + #expr);           \
  } while (0)

// Process-wide cache of device allocations.  A released buffer goes back to
// the cache once the device is idle (hipFree would wait for that too) and is
// handed to a later request of at least half its size; hipMalloc runs only
// when nothing fits (the whole cache is freed and the call retried when the
// device is out of memory).  Engines created one after another thus reuse the
// same HBM instead of mapping and unmapping it for every run.
// The streams on which a buffer released by this thread may still be in use:
// an engine names its own for the duration of its calls (StreamScope), so a
// release waits for that engine's work only and engines on other threads run
// on.  Without a scope a release waits for the whole device.
struct RelStreams {
  bool set = false;
  hipStream_t s[2] = {nullptr, nullptr};  // both null with set: nothing in flight
};
inline thread_local RelStreams tl_rel;

class StreamScope {
 public:
  StreamScope(hipStream_t a, hipStream_t b) : prev_(tl_rel) {
    tl_rel.set = true;
    tl_rel.s[0] = a;
    tl_rel.s[1] = b;
  }
  ~StreamScope() { tl_rel = prev_; }
  StreamScope(const StreamScope&) = delete;
  StreamScope& operator=(const StreamScope&) = delete;

 private:
  RelStreams prev_;
};

class DevicePool {
 public:
  static DevicePool& get() {
    static DevicePool* pool = new DevicePool();  // never destroyed: outlives every engine
    return *pool;
  }
  void* alloc(size_t bytes, size_t& got) {
    void* p = nullptr;
    {
      std::lock_guard<std::mutex> lk(m_);
      auto it = free_.lower_bound(bytes);
      if (it != free_.end() && it->first / 2 <= bytes) {
        got = it->first;
        p = it->second;
        free_.erase(it);
      }
    }
    if (p) {
      poison(p, got);
      return p;
    }
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      trim();
      e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess) throw Error(std::string("HIP error ") + hipGetErrorString(e) + " at hipMalloc");
    got = bytes;
    poison(p, got);
    return p;
  }
  // RSA_GPU_POISON=<byte>: every buffer handed out is first filled with that
  // byte, so that a kernel reading memory nobody wrote in this engine reads the
  // same garbage whatever ran there before (a test setting).
  static int poison_byte() {
    static const int v = [] {
      const char* e = std::getenv("RSA_GPU_POISON");
      return e ? (std::atoi(e) & 0xff) : -1;
    }();
    return v;
  }
  static void poison(void* p, size_t bytes) {
    const int b = poison_byte();
    if (b < 0) return;
    HIP_OK(hipMemsetAsync(p, b, bytes, nullptr));
    HIP_OK(hipStreamSynchronize(nullptr));
  }
  void release(void* p, size_t bytes) {
    if (!tl_rel.set) {
      (void)hipDeviceSynchronize();  // kernels queued earlier may still use it
    } else {
      for (hipStream_t st : tl_rel.s)
        if (st) (void)hipStreamSynchronize(st);
    }
    std::lock_guard<std::mutex> lk(m_);
    free_.emplace(bytes, p);
  }
  void trim() {
    (void)hipDeviceSynchronize();
    std::lock_guard<std::mutex> lk(m_);
    for (auto& kv : free_) (void)hipFree(kv.second);
    free_.clear();
  }

 private:
  std::mutex m_;
  std::multimap<size_t, void*> free_;  // by size in bytes
};

// Process-wide cache of pinned host allocations, by flags and size.  A new
// block is written once when allocated: the first CPU pass over fresh pinned
// pages is several times slower than later ones, and the MT ring is filled by
// a thread that must keep ahead of the draws.
class PinnedPool {
 public:
  static PinnedPool& get() {
    static PinnedPool* pool = new PinnedPool();  // never destroyed: outlives every engine
    return *pool;
  }
  void* alloc(size_t bytes, unsigned flags) {
    {
      std::lock_guard<std::mutex> lk(m_);
      for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->flags == flags && it->bytes >= bytes && it->bytes / 2 <= bytes) {
          void* p = it->p;
          used_[p] = *it;
          free_.erase(it);
          return p;
        }
      }
    }
    void* p = nullptr;
    HIP_OK(hipHostMalloc(&p, bytes, flags));
    std::memset(p, 0, bytes);
    std::lock_guard<std::mutex> lk(m_);
    used_[p] = Block{p, bytes, flags};
    return p;
  }
  void release(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(m_);
    auto it = used_.find(p);
    if (it == used_.end()) return;
    free_.push_back(it->second);
    used_.erase(it);
  }

 private:
  struct Block {
    void* p;
    size_t bytes;
    unsigned flags;
  };
  std::mutex m_;
  std::vector<Block> free_;
  std::map<void*, Block> used_;
};

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;  // elements
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void reserve(size_t n) {  // contents are not preserved; grows by at least 1.5x
    if (n <= cap && p) return;
    const size_t want = std::max<size_t>({n, cap + cap / 2, 1});
    release();
    size_t got = 0;
    p = static_cast<T*>(DevicePool::get().alloc(want * sizeof(T), got));
    cap = got / sizeof(T);
  }
  void release() {
    if (p) DevicePool::get().release(p, cap * sizeof(T));
    p = nullptr;
    cap = 0;
  }
};

constexpr int WG = 256;
static inline unsigned nblocks(int64_t n, int per = WG) { return (unsigned)std::max<int64_t>(1, (n + per - 1) / per); }

// Counters written by kernels and read by the host after a draw step.
enum Ctr : int {
  C_NTOUCHED = 0,
  C_CURSOR,       // partner-list allocation cursor
  C_ACCEPTED,     // Undecided after a sweep (how_many<Real, Undecided>)
  C_CONFLICT,     // Conflict after a sweep
  C_REMAIN,       // Not_Validated after the tag update
  C_OVERFLOW,     // cell capacity exceeded
  C_CHAIN_OVF,    // word window too short for the shots
  C_LABEL_STUCK,  // label passes made no progress
  C_DROPPED,      // centres in ghost cells
  C_IMG_MISMATCH, // periodic image whose computed cell is not the mirrored ghost cell
  C_END_REL,      // word offset after the last shot's coordinates
  C_KEEP,         // uncovered voxels
  C_NCHILD,       // children kept by subdivide
  C_GROUPED,      // voxels binned by cell (k_cover's processing order)
  C_SCATTER_ERR,  // k_scatter found a workgroup's children inconsistent with the counting pass
  C_COMPACT_ERR,  // k_compact found more stored spheres than the packed copy holds
  C_MOVED,        // voxels k_move moved into holes (the compaction's traffic)
  C_OVERLAP,      // end-of-run check: colliding (sphere, neighbour) pairs, each pair seen from both sides
  C_UNCOVERED,    // end-of-run check: sample points a sphere of the smallest radius could still take
  C_L0,           // k_outcomes: predicted, mispredicted, covered, predicted covered, waves, uniform waves
  C_L5 = C_L0 + 5,
  C_NUM
};

// ============================================================== kernels: word stream

// A load issued where it is written: relaxed atomic loads are not moved past
// the branches that follow them (the compiler sinks plain loads into the block
// that uses them, behind the test that guards it).  Workgroup scope: an
// ordinary global load, for data written by earlier kernels.
template <class T>
__device__ __forceinline__ T load_now(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The packed stencil entries of a lane's CPL cells (lane `sub` takes entries
// sub * CPL ..), read at kernel start: they depend on the lane only.
template <int D, int CPL, int NST>
__device__ __forceinline__ void lane_stencil(const Geo<D>& g, int sub, int (&pk)[CPL]) {
#pragma unroll
  for (int jj = 0; jj < CPL; jj++) {
    const int s = sub * CPL + jj;
    pk[jj] = load_now(g.stab + (s < NST ? s : 0));
  }
}

// Kernels never read device memory that a copy engine wrote through a plain
// load: an XCD's L2 may still hold an older line of it (the ring's slots are
// rewritten every kRing words, and pool memory changes hands).  Ring words are
// read with system-scope loads, which see what the copy wrote; the other
// host-provided tables are written by kernels (k_set_table, k_insert).
__device__ __forceinline__ uint32_t ring_at(const uint32_t* ring, uint64_t mask, uint64_t pos) {
  return __hip_atomic_load(ring + (pos & mask), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Voxel-index words drawn by uniform_int_distribution<int>(0, N-1) use Lemire's
// multiply-shift with rejection of low 32 bits below (2^32 - N) mod N
// (uniform_int_dist.h:246-268, 313-317).  One workgroup builds the rejection
// bitmap of the word window in LDS; wave 0 then walks the shots: a shot takes
// the first accepted word at or after its start, then 2*D coordinate words.
template <int D>
__global__ void __launch_bounds__(1024) k_chain(const uint32_t* __restrict__ ring, uint64_t mask, uint64_t w0,
                                                int nwin, int nshots, uint32_t N, uint32_t thr,
                                                uint32_t* __restrict__ vpos, int* __restrict__ ctr) {
  extern __shared__ uint32_t bits[];  // nwin/32 + 1 words
  const int nb = (nwin + 31) / 32;
  for (int w = threadIdx.x; w < nb; w += blockDim.x) {
    uint32_t b = 0;
    for (int k = 0; k < 32; k++) {
      const int pos = w * 32 + k;
      if (pos < nwin) {
        const uint64_t prod = (uint64_t)ring_at(ring, mask, w0 + pos) * (uint64_t)N;
        if ((uint32_t)prod < thr) b |= (1u << k);
      }
    }
    bits[w] = b;
  }
  __syncthreads();
  if (threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  const int stride = 2 * D + 1;
  auto rej = [&](int p) -> bool { return (bits[p >> 5] >> (p & 31)) & 1u; };
  int cur = 0, base = 0;
  bool ovf = false;
  while (base < nshots) {
    const int i = base + lane;
    const int s = cur + lane * stride;
    const bool valid = i < nshots;
    const bool out = valid && (s + stride > nwin);
    const bool r = valid && !out && rej(s);
    const uint64_t bo = __ballot(out);
    const uint64_t br = __ballot(r);
    if (bo) {  // window too short: let the host retry with a larger one
      ovf = true;
      break;
    }
    if (br == 0) {
      if (valid) vpos[i] = (uint32_t)s;
      const int cnt = min(64, nshots - base);
      cur += cnt * stride;
      base += cnt;
    } else {
      const int f = __ffsll((unsigned long long)br) - 1;
      if (lane < f) vpos[i] = (uint32_t)s;
      int p = cur + f * stride;
      while (p < nwin && rej(p)) p++;
      if (p + stride > nwin) {
        ovf = true;
        break;
      }
      if (lane == f) vpos[i] = (uint32_t)p;
      cur = p + stride;
      base += f + 1;
    }
  }
  if (lane == 0) {
    if (ovf)
      ctr[C_CHAIN_OVF] = 1;
    else
      ctr[C_END_REL] = cur;
  }
}

template <int D>
struct VoxView {
  // lattice coordinates, SoA: 32-bit while every coordinate fits, 64-bit after
  // the refinement that would overflow them (the reference keeps int64)
  const int32_t* c[D];
  const int64_t* w[D];
  int wide;
  double origin[D], len[D];
  __device__ __forceinline__ int64_t coord(int d, int64_t i) const { return wide ? w[d][i] : (int64_t)c[d][i]; }
  __device__ __forceinline__ void set(int d, int64_t i, int64_t v) const {
    if (wide)
      const_cast<int64_t*>(w[d])[i] = v;
    else
      const_cast<int32_t*>(c[d])[i] = (int32_t)v;
  }
};

struct ClassDev {
  double radius;
  uint32_t phase;
};

// pick_points + generate_priority + constant-radius class (list_of_voxels.ixx:107-129,
// operator_algorithm.ixx:373-381, radius_generator.ixx:50-60), then the cell of
// each centre (build_grid, rsa_grid.ixx:182-187).
template <int D>
__global__ void k_decode(const uint32_t* __restrict__ ring, uint64_t mask, uint64_t w0, int nshots, uint32_t N,
                         const uint32_t* __restrict__ vpos, const double* __restrict__ hrad, VoxView<D> vx, ClassDev cls,
                         Geo<D> g, double* __restrict__ cpos, double* __restrict__ crad,
                         uint32_t* __restrict__ cprio, uint32_t* __restrict__ cphase, int* __restrict__ ccell,
                         uint8_t* __restrict__ ctag, int* ctr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nshots) return;
  const uint32_t v = vpos[i];
  const uint32_t gw = ring_at(ring, mask, w0 + v);
  const uint32_t id = (uint32_t)(((uint64_t)gw * (uint64_t)N) >> 32);
  double p[D];
  for (int d = 0; d < D; d++) {
    const double o = vx.origin[d] + (double)vx.coord(d, id) * vx.len[d];
    const double u = canonical(ring_at(ring, mask, w0 + v + 1 + 2 * d), ring_at(ring, mask, w0 + v + 2 + 2 * d));
    p[d] = o + vx.len[d] * u;
    cpos[(size_t)d * nshots + i] = p[d];
  }
  const int end_rel = ctr[C_END_REL];
  cprio[i] = ring_at(ring, mask, w0 + (uint64_t)end_rel + i) >> 1;
  crad[i] = hrad ? hrad[i] : cls.radius;
  cphase[i] = cls.phase;
  const int64_t cell = cell_index<D>(g, p);
  bool ok = cell >= 0 && cell < g.ncells;
  if (ok) {
    int cc[D];
    cell_coords<D>(g, cell, cc);
    ok = !coords_ghost<D>(g, cc);
  }
  if (ok) {
    ccell[i] = (int)cell;
    ctag[i] = T_UNDECIDED;
  } else {
    ccell[i] = -1;
    ctag[i] = T_DROPPED;
    atomicAdd(&ctr[C_DROPPED], 1);
  }
}

// ============================================================== kernels: binning in shot order

__global__ void k_bin_count(int nshots, const int* __restrict__ ccell, int* __restrict__ cellcc,
                            int* __restrict__ carr, int* __restrict__ cell2t, int* __restrict__ touched,
                            int* __restrict__ ctr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nshots) return;
  const int c = ccell[i];
  if (c < 0) return;
  const int a = atomicAdd(&cellcc[c], 1);
  carr[i] = a;
  if (a == 0) {
    const int t = atomicAdd(&ctr[C_NTOUCHED], 1);
    touched[t] = c;
    cell2t[c] = t;
  }
}

// single workgroup: exclusive scan of the touched cells' candidate counts
__global__ void __launch_bounds__(1024) k_bin_scan(const int* __restrict__ touched, const int* __restrict__ cellcc,
                                                   int* __restrict__ seg, int* __restrict__ len,
                                                   const int* __restrict__ ctr) {
  __shared__ int sh[1024];
  const int nt = ctr[C_NTOUCHED];
  int carry = 0;
  for (int base = 0; base < nt; base += 1024) {
    const int t = base + threadIdx.x;
    const int v = t < nt ? cellcc[touched[t]] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int add = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += add;
      __syncthreads();
    }
    if (t < nt) {
      seg[t] = carry + sh[threadIdx.x] - v;
      len[t] = v;
    }
    const int tot = sh[1023];
    __syncthreads();
    carry += tot;
  }
}

__global__ void k_bin_place(int nshots, const int* __restrict__ ccell, const int* __restrict__ carr,
                            const int* __restrict__ cell2t, const int* __restrict__ seg, int* __restrict__ ord) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nshots) return;
  const int c = ccell[i];
  if (c < 0) return;
  ord[seg[cell2t[c]] + carr[i]] = i;
}

// build_grid appends in candidate order: sort each cell's arrivals by shot index
__global__ void k_bin_sort(const int* __restrict__ seg, const int* __restrict__ len, int* __restrict__ ord,
                           const int* __restrict__ ctr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ctr[C_NTOUCHED]) return;
  int* a = ord + seg[t];
  const int n = len[t];
  for (int k = 1; k < n; k++) {
    const int v = a[k];
    int j = k - 1;
    while (j >= 0 && a[j] > v) {
      a[j + 1] = a[j];
      j--;
    }
    a[j + 1] = v;
  }
}

// ============================================================== kernels: conflict rounds

template <int D>
__device__ __forceinline__ void cand_pos(const double* cpos, int n, int i, double* p) {
  for (int d = 0; d < D; d++) p[d] = cpos[(size_t)d * n + i];
}

// remove_collisions_with_validated_spheres (operator_cells_and_data_storage.ixx:54-76):
// an Undecided candidate colliding with any Validated sphere of its 3^D stencil
// (placed spheres in slots, spheres validated earlier in this draw, and the
// periodic images of both) becomes Skip.
template <int D>
__global__ void k_filter(Geo<D> g, CellStore<D> cs, int nshots, const double* __restrict__ cpos,
                         const double* __restrict__ crad, const int* __restrict__ ccell, uint8_t* __restrict__ ctag,
                         const int* __restrict__ cell2t, const int* __restrict__ seg, const int* __restrict__ len,
                         const int* __restrict__ ord) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nshots) return;
  if (ctag[i] != T_UNDECIDED) return;
  double p[D];
  cand_pos<D>(cpos, nshots, i, p);
  const double r = crad[i];
  int cc[D];
  cell_coords<D>(g, ccell[i], cc);
  for (int s = 0; s < g.nst; s++) {
    Nb<D> nb;
    if (!neighbour<D>(g, cc, s, nb)) continue;
    const int nv = cs.cnt[nb.src];
    const Slot<D>* sl = cs.sph + (size_t)nb.src * cs.cap;
    for (int k = 0; k < nv; k++) {
      const Slot<D> q = sl[k];
      double qp[D];
      if (nb.ghost) {
        if (!image_of<D>(g, nb, q.c, qp)) continue;
      } else {
        for (int d = 0; d < D; d++) qp[d] = q.c[d];
      }
      if (collide<D>(r, p, q.r, qp)) {
        ctag[i] = T_SKIP;
        return;
      }
    }
    const int t = cell2t[nb.src];
    if (t < 0) continue;
    const int* o = ord + seg[t];
    const int ln = len[t];
    for (int k = 0; k < ln; k++) {
      const int j = o[k];
      if (ctag[j] != T_VALIDATED) continue;
      double jp[D], qp[D];
      cand_pos<D>(cpos, nshots, j, jp);
      if (nb.ghost) {
        if (!image_of<D>(g, nb, jp, qp)) continue;
      } else {
        for (int d = 0; d < D; d++) qp[d] = jp[d];
      }
      if (collide<D>(r, p, crad[j], qp)) {
        ctag[i] = T_SKIP;
        return;
      }
    }
  }
}

// remove_tagged_spheres on one cell's candidate tail: from the back, swap with
// last (rsa_data_storage.ixx:108-116, 305-315).  mode 0: Skip, 1: Not_Validated.
__global__ void k_list_remove(int mode, const int* __restrict__ seg, int* __restrict__ len, int* __restrict__ ord,
                              const uint8_t* __restrict__ ctag, const int* __restrict__ ctr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ctr[C_NTOUCHED]) return;
  int* a = ord + seg[t];
  int n = len[t];
  for (int k = n - 1; k >= 0; k--) {
    const uint8_t tg = ctag[a[k]];
    const bool rm = mode == 0 ? (tg == T_SKIP) : (tg != T_VALIDATED);
    if (rm) {
      a[k] = a[n - 1];
      n--;
    }
  }
  len[t] = n;
}

// Ordered partner lists for the sweep (find_solve_conflicts,
// operator_cells_and_data_storage.ixx:127-142 with detection :16-52): for each
// Undecided candidate, every Undecided candidate with a strictly smaller
// priority that collides with it, in stencil order then in-cell order; a
// partner seen through a ghost neighbour is its periodic image (bit 31 set).
template <int D>
__device__ __forceinline__ int scan_partners(const Geo<D>& g, int nshots, int i, const double* cpos, const double* crad,
                                    const uint32_t* cprio, const int* ccell, const uint8_t* ctag,
                                    const int* cell2t, const int* seg, const int* len, const int* ord, int* out,
                                    int* mismatch) {
  double p[D];
  cand_pos<D>(cpos, nshots, i, p);
  const double r = crad[i];
  const uint32_t pi = cprio[i];
  int cc[D];
  cell_coords<D>(g, ccell[i], cc);
  int n = 0;
  for (int s = 0; s < g.nst; s++) {
    Nb<D> nb;
    if (!neighbour<D>(g, cc, s, nb)) continue;
    const int t = cell2t[nb.src];
    if (t < 0) continue;
    const int* o = ord + seg[t];
    const int ln = len[t];
    for (int k = 0; k < ln; k++) {
      const int j = o[k];
      if (ctag[j] != T_UNDECIDED) continue;
      if (!(cprio[j] < pi)) continue;
      double jp[D], qp[D];
      cand_pos<D>(cpos, nshots, j, jp);
      if (nb.ghost) {
        if (!image_of<D>(g, nb, jp, qp)) continue;
        // the reference files the image under compute_cell_idx(image); it must be this ghost cell
        if (out == nullptr) {
          int64_t ic = cell_index<D>(g, qp);
          int64_t want = 0;
          for (int d = D - 1; d >= 0; d--) want = want * g.n[d] + (cc[d] + stencil_off<D>(g.stab[s], d));
          if (ic != want) atomicAdd(mismatch, 1);
        }
      } else {
        for (int d = 0; d < D; d++) qp[d] = jp[d];
      }
      if (collide<D>(r, p, crad[j], qp)) {
        if (out) out[n] = j | (nb.ghost ? (int)0x80000000 : 0);
        n++;
      }
    }
  }
  return n;
}

template <int D>
__global__ void k_partners(Geo<D> g, int nshots, const double* __restrict__ cpos, const double* __restrict__ crad,
                           const uint32_t* __restrict__ cprio, const int* __restrict__ ccell,
                           const uint8_t* __restrict__ ctag, const int* __restrict__ cell2t,
                           const int* __restrict__ seg, const int* __restrict__ len, const int* __restrict__ ord,
                           int* __restrict__ poff, int* __restrict__ pcnt, int* __restrict__ plist, int pcap,
                           int* __restrict__ ctr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nshots) return;
  if (ctag[i] != T_UNDECIDED) {
    pcnt[i] = 0;
    return;
  }
  const int n = scan_partners<D>(g, nshots, i, cpos, crad, cprio, ccell, ctag, cell2t, seg, len, ord, nullptr,
                                 &ctr[C_IMG_MISMATCH]);
  pcnt[i] = n;
  if (n == 0) return;
  const int off = atomicAdd(&ctr[C_CURSOR], n);
  poff[i] = off;
  if (off + n > pcap) {
    atomicMax(&ctr[C_OVERFLOW], 2);
    return;
  }
  scan_partners<D>(g, nshots, i, cpos, crad, cprio, ccell, ctag, cell2t, seg, len, ord, plist + off,
                   &ctr[C_IMG_MISMATCH]);
}

// The sequential priority sweep as a dependency evaluation: a candidate's
// label is fixed by its first partner (in detection order) that is not Skip:
// an image or a Conflict partner gives Conflict, an Undecided real partner
// gives Skip, none leaves it Undecided.  Partners have strictly smaller
// priorities, so they are labelled first in the reference's sorted sweep and
// the passes below reach the same labels.  One workgroup.
__global__ void __launch_bounds__(1024) k_label(int nshots, uint8_t* __restrict__ ctag, uint8_t* __restrict__ lab,
                                                const int* __restrict__ poff, const int* __restrict__ pcnt,
                                                const int* __restrict__ plist, int* __restrict__ ctr) {
  __shared__ int s_prog, s_left;
  for (int i = threadIdx.x; i < nshots; i += blockDim.x) lab[i] = ctag[i] == T_UNDECIDED ? T_UNRESOLVED : ctag[i];
  __syncthreads();
  for (int pass = 0; pass < nshots + 2; pass++) {
    if (threadIdx.x == 0) {
      s_prog = 0;
      s_left = 0;
    }
    __syncthreads();
    int prog = 0, left = 0;
    for (int i = threadIdx.x; i < nshots; i += blockDim.x) {
      if (__atomic_load_n(&lab[i], __ATOMIC_RELAXED) != T_UNRESOLVED) continue;
      const int n = pcnt[i];
      const int* pl = plist + (n ? poff[i] : 0);
      uint8_t res = T_UNDECIDED;
      bool wait = false;
      for (int k = 0; k < n; k++) {
        const int e = pl[k];
        if (e < 0) {
          res = T_CONFLICT;
          break;
        }
        const uint8_t lj = __atomic_load_n(&lab[e], __ATOMIC_RELAXED);
        if (lj == T_UNRESOLVED) {
          wait = true;
          break;
        }
        if (lj == T_SKIP) continue;
        res = (lj == T_CONFLICT) ? T_CONFLICT : T_SKIP;
        break;
      }
      if (wait) {
        left++;
      } else {
        __atomic_store_n(&lab[i], res, __ATOMIC_RELAXED);
        prog++;
      }
    }
    if (prog) atomicAdd(&s_prog, prog);
    if (left) atomicAdd(&s_left, left);
    __syncthreads();
    const int sp = s_prog, sl = s_left;
    __syncthreads();
    if (sl == 0) break;
    if (sp == 0) {
      if (threadIdx.x == 0) ctr[C_LABEL_STUCK] = 1;
      break;
    }
  }
  __syncthreads();
  int acc = 0, con = 0;
  for (int i = threadIdx.x; i < nshots; i += blockDim.x) {
    if (ctag[i] != T_UNDECIDED) continue;
    const uint8_t l = lab[i];
    ctag[i] = l;
    acc += l == T_UNDECIDED;
    con += l == T_CONFLICT;
  }
  if (acc) atomicAdd(&ctr[C_ACCEPTED], acc);
  if (con) atomicAdd(&ctr[C_CONFLICT], con);
}

// retain_only_first_elements (operator_algorithm.ixx:475-519) after the host
// found the pivot: Undecided above it and every other non-validated candidate
// become Skip.
__global__ void k_retain(int nshots, double pivot, const uint32_t* __restrict__ cprio, uint8_t* __restrict__ ctag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nshots) return;
  const uint8_t t = ctag[i];
  if (t == T_DROPPED || t == T_SKIP || t == T_VALIDATED) return;
  if (t == T_UNDECIDED) {
    if ((double)cprio[i] > pivot) ctag[i] = T_SKIP;
  } else {
    ctag[i] = T_SKIP;
  }
}

__global__ void k_gather_undecided(int nshots, const uint32_t* __restrict__ cprio, const uint8_t* __restrict__ ctag,
                                   uint32_t* __restrict__ out, int* __restrict__ ctr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nshots) return;
  if (ctag[i] != T_UNDECIDED) return;
  out[atomicAdd(&ctr[C_CURSOR], 1)] = cprio[i];
}

// update_sphere_tags on real cells (operator_algorithm.ixx:446-473): Skip
// removed from the tail (swap with last, from the back), Undecided ->
// Validated, Conflict -> Undecided.
__global__ void k_update(const int* __restrict__ seg, int* __restrict__ len, int* __restrict__ ord,
                         uint8_t* __restrict__ ctag, int* __restrict__ ctr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ctr[C_NTOUCHED]) return;
  int* a = ord + seg[t];
  int n = len[t];
  for (int k = n - 1; k >= 0; k--) {
    if (ctag[a[k]] == T_SKIP) {
      a[k] = a[n - 1];
      n--;
    }
  }
  len[t] = n;
  int rem = 0;
  for (int k = 0; k < n; k++) {
    const int j = a[k];
    if (ctag[j] == T_UNDECIDED)
      ctag[j] = T_VALIDATED;
    else if (ctag[j] == T_CONFLICT) {
      ctag[j] = T_UNDECIDED;
      rem++;
    }
  }
  if (rem) atomicAdd(&ctr[C_REMAIN], rem);
}

// End of draw: the tail of each touched cell holds the spheres validated by
// this draw in the reference's in-cell order; append them after the cell's
// earlier spheres and clear the per-draw cell maps.
template <int D>
__global__ void k_commit(CellStore<D> cs, int nshots, const double* __restrict__ cpos, const double* __restrict__ crad,
                         const uint32_t* __restrict__ cphase, const uint8_t* __restrict__ ctag,
                         const int* __restrict__ touched, const int* __restrict__ seg, const int* __restrict__ len,
                         const int* __restrict__ ord, int* __restrict__ cell2t, int* __restrict__ cellcc,
                         int* __restrict__ ctr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ctr[C_NTOUCHED]) return;
  const int c = touched[t];
  const int* a = ord + seg[t];
  const int n = len[t];
  int k0 = cs.cnt[c];
  for (int k = 0; k < n; k++) {
    const int j = a[k];
    if (ctag[j] != T_VALIDATED) continue;
    if (k0 >= cs.cap) {
      atomicMax(&ctr[C_OVERFLOW], 1);
      break;
    }
    Slot<D> s;
    for (int d = 0; d < D; d++) s.c[d] = cpos[(size_t)d * nshots + j];
    s.r = crad[j];
    cs.sph[(size_t)c * cs.cap + k0] = s;
    cs.phase[(size_t)c * cs.cap + k0] = cphase[j];
    k0++;
  }
  cs.cnt[c] = (uint8_t)k0;
  cell2t[c] = -1;
  cellcc[c] = 0;
}

// ============================================================== kernels: one draw, one host round trip
//
// A draw is four kernels: decode + validated filter + cell binning
// (k_decode_filter), partner lists (k_partner_lists, which builds each
// neighbour cell's ordered list from its arrival slots in the first round),
// sweep labels (k_sweep_labels, one workgroup), tag update + commit + the
// round's publication to the host (k_update_commit, which keeps the cell lists
// for later rounds).  k_lists builds the lists instead when a cell got more
// than CSLOT candidates (the host redoes the round).  Gathers over the 3^D
// stencil are spread over a lane group per candidate (one lane per stencil
// cell) so the loads of all neighbour cells are in flight together.  No kernel
// uses a device-wide fence or a counter shared by all lanes: touched cells are
// identified by their first candidate ("owner"), and partner lists live in a
// fixed slab per candidate.

template <int D>
struct LaneGroup {
  static constexpr int NST = ipow_c(3, D);
  static constexpr int LPC = D == 2 ? 16 : (D == 3 ? 32 : 64);  // lanes per candidate
  // Stencil cells per lane.  Lane `sub` takes cells sub * CPL .. sub * CPL + CPL - 1,
  // so lane order is stencil order: a partner list assembled lane by lane is in
  // the reference's detection order (4D: 81 cells on 64 lanes, two per lane).
  static constexpr int CPL = (NST + LPC - 1) / LPC;
  __device__ static inline uint64_t mask(int lane) {
    if constexpr (LPC == 64) {
      return ~0ull;
    } else {
      return ((1ull << LPC) - 1ull) << ((lane / LPC) * LPC);
    }
  }
};

constexpr int CSLOT = 8;  // candidates of a cell kept in its fixed slots; more go to an overflow list
constexpr int KMAX = 16;  // partners kept in a candidate's slab; longer lists go to a pool
constexpr int PKEEP = 4;  // partners a lane keeps in registers between counting and writing
constexpr int KREJ = 96;  // rejected voxel words passed with the launch; more: offsets read from host memory

// The Lemire rejection chain of a draw as launch arguments: shot i's voxel word
// sits at i * (2D + 1) + #{k : idx[k] <= i} (idx = shot index of each rejected
// word, non-decreasing; entries n .. KREJ-1 hold INT_MAX).  n < 0: read the
// offsets from the pinned array.
struct RejList {
  int n;
  int idx[KREJ];
};

// Control block of the fast draw path (device side).
struct DrawCtl {
  int dropped;
  int img_mismatch;
  int seg_cursor;
  int pool_cursor;
  int pool_overflow;
  int need_pool;
  int acc, con;
  int label_stuck;
  int cell_overflow;
  // Draw gate.  The first-round kernels of draw `id` run only while gate == id
  // (a forced launch of k_decode_filter sets it).  A draw's last sweep opens
  // the gate for draw id + 1 when the draw is over and its miss rate asks for
  // no refinement; otherwise it closes it (-1).  The host may thus queue the
  // next draw before it has seen this one's results: when the gate stays
  // closed, every kernel of that draw returns at once.
  int gate;
  int swept;     // seq of the last sweep that ran (k_update_commit of that round runs)
  int nunres;    // candidates k_partner_lists left unresolved (one URec each)
  int acc_draw;  // spheres accepted so far in the current draw
  int long_list; // a cell got more than CSLOT candidates this draw (set by k_decode_filter)
  int pubv[10];  // the round's publication (DrawPub fields), left by k_sweep_labels for k_update_commit
};
// One unresolved candidate as k_sweep_labels reads it, written by
// k_partner_lists at the candidate's place in the unresolved list: index,
// partner count, list offset (-1: the candidate's slab of KMAX entries, -2: the
// pool overflowed) and its first PKEEP partners (the list itself holds the rest).
struct alignas(32) URec {
  int q, m, o, pad;
  int e[PKEEP];
};
static_assert(sizeof(URec) == 32, "two 16-byte loads per record");

// A round's results in pinned host memory (seq last): computed by
// k_sweep_labels, written to the host by the first thread of k_update_commit,
// so that the PCIe writes overlap the tag update instead of ending the sweep.
struct DrawPub {
  int acc, con, label_stuck, dropped, img_mismatch, pool_overflow, need_pool, cell_overflow, gate_open;
  int need_lists;  // the first round found a cell of more than CSLOT candidates: build the lists, redo it
  int seq;
};
constexpr int kPubFields = 10;
static_assert(sizeof(((DrawCtl*)nullptr)->pubv) == kPubFields * sizeof(int), "DrawCtl::pubv holds a DrawPub");
// Publications go to slot seq % kPubSlots: a queued next draw may publish
// before the host has read the current one.
constexpr int kPubSlots = 8;

// pick_points + generate_priority + constant-radius class (list_of_voxels.ixx:107-129,
// operator_algorithm.ixx:373-381, radius_generator.ixx:50-60); build_grid's
// cell (rsa_grid.ixx:182-187); round-0 remove_collisions_with_validated_spheres
// (operator_cells_and_data_storage.ixx:54-76: at this point only placed spheres
// and their images are validated); each candidate takes an arrival slot in its
// cell (shot order is restored in k_lists).
template <int D, bool WIDE>
__global__ void __launch_bounds__(WG) k_decode_filter(const uint32_t* __restrict__ ring, uint64_t mask, uint64_t w0,
                                                      int nshots, uint32_t N, const uint32_t* __restrict__ vpos,
                                                      const double* __restrict__ hrad, RejList rl,
                                                      int end_rel, VoxView<D> vx, ClassDev cls, Geo<D> g,
                                                      CellStore<D> cs, double* __restrict__ cpos,
                                                      double* __restrict__ crad, uint32_t* __restrict__ cprio,
                                                      uint32_t* __restrict__ cphase, int* __restrict__ ccell,
                                                      uint8_t* __restrict__ ctag, uint8_t* __restrict__ owner,
                                                      int* __restrict__ cellcc, int* __restrict__ cslot,
                                                      int* __restrict__ ovhead, int* __restrict__ next,
                                                      DrawCtl* __restrict__ ctl, int draw_id, int forced,
                                                      int list_thr) {
  using LG = LaneGroup<D>;
  // The gate does not change during this kernel.  It is read first but tested
  // only before the first store, so that its load overlaps the reads below
  // (all of which stay in bounds whatever the gate says).
  const int gate = forced ? draw_id : ctl->gate;
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int i = gtid / LG::LPC;
  const int sub = gtid % LG::LPC;
  const bool valid = i < nshots;
  double p[D];
  int64_t cell = -1;
  bool ok = false;
  uint32_t prio = 0;
  // radius: the class radius, or the shot's value of a custom law (host-evaluated)
  double r = cls.radius;
  if (valid) {
    uint32_t v;
    if (rl.n >= 0) {
      // rejected words before shot i, the list read 16 entries (one scalar load) at a time
      int k = 0;
#pragma unroll
      for (int c0 = 0; c0 < KREJ; c0 += 16) {
        if (c0 >= rl.n) break;
#pragma unroll
        for (int t = 0; t < 16; t++) k += rl.idx[c0 + t] <= i ? 1 : 0;
      }
      v = (uint32_t)i * (uint32_t)(2 * D + 1) + (uint32_t)k;
    } else {
      v = vpos[i];
    }
    if (hrad) r = hrad[i];
    // every ring word of the shot is requested before any of them is used
    const uint32_t pw = ring_at(ring, mask, w0 + (uint64_t)end_rel + i);
    const uint32_t gw = ring_at(ring, mask, w0 + v);
    uint32_t uw[2 * D];
#pragma unroll
    for (int q = 0; q < 2 * D; q++) uw[q] = ring_at(ring, mask, w0 + v + 1 + q);
    prio = pw >> 1;
    const uint32_t id = (uint32_t)(((uint64_t)gw * (uint64_t)N) >> 32);
    // the voxel's coordinates (the launch picks the instance for the lattice's
    // width, so the D loads are issued together)
    int64_t vc[D];
    if constexpr (WIDE) {
#pragma unroll
      for (int d = 0; d < D; d++) vc[d] = vx.w[d][id];
    } else {
#pragma unroll
      for (int d = 0; d < D; d++) vc[d] = (int64_t)vx.c[d][id];
    }
#pragma unroll
    for (int d = 0; d < D; d++) {
      const double o = vx.origin[d] + (double)vc[d] * vx.len[d];
      const double u = canonical(uw[2 * d], uw[2 * d + 1]);
      p[d] = o + vx.len[d] * u;
    }
    cell = cell_index<D>(g, p);
    ok = cell >= 0 && cell < g.ncells;
    if (ok) {
      int cc[D];
      cell_coords<D>(g, cell, cc);
      ok = !coords_ghost<D>(g, cc);
    }
  }
  bool hit = false;
  if (valid && ok) {
    int cc[D];
    cell_coords<D>(g, cell, cc);
#pragma unroll
    for (int jj = 0; jj < LG::CPL; jj++) {
      const int s = sub * LG::CPL + jj;
      if (s >= LG::NST || hit) continue;
      Nb<D> nb;
      if (!neighbour<D>(g, cc, s, nb)) continue;
      // slots are read only for occupied cells: most neighbour cells of a late
      // draw are empty, and a slot is a cache line of its own
      const int nv = cs.cnt[nb.src];
      const Slot<D>* sl = cs.sph + (size_t)nb.src * cs.cap;
      for (int k = 0; k < nv && !hit; k++) {
        const Slot<D> q = sl[k];
        double qp[D];
        if (nb.ghost) {
          if (!image_of<D>(g, nb, q.c, qp)) continue;
        } else {
#pragma unroll
          for (int d = 0; d < D; d++) qp[d] = q.c[d];
        }
        if (collide<D>(r, p, q.r, qp)) hit = true;
      }
    }
  }
  const bool anyhit = (__ballot(hit) & LG::mask(lane)) != 0;
  if (gate != draw_id) return;  // whole grid
  if (gtid == 0) {
    ctl->acc_draw = 0;
    if (forced) ctl->gate = draw_id;
  }
  if (valid && sub == 0) {
#pragma unroll
    for (int d = 0; d < D; d++) cpos[(size_t)d * nshots + i] = p[d];
    cprio[i] = prio;
    crad[i] = r;
    cphase[i] = cls.phase;
    if (!ok) {
      ccell[i] = -1;
      ctag[i] = T_DROPPED;
      owner[i] = 0;
      atomicAdd(&ctl->dropped, 1);
    } else {
      ccell[i] = (int)cell;
      ctag[i] = anyhit ? T_SKIP : T_UNDECIDED;
      const int a = atomicAdd(&cellcc[cell], 1);
      if (a < CSLOT)
        cslot[(size_t)cell * CSLOT + a] = i;
      else
        next[i] = atomicExch(&ovhead[cell], i);
      // more than list_thr (<= CSLOT) candidates in a cell: the first round's
      // kernels give up and the host builds the lists (k_lists) and redoes it
      if (a == list_thr) atomicOr(&ctl->long_list, 1);
      owner[i] = a == 0;
    }
  }
}

// The list of a cell of n <= CSLOT candidates as k_lists builds it: the arrival
// slots v (n entries) in shot order (a sorting network), then the round-0 Skips
// removed from the back with swap-with-last (rsa_data_storage.ixx:108-116,
// 305-315).  The list is left in v; returns its length.  self: any valid
// candidate index (the tag loads of unused entries read it, without a branch).
// The first n arrival slots in shot order (entries n.. become INT_MAX):
// 19-comparator network for 8 keys (Batcher odd-even merge).
__device__ __forceinline__ void sort_slots(int n, int (&v)[CSLOT]) {
#pragma unroll
  for (int k = 0; k < CSLOT; k++)
    if (k >= n) v[k] = INT_MAX;
  constexpr int NET[19][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6}, {5, 7}, {1, 2}, {5, 6},
                              {0, 4}, {3, 7}, {1, 5}, {2, 6}, {1, 4}, {3, 6}, {2, 4}, {3, 5}, {3, 4}};
#pragma unroll
  for (int e = 0; e < 19; e++) {
    const int x = v[NET[e][0]], y = v[NET[e][1]];
    v[NET[e][0]] = min(x, y);
    v[NET[e][1]] = max(x, y);
  }
}

__device__ __forceinline__ int short_list(int n, int (&v)[CSLOT], const uint8_t* __restrict__ ctag, int self) {
  sort_slots(n, v);
  uint8_t tg[CSLOT];
#pragma unroll
  for (int k = 0; k < CSLOT; k++) tg[k] = ctag[k < n ? v[k] : self];
  int m = n;
#pragma unroll
  for (int q = CSLOT - 1; q >= 0; q--) {
    if (q < n && tg[q] == T_SKIP) {
      int x = v[q];
#pragma unroll
      for (int k = q + 1; k < CSLOT; k++) x = (m - 1 == k) ? v[k] : x;
      v[q] = x;
      m--;
    }
  }
  return m;
}

// One thread per touched cell (its owner candidate): the cell's candidates in
// shot order (build_grid appends in candidate order,
// operator_cells_and_data_storage.ixx:153-159), then the round-0 Skip removal
// from the back with swap-with-last (rsa_data_storage.ixx:108-116, 305-315).
__global__ void __launch_bounds__(WG) k_lists(int nshots, const int* __restrict__ ccell,
                                              const uint8_t* __restrict__ owner, const uint8_t* __restrict__ ctag,
                                              int* __restrict__ cellcc, const int* __restrict__ cslot,
                                              int* __restrict__ ovhead, const int* __restrict__ next,
                                              int* __restrict__ seg, int* __restrict__ len, int* __restrict__ ord,
                                              int* __restrict__ cell2t, DrawCtl* __restrict__ ctl, int draw_id,
                                              int ncells) {
  const int gate = ctl->gate;  // tested before the first store (loads overlap)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = i < nshots;
  // Every thread reads, without a branch (index clamped to candidate 0), its
  // owner flag, cell, the cell's count and fixed slots, so that these loads are
  // issued together.  With a closed gate k_decode_filter did not run, and
  // owner/ccell hold stale or never-written values: the clamped cell keeps these
  // reads in bounds.
  const int ic = in ? i : 0;
  const uint8_t own = load_now(owner + ic);
  const bool on = in && own;
  const int c0 = load_now(ccell + ic);
  const int c = (c0 >= 0 && c0 < ncells) ? c0 : 0;
  const int n = load_now(cellcc + c);
  const int64_t* row = reinterpret_cast<const int64_t*>(cslot + (size_t)c * CSLOT);
  const int64_t q0 = load_now(row), q1 = load_now(row + 1), q2 = load_now(row + 2), q3 = load_now(row + 3);
  const int4 r0 = make_int4((int)q0, (int)(q0 >> 32), (int)q1, (int)(q1 >> 32));
  const int4 r1 = make_int4((int)q2, (int)(q2 >> 32), (int)q3, (int)(q3 >> 32));
  // draw_id < 0: a forced launch (the host's redo of a first round), no gate test
  if ((draw_id >= 0 && gate != draw_id) || !on) return;
  // the owner's list: CSLOT entries at i * CSLOT, longer lists after all of those
  const int base = n <= CSLOT ? i * CSLOT : nshots * CSLOT + atomicAdd(&ctl->seg_cursor, n);
  int* a = ord + base;
  int m = n;
  if (n <= CSLOT) {  // in registers
    int v[CSLOT] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
    m = short_list(n, v, ctag, i);
#pragma unroll
    for (int k = 0; k < CSLOT; k++)
      if (k < m) a[k] = v[k];
  } else {
    for (int k = 0; k < CSLOT; k++) a[k] = cslot[(size_t)c * CSLOT + k];
    int k = CSLOT;
    for (int j = ovhead[c]; j != -1 && k < n; j = next[j]) a[k++] = j;
    // shot order (Shell sort: such long lists occur only in clustered late draws)
    for (int gap = n / 2; gap > 0; gap /= 2)
      for (int q = gap; q < n; q++) {
        const int v = a[q];
        int j = q;
        while (j >= gap && a[j - gap] > v) {
          a[j] = a[j - gap];
          j -= gap;
        }
        a[j] = v;
      }
    for (int q = m - 1; q >= 0; q--) {
      if (ctag[a[q]] == T_SKIP) {
        a[q] = a[m - 1];
        m--;
      }
    }
  }
  seg[i] = base;
  len[i] = m;
  cell2t[c] = i;
  cellcc[c] = 0;
  ovhead[c] = -1;
}

// Rounds >= 1 filter (operator_cells_and_data_storage.ixx:54-76): placed
// spheres, spheres validated by earlier rounds of this draw, and their images.
template <int D>
__global__ void __launch_bounds__(WG) k_refilter(Geo<D> g, CellStore<D> cs, int nshots,
                                                 const double* __restrict__ cpos, const double* __restrict__ crad,
                                                 const int* __restrict__ ccell, uint8_t* __restrict__ ctag,
                                                 const int* __restrict__ cell2t, const int* __restrict__ seg,
                                                 const int* __restrict__ len, const int* __restrict__ ord) {
  using LG = LaneGroup<D>;
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int i = gtid / LG::LPC;
  const int sub = gtid % LG::LPC;
  int pk[LG::CPL];
  lane_stencil<D, LG::CPL, LG::NST>(g, sub, pk);
  const bool active = i < nshots && ctag[i] == T_UNDECIDED;
  bool hit = false;
  if (active) {
    double p[D];
    cand_pos<D>(cpos, nshots, i, p);
    const double r = crad[i];
    int cc[D];
    cell_coords<D>(g, ccell[i], cc);
#pragma unroll
    for (int jj = 0; jj < LG::CPL; jj++) {
      const int s = sub * LG::CPL + jj;
      if (s >= LG::NST || hit) continue;
      Nb<D> nb;
      if (!neighbour_at<D>(g, cc, pk[jj], nb)) continue;
      // slots are read only for occupied cells: most neighbour cells of a late
      // draw are empty, and a slot is a cache line of its own
      const int nv = cs.cnt[nb.src];
      const Slot<D>* sl = cs.sph + (size_t)nb.src * cs.cap;
      for (int k = 0; k < nv && !hit; k++) {
        const Slot<D> q = sl[k];
        double qp[D];
        if (nb.ghost) {
          if (!image_of<D>(g, nb, q.c, qp)) continue;
        } else {
#pragma unroll
          for (int d = 0; d < D; d++) qp[d] = q.c[d];
        }
        if (collide<D>(r, p, q.r, qp)) hit = true;
      }
      const int t = cell2t[nb.src];
      if (t < 0 || hit) continue;
      const int* o = ord + seg[t];
      const int ln = len[t];
      for (int k = 0; k < ln && !hit; k++) {
        const int j = o[k];
        if (ctag[j] != T_VALIDATED) continue;
        double jp[D], qp[D];
        cand_pos<D>(cpos, nshots, j, jp);
        if (nb.ghost) {
          if (!image_of<D>(g, nb, jp, qp)) continue;
        } else {
#pragma unroll
          for (int d = 0; d < D; d++) qp[d] = jp[d];
        }
        if (collide<D>(r, p, crad[j], qp)) hit = true;
      }
    }
  }
  const bool anyhit = (__ballot(hit) & LG::mask(lane)) != 0;
  if (active && sub == 0 && anyhit) ctag[i] = T_SKIP;
}

__global__ void k_remove_skips(int nshots, const uint8_t* __restrict__ owner, const int* __restrict__ seg,
                               int* __restrict__ len, int* __restrict__ ord, const uint8_t* __restrict__ ctag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nshots || !owner[i]) return;
  int* a = ord + seg[i];
  int n = len[i];
  for (int q = n - 1; q >= 0; q--) {
    if (ctag[a[q]] == T_SKIP) {
      a[q] = a[n - 1];
      n--;
    }
  }
  len[i] = n;
}

// Partners of one candidate held by one lane's stencil cells, in stencil then
// in-cell order (find_solve_conflicts' detection, operator_cells_and_data_storage.ixx:16-52,
// 127-142): Undecided candidates of strictly smaller priority that collide;
// through a ghost neighbour the partner is its periodic image (bit 31 set).
template <int D>
__device__ __forceinline__ int lane_partners(const Geo<D>& g, const int* cc, int sub, int nshots, const double* p,
                                             double r, uint32_t pi, const double* cpos, const double* crad,
                                             const uint32_t* cprio, const uint8_t* ctag, const int* cell2t,
                                             const int* seg, const int* len, const int* ord, const int* pk,
                                             const int* cellcc, const int* cslot, int self, int* out,
                                             bool& first_ghost, int* mismatch, int (*keep)[PKEEP] = nullptr) {
  using LG = LaneGroup<D>;
  int n = 0;
  first_ghost = false;
#pragma unroll
  for (int jj = 0; jj < LG::CPL; jj++) {
    const int s = sub * LG::CPL + jj;
    if (s >= LG::NST) continue;
    Nb<D> nb;
    if (!neighbour_at<D>(g, cc, pk[jj], nb)) continue;
    int v[CSLOT];
    int ln;
    bool in_row;
    const int* o;
    if (cellcc != nullptr) {
      // first round: the cell's list built here from its count and arrival slots,
      // as k_lists would (the kernel gives up when some cell has more than CSLOT)
      const int na = load_now(cellcc + nb.src);
      const int64_t* row = reinterpret_cast<const int64_t*>(cslot + (size_t)nb.src * CSLOT);
      const int64_t q0 = load_now(row), q1 = load_now(row + 1), q2 = load_now(row + 2), q3 = load_now(row + 3);
      if (na == 0) continue;
      v[0] = (int)q0; v[1] = (int)(q0 >> 32); v[2] = (int)q1; v[3] = (int)(q1 >> 32);
      v[4] = (int)q2; v[5] = (int)(q2 >> 32); v[6] = (int)q3; v[7] = (int)(q3 >> 32);
      ln = short_list(na < CSLOT ? na : CSLOT, v, ctag, self);
      in_row = true;
      o = ord;  // not read: the list is in v (ln <= CSLOT)
    } else {
      const int t = cell2t[nb.src];
      if (t < 0) continue;
      // A list of at most CSLOT entries starts at t * CSLOT (k_lists): that row is
      // read with the list's offset and length (t < nshots keeps it in bounds) and
      // used whenever the offset says the list starts there.
      const int sg = load_now(seg + t);
      ln = load_now(len + t);
      const int64_t* row = reinterpret_cast<const int64_t*>(ord + (size_t)t * CSLOT);
      const int64_t q0 = load_now(row), q1 = load_now(row + 1), q2 = load_now(row + 2), q3 = load_now(row + 3);
      v[0] = (int)q0; v[1] = (int)(q0 >> 32); v[2] = (int)q1; v[3] = (int)(q1 >> 32);
      v[4] = (int)q2; v[5] = (int)(q2 >> 32); v[6] = (int)q3; v[7] = (int)(q3 >> 32);
      in_row = sg == t * CSLOT;
      o = ord + sg;
    }
    for (int k = 0; k < ln; k++) {
      int j;
      if (in_row && k < CSLOT) {
        j = v[0];
#pragma unroll
        for (int q = 1; q < CSLOT; q++) j = k == q ? v[q] : j;
      } else {
        j = o[k];
      }
      // a candidate's tag, priority, position and radius are loaded together
      const uint8_t tj = load_now(ctag + j);
      const uint32_t pj = load_now(cprio + j);
      const double rj = load_now(crad + j);
      double jp[D], qp[D];
#pragma unroll
      for (int d = 0; d < D; d++) jp[d] = load_now(cpos + (size_t)d * nshots + j);
      if (tj != T_UNDECIDED) continue;
      if (!(pj < pi)) continue;
      if (nb.ghost) {
        if (!image_of<D>(g, nb, jp, qp)) continue;
        if (out == nullptr && mismatch) {  // the image must fall in this ghost cell
          const int64_t ic = cell_index<D>(g, qp);
          int64_t want = 0;
          for (int d = D - 1; d >= 0; d--) want = want * g.n[d] + (cc[d] + stencil_off<D>(g.stab[s], d));
          if (ic != want) (*mismatch)++;  // added by the caller once its gate is open
        }
      } else {
#pragma unroll
        for (int d = 0; d < D; d++) qp[d] = jp[d];
      }
      if (collide<D>(r, p, rj, qp)) {
        if (n == 0) first_ghost = nb.ghost;
        const int e = j | (nb.ghost ? (int)0x80000000 : 0);
        if (out) out[n] = e;
        if (keep) {  // the first PKEEP partners in registers (constant indices)
#pragma unroll
          for (int t = 0; t < PKEEP; t++) (*keep)[t] = n == t ? e : (*keep)[t];
        }
        n++;
      }
    }
  }
  return n;
}

// Sweep labels, first step (see k_sweep_labels for the semantics): candidates without
// partners stay Undecided, those whose first partner is an image become
// Conflict; the others get their ordered partner list (slab of KMAX entries,
// longer lists from a pool) and are left unresolved for k_sweep_labels.
template <int D>
__global__ void __launch_bounds__(WG) k_partner_lists(Geo<D> g, int nshots, const double* __restrict__ cpos,
                                                 const double* __restrict__ crad, const uint32_t* __restrict__ cprio,
                                                 const int* __restrict__ ccell, const uint8_t* __restrict__ ctag,
                                                 uint8_t* __restrict__ lab, const int* __restrict__ cell2t,
                                                 const int* __restrict__ seg, const int* __restrict__ len,
                                                 const int* __restrict__ ord, int* __restrict__ slab,
                                                 int* __restrict__ pool, int pool_cap, DrawCtl* ctl, int gate_id,
                                                 URec* __restrict__ urec, const int* __restrict__ cellcc,
                                                 const int* __restrict__ cslot) {
  using LG = LaneGroup<D>;
  // A draw's first round (gate_id >= 0) builds the cell lists from the arrival
  // slots; the later ones read the lists k_update_commit kept.
  const bool first = gate_id >= 0;
  const int gate = first ? ctl->gate : gate_id;  // tested before the first store (loads overlap)
  const int long_list = first ? ctl->long_list : 0;
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int i = gtid / LG::LPC;
  const int sub = gtid % LG::LPC;
  int pk[LG::CPL];
  lane_stencil<D, LG::CPL, LG::NST>(g, sub, pk);
  // read before the gate test: with a closed gate the candidate arrays are stale
  // or never written, so a cell outside the grid disables the lane
  const int ci = i < nshots ? ccell[i] : -1;
  const bool active = i < nshots && ctag[i] == T_UNDECIDED && ci >= 0 && ci < g.ncells;
  int n = 0;
  bool fg = false;
  double p[D];
  double r = 0;
  uint32_t pi = 0;
  int cc[D];
  int keep[PKEEP] = {};
  int mism = 0;
  if (active) {
    cand_pos<D>(cpos, nshots, i, p);
    r = crad[i];
    pi = cprio[i];
    cell_coords<D>(g, ccell[i], cc);
    n = lane_partners<D>(g, cc, sub, nshots, p, r, pi, cpos, crad, cprio, ctag, cell2t, seg, len, ord, pk,
                         first ? cellcc : nullptr, cslot, i, nullptr, fg, &mism, &keep);
  }
  if (gate != gate_id) return;  // whole grid
  if (long_list) return;  // a cell of more than CSLOT candidates: the host builds the lists and redoes the round
  if (mism) atomicAdd(&ctl->img_mismatch, mism);
  int incl = n;  // inclusive scan over the lane group (stencil order)
#pragma unroll
  for (int off = 1; off < LG::LPC; off <<= 1) {
    const int v = __shfl_up(incl, off, LG::LPC);
    if (sub >= off) incl += v;
  }
  const int total = __shfl(incl, LG::LPC - 1, LG::LPC);
  const uint64_t has = __ballot(n > 0) & LG::mask(lane);
  const int first_lane = has ? (__ffsll((unsigned long long)has) - 1) : 0;
  const int fgi = __shfl(fg ? 1 : 0, first_lane, 64);
  const bool pend = active && total > 0 && !fgi;
  int off = -2;
  int u = 0;
  if (active && sub == 0) {
    if (total == 0) {
      lab[i] = T_UNDECIDED;
    } else if (fgi) {
      lab[i] = T_CONFLICT;
    } else {
      lab[i] = T_UNRESOLVED;
      u = atomicAdd(&ctl->nunres, 1);
      if (total <= KMAX) {
        off = -1;
      } else {
        off = atomicAdd(&ctl->pool_cursor, total);
        if (off + total > pool_cap) {
          atomicMax(&ctl->pool_overflow, 1);
          atomicMax(&ctl->need_pool, off + total);
          off = -2;
        }
      }
      *reinterpret_cast<int4*>(urec + u) = make_int4(i, total, off, 0);
    }
  }
  const int lead = (lane / LG::LPC) * LG::LPC;
  off = __shfl(off, lead, 64);
  u = __shfl(u, lead, 64);
  if (pend && off != -2) {
    // positions < PKEEP go to the record, the others to the slab or pool
    int* dst = off == -1 ? slab + (size_t)i * KMAX : pool + off;
    int* head = urec[u].e;
    const int b = incl - n;  // this lane's first position in the candidate's list
    if (n > 0 && n <= PKEEP) {
#pragma unroll
      for (int t = 0; t < PKEEP; t++) {
        if (t < n) {
          if (b + t < PKEEP)
            head[b + t] = keep[t];
          else
            dst[b + t] = keep[t];
        }
      }
    } else if (n > PKEEP) {  // rare: walk the cells again, writing as it goes
      bool dummy;
      lane_partners<D>(g, cc, sub, nshots, p, r, pi, cpos, crad, cprio, ctag, cell2t, seg, len, ord, pk,
                       first ? cellcc : nullptr, cslot, i, dst + b, dummy, nullptr);
      for (int t = b; t < PKEEP; t++) head[t] = dst[t];
    }
  }
}

// The sequential priority sweep (find_solve_conflicts,
// operator_cells_and_data_storage.ixx:127-142) as a dependency evaluation: a
// candidate's label is fixed by its first partner (detection order) that is not
// Skip -- an image or a Conflict partner gives Conflict, an Undecided real
// partner gives Skip, none leaves it Undecided.  Partners have strictly smaller
// priorities, so in the reference's sorted sweep they are labelled first; the
// passes reach the same labels.  One workgroup, labels in LDS (nshots bytes of
// dynamic shared memory).  The round's counts are left in DrawCtl::pubv, and
// k_update_commit writes them to pinned host memory.
__global__ void __launch_bounds__(1024) k_sweep_labels(int nshots, const uint8_t* __restrict__ ctag, uint8_t* lab,
                                                const int* __restrict__ slab, const int* __restrict__ pool,
                                                DrawCtl* ctl, int seq, int gate_id, int draw_id,
                                                double aimed, uint64_t miss_base, const URec* __restrict__ urec) {
  extern __shared__ __align__(16) uint8_t s_lab[];
  __shared__ int s_prog, s_left, s_acc, s_con, s_stuck;
  // everything below reads first and stores only once the gate is known open
  const int gate = gate_id >= 0 ? ctl->gate : gate_id;
  const bool ovf = ctl->pool_overflow != 0;
  // a first round that met a cell of more than CSLOT candidates did nothing: the
  // host builds the lists and redoes it (like a pool overflow)
  const bool longl = gate_id >= 0 && ctl->long_list != 0;
  const bool redo = ovf || longl;
  const int nu = ctl->nunres;
  const int tid = threadIdx.x;
  // the counters the round publishes: no kernel of this round changes them
  // after k_partner_lists, so they are read now, with the loads above
  int pub_dropped = 0, pub_mism = 0, pub_povf = 0, pub_need = 0, pub_cell = 0, acc_before = 0;
  if (tid == 0) {
    pub_dropped = ctl->dropped;
    pub_mism = ctl->img_mismatch;
    pub_povf = ctl->pool_overflow;
    pub_need = ctl->need_pool;
    pub_cell = ctl->cell_overflow;
    acc_before = ctl->acc_draw;
    s_acc = 0;
    s_con = 0;
    s_stuck = 0;
  }
  // labels into LDS, 8 candidates per thread and load (the buffers start 256-byte aligned)
  for (int q8 = tid * 8; q8 < nshots; q8 += 8 * (int)blockDim.x) {
    if (q8 + 8 <= nshots) {
      const uint2 t = *reinterpret_cast<const uint2*>(ctag + q8);
      const uint2 l = *reinterpret_cast<const uint2*>(lab + q8);
      auto merge = [](uint32_t tw, uint32_t lw) {
        uint32_t o = 0;
#pragma unroll
        for (int b = 0; b < 32; b += 8) {
          const uint32_t tb = (tw >> b) & 0xffu;
          o |= (tb == T_UNDECIDED ? (lw >> b) & 0xffu : (uint32_t)T_DROPPED) << b;
        }
        return o;
      };
      *reinterpret_cast<uint2*>(s_lab + q8) = make_uint2(merge(t.x, l.x), merge(t.y, l.y));
    } else {
      for (int q = q8; q < nshots; q++) s_lab[q] = ctag[q] == T_UNDECIDED ? lab[q] : T_DROPPED;
    }
  }
  // each thread keeps up to two unresolved candidates and their first PKEEP
  // partners in registers for all passes; further ones are re-read per pass.
  // Records below nshots are read whatever the count says (they are in bounds)
  // and used below the count only, so their loads overlap the count's.
  constexpr int NC = 2;
  int cq[NC], cm[NC];
  int ce[NC][PKEEP];
  const int* cpl[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const int u = tid + c * (int)blockDim.x;
    int4 h = make_int4(-1, 0, 0, 0), e = make_int4(0, 0, 0, 0);
    if (u < nshots) {
      const int4* rc = reinterpret_cast<const int4*>(urec + u);
      h = rc[0];
      e = rc[1];
    }
    const bool use = !redo && u < nu;
    cq[c] = use ? h.x : -1;
    cm[c] = use ? h.y : 0;
    cpl[c] = use ? (h.z == -1 ? slab + (size_t)h.x * KMAX : pool + h.z) : nullptr;
    ce[c][0] = e.x;
    ce[c][1] = e.y;
    ce[c][2] = e.z;
    ce[c][3] = e.w;
  }
  __syncthreads();
  if (gate != gate_id) return;  // the whole workgroup
  // one candidate's label from its partners (first non-Skip one decides), or
  // "wait" while that partner is unresolved
  auto resolve = [&](int m, const int* pl, const int* cache, uint8_t& res) -> bool {
    res = T_UNDECIDED;
    for (int k = 0; k < m; k++) {
      int e;
      if (k < PKEEP) {
        e = cache[0];
#pragma unroll
        for (int t = 1; t < PKEEP; t++) e = k == t ? cache[t] : e;
      } else {
        e = pl[k];
      }
      if (e < 0) {
        res = T_CONFLICT;
        return true;
      }
      const uint8_t lj = s_lab[e];
      if (lj == T_UNRESOLVED) return false;
      if (lj == T_SKIP) continue;
      res = (lj == T_CONFLICT) ? T_CONFLICT : T_SKIP;
      return true;
    }
    return true;
  };
  if (!redo) {
    // with no unresolved candidate (nu is the same for the whole workgroup) no pass is needed
    for (int pass = 0; nu > 0 && pass < nshots + 2; pass++) {
      if (tid == 0) {
        s_prog = 0;
        s_left = 0;
      }
      __syncthreads();
      int prog = 0, left = 0;
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const int q = cq[c];
        if (q < 0 || s_lab[q] != T_UNRESOLVED) continue;
        uint8_t res;
        if (resolve(cm[c], cpl[c], ce[c], res)) {
          s_lab[q] = res;
          lab[q] = res;
          prog++;
        } else {
          left++;
        }
      }
      for (int u = tid + NC * (int)blockDim.x; u < nu; u += blockDim.x) {
        const int4* rc = reinterpret_cast<const int4*>(urec + u);
        const int4 h = rc[0], e = rc[1];
        const int q = h.x;
        if (s_lab[q] != T_UNRESOLVED) continue;
        const int m = h.y;
        const int* pl = h.z == -1 ? slab + (size_t)q * KMAX : pool + h.z;
        const int cache[PKEEP] = {e.x, e.y, e.z, e.w};
        uint8_t res;
        if (resolve(m, pl, cache, res)) {
          s_lab[q] = res;
          lab[q] = res;
          prog++;
        } else {
          left++;
        }
      }
      if (prog) atomicAdd(&s_prog, prog);
      if (left) atomicAdd(&s_left, left);
      __syncthreads();
      const int sp = s_prog, sl = s_left;
      __syncthreads();
      if (sl == 0) break;
      if (sp == 0) {
        if (tid == 0) s_stuck = 1;
        break;
      }
    }
    // counts over 8 labels per thread, summed over the wave, one LDS atomic per wave
    int acc = 0, con = 0;
    for (int q8 = tid * 8; q8 < nshots; q8 += 8 * (int)blockDim.x) {
      if (q8 + 8 <= nshots) {
        const uint2 w = *reinterpret_cast<const uint2*>(s_lab + q8);
#pragma unroll
        for (int b = 0; b < 64; b += 8) {
          const uint32_t v = ((b < 32 ? w.x >> b : w.y >> (b - 32))) & 0xffu;
          acc += v == T_UNDECIDED;
          con += v == T_CONFLICT;
        }
      } else {
        for (int q = q8; q < nshots; q++) {
          acc += s_lab[q] == T_UNDECIDED;
          con += s_lab[q] == T_CONFLICT;
        }
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      acc += __shfl_down(acc, off, 64);
      con += __shfl_down(con, off, 64);
    }
    if ((tid & 63) == 0) {
      if (acc) atomicAdd(&s_acc, acc);
      if (con) atomicAdd(&s_con, con);
    }
  }
  __syncthreads();
  if (tid != 0) return;
  ctl->acc = s_acc;
  ctl->con = s_con;
  ctl->label_stuck = s_stuck;
  // counters whose values go into the publication below
  ctl->dropped = 0;
  ctl->img_mismatch = 0;
  ctl->seg_cursor = 0;
  ctl->nunres = 0;
  if (!ovf) ctl->pool_cursor = 0;
  // the draw's refinement decision, as the host makes it after the draw
  // (operator_algorithm.ixx:170-180): miss rate = misses / (1e-6 + shots),
  // misses = shots - (spheres now - spheres after the previous draw of this
  // proceed) in unsigned 64-bit arithmetic; miss_base carries the part of the
  // difference placed before the draw (non-zero on a proceed's first draw).
  int open = 0;
  const int acc_draw = redo ? acc_before : acc_before + s_acc;
  ctl->acc_draw = acc_draw;
  if (!redo && !s_stuck && s_con == 0) {
    const uint64_t miss = (uint64_t)nshots - (uint64_t)acc_draw - miss_base;
    const double rate = (double)miss / (0.000001 + (double)(uint64_t)nshots);
    open = !(rate > aimed);
  }
  ctl->gate = open ? draw_id + 1 : -1;
  ctl->swept = seq;
  const int v[kPubFields] = {s_acc,    s_con,    s_stuck, pub_dropped, pub_mism,
                             pub_povf, pub_need, pub_cell, open,        longl ? 1 : 0};
  for (int k = 0; k < kPubFields; k++) ctl->pubv[k] = v[k];  // published by k_update_commit
}

// update_sphere_tags on the touched real cells (operator_algorithm.ixx:446-473),
// one thread per cell owner: Skip removed from the back, Undecided ->
// Validated, Conflict -> Undecided.  When the round left no Conflict the draw
// is over and the survivors are appended to their cells in tail order.
template <int D>
__global__ void __launch_bounds__(WG) k_update_commit(CellStore<D> cs, int nshots, const double* __restrict__ cpos,
                                                      const double* __restrict__ crad,
                                                      const uint32_t* __restrict__ cphase,
                                                      const int* __restrict__ ccell, const uint8_t* __restrict__ owner,
                                                      uint8_t* __restrict__ ctag, const uint8_t* __restrict__ lab,
                                                      int* __restrict__ seg, int* __restrict__ len,
                                                      int* __restrict__ ord, int* __restrict__ cell2t,
                                                      DrawCtl* __restrict__ ctl, int seq, DrawPub* host, int first,
                                                      int* __restrict__ cellcc, const int* __restrict__ cslot,
                                                      int ncells) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int swept = ctl->swept;  // all reads are issued before the tests
  if (i == 0 && swept == seq) {
    // the round's results to the host, as the sweep left them (seq last)
    int v[kPubFields];
#pragma unroll
    for (int k = 0; k < kPubFields; k++) v[k] = ctl->pubv[k];
    DrawPub* hp = host + (seq & (kPubSlots - 1));
#pragma unroll
    for (int k = 0; k < kPubFields; k++) __atomic_store_n(&reinterpret_cast<int*>(hp)[k], v[k], __ATOMIC_RELAXED);
    __threadfence_system();
    __atomic_store_n(&hp->seq, seq, __ATOMIC_RELEASE);
  }
  const bool skip_round = ctl->pool_overflow || ctl->label_stuck || (first && ctl->long_list);
  const bool commit = ctl->con == 0;
  const bool in = i < nshots;
  // Every thread also reads, without a branch (index clamped to candidate 0), its
  // owner flag and cell, and its list: in a draw's first round the cell's count
  // and arrival slots (cell clamped into the grid), later the list kept by the
  // previous round (offset, length, and the row at i * CSLOT where a list of at
  // most CSLOT entries starts).  The loads are issued with the ones above, all in
  // bounds whatever they hold.
  const int ic = in ? i : 0;
  const uint8_t own = load_now(owner + ic);
  const bool on = in && own;
  const int c = load_now(ccell + ic);
  const int cl = (c >= 0 && c < ncells) ? c : 0;
  const int sg = first ? ic * CSLOT : load_now(seg + ic);
  const int n0 = first ? load_now(cellcc + cl) : load_now(len + ic);
  const int64_t* row =
      reinterpret_cast<const int64_t*>(first ? cslot + (size_t)cl * CSLOT : ord + (size_t)ic * CSLOT);
  const int64_t q0 = load_now(row), q1 = load_now(row + 1), q2 = load_now(row + 2), q3 = load_now(row + 3);
  const int4 r0 = make_int4((int)q0, (int)(q0 >> 32), (int)q1, (int)(q1 >> 32));
  const int4 r1 = make_int4((int)q2, (int)(q2 >> 32), (int)q3, (int)(q3 >> 32));
  if (swept != seq || !on || skip_round) return;  // its sweep did not run (closed gate), or no owner
  // an owner's cell is in the grid
  const int cnt0 = load_now(cs.cnt + c);
  if (first || (sg == i * CSLOT && n0 <= CSLOT)) {
    // the list in registers: its tags are loaded together, the removal (same
    // steps as the loop below) works on the registers
    int v[CSLOT] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
    int nn = n0;
    // tags and labels without a branch (entries past the list read the owner's own)
    uint8_t tg[CSLOT], lb[CSLOT];
    if (first) {
      // the list k_lists would build (shot order, round-0 Skips removed from the
      // back with swap-with-last), its tags and labels loaded once and moved with
      // their entries; kept at i * CSLOT for the later rounds; the cell's
      // arrivals restart at 0
      const int na = n0 < CSLOT ? n0 : CSLOT;
      sort_slots(na, v);
#pragma unroll
      for (int k = 0; k < CSLOT; k++) {
        const int j = k < na ? v[k] : i;
        tg[k] = ctag[j];
        lb[k] = lab[j];
      }
      nn = na;
#pragma unroll
      for (int q = CSLOT - 1; q >= 0; q--) {
        if (q < na && tg[q] == T_SKIP) {
          int x = v[q];
          uint8_t xt = tg[q], xl = lb[q];
#pragma unroll
          for (int k = q + 1; k < CSLOT; k++) {
            const bool last = nn - 1 == k;
            x = last ? v[k] : x;
            xt = last ? tg[k] : xt;
            xl = last ? lb[k] : xl;
          }
          v[q] = x;
          tg[q] = xt;
          lb[q] = xl;
          nn--;
        }
      }
      cellcc[c] = 0;
      seg[i] = i * CSLOT;
    } else {
#pragma unroll
      for (int k = 0; k < CSLOT; k++) {
        const int j = k < nn ? v[k] : i;
        tg[k] = ctag[j];
        lb[k] = lab[j];
      }
    }
    uint8_t e[CSLOT];
    bool und[CSLOT];
#pragma unroll
    for (int k = 0; k < CSLOT; k++) {
      e[k] = k < nn ? (tg[k] == T_UNDECIDED ? lb[k] : tg[k]) : T_DROPPED;
      und[k] = k < nn && tg[k] == T_UNDECIDED && lb[k] == T_UNDECIDED;
    }
    int m = nn;
#pragma unroll
    for (int q = CSLOT - 1; q >= 0; q--) {
      if (q < nn && e[q] == T_SKIP) {
        ctag[v[q]] = T_SKIP;
        int x = v[q];
        bool ux = und[q];
#pragma unroll
        for (int k = q + 1; k < CSLOT; k++) {
          x = (m - 1 == k) ? v[k] : x;
          ux = (m - 1 == k) ? und[k] : ux;
        }
        v[q] = x;
        und[q] = ux;
        m--;
      }
    }
    len[i] = m;
    int* a = ord + sg;
#pragma unroll
    for (int k = 0; k < CSLOT; k++) {
      if (k < m) {
        a[k] = v[k];
        if (und[k]) ctag[v[k]] = T_VALIDATED;
        // Conflict candidates keep the Undecided tag for the next round
      }
    }
    if (!commit) {
      if (first) cell2t[c] = i;  // the cell's list for the later rounds
      return;
    }
    Slot<D> sl[CSLOT];
    uint32_t ph[CSLOT];
#pragma unroll
    for (int k = 0; k < CSLOT; k++) {
      if (k < m) {  // m is small: the loads of the kept entries, in one block each
#pragma unroll
        for (int d = 0; d < D; d++) sl[k].c[d] = cpos[(size_t)d * nshots + v[k]];
        sl[k].r = crad[v[k]];
        ph[k] = cphase[v[k]];
      }
    }
    int k0 = cnt0;
#pragma unroll
    for (int k = 0; k < CSLOT; k++) {
      if (k < m) {
        if (k0 >= cs.cap) {
          atomicMax(&ctl->cell_overflow, 1);
          break;
        }
        cs.sph[(size_t)c * cs.cap + k0] = sl[k];
        cs.phase[(size_t)c * cs.cap + k0] = ph[k];
        k0++;
      }
    }
    cs.cnt[c] = (uint8_t)k0;
    cell2t[c] = -1;
    return;
  }
  int* a = ord + sg;
  int n = n0;
  for (int q = n - 1; q >= 0; q--) {
    const int j = a[q];
    const uint8_t e = ctag[j] == T_UNDECIDED ? lab[j] : ctag[j];
    if (e == T_SKIP) {
      ctag[j] = T_SKIP;
      a[q] = a[n - 1];
      n--;
    }
  }
  len[i] = n;
  for (int q = 0; q < n; q++) {
    const int j = a[q];
    if (ctag[j] == T_UNDECIDED && lab[j] == T_UNDECIDED) ctag[j] = T_VALIDATED;
    // Conflict candidates keep the Undecided tag for the next round
  }
  if (!commit) return;
  int k0 = cnt0;
  for (int q = 0; q < n; q++) {
    const int j = a[q];
    if (k0 >= cs.cap) {
      atomicMax(&ctl->cell_overflow, 1);
      break;
    }
    Slot<D> sl;
#pragma unroll
    for (int d = 0; d < D; d++) sl.c[d] = cpos[(size_t)d * nshots + j];
    sl.r = crad[j];
    cs.sph[(size_t)c * cs.cap + k0] = sl;
    cs.phase[(size_t)c * cs.cap + k0] = cphase[j];
    k0++;
  }
  cs.cnt[c] = (uint8_t)k0;
  cell2t[c] = -1;
}

// ============================================================== kernels: voxel refine

// path::TabCorner (auxiliary/Path.ixx:85-116) as the reference builds it: the
// children / corner order {00,11,01,10} (2D), {000,111,001,110,010,101,011,100}
// (3D); for D >= 4 the reference's loop never sets component 0 (kept for parity).
template <int D>
struct TabCorner {
  __host__ __device__ static constexpr int at(int k, int d) {
    if constexpr (D == 2) {
      constexpr int t[4][2] = {{0, 0}, {1, 1}, {0, 1}, {1, 0}};
      return t[k][d];
    } else if constexpr (D == 3) {
      constexpr int t[8][3] = {{0, 0, 0}, {1, 1, 1}, {0, 0, 1}, {1, 1, 0},
                               {0, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 0, 0}};
      return t[k][d];
    } else {
      return d == 0 ? 0 : ((k >> d) & 1);
    }
  }
  // whether component d takes both values over the table (the corner set is
  // then the product over d of the values taken)
  __host__ __device__ static constexpr bool both(int d) { return D <= 3 || d != 0; }
};

template <int D>
struct VoxParams {
  static constexpr int NCH = ipow_c(2, D);
  double origin[D], len[D], nl[D];
  double rmin;
  // h2 = sum_d (len_d / 2)^2 when every dimension has both corner values, else
  // a negative sentinel.  Then the farthest corner of a box is at squared
  // distance >= |c - centre|^2 + h2 from any point c (per dimension
  // max(a^2, b^2) >= (a^2 + b^2) / 2 = (delta - len/2)^2 + len^2/4), so a sphere
  // with |c - centre|^2 > R^2 - h2 cannot cover the box.
  double h2;
  // The pre-check evaluates |c - centre|^2 from the rounded centre() of the
  // box, the full test from its rounded origin: their positions differ by a few
  // roundings of the coordinates' magnitude X (per dimension at most 8 eps X
  // with eps = 2^-53, over D dimensions |e| <= 8 sqrt(D) eps X).  With a
  // covering sphere |c - centre| <= R, so the computed |c - centre|^2 exceeds
  // the exact one by at most 2 R |e| + |e|^2; the sums' own roundings stay
  // below 1e-9 R^2.  slack() is that bound times 4:
  // 1e-9 R^2 + xs R + xs2 with xs = 64 sqrt(D) eps X, xs2 = 256 D eps^2 X^2.
  double xs, xs2;
  __device__ __forceinline__ double slack(double R2, double rr) const { return 1e-9 * R2 + xs * rr + xs2; }
};

// auxi::is_in_sphere (list_of_voxels.ixx:225-238) asks whether every corner
// tab[i] * len of a box lies within r + rmin of the sphere centre, the corner
// term of dimension d being t = tab[i][d] * len[d] - delta[d], summed as
// s = ((0 + t0^2) + t1^2) + ... in dimension order.  The corner set is a
// product set, and IEEE addition and squaring are monotone (in |t| for the
// square), so the largest corner sum is the same sum over the per-dimension
// largest terms: testing that one sum gives the all-corner loop's answer bit
// for bit.  far_term is that largest term of one dimension (both: the
// dimension takes tab values 0 and 1; otherwise only 0).
template <bool BOTH>
__device__ __forceinline__ double far_term(double delta, double len) {
  const double aa = delta * delta;  // tab 0: (0 * len - delta)^2; 0 - delta = -delta exactly
  if constexpr (!BOTH) {
    return aa;
  } else {
    const double b = len - delta;  // tab 1: 1 * len - delta
    const double bb = b * b;
    return bb > aa ? bb : aa;  // squares are >= +0: the larger one, as fmax
  }
}

// The placed spheres packed in cell order, for k_cover: off[c] = cnt[c] here,
// an exclusive scan (with off[ncells] = total) next, then k_compact copies
// cell c's spheres to off[c] .. off[c + 1] - 1.  The cells along dimension 0
// are consecutive, so three neighbours in a row are one contiguous range.
__global__ void __launch_bounds__(WG) k_cell_cnt(int64_t ncells, const uint8_t* __restrict__ cnt,
                                                 int* __restrict__ off) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncells) off[c] = cnt[c];
}

template <int D>
__global__ void __launch_bounds__(WG) k_compact(int64_t ncells, CellStore<D> cs, const int* __restrict__ off,
                                                Slot<D>* __restrict__ out, int64_t cap, int* __restrict__ err) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int n = cs.cnt[c], o = off[c];
  if ((int64_t)o + n > cap) {
    atomicOr(err, 1);
    return;
  }
  for (int k = 0; k < n; k++) out[o + k] = cs.sph[(size_t)c * cs.cap + k];
}

// remove_covered's coverage flag (list_of_voxels.ixx:30-40, 240-259) and
// subdivide_uncovered's children test (:56-100, 261-307) in one pass:
// out[i] = 0 if voxel i is covered, else 0x10000 | (bit k set when child k in
// TabCorner order stays uncovered).  blk[b] counts uncovered voxels per block.
// Interior voxels of D <= 3 read the packed copy (coff, csph) by stencil rows
// when the host made one (coff not null); the others read the cell store.
template <int D, bool WIDE>
__global__ void __launch_bounds__(WG) k_cover(Geo<D> g, CellStore<D> cs, VoxView<D> vx, int64_t nvox,
                                              VoxParams<D> vp, const uint32_t* __restrict__ perm,
                                              const int* __restrict__ coff, const Slot<D>* __restrict__ csph,
                                              int ncs, uint32_t* __restrict__ out, int* __restrict__ blk) {
  using TC = TabCorner<D>;
  constexpr int NCH = ipow_c(2, D);
  constexpr int NST = ipow_c(3, D);
  constexpr uint32_t FULL = (NCH == 32) ? 0xffffffffu : ((1u << NCH) - 1u);
  // thread t takes voxel perm[t] (the cell-grouped order) or, without perm, voxel t
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = perm == nullptr ? t : (t < nvox ? (int64_t)perm[t] : nvox);
  // The stencil as NR rows of three cells along dimension 0 (any order:
  // coverage is an OR): the linear offset of each row's first cell.
  constexpr int NR = NST / 3;
  constexpr int NRL = NST <= 32 ? NR : 1;
  __shared__ int s_row[NRL];
  __shared__ uint8_t s_last[NRL];  // 1 where the next row is farther from the centre row
  constexpr int KLIST = 16;  // listed spheres per lane before a flush
  __shared__ uint32_t s_list[KLIST][WG];
  if (NST <= 32 && threadIdx.x < NRL) {
    // centre row first, then by the number of non-zero offsets in dimensions
    // 1 .. D-1: a covering sphere is met early and a covered voxel stops there
    auto nz = [](int r) {
      int n = 0;
      for (int d = 1; d < D; d++, r /= 3) n += (r % 3 != 1);
      return n;
    };
    const int t = threadIdx.x, nt = nz(t);
    int pos = 0, same_after = 0;
    for (int u = 0; u < NR; u++) {
      const int nu = nz(u);
      pos += (nu < nt) || (nu == nt && u < t);
      same_after += (nu == nt && u > t);
    }
    int o = -1, r = t, stride = g.n[0];
#pragma unroll
    for (int d = 1; d < D; d++) {
      o += (r % 3 - 1) * stride;
      r /= 3;
      stride *= g.n[d];
    }
    s_row[pos] = o;
    s_last[pos] = same_after == 0;
  }
  __syncthreads();
  uint32_t code = 0;
  if (i < nvox) {
    // ov (orig_voxel, list_of_voxels.ixx:146-153) = origin + c * len equals the
    // tab-0 child origin origin + (2c) * (len / 2) bit for bit (same real
    // product, len / 2 exact), so ok0 serves as both
    double ca[D], cb[D], ok0[D], ok1[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int64_t c = WIDE ? vx.w[d][i] : (int64_t)vx.c[d][i];
      ok0[d] = vp.origin[d] + (double)(2 * c) * vp.nl[d];      // children origins (:82-86), tab 0
      ok1[d] = vp.origin[d] + (double)(2 * c + 1) * vp.nl[d];  // and tab 1
      ca[d] = ok0[d] + 0.5 * vp.len[d];                         // is_covered's centre (:245)
      cb[d] = vp.origin[d] + ((double)c + 0.5) * vp.len[d];  // center() used by subdivide (:155-162)
    }
    const double* ov = ok0;
    // voxel centres lie inside the real cells
    int cca[D], ccb[D];
    const int cella = cell_of_point<D>(g, ca, cca);
    const int cellb = cell_of_point<D>(g, cb, ccb);
    uint32_t ccov = 0;
    bool covered = false;
    // One sphere (centre qp, radius qr) against the voxel and its children.
    // Coverage and each child's flag are ORs over spheres: visiting order is free.
    auto test = [&](const double* qp, double qr, bool do_cov, bool do_child) {
      const double rr = qr + vp.rmin;
      const double R2 = rr * rr;
      double s2 = 0.;  // get_list_of_spheres pre-filter (:297-298), from the subdivide centre
#pragma unroll
      for (int d = 0; d < D; d++) {
        const double t = qp[d] - cb[d];
        s2 += t * t;
      }
      // necessary for coverage (VoxParams::h2); the slack absorbs rounding
      if (do_cov && (vp.h2 < 0. || s2 <= R2 - vp.h2 + vp.slack(R2, rr))) {
        double sf = 0.;
#pragma unroll
        for (int d = 0; d < D; d++) sf += TC::both(d) ? far_term<true>(qp[d] - ov[d], vp.len[d])
                                                      : far_term<false>(qp[d] - ov[d], vp.len[d]);
        if (!(sf > R2)) {
          covered = true;
          return;
        }
      }
      if (do_child && s2 < R2 && ccov != FULL) {
        double f0[D], f1[D];
#pragma unroll
        for (int d = 0; d < D; d++) {
          f0[d] = TC::both(d) ? far_term<true>(qp[d] - ok0[d], vp.nl[d]) : far_term<false>(qp[d] - ok0[d], vp.nl[d]);
          f1[d] = TC::both(d) ? far_term<true>(qp[d] - ok1[d], vp.nl[d]) : far_term<false>(qp[d] - ok1[d], vp.nl[d]);
        }
#pragma unroll
        for (int ch = 0; ch < NCH; ch++) {
          double sf = 0.;
#pragma unroll
          for (int d = 0; d < D; d++) sf += TC::at(ch, d) ? f1[d] : f0[d];
          if (!(sf > R2)) ccov |= (1u << ch);
        }
      }
    };
    // pass 0: coverage around cella (and children when cellb == cella); pass 1: children around cellb
    for (int pass = 0; pass < 2 && !covered; pass++) {
      const int cell = pass == 0 ? cella : cellb;
      const bool do_cov = pass == 0;
      const bool do_child = (pass == 0) ? (cellb == cella) : (cellb != cella);
      if (!do_cov && !do_child) break;
      int cc[D];
#pragma unroll
      for (int d = 0; d < D; d++) cc[d] = pass == 0 ? cca[d] : ccb[d];
      bool interior = true;  // the whole stencil lies in real cells: plain index offsets
#pragma unroll
      for (int d = 0; d < D; d++) interior = interior && cc[d] > g.gl && cc[d] < g.n[d] - g.gl - 1;
      if (interior && NST <= 32 && coff != nullptr) {
        // Two phases against divergence: every sphere of the stencil gets the
        // cheap distance test and the ones either test can use are listed (per
        // lane, in LDS); the full tests run over the short lists at the end of
        // each distance shell (and when a list is full), so that a voxel covered
        // by a sphere of the nearer cells stops there.  A sphere is listed when
        // it passes the subdivide pre-filter or the coverage pre-check, exactly
        // the gates of the full test.
        int nl = 0;
        auto flush = [&]() {
          for (int t = 0; t < nl && !covered; t++) {
            const Slot<D> q = csph[s_list[t][threadIdx.x]];
            test(q.c, q.r, do_cov, do_child);
          }
          nl = 0;
        };
        // a row's spheres are csph[coff[s] .. coff[s + 3]); the bounds are read
        // one row ahead of the tests (the stencil lies in real cells: s >= 0,
        // s + 3 <= ncells), and kept within the ncs spheres the copy holds
        int src = (int)cell + s_row[0];
        int an = coff[src], bn = coff[src + 3];
        for (int rw = 0; rw < NR && !covered; rw++) {
          const int a = an, b = min(bn, ncs);
          if (rw + 1 < NR) {
            src = (int)cell + s_row[rw + 1];
            an = coff[src];
            bn = coff[src + 3];
          }
          auto consider = [&](const Slot<D>& q, uint32_t idx) {
            const double rr = q.r + vp.rmin;
            const double R2 = rr * rr;
            double s2 = 0.;
#pragma unroll
            for (int d = 0; d < D; d++) {
              const double t = q.c[d] - cb[d];
              s2 += t * t;
            }
            const bool use =
                (do_child && s2 < R2) || (do_cov && (vp.h2 < 0. || s2 <= R2 - vp.h2 + vp.slack(R2, rr)));
            if (use) {
              if (nl == KLIST) flush();
              s_list[nl++][threadIdx.x] = idx;
            }
          };
          for (int k = a; k < b; k++) consider(csph[k], (uint32_t)k);
          if (s_last[rw]) flush();  // end of a distance shell: test what was listed
        }
        flush();
        continue;
      }
      if (interior) {
        for (int st = 0; st < g.nst && !covered; st++) {
          const int src = (int)cell + g.slin[st];
          const int nv = cs.cnt[src];
          const Slot<D>* sl = cs.sph + (size_t)src * cs.cap;
          for (int k = 0; k < nv && !covered; k++) {
            const Slot<D> q = sl[k];
            test(q.c, q.r, do_cov, do_child);
          }
        }
        continue;
      }
      for (int st = 0; st < g.nst && !covered; st++) {
        Nb<D> nb;
        if (!neighbour<D>(g, cc, st, nb)) continue;
        const int nv = cs.cnt[nb.src];
        const Slot<D>* sl = cs.sph + (size_t)nb.src * cs.cap;
        for (int k = 0; k < nv && !covered; k++) {
          const Slot<D> q = sl[k];
          double qp[D];
          if (nb.ghost) {
            if (!image_of<D>(g, nb, q.c, qp)) continue;
          } else {
#pragma unroll
            for (int d = 0; d < D; d++) qp[d] = q.c[d];
          }
          test(qp, q.r, do_cov, do_child);
        }
      }
    }
    code = covered ? 0u : (0x10000u | (~ccov & FULL & 0xffffu));
    out[i] = code;
  }
  if (perm == nullptr) {  // uniform: per-block counts of uncovered voxels
    const int u = __syncthreads_count(code != 0);
    if (threadIdx.x == 0) blk[blockIdx.x] = u;
  }
}

// k_cover's processing order: the voxels grouped by the cell of their centre
// (a counting sort), so that the lanes of a wave walk the same stencil cells.
// The flags do not depend on the order; the grouping only makes the walk
// uniform across lanes and the cell reads shared.
// With a prediction (pred not null, one byte per voxel, 1 = predicted covered)
// the key also puts every predicted-uncovered voxel after the predicted-covered
// ones (keys ncells .. 2 ncells - 1), so that a wave's voxels tend to share their
// outcome as well as their cells.
template <int D, bool WIDE>
__global__ void __launch_bounds__(WG) k_bin(Geo<D> g, VoxView<D> vx, int64_t nvox, VoxParams<D> vp,
                                            int* __restrict__ ccount, int* __restrict__ vcell,
                                            int* __restrict__ vrank, const uint8_t* __restrict__ pred) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool valid = i < nvox;
  int key = -1 - lane;  // distinct for lanes past the end
  if (valid) {
    double p[D];
    int cc[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int64_t c = WIDE ? vx.w[d][i] : (int64_t)vx.c[d][i];
      p[d] = vp.origin[d] + ((double)c + 0.5) * vp.len[d];
    }
    const int cell = cell_of_point<D>(g, p, cc);
    key = (cell >= 0 && cell < g.ncells) ? cell : 0;  // a key only: any value is correct
    if (pred && !pred[i]) key += g.ncells;
  }
  // one atomic per run of equal keys in adjacent lanes (siblings share a cell)
  const int prev = __shfl_up(key, 1, 64);
  const bool head = lane == 0 || prev != key;
  const uint64_t heads = __ballot(head);
  const uint64_t le = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const int start = 63 - __clzll(heads & le);
  const uint64_t above = heads & ~le;
  int base = 0;
  if (head && valid) {
    const int end = above ? __ffsll((unsigned long long)above) - 1 : 64;
    base = atomicAdd(&ccount[key], end - lane);
  }
  base = __shfl(base, start, 64);
  if (valid) {
    vcell[i] = key;
    vrank[i] = base + lane - start;
  }
}

__global__ void __launch_bounds__(WG) k_place(int64_t nvox, const int* __restrict__ vcell,
                                              const int* __restrict__ vrank, const int* __restrict__ cstart,
                                              uint32_t* __restrict__ perm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nvox) return;
  perm[cstart[vcell[i]] + vrank[i]] = (uint32_t)i;
}

// ---------------------------------------------------------------- learned predictor: voxels
//
// The 16 features of a voxel (the predictor's inputs), from its coordinates and
// the cell store only: the position of its centre in its cell, its size
// relative to the cell, the occupancy of its cell and of the cell's face
// neighbours, the distance to the nearest sphere of its own cell, the
// refinement level and two per-pass constants (placed spheres per cell, the
// last miss rate); the last feature is 0 for voxel samples.
template <int D, bool WIDE>
__device__ __forceinline__ void voxel_features(const Geo<D>& g, const CellStore<D>& cs, const VoxView<D>& vx,
                                               const VoxParams<D>& vp, int64_t i, float level, float density,
                                               float miss, float* f) {
  double p[D];
  int cc[D];
#pragma unroll
  for (int d = 0; d < D; d++) {
    const int64_t c = WIDE ? vx.w[d][i] : (int64_t)vx.c[d][i];
    p[d] = vp.origin[d] + ((double)c + 0.5) * vp.len[d];
  }
  int cell = cell_of_point<D>(g, p, cc);
  if (cell < 0 || cell >= g.ncells) cell = 0;
  f[0] = f[1] = f[2] = 0.5f;
#pragma unroll
  for (int d = 0; d < (D < 3 ? D : 3); d++) f[d] = (float)((p[d] - g.cshift[d]) * g.cinv[d] - cc[d]);
  f[3] = (float)(vp.len[0] * g.cinv[0]);
  const int own = cs.cnt[cell];
  f[4] = own * 0.125f;
  int stride = 1, k = 5;
#pragma unroll
  for (int d = 0; d < D; d++) {
    const int lo = cc[d] > 0 ? cell - stride : cell, hi = cc[d] < g.n[d] - 1 ? cell + stride : cell;
    if (k < 11) f[k++] = cs.cnt[lo] * 0.125f;
    if (k < 11) f[k++] = cs.cnt[hi] * 0.125f;
    stride *= g.n[d];
  }
  while (k < 11) f[k++] = 0.f;
  double best = 4.;
  for (int q = 0; q < own; q++) {
    const Slot<D> sl = cs.sph[(size_t)cell * cs.cap + q];
    double s2 = 0.;
#pragma unroll
    for (int d = 0; d < D; d++) {
      const double t = sl.c[d] - p[d];
      s2 += t * t;
    }
    const double rr = sl.r + vp.rmin;
    best = fmin(best, s2 / (rr * rr));
  }
  f[11] = (float)(best < 4. ? sqrt(best) : 2.);
  f[12] = level * (1.f / 32.f);
  f[13] = density;
  f[14] = miss;
  f[15] = 0.f;
}

// Trace of a refinement pass: every stride-th voxel's features and its
// outcome (covered: code 0), appended to the model's trace store.
template <int D, bool WIDE>
__global__ void __launch_bounds__(WG) k_trace_voxels(Geo<D> g, CellStore<D> cs, VoxView<D> vx, VoxParams<D> vp,
                                                     int64_t nvox, int64_t stride, const uint32_t* __restrict__ code,
                                                     float level, float density, float miss,
                                                     uint16_t* __restrict__ tx, uint8_t* __restrict__ tt,
                                                     int* __restrict__ cursor, int cap) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = k * stride;
  const bool valid = i < nvox;
  const uint64_t b = __ballot(valid);
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (lane == 0 && b) base = atomicAdd(cursor, (int)__popcll(b));
  base = __shfl(base, 0, 64);
  if (!valid) return;
  const int slot = base + (int)__popcll(b & ((1ull << lane) - 1ull));
  if (slot >= cap) return;
  float f[mlp::NF];
  voxel_features<D, WIDE>(g, cs, vx, vp, i, level, density, miss, f);
#pragma unroll
  for (int q = 0; q < mlp::NF; q++) tx[(size_t)slot * mlp::NF + q] = mlp::f2bf(f[q]);
  tt[(size_t)slot * 4 + 0] = code[i] == 0 ? 1 : 0;
  tt[(size_t)slot * 4 + 1] = 1;
  tt[(size_t)slot * 4 + 2] = 0;
  tt[(size_t)slot * 4 + 3] = 0;
}

// The prediction pass: each wave takes 64 voxels (4 tiles of 16), computes their
// features lane by lane into LDS and runs the MLP on the matrix cores;
// pred[i] = 1 when the coverage logit is positive.
constexpr int PW = WG / 64;  // waves per workgroup
template <int D, bool WIDE>
__global__ void __launch_bounds__(WG) k_predict(Geo<D> g, CellStore<D> cs, VoxView<D> vx, VoxParams<D> vp,
                                                int64_t nvox, float level, float density, float miss,
                                                const float* __restrict__ P, const uint16_t* __restrict__ Wq,
                                                uint8_t* __restrict__ pred) {
  using namespace mlp;
  __shared__ uint16_t sX[PW][64 * LDS_STRIDE];
  __shared__ uint16_t sH1[PW][16 * LDS_STRIDE], sH2[PW][16 * LDS_STRIDE];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t base = ((int64_t)blockIdx.x * PW + w) * 64;
  if (base >= nvox) return;  // whole waves only: the MFMA passes below need all 64 lanes
  Frags fr;
  load_frags(Wq, fr, lane);
  {
    float f[NF];
    const int64_t i = base + lane;
    if (i < nvox) {
      voxel_features<D, WIDE>(g, cs, vx, vp, i, level, density, miss, f);
    } else {
#pragma unroll
      for (int q = 0; q < NF; q++) f[q] = 0.f;
    }
#pragma unroll
    for (int q = 0; q < NF; q++) sX[w][lane * LDS_STRIDE + q] = f2bf(f[q]);
  }
  wave_lds_sync();
  const int gq = lane >> 4, c = lane & 15;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const floatx4 z3 = forward_tile(fr, P, sX[w] + t * 16 * LDS_STRIDE, sH1[w], sH2[w], lane);
    if (c == 0) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int64_t i = base + t * 16 + 4 * gq + j;
        if (i < nvox) pred[i] = z3[j] > 0.f ? 1 : 0;
      }
    }
    wave_lds_sync();
  }
}

// Outcome counters of a pass (learned mode or measured passes): per workgroup of
// 1024 threads x 16 voxels, the confusion counts of pred against the outcome
// (when pred is not null) and, over k_cover's processing order (perm, or the
// list order), the waves of 64 voxels and those whose voxels all share one
// outcome.  ctr: [0] predicted, [1] mispredicted, [2] covered, [3] predicted
// covered, [4] waves, [5] uniform waves.
__global__ void __launch_bounds__(1024) k_outcomes(int64_t nvox, const uint32_t* __restrict__ code,
                                                   const uint8_t* __restrict__ pred,
                                                   const uint32_t* __restrict__ perm, int* __restrict__ ctr) {
  int wrong = 0, cov = 0, pcov = 0, n = 0, waves = 0, uniform = 0;
  const int lane = threadIdx.x & 63;
  for (int r = 0; r < 16; r++) {
    const int64_t t = ((int64_t)blockIdx.x * 16 + r) * 1024 + threadIdx.x;
    const bool valid = t < nvox;
    bool covered = false;
    if (valid) {
      const int64_t i = perm ? (int64_t)perm[t] : t;
      covered = code[i] == 0;
      n++;
      cov += covered;
      if (pred) {
        pcov += pred[i] != 0;
        wrong += (pred[i] != 0) != covered;
      }
    }
    const uint64_t vb = __ballot(valid), cb = __ballot(covered);
    if (lane == 0 && vb == ~0ull) {
      waves++;
      uniform += (cb == 0 || cb == ~0ull);
    }
  }
  // the workgroup's sums, then one atomic per counter and workgroup
  __shared__ int s_sum[6];
  if (threadIdx.x < 6) s_sum[threadIdx.x] = 0;
  __syncthreads();
  const int v[6] = {pred ? n : 0, wrong, cov, pcov, waves, uniform};
#pragma unroll
  for (int k = 0; k < 6; k++) {
    int x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0 && x) atomicAdd(&s_sum[k], x);
  }
  __syncthreads();
  if (threadIdx.x < 6 && s_sum[threadIdx.x]) atomicAdd(&ctr[threadIdx.x], s_sum[threadIdx.x]);
}

// uncovered voxels per block of WG (the blocks of the list order)
__global__ void __launch_bounds__(WG) k_block_count(int64_t nvox, const uint32_t* __restrict__ code,
                                                    int* __restrict__ blk) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int u = __syncthreads_count(i < nvox && code[i] != 0);
  if (threadIdx.x == 0) blk[blockIdx.x] = u;
}

// Device-wide exclusive scan in three launches: tile sums, a scan of the
// sums (k_scan_blocks), then each tile's scan plus its offset.
constexpr int SCAN_T = 4096;  // elements per tile: 1024 threads x 4

__device__ __forceinline__ int wave_incl_scan(int x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  return x;
}

__global__ void __launch_bounds__(1024) k_tile_sums(const int* __restrict__ a, int64_t n, int* __restrict__ sums) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_T + threadIdx.x * 4;
  int s = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) s += base + k < n ? a[base + k] : 0;
  __shared__ int sh[16];
  s = wave_incl_scan(s);
  if ((threadIdx.x & 63) == 63) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int w = 0; w < 16; w++) tot += sh[w];
    sums[blockIdx.x] = tot;
  }
}

__global__ void __launch_bounds__(1024) k_tile_scan(int* __restrict__ a, int64_t n, const int* __restrict__ sums) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_T + threadIdx.x * 4;
  int v[4], s = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    v[k] = base + k < n ? a[base + k] : 0;
    s += v[k];
  }
  __shared__ int sh[16];
  const int x = wave_incl_scan(s);
  if ((threadIdx.x & 63) == 63) sh[threadIdx.x >> 6] = x;
  __syncthreads();
  int run = sums[blockIdx.x] + x - s;
  for (int w = 0; w < (int)(threadIdx.x >> 6); w++) run += sh[w];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (base + k < n) a[base + k] = run;
    run += v[k];
  }
}

// single workgroup exclusive scan over block counts; writes the total to *total
// (*total = -1 when the sum exceeds INT_MAX: the offsets would not fit)
__global__ void __launch_bounds__(1024) k_scan_blocks(int* __restrict__ blk, int nb, int* __restrict__ total) {
  __shared__ int sh[1024];
  long long carry = 0;
  for (int base = 0; base < nb; base += 1024) {
    const int t = base + threadIdx.x;
    const int v = t < nb ? blk[t] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int add = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += add;
      __syncthreads();
    }
    if (t < nb) blk[t] = (int)(carry + sh[threadIdx.x] - v);
    const int tot = sh[1023];
    __syncthreads();
    carry += tot;
  }
  if (threadIdx.x == 0) *total = carry > INT_MAX ? -1 : (int)carry;
}

// exclusive rank of element i among flagged elements: block offset + in-block prefix
__device__ __forceinline__ int block_rank(bool f, int blkoff) {
  __shared__ int wsum[WG / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t b = __ballot(f);
  const int inwave = __popcll(b & ((1ull << lane) - 1ull));
  if (lane == 0) wsum[w] = __popcll(b);
  __syncthreads();
  int before = 0;
  for (int k = 0; k < w; k++) before += wsum[k];
  __syncthreads();
  return blkoff + before + inwave;
}

// swap-with-last compaction of remove_covered (list_of_voxels.ixx:42-53) in
// closed form: with keep uncovered voxels, the covered positions below keep,
// in increasing order, receive the uncovered voxels at or above keep in
// decreasing order.  Step 1 lists the holes.
// The voxel at keep - 1 also leaves the number of holes (= voxels step 2 moves).
__global__ void __launch_bounds__(WG) k_holes(int64_t nvox, const uint32_t* __restrict__ code,
                                              const int* __restrict__ blk, int* __restrict__ ctr,
                                              int* __restrict__ holes) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < nvox;
  const bool unc = valid && code[i] != 0;
  const int U = block_rank(unc, blk[blockIdx.x]);
  const int keep = ctr[C_KEEP];
  if (valid && !unc && i < keep) holes[i - U] = (int)i;
  if (valid && i == (int64_t)keep - 1) ctr[C_MOVED] = keep - U - (unc ? 1 : 0);
}

// Step 2: the uncovered voxels at or above keep move into the holes.  Their
// codes go to `moved` (at the hole's index), not into `code`: the workgroup
// that straddles keep reads code[] below keep for its ranks while other
// workgroups fill holes, so code[] must not change during this kernel.
// Afterwards the code of voxel i < keep is code[i], or moved[i] where code[i]
// is 0 (every hole is filled).
template <int D>
__global__ void __launch_bounds__(WG) k_move(int64_t nvox, const uint32_t* __restrict__ code,
                                             uint32_t* __restrict__ moved, const int* __restrict__ blk,
                                             const int* __restrict__ ctr, const int* __restrict__ holes,
                                             VoxView<D> vx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < nvox;
  const bool unc = valid && code[i] != 0;
  const int U = block_rank(unc, blk[blockIdx.x]);
  const int keep = ctr[C_KEEP];
  if (valid && unc && i >= keep) {
    const int dst = holes[keep - U - 1];
    for (int d = 0; d < D; d++) vx.set(d, dst, vx.coord(d, i));
    moved[dst] = code[i];
  }
}

__device__ __forceinline__ uint32_t kept_code(const uint32_t* code, const uint32_t* moved, int64_t i) {
  const uint32_t c = code[i];
  return c != 0 ? c : moved[i];
}

// Exclusive prefix of one int per thread over a workgroup of WG threads, and
// the workgroup's total: a shuffle scan per wave, then the wave totals through
// LDS (one barrier).
__device__ __forceinline__ int block_excl_scan(int c, int& total) {
  __shared__ int s_w[WG / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int incl = c;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int v = __shfl_up(incl, off, 64);
    if (lane >= off) incl += v;
  }
  if (lane == 63) s_w[wv] = incl;
  __syncthreads();
  int base = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < WG / 64; w++) {
    const int x = s_w[w];
    base += w < wv ? x : 0;
    total += x;
  }
  return base + incl - c;
}

// kept_code with both words requested together (moved is read whatever code says)
__device__ __forceinline__ uint32_t kept_code_now(const uint32_t* code, const uint32_t* moved, int64_t i) {
  const uint32_t c = code[i], m = moved[i];
  return c != 0 ? c : m;
}

__global__ void __launch_bounds__(WG) k_child_count(int64_t n, const int* __restrict__ ctr,
                                                    const uint32_t* __restrict__ code,
                                                    const uint32_t* __restrict__ moved, int* __restrict__ blk) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t keep = ctr[C_KEEP];
  const uint32_t k = i < n && i < keep ? kept_code(code, moved, i) : 0u;
  const int c = __popc(k & 0xffffu);
  int total;
  block_excl_scan(c, total);
  if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

// stable append of the kept children, parents in list order, children in
// TabCorner order (list_of_voxels.ixx:76-99): coordinates 2c + tab[k]
// SW / DW: the source and destination lattices hold 64-bit coordinates (the
// launch picks the instance, so coordinate loads and stores carry no branch).
template <int D, bool SW, bool DW>
__global__ void __launch_bounds__(WG) k_scatter(int64_t n, const uint32_t* __restrict__ code,
                                                const uint32_t* __restrict__ moved, const int* __restrict__ blk,
                                                VoxView<D> src, VoxView<D> dst, int64_t nchild,
                                                int* __restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = i < n;
  const int64_t ic = in ? i : 0;
  const uint32_t m = in ? (kept_code_now(code, moved, ic) & 0xffffu) : 0u;
  int64_t pc[D];
#pragma unroll
  for (int d = 0; d < D; d++) pc[d] = SW ? src.w[d][ic] : (int64_t)src.c[d][ic];
  const int c = __popc(m);
  int total;
  const int excl = block_excl_scan(c, total);
  const int base = blk[blockIdx.x];  // the workgroup's children are base .. base + total - 1
  constexpr int NCH = ipow_c(2, D);
  constexpr uint32_t FULLM = (1u << NCH) - 1u;  // NCH <= 16
  // The children must agree with the counting pass (k_child_count + scan): the
  // workgroup's range inside the new list and masks of NCH bits.  Otherwise
  // nothing is written and the host is told (base and total are uniform).
  if (__syncthreads_or((m & ~FULLM) != 0) || (int64_t)base + total > nchild) {
    if (threadIdx.x == 0) atomicOr(err, 1);
    return;
  }
  if constexpr (!DW && D <= 3) {
    // 32-bit children staged in LDS, then stored as the workgroup's contiguous
    // range with consecutive lanes on consecutive entries
    __shared__ int32_t s_ch[D][WG * NCH];
    if (in && m != 0) {
      int o = excl;
#pragma unroll
      for (int k = 0; k < NCH; k++) {
        if (!(m & (1u << k))) continue;
#pragma unroll
        for (int d = 0; d < D; d++) s_ch[d][o] = (int32_t)(2 * pc[d] + TabCorner<D>::at(k, d));
        o++;
      }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < total; q += WG) {
#pragma unroll
      for (int d = 0; d < D; d++) const_cast<int32_t*>(dst.c[d])[base + q] = s_ch[d][q];
    }
  } else {
    if (!in || m == 0) return;
    int o = base + excl;
#pragma unroll
    for (int k = 0; k < NCH; k++) {
      if (!(m & (1u << k))) continue;
#pragma unroll
      for (int d = 0; d < D; d++) {
        const int64_t v = 2 * pc[d] + TabCorner<D>::at(k, d);
        if constexpr (DW)
          const_cast<int64_t*>(dst.w[d])[o] = v;
        else
          const_cast<int32_t*>(dst.c[d])[o] = (int32_t)v;
      }
      o++;
    }
  }
}

// initial lattice, x fastest (list_of_voxels.ixx:11-28, Loops.ixx:44-74)
template <int D>
__global__ void k_lattice(int64_t total, int n0, int n1, int n2, int n3, int32_t* c0, int32_t* c1, int32_t* c2,
                          int32_t* c3) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int nn[4] = {n0, n1, n2, n3};
  int32_t* cc[4] = {c0, c1, c2, c3};
  int64_t r = i;
  for (int d = 0; d < D; d++) {
    cc[d][i] = (int32_t)(r % nn[d]);
    r /= nn[d];
  }
}

// ============================================================== kernels: extraction

template <int D>
__global__ void k_cell_counts(Geo<D> g, const uint8_t* __restrict__ cnt, int* __restrict__ blk) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int v = 0;
  if (c < g.ncells) v = cnt[c];
  __shared__ int sh[WG];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = WG / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) blk[blockIdx.x] = sh[0];
}

template <int D, class PH>
__global__ void k_extract(Geo<D> g, CellStore<D> cs, const int* __restrict__ blk, double* __restrict__ pos,
                          double* __restrict__ rad, PH* __restrict__ ph) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int v = c < g.ncells ? cs.cnt[c] : 0;
  __shared__ int sh[WG];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < WG; off <<= 1) {
    const int add = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += add;
    __syncthreads();
  }
  if (v == 0) return;
  int o = blk[blockIdx.x] + sh[threadIdx.x] - v;
  for (int k = 0; k < v; k++, o++) {
    const Slot<D> s = cs.sph[(size_t)c * cs.cap + k];
    for (int d = 0; d < D; d++) pos[(size_t)o * D + d] = s.c[d];
    rad[o] = s.r;
    if (ph) ph[o] = (PH)cs.phase[(size_t)c * cs.cap + k];
  }
}

// Scratch of the allocation-free extraction (Engine::extract_noalloc): the
// sphere count of each block of WG cells, then their total.  One copy per GPU
// (each device holds its own instance of this array), shared by every call on
// that GPU, so calls that overlap on one GPU race.  The host reaches it only
// through the address hipGetSymbolAddress returns for the current device.
constexpr int kNoAllocBlocks = 1 << 20;  // cell grids of up to 2^28 cells (4 MiB of scratch)
static __device__ __align__(256) int scratch[kNoAllocBlocks + 64];

static int* scratch_ptr() {
  static std::mutex m;
  static int* ptr[64] = {};
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw Error("extract_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(m);
  if (!ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error(std::string("extract_noalloc: hipGetSymbolAddress of the scratch array failed: ") +
                  hipGetErrorString(e));
    ptr[dev] = static_cast<int*>(p);
  }
  return ptr[dev];
}

// The stencil tables (packed offsets, then linear offsets) from the launch arguments.
struct StencilTable {
  int n;  // entries in v: 2 * 3^D
  int v[2 * 81];
};
// The cell store at a larger per-cell capacity: slot k of cell c moves from
// c * cap0 + k to c * cap1 + k (grow_cells).
template <int D>
__global__ void k_grow_cells(int64_t total, int cap0, int cap1, const Slot<D>* __restrict__ src,
                             const uint32_t* __restrict__ srcph, Slot<D>* __restrict__ dst,
                             uint32_t* __restrict__ dstph) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int64_t c = t / cap0, k = t % cap0;
  dst[c * cap1 + k] = src[t];
  dstph[c * cap1 + k] = srcph[t];
}

__global__ void k_set_table(int* __restrict__ dst, StencilTable t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < t.n) dst[i] = t.v[i];
}

// rsa_domain::add's spheres into their cells, read from pinned host memory:
// sphere i goes to slot slot[i] of cell cell[i]; the sphere that fills a cell's
// last new slot (last[i]) also writes the cell's count.
template <int D>
__global__ void k_insert(CellStore<D> cs, int64_t n, const int* __restrict__ cell, const int* __restrict__ slot,
                         const uint8_t* __restrict__ last, const double* __restrict__ pos,
                         const double* __restrict__ rad, const uint32_t* __restrict__ ph) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Slot<D> s;
  for (int d = 0; d < D; d++) s.c[d] = pos[i * D + d];
  s.r = rad[i];
  const int c = cell[i], k = slot[i];
  cs.sph[(size_t)c * cs.cap + k] = s;
  cs.phase[(size_t)c * cs.cap + k] = ph[i];
  if (last[i]) cs.cnt[c] = (uint8_t)(k + 1);
}

// ============================================================== kernels: end-of-run checks

// A wave's count added to one counter, by one lane, only when non-zero (the
// checks almost always count nothing, so nearly no atomics are issued).
__device__ __forceinline__ void wave_count(int v, int* counter) {
  if (__ballot(v != 0) == 0) return;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(counter, v);
}

// No overlap: the reference's collision predicate (is_there_collision,
// operator_cells_and_data_storage.ixx:5-14) between every placed sphere and
// every other sphere or periodic image of its 3^D stencil -- the pairs its
// debug check tests (rsa_data_storage.ixx:229-243, asserted per draw at
// operator_algorithm.ixx:269,563).  One thread per cell; a colliding pair is
// seen from both of its spheres.
template <int D>
__global__ void __launch_bounds__(WG) k_check_overlap(Geo<D> g, CellStore<D> cs, int* __restrict__ ctr) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int hits = 0;
  const int n = c < g.ncells ? cs.cnt[c] : 0;  // ghost cells hold none
  if (n > 0) {
    int cc[D];
    cell_coords<D>(g, c, cc);
    for (int k = 0; k < n; k++) {
      const Slot<D> a = cs.sph[(size_t)c * cs.cap + k];
      for (int st = 0; st < g.nst; st++) {
        Nb<D> nb;
        if (!neighbour<D>(g, cc, st, nb)) continue;
        const int nv = cs.cnt[nb.src];
        for (int j = 0; j < nv; j++) {
          if (!nb.ghost && nb.src == c && j == k) continue;
          const Slot<D> b = cs.sph[(size_t)nb.src * cs.cap + j];
          double q[D];
          if (nb.ghost) {
            if (!image_of<D>(g, nb, b.c, q)) continue;
          } else {
#pragma unroll
            for (int d = 0; d < D; d++) q[d] = b.c[d];
          }
          if (collide<D>(a.r, a.c, b.r, q)) hits++;
        }
      }
    }
  }
  wave_count(hits, ctr + C_OVERLAP);
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {  // Steele, Lea & Flood (2014)
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Maximality: nsamp points spread over the real box (a counter-based hash, not
// the run's word stream); a point is still free when every sphere and image of
// its stencil is farther than r + rmin, the voxel method's coverage margin
// (list_of_voxels.ixx:225-238).  A maximal packing leaves none free.
template <int D>
__global__ void __launch_bounds__(WG) k_check_cover(Geo<D> g, CellStore<D> cs, int64_t nsamp, uint64_t key,
                                                    double rmin, int* __restrict__ ctr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int free_pt = 0;
  if (i < nsamp) {
    double p[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
      const double u = (double)(splitmix64(key ^ ((uint64_t)i * D + d)) >> 11) * 0x1p-53;
      p[d] = g.inf[d] + (g.sup[d] - g.inf[d]) * u;
    }
    int cc[D];
    const int64_t cell = cell_index<D>(g, p);
    cell_coords<D>(g, cell, cc);
    if (cell >= 0 && cell < g.ncells && !coords_ghost<D>(g, cc)) {
      bool covered = false;
      for (int st = 0; st < g.nst && !covered; st++) {
        Nb<D> nb;
        if (!neighbour<D>(g, cc, st, nb)) continue;
        const int nv = cs.cnt[nb.src];
        for (int j = 0; j < nv && !covered; j++) {
          const Slot<D> b = cs.sph[(size_t)nb.src * cs.cap + j];
          double q[D];
          if (nb.ghost) {
            if (!image_of<D>(g, nb, b.c, q)) continue;
          } else {
#pragma unroll
            for (int d = 0; d < D; d++) q[d] = b.c[d];
          }
          double s2 = 0.;
#pragma unroll
          for (int d = 0; d < D; d++) {
            const double t = q[d] - p[d];
            s2 += t * t;
          }
          const double rr = b.r + rmin;
          covered = s2 <= rr * rr;
        }
      }
      free_pt = covered ? 0 : 1;
    }
  }
  wave_count(free_pt, ctr + C_UNCOVERED);
}

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
  Engine(const double* ginf_, const double* gsup_, int ghost_layer, double domain_rad) {
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
    HIP_OK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
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
      (void)hipStreamDestroy(st_);
      throw;
    }
    hctr_ = static_cast<int*>(PinnedPool::get().alloc(sizeof(int) * C_NUM, hipHostMallocDefault));
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
    if (st_) (void)hipStreamDestroy(st_);
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

  uint64_t extract(double* pos, double* rad, uint64_t* phase, uint64_t cap) override {
    StreamScope scope(st_, cst_);
    if (!algo_ready_) return 0;
    const unsigned nb = nblocks(geo_.ncells);
    DevBuf<int> blk;
    blk.reserve(nb);
    hipLaunchKernelGGL(k_cell_counts<D>, dim3(nb), dim3(WG), 0, st_, geo_, cells_cnt_.p, blk.p);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st_, blk.p, (int)nb, dctr_.p + C_CURSOR);
    int total = 0;
    HIP_OK(hipMemcpyAsync(&total, dctr_.p + C_CURSOR, sizeof(int), hipMemcpyDeviceToHost, st_));
    HIP_OK(hipStreamSynchronize(st_));
    if (pos == nullptr) return (uint64_t)total;
    if ((uint64_t)total > cap) throw Error("extract: output capacity too small");
    DevBuf<double> dp, dr;
    DevBuf<uint32_t> dph;
    dp.reserve((size_t)total * D);
    dr.reserve(total);
    dph.reserve(total);
    hipLaunchKernelGGL((k_extract<D, uint32_t>), dim3(nb), dim3(WG), 0, st_, geo_, cs(), blk.p, dp.p, dr.p,
                       dph.p);
    HIP_OK(hipStreamSynchronize(st_));
    HIP_OK(hipMemcpyAsync(pos, dp.p, (size_t)total * D * sizeof(double), hipMemcpyDeviceToHost, st_));
    HIP_OK(hipStreamSynchronize(st_));
    HIP_OK(hipMemcpyAsync(rad, dr.p, (size_t)total * sizeof(double), hipMemcpyDeviceToHost, st_));
    HIP_OK(hipStreamSynchronize(st_));
    if (phase) {
      std::vector<uint32_t> ph(total);
      HIP_OK(hipMemcpyAsync(ph.data(), dph.p, (size_t)total * sizeof(uint32_t), hipMemcpyDeviceToHost, st_));
      HIP_OK(hipStreamSynchronize(st_));
      for (int i = 0; i < total; i++) phase[i] = ph[i];
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
    hipLaunchKernelGGL((k_extract<D, uint64_t>), dim3(nb), dim3(WG), 0, st_, geo_, cs(), s, pos, rad, phase);
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
  DevBuf<Slot<D>> csph_;     // and the spheres in cell order

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

  CellStore<D> cs() { return CellStore<D>{cells_cnt_.p, cells_sph_.p, cells_ph_.p, cap_}; }

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

  void alloc_cells() {
    store_rmin_ = gmin_;
    cap_ = cell_cap(gmin_, gmax_);
    cells_cnt_.reserve(geo_.ncells);
    cells_sph_.reserve((size_t)geo_.ncells * cap_);
    cells_ph_.reserve((size_t)geo_.ncells * cap_);
    cell2t_.reserve(geo_.ncells);
    cellcc_.reserve(geo_.ncells);
    ovhead_.reserve(geo_.ncells);
    HIP_OK(hipMemsetAsync(ovhead_.p, 0xff, (size_t)geo_.ncells * sizeof(int), st_));
    cslot_.reserve((size_t)geo_.ncells * CSLOT);
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
    // kernels of a draw whose gate stays closed read these before returning
    HIP_OK(hipMemsetAsync(owner_.p, 0, c * sizeof(uint8_t), st_));
    HIP_OK(hipMemsetAsync(ccell_.p, 0xff, c * sizeof(int), st_));
    HIP_OK(hipMemsetAsync(ctag_.p, T_DROPPED, c * sizeof(uint8_t), st_));
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
  static constexpr int64_t kGroupMin = int64_t(1) << 20;

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
    const bool grouped = nvox_ >= kGroupMin;
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
          hipLaunchKernelGGL((k_bin<D, true>), dim3(nb), dim3(WG), 0, st_, geo_, vx, nvox_, vp, ccount_.p, vcell_.p,
                             vrank_.p, pr);
        else
          hipLaunchKernelGGL((k_bin<D, false>), dim3(nb), dim3(WG), 0, st_, geo_, vx, nvox_, vp, ccount_.p, vcell_.p,
                             vrank_.p, pr);
      }, 1);
      scan_excl(ccount_.p, nkeys, dctr_.p + C_GROUPED);
      timed("k_place", (double)nvox_ * 12.0, [&] {
        hipLaunchKernelGGL(k_place, dim3(nb), dim3(WG), 0, st_, nvox_, vcell_.p, vrank_.p, ccount_.p, perm_.p);
      }, 1);
    }
    const uint32_t* perm = grouped ? perm_.p : nullptr;
    // the placed spheres packed in cell order (k_cover's stencil rows, D <= 3),
    // for long lists only: below kGroupMin voxels the copy costs more than it saves
    const int* coff = nullptr;
    const Slot<D>* csph = nullptr;
    int ncs = 0;
    if (ipow_c(3, D) <= 32 && nvox_ >= kGroupMin) {
      const int64_t nc = geo_.ncells;
      coff_.reserve((size_t)nc + 1);
      csph_.reserve((size_t)(nb_fast_ + added_) + 1);
      timed("k_compact", (double)nc * 9.0 + (double)(nb_fast_ + added_) * 2.0 * sizeof(Slot<D>), [&] {
        hipLaunchKernelGGL(k_cell_cnt, dim3(nblocks(nc)), dim3(WG), 0, st_, nc, cells_cnt_.p, coff_.p);
        scan_excl(coff_.p, nc, coff_.p + nc);
        hipLaunchKernelGGL(k_compact<D>, dim3(nblocks(nc)), dim3(WG), 0, st_, nc, cs(), coff_.p, csph_.p,
                           (int64_t)(nb_fast_ + added_), dctr_.p + C_COMPACT_ERR);
      }, 1);
      coff = coff_.p;
      csph = csph_.p;
      ncs = (int)std::min<uint64_t>(nb_fast_ + added_, (uint64_t)INT32_MAX);
    }
    timed("k_cover", cover_bytes, [&] {
      if (wide_[vcur_])
        hipLaunchKernelGGL((k_cover<D, true>), dim3(nb), dim3(WG), 0, st_, geo_, cs(), vx, nvox_, vp, perm, coff,
                           csph, ncs, code_.p, blk_.p);
      else
        hipLaunchKernelGGL((k_cover<D, false>), dim3(nb), dim3(WG), 0, st_, geo_, cs(), vx, nvox_, vp, perm, coff,
                           csph, ncs, code_.p, blk_.p);
    }, 1);
    if (grouped) hipLaunchKernelGGL(k_block_count, dim3(nb), dim3(WG), 0, st_, nvox_, code_.p, blk_.p);
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
      hipLaunchKernelGGL(k_holes, dim3(nb), dim3(WG), 0, st_, nvox_, code_.p, blk_.p, dctr_.p, holes_.p);
    }, 1);
    timed("k_move", 0., [&] {
      hipLaunchKernelGGL(k_move<D>, dim3(nb), dim3(WG), 0, st_, nvox_, code_.p, moved_.p, blk_.p, dctr_.p,
                         holes_.p, vx);
    }, 1);
    // children per block of the kept prefix, counted against the device's keep
    // (blocks past it count 0), so that one host round trip returns both totals
    hipLaunchKernelGGL(k_child_count, dim3(nb), dim3(WG), 0, st_, nvox_, dctr_.p, code_.p, moved_.p, blk_.p);
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
  void enqueue_first(Prep& pr, bool forced, uint64_t miss_base);
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
  DrawPub* hpub_ = nullptr;      // pinned, coherent: kPubSlots slots written by k_sweep_labels
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
typename Engine<D>::Prep Engine<D>::prep_get(uint64_t w, int64_t nv) {
  std::unique_lock<std::mutex> lk(pm_);
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
// for draws that cannot overshoot the class target: four kernels per round and
// one host round trip (the round's counts reach pinned host memory from the
// first thread of k_update_commit; the tag update and commit run while the
// host moves on).
// enqueue_first queues a draw's first round (gated unless forced);
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
  const unsigned nbs = nblocks(nshots);
  timed("k_partner_lists", (double)nshots * (8.0 * (D + 1) + 10.0), [&] {
    hipLaunchKernelGGL(k_partner_lists<D>, dim3(nbl), dim3(WG), 0, st_, geo_, nshots, cpos_.p, crad_.p, cprio_.p,
                       ccell_.p, ctag_.p, lab_.p, cell2t_.p, seg_.p, len_.p, ord_.p, slab_.p, pool_.p,
                       (int)pool_.cap, dctl_.p, gate_id, urec_.p, cellcc_.p, cslot_.p);
  });
  timed("k_sweep_labels", (double)nshots * 4.0, [&] {
    hipLaunchKernelGGL(k_sweep_labels, dim3(1), dim3(1024), (size_t)nshots, st_, nshots, ctag_.p, lab_.p, slab_.p,
                       pool_.p, dctl_.p, seq, gate_id, pr.id, aimed_, pr.miss_base, urec_.p);
  });
  timed("k_update_commit", (double)nshots * (8.0 * (D + 1) + 12.0), [&] {
    hipLaunchKernelGGL(k_update_commit<D>, dim3(nbs), dim3(WG), 0, st_, cs(), nshots, cpos_.p, crad_.p, cphase_.p,
                       ccell_.p, owner_.p, ctag_.p, lab_.p, seg_.p, len_.p, ord_.p, cell2t_.p, dctl_.p, seq,
                       hpub_dev_, gate_id >= 0 ? 1 : 0, cellcc_.p, cslot_.p, geo_.ncells);
  });
  HIP_OK(hipGetLastError());
}

template <int D>
void Engine<D>::enqueue_first(Prep& pr, bool forced, uint64_t miss_base) {
  const auto tq = std::chrono::steady_clock::now();
  const int nshots = pr.nshots;
  const uint64_t w0 = pr.w_shots;
  pr.id = ++draw_id_;
  pr.seq = ++seq_;
  pr.miss_base = miss_base;
  // k_decode_filter reads the shot words from the device ring and the chain
  // offsets from pinned host memory (written before this launch, untouched
  // until the draw's results have been published).
  host_timed("host_dev_words", [&] { dev_words(w0, w0 + (uint64_t)pr.end_rel + (uint64_t)nshots); });
  const uint32_t N = (uint32_t)nvox_;
  const ClassDev cls = class_dev();
  using LG = LaneGroup<D>;
  const unsigned nbl = nblocks((int64_t)nshots * LG::LPC);
  timed("k_decode_filter", (double)nshots * (4.0 * (2 * D + 2) + 4.0 * D + 8.0 * (D + 1) + 13.0), [&] {
    const VoxView<D> vv = vview();
    auto kern = vv.wide ? k_decode_filter<D, true> : k_decode_filter<D, false>;
    hipLaunchKernelGGL(kern, dim3(nbl), dim3(WG), 0, st_, ring_.p, kRing - 1, w0, nshots, N, hvpos_dev_[pr.buf],
                       pr.law ? hrad_dev_[pr.buf] : nullptr, pr.rej, pr.end_rel, vv, cls, geo_, cs(), cpos_.p,
                       crad_.p, cprio_.p, cphase_.p, ccell_.p, ctag_.p, owner_.p, cellcc_.p, cslot_.p, ovhead_.p,
                       next_.p, dctl_.p, pr.id, forced ? 1 : 0, list_thr_);
  });
  // the first round's partner and update kernels build the cell lists themselves
  // (k_lists runs only when a cell got more than CSLOT candidates)
  enqueue_round(pr, pr.seq, pr.id);
  if (prof_)
    prof_add("host_enqueue_draw", std::chrono::duration<double>(std::chrono::steady_clock::now() - tq).count(), 0);
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
  int seq = pr.seq;
  int64_t local_new = 0;
  int r = 0;
  // whether the cell lists of this draw are kept in seg/len/ord/cell2t: not after
  // a first round that gave up (those lists are built by its update kernel)
  bool lists_kept = false;
  auto build_lists = [&]() {  // k_lists, forced: once per draw, before the cell counts are reset
    timed("k_lists", (double)nshots * 24.0, [&] {
      hipLaunchKernelGGL(k_lists, dim3(nbs), dim3(WG), 0, st_, nshots, ccell_.p, owner_.p, ctag_.p, cellcc_.p,
                         cslot_.p, ovhead_.p, next_.p, seg_.p, len_.p, ord_.p, cell2t_.p, dctl_.p, -1, geo_.ncells);
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
      HIP_OK(hipMemsetAsync(&dctl_.p->long_list, 0, sizeof(int), st_));
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
      hipLaunchKernelGGL(k_refilter<D>, dim3(nbl), dim3(WG), 0, st_, geo_, cs(), nshots, cpos_.p, crad_.p, ccell_.p,
                         ctag_.p, cell2t_.p, seg_.p, len_.p, ord_.p);
    });
    hipLaunchKernelGGL(k_remove_skips, dim3(nbs), dim3(WG), 0, st_, nshots, owner_.p, seg_.p, len_.p, ord_.p, ctag_.p);
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
      if (!cur_live) {
        ensure_cand(cur.nshots);
        enqueue_first(cur, true, (uint64_t)(int)nb_fast_ - old_nb);
      }
      // Queue up to kQueueDepth draws behind this one, each gated on the one
      // before it.  Their inputs hold unless a draw before them triggers a
      // refinement (the device then keeps the gate closed and every later queued
      // draw returns at once); they cannot change the class or overshoot its
      // target: the shots of cur and of every queued draw sum below maximal.
      // With the draw head predicting a refinement, nothing more is queued ahead
      // (the gates would stay closed); the device still decides, as always.
      if (predicted_refine) lst_.speculation_skipped++;
      uint64_t queued_shots = nshots;
      for (int k = 0; k < live; k++) queued_shots += (uint64_t)ahead[k].nshots;
      while (live < kQueueDepth && !predicted_refine && (max_draws < 0 || st->draws + 2 + live <= max_draws)) {
        if ((int)ahead.size() == live) {
          const uint64_t wn = live == 0 ? cur.w_end() : ahead[live - 1].w_end();
          ahead.push_back(host_timed("host_prepare_next", [&] { return prep_get(wn, nvox_); }));
        }
        Prep& n = ahead[live];
        if (!(n.nshots <= 48000 && (size_t)n.nshots <= cand_cap_ && queued_shots + (uint64_t)n.nshots < maximal))
          break;  // it runs as a forced draw later
        enqueue_first(n, false, 0);  // the miss count restarts after each draw
        queued_shots += (uint64_t)n.nshots;
        live++;
      }
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
