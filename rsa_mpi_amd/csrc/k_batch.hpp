// Batched launches: one launch runs the same step for many independent
// realizations (uniform_generate<D, 1> each, operator_algorithm.ixx:275-292).
// Every realization keeps its own buffers; a launch reads one descriptor per
// realization from a device table (staged from pinned host memory by
// k_stage) and runs the single-realization kernel body on its share of the
// grid.  Draw steps give every realization the same number of workgroups
// (`bpr`, from the largest shot count of the launch), so a workgroup finds its
// realization by a division; refinement steps have very different sizes, so
// their workgroups find theirs in a prefix of block counts.
#pragma once

#include "device_common.hpp"
#include "k_common.hpp"
#include "k_draw.hpp"
#include "k_refine.hpp"

namespace rsa_amd {

constexpr int kMaxBatch = 256;  // realizations in one launch (one workgroup lane each in the prefix search)

// Everything the kernels of one draw round need for one realization.
template <int D>
struct DrawDesc {
  // k_decode_filter
  const uint32_t* ring;
  uint64_t mask, w0;
  int nshots;
  uint32_t N;
  const uint32_t* vpos;
  const double* hrad;
  int end_rel;
  VoxView<D> vx;
  ClassDev cls;
  Geo<D> g;
  CellStore<D> cs;
  double* cpos;
  double* crad;
  uint32_t* cprio;
  uint32_t* cphase;
  int* ccell;
  uint8_t* ctag;
  uint8_t* owner;
  int* cellcc;
  int* cslot;
  int* ovhead;
  int* next;
  DrawCtl* ctl;
  int draw_id, forced, list_thr;
  // k_partner_lists, k_refilter, k_remove_skips
  uint8_t* lab;
  int* cell2t;
  int* seg;
  int* len;
  int* ord;
  int* slab;
  int* pool;
  int pool_cap;
  int gate_id;
  URec* urec;
  // k_sweep_commit (also gate_id, ctl, lab ... above)
  int seq;
  double aimed;
  uint64_t miss_base;
  DrawPub* host;
  int ncells;
  int refilter;  // a conflict round: k_refilter and k_remove_skips run first
  int lists;     // a redone round: k_lists builds the cell lists first (and clears the long-list flag)
  RejList rl;
};

// Everything the kernels of one refinement need for one realization.  Long
// lists (grouped) are walked in cell-grouped order (perm) and read the placed
// spheres from a packed copy in cell order (packed), as Engine::refine does.
template <int D>
struct RefDesc {
  Geo<D> g;
  CellStore<D> cs;
  VoxView<D> vx;   // the list being refined
  VoxView<D> dst;  // its children
  VoxParams<D> vp;
  int64_t nvox;
  int64_t keep, nchild;  // known to the host after the counting passes
  uint32_t* code;
  uint32_t* moved;
  int* blk;
  int* holes;
  int* ctr;
  // cell-grouped order (null when not grouped)
  uint32_t* perm;
  int* vcell;
  int* vrank;
  int* ccount;
  // packed copy (null when not packed)
  int* coff;
  Slot<D>* csph;
  double* crad;
  int ncs;
  int64_t ncap;  // spheres the packed copy holds
};

// Block-count prefixes of a refinement launch group (entries 0 .. n; a step a
// realization skips has no blocks).
struct RefHead {
  int n;
  int nb[kMaxBatch + 1];    // k_cover: one block of WG voxels
  int nbw[kMaxBatch + 1];   // streaming passes: four waves of WG voxels per workgroup
  int nbg[kMaxBatch + 1];   // the same for grouped lists only (k_bin, k_place, k_block_count)
  int ncg[kMaxBatch + 1];   // cells of grouped lists (clearing the cell counts)
  int ncp[kMaxBatch + 1];   // cells of packed lists (k_cell_cnt, k_compact)
  int nsc[kMaxBatch + 1];   // k_scatter: one block of WG kept voxels
};

// One exclusive scan per realization: a[0, n) in place, the total to *total
// (-1 above INT_MAX), tile sums in tsum.
struct ScanDesc {
  int* a;
  int64_t n;
  int* tsum;
  int* total;
};
struct ScanHead {
  int n;
  int nt[kMaxBatch + 1];  // tiles of SCAN_T elements
};

// Staging copy from pinned (coherent, mapped) host memory into a device table:
// the table is then read from device memory by every workgroup of the launches
// that follow on the same stream.
__global__ void __launch_bounds__(WG) k_stage(const int4* __restrict__ src, int4* __restrict__ dst, int n16) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n16) dst[i] = src[i];
}

// The realization of workgroup `bid` in a prefix of block counts (pre[0] = 0,
// pre[n] = total): every thread below n tests one interval, the match is shared.
__device__ __forceinline__ int batch_find(const int* __restrict__ pre, int n, int bid) {
  __shared__ int s_r;
  for (int t = threadIdx.x; t < n; t += blockDim.x)
    if (pre[t] <= bid && bid < pre[t + 1]) s_r = t;
  __syncthreads();
  const int r = s_r;
  __syncthreads();  // s_r may be reused by a later call in the same workgroup
  return r;
}

// ---------------------------------------------------------------- draw steps

template <int D>
__global__ void __launch_bounds__(WG) kb_decode_filter(const DrawDesc<D>* __restrict__ tab, int bpr) {
  const int r = blockIdx.x / bpr, bid = blockIdx.x % bpr;
  const DrawDesc<D>& a = tab[r];
  if (a.vx.wide)
    k_decode_filter_body<D, true, true>(a.ring, a.mask, a.w0, a.nshots, a.N, a.vpos, a.hrad, a.rl, a.end_rel, a.vx, a.cls,
                                  a.g, a.cs, a.cpos, a.crad, a.cprio, a.cphase, a.ccell, a.ctag, a.owner, a.cellcc,
                                  a.cslot, a.ovhead, a.next, a.ctl, a.draw_id, a.forced, a.list_thr, bid);
  else
    k_decode_filter_body<D, false, true>(a.ring, a.mask, a.w0, a.nshots, a.N, a.vpos, a.hrad, a.rl, a.end_rel, a.vx, a.cls,
                                   a.g, a.cs, a.cpos, a.crad, a.cprio, a.cphase, a.ccell, a.ctag, a.owner, a.cellcc,
                                   a.cslot, a.ovhead, a.next, a.ctl, a.draw_id, a.forced, a.list_thr, bid);
}

template <int D>
__global__ void __launch_bounds__(WG) kb_partner_lists(const DrawDesc<D>* __restrict__ tab, int bpr) {
  const int r = blockIdx.x / bpr, bid = blockIdx.x % bpr;
  const DrawDesc<D>& a = tab[r];
  k_partner_lists_body<D>(a.g, a.nshots, a.cpos, a.crad, a.cprio, a.ccell, a.ctag, a.lab, a.cell2t, a.seg, a.len,
                          a.ord, a.slab, a.pool, a.pool_cap, a.ctl, a.seq, a.gate_id, a.urec, a.cellcc, a.cslot, bid);
}

// the sweep and the commit: blocks per realization as launched (those past a
// realization's owners and its publisher return); dynamic LDS of the largest
// shot count
template <int D>
__global__ void __launch_bounds__(kSweepThreads) kb_sweep_commit(const DrawDesc<D>* __restrict__ tab, int bpr) {
  const int r = blockIdx.x / bpr, bid = blockIdx.x % bpr;
  const DrawDesc<D>& a = tab[r];
  if (bid > (a.nshots + kOwnersPerBlock - 1) / kOwnersPerBlock) return;
  k_sweep_commit_body<D>(a.cs, a.nshots, a.cpos, a.crad, a.cphase, a.ccell, a.owner, a.ctag, a.lab, a.slab, a.pool,
                         a.urec, a.seg, a.len, a.ord, a.cell2t, a.ctl, a.seq, a.gate_id, a.draw_id, a.aimed,
                         a.miss_base, a.host, a.cellcc, a.cslot, a.ncells, bid);
}

// k_lists, forced, for the realizations that redo a round with host-built lists
template <int D>
__global__ void __launch_bounds__(WG) kb_lists(const DrawDesc<D>* __restrict__ tab, int bpr) {
  const int r = blockIdx.x / bpr, bid = blockIdx.x % bpr;
  const DrawDesc<D>& a = tab[r];
  if (!a.lists) return;
  if (bid == 0 && threadIdx.x == 0) a.ctl->long_list[a.draw_id & 1] = 0;  // no kernel of the redone round reads it
  k_lists_body(a.nshots, a.ccell, a.owner, a.ctag, a.cellcc, a.cslot, a.ovhead, a.next, a.seg, a.len, a.ord,
               a.cell2t, a.ctl, -1, a.ncells, bid);
}

template <int D>
__global__ void __launch_bounds__(WG) kb_refilter(const DrawDesc<D>* __restrict__ tab, int bpr) {
  const int r = blockIdx.x / bpr, bid = blockIdx.x % bpr;
  const DrawDesc<D>& a = tab[r];
  if (!a.refilter) return;
  k_refilter_body<D>(a.g, a.cs, a.nshots, a.cpos, a.crad, a.ccell, a.ctag, a.cell2t, a.seg, a.len, a.ord, bid);
}

template <int D>
__global__ void __launch_bounds__(WG) kb_remove_skips(const DrawDesc<D>* __restrict__ tab, int bpr) {
  const int r = blockIdx.x / bpr, bid = blockIdx.x % bpr;
  const DrawDesc<D>& a = tab[r];
  if (!a.refilter) return;
  k_remove_skips_body(a.nshots, a.owner, a.seg, a.len, a.ord, a.ctag, bid);
}

// ---------------------------------------------------------------- refinement steps

// the realization and local block of workgroup blockIdx.x in prefix `pre`
#define KB_FIND(pre)                                   \
  const int r = batch_find(pre, h->n, blockIdx.x);     \
  const RefDesc<D>& a = tab[r];                        \
  const int bid = (int)blockIdx.x - pre[r];            \
  (void)bid

template <int D>
__global__ void __launch_bounds__(WG) kb_clear_ccount(const RefHead* __restrict__ h, const RefDesc<D>* __restrict__ tab) {
  KB_FIND(h->ncg);
  const int64_t c = (int64_t)bid * WG + threadIdx.x;
  if (c < a.g.ncells) a.ccount[c] = 0;
}

// The kernels that read the voxel coordinates come in one instance per list
// width, each acting on the realizations of its width (one instance with both
// bodies would need the registers of both, and fewer waves would fit).
template <int D, bool WIDE>
__global__ void __launch_bounds__(WG) kb_bin(const RefHead* __restrict__ h, const RefDesc<D>* __restrict__ tab) {
  KB_FIND(h->nbg);
  if (a.vx.wide != (WIDE ? 1 : 0)) return;
  k_bin_body<D, WIDE>(a.g, a.vx, a.nvox, a.vp, a.ccount, a.vcell, a.vrank, nullptr, bid);
}

template <int D>
__global__ void __launch_bounds__(WG) kb_place(const RefHead* __restrict__ h, const RefDesc<D>* __restrict__ tab) {
  KB_FIND(h->nbg);
  k_place_body(a.nvox, a.vcell, a.vrank, a.ccount, a.perm, bid);
}

template <int D>
__global__ void __launch_bounds__(WG) kb_cell_cnt(const RefHead* __restrict__ h, const RefDesc<D>* __restrict__ tab) {
  KB_FIND(h->ncp);
  k_cell_cnt_body(a.g.ncells, a.cs.cnt, a.coff, bid);
}

template <int D>
__global__ void __launch_bounds__(WG) kb_compact(const RefHead* __restrict__ h, const RefDesc<D>* __restrict__ tab) {
  KB_FIND(h->ncp);
  k_compact_body<D>(a.g.ncells, a.cs, a.coff, a.csph, a.crad, a.vp, a.ncap, a.ctr + C_COMPACT_ERR, bid);
}

// k_cover over every realization's list; the realization's first workgroup
// also clears the moved-voxel count
template <int D, bool WIDE>
__global__ void __launch_bounds__(WG) kb_cover(const RefHead* __restrict__ h, const RefDesc<D>* __restrict__ tab) {
  KB_FIND(h->nb);
  if (a.vx.wide != (WIDE ? 1 : 0)) return;
  if (bid == 0 && threadIdx.x == 0) a.ctr[C_MOVED] = 0;
  k_cover_body<D, WIDE>(a.g, a.cs, a.vx, a.nvox, a.vp, a.perm, a.coff, a.csph, a.crad, a.ncs, a.code, a.blk, bid);
}

template <int D>
__global__ void __launch_bounds__(WG) kb_block_count(const RefHead* __restrict__ h, const RefDesc<D>* __restrict__ tab) {
  KB_FIND(h->nbg);
  k_block_count_body(a.nvox, a.code, a.blk, bid);
}

// the three launches of one exclusive scan per realization
__global__ void __launch_bounds__(1024) kbs_tile_sums(const ScanHead* __restrict__ h, const ScanDesc* __restrict__ t) {
  const int r = batch_find(h->nt, h->n, blockIdx.x);
  k_tile_sums_body(t[r].a, t[r].n, t[r].tsum, (int)blockIdx.x - h->nt[r]);
}

__global__ void __launch_bounds__(1024) kbs_scan_blocks(const ScanHead* __restrict__ h, const ScanDesc* __restrict__ t) {
  const int r = blockIdx.x;
  k_scan_blocks_body(t[r].tsum, h->nt[r + 1] - h->nt[r], t[r].total, 0);
}

__global__ void __launch_bounds__(1024) kbs_tile_scan(const ScanHead* __restrict__ h, const ScanDesc* __restrict__ t) {
  const int r = batch_find(h->nt, h->n, blockIdx.x);
  k_tile_scan_body(t[r].a, t[r].n, t[r].tsum, (int)blockIdx.x - h->nt[r]);
}

template <int D>
__global__ void __launch_bounds__(WG) kb_holes(const RefHead* __restrict__ h, const RefDesc<D>* __restrict__ tab) {
  KB_FIND(h->nbw);
  k_holes_body(a.nvox, a.code, a.blk, a.ctr, a.holes, bid);
}

template <int D, bool WIDE>
__global__ void __launch_bounds__(WG) kb_move(const RefHead* __restrict__ h, const RefDesc<D>* __restrict__ tab) {
  KB_FIND(h->nbw);
  if (a.vx.wide != (WIDE ? 1 : 0)) return;
  k_move_body<D, WIDE>(a.nvox, a.code, a.moved, a.blk, a.ctr, a.holes, a.vx, bid);
}

template <int D>
__global__ void __launch_bounds__(WG) kb_child_count(const RefHead* __restrict__ h, const RefDesc<D>* __restrict__ tab) {
  KB_FIND(h->nbw);
  k_child_count_body(a.nvox, a.ctr, a.code, a.moved, a.blk, bid);
}

template <int D, bool SW, bool DW>
__global__ void __launch_bounds__(WG) kb_scatter(const RefHead* __restrict__ h, const RefDesc<D>* __restrict__ tab) {
  KB_FIND(h->nsc);
  if (a.vx.wide != (SW ? 1 : 0) || a.dst.wide != (DW ? 1 : 0)) return;
  k_scatter_body<D, SW, DW>(a.keep, a.code, a.moved, a.blk, a.vx, a.dst, a.nchild, a.ctr + C_SCATTER_ERR, bid);
}

#undef KB_FIND

// every realization's counters to pinned host memory (C_NUM ints each): one
// host wait serves the whole group
template <int D>
__global__ void kb_gather_ctr(const RefDesc<D>* __restrict__ tab, int n, int* __restrict__ out) {
  const int r = blockIdx.x;
  if (r >= n) return;
  for (int k = threadIdx.x; k < C_NUM; k += blockDim.x) out[(size_t)r * C_NUM + k] = tab[r].ctr[k];
}

}  // namespace rsa_amd
