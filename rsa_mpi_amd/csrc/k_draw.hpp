// Kernels of a draw: the MT19937 word stream decoded into candidates, their
// binning in shot order, the conflict rounds (validated filter, priority sweep,
// tag update) and the commit of the survivors to their cells.
#pragma once

#include "device_common.hpp"
#include "k_common.hpp"

namespace rsa_amd {

// ============================================================== kernels: word stream

// A load issued where it is written: relaxed atomic loads are not moved past
// the branches that follow them (the compiler sinks plain loads into the block
// that uses them, behind the test that guards it).  Workgroup scope: an
// ordinary global load, for data written by earlier kernels.
template <class T>
__device__ __forceinline__ T load_now(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// A word every lane reads at the same address (a control word), issued where
// written and waited for where first used.  Read as a uniform value, the
// compiler copies it to a scalar register right after the load and so waits for
// it before any later load is issued; indexed by a zero it cannot see through,
// it stays a per-lane value.
__device__ __forceinline__ int opaque_zero() {
  int z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  return z;
}
template <class T>
__device__ __forceinline__ T load_late(const T* p) {
  return load_now(p + opaque_zero());
}

// A workgroup barrier for LDS data only: global loads in flight stay in flight
// (__syncthreads also waits for every outstanding global access).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
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
// A draw is three kernels: decode + validated filter + cell binning
// (k_decode_filter), partner lists (k_partner_lists, which builds each
// neighbour cell's ordered list from its arrival slots in the first round),
// then sweep labels + tag update + commit + the round's publication to the
// host (k_sweep_commit, whose workgroups each run the sweep and which keeps the
// cell lists for later rounds).  k_lists builds the lists instead when a cell got more
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
constexpr int kSweepThreads = 1024;  // workgroup of the sweep (a constant: reading the launch size is a load)
constexpr int kOwnersPerBlock = 256;  // cell owners per workgroup of k_sweep_commit
constexpr int kSlotsAtOnce = 4;  // slots of a stencil cell requested together by the filter (one 128-byte line in 3D)

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
  int dropped[2];  // per parity of the draw id, like the other words a draw's decode writes
  int img_mismatch;
  int seg_cursor;
  int pool_cursor;
  int pool_overflow;
  int need_pool;
  int cell_overflow;
  // Draw gate, per parity of the draw id.  The first-round kernels of draw `id`
  // run only while gate[id & 1] == id (a forced launch of k_decode_filter sets
  // it).  A draw's last sweep opens the gate for draw id + 1 when the draw is
  // over and its miss rate asks for no refinement; otherwise it closes it (-1).
  // The host may thus queue the next draw before it has seen this one's
  // results: when the gate stays closed, every kernel of that draw returns at
  // once.
  int gate[2];
  int nunres[2];  // candidates k_partner_lists left unresolved (one URec each), per parity of seq
  int acc_draw[2];   // spheres accepted so far in the draw, per parity of its id
  int long_list[2];  // a cell got more than CSLOT candidates in the draw (set by its decode), per parity
};
// One unresolved candidate as k_sweep_commit reads it, written by
// k_partner_lists at the candidate's place in the unresolved list: index,
// partner count, list offset (-1: the candidate's slab of KMAX entries, -2: the
// pool overflowed) and its first PKEEP partners (the list itself holds the rest).
struct alignas(32) URec {
  int q, m, o, pad;
  int e[PKEEP];
};
static_assert(sizeof(URec) == 32, "two 16-byte loads per record");

// A round's results in pinned host memory (seq last), written by the
// publishing workgroup of k_sweep_commit while the owners update their cells.
struct DrawPub {
  int acc, con, label_stuck, dropped, img_mismatch, pool_overflow, need_pool, cell_overflow, gate_open;
  int need_lists;  // the first round found a cell of more than CSLOT candidates: build the lists, redo it
  int seq;
};
constexpr int kPubFields = 10;
static_assert(offsetof(DrawPub, seq) == kPubFields * sizeof(int), "the fields before seq");
// Publications go to slot seq % kPubSlots: a queued next draw may publish
// before the host has read the current one.
constexpr int kPubSlots = 8;

// pick_points + generate_priority + constant-radius class (list_of_voxels.ixx:107-129,
// operator_algorithm.ixx:373-381, radius_generator.ixx:50-60); build_grid's
// cell (rsa_grid.ixx:182-187); round-0 remove_collisions_with_validated_spheres
// (operator_cells_and_data_storage.ixx:54-76: at this point only placed spheres
// and their images are validated); each candidate takes an arrival slot in its
// cell (shot order is restored in k_lists).
// The validated filter of one lane of a candidate's lane group (p, r in cell
// `cell`): whether a placed sphere of the lane's stencil cells collides with it
// (round-0 remove_collisions_with_validated_spheres,
// operator_cells_and_data_storage.ixx:54-76).  With take, the candidate's
// arrival slot is taken (into a) after the walk's first loads (a cell's count,
// then its first slots) and waited for by the caller's first use: loads and
// atomics complete in issue order for the wave's counter, so an atomic issued
// before them would hold back every wait of the walk.
template <int D>
__device__ __forceinline__ bool filter_walk(const Geo<D>& g, const CellStore<D>& cs, int64_t cell, int sub,
                                            const int* pk, const double* p, double r, bool take,
                                            int* __restrict__ cellcc, int& a) {
  using LG = LaneGroup<D>;
  bool hit = false;
  int cc[D];
  cell_coords<D>(g, cell, cc);
#pragma unroll
  for (int jj = 0; jj < LG::CPL; jj++) {
    const int s = sub * LG::CPL + jj;
    Nb<D> nb;
    const bool use = s < LG::NST && !hit && neighbour_at<D>(g, cc, pk[jj], nb);
    // slots are read only for occupied cells (most neighbour cells of a late
    // draw are empty); the first kSlotsAtOnce of them are requested together
    const int nv = use ? (int)cs.cnt[nb.src] : 0;
    const Slot<D>* sl = cs.sph + (size_t)(use ? nb.src : 0) * cs.cap;
    Slot<D> q4[kSlotsAtOnce];
#pragma unroll
    for (int k = 0; k < kSlotsAtOnce; k++)
      if (k < nv) q4[k] = sl[k];
    if (jj == 0 && take) a = atomicAdd(&cellcc[cell], 1);
    auto collides = [&](const Slot<D>& q) {
      double qp[D];
      if (nb.ghost) {
        if (!image_of<D>(g, nb, q.c, qp)) return false;
      } else {
#pragma unroll
        for (int d = 0; d < D; d++) qp[d] = q.c[d];
      }
      return collide<D>(r, p, q.r, qp);
    };
#pragma unroll
    for (int k = 0; k < kSlotsAtOnce; k++)
      if (k < nv && !hit && collides(q4[k])) hit = true;
    for (int k = kSlotsAtOnce; k < nv && !hit; k++)
      if (collides(sl[k])) hit = true;
  }
  return hit;
}

template <int D, bool WIDE, bool FILTER>
__device__ __forceinline__ void k_decode_filter_body(const uint32_t* __restrict__ ring, uint64_t mask, uint64_t w0,
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
                                                      int list_thr, const int bid) {
  using LG = LaneGroup<D>;
  const int gtid = bid * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int i = gtid / LG::LPC;
  const int sub = gtid % LG::LPC;
  const bool valid = i < nshots;
  // The gate does not change during this kernel.  It is read with the ring words
  // but tested only before the first store, so that its load overlaps the reads
  // below (all of which stay in bounds whatever the gate says).  Lanes past the
  // shots store nothing, whatever the gate.
  int gate = draw_id;
  int pk[LG::CPL] = {};  // the lane's stencil entries, read with the ring words
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
    // The gate and the stencil entries are requested here, after the shot's
    // offset: on the rejection-list path no load is then in flight, so the
    // offset's computation waits for none.  The gate is a vector load whatever
    // `forced` says (a read under that test is a scalar load waited for at once).
    const int gate_ld = load_late(&ctl->gate[draw_id & 1]);
    gate = forced ? draw_id : gate_ld;
    lane_stencil<D, LG::CPL, LG::NST>(g, sub, pk);
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
  // The arrival slot does not depend on the filter's outcome (filter_walk takes
  // it after its first loads; FILTER false: no walk).  With a closed gate
  // nothing is written (the test below returns).  forced: 0, gated on
  // gate[draw_id & 1]; 1, forced (the launch opens the gate); 2, run
  // speculatively inside the previous draw's k_sweep_commit_decode (its walk
  // may miss spheres that draw is placing, which k_filter tests; k_filter also
  // checks the gate and undoes the arrivals of a draw not taken).
  const bool take = valid && ok && sub == 0 && gate == draw_id;
  int a = -1;
  bool hit = false;
  if (valid && ok) {
    if constexpr (FILTER) {
      hit = filter_walk<D>(g, cs, cell, sub, pk, p, r, take, cellcc, a);
    } else if (take) {
      a = atomicAdd(&cellcc[cell], 1);
    }
  }
  const bool anyhit = (__ballot(hit) & LG::mask(lane)) != 0;
  if (gate != draw_id) return;  // whole grid
  if (gtid == 0) {
    ctl->acc_draw[draw_id & 1] = 0;
    if (forced == 1) ctl->gate[draw_id & 1] = draw_id;
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
      atomicAdd(&ctl->dropped[draw_id & 1], 1);
    } else {
      ccell[i] = (int)cell;
      ctag[i] = FILTER && anyhit ? T_SKIP : T_UNDECIDED;
      if (a < CSLOT)
        cslot[(size_t)cell * CSLOT + a] = i;
      else
        next[i] = atomicExch(&ovhead[cell], i);
      // more than list_thr (<= CSLOT) candidates in a cell: the first round's
      // kernels give up and the host builds the lists (k_lists) and redoes it
      if (a == list_thr) atomicOr(&ctl->long_list[draw_id & 1], 1);
      owner[i] = a == 0;
    }
  }
}

template <int D, bool WIDE, bool FILTER>
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
  k_decode_filter_body<D, WIDE, FILTER>(ring, mask, w0, nshots, N, vpos, hrad, rl, end_rel, vx, cls, g, cs, cpos, crad, cprio, cphase, ccell, ctag, owner, cellcc, cslot, ovhead, next, ctl, draw_id, forced, list_thr, (int)blockIdx.x);
}

// A first round's validated filter, completed: the decode ran its walk while
// the draw before it committed (k_sweep_commit_decode), so it saw every sphere
// placed before that draw and may have missed the ones that draw placed (a
// count read before the commit, or slots not yet written: they hold NaN, which
// collides with nothing).  Here each candidate not yet Skip is tested against
// the spheres the draw before it appended to its stencil cells (the cells whose
// tag names that draw, slots from the count before it); a hit makes it Skip.
// Testing more spheres than missed is harmless: they are all placed.  One lane
// per stencil cell of each candidate, as in k_decode_filter.  When the draw's
// gate is closed (the decode ran speculatively and the draw before it ended
// otherwise), the decode's arrivals are undone instead: every cell a candidate
// arrived in gets its count and overflow list reset, and the draw's per-parity
// words are cleared, so the buffers serve the draw after next.
template <int D>
__global__ void __launch_bounds__(WG) k_filter(Geo<D> g, CellStore<D> cs, int nshots, const double* __restrict__ cpos,
                                               const double* __restrict__ crad, const int* __restrict__ ccell,
                                               uint8_t* __restrict__ ctag, int* __restrict__ cellcc,
                                               int* __restrict__ ovhead, DrawCtl* __restrict__ ctl, int draw_id) {
  using LG = LaneGroup<D>;
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int i = gtid / LG::LPC;
  const int sub = gtid % LG::LPC;
  const bool valid = i < nshots;
  const int ic = valid ? i : 0;
  // every load issued before any test (indices clamped: in bounds whatever they hold)
  const int gate = load_late(&ctl->gate[draw_id & 1]);
  int pk[LG::CPL];
  lane_stencil<D, LG::CPL, LG::NST>(g, sub, pk);
  const int c = load_now(ccell + ic);
  const uint8_t t = load_now(ctag + ic);
  double p[D];
#pragma unroll
  for (int d = 0; d < D; d++) p[d] = load_now(cpos + (size_t)d * nshots + ic);
  const double r = load_now(crad + ic);
  const bool live = gate == draw_id;
  const bool in_cell = valid && c >= 0 && c < g.ncells;
  const int prev = (draw_id - 1) & 0x7fffff;  // the draw whose commit ran beside the decode
  bool hit = false;
  if (live && in_cell && t == T_UNDECIDED) {
    int cc[D];
    cell_coords<D>(g, c, cc);
#pragma unroll
    for (int jj = 0; jj < LG::CPL; jj++) {
      const int s = sub * LG::CPL + jj;
      Nb<D> nb;
      if (s >= LG::NST || hit || !neighbour_at<D>(g, cc, pk[jj], nb)) continue;
      const int m = load_now(cs.mod + nb.src);
      const int nv = cs.cnt[nb.src];
      if ((m >> 8) != prev) continue;
      const Slot<D>* sl = cs.sph + (size_t)nb.src * cs.cap;
      for (int k = m & 0xff; k < nv && !hit; k++) {
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
  if (live) {
    if (in_cell && sub == 0 && anyhit) ctag[i] = T_SKIP;
    return;
  }
  if (in_cell && sub == 0) {
    cellcc[c] = 0;
    ovhead[c] = -1;
  }
  if (gtid == 0) {
    ctl->long_list[draw_id & 1] = 0;
    ctl->dropped[draw_id & 1] = 0;
  }
}

// The list of a cell of n <= CSLOT candidates as k_lists builds it: the arrival
// slots v (n entries) in shot order (a sorting network), then the round-0 Skips
// removed from the back with swap-with-last (rsa_data_storage.ixx:108-116,
// 305-315).  The list is left in v; returns its length.  self: any valid
// candidate index (the tag loads of unused entries read it, without a branch).
// The first n arrival slots in shot order (entries n.. become INT_MAX):
// 19-comparator network for 8 keys (Batcher odd-even merge).
// 63-comparator network for 16 keys (Batcher odd-even merge; it sorts every
// 0-1 input, hence every input)
constexpr int kSortNet16[63][2] = {
    {0, 1},  {2, 3},   {0, 2},   {1, 3},   {1, 2},  {4, 5},   {6, 7},  {4, 6},   {5, 7},   {5, 6},   {0, 4},
    {2, 6},  {2, 4},   {1, 5},   {3, 7},   {3, 5},  {1, 2},   {3, 4},  {5, 6},   {8, 9},   {10, 11}, {8, 10},
    {9, 11}, {9, 10},  {12, 13}, {14, 15}, {12, 14}, {13, 15}, {13, 14}, {8, 12}, {10, 14}, {10, 12}, {9, 13},
    {11, 15}, {11, 13}, {9, 10}, {11, 12}, {13, 14}, {0, 8},  {4, 12}, {4, 8},   {2, 10},  {6, 14},  {6, 10},
    {2, 4},  {6, 8},   {10, 12}, {1, 9},   {5, 13}, {5, 9},   {3, 11}, {7, 15},  {7, 11},  {3, 5},   {7, 9},
    {11, 13}, {1, 2},  {3, 4},   {5, 6},   {7, 8},  {9, 10},  {11, 12}, {13, 14}};
constexpr int kSortNet[19][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6}, {5, 7}, {1, 2}, {5, 6},
                                 {0, 4}, {3, 7}, {1, 5}, {2, 6}, {1, 4}, {3, 6}, {2, 4}, {3, 5}, {3, 4}};
__device__ __forceinline__ void sort_slots(int n, int (&v)[CSLOT]) {
#pragma unroll
  for (int k = 0; k < CSLOT; k++)
    if (k >= n) v[k] = INT_MAX;
#pragma unroll
  for (int e = 0; e < 19; e++) {
    const int x = v[kSortNet[e][0]], y = v[kSortNet[e][1]];
    v[kSortNet[e][0]] = min(x, y);
    v[kSortNet[e][1]] = max(x, y);
  }
}
// the same network, each entry's tag moved with it (the first n entries are
// distinct candidates; the padding entries' tags are never read)
__device__ __forceinline__ void sort_slots(int n, int (&v)[CSLOT], uint8_t (&t)[CSLOT]) {
#pragma unroll
  for (int k = 0; k < CSLOT; k++)
    if (k >= n) v[k] = INT_MAX;
#pragma unroll
  for (int e = 0; e < 19; e++) {
    const int a = kSortNet[e][0], b = kSortNet[e][1];
    const int x = v[a], y = v[b];
    const uint8_t tx = t[a], ty = t[b];
    const bool sw = y < x;
    v[a] = sw ? y : x;
    v[b] = sw ? x : y;
    t[a] = sw ? ty : tx;
    t[b] = sw ? tx : ty;
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
__device__ __forceinline__ void k_lists_body(int nshots, const int* __restrict__ ccell,
                                              const uint8_t* __restrict__ owner, const uint8_t* __restrict__ ctag,
                                              int* __restrict__ cellcc, const int* __restrict__ cslot,
                                              int* __restrict__ ovhead, const int* __restrict__ next,
                                              int* __restrict__ seg, int* __restrict__ len, int* __restrict__ ord,
                                              int* __restrict__ cell2t, DrawCtl* __restrict__ ctl, int draw_id,
                                              int ncells, const int bid) {
  const int gate = load_now(&ctl->gate[draw_id & 1]);  // tested before the first store (loads overlap)
  const int i = bid * blockDim.x + threadIdx.x;
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
  } else if (n <= 2 * CSLOT) {
    // in registers too (a long list of a clustered late draw): the fixed slots,
    // then the overflow list (the count says how long it is), a 63-comparator
    // network (Batcher odd-even merge for 16 keys), the tags together, and the
    // round-0 Skips removed from the back with swap-with-last
    constexpr int L2S = 2 * CSLOT;
    int v[L2S] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
#pragma unroll
    for (int q = CSLOT; q < L2S; q++) v[q] = INT_MAX;
    int j = ovhead[c];
    for (int k = CSLOT; k < n && j != -1; k++) {
#pragma unroll
      for (int q = CSLOT; q < L2S; q++) v[q] = k == q ? j : v[q];
      j = next[j];
    }
#pragma unroll
    for (int e = 0; e < 63; e++) {
      const int x = v[kSortNet16[e][0]], y = v[kSortNet16[e][1]];
      v[kSortNet16[e][0]] = min(x, y);
      v[kSortNet16[e][1]] = max(x, y);
    }
    uint8_t tg[L2S];
#pragma unroll
    for (int q = 0; q < L2S; q++) tg[q] = ctag[(q < n && (unsigned)v[q] < (unsigned)nshots) ? v[q] : i];
#pragma unroll
    for (int q = L2S - 1; q >= 0; q--) {
      if (q < n && tg[q] == T_SKIP) {
        int x = v[q];
#pragma unroll
        for (int k = q + 1; k < L2S; k++) x = (m - 1 == k) ? v[k] : x;
        v[q] = x;
        m--;
      }
    }
#pragma unroll
    for (int q = 0; q < L2S; q++)
      if (q < m) a[q] = v[q];
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

__global__ void __launch_bounds__(WG) k_lists(int nshots, const int* __restrict__ ccell,
                                              const uint8_t* __restrict__ owner, const uint8_t* __restrict__ ctag,
                                              int* __restrict__ cellcc, const int* __restrict__ cslot,
                                              int* __restrict__ ovhead, const int* __restrict__ next,
                                              int* __restrict__ seg, int* __restrict__ len, int* __restrict__ ord,
                                              int* __restrict__ cell2t, DrawCtl* __restrict__ ctl, int draw_id,
                                              int ncells) {
  k_lists_body(nshots, ccell, owner, ctag, cellcc, cslot, ovhead, next, seg, len, ord, cell2t, ctl, draw_id, ncells, (int)blockIdx.x);
}

// Rounds >= 1 filter (operator_cells_and_data_storage.ixx:54-76): placed
// spheres, spheres validated by earlier rounds of this draw, and their images.
template <int D>
__device__ __forceinline__ void k_refilter_body(Geo<D> g, CellStore<D> cs, int nshots,
                                                 const double* __restrict__ cpos, const double* __restrict__ crad,
                                                 const int* __restrict__ ccell, uint8_t* __restrict__ ctag,
                                                 const int* __restrict__ cell2t, const int* __restrict__ seg,
                                                 const int* __restrict__ len, const int* __restrict__ ord, const int bid) {
  using LG = LaneGroup<D>;
  const int gtid = bid * blockDim.x + threadIdx.x;
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

template <int D>
__global__ void __launch_bounds__(WG) k_refilter(Geo<D> g, CellStore<D> cs, int nshots,
                                                 const double* __restrict__ cpos, const double* __restrict__ crad,
                                                 const int* __restrict__ ccell, uint8_t* __restrict__ ctag,
                                                 const int* __restrict__ cell2t, const int* __restrict__ seg,
                                                 const int* __restrict__ len, const int* __restrict__ ord) {
  k_refilter_body<D>(g, cs, nshots, cpos, crad, ccell, ctag, cell2t, seg, len, ord, (int)blockIdx.x);
}

__device__ __forceinline__ void k_remove_skips_body(int nshots, const uint8_t* __restrict__ owner, const int* __restrict__ seg,
                               int* __restrict__ len, int* __restrict__ ord, const uint8_t* __restrict__ ctag, const int bid) {
  const int i = bid * blockDim.x + threadIdx.x;
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

__global__ void k_remove_skips(int nshots, const uint8_t* __restrict__ owner, const int* __restrict__ seg,
                               int* __restrict__ len, int* __restrict__ ord, const uint8_t* __restrict__ ctag) {
  k_remove_skips_body(nshots, owner, seg, len, ord, ctag, (int)blockIdx.x);
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
    // entry j of the list (tag, priority, radius, position), in detection order
    auto consider = [&](int j, uint8_t tj, uint32_t pj, double rj, const double* jp) {
      if (tj != T_UNDECIDED) return;
      if (!(pj < pi)) return;
      double qp[D];
      if (nb.ghost) {
        if (!image_of<D>(g, nb, jp, qp)) return;
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
    };
    int k0 = 0;
    if (in_row) {
      // the first kSlotsAtOnce entries are requested together (entries past the
      // list read candidate `self`, in bounds, and are not considered)
      int je[kSlotsAtOnce];
      uint8_t te[kSlotsAtOnce];
      uint32_t pe[kSlotsAtOnce];
      double re[kSlotsAtOnce], xe[kSlotsAtOnce][D];
#pragma unroll
      for (int k = 0; k < kSlotsAtOnce; k++) {
        je[k] = k < ln ? v[k] : self;
        te[k] = load_now(ctag + je[k]);
        pe[k] = load_now(cprio + je[k]);
        re[k] = load_now(crad + je[k]);
#pragma unroll
        for (int d = 0; d < D; d++) xe[k][d] = load_now(cpos + (size_t)d * nshots + je[k]);
      }
#pragma unroll
      for (int k = 0; k < kSlotsAtOnce; k++)
        if (k < ln) consider(je[k], te[k], pe[k], re[k], xe[k]);
      k0 = kSlotsAtOnce;
    }
    for (int k = k0; k < ln; k++) {
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
      double jp[D];
#pragma unroll
      for (int d = 0; d < D; d++) jp[d] = load_now(cpos + (size_t)d * nshots + j);
      consider(j, tj, pj, rj, jp);
    }
  }
  return n;
}

// Sweep labels, first step (see k_sweep_commit for the semantics): candidates without
// partners stay Undecided, those whose first partner is an image become
// Conflict; the others get their ordered partner list (slab of KMAX entries,
// longer lists from a pool) and are left unresolved for k_sweep_commit.  Every
// other candidate's label is Dropped, so that the sweep reads labels only.
template <int D>
__device__ __forceinline__ void k_partner_lists_body(Geo<D> g, int nshots, const double* __restrict__ cpos,
                                                 const double* __restrict__ crad, const uint32_t* __restrict__ cprio,
                                                 const int* __restrict__ ccell, const uint8_t* __restrict__ ctag,
                                                 uint8_t* __restrict__ lab, const int* __restrict__ cell2t,
                                                 const int* __restrict__ seg, const int* __restrict__ len,
                                                 const int* __restrict__ ord, int* __restrict__ slab,
                                                 int* __restrict__ pool, int pool_cap, DrawCtl* ctl, int seq,
                                                 int gate_id, URec* __restrict__ urec,
                                                 const int* __restrict__ cellcc, const int* __restrict__ cslot,
                                                 const int bid) {
  using LG = LaneGroup<D>;
  // A draw's first round (gate_id >= 0) builds the cell lists from the arrival
  // slots; the later ones read the lists k_sweep_commit kept.
  const bool first = gate_id >= 0;
  // tested before the first store: read whatever `first` says, so that the loads overlap
  const int gate_ld = load_late(&ctl->gate[gate_id & 1]), long_ld = load_late(&ctl->long_list[gate_id & 1]);
  const int gate = first ? gate_ld : gate_id;
  const int long_list = first ? long_ld : 0;
  const int gtid = bid * blockDim.x + threadIdx.x;
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
  if (gate != gate_id) {  // whole grid: the round is not taken
    // its sweep will not run either, so the unresolved count of the next round
    // (the parity of the round before this one) is cleared here
    if (gtid == 0) ctl->nunres[(seq + 1) & 1] = 0;
    return;
  }
  if (long_list) return;  // a cell of more than CSLOT candidates: the host builds the lists and redoes the round
  if (!active && sub == 0 && i < nshots) lab[i] = T_DROPPED;
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
      u = atomicAdd(&ctl->nunres[seq & 1], 1);
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

template <int D>
__global__ void __launch_bounds__(WG) k_partner_lists(Geo<D> g, int nshots, const double* __restrict__ cpos,
                                                 const double* __restrict__ crad, const uint32_t* __restrict__ cprio,
                                                 const int* __restrict__ ccell, const uint8_t* __restrict__ ctag,
                                                 uint8_t* __restrict__ lab, const int* __restrict__ cell2t,
                                                 const int* __restrict__ seg, const int* __restrict__ len,
                                                 const int* __restrict__ ord, int* __restrict__ slab,
                                                 int* __restrict__ pool, int pool_cap, DrawCtl* ctl, int seq,
                                                 int gate_id, URec* __restrict__ urec,
                                                 const int* __restrict__ cellcc, const int* __restrict__ cslot) {
  k_partner_lists_body<D>(g, nshots, cpos, crad, cprio, ccell, ctag, lab, cell2t, seg, len, ord, slab, pool, pool_cap,
                          ctl, seq, gate_id, urec, cellcc, cslot, (int)blockIdx.x);
}

// The sequential priority sweep (find_solve_conflicts,
// operator_cells_and_data_storage.ixx:127-142) as a dependency evaluation, then
// update_sphere_tags and the commit (operator_algorithm.ixx:446-473, 383-443),
// in one kernel of kSweepThreads-thread workgroups.
//
// Sweep: a candidate's label is fixed by its first partner (detection order)
// that is not Skip -- an image or a Conflict partner gives Conflict, an
// Undecided real partner gives Skip, none leaves it Undecided.  Partners have
// strictly smaller priorities, so in the reference's sorted sweep they are
// labelled first; the passes reach the same labels.  Every workgroup runs the
// whole sweep over the round's labels in its LDS (nshots bytes of dynamic
// shared memory): the labels are a unique fixed point, so all reach the same
// ones, and none waits for another.
//
// Update and commit: one thread per cell owner, the first kOwnersPerBlock
// threads of each workgroup (candidate bid * kOwnersPerBlock + thread; the
// owners' gathers are spread over as many compute units as kernels of WG
// threads would use): Skip removed from the back, Undecided -> Validated, Conflict ->
// Undecided; when the round left no Conflict the draw is over and the
// survivors are appended to their cells in tail order.  Owners touch disjoint
// cells and candidates.
//
// The workgroup after the candidates' (the launch has one more) publishes the
// round: the control words for the next kernels and the counts to pinned host
// memory (seq last), its system-scope fence delaying no owner.  Control words
// that a workgroup reads are never written in the same round: the gate and the
// unresolved count are kept per parity (of the draw id and of seq), the
// publisher writing the next draw's gate and clearing the next round's count
// (also when the gate kept this round from running: that count was the
// previous round's).
template <int D>
__device__ __forceinline__ void k_sweep_commit_body(
    CellStore<D> cs, int nshots, const double* __restrict__ cpos, const double* __restrict__ crad,
    const uint32_t* __restrict__ cphase, const int* __restrict__ ccell, const uint8_t* __restrict__ owner,
    uint8_t* __restrict__ ctag, const uint8_t* __restrict__ lab, const int* __restrict__ slab,
    const int* __restrict__ pool, const URec* __restrict__ urec, int* __restrict__ seg, int* __restrict__ len,
    int* __restrict__ ord, int* __restrict__ cell2t, DrawCtl* __restrict__ ctl, int seq, int gate_id, int draw_id,
    double aimed, uint64_t miss_base, DrawPub* host, int* __restrict__ cellcc, const int* __restrict__ cslot,
    int ncells, const int bid) {
  extern __shared__ __align__(16) uint8_t s_lab[];
  __shared__ int s_prog, s_left, s_acc, s_con, s_stuck;
  const int tid = threadIdx.x;
  // a draw's first round (gate_id >= 0) built the cell lists from the arrival slots
  const bool first = gate_id >= 0;
  const int pub_bid = (nshots + kOwnersPerBlock - 1) / kOwnersPerBlock;
  const bool pub = bid == pub_bid;
  // Control words: every thread reads them, whatever the tests on them say (a
  // read under a test or a || is waited for at once); stores come only once the
  // gate is known open.
  const int gate_ld = load_late(&ctl->gate[gate_id & 1]), povf_ld = load_late(&ctl->pool_overflow);
  const int long_ld = load_late(&ctl->long_list[gate_id & 1]), nu_ld = load_late(&ctl->nunres[seq & 1]);
  const int gate = first ? gate_ld : gate_id;
  const bool ovf = povf_ld != 0;
  // a first round that met a cell of more than CSLOT candidates did nothing: the
  // host builds the lists and redoes it (like a pool overflow)
  const bool longl = first && long_ld != 0;
  const bool redo_l = ovf || longl;
  // the publisher's counters: no kernel of this round changes them after
  // k_partner_lists, so they are read now, with the loads above
  int dropped_ld = 0, mism_ld = 0, need_ld = 0, cell_ld = 0, accd_ld = 0;
  if (pub) {
    dropped_ld = load_late(&ctl->dropped[draw_id & 1]);
    mism_ld = load_late(&ctl->img_mismatch);
    need_ld = load_late(&ctl->need_pool);
    cell_ld = load_late(&ctl->cell_overflow);
    accd_ld = load_late(&ctl->acc_draw[draw_id & 1]);
  }
  // The owner's inputs, read without a branch (index clamped to candidate 0):
  // its owner flag and cell, and its list -- in a draw's first round the cell's
  // count and arrival slots (cell clamped into the grid), later the list kept by
  // the previous round (offset, length, and the row at i * CSLOT where a list of
  // at most CSLOT entries starts) -- and the cell's sphere count, all in bounds
  // whatever they hold.
  // Those that depend on the owner's index only are issued here, the ones that
  // depend on its cell after the sweep's loads: the sweep's barriers wait for
  // LDS only, so the owner's loads stay in flight across them.
  // (Both rounds' inputs are read, without a branch, and chosen only where the
  // update begins: a value loaded on one side of a branch is waited for where
  // the branches meet.)
  const int i = bid * kOwnersPerBlock + tid;
  const bool in = !pub && tid < kOwnersPerBlock && i < nshots;
  const int ic = in ? i : 0;
  const uint8_t own = load_now(owner + ic);
  const bool on = in && own;
  const int c = load_now(ccell + ic);
  const int sg_l = load_now(seg + ic), n0_l = load_now(len + ic);
  const int64_t* row_l = reinterpret_cast<const int64_t*>(ord + (size_t)ic * CSLOT);
  const int64_t l0 = load_now(row_l), l1 = load_now(row_l + 1), l2 = load_now(row_l + 2), l3 = load_now(row_l + 3);
  if (tid == 0) {
    s_acc = 0;
    s_con = 0;
    s_stuck = 0;
  }
  // the unresolved records a thread keeps (see below), requested with the labels
  constexpr int NCR = 2;
  int4 uh[NCR], ue[NCR];
#pragma unroll
  for (int c = 0; c < NCR; c++) {
    const int u = tid + c * kSweepThreads;
    const int4* rc = reinterpret_cast<const int4*>(urec + (u < nshots ? u : 0));
    uh[c] = rc[0];
    ue[c] = rc[1];
  }
  // the round's labels into LDS, 8 candidates per thread and load (k_partner_lists
  // wrote one for every candidate; the buffer starts 256-byte aligned)
  for (int q8 = tid * 8; q8 < nshots; q8 += 8 * kSweepThreads) {
    if (q8 + 8 <= nshots) {
      *reinterpret_cast<uint2*>(s_lab + q8) = *reinterpret_cast<const uint2*>(lab + q8);
    } else {
      for (int q = q8; q < nshots; q++) s_lab[q] = lab[q];
    }
  }
  // each thread keeps up to two unresolved candidates and their first PKEEP
  // partners in registers for all passes; further ones are re-read per pass.
  // The records are read (below) with the labels, whatever the count says
  // (index clamped below nshots), and used below the count only.
  constexpr int NC = 2;
  int cq[NC], cm[NC];
  int ce[NC][PKEEP];
  const int* cpl[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const int u = tid + c * kSweepThreads;
    const int4 h = uh[c], e = ue[c];
    const bool use = !redo_l && u < nu_ld;
    cq[c] = use ? h.x : -1;
    cm[c] = use ? h.y : 0;
    cpl[c] = use ? (h.z == -1 ? slab + (size_t)h.x * KMAX : pool + h.z) : nullptr;
    ce[c][0] = e.x;
    ce[c][1] = e.y;
    ce[c][2] = e.z;
    ce[c][3] = e.w;
  }
  // the owner's cell: its count and arrival slots (a first round's list), and its sphere count
  const int cl = (c >= 0 && c < ncells) ? c : 0;
  const int n0_f = load_now(cellcc + cl);
  const int64_t* row_f = reinterpret_cast<const int64_t*>(cslot + (size_t)cl * CSLOT);
  const int64_t f0 = load_now(row_f), f1 = load_now(row_f + 1), f2 = load_now(row_f + 2), f3 = load_now(row_f + 3);
  const int cnt0 = load_now(cs.cnt + cl);  // an owner's cell is in the grid (cl == c)
  // The owner's list and its entries' tags, requested before the sweep's
  // barrier (the labels come from the sweep): a first round's arrival slots in
  // arrival order, or the list the previous round kept at i * CSLOT.  Only an
  // owner of a running round with its list in registers reads its entries; the
  // other reads (entries past the list too) take candidate ic's tag, and every
  // index is kept below nshots.
  const int sg = first ? ic * CSLOT : sg_l;
  const int n0 = first ? n0_f : n0_l;
  const int64_t q0 = first ? f0 : l0, q1 = first ? f1 : l1, q2 = first ? f2 : l2, q3 = first ? f3 : l3;
  int ve[CSLOT] = {(int)q0, (int)(q0 >> 32), (int)q1, (int)(q1 >> 32),
                   (int)q2, (int)(q2 >> 32), (int)q3, (int)(q3 >> 32)};
  const bool reg = on && gate == gate_id && (first || (sg == i * CSLOT && n0 <= CSLOT));
  uint8_t te[CSLOT];
#pragma unroll
  for (int k = 0; k < CSLOT; k++) {
    const int j = (reg && k < n0 && (unsigned)ve[k] < (unsigned)nshots) ? ve[k] : ic;
    te[k] = load_now(ctag + j);
  }
  // the same for the whole workgroup: scalar values from here on (their waits
  // come after the loads above are issued)
  const int nu = __builtin_amdgcn_readfirstlane(nu_ld);
  const bool redo = __builtin_amdgcn_readfirstlane(redo_l ? 1 : 0) != 0;
  const bool ran = __builtin_amdgcn_readfirstlane(gate == gate_id ? 1 : 0) != 0;
  lds_barrier();
  if (!ran) {  // the whole grid
    if (pub && tid == 0) ctl->nunres[(seq + 1) & 1] = 0;
    return;
  }
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
      lds_barrier();
      int prog = 0, left = 0;
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const int q = cq[c];
        if (q < 0 || s_lab[q] != T_UNRESOLVED) continue;
        uint8_t res;
        if (resolve(cm[c], cpl[c], ce[c], res)) {
          s_lab[q] = res;
          prog++;
        } else {
          left++;
        }
      }
      for (int u = tid + NC * kSweepThreads; u < nu; u += kSweepThreads) {
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
          prog++;
        } else {
          left++;
        }
      }
      if (prog) atomicAdd(&s_prog, prog);
      if (left) atomicAdd(&s_left, left);
      lds_barrier();
      const int sp = s_prog, sl = s_left;
      lds_barrier();
      if (sl == 0) break;
      if (sp == 0) {
        if (tid == 0) s_stuck = 1;
        break;
      }
    }
    // counts over 8 labels per thread
    int acc = 0, con = 0;
    for (int q8 = tid * 8; q8 < nshots; q8 += 8 * kSweepThreads) {
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
    if (pub) {
      // both counts in one word (at most 8 per thread and 64 threads: no carry
      // into the high half), summed over the wave
      int both = acc + (con << 16);
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) both += __shfl_down(both, off, 64);
      if ((tid & 63) == 0 && both) {
        atomicAdd(&s_acc, both & 0xffff);
        atomicAdd(&s_con, both >> 16);
      }
    } else if (__ballot(con != 0) != 0 && (tid & 63) == 0) {
      s_con = 1;  // the owners need only whether some candidate is in Conflict
    }
  }
  lds_barrier();
  const int acc = s_acc, con = s_con, stuck = s_stuck;
  if (pub) {
    if (tid != 0) return;
    // counters whose values go into the publication below, cleared for the next round
    ctl->dropped[draw_id & 1] = 0;
    ctl->img_mismatch = 0;
    ctl->seg_cursor = 0;
    ctl->nunres[(seq + 1) & 1] = 0;
    if (!ovf) ctl->pool_cursor = 0;
    // the draw's refinement decision, as the host makes it after the draw
    // (operator_algorithm.ixx:170-180): miss rate = misses / (1e-6 + shots),
    // misses = shots - (spheres now - spheres after the previous draw of this
    // proceed) in unsigned 64-bit arithmetic; miss_base carries the part of the
    // difference placed before the draw (non-zero on a proceed's first draw).
    int open = 0;
    const int acc_draw = redo ? accd_ld : accd_ld + acc;
    ctl->acc_draw[draw_id & 1] = acc_draw;
    if (!redo && !stuck && con == 0) {
      const uint64_t miss = (uint64_t)nshots - (uint64_t)acc_draw - miss_base;
      const double rate = (double)miss / (0.000001 + (double)(uint64_t)nshots);
      open = !(rate > aimed);
    }
    ctl->gate[(draw_id + 1) & 1] = open ? draw_id + 1 : -1;
    // the round's results to the host (seq last)
    const int v[kPubFields] = {acc,     con,    stuck,   dropped_ld, mism_ld,
                               povf_ld, need_ld, cell_ld, open,       longl ? 1 : 0};
    DrawPub* hp = host + (seq & (kPubSlots - 1));
#pragma unroll
    for (int k = 0; k < kPubFields; k++) __atomic_store_n(&reinterpret_cast<int*>(hp)[k], v[k], __ATOMIC_RELAXED);
    __threadfence_system();
    __atomic_store_n(&hp->seq, seq, __ATOMIC_RELEASE);
    return;
  }
  const bool skip_round = ovf || stuck || longl;
  const bool commit = con == 0;
  if (!on || skip_round) return;  // no owner, or a round the host redoes
  if (first || (sg == i * CSLOT && n0 <= CSLOT)) {
    // the list in registers: its tags are loaded together, the removal (same
    // steps as the loop below) works on the registers
    int v[CSLOT], nn = n0;
    // tags (requested above) and labels, without a branch (entries past the list
    // read the owner's own)
    uint8_t tg[CSLOT], lb[CSLOT];
#pragma unroll
    for (int k = 0; k < CSLOT; k++) {
      v[k] = ve[k];
      tg[k] = te[k];
    }
    if (first) {
      // the list k_lists would build (shot order, round-0 Skips removed from the
      // back with swap-with-last), its tags and labels loaded once and moved with
      // their entries; kept at i * CSLOT for the later rounds; the cell's
      // arrivals restart at 0
      const int na = n0 < CSLOT ? n0 : CSLOT;
      sort_slots(na, v, tg);
#pragma unroll
      for (int k = 0; k < CSLOT; k++) lb[k] = s_lab[k < na ? v[k] : i];
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
      for (int k = 0; k < CSLOT; k++) lb[k] = s_lab[k < nn ? v[k] : i];
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
    if (cs.mod) cs.mod[c] = ((draw_id & 0x7fffff) << 8) | cnt0;
    cell2t[c] = -1;
    return;
  }
  int* a = ord + sg;
  int n = n0;
  for (int q = n - 1; q >= 0; q--) {
    const int j = a[q];
    const uint8_t e = ctag[j] == T_UNDECIDED ? s_lab[j] : ctag[j];
    if (e == T_SKIP) {
      ctag[j] = T_SKIP;
      a[q] = a[n - 1];
      n--;
    }
  }
  len[i] = n;
  for (int q = 0; q < n; q++) {
    const int j = a[q];
    if (ctag[j] == T_UNDECIDED && s_lab[j] == T_UNDECIDED) ctag[j] = T_VALIDATED;
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
  if (cs.mod) cs.mod[c] = ((draw_id & 0x7fffff) << 8) | cnt0;
  cell2t[c] = -1;
}

template <int D>
__global__ void __launch_bounds__(kSweepThreads) k_sweep_commit(
    CellStore<D> cs, int nshots, const double* __restrict__ cpos, const double* __restrict__ crad,
    const uint32_t* __restrict__ cphase, const int* __restrict__ ccell, const uint8_t* __restrict__ owner,
    uint8_t* __restrict__ ctag, const uint8_t* __restrict__ lab, const int* __restrict__ slab,
    const int* __restrict__ pool, const URec* __restrict__ urec, int* __restrict__ seg, int* __restrict__ len,
    int* __restrict__ ord, int* __restrict__ cell2t, DrawCtl* __restrict__ ctl, int seq, int gate_id, int draw_id,
    double aimed, uint64_t miss_base, DrawPub* host, int* __restrict__ cellcc, const int* __restrict__ cslot,
    int ncells) {
  k_sweep_commit_body<D>(cs, nshots, cpos, crad, cphase, ccell, owner, ctag, lab, slab, pool, urec, seg, len, ord,
                         cell2t, ctl, seq, gate_id, draw_id, aimed, miss_base, host, cellcc, cslot, ncells,
                         (int)blockIdx.x);
}

// A draw's sweep and commit together with the next draw's decode, in one
// launch: the next draw's shots are decoded (whatever this draw decides:
// k_filter checks the gate, and tests the spheres this draw places, which the
// decode's filter may miss) while this one's owners commit.
// The two use different buffers (per parity of the draw id) and different
// control words; workgroups 0 .. nsc - 1 run the sweep and commit, the others the
// decode, kSweepThreads threads each.
template <int D>
struct SweepArgs {
  CellStore<D> cs;
  int nshots;
  const double* cpos;
  const double* crad;
  const uint32_t* cphase;
  const int* ccell;
  const uint8_t* owner;
  uint8_t* ctag;
  const uint8_t* lab;
  const int* slab;
  const int* pool;
  const URec* urec;
  int* seg;
  int* len;
  int* ord;
  int* cell2t;
  DrawCtl* ctl;
  int seq, gate_id, draw_id;
  double aimed;
  uint64_t miss_base;
  DrawPub* host;
  int* cellcc;
  const int* cslot;
  int ncells;
};
template <int D>
struct DecodeArgs {
  const uint32_t* ring;
  uint64_t mask, w0;
  int nshots;
  uint32_t N;
  const uint32_t* vpos;
  const double* hrad;
  RejList rl;
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
};
template <int D, bool WIDE>
__global__ void __launch_bounds__(kSweepThreads) k_sweep_commit_decode(SweepArgs<D> s, int nsc, DecodeArgs<D> d) {
  const int b = (int)blockIdx.x;
  if (b < nsc) {
    k_sweep_commit_body<D>(s.cs, s.nshots, s.cpos, s.crad, s.cphase, s.ccell, s.owner, s.ctag, s.lab, s.slab, s.pool,
                           s.urec, s.seg, s.len, s.ord, s.cell2t, s.ctl, s.seq, s.gate_id, s.draw_id, s.aimed,
                           s.miss_base, s.host, s.cellcc, s.cslot, s.ncells, b);
  } else {
    k_decode_filter_body<D, WIDE, true>(d.ring, d.mask, d.w0, d.nshots, d.N, d.vpos, d.hrad, d.rl, d.end_rel, d.vx,
                                         d.cls, d.g, d.cs, d.cpos, d.crad, d.cprio, d.cphase, d.ccell, d.ctag,
                                         d.owner, d.cellcc, d.cslot, d.ovhead, d.next, d.ctl, d.draw_id, d.forced,
                                         d.list_thr, b - nsc);
  }
}

}  // namespace rsa_amd
