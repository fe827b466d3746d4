// Kernels of a voxel refinement: the coverage flag and children masks
// (remove_covered + subdivide_uncovered), the cell-grouped processing order,
// the device-wide scans, the swap-with-last compaction in closed form, the
// children scatter and the initial lattice.
#pragma once

#include "device_common.hpp"
#include "k_draw.hpp"

namespace rsa_amd {

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
__device__ __forceinline__ void k_cell_cnt_body(int64_t ncells, const uint8_t* __restrict__ cnt,
                                                 int* __restrict__ off, const int bid) {
  const int64_t c = (int64_t)bid * blockDim.x + threadIdx.x;
  if (c < ncells) off[c] = cnt[c];
}

__global__ void __launch_bounds__(WG) k_cell_cnt(int64_t ncells, const uint8_t* __restrict__ cnt,
                                                 int* __restrict__ off) {
  k_cell_cnt_body(ncells, cnt, off, (int)blockIdx.x);
}

template <int D>
__device__ __forceinline__ void k_compact_body(int64_t ncells, CellStore<D> cs, const int* __restrict__ off,
                                                Slot<D>* __restrict__ out, double* __restrict__ rad, VoxParams<D> vp,
                                                int64_t cap, int* __restrict__ err, const int bid) {
  const int64_t c = (int64_t)bid * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int n = cs.cnt[c], o = off[c];
  if ((int64_t)o + n > cap) {
    atomicOr(err, 1);
    return;
  }
  for (int k = 0; k < n; k++) {
    // the centre with k_cover's listing threshold U = R^2 + slack(R^2, r + r_min) in
    // place of the radius (the radius goes to rad)
    Slot<D> q = cs.sph[(size_t)c * cs.cap + k];
    const double rr = q.r + vp.rmin;
    const double R2 = rr * rr;
    rad[o + k] = q.r;
    q.r = R2 + vp.slack(R2, rr);
    out[o + k] = q;
  }
}

template <int D>
__global__ void __launch_bounds__(WG) k_compact(int64_t ncells, CellStore<D> cs, const int* __restrict__ off,
                                                Slot<D>* __restrict__ out, double* __restrict__ rad, VoxParams<D> vp,
                                                int64_t cap, int* __restrict__ err) {
  k_compact_body<D>(ncells, cs, off, out, rad, vp, cap, err, (int)blockIdx.x);
}

// remove_covered's coverage flag (list_of_voxels.ixx:30-40, 240-259) and
// subdivide_uncovered's children test (:56-100, 261-307) in one pass:
// out[i] = 0 if voxel i is covered, else 0x10000 | (bit k set when child k in
// TabCorner order stays uncovered).  blk[b] counts uncovered voxels per block.
// Interior voxels of D <= 3 read the packed copy (coff, csph) by stencil rows
// when the host made one (coff not null); the others read the cell store.
template <int D, bool WIDE>
__device__ __forceinline__ void k_cover_body(Geo<D> g, CellStore<D> cs, VoxView<D> vx, int64_t nvox,
                                              VoxParams<D> vp, const uint32_t* __restrict__ perm,
                                              const int* __restrict__ coff, const Slot<D>* __restrict__ csph,
                                              const double* __restrict__ crad, int ncs, uint32_t* __restrict__ out,
                                              int* __restrict__ blk, const int bid) {
  using TC = TabCorner<D>;
  constexpr int NCH = ipow_c(2, D);
  constexpr int NST = ipow_c(3, D);
  constexpr uint32_t FULL = (NCH == 32) ? 0xffffffffu : ((1u << NCH) - 1u);
  // thread t takes voxel perm[t] (the cell-grouped order) or, without perm, voxel t
  const int64_t t = (int64_t)bid * blockDim.x + threadIdx.x;
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
            const uint32_t k = s_list[t][threadIdx.x];
            const Slot<D> q = csph[k];
            test(q.c, crad[k], do_cov, do_child);
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
          // The copy holds U = R^2 + slack(R^2, r + r_min) in place of the radius
          // (k_compact).  Both gates imply s2 <= U: the subdivide pre-filter s2 < R^2
          // (slack >= 0), and the coverage pre-check s2 <= (R^2 - h2) + slack with
          // h2 >= 0 (D <= 3), rounding being monotone.  A sphere with s2 <= U is listed
          // and gets the exact gates in test().
          // The entry is written at position nl whether or not it is listed (nl < KLIST
          // there: a full list is flushed at once), so that listing is a count.
          const Slot<D>* const qe = csph + b;
          uint32_t idx = (uint32_t)a;
          for (const Slot<D>* qp = csph + a; qp < qe; qp++, idx++) {
            const Slot<D> q = *qp;
            double s2 = 0.;
#pragma unroll
            for (int d = 0; d < D; d++) {
              const double t = q.c[d] - cb[d];
              s2 += t * t;
            }
            s_list[nl][threadIdx.x] = idx;
            nl += s2 <= q.r ? 1 : 0;
            if (nl == KLIST) flush();
          }
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
    if (threadIdx.x == 0) blk[bid] = u;
  }
}

template <int D, bool WIDE>
__global__ void __launch_bounds__(WG) k_cover(Geo<D> g, CellStore<D> cs, VoxView<D> vx, int64_t nvox,
                                              VoxParams<D> vp, const uint32_t* __restrict__ perm,
                                              const int* __restrict__ coff, const Slot<D>* __restrict__ csph,
                                              const double* __restrict__ crad, int ncs, uint32_t* __restrict__ out,
                                              int* __restrict__ blk) {
  k_cover_body<D, WIDE>(g, cs, vx, nvox, vp, perm, coff, csph, crad, ncs, out, blk, (int)blockIdx.x);
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
__device__ __forceinline__ void k_bin_body(Geo<D> g, VoxView<D> vx, int64_t nvox, VoxParams<D> vp,
                                            int* __restrict__ ccount, int* __restrict__ vcell,
                                            int* __restrict__ vrank, const uint8_t* __restrict__ pred, const int bid) {
  // each wave takes 4 x 64 consecutive voxels, in four passes of 64 whose loads
  // are all issued first (launch nblocks(n, 4 * WG))
  constexpr int NP = 4;
  const int lane = threadIdx.x & 63;
  const int64_t w0 = ((int64_t)bid * (WG / 64) + (threadIdx.x >> 6)) * (64 * NP);
  int64_t cv[NP][D];
#pragma unroll
  for (int q = 0; q < NP; q++) {
    const int64_t i = w0 + q * 64 + lane;
    const int64_t ic = i < nvox ? i : 0;
#pragma unroll
    for (int d = 0; d < D; d++) cv[q][d] = WIDE ? vx.w[d][ic] : (int64_t)vx.c[d][ic];
  }
#pragma unroll
  for (int q = 0; q < NP; q++) {
    const int64_t i = w0 + q * 64 + lane;
    const bool valid = i < nvox;
    int key = -1 - lane;  // distinct for lanes past the end
    if (valid) {
      double p[D];
      int cc[D];
#pragma unroll
      for (int d = 0; d < D; d++) p[d] = vp.origin[d] + ((double)cv[q][d] + 0.5) * vp.len[d];
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
}

template <int D, bool WIDE>
__global__ void __launch_bounds__(WG) k_bin(Geo<D> g, VoxView<D> vx, int64_t nvox, VoxParams<D> vp,
                                            int* __restrict__ ccount, int* __restrict__ vcell,
                                            int* __restrict__ vrank, const uint8_t* __restrict__ pred) {
  k_bin_body<D, WIDE>(g, vx, nvox, vp, ccount, vcell, vrank, pred, (int)blockIdx.x);
}

__device__ __forceinline__ void k_place_body(int64_t nvox, const int* __restrict__ vcell,
                                              const int* __restrict__ vrank, const int* __restrict__ cstart,
                                              uint32_t* __restrict__ perm, const int bid) {
  // four voxels per thread, 64 apart within the wave's 256 (launch nblocks(n, 4 * WG))
  constexpr int NP = 4;
  const int lane = threadIdx.x & 63;
  const int64_t w0 = ((int64_t)bid * (WG / 64) + (threadIdx.x >> 6)) * (64 * NP);
  int c[NP], r[NP], o[NP];
#pragma unroll
  for (int q = 0; q < NP; q++) {
    const int64_t i = w0 + q * 64 + lane;
    c[q] = i < nvox ? vcell[i] : 0;
    r[q] = i < nvox ? vrank[i] : 0;
  }
#pragma unroll
  for (int q = 0; q < NP; q++) o[q] = cstart[c[q]];
#pragma unroll
  for (int q = 0; q < NP; q++) {
    const int64_t i = w0 + q * 64 + lane;
    if (i < nvox) perm[o[q] + r[q]] = (uint32_t)i;
  }
}

__global__ void __launch_bounds__(WG) k_place(int64_t nvox, const int* __restrict__ vcell,
                                              const int* __restrict__ vrank, const int* __restrict__ cstart,
                                              uint32_t* __restrict__ perm) {
  k_place_body(nvox, vcell, vrank, cstart, perm, (int)blockIdx.x);
}

// The streaming passes over the list (k_block_count, k_holes, k_move,
// k_child_count) keep the blocks of WG voxels as their unit, but one wave takes
// a whole block: each lane four consecutive voxels with one 16-byte load, so
// that four times the bytes are in flight per wave.  A workgroup of WG threads
// thus covers blocks 4 * blockIdx.x .. 4 * blockIdx.x + 3 (launch nblocks(n, 4 * WG)).
constexpr int VPL = 4;  // voxels per lane
__device__ __forceinline__ void load4(const uint32_t* __restrict__ a, int64_t i0, int64_t n, uint32_t (&v)[VPL]) {
  if (i0 + VPL <= n) {
    const uint4 q = *reinterpret_cast<const uint4*>(a + i0);  // i0 is a multiple of 4: 16-byte aligned
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
    v[3] = q.w;
  } else {
#pragma unroll
    for (int k = 0; k < VPL; k++) v[k] = i0 + k < n ? a[i0 + k] : 0u;
  }
}
__device__ __forceinline__ int wave_sum(int x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// uncovered voxels per block of WG (the blocks of the list order)
__device__ __forceinline__ void k_block_count_body(int64_t nvox, const uint32_t* __restrict__ code,
                                                    int* __restrict__ blk, const int bid) {
  const int64_t b = (int64_t)bid * (WG / 64) + (threadIdx.x >> 6);
  const int64_t i0 = b * WG + (threadIdx.x & 63) * VPL;
  uint32_t c[VPL];
  load4(code, i0, nvox, c);
  int u = 0;
#pragma unroll
  for (int k = 0; k < VPL; k++) u += c[k] != 0;
  u = wave_sum(u);
  if ((threadIdx.x & 63) == 0 && b * WG < nvox) blk[b] = u;
}

__global__ void __launch_bounds__(WG) k_block_count(int64_t nvox, const uint32_t* __restrict__ code,
                                                    int* __restrict__ blk) {
  k_block_count_body(nvox, code, blk, (int)blockIdx.x);
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

__device__ __forceinline__ void k_tile_sums_body(const int* __restrict__ a, int64_t n, int* __restrict__ sums, const int bid) {
  const int64_t base = (int64_t)bid * SCAN_T + threadIdx.x * 4;
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
    sums[bid] = tot;
  }
}

__global__ void __launch_bounds__(1024) k_tile_sums(const int* __restrict__ a, int64_t n, int* __restrict__ sums) {
  k_tile_sums_body(a, n, sums, (int)blockIdx.x);
}

__device__ __forceinline__ void k_tile_scan_body(int* __restrict__ a, int64_t n, const int* __restrict__ sums, const int bid) {
  const int64_t base = (int64_t)bid * SCAN_T + threadIdx.x * 4;
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
  int run = sums[bid] + x - s;
  for (int w = 0; w < (int)(threadIdx.x >> 6); w++) run += sh[w];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (base + k < n) a[base + k] = run;
    run += v[k];
  }
}

__global__ void __launch_bounds__(1024) k_tile_scan(int* __restrict__ a, int64_t n, const int* __restrict__ sums) {
  k_tile_scan_body(a, n, sums, (int)blockIdx.x);
}

// single workgroup exclusive scan over block counts; writes the total to *total
// (*total = -1 when the sum exceeds INT_MAX: the offsets would not fit)
__device__ __forceinline__ void k_scan_blocks_body(int* __restrict__ blk, int nb, int* __restrict__ total, const int bid) {
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

__global__ void __launch_bounds__(1024) k_scan_blocks(int* __restrict__ blk, int nb, int* __restrict__ total) {
  k_scan_blocks_body(blk, nb, total, (int)blockIdx.x);
}

// swap-with-last compaction of remove_covered (list_of_voxels.ixx:42-53) in
// closed form: with keep uncovered voxels, the covered positions below keep,
// in increasing order, receive the uncovered voxels at or above keep in
// decreasing order.  Step 1 lists the holes.
// The voxel at keep - 1 also leaves the number of holes (= voxels step 2 moves).
__device__ __forceinline__ void k_holes_body(int64_t nvox, const uint32_t* __restrict__ code,
                                              const int* __restrict__ blk, int* __restrict__ ctr,
                                              int* __restrict__ holes, const int bid) {
  const int64_t b = (int64_t)bid * (WG / 64) + (threadIdx.x >> 6);
  if (b * WG >= nvox) return;  // whole waves past the list
  const int64_t i0 = b * WG + (threadIdx.x & 63) * VPL;
  uint32_t c[VPL];
  load4(code, i0, nvox, c);
  int n = 0;
#pragma unroll
  for (int k = 0; k < VPL; k++) n += c[k] != 0;
  // uncovered voxels before i0: the block's offset plus the lanes before this one
  int U = blk[b] + wave_incl_scan(n) - n;
  const int keep = ctr[C_KEEP];
#pragma unroll
  for (int k = 0; k < VPL; k++) {
    const int64_t i = i0 + k;
    const bool valid = i < nvox, unc = c[k] != 0;
    if (valid && !unc && i < keep) holes[i - U] = (int)i;
    if (valid && i == (int64_t)keep - 1) ctr[C_MOVED] = keep - U - (unc ? 1 : 0);
    U += unc;
  }
}

__global__ void __launch_bounds__(WG) k_holes(int64_t nvox, const uint32_t* __restrict__ code,
                                              const int* __restrict__ blk, int* __restrict__ ctr,
                                              int* __restrict__ holes) {
  k_holes_body(nvox, code, blk, ctr, holes, (int)blockIdx.x);
}

// Step 2: the uncovered voxels at or above keep move into the holes.  Their
// codes go to `moved` (at the hole's index), not into `code`: the workgroup
// that straddles keep reads code[] below keep for its ranks while other
// workgroups fill holes, so code[] must not change during this kernel.
// Afterwards the code of voxel i < keep is code[i], or moved[i] where code[i]
// is 0 (every hole is filled).
template <int D, bool WIDE>
__device__ __forceinline__ void k_move_body(int64_t nvox, const uint32_t* __restrict__ code,
                                             uint32_t* __restrict__ moved, const int* __restrict__ blk,
                                             const int* __restrict__ ctr, const int* __restrict__ holes,
                                             VoxView<D> vx, const int bid) {
  const int64_t b = (int64_t)bid * (WG / 64) + (threadIdx.x >> 6);
  if (b * WG >= nvox) return;  // whole waves past the list
  const int64_t i0 = b * WG + (threadIdx.x & 63) * VPL;
  uint32_t c[VPL];
  load4(code, i0, nvox, c);
  int n = 0;
#pragma unroll
  for (int k = 0; k < VPL; k++) n += c[k] != 0;
  int U = blk[b] + wave_incl_scan(n) - n;
  const int keep = ctr[C_KEEP];
  // the lane's moves: destinations first, then every coordinate load, then the stores
  bool mv[VPL];
  int dst[VPL];
#pragma unroll
  for (int k = 0; k < VPL; k++) {
    const int64_t i = i0 + k;
    const bool unc = c[k] != 0;
    mv[k] = i < nvox && unc && i >= keep;
    dst[k] = mv[k] ? holes[keep - U - 1] : 0;
    U += unc;
  }
  int64_t cv[VPL][D];
#pragma unroll
  for (int k = 0; k < VPL; k++)
#pragma unroll
    for (int d = 0; d < D; d++) cv[k][d] = !mv[k] ? 0 : (WIDE ? vx.w[d][i0 + k] : (int64_t)vx.c[d][i0 + k]);
#pragma unroll
  for (int k = 0; k < VPL; k++) {
    if (!mv[k]) continue;
#pragma unroll
    for (int d = 0; d < D; d++) {
      if constexpr (WIDE)
        const_cast<int64_t*>(vx.w[d])[dst[k]] = cv[k][d];
      else
        const_cast<int32_t*>(vx.c[d])[dst[k]] = (int32_t)cv[k][d];
    }
    moved[dst[k]] = c[k];
  }
}

template <int D, bool WIDE>
__global__ void __launch_bounds__(WG) k_move(int64_t nvox, const uint32_t* __restrict__ code,
                                             uint32_t* __restrict__ moved, const int* __restrict__ blk,
                                             const int* __restrict__ ctr, const int* __restrict__ holes,
                                             VoxView<D> vx) {
  k_move_body<D, WIDE>(nvox, code, moved, blk, ctr, holes, vx, (int)blockIdx.x);
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

__device__ __forceinline__ void k_child_count_body(int64_t n, const int* __restrict__ ctr,
                                                    const uint32_t* __restrict__ code,
                                                    const uint32_t* __restrict__ moved, int* __restrict__ blk, const int bid) {
  const int64_t b = (int64_t)bid * (WG / 64) + (threadIdx.x >> 6);
  const int64_t i0 = b * WG + (threadIdx.x & 63) * VPL;
  const int64_t keep = ctr[C_KEEP];
  const int64_t lim = n < keep ? n : keep;
  uint32_t c[VPL], m[VPL];
  load4(code, i0, lim, c);
  load4(moved, i0, lim, m);
  int t = 0;
#pragma unroll
  for (int k = 0; k < VPL; k++) t += __popc((c[k] != 0 ? c[k] : m[k]) & 0xffffu);
  t = wave_sum(t);
  if ((threadIdx.x & 63) == 0 && b * WG < n) blk[b] = t;
}

__global__ void __launch_bounds__(WG) k_child_count(int64_t n, const int* __restrict__ ctr,
                                                    const uint32_t* __restrict__ code,
                                                    const uint32_t* __restrict__ moved, int* __restrict__ blk) {
  k_child_count_body(n, ctr, code, moved, blk, (int)blockIdx.x);
}

// stable append of the kept children, parents in list order, children in
// TabCorner order (list_of_voxels.ixx:76-99): coordinates 2c + tab[k]
// SW / DW: the source and destination lattices hold 64-bit coordinates (the
// launch picks the instance, so coordinate loads and stores carry no branch).
template <int D, bool SW, bool DW>
__device__ __forceinline__ void k_scatter_body(int64_t n, const uint32_t* __restrict__ code,
                                                const uint32_t* __restrict__ moved, const int* __restrict__ blk,
                                                VoxView<D> src, VoxView<D> dst, int64_t nchild,
                                                int* __restrict__ err, const int bid) {
  const int64_t i = (int64_t)bid * blockDim.x + threadIdx.x;
  const bool in = i < n;
  const int64_t ic = in ? i : 0;
  const uint32_t m = in ? (kept_code_now(code, moved, ic) & 0xffffu) : 0u;
  int64_t pc[D];
#pragma unroll
  for (int d = 0; d < D; d++) pc[d] = SW ? src.w[d][ic] : (int64_t)src.c[d][ic];
  const int c = __popc(m);
  int total;
  const int excl = block_excl_scan(c, total);
  const int base = blk[bid];  // the workgroup's children are base .. base + total - 1
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

template <int D, bool SW, bool DW>
__global__ void __launch_bounds__(WG) k_scatter(int64_t n, const uint32_t* __restrict__ code,
                                                const uint32_t* __restrict__ moved, const int* __restrict__ blk,
                                                VoxView<D> src, VoxView<D> dst, int64_t nchild,
                                                int* __restrict__ err) {
  k_scatter_body<D, SW, DW>(n, code, moved, blk, src, dst, nchild, err, (int)blockIdx.x);
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

}  // namespace rsa_amd
