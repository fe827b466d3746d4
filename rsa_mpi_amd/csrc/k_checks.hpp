// End-of-run checks on the device: no two placed spheres (images included)
// overlap, and no sampled point is left free for a sphere of the smallest
// radius.
#pragma once

#include "device_common.hpp"
#include "k_refine.hpp"

namespace rsa_amd {

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

}  // namespace rsa_amd
