// Kernels of the learned predictor on voxel lists: per-voxel features, trace
// logging, the MFMA prediction pass and the outcome counters.
#pragma once

#include "device_common.hpp"
#include "mlp.hpp"
#include "k_refine.hpp"

namespace rsa_amd {

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

}  // namespace rsa_amd
