// Extraction of the placed spheres in the reference's order (real cells by
// index, then in-cell order), into allocated buffers or, for callers that
// cannot allocate, with the block counts in a static device scratch array.
#pragma once

#include "device_common.hpp"
#include "k_refine.hpp"

namespace rsa_amd {

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

// xsum (may be null): the set's checksum (word_mix sum over the spheres written)
template <int D, class PH>
__global__ void k_extract(Geo<D> g, CellStore<D> cs, const int* __restrict__ blk, double* __restrict__ pos,
                          double* __restrict__ rad, PH* __restrict__ ph, unsigned long long* __restrict__ xsum) {
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
  uint64_t h = 0;
  int o = blk[blockIdx.x] + sh[threadIdx.x] - v;
  for (int k = 0; k < v; k++, o++) {
    const Slot<D> s = cs.sph[(size_t)c * cs.cap + k];
    for (int d = 0; d < D; d++) {
      pos[(size_t)o * D + d] = s.c[d];
      h += word_mix(bits_of(s.c[d]), d);
    }
    rad[o] = s.r;
    h += word_mix(bits_of(s.r), D);
    if (ph) ph[o] = (PH)cs.phase[(size_t)c * cs.cap + k];
  }
  if (xsum == nullptr) return;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) h += (uint64_t)__shfl_xor((unsigned long long)h, off, 64);
  if ((threadIdx.x & 63) == 0 && h) atomicAdd(xsum, (unsigned long long)h);
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

// n doubles set to a quiet NaN: sphere slots not yet written (a NaN centre or
// radius collides with nothing, so a decode that reads a slot before the commit
// writing it has finished misses that sphere and k_filter tests it)
__global__ void k_fill_nan(double* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = __longlong_as_double(0x7ff8000000000000ll);
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

}  // namespace rsa_amd
