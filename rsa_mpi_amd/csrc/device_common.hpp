// Device-side data layout and exact FP64 predicates shared by the kernels.
//
// Every floating-point expression below is evaluated in the order the
// reference writes it; the library is built with -ffp-contract=off so no
// multiply-add is fused.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rsa_amd {

constexpr int ipow_c(int b, int e) { return e == 0 ? 1 : b * ipow_c(b, e - 1); }

// Order-independent checksum of a sphere set (extraction check): the sum mod
// 2^64 over spheres and values (D coordinates, then the radius) of word_mix of
// each value's bits.  word_mix(., k) is a bijection of 64-bit words (an
// xor-shift, then a product with an odd constant), so changing any one value
// always changes the sum.
__host__ __device__ __forceinline__ uint64_t word_mix(uint64_t x, int k) {
  x ^= x >> 29;
  return x * ((0x9E3779B97F4A7C15ull ^ (0xC2B2AE3D27D4EB4Full * (uint64_t)(k + 1))) | 1ull);
}
__host__ __device__ __forceinline__ uint64_t bits_of(double v) {
  union {
    double d;
    uint64_t u;
  } w{v};
  return w.u;
}

// Tags of rsa_data_storage_with_tags (rsa_data_storage.hxx:141-151) for the
// candidates of the current draw; DROPPED marks a centre that rounded into a
// ghost cell (never placed).
enum : uint8_t { T_VALIDATED = 0, T_UNDECIDED = 1, T_CONFLICT = 2, T_SKIP = 3, T_DROPPED = 4, T_UNRESOLVED = 255 };

// One placed sphere in a cell slot: centre then radius, 16-byte aligned so a
// slot is fetched with vector loads.
template <int D>
struct alignas(16) Slot {
  double c[D];
  double r;
};
template <>
struct alignas(16) Slot<2> {
  double c[2];
  double r;
  double pad;
};
template <>
struct alignas(16) Slot<4> {
  double c[4];
  double r;
  double pad;
};

// Geometry of the cell grid (rsa_grid_traversal.ixx:11-73) and of the periodic
// ghost areas (rsa_domain.ixx:31-133), all host-computed exactly as the
// reference computes them.
template <int D>
struct Geo {
  int n[D];          // cells per dimension including ghost layers
  int gl;            // ghost layer count
  int ncells;
  double cshift[D];  // m_shift
  double cinv[D];    // m_cell_inverse_length
  double inf[D], sup[D];
  double lo_band_hi[D];  // inf + 2*rad : top of the band imaged to the upper ghost layer
  double hi_band_lo[D];  // sup - 2*rad : bottom of the band imaged to the lower ghost layer
  double L[D];           // global_sup - global_inf : periodic shift
  int nst;               // 3^D
  // stencil offsets in the reference's visiting order, one int per entry with
  // (offset_d + 1) in bits 2d..2d+1; device memory, so that a lane-dependent
  // entry is a cached load rather than a private copy of this struct
  const int* stab;
  // the same offsets as linear cell-index differences (valid where the whole
  // stencil lies inside the grid)
  const int* slin;
};

template <int D>
__device__ __host__ __forceinline__ int stencil_off(int packed, int d) {
  return ((packed >> (2 * d)) & 3) - 1;
}

template <int D>
struct CellStore {
  uint8_t* cnt;     // validated spheres per cell
  Slot<D>* sph;     // cell * cap + k (slots past the count hold NaN until written)
  uint32_t* phase;  // cell * cap + k
  int cap;
  // per cell, the last draw that appended to it and the count before it:
  // ((draw id mod 2^23) << 8) | count; -1 before any (null: not kept)
  int* mod = nullptr;
};

// compute_cell_idx (rsa_grid_traversal.ixx:373-382)
template <int D>
__device__ __host__ __forceinline__ int64_t cell_index(const Geo<D>& g, const double* p) {
  uint64_t ret = 0;
  for (int d = D - 1; d >= 0; d--) {
    const int64_t ijk = (int64_t)((p[d] - g.cshift[d]) * g.cinv[d]);
    ret = ret * (uint64_t)g.n[d] + (uint64_t)ijk;
  }
  return (int64_t)ret;
}

// compute_cell_idx's per-dimension terms for a point known to lie in the grid
// (every term in [0, n[d]), so the 32-bit truncation equals the 64-bit one):
// the coordinates c and the linear index (which fits 32 bits as ncells does)
template <int D>
__device__ __host__ __forceinline__ int cell_of_point(const Geo<D>& g, const double* p, int* c) {
  int ret = 0;
  for (int d = D - 1; d >= 0; d--) {
    c[d] = (int)((p[d] - g.cshift[d]) * g.cinv[d]);
    ret = ret * g.n[d] + c[d];
  }
  return ret;
}

template <int D>
__device__ __host__ __forceinline__ void cell_coords(const Geo<D>& g, int64_t idx, int* c) {
  for (int d = 0; d < D; d++) {
    c[d] = (int)(idx % g.n[d]);
    idx /= g.n[d];
  }
}

template <int D>
__device__ __host__ __forceinline__ bool coords_ghost(const Geo<D>& g, const int* c) {
  for (int d = 0; d < D; d++)
    if (c[d] < g.gl || c[d] > g.n[d] - 1 - g.gl) return true;
  return false;
}

// A stencil neighbour seen from real cell coords c: the real cell holding its
// spheres and, per dimension, whether it is a mirrored ghost (-1: lower ghost
// layer, images of the upper band shifted by -L; +1: upper ghost layer,
// images of the lower band shifted by +L).  Returns false when the neighbour
// lies beyond the first ghost layer (never happens for ghost_layer >= 1).
template <int D>
struct Nb {
  int src;
  signed char m[D];
  bool ghost;
};

// neighbour() with stencil entry s's packed offsets already loaded (packed = g.stab[s])
template <int D>
__device__ __forceinline__ bool neighbour_at(const Geo<D>& g, const int* c, int packed, Nb<D>& nb) {
  int64_t idx = 0;
  nb.ghost = false;
#pragma unroll
  for (int d = D - 1; d >= 0; d--) {
    int k = c[d] + stencil_off<D>(packed, d);
    signed char m = 0;
    const int inner = g.n[d] - 2 * g.gl;
    if (k < g.gl) {
      if (k != g.gl - 1) return false;
      k += inner;
      m = -1;
    } else if (k > g.n[d] - 1 - g.gl) {
      if (k != g.n[d] - g.gl) return false;
      k -= inner;
      m = 1;
    }
    nb.m[d] = m;
    if (m) nb.ghost = true;
    idx = idx * g.n[d] + k;
  }
  nb.src = (int)idx;
  return true;
}

template <int D>
__device__ __forceinline__ bool neighbour(const Geo<D>& g, const int* c, int s, Nb<D>& nb) {
  return neighbour_at<D>(g, c, g.stab[s], nb);
}

// Ghost-area membership of a real sphere for the mirrored neighbour nb
// (helper_check_interval over the area's closed box, helper.ixx:26-34) and its
// shifted image (rsa_ghost_area.ixx:22-30).
template <int D>
__device__ __host__ __forceinline__ bool image_of(const Geo<D>& g, const Nb<D>& nb, const double* p, double* q) {
  bool in = true;
#pragma unroll
  for (int d = 0; d < D; d++) {
    // values first, then selects: a selected address would force the struct into scratch
    const double inf = g.inf[d], sup = g.sup[d], lbh = g.lo_band_hi[d], hbl = g.hi_band_lo[d], L = g.L[d];
    const int m = nb.m[d];
    const double lo = m == -1 ? hbl : inf;  // m == -1: area at relative position +1, [sup - 2 rad, sup]
    const double hi = m == 1 ? lbh : sup;   // m == +1: relative position -1, [inf, inf + 2 rad]
    in = in && !(p[d] < lo) && !(p[d] > hi);
    const double sh = m == -1 ? -L : L;
    q[d] = (m != 0 && fabs(sh) >= 1E-16) ? p[d] + sh : p[d];
  }
  return in;
}

// is_there_collision (operator_cells_and_data_storage.ixx:5-14)
template <int D>
__device__ __host__ __forceinline__ bool collide(double r1, const double* p1, double r2, const double* p2) {
  double s = 0;
  for (int d = 0; d < D; d++) {
    const double dr = p1[d] - p2[d];
    s += dr * dr;
  }
  return s < (r1 + r2) * (r1 + r2);
}

// libstdc++ generate_canonical<double, 53> from two MT words (random.tcc:3346-3379):
// (g1 + g2 * 2^32) / 2^64, values >= 1 replaced by the largest double below 1.
__device__ __host__ __forceinline__ double canonical(uint32_t g1, uint32_t g2) {
  double sum = (double)g1;
  sum += (double)g2 * 4294967296.0;
  double r = sum * 0x1p-64;  // exact division by 2^64
  if (r >= 1.0) r = 0x1.fffffffffffffp-1;
  return r;
}

}  // namespace rsa_amd
