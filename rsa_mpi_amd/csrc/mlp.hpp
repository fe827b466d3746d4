// The learned voxel/phase predictor: a 16 -> 64 -> 64 -> 2 multilayer
// perceptron with ReLU, BF16 weights and activations, FP32 accumulation on the
// gfx950 matrix cores (v_mfma_f32_16x16x16_bf16).  Device-side pieces shared by
// the training kernels (learn.hip) and the engine's prediction pass.
//
// MFMA 16x16x16 lane layout (wave64): lane l = 16 * g + c (g = l / 16, c = l % 16)
//   A (16 x 16, M x K): lane l holds A[c][4g + j], j = 0..3
//   B (16 x 16, K x N): lane l holds B[4g + j][c]
//   D (16 x 16, M x N): lane l holds D[4g + j][c]  (float)
// Activation tiles of 16 samples live in LDS, row-major with a padded row
// stride, and are re-read in whichever layout the next product needs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rsa_amd {
namespace mlp {

constexpr int NF = 16;  // input features
constexpr int NH = 64;  // hidden units per layer
constexpr int NO = 2;   // outputs: 0 = the voxel's coverage logit, 1 = the draw's refinement logit
constexpr int NOP = 16; // outputs padded to one MFMA tile
// FP32 parameters, in this order: W1[NF][NH], b1[NH], W2[NH][NH], b2[NH], W3[NH][NO], b3[NO]
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NF * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NO, NPAR = OFF_B3 + NO;
// BF16 weights as the products read them: W1q[NF][NH], W2q[NH][NH], W3q[NH][NOP] (outputs >= NO zero)
constexpr int QOFF_W1 = 0, QOFF_W2 = QOFF_W1 + NF * NH, QOFF_W3 = QOFF_W2 + NH * NH, NQ = QOFF_W3 + NH * NOP;
constexpr int LDS_STRIDE = NH + 8;  // bf16 per LDS row: 144 bytes, so 16 rows spread over the banks

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (finite inputs; NaN stays NaN)
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
  union {
    float f;
    uint32_t u;
  } v{f};
  if ((v.u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((v.u >> 16) | ((v.u & 0xffffu) ? 0x40u : 0u));
  const uint32_t r = v.u + 0x7fffu + ((v.u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h) {
  union {
    uint32_t u;
    float f;
  } v{(uint32_t)h << 16};
  return v.f;
}

// LDS written by some lanes of a wave and read by others: a wave's LDS
// operations execute in order, so only the compiler must keep them in order
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ floatx4 mfma(bf16x4 a, bf16x4 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// A fragment from a row-major LDS tile T (16 rows): rows c, columns col0 + 4g .. + 3
__device__ __forceinline__ bf16x4 lds_a(const uint16_t* T, int stride, int col0, int lane) {
  const int g = lane >> 4, c = lane & 15;
  const uint16_t* p = T + c * stride + col0 + 4 * g;
  bf16x4 r;
  r[0] = (short)p[0];
  r[1] = (short)p[1];
  r[2] = (short)p[2];
  r[3] = (short)p[3];
  return r;
}
// A fragment of the transpose of an LDS tile: A[m][k] = T[k][col0 + m], k = sample row
__device__ __forceinline__ bf16x4 lds_at(const uint16_t* T, int stride, int col0, int lane) {
  const int g = lane >> 4, c = lane & 15;
  bf16x4 r;
#pragma unroll
  for (int j = 0; j < 4; j++) r[j] = (short)T[(4 * g + j) * stride + col0 + c];
  return r;
}
// B fragment from a row-major tile (K rows): B[k][n] = T[k0 + 4g + j][col0 + c]
__device__ __forceinline__ bf16x4 lds_b(const uint16_t* T, int stride, int k0, int col0, int lane) {
  const int g = lane >> 4, c = lane & 15;
  bf16x4 r;
#pragma unroll
  for (int j = 0; j < 4; j++) r[j] = (short)T[(k0 + 4 * g + j) * stride + col0 + c];
  return r;
}
// B fragment from a row-major weight matrix W[K][N] in global memory
__device__ __forceinline__ bf16x4 glb_b(const uint16_t* W, int ncol, int k0, int col0, int lane) {
  const int g = lane >> 4, c = lane & 15;
  bf16x4 r;
#pragma unroll
  for (int j = 0; j < 4; j++) r[j] = (short)W[(k0 + 4 * g + j) * ncol + col0 + c];
  return r;
}
// B fragment of a transposed weight matrix: B[k][n] = W[col0 + n][k0 + k] (W row-major, ncol columns)
__device__ __forceinline__ bf16x4 glb_bt(const uint16_t* W, int ncol, int k0, int col0, int lane) {
  const int g = lane >> 4, c = lane & 15;
  const uint16_t* p = W + (col0 + c) * ncol + k0 + 4 * g;
  bf16x4 r;
  r[0] = (short)p[0];
  r[1] = (short)p[1];
  r[2] = (short)p[2];
  r[3] = (short)p[3];
  return r;
}
// a D tile (+ bias per column, ReLU when asked) into an LDS tile at column col0, as bf16
__device__ __forceinline__ void d_to_lds(floatx4 d, const float* bias, bool relu, uint16_t* T, int stride, int col0,
                                         int lane) {
  const int g = lane >> 4, c = lane & 15;
  const float b = bias ? bias[col0 + c] : 0.f;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    float v = d[j] + b;
    if (relu) v = v > 0.f ? v : 0.f;
    T[(4 * g + j) * stride + col0 + c] = f2bf(v);
  }
}

// The weights one wave keeps in registers for forward passes: B fragments of
// W1 (4 column tiles), W2 (4 K steps x 4 column tiles) and W3 (4 K steps).
struct Frags {
  bf16x4 w1[4], w2[4][4], w3[4];
};
__device__ __forceinline__ void load_frags(const uint16_t* Wq, Frags& f, int lane) {
#pragma unroll
  for (int nt = 0; nt < 4; nt++) f.w1[nt] = glb_b(Wq + QOFF_W1, NH, 0, nt * 16, lane);
#pragma unroll
  for (int ks = 0; ks < 4; ks++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) f.w2[ks][nt] = glb_b(Wq + QOFF_W2, NH, ks * 16, nt * 16, lane);
#pragma unroll
  for (int ks = 0; ks < 4; ks++) f.w3[ks] = glb_b(Wq + QOFF_W3, NOP, ks * 16, 0, lane);
}

// Forward pass of one tile of 16 samples: X (16 x NF bf16, row stride LDS_STRIDE)
// in LDS; H1, H2 are LDS tiles the pass fills (16 x NH); returns the output
// tile Z3 in D layout (columns 0 and 1 hold the two logits).  bias: the FP32
// parameter array (b1, b2, b3 read from it).  Call with the whole wave.
__device__ __forceinline__ floatx4 forward_tile(const Frags& f, const float* P, const uint16_t* X, uint16_t* H1,
                                                uint16_t* H2, int lane) {
  const floatx4 z = {0.f, 0.f, 0.f, 0.f};
  const bf16x4 ax = lds_a(X, LDS_STRIDE, 0, lane);
#pragma unroll
  for (int nt = 0; nt < 4; nt++) d_to_lds(mfma(ax, f.w1[nt], z), P + OFF_B1, true, H1, LDS_STRIDE, nt * 16, lane);
  wave_lds_sync();
  bf16x4 a1[4];
#pragma unroll
  for (int ks = 0; ks < 4; ks++) a1[ks] = lds_a(H1, LDS_STRIDE, ks * 16, lane);
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    floatx4 acc = z;
#pragma unroll
    for (int ks = 0; ks < 4; ks++) acc = mfma(a1[ks], f.w2[ks][nt], acc);
    d_to_lds(acc, P + OFF_B2, true, H2, LDS_STRIDE, nt * 16, lane);
  }
  wave_lds_sync();
  floatx4 acc = z;
#pragma unroll
  for (int ks = 0; ks < 4; ks++) acc = mfma(lds_a(H2, LDS_STRIDE, ks * 16, lane), f.w3[ks], acc);
  const int c = lane & 15;
  const float b = c < NO ? P[OFF_B3 + c] : 0.f;
#pragma unroll
  for (int j = 0; j < 4; j++) acc[j] += b;
  return acc;
}

}  // namespace mlp
}  // namespace rsa_amd
