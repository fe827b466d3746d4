// Training of the learned predictor on the GPU: hand-written forward and
// backward passes of the 16 -> 64 -> 64 -> 2 MLP on the matrix cores
// (v_mfma_f32_16x16x16_bf16, BF16 operands, FP32 accumulation), Adam on FP32
// master parameters, and the process-wide trace store the engine appends to.
//
// Loss: binary cross-entropy with logits on each output whose mask is set,
// averaged over the batch.  Output 0 is trained on voxel samples (covered at
// the refinement pass), output 1 on draw samples (the draw triggered a
// refinement); a sample carries one target or the other.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>

#include "engine.hpp"
#include "learn.hpp"

namespace rsa_amd {
namespace learn {

using namespace mlp;

#define LHIP(expr)                                                                               \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) throw Error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + #expr); \
  } while (0)

// ---------------------------------------------------------------- kernels

__global__ void k_copy_host(float* __restrict__ dst, const float* __restrict__ src, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
__global__ void k_copy_host_bytes(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
__global__ void k_set_i(int* __restrict__ p, int v) { *p = v; }
__global__ void k_fill_f(float* __restrict__ p, int n, float v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// the BF16 weight copy the products read, W3 padded with zero outputs
__global__ void k_quantize(const float* __restrict__ P, uint16_t* __restrict__ Wq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < NF * NH) Wq[QOFF_W1 + i] = f2bf(P[OFF_W1 + i]);
  if (i < NH * NH) Wq[QOFF_W2 + i] = f2bf(P[OFF_W2 + i]);
  if (i < NH * NOP) {
    const int h = i / NOP, o = i % NOP;
    Wq[QOFF_W3 + i] = o < NO ? f2bf(P[OFF_W3 + h * NO + o]) : (uint16_t)0;
  }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {  // splitmix64 (Steele, Lea & Flood 2014)
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// sum of v over the 4 lanes c, c + 16, c + 32, c + 48 (the 16 rows of a D-layout column)
__device__ __forceinline__ float col_sum(float v) {
  v += __shfl_xor(v, 16, 64);
  v += __shfl_xor(v, 32, 64);
  return v;
}

// One training pass over ntiles tiles of 16 samples: forward, loss, backward;
// the gradients of the mean loss (1/n scaling by inv_n) are added to G, the
// loss to *loss.  Samples: the trace store rows s (seq: s = base + 16 t + row;
// otherwise s drawn uniformly from [0, nsamp) by a hash of (seed, t, row)).
constexpr int TW = 4;  // waves per workgroup
__global__ void __launch_bounds__(64 * TW) k_train_step(const uint16_t* __restrict__ TX, const uint8_t* __restrict__ TT,
                                                        int64_t nsamp, int ntiles, int seq, int64_t base,
                                                        uint64_t seed,
                                                        const float* __restrict__ P, const uint16_t* __restrict__ Wq,
                                                        float* __restrict__ G, float* __restrict__ loss_out,
                                                        float inv_n) {
  __shared__ uint16_t sX[TW][16 * LDS_STRIDE], sH1[TW][16 * LDS_STRIDE], sH2[TW][16 * LDS_STRIDE],
      sG[TW][16 * LDS_STRIDE];
  __shared__ int sS[TW][16];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  const int gw = blockIdx.x * TW + w, nw = gridDim.x * TW;
  Frags f;
  load_frags(Wq, f, lane);
  const floatx4 z = {0.f, 0.f, 0.f, 0.f};
  floatx4 gW1[4], gW2[4][4], gW3[4];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    gW1[a] = z;
    gW3[a] = z;
#pragma unroll
    for (int b = 0; b < 4; b++) gW2[a][b] = z;
  }
  float db1[4] = {0.f, 0.f, 0.f, 0.f}, db2[4] = {0.f, 0.f, 0.f, 0.f}, db3 = 0.f, loss = 0.f;
  uint16_t* X = sX[w];
  uint16_t* H1 = sH1[w];
  uint16_t* H2 = sH2[w];
  uint16_t* GT = sG[w];
  for (int t = gw; t < ntiles; t += nw) {
    // the tile's 16 samples: lane row c loads features 4g .. 4g + 3
    int64_t s;
    if (seq) {
      s = base + (int64_t)t * 16 + c;
    } else {
      s = (int64_t)(mix64(seed ^ ((uint64_t)t * 16 + c) * 0x2545f4914f6cdd1dull) % (uint64_t)nsamp);
    }
    if (g == 0) sS[w][c] = (int)s;
#pragma unroll
    for (int j = 0; j < 4; j++) X[c * LDS_STRIDE + 4 * g + j] = TX[s * NF + 4 * g + j];
    wave_lds_sync();
    const floatx4 z3 = forward_tile(f, P, X, H1, H2, lane);
    // dL/dz3 for the mean binary cross-entropy over masked outputs
    floatx4 d3 = z;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int r = 4 * g + j;
      const int sr = sS[w][r];
      if (c < NO) {
        const float y = (float)TT[(int64_t)sr * 4 + 2 * c];
        const float m = (float)TT[(int64_t)sr * 4 + 2 * c + 1];
        const float zz = z3[j];
        const float p = 1.f / (1.f + __expf(-zz));
        d3[j] = m * (p - y) * inv_n;
        const float sp = zz > 0.f ? zz + log1pf(__expf(-zz)) : log1pf(__expf(zz));
        loss += m * (sp - y * zz) * inv_n;
      }
    }
    db3 += col_sum(d3[0] + d3[1] + d3[2] + d3[3]);
    d_to_lds(d3, nullptr, false, GT, LDS_STRIDE, 0, lane);
    wave_lds_sync();
    // dW3 += H2^T dZ3;  dH2 = dZ3 W3^T
    const bf16x4 bG3 = lds_b(GT, LDS_STRIDE, 0, 0, lane);
    const bf16x4 aG3 = lds_a(GT, LDS_STRIDE, 0, lane);
#pragma unroll
    for (int mt = 0; mt < 4; mt++) gW3[mt] = mfma(lds_at(H2, LDS_STRIDE, mt * 16, lane), bG3, gW3[mt]);
    floatx4 dz2[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) dz2[nt] = mfma(aG3, glb_bt(Wq + QOFF_W3, NOP, 0, nt * 16, lane), z);
    wave_lds_sync();
    // dZ2 = dH2 * (H2 > 0)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (!(bf2f(H2[(4 * g + j) * LDS_STRIDE + nt * 16 + c]) > 0.f)) dz2[nt][j] = 0.f;
      db2[nt] += col_sum(dz2[nt][0] + dz2[nt][1] + dz2[nt][2] + dz2[nt][3]);
      d_to_lds(dz2[nt], nullptr, false, GT, LDS_STRIDE, nt * 16, lane);
    }
    wave_lds_sync();
    // dW2 += H1^T dZ2;  dH1 = dZ2 W2^T
    bf16x4 bG2[4], aG2[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      bG2[nt] = lds_b(GT, LDS_STRIDE, 0, nt * 16, lane);
      aG2[nt] = lds_a(GT, LDS_STRIDE, nt * 16, lane);
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      const bf16x4 a = lds_at(H1, LDS_STRIDE, mt * 16, lane);
#pragma unroll
      for (int nt = 0; nt < 4; nt++) gW2[mt][nt] = mfma(a, bG2[nt], gW2[mt][nt]);
    }
    floatx4 dz1[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      floatx4 acc = z;
#pragma unroll
      for (int ks = 0; ks < 4; ks++) acc = mfma(aG2[ks], glb_bt(Wq + QOFF_W2, NH, ks * 16, nt * 16, lane), acc);
      dz1[nt] = acc;
    }
    wave_lds_sync();
    // dZ1 = dH1 * (H1 > 0);  dW1 += X^T dZ1
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (!(bf2f(H1[(4 * g + j) * LDS_STRIDE + nt * 16 + c]) > 0.f)) dz1[nt][j] = 0.f;
      db1[nt] += col_sum(dz1[nt][0] + dz1[nt][1] + dz1[nt][2] + dz1[nt][3]);
      d_to_lds(dz1[nt], nullptr, false, GT, LDS_STRIDE, nt * 16, lane);
    }
    wave_lds_sync();
    const bf16x4 aX = lds_at(X, LDS_STRIDE, 0, lane);
#pragma unroll
    for (int nt = 0; nt < 4; nt++) gW1[nt] = mfma(aX, lds_b(GT, LDS_STRIDE, 0, nt * 16, lane), gW1[nt]);
    wave_lds_sync();  // the next tile overwrites every LDS tile
  }
  // the wave's partial gradients into G
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int r = 4 * g + j;
#pragma unroll
    for (int nt = 0; nt < 4; nt++) atomicAdd(&G[OFF_W1 + r * NH + nt * 16 + c], gW1[nt][j]);
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
#pragma unroll
      for (int nt = 0; nt < 4; nt++) atomicAdd(&G[OFF_W2 + (mt * 16 + r) * NH + nt * 16 + c], gW2[mt][nt][j]);
      if (c < NO) atomicAdd(&G[OFF_W3 + (mt * 16 + r) * NO + c], gW3[mt][j]);
    }
  }
  if (g == 0) {
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      atomicAdd(&G[OFF_B1 + nt * 16 + c], db1[nt]);
      atomicAdd(&G[OFF_B2 + nt * 16 + c], db2[nt]);
    }
    if (c < NO) atomicAdd(&G[OFF_B3 + c], db3);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
  if (lane == 0) atomicAdd(loss_out, loss);
}

// Adam on the FP32 parameters (bias-corrected), then the BF16 copy; clears the gradients
__global__ void k_adam(float* __restrict__ P, float* __restrict__ M, float* __restrict__ V, float* __restrict__ G,
                       uint16_t* __restrict__ Wq, float lr, float b1, float b2, float eps, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NPAR) return;
  const float gr = G[i];
  G[i] = 0.f;
  const float m = b1 * M[i] + (1.f - b1) * gr;
  const float v = b2 * V[i] + (1.f - b2) * gr * gr;
  M[i] = m;
  V[i] = v;
  const float p = P[i] - lr * (m / bc1) / (sqrtf(v / bc2) + eps);
  P[i] = p;
  if (i < OFF_B1) {
    Wq[QOFF_W1 + i] = f2bf(p);
  } else if (i >= OFF_W2 && i < OFF_B2) {
    Wq[QOFF_W2 + (i - OFF_W2)] = f2bf(p);
  } else if (i >= OFF_W3 && i < OFF_B3) {
    const int k = i - OFF_W3;
    Wq[QOFF_W3 + (k / NO) * NOP + k % NO] = f2bf(p);
  }
}

// forward of n samples (n a multiple of 16) from the trace-store layout to logits[n][NO]
__global__ void __launch_bounds__(64 * TW) k_forward(const uint16_t* __restrict__ TX, int ntiles,
                                                     const float* __restrict__ P, const uint16_t* __restrict__ Wq,
                                                     float* __restrict__ out) {
  __shared__ uint16_t sX[TW][16 * LDS_STRIDE], sH1[TW][16 * LDS_STRIDE], sH2[TW][16 * LDS_STRIDE];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  Frags f;
  load_frags(Wq, f, lane);
  for (int t = blockIdx.x * TW + w; t < ntiles; t += gridDim.x * TW) {
#pragma unroll
    for (int j = 0; j < 4; j++) sX[w][c * LDS_STRIDE + 4 * g + j] = TX[((int64_t)t * 16 + c) * NF + 4 * g + j];
    wave_lds_sync();
    const floatx4 z3 = forward_tile(f, P, sX[w], sH1[w], sH2[w], lane);
    if (c < NO)
#pragma unroll
      for (int j = 0; j < 4; j++) out[((int64_t)t * 16 + 4 * g + j) * NO + c] = z3[j];
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------- host

Model& Model::get() {
  static Model* m = new Model();  // never destroyed: engines may use it until the process ends
  return *m;
}

void* Model::stage(size_t bytes) {
  if (bytes > hstage_bytes_) {
    if (hstage_) LHIP(hipHostFree(hstage_));
    hstage_ = nullptr;
    LHIP(hipHostMalloc((void**)&hstage_, bytes, hipHostMallocCoherent | hipHostMallocMapped));
    hstage_bytes_ = bytes;
  }
  return hstage_;
}

void Model::ensure() {
  if (P_) return;
  LHIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
  LHIP(hipMalloc((void**)&P_, sizeof(float) * NPAR));
  LHIP(hipMalloc((void**)&M_, sizeof(float) * NPAR));
  LHIP(hipMalloc((void**)&V_, sizeof(float) * NPAR));
  LHIP(hipMalloc((void**)&G_, sizeof(float) * NPAR));
  LHIP(hipMalloc((void**)&loss_, sizeof(float) * 4));
  LHIP(hipMalloc((void**)&Wq_, sizeof(uint16_t) * NQ));
  LHIP(hipMalloc((void**)&tx_, sizeof(uint16_t) * NF * (size_t)kTraceCap));
  LHIP(hipMalloc((void**)&tt_, 4 * (size_t)kTraceCap));
  LHIP(hipMalloc((void**)&cur_, sizeof(int) * 4));
  reset(1);
}

void Model::upload_params(const float* p) {
  float* h = static_cast<float*>(stage(sizeof(float) * NPAR));
  std::memcpy(h, p, sizeof(float) * NPAR);
  float* hd = nullptr;
  LHIP(hipHostGetDevicePointer((void**)&hd, h, 0));
  hipLaunchKernelGGL(k_copy_host, dim3((NPAR + 255) / 256), dim3(256), 0, st_, P_, hd, NPAR);
  hipLaunchKernelGGL(k_quantize, dim3((NH * NH + 255) / 256), dim3(256), 0, st_, P_, Wq_);
  for (float* b : {M_, V_, G_}) hipLaunchKernelGGL(k_fill_f, dim3((NPAR + 255) / 256), dim3(256), 0, st_, b, NPAR, 0.f);
  LHIP(hipGetLastError());
  LHIP(hipStreamSynchronize(st_));
  adam_t_ = 0;
  hp_.assign(p, p + NPAR);
}

void Model::reset(uint64_t seed) {
  std::lock_guard<std::mutex> lk(m_);
  // He-normal weights, zero biases
  std::vector<float> p(NPAR, 0.f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int i = 0; i < NF * NH; i++) p[OFF_W1 + i] = (float)(nd(rng) * std::sqrt(2.0 / NF));
  for (int i = 0; i < NH * NH; i++) p[OFF_W2 + i] = (float)(nd(rng) * std::sqrt(2.0 / NH));
  for (int i = 0; i < NH * NO; i++) p[OFF_W3 + i] = (float)(nd(rng) * std::sqrt(1.0 / NH));
  upload_params(p.data());
  LHIP(hipMemsetAsync(cur_, 0, sizeof(int) * 4, st_));
  LHIP(hipStreamSynchronize(st_));
  trained_ = false;
}

void Model::set_params(const float* p) {
  ensure();
  std::lock_guard<std::mutex> lk(m_);
  upload_params(p);
  trained_ = true;
}

void Model::get_params(float* p) {
  ensure();
  std::lock_guard<std::mutex> lk(m_);
  LHIP(hipMemcpyAsync(p, P_, sizeof(float) * NPAR, hipMemcpyDeviceToHost, st_));
  LHIP(hipStreamSynchronize(st_));
}

void Model::refresh_host() {
  hp_.resize(NPAR);
  LHIP(hipMemcpyAsync(hp_.data(), P_, sizeof(float) * NPAR, hipMemcpyDeviceToHost, st_));
  LHIP(hipStreamSynchronize(st_));
}

int64_t Model::samples() {
  ensure();
  int n = 0;
  LHIP(hipMemcpyAsync(&n, cur_, sizeof(int), hipMemcpyDeviceToHost, st_));
  LHIP(hipStreamSynchronize(st_));
  return std::min<int64_t>(n, kTraceCap);
}

void Model::clear_samples() {
  ensure();
  LHIP(hipMemsetAsync(cur_, 0, sizeof(int), st_));
  LHIP(hipStreamSynchronize(st_));
}

void Model::add_samples(int64_t n, const float* x, const uint8_t* t) {
  ensure();
  std::lock_guard<std::mutex> lk(m_);
  const int64_t have = samples();
  if (have + n > kTraceCap) n = kTraceCap - have;
  if (n <= 0) return;
  const size_t bx = sizeof(uint16_t) * NF * (size_t)n, bt = 4 * (size_t)n;
  uint8_t* h = static_cast<uint8_t*>(stage(bx + bt));
  uint16_t* hx = reinterpret_cast<uint16_t*>(h);
  for (int64_t i = 0; i < n * NF; i++) hx[i] = f2bf(x[i]);
  std::memcpy(h + bx, t, bt);
  uint8_t* hd = nullptr;
  LHIP(hipHostGetDevicePointer((void**)&hd, h, 0));
  hipLaunchKernelGGL(k_copy_host_bytes, dim3((unsigned)((bx + 255) / 256)), dim3(256), 0, st_,
                     reinterpret_cast<uint8_t*>(tx_ + (size_t)have * NF), hd, (int64_t)bx);
  hipLaunchKernelGGL(k_copy_host_bytes, dim3((unsigned)((bt + 255) / 256)), dim3(256), 0, st_, tt_ + (size_t)have * 4,
                     hd + bx, (int64_t)bt);
  hipLaunchKernelGGL(k_set_i, dim3(1), dim3(1), 0, st_, cur_, (int)(have + n));
  LHIP(hipGetLastError());
  LHIP(hipStreamSynchronize(st_));
}

int64_t Model::train(int steps, int batch, double lr, uint64_t seed, double* loss_first, double* loss_last) {
  ensure();
  std::lock_guard<std::mutex> lk(m_);
  int nsamp = 0;
  LHIP(hipMemcpyAsync(&nsamp, cur_, sizeof(int), hipMemcpyDeviceToHost, st_));
  LHIP(hipStreamSynchronize(st_));
  nsamp = std::min(nsamp, kTraceCap);
  if (nsamp < 16) throw Error("learn: fewer than 16 traced samples to train on");
  if (batch < 16 || batch % 16) throw Error("learn: the batch must be a positive multiple of 16");
  const int ntiles = batch / 16;
  const unsigned grid = (unsigned)std::max(1, std::min(64, ntiles / TW));
  const float b1 = 0.9f, b2 = 0.999f;
  for (int s = 0; s < steps; s++) {
    LHIP(hipMemsetAsync(loss_, 0, sizeof(float), st_));
    hipLaunchKernelGGL(k_train_step, dim3(grid), dim3(64 * TW), 0, st_, tx_, tt_, (int64_t)nsamp, ntiles, 0,
                       (int64_t)0, seed + 0x9e3779b97f4a7c15ull * (uint64_t)(adam_t_ + 1), P_, Wq_, G_, loss_,
                       1.f / batch);
    adam_t_++;
    const float bc1 = 1.f - std::pow(b1, (float)adam_t_), bc2 = 1.f - std::pow(b2, (float)adam_t_);
    hipLaunchKernelGGL(k_adam, dim3((NPAR + 255) / 256), dim3(256), 0, st_, P_, M_, V_, G_, Wq_, (float)lr, b1, b2,
                       1e-8f, bc1, bc2);
    LHIP(hipGetLastError());
    if (s == 0 || s == steps - 1) {
      float l = 0;
      LHIP(hipMemcpyAsync(&l, loss_, sizeof(float), hipMemcpyDeviceToHost, st_));
      LHIP(hipStreamSynchronize(st_));
      if (s == 0 && loss_first) *loss_first = l;
      if (s == steps - 1 && loss_last) *loss_last = l;
    }
  }
  LHIP(hipStreamSynchronize(st_));
  refresh_host();
  trained_ = true;
  return nsamp;
}

void Model::gradcheck(int n, const float* x, const uint8_t* t, const float* params, float* grad, double* loss) {
  ensure();
  if (n < 16 || n % 16) throw Error("learn: gradcheck needs a positive multiple of 16 samples");
  if (n > kTraceCap) throw Error("learn: too many samples");
  std::vector<float> saved(NPAR);
  get_params(saved.data());
  const bool was_trained = trained_;
  const int64_t kept = samples();
  if (kept + n > kTraceCap) throw Error("learn: trace store full");
  add_samples(n, x, t);  // after the stored samples, which stay as they are
  std::lock_guard<std::mutex> lk(m_);
  upload_params(params);
  LHIP(hipMemsetAsync(loss_, 0, sizeof(float), st_));
  hipLaunchKernelGGL(k_train_step, dim3((unsigned)std::max(1, std::min(64, n / 16 / TW))), dim3(64 * TW), 0, st_,
                     tx_, tt_, (int64_t)(kept + n), n / 16, 1, kept, 0ull, P_, Wq_, G_, loss_, 1.f / n);
  LHIP(hipGetLastError());
  float l = 0;
  LHIP(hipMemcpyAsync(grad, G_, sizeof(float) * NPAR, hipMemcpyDeviceToHost, st_));
  LHIP(hipMemcpyAsync(&l, loss_, sizeof(float), hipMemcpyDeviceToHost, st_));
  LHIP(hipStreamSynchronize(st_));
  *loss = l;
  upload_params(saved.data());  // gradients cleared, the model as it was
  trained_ = was_trained;
  hipLaunchKernelGGL(k_set_i, dim3(1), dim3(1), 0, st_, cur_, (int)kept);
  LHIP(hipStreamSynchronize(st_));
}

void Model::forward(int n, const float* x, const float* params, float* logits) {
  ensure();
  if (n < 16 || n % 16) throw Error("learn: forward needs a positive multiple of 16 samples");
  std::vector<float> saved(NPAR);
  get_params(saved.data());
  std::vector<uint8_t> t((size_t)n * 4, 0);
  const bool was_trained = trained_;
  const int64_t kept = samples();
  if (kept + n > kTraceCap) throw Error("learn: trace store full");
  add_samples(n, x, t.data());
  std::lock_guard<std::mutex> lk(m_);
  upload_params(params);
  float* out = nullptr;
  LHIP(hipMalloc((void**)&out, sizeof(float) * NO * (size_t)n));
  hipLaunchKernelGGL(k_forward, dim3((unsigned)std::max(1, std::min(256, n / 16 / TW))), dim3(64 * TW), 0, st_,
                     tx_ + (size_t)kept * NF, n / 16, P_, Wq_, out);
  LHIP(hipGetLastError());
  LHIP(hipMemcpyAsync(logits, out, sizeof(float) * NO * (size_t)n, hipMemcpyDeviceToHost, st_));
  LHIP(hipStreamSynchronize(st_));
  LHIP(hipFree(out));
  upload_params(saved.data());
  trained_ = was_trained;
  hipLaunchKernelGGL(k_set_i, dim3(1), dim3(1), 0, st_, cur_, (int)kept);
  LHIP(hipStreamSynchronize(st_));
}

float host_logit(const std::vector<float>& P, const float* x, int out) {
  auto q = [](float v) { return bf2f(f2bf(v)); };
  float h1[NH], h2[NH];
  for (int o = 0; o < NH; o++) {
    float s = 0.f;
    for (int i = 0; i < NF; i++) s += q(x[i]) * q(P[OFF_W1 + i * NH + o]);
    s += P[OFF_B1 + o];
    h1[o] = q(s > 0.f ? s : 0.f);
  }
  for (int o = 0; o < NH; o++) {
    float s = 0.f;
    for (int i = 0; i < NH; i++) s += h1[i] * q(P[OFF_W2 + i * NH + o]);
    s += P[OFF_B2 + o];
    h2[o] = q(s > 0.f ? s : 0.f);
  }
  float s = 0.f;
  for (int i = 0; i < NH; i++) s += h2[i] * q(P[OFF_W3 + i * NO + out]);
  return s + P[OFF_B3 + out];
}

}  // namespace learn
}  // namespace rsa_amd
