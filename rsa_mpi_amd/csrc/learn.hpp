// Host interface of the learned predictor (learn.hip): one model per process,
// its trace store, training, and the pointers the engine's kernels read.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <vector>

#include "mlp.hpp"

namespace rsa_amd {
namespace learn {

// A training sample as stored on the device: NF bf16 features, then the two
// targets with their masks (y0, m0: the voxel is covered at this refinement;
// y1, m1: this draw triggers a refinement).
constexpr int kTraceCap = 1 << 20;

class Model {
 public:
  static Model& get();
  // device buffers (allocated on first use, kept for the process)
  void ensure();
  void reset(uint64_t seed);  // fresh parameters, empty trace store
  void set_params(const float* p);
  void get_params(float* p);
  // host copy of the parameters, refreshed after training and set_params (for the draw head)
  const std::vector<float>& host_params() const { return hp_; }
  bool trained() const { return trained_; }
  // trains on the stored samples with the BF16 MFMA forward/backward kernels and Adam;
  // returns the samples it drew from
  int64_t train(int steps, int batch, double lr, uint64_t seed, double* loss_first, double* loss_last);
  // one forward + backward over n samples (n a multiple of 16), gradients of the mean loss
  void gradcheck(int n, const float* x, const uint8_t* t, const float* params, float* grad, double* loss);
  void forward(int n, const float* x, const float* params, float* logits);
  void add_samples(int64_t n, const float* x, const uint8_t* t);  // appended to the trace store
  int64_t samples();  // samples in the trace store
  void clear_samples();

  const float* P() const { return P_; }
  const uint16_t* Wq() const { return Wq_; }
  uint16_t* tx() { return tx_; }
  uint8_t* tt() { return tt_; }
  int* cursor() { return cur_; }
  std::mutex& mutex() { return m_; }

 private:
  Model() = default;
  void upload_params(const float* p);
  void refresh_host();
  std::mutex m_;
  hipStream_t st_ = nullptr;
  float *P_ = nullptr, *M_ = nullptr, *V_ = nullptr, *G_ = nullptr, *loss_ = nullptr;
  uint16_t* Wq_ = nullptr;
  uint16_t* tx_ = nullptr;
  uint8_t* tt_ = nullptr;
  int* cur_ = nullptr;
  float* hstage_ = nullptr;  // pinned, mapped staging for host -> device copies by kernels
  size_t hstage_bytes_ = 0;
  std::vector<float> hp_;
  bool trained_ = false;
  int64_t adam_t_ = 0;
  void* stage(size_t bytes);
};

// CPU evaluation of the model with its BF16 weights (the draw head, one sample
// per draw): the same products the MFMA kernels make, accumulated in float.
float host_logit(const std::vector<float>& P, const float* x, int out);

}  // namespace learn
}  // namespace rsa_amd
