// extern "C" entry points of librsa_mpi_amd.so (declared in include/rsa_mpi_amd.h).
// No C++ exception crosses this boundary: failures return non-zero and leave
// the message in a thread-local string.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rsa_mpi_amd.h"
#include "engine.hpp"
#include "learn.hpp"
#include "mt19937.hpp"

struct rsa_gpu_ctx {
  std::unique_ptr<rsa_amd::IEngine> eng;
};

namespace {
thread_local std::string g_err;
std::string g_info;

template <class F>
int guarded(F&& f) {
  try {
    f();
    g_err.clear();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  } catch (...) {
    g_err = "unknown error";
    return 1;
  }
}

std::vector<rsa_amd::RadiusClass> make_classes(int n, const double* radius, const uint64_t* count, const int* phase) {
  std::vector<rsa_amd::RadiusClass> c;
  for (int i = 0; i < n; i++) c.push_back({radius[i], count[i], phase ? phase[i] : 0});
  return c;
}

int need_ctx(rsa_gpu_ctx* ctx) {
  if (!ctx || !ctx->eng) {
    g_err = "null context";
    return 1;
  }
  return 0;
}
}  // namespace

extern "C" {

const char* rsa_gpu_last_error(void) { return g_err.c_str(); }

const char* rsa_gpu_device_info(void) {
  int dev = 0;
  hipDeviceProp_t p;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) {
    g_info = "no HIP device";
  } else {
    g_info = std::string(p.gcnArchName) + " / " + std::to_string(p.multiProcessorCount) + " CUs / " +
             std::to_string(p.totalGlobalMem >> 30) + " GiB";
  }
  return g_info.c_str();
}

int rsa_gpu_create(int dim, const double* inf, const double* sup, int ghost_layer, double domain_rad,
                   rsa_gpu_ctx** out) {
  return guarded([&] {
    if (!out) throw rsa_amd::Error("null output pointer");
    auto* c = new rsa_gpu_ctx();
    try {
      c->eng.reset(rsa_amd::make_engine(dim, inf, sup, ghost_layer, domain_rad));
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

void rsa_gpu_destroy(rsa_gpu_ctx* ctx) { delete ctx; }

int rsa_gpu_algo_init(rsa_gpu_ctx* ctx, int n, const double* radius, const uint64_t* count, const int* phase,
                      double exclusion, int nbshots_singledraw) {
  if (need_ctx(ctx)) return 1;
  return guarded([&] { ctx->eng->algo_init(make_classes(n, radius, count, phase), exclusion, nbshots_singledraw); });
}

int rsa_gpu_algo_init_fractions(rsa_gpu_ctx* ctx, int n, const double* radius, const double* fraction,
                                const int* phase, double volume, double exclusion, int nbshots_singledraw) {
  if (need_ctx(ctx)) return 1;
  return guarded([&] {
    // compute_desired_radius_nb_phase (radius_generator.ixx:71-86)
    std::vector<uint64_t> cnt(n);
    const int dim = ctx->eng->dim();
    double mult = 1;  // volumeUnitSphere (auxiliary/Geometry/Sphere.ixx:31-44)
    const double pi = M_PI;
    if (dim % 2 == 0) {
      for (int k = 0; k <= dim / 2 - 1; k++) mult *= (2 * pi) / (dim - 2 * k);
    } else {
      mult = 1. / pi;
      for (int k = 0; k <= (dim - 1) / 2; k++) mult *= (2 * pi) / (dim - 2 * k);
    }
    for (int i = 0; i < n; i++) {
      const double r = radius[i];
      double p = 0;  // puissance<D> (AuxiFunctions.ixx:34-44)
      if (dim == 2) p = r * r;
      if (dim == 3) p = r * (r * r);
      if (dim == 4) p = (r * r) * (r * r);
      cnt[i] = static_cast<uint64_t>(volume * fraction[i] / (mult * p));
    }
    ctx->eng->algo_init(make_classes(n, radius, cnt.data(), phase), exclusion, nbshots_singledraw);
  });
}

int rsa_gpu_algo_init_law(rsa_gpu_ctx* ctx, double min_radius, double max_radius, rsa_gpu_radius_law law,
                          void* user, uint64_t phase, uint64_t count, double exclusion, int nbshots_singledraw) {
  if (need_ctx(ctx)) return 1;
  return guarded(
      [&] { ctx->eng->algo_init_law(min_radius, max_radius, law, user, phase, count, exclusion, nbshots_singledraw); });
}

int rsa_gpu_reset_classes(rsa_gpu_ctx* ctx, int n, const double* radius, const uint64_t* count, const int* phase,
                          double exclusion) {
  if (need_ctx(ctx)) return 1;
  return guarded([&] { ctx->eng->reset_classes(make_classes(n, radius, count, phase), exclusion); });
}

int rsa_gpu_reset_law(rsa_gpu_ctx* ctx, double min_radius, double max_radius, rsa_gpu_radius_law law, void* user,
                      uint64_t phase, uint64_t count, double exclusion) {
  if (need_ctx(ctx)) return 1;
  return guarded([&] { ctx->eng->reset_law(min_radius, max_radius, law, user, phase, count, exclusion); });
}

int rsa_gpu_set_miss_rate(rsa_gpu_ctx* ctx, double miss_rate) {
  if (need_ctx(ctx)) return 1;
  return guarded([&] { ctx->eng->set_miss_rate(miss_rate); });
}

int rsa_gpu_proceed(rsa_gpu_ctx* ctx, uint64_t seed, int64_t max_draws, rsa_gpu_stats* stats) {
  if (need_ctx(ctx)) return 1;
  return guarded([&] {
    rsa_gpu_stats local;
    ctx->eng->proceed(seed, max_draws, stats ? stats : &local);
  });
}

int rsa_gpu_add_spheres(rsa_gpu_ctx* ctx, int64_t n, const double* pos, const double* radius,
                        const uint64_t* phase) {
  if (need_ctx(ctx)) return 1;
  return guarded([&] {
    std::vector<uint64_t> ph;
    if (!phase) {
      ph.assign(n, 0);
      phase = ph.data();
    }
    ctx->eng->add_spheres(n, pos, radius, phase);
  });
}

int rsa_gpu_count(rsa_gpu_ctx* ctx, uint64_t* n) {
  if (need_ctx(ctx)) return 1;
  return guarded([&] { *n = ctx->eng->count(); });
}

int rsa_gpu_extract(rsa_gpu_ctx* ctx, double* pos, double* radius, uint64_t* phase, uint64_t cap, uint64_t* n) {
  if (need_ctx(ctx)) return 1;
  return guarded([&] { *n = ctx->eng->extract(pos, radius, phase, cap); });
}

int rsa_gpu_extract_noalloc(rsa_gpu_ctx* ctx, double* pos, double* radius, uint64_t* phase, uint64_t cap,
                            uint64_t* n) {
  if (need_ctx(ctx)) return 1;
  return guarded([&] { *n = ctx->eng->extract_noalloc(pos, radius, phase, cap); });
}

int rsa_gpu_volume_of_spheres(rsa_gpu_ctx* ctx, double* volume) {
  if (need_ctx(ctx)) return 1;
  return guarded([&] { *volume = ctx->eng->volume_of_spheres(); });
}

int rsa_gpu_voxels(rsa_gpu_ctx* ctx, int64_t* coords, double* len, uint64_t cap, uint64_t* n) {
  if (need_ctx(ctx)) return 1;
  return guarded([&] { *n = ctx->eng->voxels(coords, len, cap); });
}

int rsa_gpu_profile_enable(rsa_gpu_ctx* ctx, int on) {
  if (need_ctx(ctx)) return 1;
  return guarded([&] { ctx->eng->profile_enable(on); });
}

int rsa_gpu_kernel_profile(rsa_gpu_ctx* ctx, int k, const char** name, int64_t* launches, double* seconds,
                           double* bytes) {
  if (need_ctx(ctx)) return 1;
  int found = 0;
  int rc = guarded([&] { found = ctx->eng->profile_get(k, name, launches, seconds, bytes) ? 1 : 0; });
  if (rc) return rc;
  if (!found) {
    g_err = "no such profile entry";
    return 2;
  }
  return 0;
}

int rsa_gpu_learn_reset(uint64_t seed) {
  return guarded([&] {
    rsa_amd::learn::Model::get().ensure();
    rsa_amd::learn::Model::get().reset(seed);
  });
}

int rsa_gpu_set_tuning(rsa_gpu_ctx* ctx, const rsa_gpu_tuning* t) {
  if (need_ctx(ctx)) return 1;
  return guarded([&] {
    if (!t) throw rsa_amd::Error("rsa_gpu_set_tuning: null tuning");
    ctx->eng->set_tuning(*t);
  });
}

int rsa_gpu_get_tuning(rsa_gpu_ctx* ctx, rsa_gpu_tuning* out) {
  if (need_ctx(ctx)) return 1;
  return guarded([&] {
    if (!out) throw rsa_amd::Error("rsa_gpu_get_tuning: null output");
    *out = ctx->eng->tuning();
  });
}

int rsa_gpu_learn_set_mode(rsa_gpu_ctx* ctx, int trace, int use, int measure) {
  if (need_ctx(ctx)) return 1;
  return guarded([&] { ctx->eng->learn_mode(trace, use, measure); });
}

int rsa_gpu_learn_get_stats(rsa_gpu_ctx* ctx, rsa_gpu_learn_stats* out) {
  if (need_ctx(ctx)) return 1;
  return guarded([&] { ctx->eng->learn_stats(out); });
}

int rsa_gpu_learn_samples(int64_t* n) {
  return guarded([&] { *n = rsa_amd::learn::Model::get().samples(); });
}

int rsa_gpu_learn_train(int steps, int batch, double lr, uint64_t seed, double* loss_first, double* loss_last,
                        int64_t* samples) {
  return guarded([&] {
    const int64_t n = rsa_amd::learn::Model::get().train(steps, batch, lr, seed, loss_first, loss_last);
    if (samples) *samples = n;
  });
}

int rsa_gpu_learn_params(float* get, const float* set) {
  return guarded([&] {
    if (get)
      rsa_amd::learn::Model::get().get_params(get);
    else if (set)
      rsa_amd::learn::Model::get().set_params(set);
    else
      throw rsa_amd::Error("learn: get or set must be given");
  });
}

int rsa_gpu_learn_gradcheck(int n, const float* x, const uint8_t* t, const float* params, float* grad,
                            double* loss) {
  return guarded([&] { rsa_amd::learn::Model::get().gradcheck(n, x, t, params, grad, loss); });
}

int rsa_gpu_learn_forward(int n, const float* x, const float* params, float* logits) {
  return guarded([&] { rsa_amd::learn::Model::get().forward(n, x, params, logits); });
}

int rsa_gpu_mt_words(uint64_t seed, uint64_t rank, uint64_t n, uint32_t* out) {
  return guarded([&] {
    rsa_amd::MTStream s;
    s.reset(seed, rank);
    s.ensure(n);
    std::memcpy(out, s.buf.data(), n * sizeof(uint32_t));
  });
}

int rsa_gpu_batch_run(int dim, int n, const double* side, const double* radius, const uint64_t* seed, int nbshots,
                      int workers, uint64_t* count, double* volume, double* pos, double* radius_out, uint64_t cap,
                      double* seconds) {
  rsa_gpu_batch_stats st;
  const int rc = rsa_gpu_batch_queue(dim, n, side, radius, seed, nbshots, std::max(1, workers), count, volume, pos,
                                     radius_out, cap, &st);
  if (rc == 0 && seconds) seconds[0] = st.seconds;
  return rc;
}

int rsa_gpu_batch_queue(int dim, int n, const double* side, const double* radius, const uint64_t* seed, int nbshots,
                        int active, uint64_t* count, double* volume, double* pos, double* radius_out, uint64_t cap,
                        rsa_gpu_batch_stats* stats) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!side || !radius || !seed || !count)))
      throw rsa_amd::Error("batch: null or negative arguments");
    if (dim < 2 || dim > 4) throw rsa_amd::Error("batch: dim must be 2, 3 or 4");
    if ((pos == nullptr) != (radius_out == nullptr)) throw rsa_amd::Error("batch: pos and radius_out go together");
    if (active < 1) throw rsa_amd::Error("batch: active must be at least 1");
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (n == 0) return;
    rsa_amd::batch_queue_run(dim, n, side, radius, seed, nbshots, active, count, volume, pos, radius_out, cap, stats);
  });
}

}  // extern "C"
