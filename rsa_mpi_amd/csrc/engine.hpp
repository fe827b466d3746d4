// Host-side interface of one single-rank RSA context (one periodic domain on
// one GPU).  Implemented per dimension in engine.hip; the C-ABI in capi.cpp
// only sees this interface.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rsa_mpi_amd.h"

namespace rsa_amd {

struct RadiusClass {
  double radius;
  uint64_t count;
  int phase;
};

class IEngine {
 public:
  virtual ~IEngine() = default;
  virtual int dim() const = 0;
  virtual void algo_init(const std::vector<RadiusClass>& classes, double exclusion, int nbshots) = 0;
  virtual void algo_init_law(double rmin, double rmax, rsa_gpu_radius_law law, void* user, uint64_t phase,
                             uint64_t count, double exclusion, int nbshots) = 0;
  virtual void reset_classes(const std::vector<RadiusClass>& classes, double exclusion) = 0;
  virtual void reset_law(double rmin, double rmax, rsa_gpu_radius_law law, void* user, uint64_t phase,
                         uint64_t count, double exclusion) = 0;
  virtual void set_miss_rate(double m) = 0;
  virtual void proceed(uint64_t seed, int64_t max_draws, rsa_gpu_stats* st) = 0;
  virtual void add_spheres(int64_t n, const double* pos, const double* rad, const uint64_t* phase) = 0;
  virtual uint64_t count() const = 0;
  virtual uint64_t extract(double* pos, double* rad, uint64_t* phase, uint64_t cap) = 0;
  virtual uint64_t extract_noalloc(double* pos, double* rad, uint64_t* phase, uint64_t cap) = 0;
  virtual double volume_of_spheres() = 0;
  virtual uint64_t voxels(int64_t* coords, double* len, uint64_t cap) = 0;
  virtual double total_volume() const = 0;
  // learned predictor (learn.hip): trace logging, use for ordering and speculation, outcome counts
  virtual void set_tuning(const rsa_gpu_tuning& t) = 0;
  virtual rsa_gpu_tuning tuning() const = 0;
  virtual void learn_mode(int trace, int use, int measure) = 0;
  virtual void learn_stats(rsa_gpu_learn_stats* out) = 0;
  // per-kernel device timings (HIP events on the context's stream)
  virtual void profile_enable(int level) = 0;  // 0 off, 1 refine kernels, 2 all
  virtual bool profile_get(int k, const char** name, int64_t* launches, double* seconds, double* bytes) = 0;
};

IEngine* make_engine(int dim, const double* inf, const double* sup, int ghost_layer, double domain_rad);

// rsa_gpu_batch_queue: realizations in lockstep on one stream (engine.hip, batch.inl)
void batch_queue_run(int dim, int n, const double* side, const double* radius, const uint64_t* seed, int nbshots,
                     int active, uint64_t* count, double* volume, double* pos, double* radius_out, uint64_t cap,
                     rsa_gpu_batch_stats* stats);

// reference error messages (rsa_grid_traversal.ixx:22, operator_algorithm.ixx:28,246,
// radius_generator.ixx:18, list_of_voxels.ixx:113, rsa_domain.ixx:140)
struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

}  // namespace rsa_amd
