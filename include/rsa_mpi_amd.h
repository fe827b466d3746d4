/*
 * C-ABI of librsa_mpi_amd.so: the single-rank bias-free RSA hot path of
 * MarcJos/RSA_MPI on one MI355X (gfx950).
 *
 * Every entry point takes plain pointers and sizes, returns 0 on success and a
 * non-zero code on failure, never throws, and leaves a thread-local message in
 * rsa_gpu_last_error().  Host arrays are caller-owned and copied.  One context
 * owns all device buffers of one domain; it is not thread-safe.
 *
 * The C++ drop-in headers under include/rsa_mpi_amd/ forward the reference's
 * operator API to these functions; the citation on each line names the
 * reference interface it replaces.
 */
#ifndef RSA_MPI_AMD_H
#define RSA_MPI_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rsa_gpu_ctx rsa_gpu_ctx;

/* Per-proceed statistics (the quantities proceed_voxel logs or asserts,
 * operator_algorithm.ixx:104-187). */
typedef struct rsa_gpu_stats {
  int64_t loop_iterations;  /* draw-loop iterations, zero-shot ones included */
  int64_t draws;            /* iterations that drew at least one shot */
  int64_t shots;            /* total candidate centres drawn */
  int64_t accepted;         /* spheres validated by this proceed */
  int64_t refines;          /* remove_covered + subdivide_uncovered passes */
  int64_t rounds;           /* add_to_sample rounds over all draws */
  int64_t rounds_max;       /* largest number of rounds in one draw */
  int64_t nvox_max;         /* largest voxel-list length */
  int64_t nvox_final;       /* voxel-list length at exit (0 = fully packed) */
  int64_t words;            /* MT19937 words consumed */
  int64_t ghost_cell_drops; /* candidates whose centre rounded into a ghost cell */
  int64_t image_cell_mismatch; /* periodic images whose computed cell is not the mirrored one */
  int end_reason;           /* 1 fully packed, 2 lack of new radii, 3 draw limit */
  double seconds;           /* wall time of rsa_gpu_proceed */
  double voxel_seconds;     /* GPU time in refine passes (device events) */
  /* End-of-run checks, run on the device by every proceed (counterparts of the
   * reference's debug asserts, operator_algorithm.ixx:269,563,
   * rsa_data_storage.ixx:229-243, operator_check_cell.ixx:36-78): */
  int64_t overlap_pairs;    /* pairs of placed spheres (images included) that collide */
  int64_t cover_samples;    /* points sampled over the box for maximality */
  int64_t cover_free;       /* of those, points farther than r + r_min from every sphere */
} rsa_gpu_stats;

/* Counters of the learned predictor for the last proceed of a context (see
 * rsa_gpu_learn_set_mode).  No counterpart in the reference. */
typedef struct rsa_gpu_learn_stats {
  int64_t voxels_predicted;         /* voxels whose coverage the MLP predicted (learned passes) */
  int64_t voxels_mispredicted;      /* of those, predicted covered != covered by the exact test */
  int64_t voxels_covered;           /* covered voxels of the measured passes (every pass in measure
                                       or use mode) */
  int64_t voxels_predicted_covered; /* voxels predicted covered */
  int64_t voxels_covered_predicted; /* covered voxels of the learned passes (the constant "covered"
                                       predictor misses voxels_predicted - this many) */
  int64_t waves;                    /* waves of 64 voxels in the coverage kernel's walk (measured passes) */
  int64_t waves_uniform;            /* of those, waves whose 64 voxels share one outcome */
  int64_t draws_predicted;          /* draws whose refinement the MLP predicted (learned mode) */
  int64_t draws_mispredicted;       /* of those, predicted refinement != the reference's decision */
  int64_t refines;                  /* refinements of those draws */
  int64_t refines_predicted;        /* refinements predicted */
  int64_t speculation_skipped;      /* next draws not queued ahead because a refinement was predicted */
  int64_t traced_voxels;            /* voxel samples appended to the trace store */
  int64_t traced_draws;             /* draw samples appended */
} rsa_gpu_learn_stats;

/* Last error message of the calling thread ("" if none). */
const char* rsa_gpu_last_error(void);

/* Device and build description, e.g. "gfx950 / 256 CUs". */
const char* rsa_gpu_device_info(void);

/* Periodic domain [inf, sup] of dimension dim (2, 3 or 4) with ghost_layer
 * (>= 1) cell layers and implicit radius domain_rad.
 * Replaces rsa_domain<DIM>::rsa_domain (rsa_domain.hxx:53-54, rsa_domain.ixx:6-28). */
int rsa_gpu_create(int dim, const double* inf, const double* sup, int ghost_layer, double domain_rad,
                   rsa_gpu_ctx** out);

/* Releases the context and its device memory. */
void rsa_gpu_destroy(rsa_gpu_ctx* ctx);

/* rsa_algo<DIM> constructor (operator_algorithm.ixx:15-30): installs the radius
 * classes (radius, count, phase) of a sac_de_billes::RadiusGenerator<DIM>
 * (radius_generator.ixx:9-28: sorted nonincreasing, exclusion added) and
 * rebuilds the cell grid with the largest radius, which discards the spheres
 * already placed.  nbshots_singledraw is m_nbshots_singledraw. */
int rsa_gpu_algo_init(rsa_gpu_ctx* ctx, int n, const double* radius, const uint64_t* count, const int* phase,
                      double exclusion, int nbshots_singledraw);

/* RadiusGenerator (radius, volume fraction, phase) form over volume
 * (radius_generator.ixx:30-36, 71-86), followed by the rsa_algo constructor. */
int rsa_gpu_algo_init_fractions(rsa_gpu_ctx* ctx, int n, const double* radius, const double* fraction,
                                const int* phase, double volume, double exclusion, int nbshots_singledraw);

/* A custom radius law: radius[i] = law(u[i]) for i < n, u uniform in [0, 1).
 * It must be a pure function of u (prepared draws that a refinement discards
 * are evaluated again). */
typedef void (*rsa_gpu_radius_law)(const double* u, double* radius, int64_t n, void* user);

/* rsa_algo constructor with RadiusGenerator(min_radius, max_radius,
 * RandomRadiusGenerator*, count, exclusion) (radius_generator.hxx:62-70,
 * radius_generator.ixx:38-48): `count` spheres whose radii are law(u), u drawn
 * from 2 words per shot after the priorities (radius_generator.hxx:25-33), all
 * of phase `phase`.  The law runs on a host thread of the library, once per
 * draw for all its shots; its radii must lie in [min_radius, max_radius]. */
int rsa_gpu_algo_init_law(rsa_gpu_ctx* ctx, double min_radius, double max_radius, rsa_gpu_radius_law law,
                          void* user, uint64_t phase, uint64_t count, double exclusion, int nbshots_singledraw);

/* rsa_algo<DIM>::reset_radius_generator (operator_algorithm.ixx:238-250): new
 * classes, placed spheres kept. */
int rsa_gpu_reset_classes(rsa_gpu_ctx* ctx, int n, const double* radius, const uint64_t* count,
                          const int* phase, double exclusion);

/* rsa_algo<DIM>::reset_radius_generator (operator_algorithm.ixx:238-250) to a
 * RadiusGenerator(min_radius, max_radius, RandomRadiusGenerator*, count,
 * exclusion) (radius_generator.ixx:38-48): the custom law of rsa_gpu_algo_init_law
 * for the next proceed, placed spheres kept; a max_radius above the current
 * generator's fails with "Incorrect maximal radius". */
int rsa_gpu_reset_law(rsa_gpu_ctx* ctx, double min_radius, double max_radius, rsa_gpu_radius_law law, void* user,
                      uint64_t phase, uint64_t count, double exclusion);

/* rsa_algo<DIM>::set_miss_rate (operator_algorithm.hxx:66). */
int rsa_gpu_set_miss_rate(rsa_gpu_ctx* ctx, double miss_rate);

/* rsa_algo<DIM>::proceed<1>(seed) (operator_algorithm.ixx:225-236, 85-188).
 * max_draws < 0 runs to the end; otherwise stops after that many drawing
 * iterations (test hook).  stats may be NULL. */
int rsa_gpu_proceed(rsa_gpu_ctx* ctx, uint64_t seed, int64_t max_draws, rsa_gpu_stats* stats);

/* rsa_domain<DIM>::add (rsa_domain.hxx:84): n spheres (pos is n*dim, row-major)
 * inserted as validated.  Priorities are not kept, as in the reference. */
int rsa_gpu_add_spheres(rsa_gpu_ctx* ctx, int64_t n, const double* pos, const double* radius,
                        const uint64_t* phase);

/* rsa_grid<DIM>::get_number_of_spheres_fast (rsa_grid.ixx:102-121). */
int rsa_gpu_count(rsa_gpu_ctx* ctx, uint64_t* n);

/* rsa_grid<DIM>::extract_spheres / extract_data<Real> order (rsa_grid.ixx:73-98,
 * 139-160): real cells in index order, then in-cell order.  pos is cap*dim;
 * phase may be NULL.  With pos NULL only *n is set. */
int rsa_gpu_extract(rsa_gpu_ctx* ctx, double* pos, double* radius, uint64_t* phase, uint64_t cap,
                    uint64_t* n);

/* rsa_gpu_extract into caller-owned DEVICE memory, allocating nothing: pos
 * (cap*dim doubles), radius (cap doubles) and phase (cap uint64, may be NULL)
 * are device pointers on the context's GPU; with pos NULL only *n is set.  No
 * counterpart in the reference: it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its scratch (per-block counts) is a static device array of
 * 4 MiB per GPU, which limits the cell grid to 2^28 cells (larger grids are
 * refused with an error); calls that overlap on one GPU share it and race. */
int rsa_gpu_extract_noalloc(rsa_gpu_ctx* ctx, double* pos, double* radius, uint64_t* phase, uint64_t cap,
                            uint64_t* n);

/* Total volume of the placed spheres (rsa_domain<DIM>::compute_total_volume_of_spheres,
 * rsa_domain.ixx:170-175, summed per real cell in index order). */
int rsa_gpu_volume_of_spheres(rsa_gpu_ctx* ctx, double* volume);

/* Current voxel list (list_of_voxels<DIM> state after the last proceed):
 * coords is cap*dim int64 lattice coordinates, len the common voxel length. */
int rsa_gpu_voxels(rsa_gpu_ctx* ctx, int64_t* coords, double* len, uint64_t cap, uint64_t* n);

/* Per-kernel device timing, off by default (level 0).  Level 1 brackets the
 * voxel-refine kernels, level 2 also the per-draw kernels, with HIP events on
 * the context's stream, and accumulates their compulsory ("algorithmic") bytes.  No counterpart in the reference (its
 * only timers are wall clocks, rsa_test_perf.hxx:21-25). */
int rsa_gpu_profile_enable(rsa_gpu_ctx* ctx, int level);

/* Entry k of the kernel profile: name, launches, summed seconds, summed
 * algorithmic bytes.  Returns 2 when k is past the last entry. */
int rsa_gpu_kernel_profile(rsa_gpu_ctx* ctx, int k, const char** name, int64_t* launches, double* seconds,
                           double* bytes);

/* ---- Learned predictor (no counterpart in the reference).  A 16 -> 64 -> 64
 * -> 2 MLP (ReLU; BF16 weights and activations, FP32 accumulation on the
 * matrix cores) per process.  Output 0 predicts whether a voxel is covered at a
 * refinement pass, output 1 whether a draw triggers a refinement.  It only
 * orders work: the coverage kernel walks predicted-covered voxels first, and a
 * predicted refinement stops the next draw being queued ahead.  Voxels are
 * removed by the exact FP64 test and refinements follow the reference's
 * miss-rate rule, so placed sets are unchanged. */

/* Fresh parameters (seeded) and an empty trace store. */
int rsa_gpu_learn_reset(uint64_t seed);

/* Refinement thresholds of a context (no counterpart in the reference: they
 * choose how the device walks a voxel list, never what it computes).  Lists of
 * at least group_min voxels are walked in cell-grouped order; lists of at least
 * pack_min read the placed spheres from a packed copy in cell order (D <= 3).
 * Both default to 2^20; both must be >= 1 (INT64_MAX switches one off). */
typedef struct rsa_gpu_tuning {
  int64_t group_min;
  int64_t pack_min;
} rsa_gpu_tuning;
int rsa_gpu_set_tuning(rsa_gpu_ctx* ctx, const rsa_gpu_tuning* t);
int rsa_gpu_get_tuning(rsa_gpu_ctx* ctx, rsa_gpu_tuning* out);

/* Per context: trace (append voxel and draw samples of each proceed to the
 * trace store), use (the trained model orders the coverage walk and the draw
 * queue), measure (count outcomes and wave uniformity even without use). */
int rsa_gpu_learn_set_mode(rsa_gpu_ctx* ctx, int trace, int use, int measure);

/* The counters of the context's last proceed. */
int rsa_gpu_learn_get_stats(rsa_gpu_ctx* ctx, rsa_gpu_learn_stats* out);

/* Samples in the trace store. */
int rsa_gpu_learn_samples(int64_t* n);

/* Trains the model on the trace store on the GPU: steps of Adam (learning rate
 * lr) on batches of `batch` samples (a multiple of 16) drawn uniformly; the
 * loss of the first and last batch, and the samples used, are returned. */
int rsa_gpu_learn_train(int steps, int batch, double lr, uint64_t seed, double* loss_first, double* loss_last,
                        int64_t* samples);

/* The model's FP32 parameters (5378 floats: W1[16][64], b1[64], W2[64][64],
 * b2[64], W3[64][2], b3[2]), read into get when non-NULL, else set from set. */
int rsa_gpu_learn_params(float* get, const float* set);

/* Test hooks: one forward and backward pass on the GPU kernels over n samples
 * (n a multiple of 16; x is n*16 features, t is n*4 bytes y0, m0, y1, m1) with
 * the given parameters: gradients of the mean loss into grad (5378 floats)
 * and the loss; and the logits (n*2) of a forward pass.  The model and the
 * trace store are left as they were. */
int rsa_gpu_learn_gradcheck(int n, const float* x, const uint8_t* t, const float* params, float* grad,
                            double* loss);
int rsa_gpu_learn_forward(int n, const float* x, const float* params, float* logits);

/* Raw MT19937 words of std::mt19937 seeded with seed_seq{rank, seed}
 * (rsa_random.hxx:37-42), from the library's own generator. */
int rsa_gpu_mt_words(uint64_t seed, uint64_t rank, uint64_t n, uint32_t* out);

/* Independent realizations (the batched form).  Realization k is
 * uniform_generate<DIM, 1>(domain, radius[k], nbshots, 10, seed[k])
 * (operator_algorithm.ixx:275-292) on the periodic domain [0, side[k]]^dim
 * with ghost layer 1 and domain radius radius[k], run to maximal packing: what
 * the reference runs once per process.  The realizations run as the device
 * work queue of rsa_gpu_batch_queue with `workers` of them in flight; results
 * do not depend on workers.  count[k] receives the spheres of realization k,
 * volume[k] (volume may be NULL) their total volume.  With pos non-NULL, every
 * realization's spheres are extracted (extract order), realization after
 * realization in index order (pos is cap*dim, radius_out cap); their total
 * must fit cap.  seconds (may be NULL) receives the batch's wall time.  No
 * counterpart in the reference, which runs one realization per process. */
int rsa_gpu_batch_run(int dim, int n, const double* side, const double* radius, const uint64_t* seed,
                      int nbshots, int workers, uint64_t* count, double* volume, double* pos,
                      double* radius_out, uint64_t cap, double* seconds);

/* The batched form as a device work queue: realization k is the same
 * uniform_generate<DIM, 1> run as in rsa_gpu_batch_run (operator_algorithm.ixx:275-292),
 * but up to `active` realizations (at most 256) are in flight together on ONE
 * stream driven by one host thread: every draw round and every refinement step
 * of all the realizations that need it is one kernel launch over a table of
 * per-realization descriptors, and a finished realization's place is taken by
 * the next one.  Each realization's sequence of words, draws, rounds and
 * refinements is the one rsa_gpu_proceed runs, so its placed set is the same
 * (a refinement may wait for other realizations' draws: nothing it reads
 * changes meanwhile).  Arguments and outputs as rsa_gpu_batch_run; stats (may
 * be NULL) receives the launch-group counts.  No counterpart in the reference,
 * which runs one realization per process. */
typedef struct rsa_gpu_batch_stats {
  double seconds;               /* wall time of the call */
  int64_t ticks;                /* launch groups: draw rounds plus refinement groups */
  int64_t draw_launch_groups;   /* draw rounds launched together */
  int64_t refine_launch_groups; /* refinements launched together */
  int64_t max_active;           /* realizations in flight at most */
  int64_t draws;                /* draws of all realizations */
  int64_t refines;              /* refinements of all realizations */
  /* host time of the driver thread by phase (seconds): starting realizations,
   * preparing draws (prepared words, word copies, descriptors), waiting for the
   * rounds' results, refinement groups (launches and their one wait), finishing
   * realizations (end-of-run checks, extraction) */
  double t_admit, t_plan, t_wait, t_refine, t_finish;
  /* parts of t_plan: waiting for prepared draws (helper threads), making their
   * words readable on the device; part of t_refine: waiting for the counting passes */
  double t_prep, t_words, t_refine_wait;
} rsa_gpu_batch_stats;

int rsa_gpu_batch_queue(int dim, int n, const double* side, const double* radius, const uint64_t* seed,
                        int nbshots, int active, uint64_t* count, double* volume, double* pos,
                        double* radius_out, uint64_t cap, rsa_gpu_batch_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* RSA_MPI_AMD_H */
