// TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.
// C-ABI over the CPU restatement in rsa_oracle.hpp, loaded with ctypes by tests/
// and by bench.py's cpu_baseline leg only.
#include "rsa_oracle.hpp"

#include <memory>
#include <string>

using namespace rsa_oracle;

namespace {

thread_local std::string g_err;
thread_local int g_dump_limit = 0;
// re-iterations with a uniform radius law instead of a single class (k -> r0, r1, phase, count)
struct ReiterLaw {
  double r0, r1;
  uint64_t phase, count;
};
thread_local std::vector<std::pair<int, ReiterLaw>> g_reiter_law;

struct Dump {
  std::vector<double> pos, rad;
  std::vector<uint64_t> prio;
  std::vector<uint8_t> acc;
  int outreach = 0;
};

struct Result {
  int dim = 0;
  std::vector<double> pos, rad;
  std::vector<uint64_t> phase;
  std::vector<DrawTrace> trace;
  uint64_t words = 0;
  std::string msg;
  std::vector<int64_t> vox;
  std::vector<double> vox_len;
  std::vector<Dump> dumps;
};

template <int D>
RadiusGen<D> make_gen(int n, const double* radius, const double* amount, const int* phase, int mode,
                      double excl, double volume) {
  if (mode == 1) {
    std::vector<std::tuple<double, double, int>> c;
    for (int i = 0; i < n; i++) c.emplace_back(radius[i], amount[i], phase[i]);
    return RadiusGen<D>::from_fractions(c, volume, excl);
  }
  if (mode == 0) {
    std::vector<std::tuple<double, uint64_t, int>> c;
    for (int i = 0; i < n; i++) c.emplace_back(radius[i], (uint64_t)amount[i], phase[i]);
    return RadiusGen<D>::from_counts(c, excl);
  }
  // mode 2: uniform law radius = r0 + (r1 - r0) * u  (test_random_radius_generator.cpp:15-29)
  const double r0 = radius[0], r1 = radius[1];
  return RadiusGen<D>::from_law(r0, r1, [r0, r1](double x) { return r0 + (r1 - r0) * x; }, (uint64_t)phase[0],
                                (uint64_t)amount[0], excl);
}

template <int D>
Result* run_impl(const double* inf, const double* sup, int gl, double drad, int n, const double* radius,
                 const double* amount, const int* phase, int mode, double excl, int nbshots, uint64_t seed,
                 int64_t max_draws, int compat, int n_reiter, const double* reiter_radius) {
  Pt<D> a{}, b{};
  for (int d = 0; d < D; d++) {
    a[d] = inf[d];
    b[d] = sup[d];
  }
  Domain<D> dom(a, b, gl, drad);
  RadiusGen<D> gen = make_gen<D>(n, radius, amount, phase, mode, excl, dom.total_volume());
  Algo<D> algo(dom, gen, nbshots);
  algo.compat_tab = compat != 0;
  algo.max_draws = max_draws;
  algo.dump_limit = g_dump_limit;
  algo.proceed(seed);
  // re-iteration (scripts/main_max_frac.cpp:43-52): new single-class generator, same seed
  for (int k = 0; k < n_reiter; k++) {
    const ReiterLaw* law = nullptr;
    for (const auto& e : g_reiter_law)
      if (e.first == k) law = &e.second;
    if (law) {  // RadiusGenerator(min, max, RandomRadiusGenerator*, count) (radius_generator.ixx:38-48)
      const double r0 = law->r0, r1 = law->r1;
      algo.reset_radius_generator(RadiusGen<D>::from_law(
          r0, r1, [r0, r1](double x) { return r0 + (r1 - r0) * x; }, law->phase, law->count, 0.));
    } else {
      std::vector<std::tuple<double, double, int>> c = {std::make_tuple(reiter_radius[k], 1., 0)};
      algo.reset_radius_generator(RadiusGen<D>::from_fractions(c, dom.total_volume()));
    }
    algo.proceed(seed);
  }
  auto* r = new Result();
  r->dim = D;
  std::vector<Pt<D>> p;
  dom.extract(p, r->rad, r->phase);
  r->pos.resize(p.size() * D);
  for (size_t i = 0; i < p.size(); i++)
    for (int d = 0; d < D; d++) r->pos[i * D + d] = p[i][d];
  r->trace = algo.trace;
  r->words = algo.words_used;
  r->vox = algo.last_vox;
  r->vox_len.assign(algo.last_len.begin(), algo.last_len.end());
  r->msg = algo.end_message;
  for (const auto& dd : algo.dumps) {
    Dump o;
    for (const auto& p : dd.pos)
      for (int d = 0; d < D; d++) o.pos.push_back(p[d]);
    o.rad = dd.rad;
    o.prio = dd.prio;
    o.acc = dd.accepted;
    o.outreach = dd.outreach;
    r->dumps.push_back(o);
  }
  return r;
}

template <int D>
Result* voxel_impl(const double* inf, const double* sup, int gl, double drad, double rmax, int64_t ns,
                   const double* pos, const double* rad, int64_t nv, const int64_t* coords, const double* len,
                   double rmin, int mode, int compat) {
  Pt<D> a{}, b{};
  for (int d = 0; d < D; d++) {
    a[d] = inf[d];
    b[d] = sup[d];
  }
  Domain<D> dom(a, b, gl, drad);
  dom.grid.define(rmax, gl, dom.inf, dom.sup);
  // insert as candidates, image them, then validate: the path every validated sphere takes
  // (build_grid + update_sphere_tags, operator_algorithm.ixx:446-473)
  for (int64_t i = 0; i < ns; i++) {
    Pt<D> p{};
    for (int d = 0; d < D; d++) p[d] = pos[i * D + d];
    dom.grid.cells[dom.grid.checked_cell(p)].add(0, 0, rad[i], p, TAG_UNDECIDED);
  }
  dom.fill_ghost_undecided();
  dom.put_ghost_into(TAG_VALIDATED);
  for (auto& c : dom.grid.cells)
    for (auto& t : c.tag) t = TAG_VALIDATED;
  Pt<D> width{};
  for (int d = 0; d < D; d++) width[d] = dom.sup[d] - dom.inf[d];
  Voxels<D> vox(dom.inf, width, 1.0, compat != 0);
  vox.c.resize(nv);
  for (int64_t i = 0; i < nv; i++)
    for (int d = 0; d < D; d++) vox.c[i][d] = coords[i * D + d];
  Pt<D> l{};
  for (int d = 0; d < D; d++) l[d] = len[d];
  vox.set_len(l);
  auto* r = new Result();
  r->dim = D;
  if (mode == 0 || mode == 2) vox.remove_covered(dom.grid, rmin);
  if (mode == 1 || mode == 2) vox.subdivide(dom.grid, rmin);
  if (mode == 3) {  // flags only, as 0/1 in the coordinate buffer
    auto f = vox.cover_flags(dom.grid, rmin);
    r->vox.assign(f.begin(), f.end());
    return r;
  }
  r->vox.resize(vox.size() * D);
  for (size_t i = 0; i < vox.size(); i++)
    for (int d = 0; d < D; d++) r->vox[i * D + d] = vox.c[i][d];
  r->vox_len.assign(vox.len.begin(), vox.len.end());
  return r;
}

}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// Full proceed<1>(seed) run.  mode: 0 = (radius, count, phase) classes,
// 1 = (radius, volume fraction, phase) classes over the domain volume,
// 2 = uniform radius law on [radius[0], radius[1]] with amount[0] spheres.
int orc_run(int dim, const double* inf, const double* sup, int gl, double drad, int n, const double* radius,
            const double* amount, const int* phase, int mode, double excl, int nbshots, uint64_t seed,
            int64_t max_draws, int compat, int n_reiter, const double* reiter_radius, void** out) {
  try {
    Result* r = nullptr;
    switch (dim) {
      case 2:
        r = run_impl<2>(inf, sup, gl, drad, n, radius, amount, phase, mode, excl, nbshots, seed, max_draws,
                        compat, n_reiter, reiter_radius);
        break;
      case 3:
        r = run_impl<3>(inf, sup, gl, drad, n, radius, amount, phase, mode, excl, nbshots, seed, max_draws,
                        compat, n_reiter, reiter_radius);
        break;
      case 4:
        r = run_impl<4>(inf, sup, gl, drad, n, radius, amount, phase, mode, excl, nbshots, seed, max_draws,
                        compat, n_reiter, reiter_radius);
        break;
      default:
        g_err = "oracle: dimension must be 2, 3 or 4";
        return 1;
    }
    *out = r;
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// One voxel step on frozen inputs.  mode 0 remove_covered, 1 subdivide_uncovered,
// 2 both, 3 coverage flags only.
int orc_voxel_step(int dim, const double* inf, const double* sup, int gl, double drad, double rmax, int64_t ns,
                   const double* pos, const double* rad, int64_t nv, const int64_t* coords, const double* len,
                   double rmin, int mode, int compat, void** out) {
  try {
    Result* r = nullptr;
    switch (dim) {
      case 2:
        r = voxel_impl<2>(inf, sup, gl, drad, rmax, ns, pos, rad, nv, coords, len, rmin, mode, compat);
        break;
      case 3:
        r = voxel_impl<3>(inf, sup, gl, drad, rmax, ns, pos, rad, nv, coords, len, rmin, mode, compat);
        break;
      case 4:
        r = voxel_impl<4>(inf, sup, gl, drad, rmax, ns, pos, rad, nv, coords, len, rmin, mode, compat);
        break;
      default:
        g_err = "oracle: dimension must be 2, 3 or 4";
        return 1;
    }
    *out = r;
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int64_t orc_num_spheres(void* h) { return (int64_t)static_cast<Result*>(h)->rad.size(); }
void orc_spheres(void* h, double* pos, double* rad, uint64_t* phase) {
  auto* r = static_cast<Result*>(h);
  std::copy(r->pos.begin(), r->pos.end(), pos);
  std::copy(r->rad.begin(), r->rad.end(), rad);
  std::copy(r->phase.begin(), r->phase.end(), phase);
}
int64_t orc_num_draws(void* h) { return (int64_t)static_cast<Result*>(h)->trace.size(); }
// per draw: draw, nvox_before, shots, accepted, rounds, refined, nvox_after,
//           words_before, words_after_poisson, words_after_draw  (10 int64) + miss rate
void orc_trace(void* h, int64_t* ints, double* miss) {
  auto* r = static_cast<Result*>(h);
  for (size_t i = 0; i < r->trace.size(); i++) {
    const DrawTrace& t = r->trace[i];
    int64_t* o = ints + 10 * i;
    o[0] = t.draw;
    o[1] = t.nvox_before;
    o[2] = t.shots;
    o[3] = t.accepted;
    o[4] = t.rounds;
    o[5] = t.refined;
    o[6] = t.nvox_after;
    o[7] = (int64_t)t.words_before;
    o[8] = (int64_t)t.words_after_poisson;
    o[9] = (int64_t)t.words_after_draw;
    miss[i] = t.miss_rate;
  }
}
uint64_t orc_words_used(void* h) { return static_cast<Result*>(h)->words; }
const char* orc_end_message(void* h) { return static_cast<Result*>(h)->msg.c_str(); }
int64_t orc_num_voxel_values(void* h) { return (int64_t)static_cast<Result*>(h)->vox.size(); }
void orc_voxels(void* h, int64_t* coords, double* len) {
  auto* r = static_cast<Result*>(h);
  std::copy(r->vox.begin(), r->vox.end(), coords);
  if (len) std::copy(r->vox_len.begin(), r->vox_len.end(), len);
}
void orc_free(void* h) { delete static_cast<Result*>(h); }

// record the candidates of the first k draws of the next orc_run on this thread
void orc_set_dump(int k) { g_dump_limit = k; }
// re-iteration k of the next orc_run on this thread uses the law r0 + (r1 - r0) u (count spheres
// of `phase`); k < 0 clears every such setting
void orc_set_reiter_law(int k, double r0, double r1, uint64_t phase, uint64_t count) {
  if (k < 0) {
    g_reiter_law.clear();
    return;
  }
  g_reiter_law.push_back({k, ReiterLaw{r0, r1, phase, count}});
}
int64_t orc_num_dumps(void* h) { return (int64_t)static_cast<Result*>(h)->dumps.size(); }
int64_t orc_dump_size(void* h, int64_t k) { return (int64_t)static_cast<Result*>(h)->dumps[k].rad.size(); }
int orc_dump(void* h, int64_t k, double* pos, double* rad, uint64_t* prio, uint8_t* acc) {
  const Dump& d = static_cast<Result*>(h)->dumps[k];
  std::copy(d.pos.begin(), d.pos.end(), pos);
  std::copy(d.rad.begin(), d.rad.end(), rad);
  std::copy(d.prio.begin(), d.prio.end(), prio);
  std::copy(d.acc.begin(), d.acc.end(), acc);
  return d.outreach;
}

// raw words of std::mt19937 seeded with seed_seq{rank, seed} (rsa_random.hxx:37-42)
void orc_mt_words(uint64_t seed, uint64_t rank, uint64_t n, uint32_t* out) {
  CountingMT g;
  seed_engine(g, seed, rank);
  for (uint64_t i = 0; i < n; i++) out[i] = (uint32_t)g();
}

// default-constructed std::mt19937 (seed 5489), for the C++ standard's known answer
void orc_mt_default_words(uint64_t n, uint32_t* out) {
  std::mt19937 g;
  for (uint64_t i = 0; i < n; i++) out[i] = (uint32_t)g();
}

// libstdc++ distributions on that stream, for known-answer checks:
// kind 0: uniform_int_distribution<int>(0, hi), kind 1: uniform_real_distribution<double>(0,1),
// kind 2: poisson_distribution<int>(mean).  Returns the words consumed.
uint64_t orc_dist_samples(uint64_t seed, int kind, double param, int64_t n, double* out) {
  CountingMT g;
  seed_engine(g, seed);
  if (kind == 0) {
    std::uniform_int_distribution<> u(0, (int)param);
    for (int64_t i = 0; i < n; i++) out[i] = u(g);
  } else if (kind == 1) {
    std::uniform_real_distribution<> u(0, 1);
    for (int64_t i = 0; i < n; i++) out[i] = u(g);
  } else {
    for (int64_t i = 0; i < n; i++) {
      std::poisson_distribution<> p(param);
      out[i] = p(g);
    }
  }
  return g.used;
}

// median_of_medians::find_pivot_for_first_elements on one rank (the oracle's mom_pivot):
// returns 0 and the pivot, or 1 with the reference's error message
int orc_mom_pivot(const uint64_t* values, int64_t n, int64_t nb_first, double* pivot) {
  try {
    *pivot = mom_pivot(std::vector<uint64_t>(values, values + n), nb_first);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// the reference's visiting order of the 3^D stencil (rsa_grid_traversal.ixx:30-66)
int orc_neighbour_table(int dim, const int* n, int* out) {
  auto build = [&](auto tag) {
    constexpr int D = decltype(tag)::value;
    std::array<int, D> nn{};
    for (int d = 0; d < D; d++) nn[d] = n[d];
    std::vector<int> raw;
    const std::vector<int> order = Grid<D>::stencil_order(nn, raw);
    for (size_t i = 0; i < order.size(); i++) out[i] = order[i];
  };
  if (dim == 2) build(std::integral_constant<int, 2>{});
  else if (dim == 3) build(std::integral_constant<int, 3>{});
  else if (dim == 4) build(std::integral_constant<int, 4>{});
  else return 1;
  return 0;
}

}  // extern "C"
