// C++ drop-in for the operator API of MarcJos/RSA_MPI's single-rank voxel path,
// header-only, forwarding to librsa_mpi_amd.so (include/rsa_mpi_amd.h).
//
// Callers of the reference keep their code: the classes below carry the
// reference's names, namespaces, template parameters and argument meaning, and
// raise std::runtime_error with the reference's messages.  Build with
//   -I<repo>/include/rsa_mpi_amd  -L<repo>/rsa_mpi_amd/lib -lrsa_mpi_amd
// (the file names under include/rsa_mpi_amd/ are those the reference's callers
// include: rsa_domain.hxx, operator_algorithm.hxx, radius_generator.hxx,
// user_interface.hxx, test.hxx, write_xyz.hpp, rsa_decoration.hxx, ...).
//
// Scope: one process, one GPU, the whole domain (the reference's single-rank
// run); proceed<1> (the voxel method).  State lives on the device: the host
// objects are handles, and extraction copies spheres back.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iostream>
#include <memory>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../rsa_mpi_amd.h"

namespace rsa_mpi_amd_detail {
inline void check(int rc) {
  if (rc != 0) throw std::runtime_error(rsa_gpu_last_error());
}
}  // namespace rsa_mpi_amd_detail

// ------------------------------------------------------------ geometry types
namespace sac_de_billes {
using std::get;
using std::tuple;
using std::vector;

// auxiliary/Geometry/GeomTypes.hxx:19-27
using PhaseType = long;
template <unsigned short DIM>
using Point = std::array<double, DIM>;

template <unsigned short DIM>
std::array<double, DIM> create_array(double v) {
  std::array<double, DIM> a;
  a.fill(v);
  return a;
}

// auxiliary/Geometry/Sphere.hxx:18-32
template <unsigned short DIM>
class Sphere {
 public:
  Sphere() : center{create_array<DIM>(0.)}, radius{0}, phase{0} {}
  Sphere(const Point<DIM>& center_, double radius_, PhaseType phase_)
      : center(center_), radius(radius_), phase{phase_} {}
  Point<DIM> center;
  double radius;
  PhaseType phase;
  //! volume of the ball (Sphere.ixx:30-57)
  double volume() const {
    double mult = 1.;
    const double pi = M_PI;
    if (DIM % 2 == 0) {
      for (int k = 0; k <= DIM / 2 - 1; k++) mult *= (2 * pi) / (DIM - 2 * k);
    } else {
      mult = 1. / pi;
      for (int k = 0; k <= (DIM - 1) / 2; k++) mult *= (2 * pi) / (DIM - 2 * k);
    }
    return mult * std::pow(radius, (int)DIM);
  }
  //! is this point inside?
  bool isInside(const Point<DIM>& x) const {
    double s = 0.;
    for (int d = 0; d < DIM; d++) s += (x[d] - center[d]) * (x[d] - center[d]);
    return s < radius * radius;
  }
};

// radius_generator.hxx:12-38: radius = transform(u), u uniform in [0, 1]
class RandomRadiusGenerator {
 public:
  RandomRadiusGenerator(std::function<double(double)> transform_, uint64_t phase_)
      : transform{std::move(transform_)}, phase{phase_} {}
  //! radius_generator.hxx:25-33: (phases, radii) of a_size shots, u uniform in [0, 1)
  std::tuple<std::vector<int>, std::vector<double>> operator()(size_t a_size, std::mt19937& random_generator) const {
    std::uniform_real_distribution<> ulaw(0., 1.);
    std::tuple<std::vector<int>, std::vector<double>> result(std::vector<int>(a_size, (int)phase),
                                                             std::vector<double>(a_size, 0.));
    for (size_t i = 0; i < a_size; i++) std::get<1>(result)[i] = transform(ulaw(random_generator));
    return result;
  }
  std::function<double(double)> transform;
  uint64_t phase;
};

// radius_generator.hxx:89-149.  The class list is handed to the device as
// given; the device sorts it by decreasing radius and applies the exclusion
// distance exactly as radius_generator.ixx:9-28 does (on construction of the
// algorithm).  The accessors below apply the same rules on the host.
template <int DIM, class RANDOM_RADIUS_GENERATOR = RandomRadiusGenerator>
class RadiusGenerator {
 public:
  RadiusGenerator(vector<tuple<double, uint64_t, int>> desired_radius_nb_phase, double exclusion_distance = 0)
      : classes_(std::move(desired_radius_nb_phase)), exclusion_(exclusion_distance) {
    if (classes_.empty()) throw std::runtime_error("Empty radius generator!");
    normalise();
  }
  // compute_desired_radius_nb_phase (radius_generator.ixx:71-86)
  RadiusGenerator(vector<tuple<double, double, int>> desired_radius_volumeFraction_phase, double volume,
                  double exclusion_distance = 0)
      : exclusion_(exclusion_distance) {
    if (desired_radius_volumeFraction_phase.empty()) throw std::runtime_error("Empty radius generator!");
    for (const auto& c : desired_radius_volumeFraction_phase) {
      const double r = get<0>(c);
      const double p = DIM == 2 ? r * r : (DIM == 3 ? r * (r * r) : (r * r) * (r * r));
      classes_.emplace_back(r, static_cast<uint64_t>(volume * get<1>(c) / (unit_volume() * p)), get<2>(c));
    }
    normalise();
  }
  //! custom law (radius_generator.ixx:38-48): one class (0, n, 0) whose radii are
  //! law->transform(u); the library evaluates the law on a host thread, once per draw
  RadiusGenerator(double min_radius, double max_radius, const RANDOM_RADIUS_GENERATOR* randomRadiusGenerator,
                  size_t desired_nb_spheres, double exclusion_distance = 0)
      : classes_{std::make_tuple(0., (uint64_t)desired_nb_spheres, 0)},
        sorted_(classes_),
        exclusion_(exclusion_distance),
        max_radius_(max_radius),
        min_radius_(min_radius),
        law_(randomRadiusGenerator) {}
  const RANDOM_RADIUS_GENERATOR* custom_law() const { return law_; }

  bool is_there_still_radii() const { return cur_ < classes_.size(); }
  uint64_t get_current_number() const { return get<1>(sorted_[cur_]); }
  double get_max_radius() const { return max_radius_; }
  double get_min_radius() const { return min_radius_; }
  //! radius_generator.ixx:62-68
  void update_placed(size_t nb_placed_spheres) {
    get<1>(sorted_[cur_]) -= nb_placed_spheres;
    if (get<1>(sorted_[cur_]) == 0) cur_++;
  }

  // for the forwarding layer: the classes as the caller gave them
  const vector<tuple<double, uint64_t, int>>& raw_classes() const { return classes_; }
  double exclusion_distance() const { return exclusion_; }

 private:
  static double unit_volume() { return Sphere<DIM>(create_array<DIM>(0.), 1., 0).volume(); }
  void normalise() {  // radius_generator.ixx:9-28 (stable for n <= 16, as libstdc++'s insertion sort)
    sorted_ = classes_;
    std::stable_sort(sorted_.begin(), sorted_.end(),
                     [](const auto& a, const auto& b) { return get<0>(a) > get<0>(b); });
    for (size_t i = 0; i < sorted_.size(); i++) get<0>(sorted_[0]) += 0.5 * exclusion_;
    max_radius_ = get<0>(sorted_[0]);
    min_radius_ = get<0>(sorted_.back());
    cur_ = 0;
  }
  vector<tuple<double, uint64_t, int>> classes_, sorted_;
  double exclusion_ = 0, max_radius_ = 0, min_radius_ = 0;
  const RANDOM_RADIUS_GENERATOR* law_ = nullptr;
  size_t cur_ = 0;
};

namespace auxi_function {
template <class T, class C>
T productOf(const C& c) {
  T p = 1;
  for (const auto& v : c) p *= v;
  return p;
}
}  // namespace auxi_function
}  // namespace sac_de_billes

// ------------------------------------------------------------ basic_types.hxx:18-36
using vec_double = std::vector<double>;
template <int DIM>
using vec_d = sac_de_billes::Point<DIM>;
template <int DIM>
using rsa_sphere = sac_de_billes::Sphere<DIM>;

// rsa_grid_traversal.hxx:8-10, rsa_data_storage.hxx:142-150
enum class TypeCell { Edge, Ghost, Real, All };
enum class TypeTag { Validated, Undecided, Conflict, Skip, Any };

// ------------------------------------------------------------ rsa_data_storage.hxx:29-139 (host copy)
template <int DIM>
class rsa_data_storage {
 public:
  size_t size() const { return m_rad.size(); }
  const vec_d<DIM>& get_center(int a_idx) const { return m_pos[a_idx]; }
  double get_center(int a_idx, int d) const { return m_pos[a_idx][d]; }
  double get_rad(const int a_idx) const { return m_rad[a_idx]; }
  uint64_t get_priority(const int a_idx) const { return m_priority[a_idx]; }
  uint64_t get_phase(const int a_idx) const { return m_phase[a_idx]; }
  rsa_sphere<DIM> create_sphere(const size_t i) const { return rsa_sphere<DIM>(m_pos[i], m_rad[i], (long)m_phase[i]); }
  void add_sphere(const uint64_t a_priority, const uint64_t a_id, const double a_rad, const vec_d<DIM>& a_center) {
    m_priority.push_back(a_priority);
    m_phase.push_back(a_id);
    m_rad.push_back(a_rad);
    m_pos.push_back(a_center);
  }
  void add_sphere(const uint64_t a_priority, const rsa_sphere<DIM>& s) {
    add_sphere(a_priority, (uint64_t)s.phase, s.radius, s.center);
  }
  void clear() {
    m_priority.clear();
    m_phase.clear();
    m_rad.clear();
    m_pos.clear();
  }

 private:
  std::vector<uint64_t> m_priority, m_phase;
  vec_double m_rad;
  std::vector<vec_d<DIM>> m_pos;
};

template <int DIM>
class rsa_domain;

// ------------------------------------------------------------ rsa_grid.hxx:25-100 (read API)
// A view of the domain's device-resident cells.
template <int DIM>
class rsa_grid {
 public:
  explicit rsa_grid(const rsa_domain<DIM>* d) : dom_(d) {}
  //! real spheres in extraction order (rsa_grid.ixx:73-98: cells by index, in-cell order)
  std::vector<rsa_sphere<DIM>> extract_spheres() const;
  //! extract_data<TypeCell::Real> (rsa_grid.ixx:139-160); priorities are not kept on the device (0)
  template <TypeCell typeCell = TypeCell::Real, TypeTag typeTag = TypeTag::Any>
  rsa_data_storage<DIM> extract_data() const {
    static_assert(typeCell == TypeCell::Real, "only real cells are kept on the device");
    rsa_data_storage<DIM> out;
    for (const auto& s : extract_spheres()) out.add_sphere(0, s);
    return out;
  }
  //! spheres stored in the real cells (rsa_grid.ixx:73-98 counts them per cell)
  template <TypeCell typeCell = TypeCell::Real>
  uint64_t get_number_of_spheres() const;
  //! rsa_grid.hxx:64 / rsa_grid.ixx:102-121: the counter add_to_sample keeps (rsa_domain::add leaves it)
  uint64_t get_number_of_spheres_fast() const;
  double local_volume_of_spheres() const;

 private:
  const rsa_domain<DIM>* dom_;
};

// ------------------------------------------------------------ rsa_domain.hxx:40-94
template <int DIM>
class rsa_domain {
 public:
  //! rsa_domain.hxx:53-54; ghost_layer must be >= 1 for the periodic domain
  rsa_domain(const vec_d<DIM>& global_inf, const vec_d<DIM>& global_sup, const int a_ghost_layer = 0,
             const double a_rad = 0.05)
      : m_inf(global_inf), m_sup(global_sup), m_ghost_layer(a_ghost_layer), m_rad(a_rad), m_grid(this) {
    rsa_gpu_ctx* c = nullptr;
    rsa_mpi_amd_detail::check(
        rsa_gpu_create(DIM, global_inf.data(), global_sup.data(), a_ghost_layer, a_rad, &c));
    ctx_.reset(c, [](rsa_gpu_ctx* p) { rsa_gpu_destroy(p); });
  }
  double get_m_rad() const { return m_rad; }
  rsa_grid<DIM>& get_grid() { return m_grid; }
  const rsa_grid<DIM>& get_grid() const { return m_grid; }
  const vec_d<DIM>& get_inf() const { return m_inf; }
  const vec_d<DIM>& get_sup() const { return m_sup; }
  int get_ghost_layer() const { return m_ghost_layer; }
  int get_m_id() const { return 0; }
  //! rsa_domain.hxx:84: insert validated spheres
  void add(rsa_data_storage<DIM>& a_data) {
    const size_t n = a_data.size();
    std::vector<double> pos(n * DIM), rad(n);
    std::vector<uint64_t> ph(n);
    for (size_t i = 0; i < n; i++) {
      for (int d = 0; d < DIM; d++) pos[i * DIM + d] = a_data.get_center((int)i, d);
      rad[i] = a_data.get_rad((int)i);
      ph[i] = a_data.get_phase((int)i);
    }
    rsa_mpi_amd_detail::check(rsa_gpu_add_spheres(ctx_.get(), (int64_t)n, pos.data(), rad.data(), ph.data()));
  }
  //! rsa_domain.ixx:150-163
  void domain_log() const {
    std::cout << " ==== log: domain [" << 0 << "] ==== " << std::endl;
    std::cout << " == m_inf : ";
    for (double v : m_inf) std::cout << v << " ";
    std::cout << std::endl << " == m_sup : ";
    for (double v : m_sup) std::cout << v << " ";
    std::cout << std::endl;
  }
  //! rsa_domain.ixx:166-169
  double get_total_volume() const {
    double v = 1.;
    for (int d = 0; d < DIM; d++) v *= (m_sup[d] - m_inf[d]);
    return v;
  }
  //! rsa_domain.ixx:170-175 (one rank: the local sum)
  double compute_total_volume_of_spheres() const {
    double v = 0;
    rsa_mpi_amd_detail::check(rsa_gpu_volume_of_spheres(ctx_.get(), &v));
    return v;
  }
  std::vector<rsa_sphere<DIM>> extract_spheres() const { return m_grid.extract_spheres(); }

  rsa_gpu_ctx* ctx() const { return ctx_.get(); }

 private:
  vec_d<DIM> m_inf, m_sup;
  int m_ghost_layer;
  double m_rad;
  rsa_grid<DIM> m_grid;
  std::shared_ptr<rsa_gpu_ctx> ctx_;
};

template <int DIM>
std::vector<rsa_sphere<DIM>> rsa_grid<DIM>::extract_spheres() const {
  uint64_t n = 0;
  rsa_mpi_amd_detail::check(rsa_gpu_extract(dom_->ctx(), nullptr, nullptr, nullptr, 0, &n));
  std::vector<double> pos(n * DIM), rad(n);
  std::vector<uint64_t> ph(n);
  rsa_mpi_amd_detail::check(rsa_gpu_extract(dom_->ctx(), pos.data(), rad.data(), ph.data(), n, &n));
  std::vector<rsa_sphere<DIM>> out;
  out.reserve(n);
  for (uint64_t i = 0; i < n; i++) {
    vec_d<DIM> c;
    for (int d = 0; d < DIM; d++) c[d] = pos[i * DIM + d];
    out.emplace_back(c, rad[i], (long)ph[i]);
  }
  return out;
}

template <int DIM>
template <TypeCell typeCell>
uint64_t rsa_grid<DIM>::get_number_of_spheres() const {
  static_assert(typeCell == TypeCell::Real, "only real cells are kept on the device");
  uint64_t n = 0;
  rsa_mpi_amd_detail::check(rsa_gpu_extract(dom_->ctx(), nullptr, nullptr, nullptr, 0, &n));
  return n;
}

template <int DIM>
uint64_t rsa_grid<DIM>::get_number_of_spheres_fast() const {
  uint64_t n = 0;
  rsa_mpi_amd_detail::check(rsa_gpu_count(dom_->ctx(), &n));
  return n;
}

template <int DIM>
double rsa_grid<DIM>::local_volume_of_spheres() const {
  double v = 0;
  rsa_mpi_amd_detail::check(rsa_gpu_volume_of_spheres(dom_->ctx(), &v));
  return v;
}

// ------------------------------------------------------------ operator_algorithm.hxx:18-170
namespace algorithm {
using namespace sac_de_billes;

template <int DIM>
class rsa_algo {
 public:
  //! operator_algorithm.hxx:27-28 (the domain is held by reference, the generator copied)
  rsa_algo(rsa_domain<DIM>& a_domain, RadiusGenerator<DIM>& a_radius_generator, int a_nbshots_singledraw,
           int a_n_draw = 10)
      : m_nbshots_singledraw(a_nbshots_singledraw), m_n_draw(a_n_draw), m_radius_generator(a_radius_generator),
        m_domain(a_domain) {
    install(true);
  }
  //! operator_algorithm.hxx:50-51: method 1 (voxels) only
  template <int method, bool plog = false>
  void proceed(size_t seed) {
    static_assert(method == 1, "the GPU path implements the voxel method (method 1)");
    rsa_mpi_amd_detail::check(rsa_gpu_proceed(m_domain.ctx(), (uint64_t)seed, -1, &last_stats_));
    if (plog)
      std::cout << "[log rsa_mpi] draws " << last_stats_.draws << " spheres " << last_stats_.accepted
                << " refinements " << last_stats_.refines << std::endl;
  }
  //! operator_algorithm.hxx:54 (operator_algorithm.ixx:238-250)
  void reset_radius_generator(RadiusGenerator<DIM>& a_radius_generator) {
    m_radius_generator = a_radius_generator;
    install(false);
  }
  const rsa_grid<DIM>& get_grid() const { return m_domain.get_grid(); }
  rsa_grid<DIM>& get_grid() { return m_domain.get_grid(); }
  //! operator_algorithm.hxx:66
  void set_miss_rate(double miss_rate_) { rsa_mpi_amd_detail::check(rsa_gpu_set_miss_rate(m_domain.ctx(), miss_rate_)); }
  //! statistics of the last proceed (no counterpart in the reference, which logs them)
  const rsa_gpu_stats& last_stats() const { return last_stats_; }

 private:
  // the C-ABI law: the generator's transform applied to each u
  static void law_trampoline(const double* u, double* radius, int64_t n, void* user) {
    const auto* g = static_cast<const RandomRadiusGenerator*>(user);
    for (int64_t i = 0; i < n; i++) radius[i] = g->transform(u[i]);
  }
  void install(bool first) {
    if (const RandomRadiusGenerator* law = m_radius_generator.custom_law()) {
      if (!first) {
        rsa_mpi_amd_detail::check(rsa_gpu_reset_law(
            m_domain.ctx(), m_radius_generator.get_min_radius(), m_radius_generator.get_max_radius(), &law_trampoline,
            const_cast<RandomRadiusGenerator*>(law), law->phase, m_radius_generator.get_current_number(),
            m_radius_generator.exclusion_distance()));
        return;
      }
      rsa_mpi_amd_detail::check(rsa_gpu_algo_init_law(
          m_domain.ctx(), m_radius_generator.get_min_radius(), m_radius_generator.get_max_radius(), &law_trampoline,
          const_cast<RandomRadiusGenerator*>(law), law->phase, m_radius_generator.get_current_number(),
          m_radius_generator.exclusion_distance(), m_nbshots_singledraw));
      return;
    }
    const auto& cls = m_radius_generator.raw_classes();
    std::vector<double> r;
    std::vector<uint64_t> n;
    std::vector<int> ph;
    for (const auto& c : cls) {
      r.push_back(std::get<0>(c));
      n.push_back(std::get<1>(c));
      ph.push_back(std::get<2>(c));
    }
    const double excl = m_radius_generator.exclusion_distance();
    if (first)
      rsa_mpi_amd_detail::check(rsa_gpu_algo_init(m_domain.ctx(), (int)r.size(), r.data(), n.data(), ph.data(),
                                                  excl, m_nbshots_singledraw));
    else
      rsa_mpi_amd_detail::check(
          rsa_gpu_reset_classes(m_domain.ctx(), (int)r.size(), r.data(), n.data(), ph.data(), excl));
  }
  int m_nbshots_singledraw;
  int m_n_draw;
  RadiusGenerator<DIM> m_radius_generator;
  rsa_domain<DIM>& m_domain;
  rsa_gpu_stats last_stats_{};
};

//! operator_algorithm.ixx:275-292: a single class of radius a_rad at volume fraction 1
template <int DIM, int method = 1, bool plog = false>
rsa_algo<DIM> uniform_generate(rsa_domain<DIM>& a_domain, double a_rad, int a_size, int a_n_draw = 1,
                               size_t seed = 0) {
  RadiusGenerator<DIM> gen(std::vector<std::tuple<double, double, int>>{std::make_tuple(a_rad, 1., 0)},
                           a_domain.get_total_volume(), 0.);
  rsa_algo<DIM> algo(a_domain, gen, a_size, a_n_draw);
  algo.template proceed<method, plog>(seed);
  return algo;
}
template <int DIM, int method = 1, bool plog = false>
rsa_algo<DIM> uniform_generate(rsa_domain<DIM>& a_domain, sac_de_billes::RadiusGenerator<DIM>& radius_generator,
                               int a_size, int a_n_draw, size_t seed = 0) {
  rsa_algo<DIM> algo(a_domain, radius_generator, a_size, a_n_draw);
  algo.template proceed<method, plog>(seed);
  return algo;
}
}  // namespace algorithm

// ------------------------------------------------------------ rsa_decoration.hxx:84-126 (one rank)
namespace rsa_mpi {
template <typename Arg>
void message(Arg a_arg) {
  std::cout << a_arg << std::endl;
}
template <typename Arg, typename... Args>
void message(Arg a_arg, Args... a_args) {
  std::cout << a_arg;
  message(a_args...);
}
template <typename... Args>
void init(Args...) {
  message("================= ===== ================= ");
  message("==== rsa_algo on one MI355X (rank 0)  == ");
  message("================= start ================= ");
}
inline void finalize() { message("=========== end ==========="); }
}  // namespace rsa_mpi

// ------------------------------------------------------------ lossless binary dump
//! The spheres of a storage at full FP64 precision (the XYZ text above keeps 6
//! digits).  No counterpart in the reference.  Little-endian layout: the 8
//! bytes "RSAMPIB1", uint32 DIM, uint32 0, uint64 count, then per sphere DIM
//! float64 centre coordinates, float64 radius, uint64 phase.
template <int DIM>
void write_binary(std::string a_name, const rsa_data_storage<DIM>& spheres) {
  std::ofstream out(a_name, std::ios::binary);
  if (!out) throw std::runtime_error("write_binary: cannot open " + a_name);
  const char magic[8] = {'R', 'S', 'A', 'M', 'P', 'I', 'B', '1'};
  const uint32_t head[2] = {(uint32_t)DIM, 0u};
  const uint64_t n = spheres.size();
  out.write(magic, 8);
  out.write(reinterpret_cast<const char*>(head), sizeof head);
  out.write(reinterpret_cast<const char*>(&n), sizeof n);
  for (size_t i = 0; i < spheres.size(); i++) {
    double rec[DIM + 1];
    for (int d = 0; d < DIM; d++) rec[d] = spheres.get_center((int)i)[d];
    rec[DIM] = spheres.get_rad((int)i);
    const uint64_t ph = spheres.get_phase((int)i);
    out.write(reinterpret_cast<const char*>(rec), sizeof rec);
    out.write(reinterpret_cast<const char*>(&ph), sizeof ph);
  }
  if (!out) throw std::runtime_error("write_binary: write failed for " + a_name);
}

//! The storage write_binary wrote (priorities are not kept: 0).
template <int DIM>
rsa_data_storage<DIM> read_binary(std::string a_name) {
  std::ifstream in(a_name, std::ios::binary);
  if (!in) throw std::runtime_error("read_binary: cannot open " + a_name);
  char magic[8];
  uint32_t head[2];
  uint64_t n = 0;
  in.read(magic, 8);
  in.read(reinterpret_cast<char*>(head), sizeof head);
  in.read(reinterpret_cast<char*>(&n), sizeof n);
  if (!in || std::string(magic, 8) != "RSAMPIB1" || head[0] != (uint32_t)DIM)
    throw std::runtime_error("read_binary: not a " + std::to_string(DIM) + "D sphere dump: " + a_name);
  rsa_data_storage<DIM> s;
  for (uint64_t i = 0; i < n; i++) {
    double rec[DIM + 1];
    uint64_t ph = 0;
    in.read(reinterpret_cast<char*>(rec), sizeof rec);
    in.read(reinterpret_cast<char*>(&ph), sizeof ph);
    if (!in) throw std::runtime_error("read_binary: truncated file " + a_name);
    vec_d<DIM> c;
    for (int d = 0; d < DIM; d++) c[d] = rec[d];
    s.add_sphere(0, ph, rec[DIM], c);
  }
  return s;
}

// ------------------------------------------------------------ operator_paraview.hxx:5-37
// VTK XML PolyData of the real spheres (operator_paraview.ixx:12-142): per call
// `step`, a_name/rsa_mpi_<step>.pvtp naming one piece per rank (one here) and
// a_name/rsa_mpi_<step>/rsa_mpi_<step>_0.vtp with point data Id (phase), Tag
// (TypeTag::Validated) and Radius, and the centres as points, as ASCII with the
// stream's default precision.  The reference reads three coordinates whatever
// DIM is; here D = 2 writes z = 0.
namespace rsa_paraview {
inline int& output_step() {
  static int step = 0;
  return step;
}

// a file for writing, or a message on std::cerr naming it
inline bool open_for_writing(std::ofstream& out, const std::string& path) {
  out.open(path);
  if (!out) std::cerr << "rsa_paraview: cannot write " << path << std::endl;
  return bool(out);
}

inline void write_pvtp(const std::string& basedir, const std::string& basename, size_t number_of_files) {
  std::ofstream out;
  if (!open_for_writing(out, basedir + "/" + basename + ".pvtp")) return;
  // the index file's fixed part, byte for byte as the reference writes it
  static const char* const kHead =
      " <VTKFile type=\"PPolyData\"> \n"
      "   <PPolyData GhostLevel=\"0\">\n"
      "    <PPointData Scalar=\"id, tag\">\n"
      "      <PDataArray type=\"Int64\" Name=\"Id\"/>\n"
      "      <PDataArray type=\"Int32\" Name=\"Tag\"/>\n"
      "      <PDataArray type=\"Float64\" Name=\"Radius\"/>\n"
      "    </PPointData>\n"
      "     <PPoints>\n"
      "       <PDataArray type=\"Float64\" NumberOfComponents=\"3\"/>\n"
      "     </PPoints> \n";
  out << kHead;
  for (size_t piece = 0; piece < number_of_files; piece++)
    out << "     <Piece Source=\"" << basename << "/" << basename << "_" << piece << ".vtp\"/>\n";
  out << "   </PPolyData>\n </VTKFile>" << std::endl;
}

template <int DIM>
void write_vtp(const std::string& name, const std::vector<rsa_sphere<DIM>>& spheres) {
  std::ofstream out;
  if (!open_for_writing(out, name + ".vtp")) return;
  // one ASCII value list per point attribute, and the centres (z = 0 in 2D)
  std::ostringstream centres, phases, tags, radii;
  for (const rsa_sphere<DIM>& sph : spheres) {
    for (int d = 0; d < 3; d++) centres << (d < DIM ? sph.center[d] : 0.) << ' ';
    phases << sph.phase << ' ';
    tags << int(TypeTag::Validated) << ' ';
    radii << sph.radius << ' ';
  }
  auto data_array = [&out](const char* type, const char* label, const std::ostringstream& values) {
    out << "     <DataArray type=\"" << type << "\" Name=\"" << label << "\" format=\"ascii\">\n"
        << values.str() << "\n     </DataArray>\n";
  };
  out << "<VTKFile type=\"PolyData\">\n <PolyData>\n"
      << "   <Piece NumberOfPoints=\"" << spheres.size() << "\" NumberOfPolys=\"0\">\n   <PointData>\n";
  data_array("Int64", "Id", phases);
  data_array("Int32", "Tag", tags);
  data_array("Float64", "Radius", radii);
  out << "   </PointData>\n   <Points>\n"
      << "     <DataArray type=\"Float64\" Name=\"\"  NumberOfComponents=\"3\" format=\"ascii\">\n"
      << centres.str() << "\n     </DataArray>\n   </Points>\n  </Piece>\n </PolyData>\n</VTKFile>" << std::endl;
}

//! operator_paraview.ixx:98-136 on one rank
template <int DIM>
void paraview(std::string a_name, const rsa_grid<DIM>& a_grid) {
  const int step = output_step();
  const std::string basename = "rsa_mpi";
  const std::string stepname = basename + "_" + std::to_string(step);
  const std::string directory = a_name + "/" + stepname;
  std::filesystem::create_directory(a_name);
  std::filesystem::create_directory(directory);
  write_pvtp(a_name, stepname, 1);
  write_vtp<DIM>(directory + "/" + stepname + "_0", a_grid.extract_spheres());
  output_step()++;
}

//! operator_paraview.ixx:138-142
template <int DIM>
void paraview(const rsa_domain<DIM>& a_domain, std::string a_name = "ParaviewOutput") {
  paraview<DIM>(a_name, a_domain.get_grid());
}
}  // namespace rsa_paraview

// ------------------------------------------------------------ test.hxx:13-32
//! The reference passes volume 1 (not prod(L)) to the radius generator; kept.
template <int DIM, int method>
std::vector<rsa_sphere<DIM>> throw_spheres(sac_de_billes::Point<DIM> L,
                                           std::vector<std::tuple<double, double, int>> desired_radius_volumeFraction_phase,
                                           size_t seed, bool write_paraview = true) {
  vec_d<DIM> inf = sac_de_billes::create_array<DIM>(0.);
  vec_d<DIM> sup;
  for (int d = 0; d < DIM; d++) sup[d] = inf[d] + L[d];
  sac_de_billes::RadiusGenerator<DIM> radius_generator(desired_radius_volumeFraction_phase, 1);
  rsa_domain<DIM> domain(inf, sup, 1, radius_generator.get_max_radius());
  domain.domain_log();
  algorithm::uniform_generate<DIM, method>(domain, radius_generator, 6000, 10, seed);
  if (write_paraview) {
    rsa_mpi::message("====> write mpi output file <=====");
    rsa_paraview::paraview(domain);
  }
  return domain.extract_spheres();
}

// ------------------------------------------------------------ user_interface.hxx:11-124
namespace user_interface {
template <int DIM>
class rsa_domain {
 public:
  rsa_domain(const vec_d<DIM>& global_inf, const vec_d<DIM>& global_sup, const double a_rad)
      : impl_{new ::rsa_domain<DIM>(global_inf, global_sup, 1, a_rad)} {}
  const vec_d<DIM>& get_inf() const { return impl_->get_inf(); }
  const vec_d<DIM>& get_sup() const { return impl_->get_sup(); }
  double get_total_volume() const { return impl_->get_total_volume(); }
  double compute_total_volume_of_spheres() const { return impl_->compute_total_volume_of_spheres(); }
  void domain_log() const { impl_->domain_log(); }
  std::vector<rsa_sphere<DIM>> extract_spheres() const { return impl_->extract_spheres(); }
  //! user_interface.hxx:40
  void paraview() const { rsa_paraview::paraview(*impl_); }
  const ::rsa_domain<DIM>& get_pointed_to() const { return *impl_; }
  ::rsa_domain<DIM>& get_pointed_to() { return *impl_; }

 private:
  std::shared_ptr<::rsa_domain<DIM>> impl_;
};

template <int DIM>
class RadiusGenerator {
 public:
  // classes as (radius, count, phase)
  RadiusGenerator(std::vector<std::tuple<double, uint64_t, int>> classes, double exclusion)
      : impl_{std::make_shared<sac_de_billes::RadiusGenerator<DIM>>(classes, exclusion)} {}
  // classes as (radius, volume fraction, phase) of a domain of the given volume
  RadiusGenerator(std::vector<std::tuple<double, double, int>> classes, double volume, double exclusion)
      : impl_{std::make_shared<sac_de_billes::RadiusGenerator<DIM>>(classes, volume, exclusion)} {}
  sac_de_billes::RadiusGenerator<DIM>& get_pointed_to() { return *impl_; }

 private:
  std::shared_ptr<sac_de_billes::RadiusGenerator<DIM>> impl_;
};

template <int DIM>
class rsa_algo {
 public:
  rsa_algo(rsa_domain<DIM>& a_domain, user_interface::RadiusGenerator<DIM>& a_radius_generator,
           int a_nbshots_singledraw)
      : impl_{new algorithm::rsa_algo<DIM>(a_domain.get_pointed_to(), a_radius_generator.get_pointed_to(),
                                                   a_nbshots_singledraw)} {}
  void proceed(size_t seed) { impl_->template proceed<1>(seed); }
  void reset_radius_generator(RadiusGenerator<DIM>& a_radius_generator) {
    impl_->reset_radius_generator(a_radius_generator.get_pointed_to());
  }

 private:
  std::shared_ptr<algorithm::rsa_algo<DIM>> impl_;
};

template <int DIM>
std::vector<rsa_sphere<DIM>> throw_spheres(sac_de_billes::Point<DIM> L,
                                           std::vector<std::tuple<double, double, int>> desired_radius_volumeFraction_phase,
                                           size_t seed, bool write_paraview) {
  return ::throw_spheres<DIM, 1>(L, desired_radius_volumeFraction_phase, seed, write_paraview);
}
}  // namespace user_interface

// ------------------------------------------------------------ plugin/write_xyz.hpp:6-98
//! one rank: the gathered storage is the local one
template <int DIM>
rsa_data_storage<DIM> gather_data(rsa_data_storage<DIM>& a_data) {
  return a_data;
}

//! XYZ text: count, the box's sup[0..2], then "0 x y z" per sphere with the
//! stream's default 6 significant digits (plugin/write_xyz.hpp:78-98).  The
//! reference reads three coordinates whatever DIM is; here D = 2 writes z = 0.
template <int DIM>
void write_xyz(std::string a_name, rsa_data_storage<DIM>& spheres, std::array<double, DIM>& a_sup) {
  const rsa_data_storage<DIM> all = gather_data(spheres);
  std::ofstream out(a_name);
  if (!out) {
    std::cerr << "write_xyz: cannot write " << a_name << std::endl;
    return;
  }
  auto xyz = [&out](const auto& v) {  // three coordinates, z = 0 in 2D
    for (int d = 0; d < 3; d++) out << (d ? " " : "") << (d < DIM ? v[d] : 0.);
    out << std::endl;
  };
  out << all.size() << std::endl;
  xyz(a_sup);
  for (size_t k = 0; k < all.size(); k++) {
    out << 0 << " ";  // the sphere type the format carries (always 0)
    xyz(all.get_center((int)k));
  }
}
