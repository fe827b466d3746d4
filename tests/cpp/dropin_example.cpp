// A caller of the reference's C++ API, unchanged but for the build line: the
// calls mirror test/test_plugin_write_xyz.cpp:12-24, user_interface.hxx,
// scripts/main_max_frac.cpp:39-52 and test/test_random_radius_generator.cpp:5-32.  It writes the spheres it gets as binary
// records (x.., r, phase) so that the tests can compare them bit for bit.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include <operator_algorithm.hxx>
#include <operator_paraview.hxx>
#include <radius_generator.hxx>
#include <rsa_decoration.hxx>
#include <rsa_domain.hxx>
#include <test.hxx>
#include <user_interface.hxx>
#include <write_xyz.hpp>

template <int DIM>
static void dump(const std::string& path, const std::vector<rsa_sphere<DIM>>& s) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) std::exit(2);
  const uint64_t n = s.size();
  std::fwrite(&n, sizeof n, 1, f);
  for (const auto& x : s) {
    double rec[DIM + 2];
    for (int d = 0; d < DIM; d++) rec[d] = x.center[d];
    rec[DIM] = x.radius;
    rec[DIM + 1] = (double)x.phase;
    std::fwrite(rec, sizeof rec, 1, f);
  }
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s uniform3d|xyz3d|throw2d|reiter3d|reiterlaw3d|law3d|pv out [args]\n", argv[0]);
    return 1;
  }
  rsa_mpi::init(&argc, &argv);
  const std::string mode = argv[1], out = argv[2];
  if (mode == "uniform3d") {  // algorithm::uniform_generate<3, 1>(domain, rad, 6000, 10, seed)
    const double L = std::atof(argv[3]), rad = std::atof(argv[4]);
    const size_t seed = std::strtoull(argv[5], nullptr, 10);
    constexpr int DIM = 3;
    std::array<double, DIM> inf = {0., 0., 0.}, sup = {L, L, L};
    rsa_domain<DIM> domain(inf, sup, 1, rad);
    auto algo = algorithm::uniform_generate<DIM, 1>(domain, rad, 6000, 10, seed);
    dump<DIM>(out, domain.extract_spheres());
    std::printf("spheres %llu volume %.17g\n", (unsigned long long)domain.get_grid().get_number_of_spheres(),
                domain.compute_total_volume_of_spheres());
  } else if (mode == "xyz3d") {  // test/test_plugin_write_xyz.cpp:12-24
    const double rad = std::atof(argv[3]);
    constexpr int DIM = 3;
    std::array<double, DIM> inf = {0., 0., 0.};
    std::array<double, DIM> sup = {std::atof(argv[4]), std::atof(argv[5]), std::atof(argv[6])};
    rsa_domain<DIM> domain(inf, sup, 1, rad);
    auto algo = algorithm::uniform_generate<DIM, 1>(domain, rad, 6000, 10, 0);
    auto spheres = domain.get_grid().template extract_data<TypeCell::Real>();
    write_xyz<DIM>(out, spheres, sup);
  } else if (mode == "throw2d") {  // user_interface::throw_spheres<2>(L, classes, seed, false)
    const double L = std::atof(argv[3]);
    const size_t seed = std::strtoull(argv[4], nullptr, 10);
    std::vector<std::tuple<double, double, int>> cls = {{0.02, 0.2, 0}, {0.01, 1.0, 1}};
    auto s = user_interface::throw_spheres<2>({L, L}, cls, seed, false);
    dump<2>(out, s);
  } else if (mode == "reiter3d") {  // scripts/main_max_frac.cpp:39-52 through user_interface
    const double L = std::atof(argv[3]);
    const size_t seed = std::strtoull(argv[4], nullptr, 10);
    constexpr int DIM = 3;
    user_interface::rsa_domain<DIM> domain({0., 0., 0.}, {L, L, L}, 0.01);
    user_interface::RadiusGenerator<DIM> g1({{0.01, 1.0, 0}}, domain.get_total_volume(), 0.);
    user_interface::rsa_algo<DIM> algo(domain, g1, 6000);
    algo.proceed(seed);
    user_interface::RadiusGenerator<DIM> g2({{0.005, 1.0, 0}}, domain.get_total_volume(), 0.);
    algo.reset_radius_generator(g2);
    algo.proceed(seed);
    dump<DIM>(out, domain.extract_spheres());
  } else if (mode == "law3d") {  // test/test_random_radius_generator.cpp:8-30
    constexpr int DIM = 3;
    const double L = std::atof(argv[3]), r_min = std::atof(argv[4]), r_max = std::atof(argv[5]);
    const uint64_t nb_spheres_max = std::strtoull(argv[6], nullptr, 10);
    const size_t seed = std::strtoull(argv[7], nullptr, 10);
    std::array<double, DIM> domain_inf = create_array<DIM>(0.);
    std::array<double, DIM> domain_sup = create_array<DIM>(L);
    uint64_t phase = 0;
    auto nonlinear_transform = [r_min, r_max](double x) { return r_min + (r_max - r_min) * x; };
    sac_de_billes::RandomRadiusGenerator random_radius_generator(nonlinear_transform, phase);
    sac_de_billes::RadiusGenerator<DIM> radius_generator(r_min, r_max, &random_radius_generator, nb_spheres_max, 0.);
    rsa_domain<DIM> domain(domain_inf, domain_sup, 1, radius_generator.get_max_radius());
    algorithm::uniform_generate<DIM, 1, false>(domain, radius_generator, 6000, 10, seed);
    dump<DIM>(out, domain.extract_spheres());
  } else if (mode == "reiterlaw3d") {  // scripts/main_max_frac.cpp:39-52 with a law generator in the loop
    constexpr int DIM = 3;
    const double L = std::atof(argv[3]);
    const size_t seed = std::strtoull(argv[4], nullptr, 10);
    rsa_domain<DIM> domain(create_array<DIM>(0.), create_array<DIM>(L), 1, 0.01);
    sac_de_billes::RadiusGenerator<DIM> g1(std::vector<std::tuple<double, double, int>>{{0.01, 0.2, 0}},
                                           domain.get_total_volume());
    algorithm::rsa_algo<DIM> algo(domain, g1, 6000, 10);
    algo.proceed<1>(seed);
    const double r_min = 0.006, r_max = 0.008;
    auto nonlinear_transform = [r_min, r_max](double x) { return r_min + (r_max - r_min) * x; };
    sac_de_billes::RandomRadiusGenerator law(nonlinear_transform, 2);
    sac_de_billes::RadiusGenerator<DIM> g2(r_min, r_max, &law, 400, 0.);
    algo.reset_radius_generator(g2);
    algo.proceed<1>(seed);
    sac_de_billes::RadiusGenerator<DIM> g3(std::vector<std::tuple<double, double, int>>{{0.005, 1.0, 0}},
                                           domain.get_total_volume());
    algo.reset_radius_generator(g3);
    algo.proceed<1>(seed);
    dump<DIM>(out, domain.extract_spheres());
  } else if (mode == "pv") {  // uniform_generate then rsa_paraview::paraview(domain, out) (test.hxx:45-50)
    const int dim = std::atoi(argv[3]);
    const double L = std::atof(argv[4]), rad = std::atof(argv[5]);
    const size_t seed = std::strtoull(argv[6], nullptr, 10);
    if (dim == 2) {
      rsa_domain<2> domain({0., 0.}, {L, L}, 1, rad);
      auto algo = algorithm::uniform_generate<2, 1>(domain, rad, 6000, 10, seed);
      rsa_paraview::paraview(domain, out);
      rsa_paraview::paraview(domain, out);  // a second call writes the next step
      dump<2>(out + "/spheres.bin", domain.extract_spheres());
    } else {
      rsa_domain<3> domain({0., 0., 0.}, {L, L, L}, 1, rad);
      auto algo = algorithm::uniform_generate<3, 1>(domain, rad, 6000, 10, seed);
      rsa_paraview::paraview(domain, out);
      dump<3>(out + "/spheres.bin", domain.extract_spheres());
    }
  } else {
    return 1;
  }
  rsa_mpi::finalize();
  return 0;
}
