// write_binary / read_binary of the drop-in on host-built storages (no GPU):
// "w <path>" writes 3 spheres with awkward FP64 values; "r <path>" reads a
// 2D dump and prints its records with 17 significant digits.
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>

#include <rsa_data_storage.hxx>

int main(int argc, char** argv) {
  if (argc < 3) return 1;
  const std::string mode = argv[1], path = argv[2];
  if (mode == "w") {
    rsa_data_storage<3> s;
    s.add_sphere(7, 0, 0.1, {0.1, 0.2, 0.30000000000000004});
    s.add_sphere(8, 3, std::numeric_limits<double>::denorm_min(), {1e-300, 0.9999999999999999, 0.5});
    s.add_sphere(9, 18446744073709551615ull, 0.0025, {0.125, 0.25, 0.375});
    write_binary<3>(path, s);
  } else {
    auto s = read_binary<2>(path);
    std::printf("%zu\n", s.size());
    for (size_t i = 0; i < s.size(); i++)
      std::printf("%.17g %.17g %.17g %llu\n", s.get_center((int)i)[0], s.get_center((int)i)[1], s.get_rad((int)i),
                  (unsigned long long)s.get_phase((int)i));
  }
  return 0;
}
