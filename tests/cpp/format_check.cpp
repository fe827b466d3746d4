// Writes the Paraview piece and index files and an XYZ file of fixed spheres
// with the drop-in header's writers (no device involved), for a byte-for-byte
// check of the formats.  argv[1]: output directory.
#include <rsa_mpi_amd.hpp>

#include <string>
#include <vector>

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  std::vector<rsa_sphere<3>> s3;
  s3.emplace_back(sac_de_billes::Point<3>{0.125, 0.5, 0.75}, 0.0625, 2);
  s3.emplace_back(sac_de_billes::Point<3>{1.0 / 3.0, 2.5e-7, 12345.678901}, 0.01, 0);
  rsa_paraview::write_vtp<3>(dir + "/p3", s3);
  std::vector<rsa_sphere<2>> s2;
  s2.emplace_back(sac_de_billes::Point<2>{0.25, 0.375}, 0.125, 1);
  rsa_paraview::write_vtp<2>(dir + "/p2", s2);
  rsa_paraview::write_pvtp(dir, "idx", 3);
  rsa_data_storage<3> st;
  for (const auto& s : s3) st.add_sphere(0, s);
  std::array<double, 3> sup = {1.0, 2.0, 3.5};
  write_xyz<3>(dir + "/out.xyz", st, sup);
  return 0;
}
