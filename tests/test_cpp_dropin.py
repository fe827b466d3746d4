"""The C++ drop-in (include/rsa_mpi_amd/*.hxx) used the way the reference's own
programs use its API (tests/cpp/dropin_example.cpp), checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import _oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "cpp", "dropin_example.cpp")
BIN = os.path.join(HERE, "cpp", "dropin_example")


def build_cmd(out):
    return ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include", "rsa_mpi_amd"), SRC,
            "-L", os.path.join(ROOT, "rsa_mpi_amd", "lib"), "-lrsa_mpi_amd",
            "-Wl,-rpath,$ORIGIN/../../rsa_mpi_amd/lib", "-o", out]


def test_dropin_headers_compile(tmp_path):
    """The reference-style caller compiles and links with g++ against the headers and the library."""
    r = subprocess.run(build_cmd(str(tmp_path / "x")), capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "warning" not in r.stderr


def run_bin(tmp_path, *args):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built (run __graft_entry__.build())")
    out = str(tmp_path / "out.bin")
    r = subprocess.run([BIN, args[0], out] + [str(a) for a in args[1:]], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return out, r.stdout


def read_records(path, dim):
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(8), dtype=np.uint64)[0])
        rec = np.frombuffer(f.read(), dtype=np.float64).reshape(n, dim + 2)
    return rec[:, :dim], rec[:, dim], rec[:, dim + 1].astype(np.int64)


def same_as_oracle(pos, rad, ph, o):
    assert len(rad) == o["n"]
    assert np.array_equal(pos.view(np.uint64), o["pos"].view(np.uint64))
    assert np.array_equal(rad.view(np.uint64), o["rad"].view(np.uint64))
    assert np.array_equal(ph, o["phase"].astype(np.int64))


@pytest.mark.gpu
def test_uniform_generate_3d(tmp_path):
    out, log = run_bin(tmp_path, "uniform3d", 0.2, 0.005, 3)
    pos, rad, ph = read_records(out, 3)
    same_as_oracle(pos, rad, ph, O.run(3, [0.0] * 3, [0.2] * 3, [(0.005, 1.0, 0)], seed=3))
    assert f"spheres {len(rad)} " in log


@pytest.mark.gpu
def test_user_interface_throw_spheres_2d(tmp_path):
    out, _ = run_bin(tmp_path, "throw2d", 1.0, 2)
    pos, rad, ph = read_records(out, 2)
    same_as_oracle(pos, rad, ph, O.run(2, [0.0] * 2, [1.0] * 2, [(0.02, 0.2, 0), (0.01, 1.0, 1)], seed=2))


@pytest.mark.gpu
def test_user_interface_reiteration_3d(tmp_path):
    out, _ = run_bin(tmp_path, "reiter3d", 0.2, 4)
    pos, rad, ph = read_records(out, 3)
    same_as_oracle(pos, rad, ph, O.run(3, [0.0] * 3, [0.2] * 3, [(0.01, 1.0, 0)], seed=4, reiter=(0.005,)))


@pytest.mark.gpu
def test_reiteration_onto_a_radius_law_3d(tmp_path):
    """rsa_algo::reset_radius_generator to a RadiusGenerator with a RandomRadiusGenerator
    (operator_algorithm.ixx:238-250, radius_generator.ixx:38-48), then back to a class, each
    followed by proceed(seed): the packing is the oracle's bit for bit."""
    out, _ = run_bin(tmp_path, "reiterlaw3d", 0.2, 3)
    pos, rad, ph = read_records(out, 3)
    o = O.run(3, [0.0] * 3, [0.2] * 3, [(0.01, 0.2, 0)], seed=3, reiter=(("law", 0.006, 0.008, 2, 400), 0.005))
    same_as_oracle(pos, rad, ph, o)
    assert int((ph == 2).sum()) == 400


@pytest.mark.gpu
def test_write_xyz_plugin(tmp_path):
    """plugin/write_xyz.hpp:78-98: count, sup, then '0 x y z' in default stream format (6 digits)."""
    sup = (0.2, 0.3, 0.25)
    out, _ = run_bin(tmp_path, "xyz3d", 0.01, *sup)
    o = O.run(3, [0.0] * 3, list(sup), [(0.01, 1.0, 0)], seed=0)
    lines = open(out).read().splitlines()
    assert lines[0] == str(o["n"])
    assert lines[1] == " ".join("%g" % v for v in sup)
    want = ["0 " + " ".join("%g" % v for v in p) for p in o["pos"]]
    assert lines[2:] == want


@pytest.mark.gpu
def test_random_radius_generator_law_3d(tmp_path):
    """test/test_random_radius_generator.cpp: RadiusGenerator(r_min, r_max, &RandomRadiusGenerator(uniform
    transform), 1e16) in the unit cube, seed 0; radii r_min + (r_max - r_min) u from 2 words per shot."""
    out, _ = run_bin(tmp_path, "law3d", 1.0, 0.05, 0.055, 10 ** 16, 0)
    pos, rad, ph = read_records(out, 3)
    o = O.run(3, [0.0] * 3, [1.0] * 3, [(0.05, 1e16, 0), (0.055, 0.0, 0)], seed=0, mode=2)
    same_as_oracle(pos, rad, ph, o)
    assert len(np.unique(rad)) > len(rad) // 2  # the radii are drawn, not a class value


def parse_vtp(path):
    """The point data arrays and points of an ASCII VTK PolyData file."""
    import xml.etree.ElementTree as ET
    piece = ET.parse(path).getroot().find("PolyData").find("Piece")
    arrays = {a.get("Name"): a.text.split() for a in piece.find("PointData").findall("DataArray")}
    pts = piece.find("Points").find("DataArray").text.split()
    return int(piece.get("NumberOfPoints")), arrays, pts


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_paraview_output(tmp_path, dim):
    """rsa_paraview::paraview (operator_paraview.ixx:98-142): <name>/rsa_mpi_<step>.pvtp naming the piece
    rsa_mpi_<step>/rsa_mpi_<step>_0.vtp, whose points, Id (phase), Tag (Validated = 0) and Radius are the
    extracted spheres in default stream format; each call writes the next step."""
    out = str(tmp_path / "pv")
    r = subprocess.run([BIN, "pv", out, str(dim), "0.2", "0.01", "1"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    pos, rad, ph = read_records(os.path.join(out, "spheres.bin"), dim)
    pvtp = open(os.path.join(out, "rsa_mpi_0.pvtp")).read()
    assert '<Piece Source="rsa_mpi_0/rsa_mpi_0_0.vtp"/>' in pvtp
    n, arrays, pts = parse_vtp(os.path.join(out, "rsa_mpi_0", "rsa_mpi_0_0.vtp"))
    assert n == len(rad)
    want_pts = []
    for p in pos:
        want_pts += ["%g" % v for v in p] + (["0"] if dim == 2 else [])
    assert pts == want_pts
    assert arrays["Radius"] == ["%g" % v for v in rad]
    assert arrays["Id"] == [str(v) for v in ph]
    assert arrays["Tag"] == ["0"] * n
    if dim == 2:  # the second call in the same process
        assert os.path.exists(os.path.join(out, "rsa_mpi_1", "rsa_mpi_1_0.vtp"))


@pytest.mark.gpu
def test_write_xyz_plugin_reference_args(tmp_path):
    """test/CMakeLists.txt:66-67 runs test_plugin_write_xyz with 0.5 10 10 10."""
    out, _ = run_bin(tmp_path, "xyz3d", 0.5, 10, 10, 10)
    o = O.run(3, [0.0] * 3, [10.0] * 3, [(0.5, 1.0, 0)], seed=0)
    lines = open(out).read().splitlines()
    assert lines[0] == str(o["n"])
    assert lines[1] == "10 10 10"
    assert lines[2:] == ["0 " + " ".join("%g" % v for v in p) for p in o["pos"]]


def test_writers_byte_for_byte(tmp_path):
    """The Paraview piece/index writers (operator_paraview.ixx:12-77) and write_xyz
    (plugin/write_xyz.hpp:78-98) on fixed spheres, host only, against the reference's
    output restated line by line (C++ default stream format: %g with 6 digits)."""
    exe = str(tmp_path / "format_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include", "rsa_mpi_amd"),
                    os.path.join(ROOT, "tests", "cpp", "format_check.cpp"), "-o", exe], check=True)
    subprocess.run([exe, str(tmp_path)], check=True)
    g = lambda v: "%g" % v

    def vtp(spheres, dim):
        pts = "".join(" ".join(g(c[d] if d < dim else 0.0) for d in range(3)) + " " for c, _, _ in spheres)
        lines = ['<VTKFile type="PolyData">', " <PolyData>",
                 '   <Piece NumberOfPoints="%d" NumberOfPolys="0">' % len(spheres), "   <PointData>"]
        for typ, name, vals in (("Int64", "Id", [str(p) for _, _, p in spheres]), ("Int32", "Tag", ["0"] * len(spheres)),
                                ("Float64", "Radius", [g(r) for _, r, _ in spheres])):
            lines += ['     <DataArray type="%s" Name="%s" format="ascii">' % (typ, name),
                      "".join(v + " " for v in vals), "     </DataArray>"]
        lines += ["   </PointData>", "   <Points>",
                  '     <DataArray type="Float64" Name=""  NumberOfComponents="3" format="ascii">', pts,
                  "     </DataArray>", "   </Points>", "  </Piece>", " </PolyData>", "</VTKFile>"]
        return "\n".join(lines) + "\n"

    s3 = [((0.125, 0.5, 0.75), 0.0625, 2), ((1.0 / 3.0, 2.5e-7, 12345.678901), 0.01, 0)]
    assert open(tmp_path / "p3.vtp").read() == vtp(s3, 3)
    assert open(tmp_path / "p2.vtp").read() == vtp([((0.25, 0.375), 0.125, 1)], 2)
    idx = [' <VTKFile type="PPolyData"> ', '   <PPolyData GhostLevel="0">', '    <PPointData Scalar="id, tag">',
           '      <PDataArray type="Int64" Name="Id"/>', '      <PDataArray type="Int32" Name="Tag"/>',
           '      <PDataArray type="Float64" Name="Radius"/>', "    </PPointData>", "     <PPoints>",
           '       <PDataArray type="Float64" NumberOfComponents="3"/>', "     </PPoints> "]
    idx += ['     <Piece Source="idx/idx_%d.vtp"/>' % k for k in range(3)] + ["   </PPolyData>", " </VTKFile>"]
    assert open(tmp_path / "idx.pvtp").read() == "\n".join(idx) + "\n"
    xyz = ["2", "1 2 3.5"] + ["0 " + " ".join(g(v) for v in c) for c, _, _ in s3]
    assert open(tmp_path / "out.xyz").read() == "\n".join(xyz) + "\n"
