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
