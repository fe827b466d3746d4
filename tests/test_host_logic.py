"""CPU checks of the library boundary and of the identities the HIP kernels rely
on where they evaluate the reference's formulas in a different, exactly
equivalent order."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rsa_mpi_amd.h")


# ---------------------------------------------------------------- the C-ABI boundary

def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(rsa_gpu_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    """librsa_mpi_amd.so loads on a host without a GPU and exports each function
    include/rsa_mpi_amd.h declares (no compute call is made)."""
    import rsa_mpi_amd._lib as L
    if not os.path.exists(L.LIB_PATH):
        pytest.fail(f"{L.LIB_PATH} not built: run rsa_mpi_amd/build.sh")
    lib = ctypes.CDLL(L.LIB_PATH)
    names = header_functions()
    assert len(names) >= 17
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(L.EXPORTED) == names


def test_binding_fails_loudly_without_library(monkeypatch, tmp_path):
    import rsa_mpi_amd._lib as L
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(L, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        L.lib()


def test_host_only_entry_points_work_without_gpu():
    """rsa_gpu_mt_words runs on the host: it reproduces std::mt19937 seeded
    with seed_seq{rank, seed}.  Errors come back as codes with a message."""
    import rsa_mpi_amd as R
    import _oracle as O
    w = R.mt_words(123, 5000)
    assert np.array_equal(w, O.mt_words(123, 5000))
    lib = R.lib()
    assert lib.rsa_gpu_proceed(None, 0, -1, None) != 0
    assert b"null context" in lib.rsa_gpu_last_error()


# ---------------------------------------------------------------- farthest-corner test

def corner_loop(delta, length, tab, R2):
    """auxi::is_in_sphere (list_of_voxels.ixx:225-238): every corner
    tab[i] * length within R2; corner term tab*len - delta, summed in order."""
    for row in tab:
        s = 0.0
        for d in range(len(length)):
            t = float(row[d]) * length[d] - delta[d]
            s += t * t
        if s > R2:
            return False
    return True


def farthest_corner(delta, length, tab, R2):
    """The single-sum form k_cover evaluates: per dimension the larger of the
    terms over the tab values present in that dimension, summed in order."""
    s = 0.0
    for d in range(len(length)):
        vals = sorted(set(int(r[d]) for r in tab))
        terms = []
        for v in vals:
            t = float(v) * length[d] - delta[d]
            terms.append(t * t)
        s += max(terms)
    return not (s > R2)


TABS = {
    2: [[0, 0], [1, 1], [0, 1], [1, 0]],
    3: [[0, 0, 0], [1, 1, 1], [0, 0, 1], [1, 1, 0], [0, 1, 0], [1, 0, 1], [0, 1, 1], [1, 0, 0]],
    # the reference's D >= 4 table leaves component 0 at 0 (auxiliary/Path.ixx:105-115)
    4: [[0] + [(i >> d) & 1 for d in range(1, 4)] for i in range(16)],
}


@pytest.mark.parametrize("dim", [2, 3, 4])
def test_farthest_corner_equals_corner_loop(dim):
    rng = np.random.default_rng(dim)
    tab = TABS[dim]
    n_true = 0
    for trial in range(20000):
        length = rng.uniform(1e-4, 2e-2, size=dim)
        if trial % 3 == 0:
            length[:] = length[0]
        delta = rng.uniform(-0.03, 0.05, size=dim)
        if trial % 5 == 0:  # centre exactly on a face or a corner coordinate
            d = rng.integers(dim)
            delta[d] = length[d] * rng.integers(0, 2)
        # R2 at, just below and just above one corner's exact sum: ties decide
        row = tab[rng.integers(len(tab))]
        s = 0.0
        for d in range(dim):
            t = float(row[d]) * length[d] - delta[d]
            s += t * t
        for R2 in (s, np.nextafter(s, 0.0), np.nextafter(s, 1.0), rng.uniform(0, 3e-3)):
            a = corner_loop(delta, length, tab, R2)
            assert a == farthest_corner(delta, length, tab, R2), (dim, delta, length, R2)
            n_true += a
    assert n_true > 1000  # both outcomes are exercised


def test_farthest_corner_needs_the_product_set():
    """A corner set that is not a product set breaks the identity (so k_cover's
    host check on the table is needed): corners (0,0) and (1,1) only."""
    tab = [[0, 0], [1, 1]]
    length = [1.0, 1.0]
    delta = [0.0, 1.0]  # sphere centred on corner (0, 1), which is not in the set
    R2 = 1.0
    assert corner_loop(delta, length, tab, R2) is True
    assert farthest_corner(delta, length, tab, R2) is False


# ---------------------------------------------------------------- lossless binary dump

def test_binary_dump_round_trip_cpp_python(tmp_path):
    """write_binary (C++ drop-in) is read bit for bit by rsa_mpi_amd.read_binary, and the Python
    writer's file is read back by the C++ read_binary."""
    import subprocess
    import rsa_mpi_amd as R
    exe = str(tmp_path / "dump")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include", "rsa_mpi_amd"),
                        os.path.join(ROOT, "tests", "cpp", "binary_dump.cpp"), "-L",
                        os.path.join(ROOT, "rsa_mpi_amd", "lib"), "-lrsa_mpi_amd",
                        "-Wl,-rpath," + os.path.join(ROOT, "rsa_mpi_amd", "lib"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    f3 = str(tmp_path / "a.bin")
    assert subprocess.run([exe, "w", f3]).returncode == 0
    pos, rad, ph = R.read_binary(f3)
    assert pos.tolist() == [[0.1, 0.2, 0.30000000000000004], [1e-300, 0.9999999999999999, 0.5],
                            [0.125, 0.25, 0.375]]
    assert rad.tolist() == [0.1, 5e-324, 0.0025]
    assert ph.tolist() == [0, 3, 2 ** 64 - 1]
    f2 = str(tmp_path / "b.bin")
    rng = np.random.default_rng(0)
    p2 = rng.random((5, 2))
    r2 = rng.random(5) * 0.01
    ph2 = np.arange(5, dtype=np.uint64) * 11
    R.write_binary(f2, p2, r2, ph2)
    out = subprocess.run([exe, "r", f2], capture_output=True, text=True).stdout.split("\n")
    assert int(out[0]) == 5
    for k in range(5):
        x, y, rr, p = out[1 + k].split()
        assert (float(x), float(y), float(rr), int(p)) == (p2[k, 0], p2[k, 1], r2[k], int(ph2[k]))
    with pytest.raises(ValueError):
        open(str(tmp_path / "c.bin"), "wb").write(b"not a dump at all, 24+ bytes")
        R.read_binary(str(tmp_path / "c.bin"))


# ---------------------------------------------------------------- coverage pre-check (k_cover)

EPS53 = 2.0 ** -53
TAB3 = TABS[3]


def reference_covered(q, qr, rmin, origin, c, ln):
    """auxi::is_covered for one sphere (list_of_voxels.ixx:146-153, 225-259): voxel origin
    origin + c * len, delta = centre - origin, every TabCorner corner within r + rmin."""
    o = [origin[d] + float(c[d]) * ln[d] for d in range(3)]
    delta = [q[d] - o[d] for d in range(3)]
    R2 = (qr + rmin) * (qr + rmin)
    for row in TAB3:
        s = 0.0
        for d in range(3):
            t = float(row[d]) * ln[d] - delta[d]
            s += t * t
        if s > R2:
            return False
    return True


def kernel_covered(q, qr, rmin, origin, c, ln, X, slack="scaled"):
    """k_cover's decision for one sphere, step for step in IEEE double: the pre-check
    |q - centre()|^2 <= R^2 - h^2 + slack, then the farthest-corner sum from the
    children's tab-0 origin origin + (2c) * (len / 2)."""
    nl = [0.5 * v for v in ln]
    ok0 = [origin[d] + float(2 * c[d]) * nl[d] for d in range(3)]
    cb = [origin[d] + (float(c[d]) + 0.5) * ln[d] for d in range(3)]
    rr = qr + rmin
    R2 = rr * rr
    h2 = 0.0
    for d in range(3):
        h2 += (0.5 * ln[d]) * (0.5 * ln[d])
    s2 = 0.0
    for d in range(3):
        t = q[d] - cb[d]
        s2 += t * t
    if slack == "fixed":  # the earlier form: a slack relative to R^2 only
        sl = 1e-9 * R2
    else:  # VoxParams::slack
        sl = 1e-9 * R2 + 64.0 * np.sqrt(3.0) * EPS53 * X * rr + 256.0 * 3 * (EPS53 * X) * (EPS53 * X)
    if not (s2 <= R2 - h2 + sl):
        return False
    sf = 0.0
    for d in range(3):
        dl = q[d] - ok0[d]
        aa = dl * dl
        b = ln[d] - dl
        bb = b * b
        sf += bb if bb > aa else aa
    return not (sf > R2)


def test_coverage_precheck_never_skips_a_covering_sphere():
    """Boxes far from the origin (|x| up to 1e8) with radii down to 1e-9 of the box and
    voxels refined up to 60 times, each sphere placed so that the box's farthest corner
    sits at the covering distance: k_cover's pre-check plus farthest-corner test decides
    exactly as the reference's corner loop.  A slack relative to R^2 alone (the earlier
    form) skips some covering spheres here, so the test can tell the two apart."""
    rng = np.random.default_rng(0)
    wrong_fixed = wrong = covered = 0
    for _ in range(40000):
        X0 = 10.0 ** rng.uniform(0, 8)
        origin = [X0 * rng.uniform(0.5, 1.0) * rng.choice([-1, 1]) for _ in range(3)]
        W = X0 * 10.0 ** rng.uniform(-3, 0)
        R = W * 10.0 ** rng.uniform(-9, -2)
        k = int(rng.integers(0, 60))
        ln = [W / 7 * 2.0 ** -k] * 3
        c = [int(rng.integers(0, 7 * 2 ** k)) for _ in range(3)]
        o = [origin[d] + c[d] * ln[d] for d in range(3)]
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        h = np.sqrt(sum((0.5 * v) ** 2 for v in ln))
        dist = (R - h) * rng.uniform(0.999999, 1.0) if R > h else 0.0
        q = [o[d] + 0.5 * ln[d] + u[d] * dist for d in range(3)]
        X = max(max(abs(origin[d]), abs(origin[d] + W)) + W for d in range(3))
        ref = reference_covered(q, 0.5 * R, 0.5 * R, origin, c, ln)
        covered += ref
        wrong += kernel_covered(q, 0.5 * R, 0.5 * R, origin, c, ln, X) != ref
        wrong_fixed += kernel_covered(q, 0.5 * R, 0.5 * R, origin, c, ln, X, slack="fixed") != ref
    assert covered > 10000
    assert wrong == 0
    assert wrong_fixed > 0


def test_listing_threshold_contains_both_gates():
    """k_cover lists a packed-copy sphere when s2 <= U, U = R^2 + slack(R^2, r + rmin) as
    k_compact computes it; the exact gates then run on the listed spheres.  Whenever
    either gate holds (the subdivide pre-filter s2 < R^2, the coverage pre-check
    s2 <= (R^2 - h^2) + slack), s2 <= U holds too, so no sphere a gate would pass is
    left out.  Spheres are placed at the gates' boundaries (both sides), far from the
    origin and with tiny radii as in the pre-check test."""
    rng = np.random.default_rng(1)
    passes = 0
    for _ in range(60000):
        X0 = 10.0 ** rng.uniform(0, 8)
        origin = [X0 * rng.uniform(0.5, 1.0) * rng.choice([-1, 1]) for _ in range(3)]
        W = X0 * 10.0 ** rng.uniform(-3, 0)
        R = W * 10.0 ** rng.uniform(-9, -2)
        k = int(rng.integers(0, 60))
        ln = [W / 7 * 2.0 ** -k] * 3
        c = [int(rng.integers(0, 7 * 2 ** k)) for _ in range(3)]
        cb = [origin[d] + (float(c[d]) + 0.5) * ln[d] for d in range(3)]
        X = max(max(abs(origin[d]), abs(origin[d] + W)) + W for d in range(3))
        qr = rmin = 0.5 * R
        rr = qr + rmin
        R2 = rr * rr
        h2 = 0.0
        for d in range(3):
            h2 += (0.5 * ln[d]) * (0.5 * ln[d])
        sl = 1e-9 * R2 + 64.0 * np.sqrt(3.0) * EPS53 * X * rr + 256.0 * 3 * (EPS53 * X) * (EPS53 * X)
        U = R2 + sl
        # a centre at distance close to one of the two gate radii
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        target = np.sqrt(max(R2 - h2 + sl, 0.0)) if rng.uniform() < 0.5 else rr
        dist = target * rng.uniform(0.9999999, 1.0000001)
        q = [cb[d] + u[d] * dist for d in range(3)]
        s2 = 0.0
        for d in range(3):
            t = q[d] - cb[d]
            s2 += t * t
        gate = (s2 < R2) or (s2 <= R2 - h2 + sl)
        passes += gate
        assert not gate or s2 <= U, (s2, U, R2, h2, sl)
    assert passes > 20000


def _network(name):
    """A comparator table of the draw kernels, read from their source."""
    src = open(os.path.join(ROOT, "rsa_mpi_amd", "csrc", "k_draw.hpp")).read()
    m = re.search(r"constexpr int %s\[(\d+)\]\[2\] = \{(.*?)\};" % name, src, re.S)
    assert m, name
    pairs = [(int(a), int(b)) for a, b in re.findall(r"\{\s*(\d+),\s*(\d+)\s*\}", m.group(2))]
    assert len(pairs) == int(m.group(1))
    return pairs


def _sorts_all_01(net, keys, skip=-1):
    v = ((np.arange(1 << keys)[:, None] >> np.arange(keys)[None, :]) & 1).astype(np.uint8)
    for k, (a, b) in enumerate(net):
        if k == skip:
            continue
        lo, hi = np.minimum(v[:, a], v[:, b]), np.maximum(v[:, a], v[:, b])
        v[:, a], v[:, b] = lo, hi
    return bool((np.diff(v.astype(np.int8), axis=1) >= 0).all())


@pytest.mark.parametrize("name,keys", [("kSortNet", 8), ("kSortNet16", 16)])
def test_sorting_networks_sort_every_input(name, keys):
    """The cell lists are put in shot order by fixed comparator networks (k_lists,
    the partner walk, the sweep's commit).  By the 0-1 principle a comparator
    network sorts every input iff it sorts every 0-1 input: all 2^keys of them
    are run through it.  Every comparator of both networks is needed, so a table
    with any one comparator missing fails the same check."""
    net = _network(name)
    assert all(0 <= a < b < keys for a, b in net)
    assert _sorts_all_01(net, keys)
    for drop in range(len(net)):
        assert not _sorts_all_01(net, keys, drop), (name, drop)
