"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle on the
same seeded inputs.  Integer/byte/order results must be identical and every
FP64 centre bit-identical."""
import numpy as np
import pytest

import rsa_mpi_amd as R
import _oracle as O
from _common import set_hash

pytestmark = pytest.mark.gpu


def gpu_run(dim, L, classes, seed=0, mode=1, nbshots=6000, max_draws=-1, drad=None, reiter=()):
    inf = [0.0] * dim
    sup = [float(L)] * dim
    rmax = max(c[0] for c in classes)
    dom = R.Domain(inf, sup, drad if drad is not None else rmax)
    if mode == 1:
        gen = R.RadiusGenerator(classes, dom.get_total_volume(), 0.0)
    else:
        gen = R.RadiusGenerator(classes, 0.0)
    algo = R.Algo(dom, gen, nbshots)
    st = algo.proceed(seed, max_draws)
    for r in reiter:
        algo.reset_radius_generator(R.RadiusGenerator([(r, 1.0, 0)], dom.get_total_volume(), 0.0))
        st = algo.proceed(seed, max_draws)
    pos, rad, ph = dom.extract_arrays()
    return dict(pos=pos, rad=rad, phase=ph, stats=st, dom=dom)


def orc_run(dim, L, classes, **kw):
    return O.run(dim, [0.0] * dim, [float(L)] * dim, classes, **kw)


def assert_checks(st, dim=3):
    """The device's end-of-run checks: no two spheres (images included) collide and, for a
    packing that ended fully packed, no sampled point is left free for a smallest sphere.
    In 4D the reference's corner table leaves a facet of every voxel untested
    (auxiliary/Path.ixx:105-115), so its fully packed runs need not be maximal; the check
    is reported there, not required."""
    assert st["overlap_pairs"] == 0, st
    assert st["cover_samples"] >= 1 << 16
    if st["end_reason"] == 1 and dim <= 3:
        assert st["cover_free"] == 0, st


def assert_same(g, o):
    print("gpu n", len(g["rad"]), "oracle n", o["n"], "stats", g["stats"])
    assert_checks(g["stats"], g["pos"].shape[1])
    assert len(g["rad"]) == o["n"]
    assert set_hash(g["pos"], g["rad"]) == set_hash(o["pos"], o["rad"])
    # extraction order (real cells by index, then in-cell order) is the reference's
    assert np.array_equal(g["pos"].view(np.uint64), o["pos"].view(np.uint64))
    assert np.array_equal(g["rad"].view(np.uint64), o["rad"].view(np.uint64))
    assert np.array_equal(g["phase"], o["phase"])


@pytest.mark.parametrize("seed", [0, 1])
def test_2d_bit_exact(seed):
    """2D r=0.01 in the unit square to saturation (the reference's CPU case)."""
    cls = [(0.01, 1.0, 0)]
    g = gpu_run(2, 1.0, cls, seed=seed)
    o = orc_run(2, 1.0, cls, seed=seed)
    assert_same(g, o)
    tr = o["trace"]
    st = g["stats"]
    assert st["draws"] == len(tr)
    assert st["shots"] == int(tr[:, 2].sum())
    assert st["refines"] == int(tr[:, 5].sum())
    assert st["words"] == o["words"]
    assert st["end_reason"] == 1


@pytest.mark.parametrize("L,r,seed", [(0.2, 0.005, 0), (0.15, 0.01, 3)])
def test_small_3d_bit_exact(L, r, seed):
    cls = [(r, 1.0, 0)]
    g = gpu_run(3, L, cls, seed=seed)
    o = orc_run(3, L, cls, seed=seed)
    assert_same(g, o)
    assert g["stats"]["words"] == o["words"]


def test_3d_r0005_full_size_matches_golden():
    """3D r=0.005 unit cube (~734k spheres): the whole packing against the
    oracle's committed fixture, plus size-independent properties."""
    import json
    import os
    from _common import periodic_overlaps, uncovered_samples, volume_fraction
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "3d_r0.005_seed0.json")))
    g = gpu_run(3, 1.0, [(0.005, 1.0, 0)], seed=0)
    st = g["stats"]
    print(st)
    assert len(g["rad"]) == gold["n"]
    assert set_hash(g["pos"], g["rad"]) == gold["sha256"]
    assert st["words"] == gold["words"]
    assert st["draws"] == gold["draws"] and st["shots"] == gold["shots"] and st["refines"] == gold["refines"]
    assert st["rounds"] == gold["rounds"]
    assert st["nvox_max"] == gold["nvox_max"]
    assert st["end_reason"] == 1 and st["nvox_final"] == 0
    assert periodic_overlaps(g["pos"], g["rad"], [1.0] * 3) == 0
    assert uncovered_samples(g["pos"], g["rad"], [1.0] * 3, 0.005, n=200000) == 0
    assert 0.38 < volume_fraction(g["rad"], 3, 1.0) < 0.39


@pytest.mark.parametrize("dim,L,r,seed", [(2, 1.0, 0.01, 0), (3, 0.2, 0.005, 1)])
def test_stop_after_k_draws_bit_exact(dim, L, r, seed):
    """proceed(seed, max_draws=k): spheres, stream position and the voxel list
    where the run stopped, for k around the refinements (operator_algorithm.ixx:92-95)."""
    cls = [(r, 1.0, 0)]
    tr = orc_run(dim, L, cls, seed=seed)["trace"]
    ref = np.nonzero(tr[:, 5])[0]
    ks = sorted({1, 2, 3, int(ref[0]), int(ref[0]) + 1, int(ref[len(ref) // 2]) + 1, len(tr) - 1})
    for k in ks:
        print("max_draws", k, flush=True)
        g = gpu_run(dim, L, cls, seed=seed, max_draws=k)
        o = orc_run(dim, L, cls, seed=seed, max_draws=k)
        assert_same(g, o)
        st = g["stats"]
        assert st["draws"] == k and st["end_reason"] == 3
        assert st["words"] == o["words"]
        co, ln = g["dom"].voxels()
        assert np.array_equal(co, o["vox"])
        assert np.array_equal(ln.view(np.uint64), o["vox_len"].view(np.uint64))


@pytest.mark.parametrize("dim,L,classes,seed", [
    (2, 1.0, [(0.02, 60, 0), (0.01, 500, 1), (0.005, 4000, 2)], 0),
    (2, 0.5, [(0.01, 300, 3), (0.004, 500, 4)], 5),
    (3, 0.2, [(0.01, 300, 0), (0.005, 3000, 1)], 2),
    # an outreach draw whose pivot keeps none of its 4 Undecided candidates for 2 places
    # (the reference's dichotomy, median_of_medians.ixx:103-118), then a fill class
    (2, 0.3, [(0.02, 20, 0), (0.012, 50, 1), (0.008, 10 ** 9, 2)], 0),
])
def test_count_classes_bit_exact(dim, L, classes, seed):
    """(radius, count, phase) classes placed largest first: draws that may
    overshoot a class target keep the first ones by priority
    (operator_algorithm.ixx:416-421), then the next class starts."""
    g = gpu_run(dim, L, classes, seed=seed, mode=0)
    o = orc_run(dim, L, classes, seed=seed, mode=0)
    assert_same(g, o)
    assert g["stats"]["words"] == o["words"]
    for rad, cnt, ph in classes:
        got = int((g["phase"] == ph).sum())
        assert got <= cnt
        assert np.all(g["rad"][g["phase"] == ph] == rad)


@pytest.mark.parametrize("dim,L,r0,reiter,seed", [(3, 0.2, 0.01, (0.005,), 4), (2, 1.0, 0.02, (0.01, 0.005), 1)])
def test_reiteration_bit_exact(dim, L, r0, reiter, seed):
    """reset_radius_generator + proceed(seed) again on the same domain
    (scripts/main_max_frac.cpp:43-52): the new class fills the gaps."""
    cls = [(r0, 1.0, 0)]
    g = gpu_run(dim, L, cls, seed=seed, reiter=reiter)
    o = orc_run(dim, L, cls, seed=seed, reiter=reiter)
    assert_same(g, o)
    assert periodic_overlaps_ok(g, dim, L)


def periodic_overlaps_ok(g, dim, L):
    from _common import periodic_overlaps
    return periodic_overlaps(g["pos"], g["rad"], [L] * dim) == 0


def _golden(name):
    import json
    import os
    return json.load(open(os.path.join(os.path.dirname(__file__), "golden", name + ".json")))


def _check_golden(g, gold, classes_phase=None):
    st = g["stats"]
    print(st)
    assert_checks(st, g["pos"].shape[1])
    assert len(g["rad"]) == gold["n"]
    assert set_hash(g["pos"], g["rad"]) == gold["sha256"]
    assert st["words"] == gold["words"]
    assert st["draws"] == gold["draws"] and st["shots"] == gold["shots"] and st["refines"] == gold["refines"]
    assert st["rounds"] == gold["rounds"]
    assert st["nvox_max"] == gold["nvox_max"]
    assert st["end_reason"] == 1 and st["nvox_final"] == 0
    if "phases" in gold:
        for ph, cnt in gold["phases"].items():
            assert int((g["phase"] == int(ph)).sum()) == cnt


def test_3d_polydisperse_matches_golden():
    """3D polydisperse: classes r=0.01/0.005/0.0025 at volume fraction 0.10
    each, then a fill class r=0.0025 (4.35M spheres; 1.15 billion voxels at the
    peak), against the oracle's committed fixture."""
    gold = _golden("3d_poly_seed0")
    cls = [tuple(c) for c in gold["case"]["classes"]]
    g = gpu_run(3, 1.0, cls, seed=0)
    _check_golden(g, gold)


def test_3d_r0002_matches_golden():
    """3D monodisperse r=0.002 in the unit cube (11.46M spheres, 12144 draws)."""
    from _common import uncovered_samples, volume_fraction
    gold = _golden("3d_r0.002_seed0")
    g = gpu_run(3, 1.0, [(0.002, 1.0, 0)], seed=0)
    _check_golden(g, gold)
    assert uncovered_samples(g["pos"], g["rad"], [1.0] * 3, 0.002, n=100000) == 0
    assert 0.38 < volume_fraction(g["rad"], 3, 1.0) < 0.39


@pytest.mark.gpu
def test_custom_radius_law_2d_python():
    """LawRadiusGenerator (RadiusGenerator(min, max, RandomRadiusGenerator*, n)) through the Python facade:
    a uniform law on [0.008, 0.012] in the unit square, 1000 spheres wanted (the run ends on the class count)."""
    import rsa_mpi_amd as R
    dom = R.Domain([0.0, 0.0], [1.0, 1.0], 0.012)
    gen = R.LawRadiusGenerator(0.008, 0.012, lambda u: 0.008 + (0.012 - 0.008) * u, 1000, phase=3)
    algo = R.Algo(dom, gen, 6000)
    st = algo.proceed(5)
    pos, rad, ph = dom.extract_arrays()
    o = O.run(2, [0.0, 0.0], [1.0, 1.0], [(0.008, 1000.0, 3), (0.012, 0.0, 3)], seed=5, mode=2)
    assert len(rad) == o["n"] == 1000
    assert set_hash(pos, rad) == set_hash(o["pos"], o["rad"])
    assert (ph == 3).all()
    assert st["end_reason"] == 2


@pytest.mark.gpu
@pytest.mark.parametrize("L,r,seed", [(0.3, 0.03, 0), (0.5, 0.03, 1)])
def test_small_4d_bit_exact(L, r, seed):
    """4D (the reference's D >= 4 TabCorner table, kept for parity) against the oracle."""
    cls = [(r, 1.0, 0)]
    g = gpu_run(4, L, cls, seed=seed)
    o = orc_run(4, L, cls, seed=seed)
    assert_same(g, o)
    assert g["stats"]["words"] == o["words"]


@pytest.mark.gpu
@pytest.mark.parametrize("L,r,seed", [(0.2, 0.05, 0), (0.3, 0.05, 3), (0.5, 0.05, 1), (0.56, 0.04, 2)])
def test_dense_4d_conflicts_bit_exact(L, r, seed):
    """4D boxes of 2 to 7 cells per side with 6000 candidates per draw: most candidates of the early draws
    collide with each other, so the conflict sweep sees long partner lists over the 81-cell stencil, whose
    order (stencil, then in-cell) decides the labels (operator_cells_and_data_storage.ixx:16-52, 127-142)."""
    cls = [(r, 1.0, 0)]
    g = gpu_run(4, L, cls, seed=seed)
    o = orc_run(4, L, cls, seed=seed)
    assert_same(g, o)
    assert g["stats"]["words"] == o["words"]
    assert g["stats"]["draws"] == len(o["trace"])


@pytest.mark.gpu
def test_4d_r002_matches_golden():
    """4D monodisperse r=0.02 in the unit hypercube (305 580 spheres, 1584 draws) against the oracle's
    fixture.  The reference's 4D corner table leaves component 0 at 0, so the voxel path is not
    maximal (kept for parity); only the set and the trace totals are checked."""
    gold = _golden("4d_r0.02_seed0")
    g = gpu_run(4, 1.0, [(0.02, 1.0, 0)], seed=0)
    _check_golden(g, gold)
    # the device's maximality check finds the points that defect leaves free
    assert g["stats"]["cover_free"] > 0


# ------------------------------------------------------------ the reference's own test programs

@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3, 4])
def test_reference_full_voxel(dim):
    """test/test_{2D,3D,4D}_full_voxel.cpp: test_ND<DIM, 1>(0.05) (test.hxx:53-67), unit box, seed 0."""
    cls = [(0.05, 1.0, 0)]
    g = gpu_run(dim, 1.0, cls, seed=0)
    o = orc_run(dim, 1.0, cls, seed=0)
    assert_same(g, o)
    assert g["stats"]["words"] == o["words"]


@pytest.mark.gpu
def test_reference_3d_radius():
    """test/test_3D_radius.cpp: test_ND_radius<3, 1>({(0.25, 0.1, 0), (0.125, 0.1, 1)}) with volume 1
    (test.hxx:34-51): two cells per dimension for the largest radius."""
    cls = [(0.25, 0.1, 0), (0.125, 0.1, 1)]
    g = gpu_run(3, 1.0, cls, seed=0)
    o = orc_run(3, 1.0, cls, seed=0)
    assert_same(g, o)


@pytest.mark.gpu
def test_reference_user_interface_cpp():
    """test/test_python.cpp: user_interface domain [0, 3.5]^3, r=0.4 at volume fraction 1, seed 0."""
    cls = [(0.4, 1.0, 0)]
    g = gpu_run(3, 3.5, cls, seed=0)
    o = orc_run(3, 3.5, cls, seed=0)
    assert_same(g, o)


@pytest.mark.gpu
def test_reference_bug_2d_matches_golden():
    """test.hxx:104-109 test_bug_2D: throw_spheres<2, 1>({1, 1}, {(0.0005, 1, 0)}, 0), written in the
    reference to provoke subdivisions past 32-bit voxel coordinates (696 552 disks, 23 refinements)."""
    from _common import uncovered_samples
    gold = _golden("bug2d_2d_r0.0005_seed0")
    g = gpu_run(2, 1.0, [(0.0005, 1.0, 0)], seed=0)
    _check_golden(g, gold)
    assert uncovered_samples(g["pos"], g["rad"], [1.0] * 2, 0.0005, n=100000) == 0


@pytest.mark.gpu
def test_domain_add_round_trip():
    """rsa_domain::add (rsa_domain.hxx:84, build_grid with Validated tags): a packing's spheres added to a fresh
    domain in extraction order come back bit for bit in the same order; proceed on the full packing then adds
    nothing, and the add_to_sample counter (get_number_of_spheres_fast) is left as add leaves it."""
    g = gpu_run(3, 0.2, [(0.01, 1.0, 0)], seed=4)
    dom = R.Domain([0.0] * 3, [0.2] * 3, 0.01)
    gen = R.RadiusGenerator([(0.01, 1.0, 0)], dom.get_total_volume(), 0.0)
    algo = R.Algo(dom, gen, 6000)
    dom.add(g["pos"], g["rad"], g["phase"])
    pos, rad, ph = dom.extract_arrays()
    assert np.array_equal(pos.view(np.uint64), g["pos"].view(np.uint64))
    assert np.array_equal(rad.view(np.uint64), g["rad"].view(np.uint64))
    assert np.array_equal(ph, g["phase"])
    assert dom.number_of_spheres() == 0
    st = algo.proceed(4)
    assert st["accepted"] == 0 and st["nvox_final"] == 0
    assert len(dom.extract_arrays()[1]) == len(g["rad"])


@pytest.mark.gpu
@pytest.mark.parametrize("dim,L,r,seed", [(2, 1.0, 0.01, 0), (3, 0.2, 0.005, 1), (4, 0.3, 0.05, 3), (4, 0.56, 0.04, 2)])
def test_list_redo_path_bit_exact(monkeypatch, dim, L, r, seed):
    """A draw's first round builds the cell lists inside its kernels; a cell of more candidates than the threshold
    (8, lowered here to 1 through RSA_GPU_LIST_THRESHOLD) makes the host build them with k_lists and redo the round.
    With the threshold at 1 most first rounds take that path: the profile counts the redos, and the packing, word
    and draw counts are the oracle's bit for bit."""
    monkeypatch.setenv("RSA_GPU_LIST_THRESHOLD", "1")
    cls = [(r, 1.0, 0)]
    dom = R.Domain([0.0] * dim, [L] * dim, r)
    gen = R.RadiusGenerator(cls, dom.get_total_volume(), 0.0)
    algo = R.Algo(dom, gen, 6000)
    dom.profile(1)
    st = algo.proceed(seed)
    redo = dom.kernel_profile().get("host_list_redo", {}).get("launches", 0)
    print("list redos", redo, "draws", st["draws"])
    assert redo > 0
    pos, rad, ph = dom.extract_arrays()
    o = orc_run(dim, L, cls, seed=seed)
    assert_same(dict(pos=pos, rad=rad, phase=ph, stats=st), o)
    assert st["words"] == o["words"]
    assert st["draws"] == len(o["trace"])


def reference_volume(dim, L, rmax, pos, rad):
    """compute_total_volume_of_spheres on one rank, one OpenMP thread (rsa_domain.ixx:171-175,
    rsa_grid.ixx:74-86): per real cell the sum of volumeSphere(r) = mult * puissance<D>(r) in
    in-cell order, the cell sums added in cell index order; mult = volumeUnitSphere(D)
    (auxiliary/Geometry/Sphere.ixx:31-57), puissance as in AuxiFunctions.ixx:34-44.  pos and
    rad come in extraction order (cells by index, in-cell order)."""
    import math
    mult = 1.0
    if dim % 2 == 0:
        for k in range(dim // 2):
            mult *= (2 * math.pi) / (dim - 2 * k)
    else:
        mult = 1.0 / math.pi
        for k in range((dim - 1) // 2 + 1):
            mult *= (2 * math.pi) / (dim - 2 * k)
    r = np.asarray(rad, dtype=np.float64)
    p = {2: r * r, 3: r * (r * r), 4: (r * r) * (r * r)}[dim]
    v = mult * p
    # cell of each centre (compute_cell_idx with the grid of rsa_grid_traversal::define)
    nin = int(L / (2 * rmax))
    cl = L / nin
    shift = 0.0 - cl
    inv = 1.0 / cl
    n = nin + 2
    idx = np.zeros(len(r), dtype=np.int64)
    for d in range(dim - 1, -1, -1):
        idx = idx * n + ((np.asarray(pos)[:, d] - shift) * inv).astype(np.int64)
    assert np.all(np.diff(idx) >= 0)  # extraction order is cell order
    starts = np.r_[0, np.nonzero(np.diff(idx))[0] + 1]
    res = 0.0
    for a, b in zip(starts, np.r_[starts[1:], len(r)]):
        loc = 0.0
        for x in v[a:b]:
            loc += float(x)
        res += loc
    return res


@pytest.mark.parametrize("dim,L,classes,seed", [(2, 0.5, [(0.01, 1.0, 0)], 1), (3, 0.2, [(0.005, 1.0, 0)], 0),
                                                (3, 0.3, [(0.01, 0.1, 0), (0.005, 1.0, 1)], 3)])
def test_volume_of_spheres_is_the_reference_sum(dim, L, classes, seed):
    """rsa_gpu_volume_of_spheres equals, bit for bit, the reference's per-cell summation
    restated on the oracle's spheres (the placed sets are equal, checked first)."""
    g = gpu_run(dim, L, classes, seed=seed)
    o = orc_run(dim, L, classes, seed=seed)
    assert_same(g, o)
    want = reference_volume(dim, L, max(c[0] for c in classes), o["pos"], o["rad"])
    got = g["dom"].compute_total_volume_of_spheres()
    assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64), (got, want)


@pytest.mark.parametrize("dim,off,L,r,seed", [(3, 1.0e5, 0.12, 0.01, 2), (2, -7.5e6, 0.3, 0.004, 1),
                                              (3, 3.0e7, 0.1, 0.006, 0)])
def test_offset_domain_bit_exact(dim, off, L, r, seed):
    """A periodic box far from the origin ([off, off + L]^D): coordinates carry large
    exponents while radii and voxels stay small, so the coverage pre-check's slack must
    scale with the coordinates' magnitude.  The packing is the oracle's bit for bit."""
    inf = [off] * dim
    sup = [off + L] * dim
    dom = R.Domain(inf, sup, r)
    gen = R.RadiusGenerator([(r, 1.0, 0)], dom.get_total_volume(), 0.0)
    algo = R.Algo(dom, gen, 6000)
    st = algo.proceed(seed)
    pos, rad, ph = dom.extract_arrays()
    o = O.run(dim, inf, sup, [(r, 1.0, 0)], seed=seed)
    assert st["end_reason"] == 1 and o["msg"].endswith("fully packed")
    assert_checks(st, dim)
    assert len(rad) == o["n"] > 0
    assert np.array_equal(pos.view(np.uint64), o["pos"].view(np.uint64))
    assert st["words"] == o["words"]


def test_device_checks_find_planted_faults():
    """The end-of-run checks report what they are built to find: spheres handed in with
    Domain.add that overlap (each such pair counted once), and a box left almost empty
    (free points sampled)."""
    dom = R.Domain([0.0] * 3, [0.2] * 3, 0.01)
    algo = R.Algo(dom, R.RadiusGenerator([(0.01, 1.0, 0)], dom.get_total_volume(), 0.0), 6000)
    dom.add([[0.05, 0.05, 0.05], [0.06, 0.05, 0.05], [0.15, 0.15, 0.15], [0.195, 0.15, 0.15],
             [0.004, 0.15, 0.15]], [0.01] * 5)
    st = algo.proceed(0, 1)  # one draw: the box stays almost empty
    assert st["overlap_pairs"] == 2  # (0.05, 0.06) and (0.195, 0.004) across the periodic face
    assert st["end_reason"] == 3 and st["cover_free"] > st["cover_samples"] // 2


def test_reset_to_custom_law_python_bit_exact():
    """Algo.reset_radius_generator(LawRadiusGenerator) between proceeds (the re-iteration of
    scripts/main_max_frac.cpp:43-62 with a law generator), then a class again."""
    dom = R.Domain([0.0] * 3, [0.2] * 3, 0.01)
    algo = R.Algo(dom, R.RadiusGenerator([(0.01, 0.2, 0)], dom.get_total_volume(), 0.0), 6000)
    algo.proceed(3)
    algo.reset_radius_generator(R.LawRadiusGenerator(0.006, 0.008, lambda u: 0.006 + (0.008 - 0.006) * u, 400,
                                                     phase=2))
    st = algo.proceed(3)
    assert st["end_reason"] == 2 and st["accepted"] == 400
    algo.reset_radius_generator(R.RadiusGenerator([(0.005, 1.0, 0)], dom.get_total_volume(), 0.0))
    st = algo.proceed(3)
    assert_checks(st)
    pos, rad, ph = dom.extract_arrays()
    o = O.run(3, [0.0] * 3, [0.2] * 3, [(0.01, 0.2, 0)], seed=3, reiter=(("law", 0.006, 0.008, 2, 400), 0.005))
    assert_same(dict(pos=pos, rad=rad, phase=ph, stats=st), o)
    with pytest.raises(RuntimeError, match="Incorrect maximal radius"):
        algo.reset_radius_generator(R.LawRadiusGenerator(0.006, 0.02, lambda u: 0.006 + 0.014 * u, 10))


@pytest.mark.parametrize("inf,sup,r,seed,nbshots", [
    ([0.0, 0.0], [1.0, 0.4], 0.01, 2, 6000),
    ([-0.1, 0.2, 0.0], [0.2, 0.45, 0.18], 0.006, 1, 6000),
    ([0.0, 0.0, 0.0], [0.5, 0.1, 0.3], 0.008, 4, 6000),
    ([0.0, 0.0, 0.0, 0.0], [0.3, 0.4, 0.3, 0.35], 0.04, 0, 6000),
    ([0.0, 0.0], [0.6, 0.6], 0.012, 5, 100),  # many small draws
])
def test_non_cubic_boxes_bit_exact(inf, sup, r, seed, nbshots):
    """Boxes with a different side per dimension (and offsets per dimension), and a small
    nbshots_singledraw, against the oracle: same spheres, same order, same words."""
    dim = len(inf)
    dom = R.Domain(inf, sup, r)
    gen = R.RadiusGenerator([(r, 1.0, 0)], dom.get_total_volume(), 0.0)
    st = R.Algo(dom, gen, nbshots).proceed(seed)
    pos, rad, ph = dom.extract_arrays()
    o = O.run(dim, inf, sup, [(r, 1.0, 0)], seed=seed, nbshots=nbshots)
    assert st["words"] == o["words"] and st["draws"] == len(o["trace"])
    assert_same(dict(pos=pos, rad=rad, phase=ph, stats=st), o)


def test_reference_error_messages():
    """The reference's own errors: a box narrower than twice the domain radius (the grid
    is built first and refuses it, rsa_grid_traversal.ixx:18-23, before
    rsa_domain.ixx:135-142 could), an empty radius generator, a generator whose largest
    radius exceeds the domain's; a zero-draw proceed places nothing."""
    with pytest.raises(RuntimeError, match="Too small domain wrt ghost layer"):
        R.Domain([0.0, 0.0, 0.0], [0.1, 0.5, 0.5], 0.06)
    dom = R.Domain([0.0] * 3, [0.2] * 3, 0.01)
    with pytest.raises(RuntimeError, match="Empty radius generator!"):
        R.Algo(dom, R.RadiusGenerator([], 0.0), 6000)
    with pytest.raises(RuntimeError, match="Impossible to have the maximal radius of radius generator larger than "
                                           "domain implicit radius"):  # operator_algorithm.ixx:26-28
        R.Algo(dom, R.RadiusGenerator([(0.02, 10, 0)], 0.0), 6000)
    algo = R.Algo(dom, R.RadiusGenerator([(0.01, 1.0, 0)], dom.get_total_volume(), 0.0), 6000)
    st = algo.proceed(0, 0)
    assert st["draws"] == 0 and st["end_reason"] == 3 and dom.extract_arrays()[1].size == 0
