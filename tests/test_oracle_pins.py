"""Pins of the CPU oracle (oracle/rsa_oracle.hpp) by checks it cannot pass by
construction: the C++ standard's known answers, closed forms restated here from
their definitions, brute-force geometry on small inputs, and the sequential
greedy semantics of RSA.  The reference ships no golden vectors for this path
(its tests are smoke runs), so these checks are what anchors parity."""
import math

import numpy as np
import pytest

import _oracle as O
from _common import periodic_overlaps, uncovered_samples, volume_fraction


# ---------------------------------------------------------------- RNG (rsa_random.hxx:37-42)

def test_mt19937_default_seed_known_answer():
    # [rand.predef]: the 10000th consecutive output of a default-constructed mt19937
    assert int(O.mt_default_words(10000)[9999]) == 4123659995


def test_uniform_int_is_lemire_multiply_shift():
    """uniform_int_distribution<int>(0, N-1) on mt19937: index = (g*N) >> 32,
    words whose low half falls below (2^32 - N) mod N are rejected.  The HIP
    decode kernel uses exactly this rule."""
    for seed, N in [(0, 5268024), (3, 81746504), (11, 7), (5, 2 ** 31 - 1), (9, 1000003)]:
        vals, used = O.dist_samples(seed, 0, N - 1, 2000)
        w = O.mt_words(seed, used).astype(np.uint64)
        thr = (2 ** 32 - N) % N
        out, k = [], 0
        while len(out) < 2000:
            prod = int(w[k]) * N
            k += 1
            if (prod & 0xFFFFFFFF) < thr:
                continue
            out.append(prod >> 32)
        assert k == used
        assert np.array_equal(np.array(out), vals.astype(np.int64))


def test_priority_is_word_shift():
    """generate_priority<int> draws uniform_int(0, INT_MAX) = g >> 1 (never rejects)."""
    vals, used = O.dist_samples(4, 0, 2 ** 31 - 1, 5000)
    assert used == 5000
    w = O.mt_words(4, 5000)
    assert np.array_equal((w >> 1).astype(np.int64), vals.astype(np.int64))


def test_canonical_double_from_two_words():
    """uniform_real_distribution(0,1) = generate_canonical<double,53>:
    (g1 + g2*2^32)/2^64 rounded once, values >= 1 mapped below 1."""
    vals, used = O.dist_samples(2, 1, 0.0, 4000)
    assert used == 8000
    w = O.mt_words(2, used).astype(np.float64)
    s = w[0::2] + w[1::2] * 4294967296.0
    u = s / 18446744073709551616.0
    u[u >= 1.0] = np.nextafter(1.0, 0.0)
    assert np.array_equal(u.view(np.uint64), vals.view(np.uint64))


def test_poisson_consumes_the_same_stream():
    """Poisson counts on the stream: mean >= 12 uses the rejection method (words
    vary), small means the product-of-uniforms method (two words per uniform)."""
    small, used_small = O.dist_samples(8, 2, 0.05, 300)
    assert used_small % 2 == 0 and used_small >= 600
    big, used_big = O.dist_samples(8, 2, 6000.0, 300)
    assert abs(big.mean() - 6000) < 4 * math.sqrt(6000 / 300)


# ---------------------------------------------------------------- grid (rsa_grid_traversal.ixx:30-66)

@pytest.mark.parametrize("dim,n", [(2, [52, 52]), (3, [102, 102, 102]), (3, [4, 5, 6]), (4, [7, 7, 7, 7])])
def test_neighbour_order_is_a_norm_sort(dim, n):
    order = O.neighbour_table(dim, n)
    assert sorted(order.tolist()) == list(range(3 ** dim))
    vec = np.array([[(i // 3 ** d) % 3 - 1 for d in range(dim)] for i in range(3 ** dim)])
    norms = (vec[order] ** 2).sum(axis=1)
    assert np.all(np.diff(norms) >= 0)
    assert order[0] == (3 ** dim - 1) // 2  # the cell itself first
    if dim == 2:  # n <= 16: libstdc++ sorts by insertion, which is stable
        assert order.tolist() == sorted(range(9), key=lambda i: (norms[list(order).index(i)], i))


# ---------------------------------------------------------------- voxels (list_of_voxels.ixx)

def literal_remove_covered(coords, flags):
    """The two-pointer loop of list_of_voxels.ixx:42-53, verbatim semantics."""
    c = [tuple(x) for x in coords]
    f = list(flags)
    old = len(c)
    i = 0
    while i < old:
        if f[i]:
            c[i] = c[old - 1]
            f[i] = f[old - 1]
            old -= 1
            continue
        i += 1
    return np.array(c[:old], dtype=np.int64).reshape(-1, coords.shape[1])


def closed_form_remove_covered(coords, flags):
    """Closed form used by the HIP kernels: holes below keep (increasing) receive
    the uncovered voxels at or above keep in decreasing order."""
    flags = np.asarray(flags, dtype=bool)
    keep = int((~flags).sum())
    out = coords[:keep].copy()
    holes = np.nonzero(flags[:keep])[0]
    src = np.nonzero(~flags[keep:])[0][::-1] + keep
    assert len(holes) == len(src)
    out[holes] = coords[src]
    return out


@pytest.mark.parametrize("p", [0.0, 0.1, 0.5, 0.9, 1.0])
def test_swap_compaction_closed_form(p):
    rng = np.random.default_rng(int(p * 100) + 1)
    for n in (1, 2, 7, 64, 1000):
        coords = rng.integers(0, 1000, size=(n, 3))
        flags = rng.random(n) < p
        assert np.array_equal(literal_remove_covered(coords, flags), closed_form_remove_covered(coords, flags))


TAB3 = np.array([[0, 0, 0], [1, 1, 1], [0, 0, 1], [1, 1, 0], [0, 1, 0], [1, 0, 1], [0, 1, 1], [1, 0, 0]])
TAB2 = np.array([[0, 0], [1, 1], [0, 1], [1, 0]])


def brute_cover(dim, L, pos, rad, coords, length, rmin):
    """Coverage of each voxel by any sphere or periodic image (all 3^D shifts),
    with the reference's corner predicate: every corner within r + rmin."""
    tab = TAB3 if dim == 3 else TAB2
    ov = coords * length  # origin 0
    corners = tab * length
    shifts = np.array([[((i // 3 ** d) % 3 - 1) * L for d in range(dim)] for i in range(3 ** dim)])
    cov = np.zeros(len(coords), dtype=bool)
    for s in shifts:
        img = pos + s
        for j in range(len(rad)):
            delta = img[j] - ov  # [nv, dim]
            r2 = (rad[j] + rmin) ** 2
            ok = np.ones(len(coords), dtype=bool)
            for c in corners:
                t = c - delta
                ok &= (t * t).sum(axis=1) <= r2
            cov |= ok
    return cov


@pytest.mark.parametrize("dim", [2, 3])
def test_coverage_flags_match_brute_force(dim):
    rng = np.random.default_rng(dim)
    L, r = 0.3, 0.03
    n = 40 if dim == 3 else 25
    pos = rng.random((n, dim)) * L
    rad = np.full(n, r)
    nv = 1 + int(math.sqrt(dim) * L / (2 * r))
    length = np.full(dim, L / nv)
    coords = np.array(np.meshgrid(*[np.arange(nv)] * dim, indexing="ij")).reshape(dim, -1).T
    flags = O.voxel_step(dim, [0] * dim, [L] * dim, r, pos, rad, coords, length, r, mode=3)
    brute = brute_cover(dim, L, pos, rad, coords, length, r)
    assert flags.sum() > 0
    assert np.array_equal(flags, brute)


@pytest.mark.parametrize("dim", [2, 3])
def test_subdivide_matches_brute_force(dim):
    rng = np.random.default_rng(10 + dim)
    L, r = 0.3, 0.03
    n = 30
    pos = rng.random((n, dim)) * L
    rad = np.full(n, r)
    nv = 1 + int(math.sqrt(dim) * L / (2 * r))
    length = np.full(dim, L / nv)
    coords = np.array(np.meshgrid(*[np.arange(nv)] * dim, indexing="ij")).reshape(dim, -1).T
    out, nl = O.voxel_step(dim, [0] * dim, [L] * dim, r, pos, rad, coords, length, r, mode=1)
    assert np.array_equal(nl, length * 0.5)
    tab = TAB3 if dim == 3 else TAB2
    children = (2 * coords[:, None, :] + tab[None, :, :]).reshape(-1, dim)
    cov = brute_cover(dim, L, pos, rad, children, length * 0.5, r)
    # the reference prefilters spheres whose centre is within r + rmin of the
    # parent centre; a child can only be covered by such a sphere (its corners
    # lie within half a parent diagonal < r + rmin of the parent centre), so
    # the stable append is the uncovered children in (parent, TabCorner) order
    expect = children[~cov]
    assert np.array_equal(out, expect)


def test_remove_covered_matches_literal_loop():
    rng = np.random.default_rng(5)
    L, r = 0.3, 0.03
    pos = rng.random((40, 3)) * L
    rad = np.full(40, r)
    nv = 1 + int(math.sqrt(3) * L / (2 * r))
    length = np.full(3, L / nv)
    coords = np.array(np.meshgrid(*[np.arange(nv)] * 3, indexing="ij")).reshape(3, -1).T
    coords = coords[rng.permutation(len(coords))]
    flags = O.voxel_step(3, [0] * 3, [L] * 3, r, pos, rad, coords, length, r, mode=3)
    out, _ = O.voxel_step(3, [0] * 3, [L] * 3, r, pos, rad, coords, length, r, mode=0)
    assert np.array_equal(out, literal_remove_covered(coords, flags))


# ---------------------------------------------------------------- whole runs

def test_2d_saturation_properties():
    """2D r=0.01, unit square: maximal, no overlap, coverage ~0.547."""
    o = O.run(2, [0, 0], [1, 1], [(0.01, 1.0, 0)], seed=0)
    assert o["msg"].endswith("fully packed")
    phi = volume_fraction(o["rad"], 2, 1.0)
    assert 0.54 < phi < 0.555
    assert periodic_overlaps(o["pos"], o["rad"], [1, 1]) == 0
    assert uncovered_samples(o["pos"], o["rad"], [1, 1], 0.01, n=100000) == 0


def test_small_3d_saturation_properties():
    o = O.run(3, [0] * 3, [0.2] * 3, [(0.005, 1.0, 0)], seed=0)
    assert o["msg"].endswith("fully packed")
    phi = volume_fraction(o["rad"], 3, 0.2 ** 3)
    assert 0.37 < phi < 0.40
    assert periodic_overlaps(o["pos"], o["rad"], [0.2] * 3) == 0
    assert uncovered_samples(o["pos"], o["rad"], [0.2] * 3, 0.005, n=100000) == 0


def images(p, L, rad_dom, dim):
    """A centre and its periodic images as the reference's ghost areas build
    them: shifted by +-L along the dimensions where it lies in the band of width
    2*rad at that face (rsa_domain.ixx:117-133, rsa_ghost_area.ixx:22-30)."""
    out = [p]
    for i in range(3 ** dim):
        rel = [(i // 3 ** d) % 3 - 1 for d in range(dim)]
        if all(v == 0 for v in rel):
            continue
        ok = True
        q = p.copy()
        for d in range(dim):
            if rel[d] == -1:
                ok &= (p[d] >= 0.0) and (p[d] <= 0.0 + 2 * rad_dom)
                q[d] = p[d] + L
            elif rel[d] == 1:
                ok &= (p[d] >= L - 2 * rad_dom) and (p[d] <= L)
                q[d] = p[d] - L
        if ok:
            out.append(q)
    return out


def hits(a, ra, bs, rb):
    for b in bs:
        d = a - b
        if float(np.sum(d * d)) < (ra + rb) * (ra + rb):
            return True
    return False


@pytest.mark.parametrize("dim,L,r,seed", [(2, 0.25, 0.01, 0), (3, 0.12, 0.01, 1)])
def test_draws_equal_sequential_greedy_rsa(dim, L, r, seed):
    """Each non-outreach draw validates exactly the greedy-by-priority
    independent set (ties never block, strict '<' at
    operator_cells_and_data_storage.ixx:39) of the candidates that miss every
    placed sphere -- the sequential RSA the reference's conflict rounds
    emulate (operator_algorithm.ixx:383-473)."""
    o = O.run(dim, [0] * dim, [L] * dim, [(r, 1.0, 0)], seed=seed, dump=12)
    placed = []  # (pos, rad) of validated spheres
    for dd in o["dumps"]:
        assert dd["outreach"] == 0
        cand = list(range(len(dd["rad"])))
        alive = [i for i in cand
                 if not any(hits(dd["pos"][i], dd["rad"][i], images(q, L, r, dim), rq) for q, rq in placed)]
        acc = np.zeros(len(cand), dtype=bool)
        for i in sorted(alive, key=lambda k: int(dd["prio"][k])):
            blocked = False
            for j in alive:
                if acc[j] and int(dd["prio"][j]) < int(dd["prio"][i]):
                    if hits(dd["pos"][i], dd["rad"][i], images(dd["pos"][j], L, r, dim), dd["rad"][j]):
                        blocked = True
                        break
            acc[i] = not blocked
        assert np.array_equal(acc, dd["accepted"])
        placed += [(dd["pos"][i], dd["rad"][i]) for i in np.nonzero(acc)[0]]


def test_class_counts_from_volume_fractions():
    """compute_desired_radius_nb_phase: floor(V*phi / (v_D r^D)) (radius_generator.ixx:71-86);
    a 2-class run stops at the second class target."""
    o = O.run(3, [0] * 3, [0.3] * 3, [(0.03, 0.05, 0), (0.02, 0.05, 1)], seed=2)
    n0 = int(0.3 ** 3 * 0.05 / (4.0 / 3.0 * math.pi * 0.03 ** 3))
    n1 = int(0.3 ** 3 * 0.05 / (4.0 / 3.0 * math.pi * 0.02 ** 3))
    ph = o["phase"]
    assert int((ph == 0).sum()) == n0
    assert int((ph == 1).sum()) == n1
    assert o["msg"] == "End of generation, due to lack of new radii"


# ---------------------------------------------------------------- a second restatement of the rounds

def replay_in_python(dim, L, classes, seed, drad=None):
    """Run the oracle with every draw recorded, replay the draws through tests/_rsa_python.py
    (cell lists, swap-with-last removals, images, first-partner labels, outreach retention)
    and return (oracle result, per-draw accepted flags of both, Python extraction order)."""
    import _rsa_python as P
    rmax = max(c[0] for c in classes)
    drad = rmax if drad is None else drad
    o = O.run(dim, [0.0] * dim, [L] * dim, classes, seed=seed, mode=0, drad=drad, dump=1 << 30)
    g0 = P.Grid(dim, [0.0] * dim, [L] * dim, rmax, [0] * 3 ** dim)
    grid = P.Grid(dim, [0.0] * dim, [L] * dim, rmax, O.neighbour_table(dim, g0.n))
    sim = P.Sim(grid, drad)
    counts = [int(c[1]) for c in sorted(classes, key=lambda c: -c[0])]
    cur = 0
    pairs = []
    n_outreach = 0
    for dd in o["dumps"]:
        maximal = counts[cur]
        may = len(dd["rad"]) >= maximal
        assert bool(dd["outreach"]) == may
        n_outreach += may
        acc = sim.draw(dd["pos"], dd["rad"], dd["prio"], maximal, may)
        pairs.append((acc, dd["accepted"]))
        counts[cur] -= int(acc.sum())
        if counts[cur] == 0:
            cur += 1
    pos, rad = sim.extraction()
    return o, pairs, pos, rad, n_outreach


# every class target but the last is reachable (an unreachable one makes the reference
# refine forever, docs/source/algorithm.rst:57-62); the last class fills the box
@pytest.mark.parametrize("dim,L,classes,seed", [
    (2, 0.3, [(0.02, 20, 0), (0.012, 50, 1), (0.008, 10 ** 9, 2)], 0),
    (2, 0.4, [(0.02, 40, 0), (0.012, 100, 1), (0.008, 10 ** 9, 2)], 4),
    (2, 0.25, [(0.015, 40, 0), (0.01, 10 ** 9, 1)], 3),
    (3, 0.15, [(0.012, 60, 0), (0.008, 150, 1), (0.006, 10 ** 9, 2)], 1),
])
def test_rounds_and_cell_order_match_a_second_restatement(dim, L, classes, seed):
    """Every draw of a multi-class run, outreach draws included (the first-partner labels
    and the retained smallest priorities), accepts the same candidates in the Python
    restatement as in the oracle, and the final extraction order -- real cells by index,
    in-cell order left by the swap-with-last removals -- is the oracle's bit for bit."""
    o, pairs, pos, rad, n_outreach = replay_in_python(dim, L, classes, seed)
    assert o["msg"].endswith("fully packed")
    assert n_outreach >= 2
    for k, (a, b) in enumerate(pairs):
        assert np.array_equal(a, b), k
    assert len(rad) == o["n"]
    assert np.array_equal(pos.view(np.uint64), o["pos"].view(np.uint64))
    assert np.array_equal(rad.view(np.uint64), o["rad"].view(np.uint64))


def test_pivot_search_matches_a_second_transcription():
    """The oracle's find_pivot_for_first_elements against the Python transcription of
    median_of_medians.ixx:8-178 on random priority sets (distinct values, heavy ties,
    values near 0 and near 2^31), every N first from 1 to the size: the same pivot bit for
    bit, or the same error (the reference throws "Unexpected - 1" when its first candidate
    already holds too many values and there is no lower one, which ties can cause)."""
    import _rsa_python as P
    rng = np.random.default_rng(7)
    checked = errors = 0
    for trial in range(400):
        n = int(rng.integers(1, 40))
        kind = trial % 4
        if kind == 0:
            v = rng.integers(0, 2 ** 31, size=n)
        elif kind == 1:
            v = rng.integers(0, 4, size=n)  # ties everywhere
        elif kind == 2:
            v = rng.integers(0, 50, size=n)
        else:
            v = 2 ** 31 - 1 - rng.integers(0, 1000, size=n)
        for nb in range(1, n + 1):
            def outcome(f):
                try:
                    return "value", np.float64(f()).view(np.uint64)
                except RuntimeError as e:  # the reference's "Unexpected - 1" (ties at both candidates)
                    return "error", str(e)
            a = outcome(lambda: O.mom_pivot(v, nb))
            b = outcome(lambda: P.reference_pivot(v.tolist(), nb))
            assert a == b, (v.tolist(), nb, a, b)
            errors += a[0] == "error"
            checked += 1
    for bad in (0, 5):
        with pytest.raises(RuntimeError, match="Impossible to find the pivot"):
            O.mom_pivot([1, 2, 3, 4], bad)
    assert checked > 5000 and 0 < errors < checked // 4


def test_pivot_search_keeps_the_smallest_when_it_reaches_the_count():
    """What the pivot selects, stated directly: the values not above it are a prefix of
    the sorted values, and whenever that prefix has N values it is the N smallest.  The
    reference's dichotomy steps away from the count when too few values lie below
    (median_of_medians.ixx:103-118), so it can keep fewer: for these four priorities and
    N = 2 it keeps none (the 2D three-class case below meets exactly this draw)."""
    import _rsa_python as P
    rng = np.random.default_rng(3)
    exact = fewer = 0
    for _ in range(3000):
        n = int(rng.integers(2, 30))
        v = np.unique(rng.integers(0, 2 ** 31, size=n))
        nb = int(rng.integers(1, len(v) + 1))
        p = P.reference_pivot(v.tolist(), nb)
        kept = v[~(v.astype(np.float64) > p)]
        assert np.array_equal(kept, np.sort(v)[:len(kept)])  # a prefix
        exact += len(kept) == nb
        fewer += len(kept) < nb
    assert exact > 1000 and fewer > 0
    vals = [1907355172, 1641811781, 1812586757, 934655398]
    p = O.mom_pivot(vals, 2)
    assert p < min(vals)  # no priority below the pivot: the draw validates none


def test_outreach_retention_defect_in_a_whole_run():
    """2D three-class run: one outreach draw has 4 Undecided candidates for 2 places and
    validates none (the pivot case above); the class fills one draw later."""
    o, pairs, pos, rad, n_outreach = replay_in_python(2, 0.3, [(0.02, 20, 0), (0.012, 50, 1), (0.008, 10 ** 9, 2)], 0)
    tr = o["trace"]
    short = [k for k, (a, b) in enumerate(pairs) if b.sum() == 0 and len(b) > 50]
    assert 14 in short
    assert int(tr[14, 3]) == 0 and int(tr[15, 3]) == 2


@pytest.mark.parametrize("dim,L,classes", [(2, 1.0, [(0.01, 1.0, 0)]), (3, 0.37, [(0.02, 1.0, 0), (0.013, 1.0, 1)]),
                                           (4, 0.5, [(0.03, 1.0, 0)])])
def test_initial_voxel_lattice(dim, L, classes):
    """list_of_voxels(origin, width, r_max + r_min) (list_of_voxels.ixx:11-28, 192-211):
    n_d = 1 + floor(sqrt(D) * W_d / (r_max + r_min)) voxels of length W_d / n_d per
    dimension, enumerated with the first coordinate fastest (Loops.ixx:44-74)."""
    o = O.run(dim, [0.0] * dim, [L] * dim, classes, seed=0, max_draws=0)
    rmax = max(c[0] for c in classes)
    rmin = min(c[0] for c in classes)
    n = 1 + int(math.sqrt(dim) * L / (rmax + rmin))
    grid = np.stack(np.meshgrid(*[np.arange(n)] * dim, indexing="ij"), axis=-1)
    want = grid.transpose(list(range(dim))[::-1] + [dim]).reshape(-1, dim)  # coordinate 0 fastest
    assert np.array_equal(o["vox"], want)
    assert np.array_equal(o["vox_len"], np.full(dim, L / n))
