"""The batched form: 256 independent 3D realizations (box edge L ~ U[0.2, 1],
radius r ~ U[0.004, 0.008], seeds 0..255; tests/golden/batch_3d_256.json) run
as a device work queue, many in flight on one stream.  Each realization is
uniform_generate<3, 1> on [0, L]^3 (operator_algorithm.ixx:275-292); its placed
set must be the oracle's, whatever the number in flight."""
import json
import os

import numpy as np
import pytest

from _common import set_hash

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "batch_3d_256.json")


def load():
    return json.load(open(FIX))


def test_batch_fixture_regenerates_from_meta_seed():
    """The realization list is numpy PCG64(20261203): L ~ U[0.2, 1.0], then r ~ U[0.004, 0.008]."""
    fx = load()
    g = np.random.Generator(np.random.PCG64(fx["meta_seed"]))
    L = g.uniform(0.2, 1.0, size=256)
    r = g.uniform(0.004, 0.008, size=256)
    assert [c["L"] for c in fx["realizations"]] == L.tolist()
    assert [c["r"] for c in fx["realizations"]] == r.tolist()
    assert [c["seed"] for c in fx["realizations"]] == list(range(256))
    assert [c["index"] for c in fx["oracle"]] == list(range(256))
    assert all(c["end"].endswith("fully packed") for c in fx["oracle"])


def test_batch_rejects_bad_arguments():
    import rsa_mpi_amd as R
    with pytest.raises(RuntimeError, match="dim must be"):
        R.batch_run(5, [1.0], [0.01], [0])


@pytest.mark.gpu
def test_batch256_counts_match_oracle():
    """All 256 realizations at full size, 4 in flight: counts equal the oracle's."""
    import rsa_mpi_amd as R
    fx = load()
    real = fx["realizations"]
    out = R.batch_run(3, [c["L"] for c in real], [c["r"] for c in real], [c["seed"] for c in real], workers=4)
    want = np.array([c["n"] for c in fx["oracle"]], dtype=np.uint64)
    bad = np.nonzero(out["counts"] != want)[0]
    assert len(bad) == 0, (bad[:10].tolist(), out["counts"][bad[:10]].tolist(), want[bad[:10]].tolist())


@pytest.mark.gpu
@pytest.mark.parametrize("workers", [1, 3, 8])
def test_batch256_sets_match_oracle(workers):
    """Every 8th realization extracted, `workers` in flight: the placed sets are the oracle's bit for bit."""
    import rsa_mpi_amd as R
    fx = load()
    idx = list(range(workers % 8, 256, 8))
    real = [fx["realizations"][k] for k in idx]
    out = R.batch_run(3, [c["L"] for c in real], [c["r"] for c in real], [c["seed"] for c in real],
                      workers=workers, extract=True)
    for k, (pos, rad) in zip(idx, out["spheres"]):
        o = fx["oracle"][k]
        assert len(rad) == o["n"], k
        assert set_hash(pos, rad) == o["sha256"], k


@pytest.mark.gpu
def test_batch256_subset0_one_worker_matches_oracle():
    """The realizations 0, 8, .., 248 (the 8-in-flight case's subset) run one at a time: sets equal the oracle's."""
    import rsa_mpi_amd as R
    fx = load()
    idx = list(range(0, 256, 8))
    real = [fx["realizations"][k] for k in idx]
    out = R.batch_run(3, [c["L"] for c in real], [c["r"] for c in real], [c["seed"] for c in real],
                      workers=1, extract=True)
    for k, (pos, rad) in zip(idx, out["spheres"]):
        o = fx["oracle"][k]
        assert len(rad) == o["n"], k
        assert set_hash(pos, rad) == o["sha256"], k


# ---------------------------------------------------------------- the device work queue (rsa_gpu_batch_queue)

def test_batch_queue_rejects_bad_arguments():
    import rsa_mpi_amd as R
    with pytest.raises(RuntimeError, match="dim must be"):
        R.batch_queue(5, [1.0], [0.01], [0])
    with pytest.raises(RuntimeError, match="active must be"):
        R.batch_queue(3, [1.0], [0.01], [0], active=0)


@pytest.mark.gpu
@pytest.mark.parametrize("dim,sides,radii,seeds,active", [
    (2, [0.3, 0.5, 0.21, 0.4, 0.35, 0.6, 0.25], [0.01, 0.012, 0.008, 0.015, 0.011, 0.02, 0.009], [3, 1, 4, 1, 5, 9, 2], 3),
    (3, [0.12, 0.15, 0.1, 0.2, 0.11], [0.01, 0.012, 0.009, 0.015, 0.008], [0, 7, 11, 2, 5], 5),
    (3, [0.12, 0.15, 0.1, 0.2, 0.11], [0.01, 0.012, 0.009, 0.015, 0.008], [0, 7, 11, 2, 5], 1),
    (4, [0.3, 0.25], [0.04, 0.035], [1, 2], 2),
])
def test_batch_queue_small_bit_exact(dim, sides, radii, seeds, active):
    """Realizations driven together (one launch per step for all of them) give the
    oracle's uniform_generate<dim, 1> result for each, in extraction order."""
    import rsa_mpi_amd as R
    import _oracle as O
    print(f"batch_queue dim={dim} n={len(sides)} active={active}", flush=True)
    out = R.batch_queue(dim, sides, radii, seeds, active=active, extract=True)
    assert 1 <= out["stats"]["max_active"] <= min(active, len(sides))
    for k, (pos, rad) in enumerate(out["spheres"]):
        o = O.run(dim, [0.0] * dim, [sides[k]] * dim, [(radii[k], 1.0, 0)], seed=seeds[k])
        assert len(rad) == o["n"] and o["n"] > 0, k
        assert np.array_equal(pos.view(np.uint64), o["pos"].view(np.uint64)), k
        assert np.array_equal(rad.view(np.uint64), o["rad"].view(np.uint64)), k


@pytest.mark.gpu
def test_batch_queue_list_redo_path_bit_exact(monkeypatch):
    """With the first-round list limit lowered to 1 the host-built lists (k_lists)
    and the redone first rounds run inside the queue's rounds."""
    import rsa_mpi_amd as R
    import _oracle as O
    monkeypatch.setenv("RSA_GPU_LIST_THRESHOLD", "1")
    sides, radii, seeds = [0.15, 0.12, 0.2], [0.01, 0.012, 0.014], [4, 8, 15]
    out = R.batch_queue(3, sides, radii, seeds, active=3, extract=True)
    for k, (pos, rad) in enumerate(out["spheres"]):
        o = O.run(3, [0.0] * 3, [sides[k]] * 3, [(radii[k], 1.0, 0)], seed=seeds[k])
        assert np.array_equal(pos.view(np.uint64), o["pos"].view(np.uint64)), k


@pytest.mark.gpu
def test_batch_queue_256_sets_match_oracle():
    """All 256 realizations of the batched workload, 64 in flight: every placed set's
    SHA-256 equals the oracle's."""
    import rsa_mpi_amd as R
    fx = load()
    real = fx["realizations"]
    out = R.batch_queue(3, [c["L"] for c in real], [c["r"] for c in real], [c["seed"] for c in real], active=64,
                        extract=True)
    bad = [k for k, (pos, rad) in enumerate(out["spheres"])
           if len(rad) != fx["oracle"][k]["n"] or set_hash(pos, rad) != fx["oracle"][k]["sha256"]]
    assert not bad, bad[:10]
    print("batch_queue 256:", out["seconds"], out["stats"], flush=True)


@pytest.mark.gpu
def test_batch_queue_host_driven_draws_match_single_contexts():
    """Draws of more than 48 000 shots (nbshots 60 000 on a box of 10 M initial voxels)
    take the host-driven form inside the queue (alone on its stream, its words copied
    under the same lock as the helper threads'): the sets equal the single-context
    runs of the same realizations, in extraction order."""
    import rsa_mpi_amd as R
    sides, radii, seeds = [1.0, 0.3], [0.004, 0.006], [5, 6]
    out = R.batch_queue(3, sides, radii, seeds, nbshots=60000, active=2, extract=True)
    for k, (pos, rad) in enumerate(out["spheres"]):
        dom = R.Domain([0.0] * 3, [sides[k]] * 3, radii[k])
        R.uniform_generate(dom, radii[k], 60000, 10, seeds[k])
        p1, r1, _ = dom.extract_arrays()
        assert len(rad) == len(r1) > 0, k
        assert np.array_equal(pos.view(np.uint64), p1.view(np.uint64)), k
        assert np.array_equal(rad.view(np.uint64), r1.view(np.uint64)), k
