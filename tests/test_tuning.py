"""The autotuner: its cache (keys, round trip, damaged files) on the CPU, and on
the GPU the refinement thresholds set through the C-ABI, shown to leave every
placed set unchanged at their extremes, tuned, stored and picked up again."""
import json
import os

import numpy as np
import pytest

import rsa_mpi_amd.tuning as T


# ---------------------------------------------------------------- cache (no GPU)

def test_ratio_bins_and_keys():
    assert T.ratio_bin(1.0, 0.002) == 9.0  # log2(500) = 8.97
    assert T.ratio_bin(1.0, 0.005) == 7.5  # log2(200) = 7.64
    assert T.ratio_bin(0.2, 0.004) == 5.5  # log2(50) = 5.64
    assert T.key(3, 1.0, 0.002) == "context:3d:9"
    assert T.key(2, 0.2, 0.004, "batch") == "batch:2d:5.5"
    with pytest.raises(ValueError):
        T.ratio_bin(0.0, 1.0)


def test_store_load_lookup_round_trip(tmp_path):
    p = str(tmp_path / "t.json")
    assert T.lookup(3, 1.0, 0.002, path=p) is None
    T.store(T.key(3, 1.0, 0.002), {"group_min": 1 << 19, "pack_min": 1 << 21}, path=p)
    T.store(T.key(3, 1.0, 0.005), {"group_min": 7, "pack_min": 9}, path=p)
    assert T.lookup(3, 1.0, 0.002, path=p) == {"group_min": 1 << 19, "pack_min": 1 << 21}
    assert T.lookup(3, 1.01, 0.002, path=p) == {"group_min": 1 << 19, "pack_min": 1 << 21}  # same bin
    assert T.lookup(3, 1.0, 0.005, path=p) == {"group_min": 7, "pack_min": 9}
    assert T.lookup(2, 1.0, 0.002, path=p) is None
    sides, radii = np.array([0.2, 0.5, 1.0]), np.array([0.004, 0.005, 0.004])  # median ratio 100
    T.store(T.key(3, 100.0, 1.0, "batch"), {"workers": 3}, path=p)
    assert T.lookup_batch(3, sides, radii, path=p) == 3
    assert T.lookup_batch(2, sides, radii, path=p) is None


def test_damaged_or_foreign_cache_is_ignored(tmp_path):
    p = tmp_path / "t.json"
    p.write_text("{not json")
    assert T.lookup(3, 1.0, 0.002, path=str(p)) is None
    p.write_text(json.dumps({"version": 999, "entries": {T.key(3, 1.0, 0.002): {"group_min": 1, "pack_min": 1}}}))
    assert T.lookup(3, 1.0, 0.002, path=str(p)) is None
    # a store over a damaged file starts a new one
    T.store(T.key(3, 1.0, 0.002), {"group_min": 5, "pack_min": 6}, path=str(p))
    assert T.lookup(3, 1.0, 0.002, path=str(p)) == {"group_min": 5, "pack_min": 6}


def test_cache_can_be_switched_off(monkeypatch):
    monkeypatch.setenv("RSA_MPI_AMD_TUNING", "off")
    assert T.cache_path() is None
    assert T.lookup(3, 1.0, 0.002) is None
    assert T.store("x", {}) is None


# ---------------------------------------------------------------- GPU

def _golden(name):
    return json.load(open(os.path.join(os.path.dirname(__file__), "golden", name + ".json")))


def _run(dim, L, r, tuning, seed=0):
    import rsa_mpi_amd as R
    from _common import set_hash
    dom = R.Domain([0.0] * dim, [L] * dim, r, tuning=tuning)
    gen = R.RadiusGenerator([(r, 1.0, 0)], dom.get_total_volume(), 0.0)
    st = R.Algo(dom, gen, 6000).proceed(seed, -1)
    pos, rad, ph = dom.extract_arrays()
    return st, set_hash(pos, rad), pos, rad, ph, dom.tuning()


@pytest.mark.gpu
def test_gpu_tuning_set_get_and_rejected_values(monkeypatch):
    import rsa_mpi_amd as R
    monkeypatch.setenv("RSA_MPI_AMD_TUNING", "off")
    dom = R.Domain([0.0] * 3, [0.2] * 3, 0.01)
    assert dom.tuning() == {"group_min": 1 << 20, "pack_min": 1 << 20}
    dom.set_tuning(group_min=123, pack_min=456)
    assert dom.tuning() == {"group_min": 123, "pack_min": 456}
    for bad in ({"group_min": 0}, {"pack_min": -5}):
        with pytest.raises(RuntimeError):
            dom.set_tuning(**bad)
    assert dom.tuning() == {"group_min": 123, "pack_min": 456}


@pytest.mark.gpu
@pytest.mark.parametrize("gm,pm", [(1, 1), (T.NEVER, T.NEVER), (1, T.NEVER), (T.NEVER, 1)])
def test_gpu_thresholds_at_their_extremes_leave_the_set_unchanged(monkeypatch, gm, pm):
    """3D r=0.005 in the unit cube against the oracle's fixture: every list walked in
    grouped order with the packed copy (1, 1), none of them (never), and the mixed cases."""
    monkeypatch.setenv("RSA_MPI_AMD_TUNING", "off")
    gold = _golden("3d_r0.005_seed0")
    st, h, *_ , tun = _run(3, 1.0, 0.005, {"group_min": gm, "pack_min": pm})
    assert tun == {"group_min": gm, "pack_min": pm}
    assert h == gold["sha256"] and st["draws"] == gold["draws"] and st["words"] == gold["words"]
    assert st["overlap_pairs"] == 0 and st["cover_free"] == 0


@pytest.mark.gpu
@pytest.mark.parametrize("dim,L,r", [(2, 1.0, 0.01), (4, 0.3, 0.03)])
def test_gpu_grouped_order_on_every_list_matches_the_oracle(monkeypatch, dim, L, r):
    """2D and 4D (no packed copy in 4D) with every list grouped, against the oracle."""
    import _oracle as O
    monkeypatch.setenv("RSA_MPI_AMD_TUNING", "off")
    st, h, pos, rad, ph, _ = _run(dim, L, r, {"group_min": 1, "pack_min": 1}, seed=3)
    o = O.run(dim, [0.0] * dim, [L] * dim, [(r, 1.0, 0)], seed=3)
    assert len(rad) == o["n"]
    assert np.array_equal(pos.view(np.uint64), o["pos"].view(np.uint64))
    assert np.array_equal(rad.view(np.uint64), o["rad"].view(np.uint64))


@pytest.mark.gpu
def test_gpu_autotune_stores_and_domain_applies(tmp_path, monkeypatch):
    import rsa_mpi_amd as R
    p = str(tmp_path / "tuning.json")
    monkeypatch.setenv("RSA_MPI_AMD_TUNING", p)
    grid = ((1 << 16, 1 << 16), (1 << 18, T.NEVER), (T.NEVER, T.NEVER))
    res = R.autotune(3, 0.4, 0.003, seed=1, grid=grid)
    print(res)
    assert tuple(res["best"].values()) in grid and res["path"] == p
    assert set(res["seconds"]) == set(grid)
    assert T.lookup(3, 0.4, 0.003) == res["best"]
    dom = R.Domain([0.0] * 3, [0.4] * 3, 0.003)  # tuning="auto"
    assert dom.tuning() == res["best"]
    assert R.Domain([0.0] * 3, [0.4] * 3, 0.003, tuning=None).tuning() == {"group_min": 1 << 20,
                                                                              "pack_min": 1 << 20}


@pytest.mark.gpu
def test_gpu_autotune_batch_stores_the_worker_count(tmp_path, monkeypatch):
    import rsa_mpi_amd as R
    p = str(tmp_path / "tuning.json")
    monkeypatch.setenv("RSA_MPI_AMD_TUNING", p)
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "batch_3d_256.json")))
    real = fx["realizations"][:8]
    sides = [c["L"] for c in real]
    radii = [c["r"] for c in real]
    seeds = [c["seed"] for c in real]
    res = R.autotune_batch(3, sides, radii, seeds, workers=(1, 2))
    print(res)
    assert res["best"] in (1, 2) and T.lookup_batch(3, sides, radii) == res["best"]
    out = R.batch_run(3, sides, radii, seeds)  # workers from the cache
    assert out["counts"].tolist() == [fx["oracle"][k]["n"] for k in range(8)]
