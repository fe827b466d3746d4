"""The learned predictor: its CPU reference pinned by finite differences, the
GPU forward/backward kernels (BF16 on the matrix cores, FP32 accumulation)
checked against it, training on traces the method logs, and the learned mode
shown to leave every placed set unchanged while it reports its misprediction
rates and what its ordering gains."""
import numpy as np
import pytest

import _mlp_ref as M


# ---------------------------------------------------------------- CPU reference (no GPU)

def test_reference_gradient_matches_finite_differences():
    """The float64 reference's analytic gradient against central differences of its loss,
    on every parameter block (a dropped term or a transposed product fails here)."""
    p, x, t = M.random_case(48, 0)
    p = p.astype(np.float64)
    _, g = M.loss_and_grad(p, x, t)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(np.arange(a, b), 12, replace=False) for a, b in
                          [(M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                           (M.OFF_W3, M.OFF_B3)]] + [np.arange(M.OFF_B3, M.NPAR)])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        num = (M.loss_and_grad(p + e, x, t)[0] - M.loss_and_grad(p - e, x, t)[0]) / (2 * h)
        assert abs(num - g[i]) < 1e-6 + 1e-4 * abs(g[i]), (i, num, g[i])


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -3.5, 0.0],
                 dtype=np.float32)
    assert M.bf16(v).tolist() == [1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -3.5, 0.0]


# ---------------------------------------------------------------- GPU kernels

@pytest.mark.gpu
def test_gpu_forward_matches_reference():
    import rsa_mpi_amd as R
    p, x, t = M.random_case(256, 2)
    z = R.learn_forward(x, p)
    _, _, _, zr = M.forward(p, x, bf=True)
    # FP32 accumulation in another order than float64: 1e-3 relative to the logits' scale
    assert np.max(np.abs(z - zr)) < 1e-3 * max(1.0, np.max(np.abs(zr)))


@pytest.mark.gpu
def test_gpu_gradients_match_reference():
    """Gradient check of the GPU backward kernel: against the reference that rounds to BF16
    where the kernel does, each parameter block within 1e-2 relative (norm), the loss within
    1e-3; against the exact float64 model within 1e-1 -- BF16 keeps 8 significant bits, and
    the rounding of the inputs, weights, activations and backpropagated signals adds up to
    about 5 % on the first layer's gradient."""
    import rsa_mpi_amd as R
    for n, seed in ((256, 3), (1024, 4)):
        p, x, t = M.random_case(n, seed)
        g, loss = R.learn_gradcheck(x, t, p)
        lr, gr = M.loss_and_grad(p, x, t, bf=True)
        le, ge = M.loss_and_grad(p, x, t, bf=False)
        assert abs(loss - lr) < 1e-3 * abs(lr), (loss, lr)
        for a, b in [(M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                     (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.NPAR)]:
            d = np.linalg.norm(g[a:b] - gr[a:b]) / np.linalg.norm(gr[a:b])
            assert d < 1e-2, (a, d)
            de = np.linalg.norm(g[a:b] - ge[a:b]) / np.linalg.norm(ge[a:b])
            assert de < 1e-1, (a, de)


# ---------------------------------------------------------------- training on traces, learned mode

def _golden(name):
    import json
    import os
    return json.load(open(os.path.join(os.path.dirname(__file__), "golden", name + ".json")))


def _run(gold, trace=False, use=False, measure=False):
    import rsa_mpi_amd as R
    from _common import set_hash
    dom = R.Domain([0.0] * 3, [1.0] * 3, 0.005)
    dom.learn_mode(trace=trace, use=use, measure=measure)
    gen = R.RadiusGenerator([(0.005, 1.0, 0)], dom.get_total_volume(), 0.0)
    st = R.Algo(dom, gen, 6000).proceed(0, -1)
    pos, rad, _ = dom.extract_arrays()
    ls = dom.learn_stats()
    print(st, ls)
    assert len(rad) == gold["n"] and set_hash(pos, rad) == gold["sha256"]
    assert st["draws"] == gold["draws"] and st["refines"] == gold["refines"] and st["words"] == gold["words"]
    assert st["overlap_pairs"] == 0 and st["cover_free"] == 0
    return st, ls


@pytest.mark.gpu
def test_gpu_trace_train_and_learned_mode_leave_the_set_unchanged():
    """3D r=0.005 in the unit cube (the oracle's fixture): a traced run logs voxel and draw
    samples, training on them lowers the loss, and runs that use the trained model (to order
    the coverage walk and to hold back the speculative draw) place the same set bit for bit,
    with the same draws, refinements and words, while reporting their misprediction counts."""
    import rsa_mpi_amd as R
    gold = _golden("3d_r0.005_seed0")
    R.learn_reset(5)
    _, ls = _run(gold, trace=True)
    assert ls["traced_voxels"] > 0 and ls["traced_draws"] > 0
    assert R.learn_samples() == ls["traced_voxels"] + ls["traced_draws"]
    first, last, n = R.learn_train(steps=300, batch=1024, lr=3e-3, seed=1)
    print("loss", first, last, n)
    assert n == R.learn_samples() and np.isfinite(last) and last < 0.8 * first
    _, plain = _run(gold, measure=True)
    assert plain["voxels_predicted"] == 0 and plain["waves"] > 0
    _, used = _run(gold, use=True)
    # the same passes, the same outcomes: only the coverage walk's order differs
    assert used["voxels_covered"] == plain["voxels_covered"] and used["waves"] == plain["waves"]
    assert 0 < used["voxels_predicted"] and used["voxels_covered_predicted"] <= used["voxels_predicted"]
    # the model beats the constant predictor "covered" on the passes it predicted
    constant_misses = used["voxels_predicted"] - used["voxels_covered_predicted"]
    assert used["voxels_mispredicted"] < constant_misses
    assert used["draws_predicted"] == gold["draws"] and used["refines"] == gold["refines"]
    assert used["draws_mispredicted"] <= used["draws_predicted"]
    assert used["speculation_skipped"] <= used["refines_predicted"]  # fast-path draws only
    print("voxel misprediction %.4f (constant predictor %.4f); draw misprediction %.4f; "
          "uniform waves plain %.4f learned %.4f" % (
              used["voxels_mispredicted"] / used["voxels_predicted"], constant_misses / used["voxels_predicted"],
              used["draws_mispredicted"] / used["draws_predicted"], plain["waves_uniform"] / plain["waves"],
              used["waves_uniform"] / used["waves"]))
