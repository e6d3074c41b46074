"""Surrogate network on the GPU: the HIP forward and backward kernels against the float64
oracle (oracle/nn_ref.py), bitwise-reproducible training, and an end-to-end run that shows
what the trained ensemble gains over the method without it (the solver alone)."""
import numpy as np
import pytest

from oracle import nn_ref
import vortexlab_amd as vl
from vortexlab_amd import surrogate

pytestmark = pytest.mark.gpu


def test_forward_and_gradients_match_oracle(gpu):
    rng = np.random.default_rng(1)
    net = surrogate.Net(gpu, seed=3, max_batch=8)
    p = net.params().astype(np.float64)
    masks = (rng.uniform(size=(4, 64, 64)) > 0.5).astype(np.float32)
    s = rng.normal(size=4).astype(np.float32)
    t = rng.normal(size=(4, 3)).astype(np.float32)
    y_ref, _ = nn_ref.forward(p, masks.astype(np.float64), s.astype(np.float64))
    y = net.forward(masks, s)
    assert np.max(np.abs(y - y_ref)) <= 1e-4 * max(1.0, np.max(np.abs(y_ref)))
    loss_ref, g_ref = nn_ref.loss_and_grad(p, masks.astype(np.float64), s.astype(np.float64), t.astype(np.float64))
    loss, g = net.gradients(masks, s, t)
    assert abs(loss - loss_ref) <= 1e-4 * abs(loss_ref)
    off, _ = nn_ref.layout()
    for name, (o, shape) in off.items():
        n = int(np.prod(shape))
        ref, got = g_ref[o:o + n], g[o:o + n]
        scale = max(np.max(np.abs(ref)), 1e-12)
        assert np.max(np.abs(got - ref)) <= 2e-4 * scale, f"{name}: max err {np.max(np.abs(got - ref)):.3e} / {scale:.3e}"
    net.close()


def test_training_is_reproducible(gpu):
    rng = np.random.default_rng(2)
    masks = (rng.uniform(size=(96, 64, 64)) > 0.5).astype(np.float32)
    feat = rng.normal(size=96).astype(np.float32)
    y = rng.normal(size=(96, 3)).astype(np.float32)
    params = []
    for _ in range(2):
        net = surrogate.Net(gpu, seed=11, max_batch=32)
        net.set_dataset(masks, feat, y)
        for ep in range(3):
            net.train_epoch(np.random.default_rng(ep).permutation(96), 32, 1e-3)
        params.append(net.params())
        net.close()
    assert np.array_equal(params[0], params[1])


def test_surrogate_gains_over_solver_alone(gpu):
    """Small end-to-end run: solver-labelled training set from the batched path, an 8-member
    ensemble, held-out shapes.  The ensemble must predict the drag better than the training
    mean does; the fallback threshold on its spread, calibrated on half the held-out shapes for
    a 90th-percentile relative Cd error of 5 % (the full-size report's target), must avoid a share
    of the solves on the other half and hold the stated error there; and an answer must cost less
    than a
    batched solve of the same case.  At this test's 10 cells per diameter the cases are tiny
    (160 x 60 nodes) and a batched solve costs well under a millisecond; at the production
    resolution the gap is two orders of magnitude (the full-size surrogate report).

    At this size the spread ranks the answers by their error only weakly (rank correlation
    0.1-0.2), so the accepted share at a tighter target swings with the training's rounding: at
    4 % it ranged from 0 to 64 % over ensemble seeds and sizes.  At 5 % it was 56-100 % over
    the eight (seed, size) pairs tried."""
    # 10 cells per diameter is resolved up to Re ~ 90 (the solver itself diverges near 117)
    kw = dict(d=10.0, units=24.0, re=(30.0, 90.0), upstream=4.0, downstream=12.0, height=6.0)
    tr = surrogate.make_dataset(gpu, 21, 1536, **kw)
    te = surrogate.make_dataset(gpu, 22, 384, **kw)
    assert tr.ok.mean() > 0.8 and te.ok.mean() > 0.8
    ens = surrogate.Ensemble(gpu, 8, seed=3, max_batch=64)
    ens.fit(tr.masks[tr.ok], tr.re[tr.ok], tr.y[tr.ok], epochs=30, batch=32, lr=2e-3)
    import time
    t0 = time.perf_counter()
    mean, std = ens.predict(te.masks[te.ok], te.re[te.ok])
    t_pred = time.perf_counter() - t0
    truth = te.y[te.ok]
    err_net = np.sqrt(np.mean((mean[:, 0] - truth[:, 0]) ** 2))
    err_mean = np.sqrt(np.mean((ens.y_mu[0] - truth[:, 0]) ** 2))
    assert err_net < 0.6 * err_mean
    rel = np.abs(mean[:, 0] - truth[:, 0]) / np.abs(truth[:, 0])
    spread = surrogate.spread_of(mean, std, ens.y_sd)
    half = len(truth) // 2
    target = 0.05
    th = surrogate.calibrate_threshold(spread[:half], rel[:half], target)
    cal = surrogate.fallback_report(spread[:half], rel[:half], th)
    ev = surrogate.fallback_report(spread[half:], rel[half:], th)
    per_answer, per_solve = t_pred / len(truth), te.solve_seconds / len(te.ok)
    print(f"[surrogate] Cd rmse {err_net:.4f} vs mean predictor {err_mean:.4f}; threshold {th:.4f}: calibration "
          f"{cal}; evaluation {ev}; {per_answer * 1e6:.0f} us per answer, {per_solve * 1e6:.0f} us per batched solve",
          flush=True)
    assert cal["p90_rel_err_accepted"] is not None and cal["p90_rel_err_accepted"] <= target
    # the rule avoids a share of the solves on the other half...
    assert ev["accepted_fraction"] >= 0.15
    # ...and holds its stated error there: over all queries (the rejected ones are solved) and,
    # with room for 190 answers' sampling noise, over the accepted answers alone
    assert ev["p90_rel_err_all_queries"] <= target
    assert ev["p90_rel_err_accepted"] <= 1.5 * target
    assert ev["p90_rel_err_all_queries"] <= ev["p90_rel_err_surrogate_alone"]
    assert per_answer < per_solve
    ens.close()
