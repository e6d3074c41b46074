"""CPU checks of the surrogate's oracle and host logic.

oracle/nn_ref.py (the float64 restatement of the network the HIP kernels implement) is pinned
by central finite differences on every parameter block; the mask rasterizer is checked against
the solver library's own polygon predicate."""
import ctypes

import numpy as np
import pytest

from oracle import nn_ref
import vortexlab_amd as vl
from vortexlab_amd import api, surrogate, workloads


def test_parameter_count_matches_library():
    _, total = nn_ref.layout()
    lib = vl.backend().lib
    lib.vl_nn_param_count.restype = ctypes.c_int64
    assert lib.vl_nn_param_count() == total


def test_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    off, total = nn_ref.layout()
    p = rng.normal(0.0, 0.08, total)
    masks = rng.uniform(0.0, 1.0, (2, 64, 64))  # continuous inputs: no ties in the max-pools
    s = rng.normal(size=2)
    t = rng.normal(size=(2, 3))
    loss, g = nn_ref.loss_and_grad(p, masks, s, t)
    h = 1e-6
    for name, (o, shape) in off.items():
        n = int(np.prod(shape))
        for k in rng.choice(n, size=min(3, n), replace=False):
            i = o + int(k)
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (nn_ref.loss_and_grad(pp, masks, s, t)[0] - nn_ref.loss_and_grad(pm, masks, s, t)[0]) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, f"{name}[{k}]: fd {fd} vs {g[i]}"


def test_rasterizer_matches_polygon_predicate():
    rng = np.random.default_rng(4)
    be = vl.backend()
    for _ in range(3):
        d = 16.0
        outline = workloads.star_polygon(rng, d, 100.5, 60.0)
        m = surrogate.rasterize(outline, 100.5, 60.0, d)
        h = api.polygon(outline).make(be, [])
        c = (np.arange(64) + 0.5) / 64 * 1.2 * d - 0.6 * d
        for iy in range(0, 64, 3):
            for ix in range(0, 64, 3):
                assert m[iy, ix] == be.shape_inside(h, 100.5 + c[ix], 60.0 + c[iy])
        be.shape_free(h)
        assert 0.15 < m.mean() < 0.6


def test_acceptance_rule():
    mean = np.array([[1.5, 0.3, 0.2], [1.5, 0.3, 0.2]])
    std = np.array([[0.01, 0.003, 0.002], [0.05, 0.003, 0.002]])
    acc = surrogate.accept(mean, std, np.array([0.5, 0.2, 0.05]), 0.02)
    assert acc.tolist() == [True, False]


def test_calibrated_threshold_meets_the_stated_error():
    """calibrate_threshold returns the largest spread threshold whose accepted answers meet the
    target 90th-percentile error: answers sorted by spread whose errors grow with it."""
    rng = np.random.default_rng(9)
    spread = rng.uniform(0.0, 0.1, size=400)
    err = spread * rng.uniform(0.5, 1.5, size=400)  # errors grow with the spread
    th = surrogate.calibrate_threshold(spread, err, 0.05)
    acc = spread <= th
    assert np.quantile(err[acc], 0.9) <= 0.05
    # one more answer (the next larger spread) would break the target
    nxt = np.sort(spread[~acc])[0]
    assert np.quantile(err[spread <= nxt], 0.9) > 0.05
    # nothing meets an impossible target; everything meets a lax one
    assert surrogate.calibrate_threshold(spread, err + 1.0, 0.05) == 0.0
    assert surrogate.calibrate_threshold(spread, err, 1.0) == spread.max()
    rep = surrogate.fallback_report(spread, err, th)
    assert rep["accepted_fraction"] == acc.mean() and rep["p90_rel_err_all_queries"] <= rep["p90_rel_err_surrogate_alone"]
