"""CPU checks of the case drivers (vortexlab_amd/cases.py) on the oracle.

The drivers restate the reference's case modules (stopping rules, post-processing); run on
the oracle they must give the C++ oracle driver's figures bit for bit (Poiseuille) and the
reference's printed validation rows (Schaefer-Turek at 20 cells per diameter).
"""
import math

import numpy as np
import pytest

import oracle
from vortexlab_amd import api, cases

from published import ST_2D1, ST_2D2, printed

MODEL = {"bgk": (api.BGK, 0.0), "trt": (api.TRT, 3.0 / 16.0), "trt14": (api.TRT, 0.25)}


@pytest.mark.parametrize("scheme,q,rows", [("bgk", 0.5, 8), ("trt", 0.25, 16), ("trt", 0.5, 8), ("trt14", 0.5, 16)])
def test_poiseuille_driver_equals_oracle_driver(orc, scheme, q, rows):
    o = cases.poiseuille_run(orc, rows, 0.8, MODEL[scheme], q, arg=2)
    l2, balance, steps = oracle.poiseuille(rows, 0.8, *MODEL[scheme], q)
    assert (o.l2, o.force_balance, o.steps) == (l2, balance, steps)


def test_observed_order_closed_form():
    """Errors e and e/4 at h and 2h (q = 0.5: h = rows) give order 2 exactly."""
    a = cases.PoiseuilleOutcome(8, 0.5, 1e-2, 0, 1.0)
    b = cases.PoiseuilleOutcome(16, 0.5, 2.5e-3, 0, 1.0)
    assert cases.observed_orders([a, b]) == pytest.approx([2.0], abs=1e-14)


def test_2d2_driver_reproduces_the_published_row(orc):
    n, u, *row = ST_2D2[0]
    o = cases.run_2d2(orc, n, u, arg=8)
    got = (o.cd_max, o.cd_min, o.cl_max, o.cl_min, o.st)
    assert all(printed(v, p) for v, p in zip(got, row)), f"{got} vs {row}"


def test_2d1_driver_reproduces_the_published_row(orc):
    n, u, *row = ST_2D1[1]
    o = cases.run_2d1(orc, n, u, arg=8)
    assert o.converged
    got = (o.cd, o.cl, o.la, o.dp)
    assert all(printed(v, p) for v, p in zip(got, row)), f"{got} vs {row}"


def test_peaks_closed_form_and_oracle(orc):
    """Maxima of sin(0.3 i) at i = (pi/2 + 2 pi k) / 0.3; parabolic refinement of an exact
    parabola is exact; library and oracle give the same bits."""
    from vortexlab_amd import backend
    x = np.sin(np.arange(200) * 0.3)
    p = api.peaks(backend(), x)
    assert np.array_equal(p, api.peaks(orc, x))
    expect = (math.pi / 2 + 2 * math.pi * np.arange(len(p))) / 0.3
    assert np.max(np.abs(p[:, 0] - expect)) < 0.02 and np.all(np.abs(p[:, 1] - 1.0) < 1e-3)
    par = -((np.arange(11) - 4.3) ** 2)
    q = api.peaks(backend(), par)
    assert q.shape == (1, 2) and abs(q[0, 0] - 4.3) < 1e-12 and abs(q[0, 1]) < 1e-12


def test_output_rows_in_reference_formats():
    text = cases.forces_csv(np.array([1.0, 2.0]), np.array([0.5, -0.5]), 0.25)
    assert text == "t,cd,cl\n0.0000,1.000000,0.500000\n0.2500,2.000000,-0.500000\n"
    assert cases.BENCH_HEADER == ["threads", "grid", "fluid_nodes", "steps", "seconds", "mlups", "state_hash"]


def test_cavity_driver_on_the_oracle(orc):
    """tests/validation.rs:41-52 at 33 x 33, Re 100: converged; deviations from Ghia below
    0.006 / 0.008 (the reference measured 0.0039 / 0.0052); u_min within 1.5 % of the spectral
    value."""
    from published import ghia
    o = cases.cavity_run(orc, 33, 100.0, 0.1, arg=8)
    c = cases.cavity_compare(o, *ghia(100))
    assert o.converged
    assert printed(c.max_dev_u, "0.0039") and printed(c.max_dev_v, "0.0052")
    assert abs(c.u_min[1] / -0.2140424 - 1.0) < 0.015


def test_cavity_interpolation_exact_for_cubics():
    """cavity.rs:103-113: cubic Lagrange interpolation reproduces a cubic."""
    f = lambda x: 1.0 - 2.0 * x + 0.5 * x * x * x  # noqa: E731
    prof = [(x, f(x)) for x in (0.0, 0.07, 0.2, 0.31, 0.5, 0.74, 0.9, 1.0)]
    for s in (0.0, 0.01, 0.25, 0.5, 0.62, 0.95, 1.0):
        assert abs(cases.interpolate(prof, s) - f(s)) < 1e-12
