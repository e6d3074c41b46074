"""Pins the CPU oracle against what the reference itself publishes and what mathematics fixes.

* bit-exact: the FNV-1a state hashes printed in the reference report after its benchmark
  (docs/report/index.html:59) and the fluid-node counts;
* printed values: the Poiseuille L2 table (docs/report/index.html, section 1, 4 digits);
* closed forms / invariants the reference's own tests assert (tests/solver.rs,
  tests/validation.rs, src/lattice.rs and src/signal.rs unit tests);
* textbook routines: the FFT against the O(n^2) DFT.
A plausible slip in the restatement (a dropped term, a wrong sign or index, a swapped pair)
breaks the bit-exact hash and at least one of the invariants.
"""
import math

import numpy as np
import pytest

import oracle
from vortexlab_amd import api

from cases import COMOVING_SHAPES, closed_box, comoving_body, couette, open_channel, periodic_box


def test_published_hash_n40():
    h, fluid, nx, ny, _ = oracle.cylinder_bench(40, 3000, threads=4)
    assert (nx, ny, fluid) == (882, 166, 143056)
    assert h == 0x6B6814B976BBBAFF


@pytest.mark.slow
def test_published_hash_n80():
    h, fluid, nx, ny, _ = oracle.cylinder_bench(80, 1000, threads=8)
    assert (nx, ny, fluid) == (1762, 330, 572256)
    assert h == 0xC93E26BFA33474EE


def test_hash_independent_of_thread_count():
    """tests/solver.rs:271-291 (results do not depend on the row split)."""
    hashes = {oracle.cylinder_bench(10, 400, threads=t)[0] for t in (1, 2, 3, 7)}
    assert len(hashes) == 1


from published import POISEUILLE as POISEUILLE_TABLE  # noqa: E402
MODEL = {"bgk": (api.BGK, 0.0), "trt": (api.TRT, 3.0 / 16.0), "trt14": (api.TRT, 0.25)}


@pytest.mark.parametrize("scheme,q,rows,published", POISEUILLE_TABLE)
def test_poiseuille_published_l2(scheme, q, rows, published):
    l2, balance, _ = oracle.poiseuille(rows, 0.8, *MODEL[scheme], q)
    assert abs(l2 / published - 1.0) < 6e-4, f"l2 {l2:.4e} vs published {published:.3e}"
    assert abs(balance - 1.0) < 1e-9


@pytest.mark.parametrize("rows", [8, 16, 32])
def test_poiseuille_trt_magic_is_exact(rows):
    """tests/validation.rs:32-35: TRT Lambda=3/16 + halfway walls reproduce the parabola."""
    l2, balance, _ = oracle.poiseuille(rows, 0.8, api.TRT, 3.0 / 16.0, 0.5)
    assert l2 < 1e-10 and abs(balance - 1.0) < 1e-9


@pytest.mark.parametrize("model", ["bgk", "trt", "trt14"])
def test_periodic_box_conserves(orc, model):
    s = periodic_box(orc, model, 3)
    m0, p0 = s.total_mass(), s.total_momentum()
    s.step(400)
    m1, p1 = s.total_mass(), s.total_momentum()
    assert abs(m1 - m0) < 2e-12
    assert abs(p1[0] - p0[0]) < 1e-12 and abs(p1[1] - p0[1]) < 1e-12
    assert s.max_speed() < 0.07


def test_body_force_impulse(orc):
    s = periodic_box(orc, "trt", 1, nx=20, ny=16, tau=0.7, accel=(3e-5, -1e-5))
    m, p0 = s.total_mass(), s.total_momentum()
    s.step(250)
    p1 = s.total_momentum()
    assert abs(p1[0] - p0[0] - 250.0 * 3e-5 * m) < 1e-11
    assert abs(p1[1] - p0[1] + 250.0 * 1e-5 * m) < 1e-11


@pytest.mark.parametrize("model", ["bgk", "trt"])
def test_closed_box(orc, model):
    s = closed_box(orc, model, 2)
    m0 = s.total_mass()
    s.step(600)
    assert abs(s.total_mass() - m0) < 2e-12
    assert s.max_speed() < 0.05


@pytest.mark.parametrize("model", ["bgk", "trt"])
@pytest.mark.parametrize("frame", [0.0, -0.05])
def test_couette(orc, model, frame):
    shear = 0.04
    s = couette(orc, model, 1, frame, shear)
    s.step(20000)
    _, ux, _ = s.macros_all()
    dev = max(abs(ux[y, 2] - (frame + shear * (y - 0.5) / 20)) for y in range(1, 21))
    assert dev < 1e-12
    f = s.force
    expected = (0.9 - 0.5) / 3.0 * shear / 20.0
    assert abs(f[0][0] / 6 - expected) < 1e-12 * max(1.0, 1.0 / expected) * expected + 1e-13
    assert abs(f[1][0] / 6 + expected) < 1e-13
    assert abs(f[0][1] / 6 + 1.0 / 3.0) < 1e-12


@pytest.mark.parametrize("shape", sorted(COMOVING_SHAPES))
def test_comoving_body(orc, shape):
    for u in [(0.06, 0.0), (0.0, -0.05), (0.04, 0.03)]:
        s = comoving_body(orc, "trt", 1, COMOVING_SHAPES[shape](), u)
        s.step(200)
        f = s.force[0]
        assert abs(f[0]) < 1e-13 and abs(f[1]) < 1e-13
        _, ux, uy = s.macros(3, 3)
        assert abs(ux - u[0]) < 1e-14 and abs(uy - u[1]) < 1e-14


def test_uniform_stream(orc):
    s = open_channel(orc, 2, 40, 16, "uniform", 0.07, 0.55)
    s.step(500)
    for x, y in [(1, 1), (20, 8), (38, 16), (38, 1)]:
        rho, ux, uy = s.macros(x, y)
        assert abs(rho - 1.0) < 1e-13 and abs(ux - 0.07) < 1e-13 and abs(uy) < 1e-13


def test_cut_fraction_on_a_circle(orc):
    """src/geometry.rs:298-308."""
    c = api.circle(0.0, 0.0, 10.3).make(orc, [])
    assert abs(orc.cut_fraction(c, -11.0, 0.0, 1.0, 0.0) - 0.7) < 1e-9
    q = orc.cut_fraction(c, 8.0, 8.0, -1.0, -1.0)
    assert abs(q - (8.0 - 10.3 / math.sqrt(2.0))) < 1e-9
    orc.shape_free(c)


def test_fft_matches_naive_dft(orc):
    """src/signal.rs:159-172."""
    rng = np.random.default_rng(2024)
    for n in [1, 2, 4, 8, 64, 256]:
        re, im = rng.standard_normal(n), rng.standard_normal(n)
        er, ei = oracle.dft_naive(re, im)
        fr, fi = api.fft(orc, re, im)
        assert np.allclose(fr, er, atol=1e-10) and np.allclose(fi, ei, atol=1e-10)
        # and against numpy's FFT (an independent implementation)
        ref = np.fft.fft(re + 1j * im)
        assert np.allclose(fr + 1j * fi, ref, atol=1e-10)


def test_dominant_frequency_between_bins(orc):
    """src/signal.rs:187-199."""
    n, dt = 3000, 0.01
    f0 = 13.37 / (n * dt)
    x = 2.5 + 0.8 * np.sin(2 * np.pi * f0 * np.arange(n) * dt + 0.4)
    f, amp = api.dominant_frequency(orc, x, dt)
    assert abs(f / f0 - 1.0) < 2e-4 and abs(amp - 0.8) < 0.01
    fc = api.crossing_frequency(orc, x, dt)
    assert abs(fc / f0 - 1.0) < 1e-3


def test_series_stats(orc):
    """src/cases/channel.rs:213-235."""
    dt = 0.01

    def wave(f, n):
        cl = 0.4 * np.sin(2 * np.pi * f * np.arange(n) * dt)
        return np.full(n, 1.5), cl

    s = api.series_stats(orc, *wave(0.16, 10000), dt, 50.0)
    assert s.periodic and abs(s.st - 0.16) < 1e-3 and abs(s.st_crossings - 0.16) < 2e-3
    assert abs(s.cl_rms - 0.4 / math.sqrt(2)) < 5e-3 and abs(s.cd_mean - 1.5) < 1e-12
    s = api.series_stats(orc, *wave(0.03, 10000), dt, 50.0)
    assert not s.periodic and s.st == 0.0
    n = 5000
    drift = (np.arange(n) / n) ** 2
    assert not api.series_stats(orc, np.ones(n), drift, dt, 10.0).periodic
    s = api.series_stats(orc, np.ones(n), np.zeros(n), dt, 10.0)
    assert not s.periodic and s.st == 0.0 and s.cl_rms == 0.0


def test_fnv1a_known_answers():
    """64-bit FNV-1a published test vectors (empty string, "a", "foobar")."""
    from cases import fnv1a
    assert fnv1a(b"") == 0xcbf29ce484222325
    assert fnv1a(b"a") == 0xaf63dc4c8601ec8c
    assert fnv1a(b"foobar") == 0x85944171f73967e8


def test_wake_golden_fixture_is_consistent():
    """The stored wake result: its history hash covers a float64 (35660 x 2) history whose every
    10th row is stored, and the first 1000 steps of the oracle reproduce those rows exactly."""
    import json
    import os
    from cases import wake_662x302_spec
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "wake_662x302_35660.json")) as fh:
        gold = json.load(fh)
    rows = np.array(gold["coef_every_10"])
    assert rows.shape == (gold["steps"] // 10, 2) and np.all(np.isfinite(rows))
    be = oracle.backend()
    ch = api.Channel(be, wake_662x302_spec(), os.cpu_count() or 1)
    coef = ch.run(1000, gold["d"])
    assert ch.sim.fluid_nodes == gold["fluid_nodes"]
    assert np.array_equal(coef[9::10], rows[:100])
    ch.close()


def test_vorticity_closed_forms(orc):
    """Sim::vorticity (src/sim.rs:581-620) on fields whose differences are known exactly, set as
    equilibria in the closed box of tests/solver.rs:64-80 (fluid x = 1..32, y = 1..24):
    * u = (a y, b x): every difference is exact, so the vorticity is b - a at every fluid node;
    * u = (0, c x^2): central differences give 2 c x inside, the one-sided ones (2 x + 1) c at
      x = 1 (its left neighbour is a wall) and (2 x - 1) c at x = 32;
    and 0 in the walls.  Also Sim::is_fluid and Sim::periodic."""
    s = closed_box(orc, "trt", 1)
    nx, ny = 34, 26
    yy, xx = np.mgrid[0:ny, 0:nx].astype(np.float64)
    a, b = 3e-4, -7e-4
    s.init_arrays(1.0, a * yy, b * xx)
    w = s.vorticity()
    fluid = np.zeros((ny, nx), dtype=bool)
    fluid[1:25, 1:33] = True
    assert np.all(w[~fluid] == 0.0)
    assert np.max(np.abs(w[fluid] - (b - a))) < 1e-9
    c = 2e-5
    s.init_arrays(1.0, 0.0, c * xx * xx)
    w = s.vorticity()
    want = 2.0 * c * xx
    want[:, 1] = 3.0 * c
    want[:, 32] = (2.0 * 32 - 1.0) * c
    assert np.max(np.abs(w[fluid] - want[fluid])) < 1e-9
    assert s.is_fluid(1, 1) and s.is_fluid(32, 24) and not s.is_fluid(0, 5) and not s.is_fluid(33, 5)
    with pytest.raises(api.VortexError):
        s.is_fluid(34, 0)
    assert s.periodic() == (False, False)
    assert periodic_box(orc, "bgk", 1, nx=12, ny=10).periodic() == (True, True)
