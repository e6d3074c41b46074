"""CPU-side checks of the product library: the C ABI loads and exports every declared entry
point, host-side logic (shapes, cut fractions, statistics) agrees with the oracle, and the
solver refuses to run without a GPU instead of falling back to anything."""
import math
import os
import re

import numpy as np
import pytest

import vortexlab_amd as vl
from vortexlab_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(vl_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = vl.backend().lib
    names = declared_symbols()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_abi_version():
    assert vl.backend().lib.vl_abi_version() == 1


def test_no_gpu_means_a_loud_error():
    if vl.device_count() > 0:
        pytest.skip("a GPU is visible")
    b = vl.Builder(10, 10)
    b.periodic_x = b.periodic_y = True
    with pytest.raises(api.VortexError) as e:
        b.build(api.F64)
    assert e.value.status == api.EDEVICE


SHAPES = [
    api.circle(20.3, 15.7, 6.4),
    api.section("double-recessed", 0.1, 20.5, 16.5, 12.0, 0.3),
    api.section("recessed", 0.125, 20.0, 16.0, 14.0, 0.0),
    api.section("chamfered", 0.15, 19.5, 16.0, 13.0, 0.0),
    api.section("rounded", 0.2, 20.0, 16.5, 12.0, 0.4),
    api.translated(api.section("square", 0.0, 0.0, 0.0, 4.0), 1.0, -0.5),
    api.polygon([[10, 10], [30, 12], [25, 25], [12, 22]]),
]


@pytest.mark.parametrize("k", range(len(SHAPES)))
def test_shapes_and_cut_fractions_match_oracle(orc, k):
    g, o = vl.backend(), orc
    keep = []  # callbacks of the shapes, held while the handles are used
    hg, ho = SHAPES[k].make(g, keep), SHAPES[k].make(o, keep)
    rng = np.random.default_rng(k)
    pts = rng.uniform(0, 40, size=(2000, 2))
    for x, y in pts:
        assert g.shape_inside(hg, x, y) == o.shape_inside(ho, x, y)
    # every lattice link from an outside node into an inside node
    for y in range(0, 33):
        for x in range(0, 41):
            if o.shape_inside(ho, x, y):
                continue
            for j in range(1, 9):
                if o.shape_inside(ho, x + api.CX[j], y + api.CY[j]):
                    qg = g.cut_fraction(hg, x, y, api.CX[j], api.CY[j])
                    qo = o.cut_fraction(ho, x, y, api.CX[j], api.CY[j])
                    assert qg == qo
    g.shape_free(hg)
    o.shape_free(ho)


def test_cut_fraction_closed_form():
    """src/geometry.rs:298-308: the bisection finds the circle crossing."""
    g = vl.backend()
    c = api.circle(0.0, 0.0, 10.3).make(g, [])
    assert abs(g.cut_fraction(c, -11.0, 0.0, 1.0, 0.0) - 0.7) < 1e-9
    assert abs(g.cut_fraction(c, 8.0, 8.0, -1.0, -1.0) - (8.0 - 10.3 / math.sqrt(2.0))) < 1e-9
    g.shape_free(c)


def test_section_areas():
    """src/geometry.rs:310-330: areas of the corner variants."""
    g = vl.backend()
    cases = [("square", 0.0, 1.0), ("chamfered", 0.1, 1.0 - 4 * 0.005),
             ("rounded", 0.1, 1.0 - (4 - math.pi) * 0.01), ("recessed", 0.1, 1.0 - 4 * 0.01),
             ("double-recessed", 0.05, 1.0 - 4 * 3 * 0.0025)]
    n = 600
    h = 1.2 / n
    xs = -0.6 + (np.arange(n) + 0.5) * h
    for corner, b, expect in cases:
        s = api.section(corner, b, 0.0, 0.0, 1.0).make(g, [])
        count = sum(g.shape_inside(s, x, y) for y in xs for x in xs)
        assert abs(count * h * h - expect) < 3e-3, corner
        assert g.shape_inside(s, 0.499, 0.0) and g.shape_inside(s, 0.0, 0.499) and not g.shape_inside(s, 0.501, 0.0)
        g.shape_free(s)


def test_fft_against_numpy():
    rng = np.random.default_rng(5)
    for n in [1, 2, 8, 64, 1024]:
        re, im = rng.standard_normal(n), rng.standard_normal(n)
        fr, fi = api.fft(vl.backend(), re, im)
        assert np.allclose(fr + 1j * fi, np.fft.fft(re + 1j * im), atol=1e-10)


def test_statistics_match_oracle(orc):
    rng = np.random.default_rng(11)
    dt = 0.005
    t = np.arange(12000) * dt
    for f0 in [0.16, 0.2113, 0.05]:
        cl = 0.3 * np.sin(2 * np.pi * f0 * t + 0.3) + 1e-3 * rng.standard_normal(len(t))
        cd = 1.4 + 0.02 * np.sin(4 * np.pi * f0 * t)
        sg = vl.series_stats(cd, cl, dt, 10.0)
        so = api.series_stats(orc, cd, cl, dt, 10.0)
        for k in ("cd_mean", "cd_rms", "cl_mean", "cl_rms", "st", "st_crossings"):
            assert math.isclose(getattr(sg, k), getattr(so, k), rel_tol=1e-12, abs_tol=1e-15), k
        assert sg.periodic == so.periodic
