"""The split step (boundary kernel beside a vectorised bulk kernel) against the oracle and the
fused step.

Large domains run the bulk with V nodes per thread (vector loads), or with the x-shifted
populations staged through LDS; nodes whose neighbours wrap around a periodic edge, or that
share a vector with a non-interior node, take the one-node path.  The arithmetic per node is the same in every variant, so FP64 must stay
bit-identical to the oracle and to the reference's published hash, and FP32 / FP16 must give
the same bits as the fused kernel.  VL_STEP_MODE forces the mode and VL_BULK_V the vector
width (both read when a simulation is created); VL_PASS_STEPS=1 keeps every step a single step.
"""
import numpy as np
import pytest

from vortexlab_amd import api

from cases import cylinder_rig_spec, periodic_box, tunnel_spec

pytestmark = pytest.mark.gpu

# "0": one node per thread with the x-shifted populations staged through LDS
WIDTHS = {api.F64: ["0", "1", "2"], api.F32: ["0", "1", "2", "4"], api.F16: ["0", "2", "4", "8"]}


def make(monkeypatch, mode, v=None):
    monkeypatch.setenv("VL_PASS_STEPS", "1")  # single steps only (K-step passes: test_gpu_passes.py)
    monkeypatch.setenv("VL_STEP_MODE", mode)
    if v is None:
        monkeypatch.delenv("VL_BULK_V", raising=False)
    else:
        monkeypatch.setenv("VL_BULK_V", v)


@pytest.mark.parametrize("v", WIDTHS[api.F64])
def test_split_published_hash(gpu, monkeypatch, v):
    """cylinder::bench n=40, 3000 steps (docs/report/index.html:59) in split mode."""
    print(f"[split] hash n=40 V={v}", flush=True)
    make(monkeypatch, "split", v)
    ch = api.Channel(gpu, cylinder_rig_spec(40), api.F64)
    assert ch.sim.info().split == 1
    ch.run(3000, 40.0)
    assert ch.sim.state_hash() == 0x6B6814B976BBBAFF
    ch.close()


@pytest.mark.parametrize("v", WIDTHS[api.F64])
@pytest.mark.parametrize("model", ["bgk", "trt"])
def test_split_periodic_box_equals_oracle(gpu, orc, monkeypatch, model, v):
    """Periodic in x and y: every edge node wraps, so the one-node path runs beside the
    vector path (width 50 is not a multiple of the vector width either)."""
    print(f"[split] periodic {model} V={v}", flush=True)
    make(monkeypatch, "split", v)
    g = periodic_box(gpu, model, api.F64, nx=50, ny=36, accel=(1e-6, 0.0))
    o = periodic_box(orc, model, 1, nx=50, ny=36, accel=(1e-6, 0.0))
    g.step(300)
    o.step(300)
    assert np.array_equal(g.read(), o.read())


@pytest.mark.parametrize("prec", [api.F32, api.F16])
def test_split_equals_fused_reduced_precision(gpu, monkeypatch, prec):
    """A tunnel with radiating sides, a ramping inflow and a square body: the split step at
    every vector width gives the fused step's bits, populations and forces."""
    spec = tunnel_spec(d=16, re=100.0)
    make(monkeypatch, "fused")
    ref = api.Channel(gpu, spec, prec)
    ref_coef = ref.run(1200, 16.0)
    ref_pops = ref.sim.read()
    ref.close()
    for v in WIDTHS[prec]:
        print(f"[split] prec={prec} V={v}", flush=True)
        make(monkeypatch, "split", v)
        ch = api.Channel(gpu, spec, prec)
        coef = ch.run(1200, 16.0)
        assert np.array_equal(ch.sim.read(), ref_pops), f"V={v}: populations differ"
        assert np.array_equal(coef, ref_coef), f"V={v}: forces differ"
        ch.close()
