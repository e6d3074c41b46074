"""Two-step passes (temporal blocking of the bulk) against the oracle and the single-step path.

A step pair runs one kernel that takes the interior nodes of its tasks (64-column strips, one
wave each) from t to t+2, beside two single steps of the boundary nodes and of the interior
nodes no task covers.  Every node's arithmetic is the single-step bulk's, so FP64 must stay
bit-identical to the oracle and to the reference's published hash, and FP32 / FP16 must give
the single-step path's bits.  VL_TWO_STEP=1 forces the mode (and split mode with it) on small
domains, VL_TB_ROWS sets the output rows per task; short tasks multiply the task edges, the
ring nodes whose t+1 values single-step neighbours pull, and the odd remainder of a run is a
single step.
"""
import numpy as np
import pytest

from vortexlab_amd import api

from cases import cylinder_rig_spec, periodic_box, tunnel_spec

pytestmark = pytest.mark.gpu


def two_step(monkeypatch, rows=None, graphs=True, layout=None):
    """layout = (waves per workgroup, lockstep, first strip column), or None for the defaults."""
    monkeypatch.setenv("VL_TWO_STEP", "1")
    monkeypatch.delenv("VL_STEP_MODE", raising=False)
    monkeypatch.delenv("VL_BULK_V", raising=False)
    for name, value in zip(("VL_TB_WAVES", "VL_TB_SYNC", "VL_TB_X0"), layout or (None, None, None)):
        if value is None:
            monkeypatch.delenv(name, raising=False)
        else:
            monkeypatch.setenv(name, str(value))
    if rows is None:
        monkeypatch.delenv("VL_TB_ROWS", raising=False)
    else:
        monkeypatch.setenv("VL_TB_ROWS", str(rows))
    if graphs:
        monkeypatch.delenv("VL_GRAPHS", raising=False)
    else:
        monkeypatch.setenv("VL_GRAPHS", "0")


def single_step(monkeypatch, mode="split"):
    monkeypatch.setenv("VL_TWO_STEP", "0")
    monkeypatch.setenv("VL_STEP_MODE", mode)
    monkeypatch.delenv("VL_BULK_V", raising=False)
    monkeypatch.delenv("VL_GRAPHS", raising=False)


@pytest.mark.parametrize("rows,graphs,layout", [(64, True, None), (7, True, None), (1, False, None),
                                                (12, True, (4, 0, 2)), (5, True, (8, 1, 7))])
def test_two_step_published_hash(gpu, monkeypatch, rows, graphs, layout):
    """cylinder::bench n=40, 3000 steps (docs/report/index.html:59) with two-step passes, in
    the default layout (8 waves per workgroup in step, strips from x = 32) and others."""
    print(f"[two-step] hash n=40 rows={rows} graphs={graphs} layout={layout}", flush=True)
    two_step(monkeypatch, rows, graphs, layout)
    ch = api.Channel(gpu, cylinder_rig_spec(40), api.F64)
    info = ch.sim.info()
    assert info.two_step == 1 and info.split == 1
    assert info.two_step_nodes > 0.8 * info.fluid_nodes
    assert info.two_step_nodes + info.listed_nodes + info.bnode_count == info.fluid_nodes
    ch.run(3000, 40.0)
    assert ch.sim.state_hash() == 0x6B6814B976BBBAFF
    ch.close()


@pytest.mark.parametrize("model", ["bgk", "trt"])
@pytest.mark.parametrize("steps", [300, 301])
def test_two_step_periodic_box_equals_oracle(gpu, orc, monkeypatch, model, steps):
    """Periodic in x and y with a body force: the tasks stay off the wrapping edges, whose nodes
    take the listed single-step path (with the wrap), and every task edge is a ring."""
    print(f"[two-step] periodic {model} steps={steps}", flush=True)
    two_step(monkeypatch, rows=5)
    g = periodic_box(gpu, model, api.F64, nx=130, ny=40, accel=(1e-6, 0.0))
    o = periodic_box(orc, model, 1, nx=130, ny=40, accel=(1e-6, 0.0))
    assert g.info().two_step_tasks > 0
    g.step(steps)
    o.step(steps)
    assert np.array_equal(g.read(), o.read())


@pytest.mark.parametrize("prec", [api.F32, api.F16])
@pytest.mark.parametrize("rows", [64, 3])
def test_two_step_equals_single_step_reduced_precision(gpu, monkeypatch, prec, rows):
    """A tunnel with radiating sides, a ramping inflow and a square body: two-step passes give
    the single-step bits, populations and forces, over an odd number of steps."""
    print(f"[two-step] prec={prec} rows={rows}", flush=True)
    spec = tunnel_spec(d=16, re=100.0)
    single_step(monkeypatch)
    ref = api.Channel(gpu, spec, prec)
    ref_coef = ref.run(1201, 16.0)
    ref_pops = ref.sim.read()
    ref.close()
    two_step(monkeypatch, rows)
    ch = api.Channel(gpu, spec, prec)
    assert ch.sim.info().two_step_tasks > 0
    coef = ch.run(1201, 16.0)
    assert np.array_equal(ch.sim.read(), ref_pops), "populations differ"
    assert np.array_equal(coef, ref_coef), "forces differ"
    ch.close()
