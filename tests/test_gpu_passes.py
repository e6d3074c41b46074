"""K-step passes (temporal blocking of the bulk) against the oracle and the single-step path.

A group of K steps runs one kernel that takes the interior nodes of its tasks (strips of 64 V
columns, one wave each) from t to t+K, beside K single steps of the boundary nodes and of the
interior nodes no task covers.  Every node's arithmetic is the single-step bulk's, so FP64 must
stay bit-identical to the oracle and to the reference's published hash, and FP32 / FP16 must
give the single-step path's bits.  VL_PASS_STEPS=K forces the mode (and split mode with it) on
small domains, VL_TB_ROWS sets the output rows per task; short tasks multiply the task edges,
the ring nodes whose intermediate values single-step neighbours pull, and a run whose length is
not a multiple of K ends in single steps.
"""
import numpy as np
import pytest

from vortexlab_amd import api

from cases import cylinder_rig_spec, periodic_box, tunnel_spec

pytestmark = pytest.mark.gpu

KS = [2, 3, 4]


def passes(monkeypatch, k, rows=None, graphs=True, layout=None):
    """layout = (waves per workgroup, lockstep, first strip column, columns per lane, joined
    strips), or None for the defaults (trailing entries may be left out)."""
    monkeypatch.setenv("VL_PASS_STEPS", str(k))
    monkeypatch.delenv("VL_STEP_MODE", raising=False)
    monkeypatch.delenv("VL_BULK_V", raising=False)
    names = ("VL_TB_WAVES", "VL_TB_SYNC", "VL_TB_X0", "VL_TB_V", "VL_TB_JOIN")
    layout = tuple(layout or ()) + (None,) * (len(names) - len(layout or ()))
    for name, value in zip(names, layout):
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
    monkeypatch.setenv("VL_PASS_STEPS", "1")
    monkeypatch.setenv("VL_STEP_MODE", mode)
    monkeypatch.delenv("VL_BULK_V", raising=False)
    monkeypatch.delenv("VL_GRAPHS", raising=False)


@pytest.mark.parametrize("k", KS)
@pytest.mark.parametrize("rows,graphs,layout", [(64, True, None), (7, True, None), (1, False, None),
                                                (12, True, (4, 0, 2, None)), (5, True, (8, 1, 7, None)),
                                                (16, True, (8, 1, None, None, 1)), (9, True, (4, 1, 35, None, 1))])
def test_pass_published_hash(gpu, monkeypatch, k, rows, graphs, layout):
    """cylinder::bench n=40, 3000 steps (docs/report/index.html:59) with K-step passes, in the
    default layout (4 waves per workgroup for FP64, strips from x = 32) and others, joined strips
    among them."""
    print(f"[passes] hash n=40 K={k} rows={rows} graphs={graphs} layout={layout}", flush=True)
    passes(monkeypatch, k, rows, graphs, layout)
    ch = api.Channel(gpu, cylinder_rig_spec(40), api.F64)
    info = ch.sim.info()
    assert info.pass_steps == k and info.split == 1
    assert (info.pass_joined_tasks > 0) == (layout is not None and len(layout) == 5 and layout[4] == 1)
    assert info.pass_nodes > 0.7 * info.fluid_nodes
    assert info.pass_nodes + info.listed_nodes + info.bnode_count == info.fluid_nodes
    ch.run(3000, 40.0)
    assert ch.sim.state_hash() == 0x6B6814B976BBBAFF
    ch.close()


@pytest.mark.parametrize("k", KS)
@pytest.mark.parametrize("model", ["bgk", "trt"])
@pytest.mark.parametrize("steps", [300, 301])
def test_pass_periodic_box_equals_oracle(gpu, orc, monkeypatch, k, model, steps):
    """Periodic in x and y with a body force: the tasks stay off the wrapping edges, whose nodes
    take the listed single-step path (with the wrap), and every task edge is a ring."""
    print(f"[passes] periodic K={k} {model} steps={steps}", flush=True)
    passes(monkeypatch, k, rows=5)
    g = periodic_box(gpu, model, api.F64, nx=140, ny=40, accel=(1e-6, 0.0))
    o = periodic_box(orc, model, 1, nx=140, ny=40, accel=(1e-6, 0.0))
    assert g.info().pass_tasks > 0
    g.step(steps)
    o.step(steps)
    assert np.array_equal(g.read(), o.read())


@pytest.mark.parametrize("k,prec,rows,v", [(k, p, r, v) for k in KS for p in (api.F32, api.F16)
                                            for r, v in ((64, 1), (3, 1), (12, 2))]
                         + [(5, api.F32, 64, 1), (5, api.F32, 3, 1), (5, api.F32, 20, 1)])
def test_pass_equals_single_step_reduced_precision(gpu, monkeypatch, k, prec, rows, v):
    """A tunnel with radiating sides, a ramping inflow and a square body: K-step passes give the
    single-step bits, populations and forces, over a step count that is not a multiple of K
    (K = 5 is FP32 only, with one column per lane)."""
    print(f"[passes] K={k} prec={prec} rows={rows} V={v}", flush=True)
    spec = tunnel_spec(d=16, re=100.0)
    single_step(monkeypatch)
    ref = api.Channel(gpu, spec, prec)
    ref_coef = ref.run(1201, 16.0)
    ref_pops = ref.sim.read()
    ref.close()
    passes(monkeypatch, k, rows, layout=(None, None, None, v))
    ch = api.Channel(gpu, spec, prec)
    assert ch.sim.info().pass_tasks > 0
    coef = ch.run(1201, 16.0)
    assert np.array_equal(ch.sim.read(), ref_pops), "populations differ"
    assert np.array_equal(coef, ref_coef), "forces differ"
    ch.close()


@pytest.mark.parametrize("k,prec,rows,v", [(k, p, r, v) for k in KS for p in (api.F32, api.F16)
                                            for r, v in ((64, 1), (3, 1), (12, 2))]
                         + [(5, api.F32, 20, 1), (5, api.F32, 3, 1)])
def test_joined_pass_equals_single_step(gpu, monkeypatch, k, prec, rows, v):
    """Joined strips (a workgroup's waves sweep one strip and hand each other their edge values
    at every level) on the tunnel with a square body, in workgroups of 4 waves (256 V columns, so
    that they fit its width): the single-step bits, populations and forces."""
    waves = 4
    print(f"[passes] joined K={k} prec={prec} rows={rows} V={v} waves={waves}", flush=True)
    spec = tunnel_spec(d=16 if v == 1 else 32, re=100.0)
    single_step(monkeypatch)
    ref = api.Channel(gpu, spec, prec)
    ref_coef = ref.run(1201, 16.0)
    ref_pops = ref.sim.read()
    ref.close()
    passes(monkeypatch, k, rows, layout=(waves, 1, None, v, 1))
    ch = api.Channel(gpu, spec, prec)
    info = ch.sim.info()
    assert info.pass_joined_tasks > 0 and info.pass_joined_tasks % waves == 0
    coef = ch.run(1201, 16.0)
    assert np.array_equal(ch.sim.read(), ref_pops), "populations differ"
    assert np.array_equal(coef, ref_coef), "forces differ"
    ch.close()


@pytest.mark.parametrize("k,prec", [(5, api.F32), (4, api.F64), (3, api.F16), (1, api.F32)])
def test_prepared_short_runs_equal_direct_launches(gpu, monkeypatch, k, prec):
    """vl_sim_prepare instantiates the graph of a run's whole groups that do not fill a step
    window; runs that replay it give the bits of the same runs launched group by group (K-step
    passes, and fused single steps), over pieces of every kind: shorter than a group, a few
    groups, a window and more, with and without single steps left; preparing leaves the state
    as it was.  Each piece prepared just before it runs, or ahead, before the piece before it
    (vl_sim_prepare_after)."""
    spec = tunnel_spec(d=16, re=100.0)
    if k == 1:
        single_step(monkeypatch, "fused")
    else:
        passes(monkeypatch, k, rows=12)
    pieces = (5, 20, 7, 33, 3, 61, 1, 12, 40)
    print(f"[passes] prepared runs K={k} prec={prec} pieces={pieces}", flush=True)

    def go(prepare):
        ch = api.Channel(gpu, spec, prec)
        assert ch.sim.info().pass_steps == k
        coefs = []
        for i, n in enumerate(pieces):
            t = ch.sim.t
            if prepare == "now":
                ch.sim.prepare(n)
            elif prepare == "ahead" and i + 1 < len(pieces):
                ch.sim.prepare(pieces[i + 1], after=n)
            assert ch.sim.t == t
            coefs.append(ch.run(n, 16.0))
        pops = ch.sim.read()
        ch.close()
        return np.concatenate(coefs), pops

    direct = go(None)
    for mode in ("now", "ahead"):
        prepared = go(mode)
        assert np.array_equal(direct[0], prepared[0]), f"forces differ ({mode})"
        assert np.array_equal(direct[1], prepared[1]), f"populations differ ({mode})"


@pytest.mark.parametrize("workload,prec,steps", [("notched_8192x4096", api.F32, 43),
                                                 ("circle_16384x8192", api.F16, 23)])
def test_pass_full_size_equals_single_step(gpu, monkeypatch, workload, prec, steps):
    """The benchmark domains at full size (8192 x 4096 FP32 with the notched square, 16384 x 8192
    FP16 with the D = 512 circle), in their default passes (5 and 3 steps): the state after a
    step count that mixes passes with single steps has the single-step path's hash and forces."""
    import bench
    print(f"[passes] full size {workload} prec={prec} steps={steps}", flush=True)
    builder, d = bench.WORKLOADS[workload][:2]
    spec = builder(api, 1)
    single_step(monkeypatch)
    ref = api.Channel(gpu, spec, prec)
    assert ref.sim.info().pass_steps == 1
    ref_coef = ref.run(steps, d)
    ref_hash = ref.sim.state_hash()
    ref.close()
    for name in ("VL_PASS_STEPS", "VL_STEP_MODE", "VL_TB_ROWS", "VL_TB_WAVES", "VL_TB_SYNC", "VL_TB_X0", "VL_TB_V",
                 "VL_TB_JOIN"):
        monkeypatch.delenv(name, raising=False)
    ch = api.Channel(gpu, spec, prec)
    info = ch.sim.info()
    assert info.pass_steps == (3 if prec == api.F16 else 5)
    assert info.pass_nodes > 0.98 * info.fluid_nodes
    assert info.pass_joined_tasks == 0
    coef = ch.run(steps, d)
    assert ch.sim.state_hash() == ref_hash
    assert np.array_equal(coef, ref_coef)
    ch.close()
