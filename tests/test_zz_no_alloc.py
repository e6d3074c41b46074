"""vl_sim_run_noalloc (the allocation-free variant of vl_sim_run, for callers that cannot allocate
device memory): on the same simulation and the same per-step controls it must give the force
history and the populations of vl_sim_run bit for bit, on two calls in a row, the second one
shorter, in fused single steps and in K-step passes, and it must refuse calls beyond its fixed
scratch area."""
import math

import numpy as np
import pytest

from vortexlab_amd import api
from cases import small_cylinder_spec

pytestmark = pytest.mark.gpu


def controls(t0, n, ramp=200):
    """A ramping inflow with a body kick and the absorbing outlet's target, steps t0 .. t0+n-1."""
    rows, targets = [], []
    for t in range(t0, t0 + n):
        r = min(1.0, t / ramp)
        rows.append((1.0, r, 0.01 * math.sin(0.05 * t), 1.0))
        targets.append(0.05 * r)
    return api.make_ctl(rows, targets)


def run_pair(gpu, spec, prec, lengths):
    a = api.Channel(gpu, spec, prec)
    b = api.Channel(gpu, spec, prec)
    ng = int(a.sim.info().force_groups)
    t = 0
    for n in lengths:
        arr = (api.Ctl * n)(*controls(t, n))
        ha = np.zeros((n, ng, 2))
        hb = np.full((n, ng, 2), np.nan)
        gpu.check(gpu.sim_run(a.sim.h, n, arr, api._dptr(ha)))
        gpu.check(gpu.sim_run_noalloc(b.sim.h, n, arr, api._dptr(hb)))
        print(f"[noalloc] prec {prec}, steps {t}..{t + n}: {a.sim.info().nx}x{a.sim.info().ny}, "
              f"pass steps {a.sim.info().pass_steps}", flush=True)
        assert np.array_equal(ha, hb), f"force history differs at steps {t}..{t + n}"
        assert a.sim.state_hash() == b.sim.state_hash()
        assert np.array_equal(a.sim.force, b.sim.force)
        assert a.sim.outlet_density_offset == b.sim.outlet_density_offset
        t += n
    assert np.array_equal(a.sim.read(), b.sim.read())
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [api.F64, api.F32])
def test_noalloc_equals_run_small(gpu, prec):
    """222 x 43 channel: fused single steps."""
    run_pair(gpu, small_cylinder_spec(), prec, [700, 300])


def test_noalloc_equals_run_passes(gpu):
    """2304 x 960 channel (above 2^21 nodes): K-step passes beside single steps, launched directly."""
    spec = api.ChannelSpec(nx=2304, ny=960, inflow=api.INFLOW_UNIFORM_OPEN, u=0.05, nu=0.05 * 64 / 200.0,
                           body=api.circle(600.5, 480.0, 32.0), collision=api.TRT, magic=3.0 / 16.0, ramp_steps=300)
    run_pair(gpu, spec, api.F32, [203, 57])


def test_noalloc_refuses_beyond_its_scratch(gpu):
    ch = api.Channel(gpu, small_cylinder_spec(), api.F32)
    n = 4097
    arr = (api.Ctl * n)(*controls(0, n))
    with pytest.raises(api.VortexError, match="at most 4096 steps"):
        gpu.check(gpu.sim_run_noalloc(ch.sim.h, n, arr, None))
    assert ch.sim.t == 0
    ch.close()
