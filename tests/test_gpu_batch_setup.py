"""Device set-up of a batch (cell kinds, node flags and boundary tables built by the set-up
kernels) against the host builder (src/sim.rs:84-354, src/geometry.rs:16-49, 268-279 restated
in setup.cpp): every table of every case must be byte-identical, and so the runs bit-identical."""
import math

import numpy as np
import pytest

from vortexlab_amd import api, workloads

pytestmark = pytest.mark.gpu


def mixed_cases():
    """Seeded sweep cases (star polygons in towing tanks with radiating sides, ramp and kick)
    plus channels with the other inflows and a circular body."""
    cases = workloads.sweep_cases(2024, 40)
    rng = np.random.default_rng(7)
    for k in range(6):
        nx, ny = int(rng.integers(150, 400)), int(rng.integers(50, 120))
        d = float(rng.uniform(8.0, 20.0))
        cx, cy = 0.5 + round(0.3 * nx), 0.5 * (ny - 1) + rng.uniform(-2, 2)
        inflow = [api.INFLOW_PARABOLIC, api.INFLOW_UNIFORM][k % 2]
        body = api.circle(cx, cy, 0.5 * d) if k < 3 else api.polygon(workloads.star_polygon(rng, d, cx, cy))
        kick = api.Kick(cx, cy, 0.4 * 0.05 / d, 100, 80) if k % 3 == 0 else None
        spec = api.ChannelSpec(nx=nx, ny=ny, inflow=inflow, u=0.05, nu=0.05 * d / 80.0, body=body,
                               collision=api.TRT if k % 2 else api.BGK, magic=3.0 / 16.0, ramp_steps=200, kick=kick)
        cases.append(workloads.Case(spec=spec, steps=300 + 37 * k, d=d, re=80.0, outline=None))
    return cases


def build(gpu, monkeypatch, cases, where, prec=api.F32):
    monkeypatch.setenv("VL_BATCH_SETUP", where)
    b = api.Batch(gpu, prec)
    for c in cases:
        b.add(c.spec, c.steps, c.d)
    b.build()
    return b


def test_device_tables_equal_host_tables(gpu, monkeypatch):
    cases = mixed_cases()
    dev = build(gpu, monkeypatch, cases, "device")
    host = build(gpu, monkeypatch, cases, "host")
    assert dev.info().setup_on_device == 1 and host.info().setup_on_device == 0
    print(f"[setup] device {list(np.round(dev.info().setup_ms, 1))} ms, host {list(np.round(host.info().setup_ms, 1))} ms",
          flush=True)
    for k in range(len(cases)):
        dd, hd = dev.table_digests(k), host.table_digests(k)
        bad = [n for n in dd if dd[n] != hd[n]]
        assert not bad, f"case {k}: tables differ in {bad}"
        assert dev.case_info(k) == host.case_info(k)
    assert dev.info().fluid_nodes == host.info().fluid_nodes
    assert dev.info().pass_tasks == host.info().pass_tasks and dev.info().listed_nodes == host.info().listed_nodes
    # the same tables give the same run, bit for bit (a few hundred steps of every case)
    dev.run(400)
    host.run(400)
    for k in range(len(cases)):
        a, b = dev.result(k), host.result(k)
        assert np.array_equal(a[0], b[0]) and a[1:] == b[1:], f"case {k}: run differs"
    dev.close()
    host.close()


def test_other_bodies_use_the_host_builder(gpu, monkeypatch):
    """A body the set-up kernels cannot evaluate (a rounded square) keeps the whole batch on the
    host builder, with the same results as a batch of that case alone."""
    spec = api.ChannelSpec(nx=200, ny=80, inflow=api.INFLOW_UNIFORM_OPEN, u=0.05, nu=0.05 * 12 / 80.0,
                           body=api.section("rounded", 0.2, 60.5, 40.0, 12.0), collision=api.TRT, magic=3.0 / 16.0,
                           ramp_steps=100)
    cases = [workloads.Case(spec=spec, steps=200, d=12.0, re=80.0, outline=None)] + workloads.sweep_cases(5, 2)
    b = build(gpu, monkeypatch, cases, "device")
    assert b.info().setup_on_device == 0
    b.run()
    assert all(b.result(k)[2] == 0 for k in range(len(cases)))
    b.close()
