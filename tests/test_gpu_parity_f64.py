"""FP64 GPU path against the CPU oracle and the reference's published golden values.

The FP64 kernels keep the reference's operation order and the outlet controller's serial sum,
so populations are expected to be bit-identical to the oracle (which itself reproduces the
reference's published state hashes).  Forces use a fixed-order tree instead of the
reference's serial sum: they must agree to 1e-9 relative (the required tolerance) and be
bitwise reproducible run to run.
"""
import json
import os

import numpy as np
import pytest

from vortexlab_amd import api

from cases import (COMOVING_SHAPES, closed_box, comoving_body, couette, cylinder_rig_spec, open_channel,
                   periodic_box, small_cylinder_spec, tunnel_spec, wake_662x302_spec)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu

POP_RTOL = 1e-12  # required bound; the FP64 path is in fact bit-exact (asserted separately)


def assert_pops(g, o, label, exact=True):
    assert g.shape == o.shape
    if exact and np.array_equal(g, o):
        return
    rel = np.max(np.abs(g - o) / np.maximum(np.abs(o), 1e-300))
    if exact:
        raise AssertionError(f"{label}: populations not bit-identical, max rel diff {rel:.3e}")
    assert rel <= POP_RTOL, f"{label}: max rel diff {rel:.3e}"


def assert_force(g, o, label, rtol=1e-9):
    """Forces within 1e-9 of the force scale of the run (the largest |F| in the history):
    the tree and the serial sum differ only by rounding, which near a zero force is absolute."""
    g, o = np.asarray(g), np.asarray(o)
    err = np.abs(g - o)
    scale = max(float(np.max(np.abs(o))), 1e-300)
    assert np.all(err <= rtol * scale), f"{label}: max err {err.max():.3e} vs scale {scale:.3e}"


@pytest.mark.parametrize("n,steps,golden,fluid", [(40, 3000, 0x6B6814B976BBBAFF, 143056),
                                                   (80, 1000, 0xC93E26BFA33474EE, 572256)])
def test_published_state_hash(gpu, n, steps, golden, fluid):
    """cylinder::bench (src/cases/cylinder.rs:224-232) reproduces docs/report/index.html:59."""
    print(f"[hash] n={n} steps={steps}", flush=True)
    ch = api.Channel(gpu, cylinder_rig_spec(n), api.F64)
    assert ch.sim.fluid_nodes == fluid
    ch.run(steps, float(n))
    assert ch.sim.state_hash() == golden, f"hash {ch.sim.state_hash():016x} != {golden:016x}"


def test_wake_662x302_full_run_matches_oracle(gpu):
    """The 662x302 wake case (BGK, circle D=20, Re 150) over its full 35660 steps: final
    populations bit-identical to the oracle's (state hash) and the Cd/Cl history within 1e-9 of
    its scale at every 10th step."""
    with open(os.path.join(GOLDEN, "wake_662x302_35660.json")) as fh:
        gold = json.load(fh)
    print(f"[wake] 662x302 {gold['steps']} steps", flush=True)
    ch = api.Channel(gpu, wake_662x302_spec(), api.F64)
    assert ch.sim.fluid_nodes == gold["fluid_nodes"]
    coef = ch.run(gold["steps"], gold["d"])
    assert f"{ch.sim.state_hash():016x}" == gold["state_hash"]
    assert_force(coef[9::10], np.array(gold["coef_every_10"]), "wake Cd/Cl")
    ch.close()


@pytest.mark.parametrize("model", ["bgk", "trt", "trt14"])
def test_periodic_box_matches_oracle_and_conserves(gpu, orc, model):
    print(f"[periodic] {model}", flush=True)
    g = periodic_box(gpu, model, api.F64)
    o = periodic_box(orc, model, 1)
    m0, p0 = g.total_mass(), g.total_momentum()
    g.step(400)
    o.step(400)
    assert_pops(g.read(), o.read(), f"periodic {model}")
    # tests/solver.rs:36-43
    m1, p1 = g.total_mass(), g.total_momentum()
    assert abs(m1 - m0) < 2e-12
    assert abs(p1[0] - p0[0]) < 1e-12 and abs(p1[1] - p0[1]) < 1e-12
    assert g.max_speed() < 0.07


def test_body_force_impulse(gpu, orc):
    """tests/solver.rs:49-61 and the forced collision (src/lattice.rs:127-156)."""
    g = periodic_box(gpu, "trt", api.F64, nx=20, ny=16, tau=0.7, accel=(3e-5, -1e-5))
    o = periodic_box(orc, "trt", 1, nx=20, ny=16, tau=0.7, accel=(3e-5, -1e-5))
    m, p0 = g.total_mass(), g.total_momentum()
    g.step(250)
    o.step(250)
    assert_pops(g.read(), o.read(), "forced")
    p1 = g.total_momentum()
    assert abs(p1[0] - p0[0] - 250.0 * 3e-5 * m) < 1e-11
    assert abs(p1[1] - p0[1] + 250.0 * 1e-5 * m) < 1e-11


@pytest.mark.parametrize("model", ["bgk", "trt", "trt14"])
def test_closed_box(gpu, orc, model):
    g = closed_box(gpu, model, api.F64)
    o = closed_box(orc, model, 2)
    m0 = g.total_mass()
    g.step(600)
    o.step(600)
    assert_pops(g.read(), o.read(), f"closed box {model}")
    assert abs(g.total_mass() - m0) < 2e-12
    assert g.max_speed() < 0.05


@pytest.mark.parametrize("model", ["bgk", "trt"])
@pytest.mark.parametrize("frame", [0.0, 0.03, -0.05, 0.08])
def test_couette_wall_shear(gpu, orc, model, frame):
    """tests/solver.rs:104-123: wall forces are Galilean invariant; moving-wall links."""
    shear = 0.04
    g = couette(gpu, model, api.F64, frame, shear)
    o = couette(orc, model, 1, frame, shear)
    g.step(20000)
    o.step(20000)
    assert_pops(g.read(), o.read(), f"couette {model} {frame}")
    assert_force(g.force, o.force, "couette")
    nx, rows = 6, 20
    _, ux, _ = g.macros_all()
    dev = max(abs(ux[y, 2] - (frame + shear * (y - 0.5) / rows)) for y in range(1, rows + 1))
    assert dev < 1e-12
    low, top, normal = g.force[0][0] / nx, g.force[1][0] / nx, g.force[0][1] / nx
    expected = (0.9 - 0.5) / 3.0 * shear / 20.0
    assert abs(low - expected) < 1e-12 * max(1.0, 1.0 / expected) * expected + 1e-13
    assert abs(top + expected) < 1e-13
    assert abs(normal + 1.0 / 3.0) < 1e-12


@pytest.mark.parametrize("shape", sorted(COMOVING_SHAPES))
@pytest.mark.parametrize("u", [(0.06, 0.0), (-0.07, 0.02)])
@pytest.mark.parametrize("model", ["bgk", "trt"])
def test_comoving_body_feels_no_force(gpu, orc, shape, u, model):
    """tests/solver.rs:126-159: interpolated bounce-back on curved / notched bodies."""
    g = comoving_body(gpu, model, api.F64, COMOVING_SHAPES[shape](), u)
    o = comoving_body(orc, model, 1, COMOVING_SHAPES[shape](), u)
    g.step(200)
    o.step(200)
    assert_pops(g.read(), o.read(), f"comoving {shape} {u}")
    f = g.force[0]
    assert abs(f[0]) < 1e-13 and abs(f[1]) < 1e-13
    _, ux, uy = g.macros(3, 3)
    assert abs(ux - u[0]) < 1e-14 and abs(uy - u[1]) < 1e-14


def test_uniform_stream_passes_open_boundaries(gpu, orc):
    """tests/solver.rs:202-212: inlet, moving walls and pressure outlet (no controller)."""
    g = open_channel(gpu, api.F64, 40, 16, "uniform", 0.07, 0.55)
    o = open_channel(orc, 2, 40, 16, "uniform", 0.07, 0.55)
    g.step(500)
    o.step(500)
    assert_pops(g.read(), o.read(), "uniform stream")
    for x, y in [(1, 1), (20, 8), (38, 16), (38, 1)]:
        rho, ux, uy = g.macros(x, y)
        assert abs(rho - 1.0) < 1e-13 and abs(ux - 0.07) < 1e-13 and abs(uy) < 1e-13


def test_small_cylinder_forces_and_hash(gpu, orc):
    """tests/solver.rs:255-291 case: forces every step vs the oracle; reproducible run to run."""
    gch = api.Channel(gpu, small_cylinder_spec(), api.F64)
    och = api.Channel(orc, small_cylinder_spec(), 3)
    gc = gch.run(400, 10.0)
    oc = och.run(400, 10.0)
    assert_pops(gch.sim.read(), och.sim.read(), "small cylinder")
    assert gch.sim.state_hash() == och.sim.state_hash()
    assert_force(gc, oc, "cd/cl history")
    assert gc[-1, 0] > 1.0
    again = api.Channel(gpu, small_cylinder_spec(), api.F64)
    assert np.array_equal(again.run(400, 10.0), gc), "force history not bitwise reproducible"


def test_tunnel_with_radiating_sides_and_kick(gpu, orc):
    """Towing tank (src/cases/tunnel.rs:52-67): radiating sides, uniform inlet, absorbing
    outlet, ramp and the body kick (moving-wall links with a spin velocity)."""
    spec = tunnel_spec(d=10, u=0.1, upstream=4.0, downstream=8.0, height=6.0)
    steps = 900  # through the ramp (500 steps) into the kick (from step 500)
    g = api.Channel(gpu, spec, api.F64)
    o = api.Channel(orc, tunnel_spec(d=10, u=0.1, upstream=4.0, downstream=8.0, height=6.0), 2)
    gc = g.run(steps, 10.0)
    oc = o.run(steps, 10.0)
    assert_pops(g.sim.read(), o.sim.read(), "tunnel")
    assert_force(gc, oc, "tunnel cd/cl")
    assert abs(g.sim.outlet_density_offset - o.sim.outlet_density_offset) == 0.0


def test_vorticity_is_fluid_periodic_match_oracle(gpu, orc):
    """Sim::vorticity, is_fluid and periodic (src/sim.rs:581-620, 381-383, 670-672) on the GPU
    state after 400 steps of the small-cylinder channel: the same FP64 populations give the same
    float vorticity bit for bit; the FP32-storage state gives it within its storage error."""
    g = api.Channel(gpu, small_cylinder_spec(), api.F64)
    o = api.Channel(orc, small_cylinder_spec(), 2)
    g.run(400, 10.0)
    o.run(400, 10.0)
    wg, wo = g.sim.vorticity(), o.sim.vorticity()
    assert wg.dtype == np.float32 and wg.shape == (g.sim.ny, g.sim.nx)
    assert np.array_equal(wg, wo)
    assert np.max(np.abs(wo)) > 1e-3  # the wake has vorticity
    for x, y in [(0, 0), (1, 1), (g.sim.nx // 2, g.sim.ny // 2), (g.sim.nx - 1, 5), (40, 20)]:
        assert g.sim.is_fluid(x, y) == o.sim.is_fluid(x, y)
    assert g.sim.periodic() == o.sim.periodic()
    f32 = api.Channel(gpu, small_cylinder_spec(), api.F32)
    f32.run(400, 10.0)
    assert np.max(np.abs(f32.sim.vorticity() - wo)) < 1e-5
