"""Two open-boundary invariants of the reference's solver tests, restated on the oracle (CPU) and
on the GPU library in FP64 with the reference's own tolerances:

* developed channel flow between a parabolic velocity inlet and a constant-pressure outlet
  (tests/solver.rs:215-253): the mass-flux profile stays parabolic to 2e-4 of u_mean, the
  pressure falls at the Poiseuille rate to 1e-3, the outlet density stays within 6 u_max^2 of 1,
  the flux in equals the flux out to 1e-6, and the Simpson-sampled inlet passes exactly
  u_mean * H per step (1e-6);
* the absorbing outlet (tests/solver.rs:294-327): after an impulsive start the drag ripples by
  under 1 % with the absorbing outlet, and by more than 10x that with the reflecting one.

On the GPU the FP64 path is also bit-identical to the oracle on both set-ups (state hash)."""
import numpy as np
import pytest

from vortexlab_amd import api
from cases import open_channel

CS2 = 1.0 / 3.0


def viscosity(tau):
    """src/lattice.rs:53-55."""
    return CS2 * (tau - 0.5)


def developed_channel(be, arg):
    """tests/solver.rs:215-221: 82 columns, 20 fluid rows, u_mean 0.005, tau 0.8, 40000 steps."""
    nx, rows, u_mean, tau = 82, 20, 0.005, 0.8
    s = open_channel(be, arg, nx, rows, "parabolic", u_mean, tau)
    s.step(40000)
    return s, nx, rows, u_mean, tau


def check_developed_channel(s, nx, rows, u_mean, tau):
    """The assertions of tests/solver.rs:222-252 on a stepped channel."""
    rho, ux, uy = s.macros_all()
    h = float(rows)

    def flux(x):
        return float(sum(rho[y, x] * ux[y, x] for y in range(1, rows + 1))) / h / (1.0 + 0.5 / (h * h))

    mean_flux = flux(40)
    worst = 0.0
    for y in range(1, rows + 1):
        exact = 6.0 * mean_flux * (y - 0.5) * (h - y + 0.5) / (h * h)
        worst = max(worst, abs(rho[y, 40] * ux[y, 40] - exact), abs(uy[y, 40]))
    assert worst < 2e-4 * u_mean, f"profile deviates by {worst}"
    slope = CS2 * (rho[10, 60] - rho[10, 20]) / 40.0
    exact = -12.0 * viscosity(tau) * mean_flux / (h * h)
    assert abs(slope / exact - 1.0) < 1e-3, f"pressure gradient {slope} vs {exact}"
    u_max = 1.5 * u_mean
    assert abs(rho[10, nx - 2] - 1.0) < 6.0 * u_max * u_max
    assert abs(flux(2) / flux(nx - 3) - 1.0) < 1e-6
    crossing = mean_flux * (1.0 + 0.5 / (h * h))
    assert abs(crossing / u_mean - 1.0) < 1e-6, f"mass flux {crossing}"
    # the check is not vacuous: the density does fall along the channel
    assert rho[10, 20] - rho[10, 60] > 1e-6


def outlet_spec():
    """tests/solver.rs:298-308: impulsive start (ramp of one step) past a small circle."""
    return api.ChannelSpec(nx=302, ny=42, inflow=api.INFLOW_UNIFORM, u=0.05, nu=0.02,
                           body=api.circle(60.5, 20.5, 4.0), collision=api.TRT, magic=3.0 / 16.0, ramp_steps=1)


def drag_ripple(be, arg, absorbing):
    """tests/solver.rs:297-321: peak-to-peak drag over steps 6000..7200, relative to its peak."""
    ch = api.Channel(be, outlet_spec(), arg)
    ch.set_absorbing(absorbing)
    ch.run(6000, 8.0)
    cd = ch.run(1200, 8.0)[:, 0]
    h = ch.sim.state_hash()
    ch.close()
    return (cd.max() - cd.min()) / cd.max(), h


def test_developed_channel_flow_oracle(orc):
    check_developed_channel(*developed_channel(orc, 2))


def test_absorbing_outlet_removes_sound_waves_oracle(orc):
    reflecting, _ = drag_ripple(orc, 2, False)
    absorbing, _ = drag_ripple(orc, 2, True)
    assert absorbing < 0.01, f"drag still ripples by {absorbing} with the absorbing outlet"
    assert reflecting > 10.0 * absorbing, f"reflecting {reflecting} vs absorbing {absorbing}"


@pytest.mark.gpu
def test_developed_channel_flow_gpu(gpu, orc):
    g = developed_channel(gpu, api.F64)
    check_developed_channel(*g)
    o = developed_channel(orc, 2)[0]
    assert g[0].state_hash() == o.state_hash()


@pytest.mark.gpu
def test_absorbing_outlet_removes_sound_waves_gpu(gpu, orc):
    reflecting, hr = drag_ripple(gpu, api.F64, False)
    absorbing, ha = drag_ripple(gpu, api.F64, True)
    assert absorbing < 0.01, f"drag still ripples by {absorbing} with the absorbing outlet"
    assert reflecting > 10.0 * absorbing, f"reflecting {reflecting} vs absorbing {absorbing}"
    # the FP64 path is the oracle's, bit for bit, with either outlet
    assert hr == drag_ripple(orc, 2, False)[1]
    assert ha == drag_ripple(orc, 2, True)[1]
