"""Test cases stated once and built on any backend (GPU library or CPU oracle).

They restate the set-ups of the reference's own tests (tests/solver.rs, tests/validation.rs,
src/cases/*.rs) through the mirrored Builder/Channel API.
"""
import math

import numpy as np

from vortexlab_amd import api

TRT = (api.TRT, 3.0 / 16.0)
TRT14 = (api.TRT, 0.25)
BGK = (api.BGK, 0.0)
MODELS = {"bgk": BGK, "trt": TRT, "trt14": TRT14}


def swirl_field(nx, ny):
    """tests/solver.rs:14-19."""
    def f(x, y):
        a = 2.0 * math.pi * x / nx
        b = 2.0 * math.pi * y / ny
        return (1.0 + 0.02 * math.sin(a + 2.0 * b), 0.05 * math.sin(b) + 0.01,
                -0.04 * math.cos(a) + 0.02 * math.sin(a - b))
    return f


def swirl_arrays(nx, ny):
    f = swirl_field(nx, ny)
    rho = np.empty((ny, nx))
    ux = np.empty((ny, nx))
    uy = np.empty((ny, nx))
    for y in range(ny):
        for x in range(nx):
            rho[y, x], ux[y, x], uy[y, x] = f(x, y)
    return rho, ux, uy


def periodic_box(be, model, arg, nx=48, ny=36, tau=0.62, accel=(0.0, 0.0)):
    """tests/solver.rs:22-61: periodic box with the swirl field."""
    b = api.Builder(be, nx, ny)
    b.periodic_x = b.periodic_y = True
    b.tau = tau
    b.collision, b.magic = MODELS[model]
    b.accel = accel
    s = b.build(arg)
    s.init_arrays(*swirl_arrays(nx, ny))
    return s


def closed_box(be, model, arg):
    """tests/solver.rs:64-80."""
    b = api.Builder(be, 34, 26)
    b.tau = 0.58
    b.collision, b.magic = MODELS[model]
    wall = b.add_static_wall()
    b.fill_rect(0, 33, 0, 25, wall)
    b.fill_rect(1, 32, 1, 24, 0)
    s = b.build(arg)
    s.init_arrays(*swirl_arrays(34, 26))
    return s


def couette(be, model, arg, frame, shear):
    """tests/solver.rs:86-102 (set-up; returns the sim before stepping)."""
    nx, rows = 6, 20
    b = api.Builder(be, nx, rows + 2)
    b.periodic_x = True
    b.tau = 0.9
    b.collision, b.magic = MODELS[model]
    low = b.add_wall(velocity=api.vel_const(frame, 0.0), fgroup=0)
    top = b.add_wall(velocity=api.vel_const(frame + shear, 0.0), fgroup=1)
    b.fill_rect(0, nx - 1, 0, 0, low)
    b.fill_rect(0, nx - 1, rows + 1, rows + 1, top)
    s = b.build(arg)
    ux = np.array([[frame + shear * (y - 0.5) / rows for _ in range(nx)] for y in range(rows + 2)])
    s.init_arrays(1.0, ux, 0.0)
    return s


def comoving_body(be, model, arg, shape, u):
    """tests/solver.rs:126-159: a body moving with a uniform stream."""
    b = api.Builder(be, 40, 32)
    b.periodic_x = b.periodic_y = True
    b.tau = 0.6
    b.collision, b.magic = MODELS[model]
    body = b.add_wall(velocity=api.vel_const(*u), fgroup=0, shape=shape)
    b.fill_shape(body)
    s = b.build(arg)
    s.init_arrays(1.0, u[0], u[1])
    return s


COMOVING_SHAPES = {
    "circle": lambda: api.circle(20.3, 15.7, 6.4),
    "notched": lambda: api.section("double-recessed", 0.1, 20.5, 16.5, 12.0, 0.3),
}


def open_channel(be, arg, nx, rows, kind, u, tau):
    """tests/solver.rs:176-199: an empty channel between a velocity inlet and a pressure outlet."""
    b = api.Builder(be, nx, rows + 2)
    b.tau = tau
    b.collision, b.magic = TRT
    h = float(rows)
    if kind == "parabolic":
        prof = lambda y: 6.0 * u * (y - 0.5) * (h - (y - 0.5)) / (h * h)
        inlet = b.add_wall(velocity=api.vel_fn(lambda _x, y: (prof(y), 0.0)))
        walls = b.add_static_wall()
    else:
        prof = lambda y: u
        inlet = b.add_wall(velocity=api.vel_const(u, 0.0))
        walls = b.add_wall(velocity=api.vel_const(u, 0.0))
    outlet = b.add_pressure(1.0)
    b.fill_rect(0, 0, 0, rows + 1, inlet)
    b.fill_rect(nx - 1, nx - 1, 0, rows + 1, outlet)
    b.fill_rect(0, nx - 1, 0, 0, walls)
    b.fill_rect(0, nx - 1, rows + 1, rows + 1, walls)
    s = b.build(arg)
    ux = np.array([[prof(float(y)) for _ in range(nx)] for y in range(rows + 2)])
    s.init_arrays(1.0, ux, 0.0)
    return s


def cylinder_rig_spec(n, u_mean=0.05, re=100.0, ramp_units=8.0, model=TRT):
    """src/cases/cylinder.rs:35-50 (the published-hash bench uses n=40/80, TRT 3/16)."""
    nn = float(n)
    xc = yc = 0.5 + 2.0 * nn
    return api.ChannelSpec(nx=22 * n + 2, ny=41 * n // 10 + 2, inflow=api.INFLOW_PARABOLIC, u=u_mean,
                           nu=u_mean * nn / re, body=api.circle(xc, yc, 0.5 * nn), collision=model[0],
                           magic=model[1], ramp_steps=int(ramp_units * nn / u_mean))


def small_cylinder_spec(ramp_steps=100):
    """tests/solver.rs:255-268 (the thread-count determinism case)."""
    return api.ChannelSpec(nx=222, ny=43, inflow=api.INFLOW_PARABOLIC, u=0.05, nu=0.05 * 10.0 / 100.0,
                           body=api.circle(20.5, 20.5, 5.0), collision=api.TRT, magic=3.0 / 16.0,
                           ramp_steps=ramp_steps)


def tunnel_spec(d=20, re=100.0, u=0.1, upstream=6.0, downstream=12.0, height=10.0, model=TRT,
                open_sides=True, corner="square", b=0.0):
    """src/cases/tunnel.rs:35-67: towing tank with radiating (or moving) sides, ramp and kick."""
    dd = float(d)
    nx = int(round((upstream + downstream) * dd)) + 2
    ny = int(round(height * dd)) + 2
    cx = 0.5 + round(upstream * dd)
    cy = 0.5 + 0.5 * round(height * dd)
    unit = dd / u
    body = api.section(corner, b, cx, cy, dd, 0.0)
    return api.ChannelSpec(nx=nx, ny=ny, inflow=api.INFLOW_UNIFORM_OPEN if open_sides else api.INFLOW_UNIFORM,
                           u=u, nu=u * dd / re, body=body, collision=model[0], magic=model[1],
                           ramp_steps=int(5.0 * unit),
                           kick=api.Kick(cx, cy, 0.4 * u / dd, int(5.0 * unit), int(2.0 * unit)))


def wake_spec(steps_ramp=2000, kick_steps=800):
    """Wake case on 662x302: BGK, uniform inflow u=0.05, circle D=20 at (160,151), Re=150,
    radiating ("free-slip") sides, absorbing pressure outlet; 5-unit ramp and a 2-unit kick."""
    u, d, re = 0.05, 20.0, 150.0
    return api.ChannelSpec(nx=662, ny=302, inflow=api.INFLOW_UNIFORM_OPEN, u=u, nu=u * d / re,
                           body=api.circle(160.0, 151.0, 10.0), collision=api.BGK, magic=0.0,
                           ramp_steps=steps_ramp,
                           kick=api.Kick(160.0, 151.0, 0.4 * u / d, steps_ramp, kick_steps))


def wake_662x302_spec():
    """The 662x302 wake case at its full definition (ramp 2000 steps = 5 D/U, kick 800 steps)."""
    return wake_spec(2000, 800)


def fnv1a(data: bytes) -> int:
    """64-bit FNV-1a of a byte string (offset basis 0xcbf29ce484222325, prime 0x100000001b3)."""
    h = 0xcbf29ce484222325
    for b in data:
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h
