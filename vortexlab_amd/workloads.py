"""Seeded synthetic workloads: random star-polygon bodies in towing-tank channels.

A case is the reference's towing tank (src/cases/tunnel.rs:52-67: uniform inflow between
radiating side boundaries, absorbing pressure outlet, a 5-unit inflow ramp and a 2-unit body
kick) around a star-shaped polygon.  The generator fixes every random choice from one seed,
so a batch is reproducible.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import api


def star_polygon(rng: np.random.Generator, d_nominal: float, cx: float, cy: float,
                 n_min: int = 8, n_max: int = 24, r_min: float = 0.6, r_max: float = 1.0) -> np.ndarray:
    """Star-shaped polygon: n in [n_min, n_max] vertices at evenly spaced angles (random
    rotation), radius of each vertex uniform in [r_min, r_max] x d_nominal / 2."""
    n = int(rng.integers(n_min, n_max + 1))
    phase = rng.uniform(0.0, 2.0 * np.pi)
    ang = phase + 2.0 * np.pi * np.arange(n) / n
    rad = rng.uniform(r_min, r_max, size=n) * 0.5 * d_nominal
    return np.stack([cx + rad * np.cos(ang), cy + rad * np.sin(ang)], axis=1)


@dataclass
class Case:
    spec: api.ChannelSpec
    steps: int
    d: float
    re: float
    outline: np.ndarray


def tank_case(nx: int, ny: int, d: float, re: float, u: float, collision: int, outline_fn, units: float = 20.0,
              ramp_units: float = 5.0, kick_units: float = 2.0) -> Case:
    """Towing tank of nx x ny cells with the body centred a quarter of the way downstream."""
    cx = 0.5 + round(0.25 * (nx - 2))
    cy = 0.5 * (ny - 1)
    outline = outline_fn(cx, cy)
    unit = d / u
    spec = api.ChannelSpec(nx=nx, ny=ny, inflow=api.INFLOW_UNIFORM_OPEN, u=u, nu=u * d / re,
                           body=api.polygon(outline), collision=collision, magic=3.0 / 16.0,
                           ramp_steps=int(ramp_units * unit),
                           kick=api.Kick(cx, cy, 0.4 * u / d, int(ramp_units * unit), int(kick_units * unit)))
    return Case(spec=spec, steps=int(units * unit), d=d, re=re, outline=outline)


def sweep_cases(seed: int, n: int = 512, w=(200, 1400), h=(80, 600), d=(10.0, 60.0), re=(40.0, 250.0),
                u: float = 0.05, units: float = 20.0) -> list:
    """The batched sweep: n cases with grid width / height, nominal diameter and Re drawn
    uniformly, BGK or TRT (Lambda 3/16) with equal odds, `units` convective units each."""
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(n):
        nx = int(rng.integers(w[0], w[1] + 1))
        ny = int(rng.integers(h[0], h[1] + 1))
        dn = float(rng.uniform(*d))
        r = float(rng.uniform(*re))
        coll = api.BGK if rng.random() < 0.5 else api.TRT
        sub = np.random.default_rng(rng.integers(1 << 62))
        cases.append(tank_case(nx, ny, dn, r, u, coll, lambda cx, cy: star_polygon(sub, dn, cx, cy), units=units))
    return cases
