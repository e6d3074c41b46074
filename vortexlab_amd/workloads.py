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


def build_batch(be, cases: list, precision: int = api.F32):
    """All cases in one batch, built (set-up, packing, autotuning) but not run; returns (batch,
    set-up seconds).  The caller closes the batch."""
    import time
    t0 = time.perf_counter()
    b = api.Batch(be, precision)
    for c in cases:
        b.add(c.spec, c.steps, c.d)
    b.build()
    return b, time.perf_counter() - t0


def run_batch(be, cases: list, precision: int = api.F32):
    """All cases in one batch; returns (batch, set-up seconds, run seconds).  The caller closes
    the batch."""
    import time
    b, t_build = build_batch(be, cases, precision)
    t0 = time.perf_counter()
    b.run()
    return b, t_build, time.perf_counter() - t0


def run_phases(be, cases: list, precision: int, pieces: int):
    """Runs the batch with kernel timing on, in `pieces` equal spans of steps, and prints per
    span the active cases, node updates and the boundary / bulk / whole-step time per step."""
    import sys
    import time
    b = api.Batch(be, precision)
    for c in cases:
        b.add(c.spec, c.steps, c.d)
    b.build()
    info = b.info()
    at = np.array(info.autotune_ms)
    print(f"[phase] autotune trial-step ms (rows 64x4/128x2/256x1; cols one-per-tile/4/8/16 per CU):\n{np.round(at, 3)}",
          file=sys.stderr, flush=True)
    steps = np.array([c.steps for c in cases])
    fluid = np.array([b.case_fluid_nodes(k) for k in range(len(cases))])
    end = int(info.max_steps) + 1
    span = -(-end // pieces)
    t = 0
    while t < end:
        n = min(span, end - t)
        b.set_timing(True)
        w0 = time.perf_counter()
        b.run(n)
        wall = time.perf_counter() - w0
        bnd, bulk, step, ns = b.timing()
        lo, hi = t, t + n
        active = int(np.sum(steps > lo))
        lu = int(np.sum(fluid * np.clip(np.minimum(steps, hi) - lo, 0, None)))
        print(f"[phase] steps {lo}-{hi}: {active} cases active, {lu / 1e9:.1f} G updates; per step "
              f"boundary {bnd / ns * 1e3:.1f} us, bulk {bulk / ns * 1e3:.1f} us, step {step / ns * 1e3:.1f} us; "
              f"{lu / (step * 1e-3) / 1e9:.1f} GLUPS on the step time, {lu / wall / 1e9:.1f} on the wall clock",
              file=sys.stderr, flush=True)
        t += n
    b.set_timing(False)
    return b


def main(argv=None):
    """Runs one seeded sweep as a batch and prints its throughput as one JSON line."""
    import argparse
    import json
    import sys
    from . import backend
    ap = argparse.ArgumentParser(description="batched towing-tank sweep on one GPU")
    ap.add_argument("--cases", type=int, default=512)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--units", type=float, default=20.0, help="convective units per case")
    ap.add_argument("--precision", choices=["f32", "f16", "f64"], default="f32")
    ap.add_argument("--phases", type=int, default=0,
                    help="if > 0: run with kernel timing on, in this many pieces, one line per piece")
    args = ap.parse_args(argv)
    prec = {"f64": api.F64, "f32": api.F32, "f16": api.F16}[args.precision]
    cases = sweep_cases(args.seed, args.cases, units=args.units)
    if args.phases > 0:
        b = run_phases(backend(), cases, prec, args.phases)
        t_build, t_run = 0.0, b.info().last_run_seconds
    else:
        b, t_build, t_run = run_batch(backend(), cases, prec)
    info = b.info()
    status = [b.result(k)[2] for k in range(len(cases))]
    out = {"cases": len(cases), "precision": args.precision, "fluid_nodes": int(info.fluid_nodes),
           "lattice_updates": int(info.lattice_updates), "max_steps": int(info.max_steps),
           "setup_s": round(t_build, 3), "run_s": round(t_run, 3), "cases_per_s": round(len(cases) / t_run, 3),
           "glups": round(info.lattice_updates / t_run / 1e9, 2), "diverged": int(sum(1 for s in status if s)),
           "tile": f"{info.tile_w}x{info.tile_h}", "grid_bulk": int(info.grid_bulk),
           "grid_boundary": int(info.grid_boundary), "boundary_tiles": int(info.boundary_tiles),
           "bulk_tiles": int(info.bulk_tiles), "pass_steps": int(info.pass_steps),
           "pass_tasks": int(info.pass_tasks), "pass_nodes": int(info.pass_nodes),
           "listed_nodes": int(info.listed_nodes), "device_gb": round(info.device_bytes / 1e9, 2)}
    b.close()
    print(json.dumps(out), file=sys.stdout, flush=True)


if __name__ == "__main__":
    main()
