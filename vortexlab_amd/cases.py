"""Case drivers of the reference's `cases` module, on the mirrored API.

Each driver states a reference set-up once and runs it on any backend: the MI355X library
(``arg`` = storage precision) or, in tests, the CPU oracle (``arg`` = thread count).  The
step loops, stopping rules, divergence checks and post-processing follow the reference line
by line, so a driver's figures can be compared with the reference's published tables.

* Poiseuille flow (src/cases/poiseuille.rs:41-102)
* Lid-driven cavity against Ghia's tables (src/cases/cavity.rs, src/suite.rs:95-179)
* Schaefer-Turek cylinder rig, 2D-1 / 2D-2 and the bench (src/cases/cylinder.rs:35-232)
* Towing tank (src/cases/tunnel.rs:35-92)
* Output rows in the reference's CSV formats (src/main.rs:197-208, 448-476)
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import api

CS2 = 1.0 / 3.0
TRT = (api.TRT, 3.0 / 16.0)


def viscosity(tau: float) -> float:
    """lattice.rs:48-50."""
    return (tau - 0.5) * CS2


def tau_for(nu: float) -> float:
    """lattice.rs:53-55."""
    return nu / CS2 + 0.5


def _serial_sum(x) -> float:
    """Left-to-right sum, as the reference's iterator sums."""
    s = 0.0
    for v in x:
        s += float(v)
    return s


def _mean(x) -> float:
    """signal.rs:61-63."""
    return _serial_sum(x) / len(x)


def sample_at(x, pos: float) -> float:
    """signal.rs:143-147: linear interpolation at a fractional index."""
    i = min(int(math.floor(pos)), len(x) - 2)
    f = pos - i
    return x[i] * (1.0 - f) + x[i + 1] * f


# ------------------------------------------------------------------------------------------
# Poiseuille flow (src/cases/poiseuille.rs)
# ------------------------------------------------------------------------------------------
@dataclass
class PoiseuilleOutcome:
    rows: int
    q: float
    l2: float
    steps: int
    force_balance: float
    profile: list = field(default_factory=list)


def poiseuille_run(be: api.Backend, rows: int, tau: float, collision=TRT, q: float = 0.5, arg: int = api.F64,
                   max_steps: int = 4_000_000) -> PoiseuilleOutcome:
    """poiseuille::run (poiseuille.rs:41-89): a periodic channel of `rows` fluid rows driven by
    a body force, run until the profile stops changing; `q` places both walls q cells beyond
    the outermost fluid rows (0.5: halfway bounce-back, otherwise interpolated)."""
    nx, ny = 4, rows + 2
    y0, y1 = 1.0 - q, rows + q
    width = y1 - y0
    nu = viscosity(tau)
    umax = 0.1 * 16.0 / rows
    g = 8.0 * nu * umax / (width * width)
    b = api.Builder(be, nx, ny)
    b.periodic_x = True
    b.tau = tau
    b.collision, b.magic = collision
    b.accel = (g, 0.0)
    shape = None if q == 0.5 else api.outside_band(y0, y1)
    wall = b.add_wall(fgroup=0, shape=shape)
    b.fill_rect(0, nx - 1, 0, 0, wall)
    b.fill_rect(0, nx - 1, ny - 1, ny - 1, wall)
    sim = b.build(arg)

    def exact(y):
        return g / (2.0 * nu) * (y - y0) * (y1 - y)

    prev = [0.0] * ny
    check = 200
    t = 0
    while True:
        sim.step(check)
        t += check
        _, ux, _ = sim.macros_all()
        change = 0.0
        for y in range(1, rows + 1):
            u = float(ux[y, 1])
            change = max(change, abs(u - prev[y]))
            prev[y] = u
        if change < 1e-13 * umax or t > max_steps:
            break
    _, ux, _ = sim.macros_all()
    num = den = 0.0
    profile = []
    for y in range(1, rows + 1):
        u, e = float(ux[y, 1]), exact(y)
        num += (u - e) * (u - e)
        den += e * e
        profile.append(((y - y0) / width, u / umax, e / umax))
    driving = g * sim.total_mass()
    balance = float(sim.force[0][0]) / driving
    out = PoiseuilleOutcome(rows, q, math.sqrt(num / den), int(sim.info().t), balance, profile)
    sim.close()
    return out


def observed_orders(outcomes) -> list:
    """poiseuille::observed_orders (poiseuille.rs:93-102)."""
    out = []
    for a, b in zip(outcomes, outcomes[1:]):
        h0 = a.rows - 1.0 + 2.0 * a.q
        h1 = b.rows - 1.0 + 2.0 * b.q
        out.append(math.log(a.l2 / b.l2) / math.log(h1 / h0))
    return out


# ------------------------------------------------------------------------------------------
# Lid-driven cavity (src/cases/cavity.rs, src/suite.rs:95-179)
# ------------------------------------------------------------------------------------------
@dataclass
class CavityOutcome:
    n: int
    re: float
    steps: int
    converged: bool
    seconds: float
    u_vertical: list   # (y, u / U_lid) on the vertical centreline, walls included
    v_horizontal: list  # (x, v / U_lid) on the horizontal centreline, walls included


def cavity_run(be: api.Backend, n: int, re: float, u_lid: float = 0.1, collision=TRT, tol: float = 1e-9,
               max_steps: int = 3_000_000, arg: int = api.F64) -> CavityOutcome:
    """cavity::run (cavity.rs:23-79): n x n fluid cells, no-slip walls half a cell outside, the
    top one sliding at `u_lid`; runs until the largest velocity change over 1000 steps is
    below `tol` x u_lid."""
    start = time.time()
    b = api.Builder(be, n + 2, n + 2)
    b.tau = tau_for(u_lid * n / re)
    b.collision, b.magic = collision
    wall = b.add_static_wall()
    lid = b.add_wall(velocity=api.vel_const(u_lid, 0.0))
    b.fill_rect(0, n + 1, 0, n + 1, wall)
    b.fill_rect(1, n, n + 1, n + 1, lid)
    b.fill_rect(1, n, 1, n, 0)
    sim = b.build(arg)

    def snapshot():
        _, ux, uy = sim.macros_all()
        return ux[1:n + 1, 1:n + 1].copy(), uy[1:n + 1, 1:n + 1].copy()

    prev = snapshot()
    converged = False
    t = 0
    while t < max_steps:
        sim.step(1000)
        t += 1000
        now = snapshot()
        change = float(max(np.max(np.abs(now[0] - prev[0])), np.max(np.abs(now[1] - prev[1]))))
        prev = now
        if math.isnan(change):
            break
        if change < tol * u_lid:
            converged = True
            break
    _, ux, uy = sim.macros_all()
    sim.close()
    c0, c1 = ((n + 1) // 2, (n + 1) // 2) if n % 2 == 1 else (n // 2, n // 2 + 1)
    u_vertical, v_horizontal = [(0.0, 0.0)], [(0.0, 0.0)]
    for k in range(1, n + 1):
        sk = (k - 0.5) / n
        u_vertical.append((sk, 0.5 * (float(ux[k, c0]) + float(ux[k, c1])) / u_lid))
        v_horizontal.append((sk, 0.5 * (float(uy[c0, k]) + float(uy[c1, k])) / u_lid))
    u_vertical.append((1.0, 1.0))
    v_horizontal.append((1.0, 0.0))
    return CavityOutcome(n, re, t, converged, time.time() - start, u_vertical, v_horizontal)


def interpolate(profile, s: float) -> float:
    """cavity::interpolate (cavity.rs:83-99): cubic Lagrange through the four nearest samples."""
    n = len(profile)
    k = 0
    while k < n and profile[k][0] < s:
        k += 1
    k = min(max(k, 2), n - 2)
    pts = profile[k - 2:k + 2]
    out = 0.0
    for i in range(4):
        w = 1.0
        for j in range(4):
            if i != j:
                w *= (s - pts[j][0]) / (pts[i][0] - pts[j][0])
        out += w * pts[i][1]
    return out


def refine_extremum(p, k: int):
    """suite.rs:95-108: (position, value) of the parabola through samples k - 1, k, k + 1."""
    if k == 0 or k + 1 >= len(p):
        return p[k]
    (x0, y0), (x1, y1), (x2, y2) = p[k - 1], p[k], p[k + 1]
    d1 = (y1 - y0) / (x1 - x0)
    d2 = (y2 - y1) / (x2 - x1)
    c = (d2 - d1) / (x2 - x0)
    if c == 0.0:
        return p[k]
    x = 0.5 * (x0 + x1) - 0.5 * d1 / c
    return x, y0 + d1 * (x - x0) + c * (x - x0) * (x - x1)


@dataclass
class CavityComparison:
    max_dev_u: float
    max_dev_v: float
    rms_dev: float
    u_min: tuple  # (position, value)
    v_max: tuple
    v_min: tuple


def cavity_compare(o: CavityOutcome, ghia_u, ghia_v) -> CavityComparison:
    """suite::cavity_run's comparison (suite.rs:137-158): deviations at the interior points of
    Ghia's tables (lists of (coordinate, value)) and the refined centreline extrema."""
    max_u = max_v = sq = 0.0
    count = 0
    for table, profile in ((ghia_u, o.u_vertical), (ghia_v, o.v_horizontal)):
        for sk, r in table:
            if sk == 0.0 or sk == 1.0:
                continue
            d = interpolate(profile, sk) - r
            if table is ghia_u:
                max_u = max(max_u, abs(d))
            else:
                max_v = max(max_v, abs(d))
            sq += d * d
            count += 1

    def arg(p, sign):
        best = 0
        for i in range(1, len(p)):
            if sign * p[i][1] > sign * p[best][1]:
                best = i
        return best

    return CavityComparison(max_u, max_v, math.sqrt(sq / count), refine_extremum(o.u_vertical, arg(o.u_vertical, -1.0)),
                            refine_extremum(o.v_horizontal, arg(o.v_horizontal, 1.0)),
                            refine_extremum(o.v_horizontal, arg(o.v_horizontal, -1.0)))


# ------------------------------------------------------------------------------------------
# Schaefer-Turek cylinder rig (src/cases/cylinder.rs)
# ------------------------------------------------------------------------------------------
def cylinder_rig_spec(n: int, u_mean: float = 0.05, re: float = 100.0, ramp_units: float = 8.0,
                      collision=TRT) -> api.ChannelSpec:
    """rig (cylinder.rs:35-50): channel 22n x 4.1n with a parabolic inflow and a circle of
    diameter n at (2n + 0.5, 2n + 0.5)."""
    if n < 10 or n % 10:
        raise ValueError("cells per diameter must be a multiple of 10")
    nn = float(n)
    xc = yc = 0.5 + 2.0 * nn
    return api.ChannelSpec(nx=22 * n + 2, ny=41 * n // 10 + 2, inflow=api.INFLOW_PARABOLIC, u=u_mean,
                           nu=u_mean * nn / re, body=api.circle(xc, yc, 0.5 * nn), collision=collision[0],
                           magic=collision[1], ramp_steps=int(ramp_units * nn / u_mean))


def probe_density(rho: np.ndarray, solid: np.ndarray, x: float, y: float) -> float:
    """Sim::probe_density (sim.rs:624-643): bilinear density over the non-solid corners of the
    cell holding (x, y), weights renormalised."""
    ny, nx = rho.shape
    i, j = math.floor(x), math.floor(y)
    fx, fy = x - i, y - j
    s = ws = 0.0
    for di, dj, w in ((0, 0, (1.0 - fx) * (1.0 - fy)), (1, 0, fx * (1.0 - fy)), (0, 1, (1.0 - fx) * fy),
                      (1, 1, fx * fy)):
        xi, yj = i + di, j + dj
        if xi < 0 or yj < 0 or xi >= nx or yj >= ny:
            continue
        if not solid[yj, xi] and w > 0.0:
            s += w * float(rho[yj, xi])
            ws += w
    return s / ws if ws > 0.0 else math.nan


def probe_density_at_wall(rho, solid, x, y, nx, ny) -> float:
    """Sim::probe_density_at_wall (sim.rs:647-651): quadratic extrapolation from 1, 2 and 3
    lattice units along the outward normal."""
    def p(d):
        return probe_density(rho, solid, x + d * nx, y + d * ny)
    return 3.0 * p(1.0) - 3.0 * p(2.0) + p(3.0)


def _solid_mask(spec: api.ChannelSpec) -> np.ndarray:
    """Cells of the rig's circle: centre strictly inside (Circle::inside, geometry.rs:22-27, as
    Builder::fill_shape paints them).  The probes sit next to the body, far from the walls."""
    b = spec.body
    cx, cy, r = b.args
    y, x = np.mgrid[0:spec.ny, 0:spec.nx]
    return (x - cx) ** 2 + (y - cy) ** 2 < r * r


@dataclass
class Steady:
    n: int
    cd: float
    cl: float
    dp: float
    la: float
    residual: float
    steps: int
    converged: bool
    seconds: float


def _pressure_drop(rho, solid, n, xc, yc, u_ref):
    """Rig::pressure_drop (cylinder.rs:57-62), in units of rho U^2."""
    r = 0.5 * n
    front = probe_density_at_wall(rho, solid, xc - r, yc, -1.0, 0.0)
    back = probe_density_at_wall(rho, solid, xc + r, yc, 1.0, 0.0)
    return CS2 * (front - back) / (u_ref * u_ref)


def _recirculation_length(ux, n, xc, yc):
    """Rig::recirculation_length (cylinder.rs:66-82), in diameters."""
    j0 = int(math.floor(yc))
    j1 = j0 + 1

    def u(x):
        return 0.5 * (float(ux[j0, x]) + float(ux[j1, x]))

    rear = xc + 0.5 * n
    start = int(math.ceil(rear)) + 1
    if u(start) >= 0.0:
        return 0.0
    for x in range(start, ux.shape[1] - 2):
        a, b = u(x), u(x + 1)
        if a < 0.0 and b >= 0.0:
            return (x + a / (a - b) - rear) / n
    return math.nan


def run_2d1(be: api.Backend, n: int, u_mean: float = 0.04, max_units: float = 300.0, tol: float = 1e-6,
            arg: int = api.F64, collision=TRT) -> Steady:
    """Case 2D-1, Re = 20 (cylinder.rs:100-143): windows of 5 convective units until the drag
    varies by less than `tol` over a window and its window mean has stopped drifting."""
    start = time.time()
    spec = cylinder_rig_spec(n, u_mean, 20.0, 8.0, collision)
    ch = api.Channel(be, spec, arg)
    nn = float(n)
    unit = int(round(nn / u_mean))
    window = 5 * unit
    prev_mean, residual, converged = math.nan, math.nan, False
    cd = cl = 0.0
    t = 0
    while t < max_units * unit:
        coef = ch.run(window, nn)
        t += window
        lo, hi = float(np.min(coef[:, 0])), float(np.max(coef[:, 0]))
        cd = _serial_sum(coef[:, 0]) / window
        cl = _serial_sum(coef[:, 1]) / window
        residual = (hi - lo) / abs(cd)
        if not math.isfinite(cd):
            break
        if residual < tol and abs((cd - prev_mean) / cd) < tol:
            converged = True
            break
        prev_mean = cd
    rho, ux, _ = ch.sim.macros_all()
    solid = _solid_mask(spec)
    xc = yc = 0.5 + 2.0 * nn
    out = Steady(n, cd, cl, _pressure_drop(rho, solid, nn, xc, yc, u_mean) * 0.2 * 0.2,
                 _recirculation_length(ux, nn, xc, yc) * 0.1, residual, t, converged, time.time() - start)
    ch.close()
    return out


@dataclass
class Unsteady:
    n: int
    cd_max: float
    cd_min: float
    cl_max: float
    cl_min: float
    st: float
    st_crossings: float
    peak_spread: float
    periods: int
    steps: int
    seconds: float
    cd: np.ndarray
    cl: np.ndarray
    dt: float
    window_start: int


def run_2d2(be: api.Backend, n: int, u_mean: float = 0.05, total_units: float = 110.0, window_units: float = 20.0,
            arg: int = api.F64, collision=TRT) -> Unsteady:
    """Case 2D-2, Re = 100 (cylinder.rs:169-222): `total_units` convective units, the last
    `window_units` analysed.  Raises api.Diverged if the run blows up.  The pressure
    difference needs a density probe every step and is not computed."""
    start = time.time()
    spec = cylinder_rig_spec(n, u_mean, 100.0, 8.0, collision)
    ch = api.Channel(be, spec, arg)
    nn = float(n)
    unit = nn / u_mean
    steps = int(total_units * unit)
    coef = ch.run(steps, nn)  # raises api.Diverged on the first non-finite drag
    ch.close()
    dt = 1.0 / unit
    i0 = steps - min(int(window_units * unit), steps)
    cd, cl = coef[i0:, 0], coef[i0:, 1]
    f, _ = api.dominant_frequency(be, cl, dt)
    pk = api.peaks(be, cl)
    tops = pk[:, 1] if len(pk) else np.zeros(0)
    cl_max = max([-math.inf] + [float(v) for v in tops])
    spread = (cl_max - min(float(v) for v in tops)) / _mean(tops) if len(tops) else math.nan
    stc = api.crossing_frequency(be, cl, dt)
    return Unsteady(n, float(np.max(cd)), float(np.min(cd)), cl_max, float(np.min(cl)), f,
                    stc if stc is not None else math.nan, spread, len(tops), steps, time.time() - start,
                    coef[:, 0].copy(), coef[:, 1].copy(), dt, i0)


def cylinder_bench(be: api.Backend, n: int, steps: int, arg: int = api.F64):
    """cylinder::bench (cylinder.rs:224-232): `steps` steps of the 2D-2 set-up at u = 0.05,
    TRT; returns (channel, seconds).  The caller closes the channel."""
    ch = api.Channel(be, cylinder_rig_spec(n, 0.05, 100.0, 8.0), arg)
    t0 = time.perf_counter()
    ch.run(steps, float(n))
    return ch, time.perf_counter() - t0


def bench_row(ch, steps: int, seconds: float, threads: int = 1) -> list:
    """A row of the CLI's bench table (main.rs:452-474): threads, grid, fluid_nodes, steps,
    seconds, mlups, state_hash."""
    info = ch.sim.info()
    nodes = int(info.fluid_nodes)
    mlups = nodes * steps / seconds / 1e6
    return [str(threads), f"{info.nx}x{info.ny}", str(nodes), str(steps), f"{seconds:.3f}", f"{mlups:.1f}",
            f"{ch.sim.state_hash():016x}"]


BENCH_HEADER = ["threads", "grid", "fluid_nodes", "steps", "seconds", "mlups", "state_hash"]


def forces_csv(cd, cl, dt: float) -> str:
    """forces.csv as the CLI writes it (main.rs:197-208): t, cd, cl, at most ~5000 rows."""
    lines = ["t,cd,cl"]
    stride = max(len(cd) // 5000, 1)
    for i in range(0, len(cd), stride):
        lines.append(f"{i * dt:.4f},{cd[i]:.6f},{cl[i]:.6f}")
    return "\n".join(lines) + "\n"


# ------------------------------------------------------------------------------------------
# Towing tank (src/cases/tunnel.rs)
# ------------------------------------------------------------------------------------------
@dataclass
class Tunnel:
    """tunnel.rs:14-31: a body of width `d` cells in a tank `upstream + downstream` widths
    long and `height` widths high, stream speed `u`, Reynolds number `re`.  The defaults are
    the demo's tank (main.rs:245-256)."""
    d: int = 30
    re: float = 150.0
    u: float = 0.08
    upstream: float = 6.0
    downstream: float = 16.0
    height: float = 10.0
    collision: tuple = TRT
    open_sides: bool = True

    def grid(self):
        d = float(self.d)
        return int(round((self.upstream + self.downstream) * d)) + 2, int(round(self.height * d)) + 2

    def centre(self):
        d = float(self.d)
        return 0.5 + round(self.upstream * d), 0.5 + 0.5 * round(self.height * d)

    def steps_per_unit(self) -> float:
        return self.d / self.u

    def spec(self, body: api.ShapeSpec) -> api.ChannelSpec:
        """Tunnel::build (tunnel.rs:50-68): 5-unit inflow ramp, then a 2-unit twist of the body."""
        nx, ny = self.grid()
        cx, cy = self.centre()
        unit = self.steps_per_unit()
        return api.ChannelSpec(nx=nx, ny=ny, inflow=api.INFLOW_UNIFORM_OPEN if self.open_sides else api.INFLOW_UNIFORM,
                               u=self.u, nu=self.u * self.d / self.re, body=body, collision=self.collision[0],
                               magic=self.collision[1], ramp_steps=int(5.0 * unit),
                               kick=api.Kick(cx, cy, 0.4 * self.u / self.d, int(5.0 * unit), int(2.0 * unit)))

    def run(self, ch: api.Channel, units: float) -> np.ndarray:
        """Tunnel::run (tunnel.rs:72-92): `units` convective units of (cd, cl) per step;
        api.Diverged when the drag stops being finite."""
        return ch.run(int(units * self.steps_per_unit()), float(self.d))

    def experiment(self, ch: api.Channel, units: float, discard: float):
        """Tunnel::experiment's run and statistics (tunnel.rs:152-169): `units` in 20 equal
        parts, then Series::stats from time `discard` on.  Returns (stats, cd, cl)."""
        unit = self.steps_per_unit()
        parts = [ch.run(int(units / 20.0 * unit), float(self.d)) for _ in range(20)]
        coef = np.concatenate(parts)
        stats = api.series_stats(ch.be, coef[:, 0], coef[:, 1], 1.0 / unit, discard)
        return stats, coef[:, 0], coef[:, 1]


# ------------------------------------------------------------------------------------------
# Corner study (src/suite.rs:246-290; the "corners" command, main.rs:389-420)
# ------------------------------------------------------------------------------------------
CORNERS = ("square", "chamfered", "rounded", "recessed", "double-recessed")


def corner_run(be: api.Backend, corner: str, d: int = 60, re: float = 200.0, u: float = 0.1, b: float = 0.1,
               units: float = 140.0, discard: Optional[float] = None, arg: int = api.F64):
    """One section of the corner study: a square of side `d` with `corner` corners (size b D;
    the double recess uses two steps of b/2) in a tank 10 + 18 D long and 20 D high with
    radiating sides; returns Series::stats of the run after `discard` (default: half)."""
    t = Tunnel(d=d, re=re, u=u, upstream=10.0, downstream=18.0, height=20.0, collision=TRT, open_sides=True)
    cx, cy = t.centre()
    size = 0.5 * b if corner == "double-recessed" else b
    ch = api.Channel(be, t.spec(api.section(corner, size, cx, cy, float(d), 0.0)), arg)
    stats, _, _ = t.experiment(ch, units, 0.5 * units if discard is None else discard)
    ch.close()
    return stats

