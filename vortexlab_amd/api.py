"""Python mirror of the reference's solver API (Builder -> Sim -> Channel) over a C ABI.

The classes follow the reference crate's public Rust API (src/sim.rs:36-673,
src/geometry.rs, src/cases/channel.rs) and call a shared library through ctypes.  They are
written against a *backend*: the MI355X library ``libvortexlab.so`` (prefix ``vl_``) or, in
tests only, the CPU oracle (prefix ``or_``), which exports the same entry points.  That lets a
test state one case once and run it on both.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Callable, Optional, Sequence

import numpy as np

# ---------------------------------------------------------------------------------------
# constants (include/vortexlab.h)
# ---------------------------------------------------------------------------------------
OK, EINVAL, ENOMEM, EDEVICE, ENONFINITE = 0, 1, 2, 3, 4
F64, F32, F16 = 0, 1, 2
BGK, TRT = 0, 1
VEL_REST, VEL_CONST, VEL_PARABOLA, VEL_SPIN, VEL_CALLBACK = 0, 1, 2, 3, 4
INFLOW_PARABOLIC, INFLOW_UNIFORM, INFLOW_UNIFORM_OPEN = 0, 1, 2
CORNER = {"square": 0, "sharp": 0, "chamfered": 1, "rounded": 2, "recessed": 3, "double-recessed": 4}
Q = 9
CX = (0, 1, 0, -1, 0, 1, -1, -1, 1)
CY = (0, 0, 1, 0, -1, 1, 1, -1, -1)
W = (4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0)
CS2 = 1.0 / 3.0

INSIDE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.c_double)
VELOCITY_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_double))


class VelocityDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("p", C.c_double * 4), ("fn", VELOCITY_FN), ("ctx", C.c_void_p)]


class SimInfo(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("fluid_nodes", C.c_int64), ("link_count", C.c_int64),
                ("bnode_count", C.c_int64), ("pnode_count", C.c_int64), ("force_groups", C.c_int64),
                ("t", C.c_int64), ("precision", C.c_int32), ("collision", C.c_int32),
                ("omega_plus", C.c_double), ("omega_minus", C.c_double),
                ("outlet_density_offset", C.c_double), ("device_bytes", C.c_int64), ("split", C.c_int32),
                ("pass_steps", C.c_int32), ("pass_nodes", C.c_int64), ("pass_tasks", C.c_int64),
                ("listed_nodes", C.c_int64), ("pass_rows", C.c_int32), ("pass_joined_tasks", C.c_int32),
                ("pass_far_nodes", C.c_int64), ("pass_far_tasks", C.c_int64)]


class Ctl(C.Structure):
    _fields_ = [("vscale", C.c_double * 4), ("target", C.c_double), ("has_target", C.c_int32),
                ("pad_", C.c_int32)]


class KickDesc(C.Structure):
    _fields_ = [("cx", C.c_double), ("cy", C.c_double), ("omega", C.c_double), ("start", C.c_uint64),
                ("duration", C.c_uint64)]


class ChannelSpecDesc(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("inflow", C.c_int32), ("collision", C.c_int32),
                ("u", C.c_double), ("nu", C.c_double), ("magic", C.c_double), ("ramp_steps", C.c_uint64),
                ("has_kick", C.c_int32), ("pad_", C.c_int32), ("kick", KickDesc)]


class StatsDesc(C.Structure):
    _fields_ = [("cd_mean", C.c_double), ("cd_rms", C.c_double), ("cl_mean", C.c_double),
                ("cl_rms", C.c_double), ("st", C.c_double), ("st_crossings", C.c_double),
                ("periodic", C.c_int32), ("pad_", C.c_int32)]


class BatchInfoDesc(C.Structure):
    _fields_ = [("ncases", C.c_int64), ("total_nodes", C.c_int64), ("fluid_nodes", C.c_int64),
                ("lattice_updates", C.c_int64), ("max_steps", C.c_int64), ("t", C.c_int64),
                ("boundary_tiles", C.c_int32), ("bulk_tiles", C.c_int32), ("tile_w", C.c_int32),
                ("tile_h", C.c_int32), ("grid_bulk", C.c_int32), ("grid_boundary", C.c_int32),
                ("last_run_seconds", C.c_double), ("device_bytes", C.c_int64),
                ("autotune_ms", (C.c_double * 4) * 3),
                ("pass_steps", C.c_int32), ("pass_rows", C.c_int32), ("pass_tasks", C.c_int32),
                ("listed_tiles", C.c_int32), ("pass_nodes", C.c_int64), ("listed_nodes", C.c_int64),
                ("setup_on_device", C.c_int32), ("pad_", C.c_int32), ("setup_ms", C.c_double * 4)]


class VortexError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


class Diverged(VortexError):
    pass


_D = C.POINTER(C.c_double)
_VP = C.c_void_p
# (name, restype, argtypes); the prefix (vl_ / or_) is added per backend.
_SIGS = [
    ("last_error", C.c_char_p, []),
    ("shape_circle", _VP, [C.c_double] * 3),
    ("shape_polygon", _VP, [_D, C.c_size_t]),
    ("shape_rounded_square", _VP, [C.c_double] * 4),
    ("shape_rotated", _VP, [_VP, C.c_double, C.c_double, C.c_double]),
    ("shape_translated", _VP, [_VP, C.c_double, C.c_double]),
    ("shape_section", _VP, [C.c_int] + [C.c_double] * 5),
    ("shape_outside_band", _VP, [C.c_double, C.c_double]),
    ("shape_callback", _VP, [INSIDE_FN, _VP]),
    ("shape_inside", C.c_int, [_VP, C.c_double, C.c_double]),
    ("cut_fraction", C.c_double, [_VP] + [C.c_double] * 4),
    ("shape_free", None, [_VP]),
    ("builder_new", _VP, [C.c_size_t, C.c_size_t]),
    ("builder_config", C.c_int, [_VP, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double,
                                 C.c_double]),
    ("builder_add_wall", C.c_int, [_VP, C.POINTER(VelocityDesc), C.c_int, C.c_int, _VP]),
    ("builder_add_pressure", C.c_int, [_VP, C.c_double]),
    ("builder_add_radiating", C.c_int, [_VP, C.c_double]),
    ("builder_fill_rect", C.c_int, [_VP, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.c_int]),
    ("builder_fill_shape", C.c_int, [_VP, C.c_int]),
    ("builder_free", None, [_VP]),
    ("sim_create", C.c_int, [_VP, C.c_int, C.POINTER(_VP)]),
    ("sim_destroy", None, [_VP]),
    ("sim_info_get", C.c_int, [_VP, C.POINTER(SimInfo)]),
    ("plan_passes", C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _VP,
                              C.c_int64, _VP, C.c_int64, _VP]),
    ("pass_lockstep", C.c_int, [_VP, _VP, C.c_int64, _VP, C.c_int, C.c_int, C.c_int]),
    ("sim_init_macro", C.c_int, [_VP, _D, _D, _D]),
    ("sim_set_vscale", C.c_int, [_VP, _D]),
    ("sim_set_outflow_target", C.c_int, [_VP, C.c_int, C.c_double]),
    ("sim_step", C.c_int, [_VP, C.c_int64]),
    ("sim_run", C.c_int, [_VP, C.c_int64, C.POINTER(Ctl), _D]),
    ("sim_run_noalloc", C.c_int, [_VP, C.c_int64, C.POINTER(Ctl), _D]),
    ("sim_prepare", C.c_int, [_VP, C.c_int64]),
    ("sim_prepare_after", C.c_int, [_VP, C.c_int64, C.c_int64]),
    ("sim_get_force", C.c_int, [_VP, _D]),
    ("sim_read", C.c_int, [_VP, _D]),
    ("sim_write", C.c_int, [_VP, _D]),
    ("sim_macros", C.c_int, [_VP, _D, _D, _D]),
    ("sim_state_hash", C.c_int, [_VP, C.POINTER(C.c_uint64)]),
    ("sim_total_mass", C.c_int, [_VP, _D]),
    ("sim_total_momentum", C.c_int, [_VP, _D]),
    ("sim_max_speed", C.c_int, [_VP, _D]),
    ("sim_vorticity", C.c_int, [_VP, C.POINTER(C.c_float)]),
    ("sim_is_fluid", C.c_int, [_VP, C.c_int64, C.c_int64, C.POINTER(C.c_int)]),
    ("sim_periodic", C.c_int, [_VP, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("sim_set_timing", C.c_int, [_VP, C.c_int]),
    ("device_copy_bandwidth", C.c_int, [C.c_size_t, C.c_int, _D]),
    ("sim_get_timing", C.c_int, [_VP, _D, C.POINTER(C.c_int64), _D, C.POINTER(C.c_int64)]),
    ("channel_create", C.c_int, [C.POINTER(ChannelSpecDesc), _VP, C.c_int, C.POINTER(_VP)]),
    ("channel_destroy", None, [_VP]),
    ("channel_sim", _VP, [_VP]),
    ("channel_set_absorbing", C.c_int, [_VP, C.c_int]),
    ("channel_run", C.c_int, [_VP, C.c_int64, C.c_double, _D]),
    ("channel_coefficients", C.c_int, [_VP, C.c_double, _D, _D]),
    ("series_stats", C.c_int, [_D, _D, C.c_size_t, C.c_double, C.c_double, C.POINTER(StatsDesc)]),
    ("dominant_frequency", C.c_int, [_D, C.c_size_t, C.c_double, _D]),
    ("crossing_frequency", C.c_int, [_D, C.c_size_t, C.c_double, _D, C.POINTER(C.c_int)]),
    ("peaks", C.c_int, [_D, C.c_size_t, _D, _D, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("fft", C.c_int, [_D, _D, C.c_size_t]),
    ("batch_create", C.c_int, [C.c_int, C.POINTER(_VP)]),
    ("batch_add", C.c_int, [_VP, C.POINTER(ChannelSpecDesc), _VP, C.c_int64, C.c_double]),
    ("batch_build", C.c_int, [_VP]),
    ("batch_run", C.c_int, [_VP, C.c_int64]),
    ("batch_info_get", C.c_int, [_VP, C.POINTER(BatchInfoDesc)]),
    ("batch_case_result", C.c_int, [_VP, C.c_int, _D, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    ("batch_read_case", C.c_int, [_VP, C.c_int, _D]),
    ("batch_case_info", C.c_int, [_VP, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    ("batch_set_timing", C.c_int, [_VP, C.c_int]),
    ("batch_get_timing", C.c_int, [_VP, _D, _D, _D, C.POINTER(C.c_int64)]),
    ("batch_table_digests", C.c_int, [_VP, C.c_int, C.POINTER(C.c_uint64), C.c_int]),
    ("batch_destroy", None, [_VP]),
]


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_D)


class Backend:
    """A loaded shared library exposing the solver entry points under one prefix.

    ``create_arg`` names what the second argument of ``sim_create`` means for this backend:
    the storage precision for the GPU library, the worker-thread count for the oracle.
    """

    def __init__(self, lib: C.CDLL, prefix: str, create_arg: str):
        self.lib = lib
        self.prefix = prefix
        self.create_arg = create_arg
        for name, res, args in _SIGS:
            if not hasattr(lib, prefix + name):
                continue  # the oracle has no device-timing entry points
            fn = getattr(lib, prefix + name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    def check(self, status: int):
        if status != OK:
            msg = self.last_error().decode(errors="replace")
            if status == ENONFINITE:
                raise Diverged(status, msg)
            raise VortexError(status, msg)


def plan_passes(be: Backend, flags: np.ndarray, k: int, v: int = 1, rows: int = 20, x0: int = 32, waves: int = 0,
                join: bool = False):
    """The K-step pass plan of a domain (flags: ny x nx, 0 interior, 1 boundary, 2 solid), host
    only: (tasks as an (n, 4) int32 array of first column, first row, rows, columns | ring-edge
    bits << 8 | lockstep << 12 | joined << 13 | offset of the first output column in the wave's
    columns << 16, in the order far joined groups, far single-wave tasks, near joined groups, near
    single-wave tasks; the number of far tasks; listed nodes as (y << 16 | x); covered node count).
    `join` plans joined strips for workgroups of `waves` waves; waves > 0 sets the lockstep bits
    as a simulation with workgroups of that many waves does (far and near tasks launched apart)."""
    fl = np.ascontiguousarray(flags, dtype=np.uint8)
    ny, nx = fl.shape
    counts = np.zeros(6, dtype=np.int64)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    be.check(be.plan_passes(ptr(fl), nx, ny, k, v, rows, x0, waves, int(join), None, 0, None, 0, ptr(counts)))
    tasks = np.zeros((max(int(counts[0]), 1), 4), dtype=np.int32)
    listed = np.zeros(max(int(counts[2]), 1), dtype=np.uint32)
    be.check(be.plan_passes(ptr(fl), nx, ny, k, v, rows, x0, waves, int(join), ptr(tasks), int(counts[0]),
                            ptr(listed), int(counts[2]), ptr(counts)))
    return tasks[:int(counts[0])], int(counts[1]), listed[:int(counts[2])], int(counts[3])


def pass_lockstep(be: Backend, tasks: np.ndarray, ranges, waves: int, owner=None, check_only: bool = False):
    """The lockstep bits of a pass task list (host only): a copy of `tasks` with bit 12 set as the
    library sets it for launches over `ranges` ([(begin, end), ...], workgroups of `waves` tasks
    from each begin) and `owner` (the case of each task, or None); raises VortexError if the result
    (or, with check_only, the given bits) would let one workgroup's waves meet different barriers."""
    t = np.ascontiguousarray(tasks, dtype=np.int32).copy()
    r = np.ascontiguousarray(np.array(ranges, dtype=np.int64).reshape(-1, 2))
    o = None if owner is None else np.ascontiguousarray(owner, dtype=np.int32)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    be.check(be.pass_lockstep(ptr(t), None if o is None else ptr(o), len(t), ptr(r), len(r), waves, int(check_only)))
    return t


# ---------------------------------------------------------------------------------------
# Shapes (src/geometry.rs)
# ---------------------------------------------------------------------------------------
@dataclass
class ShapeSpec:
    """A shape description; materialised into a library handle when used."""

    kind: str
    args: tuple
    inner: Optional["ShapeSpec"] = None

    def make(self, be: Backend, keep: list):
        """The library handle of this shape.  A Python predicate becomes a ctypes callback, which
        is appended to `keep`: the caller holds it until the handle has been consumed (shapes
        are evaluated while a domain is built, never afterwards)."""
        k, a = self.kind, self.args
        if k == "circle":
            return be.shape_circle(*a)
        if k == "polygon":
            pts = np.ascontiguousarray(np.asarray(a[0], dtype=np.float64).reshape(-1, 2))
            return be.shape_polygon(_dptr(pts), pts.shape[0])
        if k == "rounded_square":
            return be.shape_rounded_square(*a)
        if k == "rotated":
            return be.shape_rotated(self.inner.make(be, keep), *a)
        if k == "translated":
            return be.shape_translated(self.inner.make(be, keep), *a)
        if k == "section":
            return be.shape_section(*a)
        if k == "outside_band":
            return be.shape_outside_band(*a)
        if k == "callback":
            fn = a[0]
            cb = INSIDE_FN(lambda _ctx, x, y: 1 if fn(x, y) else 0)
            keep.append(cb)
            return be.shape_callback(cb, None)
        raise ValueError(k)


def circle(cx, cy, r):
    return ShapeSpec("circle", (float(cx), float(cy), float(r)))


def polygon(pts):
    return ShapeSpec("polygon", (np.asarray(pts, dtype=np.float64),))


def rounded_square(cx, cy, half, r):
    return ShapeSpec("rounded_square", (float(cx), float(cy), float(half), float(r)))


def rotated(inner: ShapeSpec, cx, cy, angle):
    return ShapeSpec("rotated", (float(cx), float(cy), float(angle)), inner)


def translated(inner: ShapeSpec, dx, dy):
    return ShapeSpec("translated", (float(dx), float(dy)), inner)


def section(corner: str, b: float, cx, cy, d, angle=0.0):
    return ShapeSpec("section", (CORNER[corner], float(b), float(cx), float(cy), float(d), float(angle)))


def outside_band(y0, y1):
    return ShapeSpec("outside_band", (float(y0), float(y1)))


def shape_from(fn: Callable[[float, float], bool]):
    return ShapeSpec("callback", (fn,))


# ---------------------------------------------------------------------------------------
# Wall velocities (the reference's VelocityFn closures)
# ---------------------------------------------------------------------------------------
def vel_const(ux, uy):
    return ("const", float(ux), float(uy))


def vel_parabola(u_mean, h):
    return ("parabola", float(u_mean), float(h))


def vel_spin(cx, cy):
    return ("spin", float(cx), float(cy))


def vel_fn(fn: Callable[[float, float], Sequence[float]]):
    return ("fn", fn)


def _velocity_desc(v, keep: list) -> Optional[VelocityDesc]:
    """The C descriptor of a wall velocity; a Python function becomes a ctypes callback appended
    to `keep`, which the caller holds until the domain is built."""
    if v is None:
        return None
    d = VelocityDesc()
    if v[0] == "const":
        d.kind, d.p[0], d.p[1] = VEL_CONST, v[1], v[2]
    elif v[0] == "parabola":
        d.kind, d.p[0], d.p[1] = VEL_PARABOLA, v[1], v[2]
    elif v[0] == "spin":
        d.kind, d.p[0], d.p[1] = VEL_SPIN, v[1], v[2]
    elif v[0] == "fn":
        fn = v[1]

        def tramp(_ctx, x, y, out):
            u = fn(x, y)
            out[0], out[1] = float(u[0]), float(u[1])

        cb = VELOCITY_FN(tramp)
        keep.append(cb)
        d.kind, d.fn = VEL_CALLBACK, cb
    else:
        raise ValueError(v)
    return d


# ---------------------------------------------------------------------------------------
# Builder / Sim (src/sim.rs)
# ---------------------------------------------------------------------------------------
class Builder:
    """Describes a domain before the solver is built (src/sim.rs:70-137)."""

    def __init__(self, be: Backend, nx: int, ny: int):
        self.be = be
        self.h = be.builder_new(nx, ny)
        if not self.h:
            be.check(EINVAL)
        self.nx, self.ny = nx, ny
        self.periodic_x = False
        self.periodic_y = False
        self.tau = 1.0
        self.collision = TRT
        self.magic = 3.0 / 16.0
        self.accel = (0.0, 0.0)
        self._keep = []  # ctypes callbacks of this domain's shapes and velocities, until build()

    def _kind(self, k: int) -> int:
        if k < 0:
            self.be.check(EINVAL)
        return k

    def add_wall(self, velocity=None, vgroup: int = 0, fgroup: Optional[int] = None,
                 shape: Optional[ShapeSpec] = None) -> int:
        d = _velocity_desc(velocity, self._keep)
        sh = shape.make(self.be, self._keep) if shape is not None else None
        return self._kind(self.be.builder_add_wall(self.h, C.byref(d) if d is not None else None, vgroup,
                                                   -1 if fgroup is None else fgroup, sh))

    def add_static_wall(self) -> int:
        return self.add_wall()

    def add_pressure(self, rho: float) -> int:
        return self._kind(self.be.builder_add_pressure(self.h, rho))

    def add_radiating(self, rho: float) -> int:
        return self._kind(self.be.builder_add_radiating(self.h, rho))

    def fill_rect(self, x0, x1, y0, y1, kind):
        self.be.check(self.be.builder_fill_rect(self.h, x0, x1, y0, y1, kind))

    def fill_shape(self, kind):
        self.be.check(self.be.builder_fill_shape(self.h, kind))

    def build(self, arg: int = 0) -> "Sim":
        """Build the solver.  ``arg`` is the precision (GPU library) or thread count (oracle)."""
        be = self.be
        be.check(be.builder_config(self.h, int(self.periodic_x), int(self.periodic_y), self.tau, self.collision,
                                   self.magic, self.accel[0], self.accel[1]))
        out = C.c_void_p()
        h, self.h = self.h, None
        try:
            be.check(be.sim_create(h, arg, C.byref(out)))
        finally:
            self._keep = []  # the build evaluated (and freed) every shape and velocity
        return Sim(be, out.value)

    def __del__(self):
        if getattr(self, "h", None):
            self.be.builder_free(self.h)
            self.h = None


class Sim:
    """A built domain: populations, boundary tables and the step (src/sim.rs:180-673)."""

    def __init__(self, be: Backend, handle, owned: bool = True):
        self.be, self.h, self.owned = be, handle, owned
        self.vscale = [1.0, 1.0, 1.0, 1.0]
        self.outflow_target: Optional[float] = None
        i = self.info()
        self.nx, self.ny = int(i.nx), int(i.ny)
        self.n = self.nx * self.ny

    def info(self) -> SimInfo:
        i = SimInfo()
        self.be.check(self.be.sim_info_get(self.h, C.byref(i)))
        return i

    @property
    def t(self) -> int:
        return int(self.info().t)

    @property
    def fluid_nodes(self) -> int:
        return int(self.info().fluid_nodes)

    @property
    def link_count(self) -> int:
        return int(self.info().link_count)

    @property
    def outlet_density_offset(self) -> float:
        return float(self.info().outlet_density_offset)

    def _push_controls(self):
        v = np.asarray(self.vscale, dtype=np.float64)
        self.be.check(self.be.sim_set_vscale(self.h, _dptr(v)))
        tgt = self.outflow_target
        self.be.check(self.be.sim_set_outflow_target(self.h, int(tgt is not None), float(tgt or 0.0)))

    def step(self, n: int = 1):
        """``n`` time steps with the current ``vscale`` and ``outflow_target``."""
        self._push_controls()
        self.be.check(self.be.sim_step(self.h, n))

    def run(self, ctl: np.ndarray, want_force: bool = True) -> Optional[np.ndarray]:
        """Steps with per-step controls: ``ctl`` is a structured array of ``Ctl`` records."""
        k = len(ctl)
        ng = int(self.info().force_groups)
        hist = np.zeros((k, max(ng, 1), 2)) if want_force else None
        arr = (Ctl * k)(*ctl)
        self.be.check(self.be.sim_run(self.h, k, arr, _dptr(hist) if want_force else None))
        return hist[:, :ng] if want_force else None

    def prepare(self, steps: int, after: int = 0):
        """Instantiates the step graphs a run of ``steps`` steps replays, the run starting
        ``after`` steps from the current state (GPU library only; runs shorter than a step window
        replay their whole groups as one graph only once it is prepared)."""
        self.be.check(self.be.sim_prepare_after(self.h, int(after), int(steps)))

    def init(self, field: Callable[[int, int], tuple]):
        """Sim::init: equilibrium of field(x, y) -> (rho, ux, uy) on every fluid node."""
        rho = np.ones(self.n)
        ux = np.zeros(self.n)
        uy = np.zeros(self.n)
        for y in range(self.ny):
            for x in range(self.nx):
                r, u, v = field(x, y)
                k = y * self.nx + x
                rho[k], ux[k], uy[k] = r, u, v
        self.init_arrays(rho, ux, uy)

    def init_arrays(self, rho, ux, uy):
        rho, ux, uy = (np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (self.ny, self.nx)))
                       .reshape(-1).copy() for a in (rho, ux, uy))
        self.be.check(self.be.sim_init_macro(self.h, _dptr(rho), _dptr(ux), _dptr(uy)))

    @property
    def force(self) -> np.ndarray:
        ng = int(self.info().force_groups)
        out = np.zeros((max(ng, 1), 2))
        self.be.check(self.be.sim_get_force(self.h, _dptr(out)))
        return out[:ng]

    def read(self) -> np.ndarray:
        """Populations as an array [9, ny, nx] (reference layout f[i*n + y*nx + x])."""
        out = np.zeros((Q, self.ny, self.nx))
        self.be.check(self.be.sim_read(self.h, _dptr(out)))
        return out

    def write(self, f: np.ndarray):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(Q, self.ny, self.nx)
        self.be.check(self.be.sim_write(self.h, _dptr(f)))

    def macros_all(self):
        rho, ux, uy = (np.zeros((self.ny, self.nx)) for _ in range(3))
        self.be.check(self.be.sim_macros(self.h, _dptr(rho), _dptr(ux), _dptr(uy)))
        return rho, ux, uy

    def macros(self, x: int, y: int):
        rho, ux, uy = self.macros_all()
        return float(rho[y, x]), float(ux[y, x]), float(uy[y, x])

    def state_hash(self) -> int:
        h = C.c_uint64()
        self.be.check(self.be.sim_state_hash(self.h, C.byref(h)))
        return int(h.value)

    def total_mass(self) -> float:
        v = C.c_double()
        self.be.check(self.be.sim_total_mass(self.h, C.byref(v)))
        return v.value

    def total_momentum(self):
        v = np.zeros(2)
        self.be.check(self.be.sim_total_momentum(self.h, _dptr(v)))
        return float(v[0]), float(v[1])

    def max_speed(self) -> float:
        v = C.c_double()
        self.be.check(self.be.sim_max_speed(self.h, C.byref(v)))
        return v.value

    def vorticity(self) -> np.ndarray:
        """Sim::vorticity (src/sim.rs:581-620): dv/dx - du/dy per node as float32 [ny, nx]."""
        out = np.zeros((self.ny, self.nx), dtype=np.float32)
        self.be.check(self.be.sim_vorticity(self.h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def is_fluid(self, x: int, y: int) -> bool:
        """Sim::is_fluid (src/sim.rs:381-383)."""
        v = C.c_int()
        self.be.check(self.be.sim_is_fluid(self.h, int(x), int(y), C.byref(v)))
        return bool(v.value)

    def periodic(self):
        """Sim::periodic (src/sim.rs:670-672): (periodic in x, periodic in y)."""
        px, py = C.c_int(), C.c_int()
        self.be.check(self.be.sim_periodic(self.h, C.byref(px), C.byref(py)))
        return bool(px.value), bool(py.value)

    def set_timing(self, on: bool = True):
        self.be.check(self.be.sim_set_timing(self.h, int(on)))

    def timing(self):
        """(bulk_ms, bulk_launches, total_ms, steps) since set_timing(True)."""
        bms, tms = C.c_double(), C.c_double()
        bn, st = C.c_int64(), C.c_int64()
        self.be.check(self.be.sim_get_timing(self.h, C.byref(bms), C.byref(bn), C.byref(tms), C.byref(st)))
        return bms.value, bn.value, tms.value, st.value

    def close(self):
        if self.h and self.owned:
            self.be.sim_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_ctl(vscale_rows, targets=None) -> list:
    """Per-step controls: vscale_rows [k][4]; targets [k] (None / NaN = no outflow target)."""
    out = []
    for k, v in enumerate(vscale_rows):
        c = Ctl()
        for g in range(4):
            c.vscale[g] = float(v[g])
        t = None if targets is None else targets[k]
        if t is None or (isinstance(t, float) and math.isnan(t)):
            c.has_target, c.target = 0, 0.0
        else:
            c.has_target, c.target = 1, float(t)
        out.append(c)
    return out


# ---------------------------------------------------------------------------------------
# Channel (src/cases/channel.rs)
# ---------------------------------------------------------------------------------------
@dataclass
class Kick:
    cx: float
    cy: float
    omega: float
    start: int
    duration: int


@dataclass
class ChannelSpec:
    nx: int
    ny: int
    inflow: int
    u: float
    nu: float
    body: ShapeSpec
    collision: int = TRT
    magic: float = 3.0 / 16.0
    ramp_steps: int = 0
    kick: Optional[Kick] = None


def _spec_desc(spec: ChannelSpec) -> ChannelSpecDesc:
    d = ChannelSpecDesc()
    d.nx, d.ny, d.inflow, d.collision = spec.nx, spec.ny, spec.inflow, spec.collision
    d.u, d.nu, d.magic, d.ramp_steps = spec.u, spec.nu, spec.magic, spec.ramp_steps
    if spec.kick is not None:
        d.has_kick = 1
        k = spec.kick
        d.kick.cx, d.kick.cy, d.kick.omega, d.kick.start, d.kick.duration = (k.cx, k.cy, k.omega, k.start,
                                                                             k.duration)
    return d


class Channel:
    """A body in a channel with its per-step driver (src/cases/channel.rs:78-155)."""

    def __init__(self, be: Backend, spec: ChannelSpec, arg: int = 0):
        self.be = be
        d = _spec_desc(spec)
        out = C.c_void_p()
        keep = []  # callbacks of the body, held until the domain is built
        be.check(be.channel_create(C.byref(d), spec.body.make(be, keep), arg, C.byref(out)))
        del keep
        self.h = out.value
        self.u_ref = spec.u
        self.sim = Sim(be, be.channel_sim(self.h), owned=False)

    def set_absorbing(self, on: bool):
        self.be.check(self.be.channel_set_absorbing(self.h, int(on)))

    def run(self, steps: int, d: float) -> np.ndarray:
        """``steps`` x Channel::step; returns the (cd, cl) history [steps, 2]."""
        coef = np.zeros((steps, 2))
        self.be.check(self.be.channel_run(self.h, steps, d, _dptr(coef)))
        return coef

    def coefficients(self, d: float):
        cd, cl = C.c_double(), C.c_double()
        self.be.check(self.be.channel_coefficients(self.h, d, C.byref(cd), C.byref(cl)))
        return cd.value, cl.value

    def close(self):
        if self.h:
            self.sim.h = None
            self.be.channel_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many channel cases stepped together on one GPU (GPU library only)."""

    def __init__(self, be: Backend, precision: int = F32):
        self.be = be
        out = C.c_void_p()
        be.check(be.batch_create(precision, C.byref(out)))
        self.h = out.value
        self.steps = []
        self._keep = []  # callbacks of the pending bodies, held until build()

    def add(self, spec: ChannelSpec, steps: int, d: float) -> int:
        k = self.be.batch_add(self.h, C.byref(_spec_desc(spec)), spec.body.make(self.be, self._keep), int(steps),
                              float(d))
        if k < 0:
            self.be.check(EINVAL)
        self.steps.append(int(steps))
        return k

    def build(self):
        try:
            self.be.check(self.be.batch_build(self.h))
        finally:
            self._keep = []  # the build evaluated (and freed) every pending body

    def run(self, max_steps: int = -1):
        self.be.check(self.be.batch_run(self.h, int(max_steps)))

    def info(self) -> BatchInfoDesc:
        i = BatchInfoDesc()
        self.be.check(self.be.batch_info_get(self.h, C.byref(i)))
        return i

    def case_info(self, k: int):
        """(nx, ny, fluid_nodes, steps) of case k."""
        v = [C.c_int64() for _ in range(4)]
        self.be.check(self.be.batch_case_info(self.h, k, *(C.byref(x) for x in v)))
        return tuple(int(x.value) for x in v)

    def set_timing(self, on: bool):
        self.be.check(self.be.batch_set_timing(self.h, 1 if on else 0))

    def timing(self):
        """(boundary_ms, bulk_ms, step_ms, steps) summed since set_timing."""
        v = [C.c_double() for _ in range(3)]
        n = C.c_int64()
        self.be.check(self.be.batch_get_timing(self.h, *(C.byref(x) for x in v), C.byref(n)))
        return v[0].value, v[1].value, v[2].value, int(n.value)

    def case_fluid_nodes(self, k: int) -> int:
        return self.case_info(k)[2]

    def result(self, k: int):
        """(coef [steps, 2], steps_done, status) of case k."""
        coef = np.zeros((self.steps[k], 2))
        done, st = C.c_int64(), C.c_int()
        self.be.check(self.be.batch_case_result(self.h, k, _dptr(coef), C.byref(done), C.byref(st)))
        return coef, int(done.value), int(st.value)

    def read_case(self, k: int, nx: int, ny: int) -> np.ndarray:
        out = np.zeros((Q, ny, nx))
        self.be.check(self.be.batch_read_case(self.h, k, _dptr(out)))
        return out

    TABLE_NAMES = ("flags", "bx", "by", "bfirst", "dir", "kind", "vgroup", "lflags", "lnx", "lny", "fslot", "q",
                   "uwx", "uwy", "cu6", "rho", "px", "py", "pnx", "pny", "gfirst", "counts")

    def table_digests(self, k: int) -> dict:
        """FNV-1a digests of case k's device tables, by name (vl_batch_table_digests)."""
        out = (C.c_uint64 * len(self.TABLE_NAMES))()
        self.be.check(self.be.batch_table_digests(self.h, k, out, len(out)))
        return dict(zip(self.TABLE_NAMES, list(out)))

    def close(self):
        if self.h:
            self.be.batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------------------
# Statistics (src/cases/channel.rs:157-207, src/signal.rs)
# ---------------------------------------------------------------------------------------
@dataclass
class Stats:
    cd_mean: float
    cd_rms: float
    cl_mean: float
    cl_rms: float
    st: float
    st_crossings: float
    periodic: bool


def series_stats(be: Backend, cd, cl, dt: float, start: float) -> Stats:
    cd = np.ascontiguousarray(cd, dtype=np.float64)
    cl = np.ascontiguousarray(cl, dtype=np.float64)
    s = StatsDesc()
    be.check(be.series_stats(_dptr(cd), _dptr(cl), len(cd), dt, start, C.byref(s)))
    return Stats(s.cd_mean, s.cd_rms, s.cl_mean, s.cl_rms, s.st, s.st_crossings, bool(s.periodic))


def dominant_frequency(be: Backend, x, dt: float):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(2)
    be.check(be.dominant_frequency(_dptr(x), len(x), dt, _dptr(out)))
    return float(out[0]), float(out[1])


def peaks(be: Backend, x) -> np.ndarray:
    """signal::peaks: [k, 2] array of (fractional index, refined value) of the local maxima
    above the mean."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    cap = max(len(x) // 2, 1)
    pos, val = np.zeros(cap), np.zeros(cap)
    n = C.c_size_t()
    be.check(be.peaks(_dptr(x), len(x), _dptr(pos), _dptr(val), cap, C.byref(n)))
    k = min(int(n.value), cap)
    return np.stack([pos[:k], val[:k]], axis=1)


def crossing_frequency(be: Backend, x, dt: float) -> Optional[float]:
    x = np.ascontiguousarray(x, dtype=np.float64)
    f, found = C.c_double(), C.c_int()
    be.check(be.crossing_frequency(_dptr(x), len(x), dt, C.byref(f), C.byref(found)))
    return f.value if found.value else None


def fft(be: Backend, re, im):
    re = np.array(re, dtype=np.float64)
    im = np.array(im, dtype=np.float64)
    be.check(be.fft(_dptr(re), _dptr(im), len(re)))
    return re, im
