"""vortexlab_amd: the vortexlab D2Q9 time step on AMD MI355X (gfx950).

Hand-written HIP kernels behind the C ABI in include/vortexlab.h, driven from Python through
ctypes.  The GPU library is the only compute path: importing works without it (so the CPU
test suite can inspect the package), but every solver entry point raises if the HIP library
is missing or no GPU is visible -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

from . import api
from .api import (BGK, F16, F32, F64, TRT, INFLOW_PARABOLIC, INFLOW_UNIFORM, INFLOW_UNIFORM_OPEN, Channel,
                  ChannelSpec, Diverged, Kick, Stats, VortexError, circle, make_ctl, outside_band, polygon,
                  rotated, rounded_square, section, shape_from, translated, vel_const, vel_fn, vel_parabola,
                  vel_spin)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvortexlab.so")

_backend = None


def library_path() -> str:
    return LIB_PATH


def backend() -> api.Backend:
    """The loaded GPU library (raises if it has not been built)."""
    global _backend
    if _backend is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} is missing (run __graft_entry__.build())")
        _backend = api.Backend(ctypes.CDLL(LIB_PATH), "vl_", "precision")
    return _backend


def device_count() -> int:
    return int(backend().lib.vl_device_count())


def Builder(nx: int, ny: int) -> api.Builder:
    return api.Builder(backend(), nx, ny)


def make_channel(spec: ChannelSpec, precision: int = F64) -> Channel:
    return Channel(backend(), spec, precision)


def series_stats(cd, cl, dt, start) -> Stats:
    return api.series_stats(backend(), cd, cl, dt, start)
