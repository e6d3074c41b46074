"""CPU oracle loader -- TEST INFRASTRUCTURE ONLY.

Loads oracle/liboracle.so (the C++ restatement of the reference solver) behind the same
Python API as the GPU library.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module.
"""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_be = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def backend():
    global _be
    if _be is None:
        if not os.path.exists(LIB):
            build()
        from vortexlab_amd import api
        _be = api.Backend(ctypes.CDLL(LIB), "or_", "threads")
        _be.lib.or_cylinder_bench.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int] + [ctypes.c_void_p] * 5
        _be.lib.or_poiseuille.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int] + [ctypes.c_void_p] * 3
        _be.lib.or_dft_naive.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                        ctypes.c_void_p]
    return _be


def cylinder_bench(n, steps, threads=1):
    """cylinder::bench with the CLI's settings: (hash, fluid_nodes, nx, ny, seconds)."""
    be = backend()
    h = ctypes.c_uint64()
    fl, nx, ny = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    s = ctypes.c_double()
    be.check(be.lib.or_cylinder_bench(n, steps, threads, ctypes.byref(h), ctypes.byref(fl), ctypes.byref(nx),
                                      ctypes.byref(ny), ctypes.byref(s)))
    return h.value, fl.value, nx.value, ny.value, s.value


def poiseuille(rows, tau, collision, magic, q, threads=1):
    """poiseuille::run: (l2, force_balance, steps)."""
    be = backend()
    l2, fb, st = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
    be.check(be.lib.or_poiseuille(rows, tau, collision, magic, q, threads, ctypes.byref(l2), ctypes.byref(fb),
                                  ctypes.byref(st)))
    return l2.value, fb.value, st.value


def dft_naive(re, im):
    import numpy as np
    re = np.ascontiguousarray(re, dtype=np.float64)
    im = np.ascontiguousarray(im, dtype=np.float64)
    ore, oim = np.zeros_like(re), np.zeros_like(im)
    backend().lib.or_dft_naive(re.ctypes.data, im.ctypes.data, len(re), ore.ctypes.data, oim.ctypes.data)
    return ore, oim
