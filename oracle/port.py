"""ctypes driver of the C/OpenMP FP32 step in ``port.c`` (the timed CPU
baseline).  TEST INFRASTRUCTURE: used by ``bench.py``'s cpu_baseline leg and
the tests only."""

from __future__ import annotations

import ctypes
import os
import subprocess
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libport.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


class _Grid(ctypes.Structure):
    _fields_ = [
        ("nz", ctypes.c_int), ("ny", ctypes.c_int), ("nx", ctypes.c_int),
        ("inv_dx", ctypes.c_float), ("h_src", ctypes.c_float),
        ("e_src", ctypes.c_void_p * 3), ("L", ctypes.c_int),
        ("a", ctypes.c_void_p * 12), ("b", ctypes.c_void_p * 12), ("ik", ctypes.c_void_p * 12),
        ("psi", ctypes.c_void_p * 12), ("metallic", ctypes.c_int),
    ]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class PortGrid:
    """Host arrays for one grid, laid out like the GPU state."""

    def __init__(self, grid, inv_dx, h_src, e_src, L=0, terms=None, metallic=0, seed=0):
        nz, ny, nx = grid
        self.shapes = {
            "hx": (nz, ny, nx + 1), "hy": (nz, ny + 1, nx), "hz": (nz + 1, ny, nx),
            "ex": (nz + 1, ny + 1, nx), "ey": (nz + 1, ny, nx + 1), "ez": (nz, ny + 1, nx + 1),
        }
        rng = np.random.default_rng(seed)
        self.f = {k: (rng.standard_normal(s, dtype=np.float32) * 1e-3) for k, s in self.shapes.items()}
        self.e_src = [np.ascontiguousarray(x, dtype=np.float32) for x in e_src]
        self.terms = terms or []
        self.psi = []
        g = _Grid()
        g.nz, g.ny, g.nx = nz, ny, nx
        g.inv_dx, g.h_src, g.L, g.metallic = inv_dx, h_src, L, metallic
        for c in range(3):
            g.e_src[c] = _ptr(self.e_src[c])
        keep = []
        for t, (a, b, ik, psi_shape) in enumerate(self.terms):
            arrs = [np.ascontiguousarray(v, dtype=np.float32) for v in (a, b, ik)]
            psi = np.zeros(psi_shape, dtype=np.float32)
            keep += arrs
            self.psi.append(psi)
            g.a[t], g.b[t], g.ik[t], g.psi[t] = _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]), _ptr(psi)
        self._keep = keep
        self.g = g
        self.lib = ctypes.CDLL(LIB)
        self.lib.port_step.argtypes = [ctypes.POINTER(_Grid)] + [ctypes.c_void_p] * 6
        self.lib.port_threads.restype = ctypes.c_int

    def step(self, n=1):
        f = self.f
        args = [_ptr(f[k]) for k in ("hx", "hy", "hz", "ex", "ey", "ez")]
        for _ in range(n):
            self.lib.port_step(ctypes.byref(self.g), *args)

    def threads(self):
        return int(self.lib.port_threads())


def time_gcups(pg: PortGrid, cells, steps):
    pg.step(1)  # warm-up (page-in)
    t0 = time.perf_counter()
    pg.step(steps)
    dt = time.perf_counter() - t0
    return cells * steps / dt / 1e9, dt
