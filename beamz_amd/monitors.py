"""DFT flux monitors: host lowering to weighted-gather plans, packing into the
ragged accumulator arenas, and flux post-processing.

References: ``beamz/devices/monitors/compiler.py:61-102, 233-636`` (specs,
arenas), ``beamz/simulation/cuda/runtime.py:530-618`` (packing),
``beamz/simulation/observe.py:166-234`` (phasors and flux),
``beamz/devices/monitors/monitors.py:61-68`` (2D line element).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .lattice import COMPONENTS, line_gather_plan, plane_gather_plan


@dataclass(frozen=True)
class FluxMonitor:
    """An axis-aligned DFT flux plane (3D) or line (2D) in index space.

    ``normal`` is ``"x"``, ``"y"`` or ``"z"``; the plane sits on the E-node plane
    ``index`` (position ``index * dx``); ``ranges`` maps each tangential axis to a
    ``(start, stop)`` cell interval (default: the whole axis).  Samples are taken
    at tangential cell centres.  ``components`` restricts the accumulated
    components (default: the four tangential ones that enter the flux)."""

    normal: str
    index: int
    freqs: np.ndarray
    ranges: dict = field(default_factory=dict)
    components: tuple = None
    normal_sign: float = 1.0
    interval: int = 1
    t_start: float = 0.0
    t_end: float = float("inf")
    window: str = "rect"
    name: str = None


FLUX_COMPONENTS = {
    "x": ("Ey", "Ez", "Hy", "Hz"),
    "y": ("Ex", "Ez", "Hx", "Hz"),
    "z": ("Ex", "Ey", "Hx", "Hy"),
}


@dataclass
class CompiledMonitor:
    name: str
    freqs: np.ndarray
    plans: dict            # component -> (index[P, K], weight[P, K]) or None
    points: int
    normal: str
    normal_sign: float
    measure: float
    interval: int
    t_start: float
    t_end: float
    window_code: int
    value_offset: int = 0
    weight_offset: int = 0
    index: int = 0             # the normal (E-node) index of the plane or line
    cell_range: tuple = ()     # 2D: the tangential cell interval of the line samples
    ranges: tuple = ()         # 3D: ((axis, (start, stop)), ...) of the plane's tangential cells


def lower_monitor(mon: FluxMonitor, grid, shapes, dx, index):
    is_3d = len(grid) == 3
    freqs = np.asarray(mon.freqs, dtype=np.float64).reshape(-1)
    if freqs.size == 0 or np.any(~np.isfinite(freqs)) or np.any(freqs <= 0):
        raise ValueError("monitor frequencies must be finite and positive")
    wanted = mon.components or FLUX_COMPONENTS[mon.normal]
    plans = {}
    if is_3d:
        axes = {"z": 0, "y": 1, "x": 2}
        tangential = {"x": ("z", "y"), "y": ("z", "x"), "z": ("y", "x")}[mon.normal]
        ranges = {a: tuple(mon.ranges.get(a, (0, grid[axes[a]]))) for a in tangential}
        for c in COMPONENTS:
            plans[c] = (plane_gather_plan(c, shapes[c], dx, mon.normal, mon.index, ranges)
                        if c in wanted else None)
        measure = dx * dx  # plane_sample_area on a uniform grid (lattice.py:357-366)
    else:
        if mon.normal not in ("x", "y"):
            raise ValueError("a 2D line monitor is normal to x or y")
        axes = {"y": 0, "x": 1}
        tang = "y" if mon.normal == "x" else "x"
        rng = tuple(mon.ranges.get(tang, (0, grid[axes[tang]])))
        for c in COMPONENTS:
            active = c in ("Ez", "Hx", "Hy") and c in wanted
            plans[c] = line_gather_plan(c, shapes[c], dx, mon.normal, mon.index, rng) if active else None
        measure = dx  # _line_integral_scale_2d (monitors.py:61-68)
    cell_range = rng if not is_3d else ()
    plane_ranges = tuple((a, tuple(int(v) for v in ranges[a])) for a in tangential) if is_3d else ()
    points = next(p[0].shape[0] for p in plans.values() if p is not None)
    return CompiledMonitor(
        name=mon.name or f"monitor_{index}", freqs=freqs, plans=plans, points=points,
        normal=mon.normal, normal_sign=float(mon.normal_sign), measure=measure,
        interval=max(1, int(mon.interval)), t_start=float(mon.t_start), t_end=float(mon.t_end),
        window_code=1 if mon.window == "hann" else 0, index=int(mon.index),
        cell_range=tuple(int(v) for v in cell_range), ranges=plane_ranges)


def assign_offsets(compiled):
    """Ragged arena offsets (``monitors/compiler.py:619-636``)."""
    v = w = 0
    for m in compiled:
        m.value_offset, m.weight_offset = v, w
        v += 6 * m.freqs.size * m.points
        w += m.freqs.size
    return v, w


def pack_monitors(compiled):
    """Dense padded gather tables (``cuda/runtime.py:530-618``): indices
    s32[M, 6, Pmax, K], weights f32[M, 6, Pmax, K], freqs f64[M, Fmax],
    masks f32[M, 6], counts s32[M, 5], codes s32[M, 2], windows f32[M, 3]."""
    M = len(compiled)
    P = max(m.points for m in compiled)
    F = max(m.freqs.size for m in compiled)
    K = max(p[0].shape[1] for m in compiled for p in m.plans.values() if p is not None)
    idx = np.zeros((M, 6, P, K), dtype=np.int32)
    wts = np.zeros((M, 6, P, K), dtype=np.float32)
    freqs = np.zeros((M, F), dtype=np.float64)
    masks = np.zeros((M, 6), dtype=np.float32)
    counts = np.zeros((M, 5), dtype=np.int32)
    codes = np.zeros((M, 2), dtype=np.int32)
    windows = np.zeros((M, 3), dtype=np.float32)
    for i, m in enumerate(compiled):
        for c, name in enumerate(COMPONENTS):
            plan = m.plans.get(name)
            if plan is None:
                continue
            pi, pw = plan
            idx[i, c, :pi.shape[0], :pi.shape[1]] = pi
            wts[i, c, :pw.shape[0], :pw.shape[1]] = pw
            masks[i, c] = 1.0
        freqs[i, :m.freqs.size] = m.freqs
        counts[i] = (m.freqs.size, m.points, m.interval, m.value_offset, m.weight_offset)
        codes[i] = (m.window_code, 0)
        windows[i] = (m.t_start, m.t_end, 1e-6)
    return idx, wts, freqs, masks, counts, codes, windows


def phasors(m: CompiledMonitor, re, im, wsum, dt):
    """``observe.py:166-187``: ``2 DFT / wsum``, H shifted by ``exp(-i pi f dt)``."""
    F, P = m.freqs.size, m.points
    w = np.maximum(np.asarray(wsum[m.weight_offset:m.weight_offset + F], dtype=np.float64), 1e-18)
    out = {}
    for c, name in enumerate(COMPONENTS):
        if m.plans.get(name) is None:
            out[name] = None
            continue
        sl = slice(m.value_offset + c * F * P, m.value_offset + (c + 1) * F * P)
        v = (np.asarray(re[sl], dtype=np.float64) + 1j * np.asarray(im[sl], dtype=np.float64)).reshape(F, P)
        v = (2.0 / w[:, None]) * v
        if name.startswith("H"):
            v = v * np.exp(-1j * np.pi * m.freqs * dt)[:, None]
        out[name] = v
    return out


def flux(m: CompiledMonitor, re, im, wsum, dt):
    """``observe.py:198-234``: ``1/2 Re sum sign (E x H*)_n * measure``."""
    ph = phasors(m, re, im, wsum, dt)
    zero = np.zeros((m.freqs.size, m.points), dtype=complex)

    def f(c):
        return ph[c] if ph.get(c) is not None else zero

    if m.normal == "x":
        s = f("Ey") * np.conj(f("Hz")) - f("Ez") * np.conj(f("Hy"))
    elif m.normal == "y":
        s = f("Ez") * np.conj(f("Hx")) - f("Ex") * np.conj(f("Hz"))
    else:
        s = f("Ex") * np.conj(f("Hy")) - f("Ey") * np.conj(f("Hx"))
    return 0.5 * np.real(np.sum(m.normal_sign * s, axis=1)) * m.measure
