"""DFT flux monitors: host lowering to weighted-gather plans, packing into the
ragged accumulator arenas, and flux post-processing.

References: ``beamz/devices/monitors/compiler.py:61-102, 233-636`` (specs,
arenas), ``beamz/simulation/cuda/runtime.py:530-618`` (packing),
``beamz/simulation/observe.py:166-234`` (phasors and flux),
``beamz/devices/monitors/monitors.py:61-68`` (2D line element).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .lattice import COMPONENTS, line_gather_plan, plane_gather_plan, plane_gather_plan_at


@dataclass(frozen=True)
class FluxMonitor:
    """An axis-aligned DFT flux plane (3D) or line (2D), placed in index space
    or, like the reference's ``FluxMonitor(center=, size=, freqs=)``
    (``monitors/monitors.py:432-450``), by a physical centre and size.

    Index space: ``normal`` is ``"x"``, ``"y"`` or ``"z"``; the plane sits on the
    E-node plane ``index`` (position ``index * dx``); ``ranges`` maps each
    tangential axis to a ``(start, stop)`` cell interval (default: the whole
    axis).  Samples are taken at tangential cell centres.

    Physical placement: ``center`` and ``size`` in metres, 2 or 3 values each (a
    2-value tuple gets a zero z entry), exactly one zero extent, which names the
    normal.  The monitor snaps to the grid when the simulation is lowered, as
    the reference does on a uniform grid at the origin (see ``lower_placed``).

    ``components`` restricts the accumulated components (default: the four
    tangential ones that enter the flux)."""

    normal: str = None
    index: int = None
    freqs: np.ndarray = None
    ranges: dict = field(default_factory=dict)
    components: tuple = None
    normal_sign: float = 1.0
    interval: int = 1
    t_start: float = 0.0
    t_end: float = float("inf")
    window: str = "rect"
    name: str = None
    center: tuple = None
    size: tuple = None

    def __post_init__(self):
        if self.freqs is None:
            raise ValueError("FluxMonitor requires frequencies.")
        if self.center is None and self.size is None:
            if self.normal not in ("x", "y", "z") or self.index is None:
                raise ValueError("an index-space FluxMonitor needs normal and index; "
                                 "otherwise give center and size")
            return
        if self.center is None or self.size is None:
            raise ValueError("FluxMonitor placement requires both center and size.")
        if self.index is not None or self.ranges:
            raise ValueError("give either center/size or index/ranges, not both")
        center, size = normalize_center_size(self.center, self.size)
        zero = np.isclose(np.asarray(size), 0.0, rtol=0.0, atol=1e-15)
        if np.count_nonzero(zero) != 1:
            raise ValueError("FluxMonitor size must contain exactly one zero extent.")
        normal = "xyz"[int(np.argmin(np.abs(np.asarray(size))))]
        if self.normal is not None and self.normal != normal:
            raise ValueError(f"normal {self.normal!r} disagrees with the zero extent of size")
        object.__setattr__(self, "center", center)
        object.__setattr__(self, "size", size)
        object.__setattr__(self, "normal", normal)

    @property
    def placed(self):
        return self.center is not None


@dataclass(frozen=True)
class FieldMonitor(FluxMonitor):
    """DFT fields of chosen components on a placed plane or line (the
    reference's ``FieldMonitor``, ``monitors/monitors.py:341-367``): at least
    one zero extent (the first names the normal), ``fields`` any of the six
    components (in 2D the TMxy ones are recorded).  Lowered as a placed
    ``FluxMonitor``; ``SimulationResults.monitor_phasors`` returns the fields."""

    fields: tuple = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")

    def __post_init__(self):
        if self.freqs is None or np.asarray(self.freqs).size == 0:
            raise ValueError("FieldMonitor requires at least one frequency.")
        if self.center is None or self.size is None:
            raise ValueError("FieldMonitor placement requires both center and size.")
        center, size = normalize_center_size(self.center, self.size)
        if not np.any(np.isclose(np.asarray(size), 0.0, rtol=0.0, atol=1e-15)):
            raise ValueError("FieldMonitor size must contain at least one zero extent.")
        fields = tuple(str(c) for c in self.fields)
        invalid = set(fields) - set(COMPONENTS)
        if not fields or invalid:
            raise ValueError(f"Unsupported FieldMonitor fields: {sorted(invalid)}")
        object.__setattr__(self, "center", center)
        object.__setattr__(self, "size", size)
        object.__setattr__(self, "normal", "xyz"[int(np.argmin(np.abs(np.asarray(size))))])
        object.__setattr__(self, "fields", fields)
        object.__setattr__(self, "components", fields)


def normalize_center_size(center, size):
    """3-tuples of a monitor's centre and size (``monitors.py:20-29``): finite
    centre, finite non-negative size, a 2-value tuple gets a zero z entry."""
    center = tuple(float(v) for v in center)
    size = tuple(float(v) for v in size)
    if not all(np.isfinite(center)):
        raise ValueError("Monitor center must be finite.")
    if not all(np.isfinite(size)) or any(v < 0.0 for v in size):
        raise ValueError("Monitor size must be finite and non-negative.")
    if len(center) == 2:
        center = center + (0.0,)
    if len(size) == 2:
        size = size + (0.0,)
    if len(center) != 3 or len(size) != 3:
        raise ValueError("center and size must be 2D or 3D coordinate tuples.")
    return center, size


def snap_cell_center(coord, step, count):
    """The cell whose centre is nearest ``coord`` (``_placement.py:151-156``):
    ``round(coord / step - 1/2)``, ties to even, clamped to ``[0, count)``.
    Returns ``(index, (index + 1/2) step)``."""
    if count <= 0:
        return 0, 0.0
    idx = int(round(float(coord) / float(step) - 0.5))
    idx = max(0, min(idx, count - 1))
    return idx, (idx + 0.5) * float(step)


def snap_edge_interval(lower, upper, step, count, min_cells=1):
    """The cells ``[start, stop)`` that ``[lower, upper]`` overlaps
    (``_placement.py:159-189``): bounds within 1e-12 (relative) of a cell edge
    count as on it; clamped to ``[0, count]``; widened to ``min_cells`` cells
    around the centre when narrower."""
    if count <= 0:
        return 0, 0
    step = float(step)
    lo, hi = sorted((float(lower), float(upper)))
    eps = 1e-12 * max(abs(lo), abs(hi), abs(step), 1.0)
    start = max(0, min(int(math.floor((lo + eps) / step)), count))
    stop = max(0, min(int(math.ceil((hi - eps) / step)), count))
    need = max(1, int(min_cells))
    if stop - start < need:
        cell = max(0, min(int(math.floor(0.5 * (lo + hi) / step)), count - 1))
        stop = min(count, cell + need)
        start = max(0, stop - need)
    return start, stop


def uniform_axis_centers(lower, upper, count):
    """``count`` equal sub-interval centres of ``[lower, upper]``
    (``lattice.py:203-211``)."""
    count = max(0, int(count))
    if count == 0:
        return np.zeros(0)
    if count == 1:
        return np.array([0.5 * (float(lower) + float(upper))])
    step = (float(upper) - float(lower)) / count
    return float(lower) + (np.arange(count) + 0.5) * step


def plane_sample_area(coords_u, coords_v, fallback):
    """Area per plane sample (``lattice.py:357-366``): per axis the median
    spacing times n / (n - 1), or ``fallback`` for fewer than two samples."""
    def step(values):
        values = np.asarray(values, dtype=np.float64).reshape(-1)
        if values.size > 1:
            spacing = float(np.median(np.abs(np.diff(values))))
            if np.isfinite(spacing) and spacing > 0.0:
                return spacing * values.size / (values.size - 1)
        return float(fallback)
    return step(coords_u) * step(coords_v)


def line_grid_points_2d(start, end, dx):
    """Lattice points of a 2D line on a uniform grid (``monitors.py:148-174``,
    the path taken without a snapped region): end points rounded to the
    nearest node, ties to even, and one point per node along the longer
    side, both ends included.  Returns ``(x, y)`` index arrays."""
    sx, sy = int(round(start[0] / dx)), int(round(start[1] / dx))
    ex, ey = int(round(end[0] / dx)), int(round(end[1] / dx))
    n = abs(ex - sx) + 1 if abs(ex - sx) > abs(ey - sy) else abs(ey - sy) + 1
    return (np.linspace(sx, ex, n, dtype=int), np.linspace(sy, ey, n, dtype=int))


def line_normal_2d(start, end):
    """The signed normal of an axis-aligned 2D line from ``start`` to ``end``
    (``monitors.py:42-55``): ``("x", +1)`` for a line along +y, ``("y", -1)``
    for a line along +x; None for a point or a slanted line."""
    ddx = float(end[0]) - float(start[0])
    ddy = float(end[1]) - float(start[1])
    tol = 1e-12 * max(abs(ddx), abs(ddy), 1.0)
    if abs(ddx) <= tol and abs(ddy) > tol:
        return "x", 1.0 if ddy >= 0.0 else -1.0
    if abs(ddy) <= tol and abs(ddx) > tol:
        return "y", -1.0 if ddx >= 0.0 else 1.0
    return None


def direct_line_plan(shape, xs, ys):
    """One lattice point per sample (``monitors/compiler.py:105-110``,
    ``219-224``): indices clipped into the component's array, weight 1 inside
    it and 0 outside."""
    h, w = shape
    valid = (xs >= 0) & (xs < w) & (ys >= 0) & (ys < h)
    x = np.clip(xs, 0, max(w - 1, 0))
    y = np.clip(ys, 0, max(h - 1, 0))
    flat = np.ravel_multi_index((y, x), shape).astype(np.int64)
    return flat[:, None], valid.astype(np.float64)[:, None]


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
    placed: bool = False       # lowered from a physical centre and size
    position: float = None     # the normal coordinate of the samples (metres)
    samples: tuple = ()        # placed 3D plane: the two tangential sample coordinate arrays
    point_measure: np.ndarray = None  # per-sample area (width in 2D) on a non-uniform grid


def _frequencies(mon):
    freqs = np.asarray(mon.freqs, dtype=np.float64).reshape(-1)
    if freqs.size == 0 or np.any(~np.isfinite(freqs)) or np.any(freqs <= 0):
        raise ValueError("monitor frequencies must be finite and positive")
    return freqs


def _points(plans, wanted, is_3d):
    """Sample count of a lowered monitor; a monitor none of whose components
    exists in this dimensionality (a 2D ``fields=("Ex",)``) is refused."""
    for p in plans.values():
        if p is not None:
            return p[0].shape[0]
    kind = "3D" if is_3d else "2D TMxy (Ez, Hx, Hy)"
    raise ValueError(f"monitor fields {tuple(wanted)} include no component of a {kind} simulation")


def lower_monitor(mon: FluxMonitor, grid, shapes, dx, index):
    if mon.placed:
        return lower_placed(mon, grid, shapes, dx, index)
    is_3d = len(grid) == 3
    freqs = _frequencies(mon)
    wanted = mon.components or FLUX_COMPONENTS[mon.normal]
    plans = {}
    if is_3d:
        axes = {"z": 0, "y": 1, "x": 2}
        tangential = {"x": ("z", "y"), "y": ("z", "x"), "z": ("y", "x")}[mon.normal]
        ranges = {a: tuple(mon.ranges.get(a, (0, grid[axes[a]]))) for a in tangential}
        for c in COMPONENTS:
            plans[c] = (plane_gather_plan(c, shapes[c], dx, mon.normal, mon.index, ranges)
                        if c in wanted else None)
        measure = dx * dx  # the cell face: one sample per cell of the plane
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
    points = _points(plans, wanted, is_3d)
    return CompiledMonitor(
        name=mon.name or f"monitor_{index}", freqs=freqs, plans=plans, points=points,
        normal=mon.normal, normal_sign=float(mon.normal_sign), measure=measure,
        interval=max(1, int(mon.interval)), t_start=float(mon.t_start), t_end=float(mon.t_end),
        window_code=1 if mon.window == "hann" else 0, index=int(mon.index),
        cell_range=tuple(int(v) for v in cell_range), ranges=plane_ranges,
        position=float(mon.index) * dx)


def lower_placed(mon: FluxMonitor, grid, shapes, dx, index):
    """A monitor placed by centre and size, lowered as the reference lowers it
    on a uniform grid at the origin (``monitors/compiler.py:469-616``).

    3D (``_placement.py:462-526``, ``lattice.py:214-238``, ``397-446``): each
    tangential extent snaps to the cells it overlaps; the samples sit at the
    centres of that many equal parts of the physical extent (the cell centres
    when the extent lies on cell edges), on the unsnapped normal coordinate
    ``center[normal]``, interpolated trilinearly; the area per sample is
    ``plane_sample_area``; the sign is +1.

    2D (``monitors.py:148-174``, ``compiler.py:113-131``, ``486-545``): the
    line from the lower to the upper corner is rounded to lattice nodes, both
    ends included, and every component is read at those nodes directly (no
    interpolation; points outside a component's array get weight 0).  The
    sign follows the line's direction: +1 for an x-normal line, -1 for a
    y-normal one; the line element is dx."""
    freqs = _frequencies(mon)
    wanted = mon.components or FLUX_COMPONENTS[mon.normal]
    center, size = mon.center, mon.size
    lower = tuple(c - 0.5 * s for c, s in zip(center, size))
    upper = tuple(c + 0.5 * s for c, s in zip(center, size))
    at = {"x": 0, "y": 1, "z": 2}
    plans = {}
    if len(grid) == 3:
        counts = {"z": grid[0], "y": grid[1], "x": grid[2]}
        tangential = {"x": ("z", "y"), "y": ("z", "x"), "z": ("y", "x")}[mon.normal]
        ranges, coords = {}, []
        for a in tangential:
            lo, hi = lower[at[a]], upper[at[a]]
            ranges[a] = snap_edge_interval(lo, hi, dx, counts[a])
            coords.append(uniform_axis_centers(lo, hi, ranges[a][1] - ranges[a][0]))
        position = center[at[mon.normal]]
        for c in COMPONENTS:
            plans[c] = (plane_gather_plan_at(c, shapes[c], dx, mon.normal, position, *coords)
                        if c in wanted else None)
        plane_index = snap_cell_center(position, dx, counts[mon.normal])[0]
        measure = plane_sample_area(coords[0], coords[1], dx)
        sign, cell_range = 1.0, ()
        plane_ranges = tuple((a, tuple(int(v) for v in ranges[a])) for a in tangential)
        samples = tuple(coords)
    else:
        if mon.normal not in ("x", "y"):
            raise ValueError("a 2D line monitor is normal to x or y")
        start = (center[0], lower[1]) if mon.normal == "x" else (lower[0], center[1])
        end = (center[0], upper[1]) if mon.normal == "x" else (upper[0], center[1])
        xs, ys = line_grid_points_2d(start, end, dx)
        for c in COMPONENTS:
            active = c in ("Ez", "Hx", "Hy") and c in wanted
            plans[c] = direct_line_plan(shapes[c], xs, ys) if active else None
        sign = (line_normal_2d(start, end) or (mon.normal, 1.0))[1]
        measure = dx
        position = center[at[mon.normal]]
        plane_index = int(xs[0]) if mon.normal == "x" else int(ys[0])
        tang = ys if mon.normal == "x" else xs
        cell_range = (int(tang[0]), int(tang[-1]) + 1)
        plane_ranges, samples = (), ()
    points = _points(plans, wanted, len(grid) == 3)
    return CompiledMonitor(
        name=mon.name or f"monitor_{index}", freqs=freqs, plans=plans, points=points,
        normal=mon.normal, normal_sign=sign * float(mon.normal_sign), measure=measure,
        interval=max(1, int(mon.interval)), t_start=float(mon.t_start), t_end=float(mon.t_end),
        window_code=1 if mon.window == "hann" else 0, index=int(plane_index),
        cell_range=cell_range, ranges=plane_ranges, placed=True, position=float(position),
        samples=samples)


def assign_offsets(compiled):
    """Ragged arena offsets (``monitors/compiler.py:619-636``)."""
    v = w = 0
    for m in compiled:
        m.value_offset, m.weight_offset = v, w
        v += 6 * m.freqs.size * m.points
        w += m.freqs.size
    return v, w


def capture_boxes(idx, wts, masks, shapes):
    """Inclusive (z0, z1, y0, y1, x0, x1) bounding boxes s32[M, 6, 6] of the
    cells each (monitor, component) gather reads: corners with a non-zero
    weight whose index lies inside the component (the kernels' gather rule),
    in the DFT component order Ex, Ey, Ez, Hx, Hy, Hz.  A component that no
    gather reads (mask zero or no live corner) has the row -1.  The pair
    kernel captures exactly these cells of both its steps, so that the
    monitors' samples are gathered from a dense buffer instead of strided
    field rows (no counterpart in the reference, which samples every step
    from the fields)."""
    M = idx.shape[0]
    out = np.full((M, 6, 6), -1, dtype=np.int32)
    for m in range(M):
        for c, name in enumerate(COMPONENTS):
            shape = shapes.get(name)
            if shape is None or len(shape) != 3 or masks[m, c] == 0.0:
                continue
            i = idx[m, c].reshape(-1).astype(np.int64)
            w = wts[m, c].reshape(-1)
            live = (w != 0.0) & (i >= 0) & (i < int(np.prod(shape)))
            if not live.any():
                continue
            z, y, x = np.unravel_index(i[live], shape)
            out[m, c] = (z.min(), z.max(), y.min(), y.max(), x.min(), x.max())
    return out


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
    if m.point_measure is not None:
        return 0.5 * np.real(np.sum(m.normal_sign * s * m.point_measure[None, :], axis=1)) * m.measure
    return 0.5 * np.real(np.sum(m.normal_sign * s, axis=1)) * m.measure
