"""Rectilinear (non-uniform) grids and the derivative metrics of the Yee step.

A ``RectilinearGrid`` holds the physical cell edges of each axis, as the
reference's realized grid does (``beamz/design/grid.py:100-398``): cell widths,
centres, the classification of the derivative metric (isotropic uniform, axis
uniform or rectilinear) and the CFL time step follow it.

``derivative_metrics`` restates ``_compile_derivative_metrics``
(``beamz/simulation/compile.py:209-240``): the inverse distances that scale
every staggered difference of the curl.  E -> H (forward) differences span one
cell, so their metric is 1 / w_i per cell; H -> E (backward) differences span
the dual cell between two cell centres, (w_{i-1} + w_i) / 2, and the two end
samples, whose ghost lies half a cell outside, take 1 / w_0 and 1 / w_{n-1}.
Axis-uniform grids carry one scalar per axis, isotropic ones none (the
kernels then use 1 / resolution).
"""

from __future__ import annotations

import numpy as np

from .constants import LIGHT_SPEED

AXES = ("x", "y", "z")


def _axis_index(axis) -> int:
    if isinstance(axis, str):
        return AXES.index(axis)
    return int(axis)


def _edges(values, name):
    arr = np.array(values, dtype=np.float64).reshape(-1)
    if arr.size < 2:
        raise ValueError(f"{name} needs at least two edges")
    if not np.all(np.isfinite(arr)) or np.any(np.diff(arr) <= 0.0):
        raise ValueError(f"{name} must be finite and strictly increasing")
    arr.setflags(write=False)
    return arr


def _widths_are_equal(widths, reference, edges) -> bool:
    """``grid.py:87-96``: equal within 1e-12 relative plus the roundoff floor of
    the coordinate magnitude."""
    roundoff = 8.0 * np.finfo(edges.dtype).eps * float(np.max(np.abs(edges), initial=0.0))
    return bool(np.max(np.abs(widths - reference), initial=0.0) <= 1e-12 * abs(reference) + roundoff)


class RectilinearGrid:
    """A Cartesian grid given by physical x, y and z cell edges in metres.

    2D (TMxy) simulations use the x and y axes; ``z_edges`` may then be
    omitted."""

    def __init__(self, x_edges, y_edges, z_edges=(0.0, 1.0)):
        self.x_edges = _edges(x_edges, "x_edges")
        self.y_edges = _edges(y_edges, "y_edges")
        self.z_edges = _edges(z_edges, "z_edges")

    @classmethod
    def from_spacing(cls, shape, spacing, *, origin=(0.0, 0.0, 0.0)):
        """Constant per-axis spacings from a lower corner (``grid.py:143-175``);
        ``shape`` is the cell count in (x, y, z) order."""
        counts = tuple(int(v) for v in shape)
        if len(counts) != 3 or any(c < 1 for c in counts):
            raise ValueError("shape must hold three positive cell counts")
        spacings = tuple(float(v) for v in spacing) if isinstance(spacing, (tuple, list)) \
            else (float(spacing),) * 3
        if len(spacings) != 3 or any(not np.isfinite(v) or v <= 0.0 for v in spacings):
            raise ValueError("grid spacing must contain three positive finite values")
        return cls(*(float(origin[i]) + spacings[i] * np.arange(counts[i] + 1, dtype=np.float64)
                     for i in range(3)))

    @classmethod
    def from_widths(cls, x_widths, y_widths, z_widths=(1.0,), *, origin=(0.0, 0.0, 0.0)):
        """Edges from per-cell widths along each axis."""
        ws = [np.asarray(w, dtype=np.float64).reshape(-1) for w in (x_widths, y_widths, z_widths)]
        return cls(*(float(origin[i]) + np.concatenate([[0.0], np.cumsum(ws[i])]) for i in range(3)))

    def axis_edges(self, axis) -> np.ndarray:
        return (self.x_edges, self.y_edges, self.z_edges)[_axis_index(axis)]

    def cell_widths(self, axis) -> np.ndarray:
        return np.diff(self.axis_edges(axis))

    def centers(self, axis) -> np.ndarray:
        e = self.axis_edges(axis)
        return 0.5 * (e[:-1] + e[1:])

    @property
    def shape(self):
        """Cell counts in (x, y, z) order."""
        return tuple(int(e.size - 1) for e in (self.x_edges, self.y_edges, self.z_edges))

    @property
    def shape_zyx(self):
        x, y, z = self.shape
        return z, y, x

    def min_spacings(self, axes=AXES):
        return tuple(float(np.min(self.cell_widths(a))) for a in axes)

    def minimum_spacing(self, axes=AXES) -> float:
        return min(self.min_spacings(axes))

    def metric_kind_for(self, axes=AXES) -> str:
        """``grid.py:250-267``: ``isotropic_uniform`` when every active axis has
        one common width, ``axis_uniform`` when each has its own, else
        ``rectilinear``."""
        widths = [self.cell_widths(a) for a in axes]
        reference = float(widths[0][0])
        if all(_widths_are_equal(w, reference, self.axis_edges(a)) for a, w in zip(axes, widths)):
            return "isotropic_uniform"
        if all(_widths_are_equal(w, float(w[0]), self.axis_edges(a)) for a, w in zip(axes, widths)):
            return "axis_uniform"
        return "rectilinear"

    def cfl_time_step(self, courant=1.0, *, axes=AXES) -> float:
        """``grid.py:358-371``: courant / (c sqrt(sum over axes of 1 / min width^2))."""
        inv = sum(1.0 / float(np.min(self.cell_widths(a))) ** 2 for a in axes)
        return float(courant) / (LIGHT_SPEED * float(np.sqrt(inv)))


METRIC_KIND_CODE = {"isotropic_uniform": 0, "axis_uniform": 1, "rectilinear": 2}


def derivative_metrics(grid: RectilinearGrid, axes=AXES):
    """``compile.py:209-240``: ``(kind, forward, backward)`` with ``forward[a]``
    and ``backward[a]`` float32 arrays per active axis name (0-d for axis-uniform
    grids); both dicts are empty for isotropic uniform grids."""
    kind = grid.metric_kind_for(axes)
    forward, backward = {}, {}
    if kind == "isotropic_uniform":
        return kind, forward, backward
    for a in axes:
        widths = grid.cell_widths(a)
        if kind == "axis_uniform":
            forward[a] = backward[a] = np.asarray(1.0 / float(widths[0]), dtype=np.float32)
            continue
        inv_back = np.empty(widths.size + 1, dtype=np.float64)
        inv_back[0] = 1.0 / widths[0]
        inv_back[-1] = 1.0 / widths[-1]
        if widths.size > 1:
            inv_back[1:-1] = 2.0 / (widths[:-1] + widths[1:])
        forward[a] = np.asarray(1.0 / widths, dtype=np.float32)
        backward[a] = np.asarray(inv_back, dtype=np.float32)
    return kind, forward, backward
