"""Yee-lattice geometry on the host: supports, material sampling, gather plans.

Host-side lowering only (NumPy, float64); nothing here runs per time step.

References: ``beamz/lattice.py:14-45`` (supports), ``:218-333`` (plane
interpolation plans), ``:732-786`` (E-site averaging), ``:919-951`` (2D site
sampling); ``beamz/devices/monitors/compiler.py:126-197`` (2D line plans).
"""

from __future__ import annotations

import numpy as np

COMPONENTS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")

# Yee offsets in (z, y, x) cell units, lattice.py:14-21.
OFFSETS_3D = {
    "Ex": (0.0, 0.0, 0.5),
    "Ey": (0.0, 0.5, 0.0),
    "Ez": (0.5, 0.0, 0.0),
    "Hx": (0.5, 0.5, 0.0),
    "Hy": (0.5, 0.0, 0.5),
    "Hz": (0.0, 0.5, 0.5),
}
# TMxy offsets in (y, x), lattice.py:905-911.
OFFSETS_TM = {"Ez": (0.0, 0.0), "Hx": (0.5, 0.0), "Hy": (0.0, 0.5)}

EDGES_3D = (("front", "back"), ("bottom", "top"), ("left", "right"))
EDGES_2D = (("bottom", "top"), ("left", "right"))
EDGE_BITS = {"front": 1, "back": 2, "bottom": 4, "top": 8, "left": 16, "right": 32}


def component_shape_3d(component, grid):
    """``lattice.py:32-45``: +1 along every axis where the offset is zero."""
    return tuple(int(n) + int(o == 0.0) for n, o in zip(grid, OFFSETS_3D[component]))


def component_shapes(grid):
    if len(grid) == 3:
        return {c: component_shape_3d(c, grid) for c in COMPONENTS}
    ny, nx = (int(v) for v in grid)
    return {"Ez": (ny + 1, nx + 1), "Hx": (ny, nx + 1), "Hy": (ny + 1, nx)}


def sample_e_site_3d(cell, component):
    """Centered average of a cell raster onto an E site (``lattice.py:732-786``)."""
    out = np.asarray(cell, dtype=np.float64)
    grid = out.shape
    shape = component_shape_3d(component, grid)
    for axis, (dim, gdim, off) in enumerate(zip(shape, grid, OFFSETS_3D[component])):
        coord = np.floor(np.arange(dim) + off).astype(np.int64)
        if off == 0.0:
            hi = np.clip(coord, 0, gdim - 1)
            lo = np.clip(coord - 1, 0, gdim - 1)
            out = 0.5 * (np.take(out, lo, axis=axis) + np.take(out, hi, axis=axis))
        else:
            out = np.take(out, np.clip(coord, 0, gdim - 1), axis=axis)
    return out


def sample_site_tm(cell, component):
    """Cell-ownership sampling onto a TMxy site (``lattice.py:919-951``)."""
    cell = np.asarray(cell, dtype=np.float64)
    ny, nx = cell.shape
    shape = component_shapes((ny, nx))[component]
    oy, ox = OFFSETS_TM[component]
    yi = np.clip(np.floor(np.arange(shape[0]) + oy).astype(np.int64), 0, ny - 1)
    xi = np.clip(np.floor(np.arange(shape[1]) + ox).astype(np.int64), 0, nx - 1)
    return cell[np.ix_(yi, xi)]


def interpolation_plan(source, target):
    """Bracketing indices and linear weights (``lattice.py:218-249``)."""
    source = np.asarray(source, dtype=np.float64).reshape(-1)
    target = np.asarray(target, dtype=np.float64).reshape(-1)
    if source.size == 1:
        zero = np.zeros(target.size, dtype=np.int64)
        return zero, zero, np.ones(target.size), np.zeros(target.size)
    high = np.clip(np.searchsorted(source, target, side="right"), 1, source.size - 1)
    low = high - 1
    below = target <= source[0]
    above = target >= source[-1]
    low[below] = 0
    high[below] = 0
    low[above] = source.size - 1
    high[above] = source.size - 1
    span = source[high] - source[low]
    frac = np.zeros_like(target)
    ok = (high != low) & (np.abs(span) > 0.0)
    frac[ok] = (target[ok] - source[low][ok]) / span[ok]
    w_low, w_high = 1.0 - frac, frac
    same = high == low
    w_low[same] = 1.0
    w_high[same] = 0.0
    return low, high, w_low, w_high


def plane_gather_plan(component, shape, dx, normal, plane_index, ranges):
    """Trilinear weighted gather of one component onto the tangential cell
    centres of an axis-aligned plane at ``plane_index * dx``
    (``lattice.py:252-333``, uniform grid).  Returns ``(index[P, K], weight[P, K])``
    with zero-weight corners compacted away."""
    tangential = {"x": ("z", "y"), "y": ("z", "x"), "z": ("y", "x")}[normal]
    u = (np.arange(*ranges[tangential[0]]) + 0.5) * dx
    v = (np.arange(*ranges[tangential[1]]) + 0.5) * dx
    return plane_gather_plan_at(component, shape, dx, normal, plane_index * dx, u, v)


def component_coords_3d(component, shape, dx=None, cell_grid=None):
    """Physical coordinates of a component's samples per axis name: nodes on
    its zero-offset axes, cell centres on the others (uniform spacing ``dx``,
    or the edges and centres of a ``RectilinearGrid``)."""
    pos = {"z": 0, "y": 1, "x": 2}
    out = {}
    for a, i in pos.items():
        off = OFFSETS_3D[component][i]
        if cell_grid is None:
            out[a] = (np.arange(shape[i]) + off) * dx
        else:
            out[a] = (cell_grid.axis_edges(a) if off == 0.0 else cell_grid.centers(a))[:shape[i]]
    return out


def plane_gather_plan_at(component, shape, dx, normal, position, u, v, coords=None):
    """The plane plan at normal coordinate ``position`` (metres) with samples
    on the tangential coordinate grid ``u x v`` (``lattice.py:252-354``): the
    stored plane order is ``(z, y)`` for x, ``(z, x)`` for y, ``(y, x)`` for z,
    and targets outside the component's coordinates take its edge value.
    ``coords`` gives the component's sample coordinates per axis (default:
    the uniform lattice of spacing ``dx``)."""
    tangential = {"x": ("z", "y"), "y": ("z", "x"), "z": ("y", "x")}[normal]
    pos = {"z": 0, "y": 1, "x": 2}
    if coords is None:
        coords = component_coords_3d(component, shape, dx)
    gu, gv = np.meshgrid(np.asarray(u, dtype=np.float64), np.asarray(v, dtype=np.float64),
                         indexing="ij")
    plans = (
        interpolation_plan(coords[tangential[0]], gu.ravel()),
        interpolation_plan(coords[tangential[1]], gv.ravel()),
        interpolation_plan(coords[normal], np.full(gu.size, float(position))),
    )
    corners = np.array([(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)])
    index = {}
    weight = np.ones((gu.size, 8))
    for d, (lo, hi, wl, wh) in enumerate(plans):
        bit = corners[:, d][None, :]
        index[d] = np.where(bit == 0, lo[:, None], hi[:, None])
        weight = weight * np.where(bit == 0, wl[:, None], wh[:, None])
    by_axis = {tangential[0]: index[0], tangential[1]: index[1], normal: index[2]}
    flat = np.ravel_multi_index((by_axis["z"], by_axis["y"], by_axis["x"]), shape)
    live = np.abs(weight) > 4.0 * np.finfo(np.float32).eps
    weight = np.where(live, weight, 0.0)
    total = weight.sum(axis=1, keepdims=True)
    weight = np.divide(weight, total, out=np.zeros_like(weight), where=np.abs(total) > 0)
    width = int(np.max(live.sum(axis=1), initial=0))
    if 0 < width < 8:
        order = np.argsort(~live, axis=1, kind="stable")[:, :width]
        flat = np.take_along_axis(flat, order, axis=1)
        weight = np.take_along_axis(weight, order, axis=1)
    return flat.astype(np.int64), weight


def line_gather_plan(component, shape, dx, normal, line_index, cell_range, cell_grid=None):
    """Bilinear gather onto the tangential cell centres of an axis-aligned 2D
    line at the E-node ``line_index`` (``monitors/compiler.py:126-197``), on
    the uniform lattice of spacing ``dx`` or on a ``RectilinearGrid``."""
    oy, ox = OFFSETS_TM[component]
    if cell_grid is None:
        ys = (np.arange(shape[0]) + oy) * dx
        xs = (np.arange(shape[1]) + ox) * dx
        t = (np.arange(*cell_range) + 0.5) * dx
        at = line_index * dx
    else:
        ys = (cell_grid.axis_edges("y") if oy == 0.0 else cell_grid.centers("y"))[:shape[0]]
        xs = (cell_grid.axis_edges("x") if ox == 0.0 else cell_grid.centers("x"))[:shape[1]]
        tang = "y" if normal == "x" else "x"
        t = cell_grid.centers(tang)[cell_range[0]:cell_range[1]]
        at = cell_grid.axis_edges(normal)[line_index]
    if normal == "x":
        px, py = np.full(t.shape, at), t
    else:
        px, py = t, np.full(t.shape, at)
    x0, x1, wx0, wx1 = interpolation_plan(xs, px)
    y0, y1, wy0, wy1 = interpolation_plan(ys, py)
    bits = np.array([(0, 0), (0, 1), (1, 0), (1, 1)])
    xi = np.where(bits[:, 1][None] == 0, x0[:, None], x1[:, None])
    yi = np.where(bits[:, 0][None] == 0, y0[:, None], y1[:, None])
    wx = np.where(bits[:, 1][None] == 0, wx0[:, None], wx1[:, None])
    wy = np.where(bits[:, 0][None] == 0, wy0[:, None], wy1[:, None])
    return np.ravel_multi_index((yi, xi), shape).astype(np.int64), wx * wy
