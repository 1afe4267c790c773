"""FP64 NumPy restatement of BeamZ's Yee/CPML time step (the oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg, never by the ``beamz_amd`` package.

Every function cites the reference file:line whose behaviour it restates.
The reference runs this arithmetic in JAX/XLA (float32 fields); here every
array is float64 so the GPU result can be compared with an exact-arithmetic
stand-in.  Differences from the reference that are deliberate:

* time is ``t0 + (step + 1) * dt`` evaluated in float64 instead of a running
  float32 sum (``execute.py:527-541``), so the DFT phase has no clock drift;
* CPML profiles are evaluated in float64 (the reference samples them in
  float32, ``_boundary_compile.py:498-614``).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

# beamz/const.py:1-5
LIGHT_SPEED = 299792458.0
EPS_0 = 8.8541878128e-12
MU_0 = 1.25663706212e-6
ETA_0 = float(np.sqrt(MU_0 / EPS_0))

COMPONENTS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
AXIS_EDGES_3D = (("front", "back"), ("bottom", "top"), ("left", "right"))
AXIS_EDGES_2D = (("bottom", "top"), ("left", "right"))

# lattice.py:14-21 -- Yee offsets in (z, y, x) cell units.
OFFSETS_3D = {
    "Ex": (0.0, 0.0, 0.5),
    "Ey": (0.0, 0.5, 0.0),
    "Ez": (0.5, 0.0, 0.0),
    "Hx": (0.5, 0.5, 0.0),
    "Hy": (0.5, 0.0, 0.5),
    "Hz": (0.0, 0.5, 0.5),
}
# lattice.py:905-911 -- TMxy offsets in (y, x) cell units.
OFFSETS_TM = {"Ez": (0.0, 0.0), "Hx": (0.5, 0.0), "Hy": (0.0, 0.5)}

# _boundary_compile.py:44-60 -- the twelve 3D derivative terms in kernel order:
# (target, source field, derivative axis); signs alternate +,- (compile.py:273-274).
H_TERMS_3D = (
    ("Hx", "Ez", 1), ("Hx", "Ey", 0),
    ("Hy", "Ex", 0), ("Hy", "Ez", 2),
    ("Hz", "Ey", 2), ("Hz", "Ex", 1),
)
E_TERMS_3D = (
    ("Ex", "Hz", 1), ("Ex", "Hy", 0),
    ("Ey", "Hx", 0), ("Ey", "Hz", 2),
    ("Ez", "Hy", 2), ("Ez", "Hx", 1),
)
# compile.py:287-295 -- 2D TMxy terms: (target, source, axis, sign) on (y, x).
H_TERMS_TM = (("Hx", "Ez", 0, 1.0), ("Hy", "Ez", 1, -1.0))
E_TERMS_TM = (("Ez", "Hy", 1, 1.0), ("Ez", "Hx", 0, -1.0))


# --------------------------------------------------------------------------
# Lattice
# --------------------------------------------------------------------------
def shapes_3d(grid):
    """Complete Yee supports, ``lattice.py:32-45``."""
    return {
        c: tuple(int(n) + int(o == 0.0) for n, o in zip(grid, OFFSETS_3D[c]))
        for c in COMPONENTS
    }


def shapes_tm(grid):
    """TMxy supports, ``lattice.py:789-813``."""
    ny, nx = grid
    return {"Ez": (ny + 1, nx + 1), "Hx": (ny, nx + 1), "Hy": (ny + 1, nx)}


def e_site_permittivity_3d(cell, component):
    """Centered average of cell values onto an E site, ``lattice.py:732-786``."""
    out = np.asarray(cell, dtype=np.float64)
    grid = out.shape
    shape = shapes_3d(grid)[component]
    for axis, (dim, gdim, off) in enumerate(zip(shape, grid, OFFSETS_3D[component])):
        coord = np.arange(dim, dtype=np.float64) + off
        if off == 0.0:
            hi = np.clip(np.floor(coord).astype(np.int64), 0, gdim - 1)
            lo = np.clip(np.floor(coord).astype(np.int64) - 1, 0, gdim - 1)
            out = 0.5 * (np.take(out, lo, axis=axis) + np.take(out, hi, axis=axis))
        else:
            idx = np.clip(np.floor(coord).astype(np.int64), 0, gdim - 1)
            out = np.take(out, idx, axis=axis)
    return out


def site_sample_tm(cell, component):
    """Cell ownership sampling onto a TMxy site, ``lattice.py:919-951``."""
    cell = np.asarray(cell, dtype=np.float64)
    ny, nx = cell.shape
    shape = shapes_tm((ny, nx))[component]
    oy, ox = OFFSETS_TM[component]
    yi = np.clip(np.floor(np.arange(shape[0]) + oy).astype(np.int64), 0, ny - 1)
    xi = np.clip(np.floor(np.arange(shape[1]) + ox).astype(np.int64), 0, nx - 1)
    return cell[np.ix_(yi, xi)]


# --------------------------------------------------------------------------
# Boundaries: PEC masks, CPML profiles and packed terms
# --------------------------------------------------------------------------
def extrude_into_cpml(cell, cells, faces, axis_edges):
    """Copy the first interior slice through each active CPML slab,
    ``_boundary_compile.py:334-385`` (count = ``cells`` on each face)."""
    out = np.array(cell, dtype=np.float64, copy=True)
    for axis, (low_edge, high_edge) in enumerate(axis_edges):
        n = out.shape[axis]
        if cells <= 0 or cells >= n:
            continue
        if low_edge in faces:
            src = np.take(out, [cells], axis=axis)
            sl = [slice(None)] * out.ndim
            sl[axis] = slice(0, cells)
            out[tuple(sl)] = src
        if high_edge in faces:
            src = np.take(out, [n - cells - 1], axis=axis)
            sl = [slice(None)] * out.ndim
            sl[axis] = slice(n - cells, n)
            out[tuple(sl)] = src
    return out


def cpml_sigma_max(m, reflection, thickness_m):
    """Automatic conductivity, ``_boundary_compile.py:244-253``."""
    return -(m + 1) * np.log(max(reflection, 1e-16)) / (2.0 * ETA_0 * thickness_m)


def cpml_profile_1d(samples, kind, domain_cells, pml_cells, low, high,
                    sigma_max, kappa_max, alpha_max, m, alpha_order=1.0):
    """Staggered graded profile, ``_boundary_compile.py:425-614`` (uniform grid)."""
    coords = np.arange(samples, dtype=np.float64) + (0.0 if kind == "E" else 0.5)
    sigma = np.zeros(samples)
    kappa = np.ones(samples)
    alpha = np.zeros(samples)
    L = float(max(pml_cells, 1))

    def side(dist):
        u = np.clip(dist / L, 0.0, 1.0)
        return (dist > 0.0, sigma_max * u ** m, 1.0 + (kappa_max - 1.0) * u ** m,
                alpha_max * (1.0 - u) ** alpha_order)

    for active, dist in (
        (low, np.clip(L - coords, 0.0, L)),
        (high, np.clip(coords - float(domain_cells - pml_cells), 0.0, L)),
    ):
        if not active:
            continue
        mask, s, k, a = side(dist)
        sigma = np.where(mask, np.maximum(sigma, s), sigma)
        kappa = np.where(mask, np.maximum(kappa, k), kappa)
        alpha = np.where(mask, np.maximum(alpha, a), alpha)
    return sigma, kappa, alpha


def cpml_coefficients(sigma, kappa, alpha, dt):
    """Recurrence coefficients ``a, b, 1/kappa``, ``kernels.py:130-144``."""
    kappa = np.maximum(kappa, 1.0)
    b = np.exp(-(sigma / kappa + alpha) * dt / EPS_0)
    a = np.nan_to_num(((b - 1.0) * sigma) / np.maximum((sigma + kappa * alpha) * kappa, 1e-30))
    return a, b, 1.0 / kappa


def active_slab_counts(a, inv_kappa):
    """Leading/trailing active run lengths, ``kernels.py:225-234`` (1-D)."""
    active = (np.abs(a) > 1e-30) | (np.abs(inv_kappa - 1.0) > 1e-7)
    low = next((i for i, v in enumerate(active) if not v), active.size)
    tail = active[low:][::-1]
    high = next((i for i, v in enumerate(tail) if not v), tail.size)
    return int(low), int(high)


@dataclass
class Term:
    """One packed CPML derivative, ``model.py:306-324`` / ``kernels.py:237-254``.

    ``a, b, ik`` are the packed 1-D coefficient vectors (length low+high) along
    ``axis``; ``n`` is the target extent along that axis."""

    target: str
    source: str
    axis: int
    sign: float
    low: int
    high: int
    n: int
    a: np.ndarray
    b: np.ndarray
    ik: np.ndarray

    def pack(self, full):
        return np.concatenate([full[: self.low], full[self.n - self.high:]])


def _make_term(target, source, axis, sign, shape, prof, dt):
    a, b, ik = cpml_coefficients(*prof, dt)
    low, high = active_slab_counts(a, ik)
    n = shape[axis]
    t = Term(target, source, axis, sign, low, high, n, a, b, ik)
    t.a, t.b, t.ik = t.pack(a), t.pack(b), t.pack(ik)
    return t


@dataclass
class CpmlSpec:
    """``PML(formulation='cpml')`` parameters, ``boundaries.py:96-186``."""

    cells: int
    kappa_max: float = 2.0
    alpha_norm: Optional[float] = None
    m: int = 3
    reflection: float = 1e-6
    faces: frozenset = frozenset(
        ("front", "back", "bottom", "top", "left", "right"))


def build_cpml_terms(grid, shapes, spec, dx, dt, is_3d):
    """Compile the H and E term lists, ``compile.py:241-305``."""
    ndim = 3 if is_3d else 2
    axis_edges = AXIS_EDGES_3D if is_3d else AXIS_EDGES_2D
    offsets = OFFSETS_3D if is_3d else OFFSETS_TM
    alpha_norm = spec.alpha_norm if spec.alpha_norm is not None else (0.05 if is_3d else 0.1)
    sigma_max = cpml_sigma_max(spec.m, spec.reflection, max(spec.cells * dx, dx))
    alpha_max = 2.0 * EPS_0 * alpha_norm / max(dt, 1e-30)

    def prof(target, axis):
        kind = "E" if offsets[target][axis] == 0.0 else "H"
        low_e, high_e = axis_edges[axis]
        return cpml_profile_1d(
            shapes[target][axis], kind, grid[axis], spec.cells,
            low_e in spec.faces, high_e in spec.faces,
            sigma_max, spec.kappa_max, alpha_max, spec.m)

    if is_3d:
        hterms = [
            _make_term(t, s, ax, 1.0 if i % 2 == 0 else -1.0, shapes[t], prof(t, ax), dt)
            for i, (t, s, ax) in enumerate(H_TERMS_3D)]
        eterms = [
            _make_term(t, s, ax, 1.0 if i % 2 == 0 else -1.0, shapes[t], prof(t, ax), dt)
            for i, (t, s, ax) in enumerate(E_TERMS_3D)]
    else:
        hterms = [_make_term(t, s, ax, sg, shapes[t], prof(t, ax), dt)
                  for (t, s, ax, sg) in H_TERMS_TM]
        eterms = [_make_term(t, s, ax, sg, shapes[t], prof(t, ax), dt)
                  for (t, s, ax, sg) in E_TERMS_TM]
    del ndim
    return hterms, eterms


def pec_masks(shapes, metallic, is_3d):
    """Tangential-E / normal-H face masks, ``_boundary_compile.py:118-160``."""
    masks = {c: np.zeros(s, dtype=bool) for c, s in shapes.items()}
    if is_3d:
        walls = {
            "front": (0, 0, ("Ex", "Ey", "Hz")), "back": (0, -1, ("Ex", "Ey", "Hz")),
            "bottom": (1, 0, ("Ex", "Ez", "Hy")), "top": (1, -1, ("Ex", "Ez", "Hy")),
            "left": (2, 0, ("Ey", "Ez", "Hx")), "right": (2, -1, ("Ey", "Ez", "Hx")),
        }
    else:
        walls = {
            "bottom": (0, 0, ("Ez", "Hy")), "top": (0, -1, ("Ez", "Hy")),
            "left": (1, 0, ("Ez", "Hx")), "right": (1, -1, ("Ez", "Hx")),
        }
    for edge in metallic:
        axis, index, comps = walls[edge]
        for c in comps:
            sl = [slice(None)] * len(shapes[c])
            sl[axis] = index
            masks[c][tuple(sl)] = True
    return masks


# --------------------------------------------------------------------------
# Sources and monitors
# --------------------------------------------------------------------------
@dataclass
class Source:
    """A slab injection ``field[start:start+shape] += coeff * w[clip(step)]``
    (``execute.py:374-433``).  ``timing`` in {"pre_e", "h", "e"}."""

    component: str
    timing: str
    start: tuple
    coeff: np.ndarray
    waveform: np.ndarray


def gaussian_pulse(times, freq0, fwidth, amplitude=1.0, offset=4.0,
                   remove_dc=True, carrier="cos"):
    """``GaussianPulse.sample``, ``sources/time.py:459-486``.

    ``carrier="sin"`` gives the Gaussian-modulated sine used by the 2D config
    (``sin`` instead of ``cos`` in the carrier)."""
    t = np.asarray(times, dtype=np.float64)
    width = 1.0 / (2.0 * np.pi * max(fwidth, 1e-30))
    peak = offset / max(fwidth, 1e-30)
    env = amplitude * np.exp(-((t - peak) ** 2) / (2.0 * width ** 2))
    ph = 2.0 * np.pi * freq0 * t
    sig = env * (np.cos(ph) if carrier == "cos" else np.sin(ph))
    if remove_dc and sig.size:
        sig = sig - np.mean(sig)
    return sig


def sample_times(t0, dt, n, half_step=True):
    """Waveform sample times ``t0 + (i + 1/2) dt``, ``sources/compiler.py:455-460``
    with ``sources/time.py:67-81``."""
    return float(t0) + (np.arange(n, dtype=np.float64) + (0.5 if half_step else 0.0)) * dt


def linear_interpolation_plan(src, dst):
    """``lattice.py:218-249``: bracketing indices and weights."""
    src = np.asarray(src, dtype=np.float64).reshape(-1)
    dst = np.asarray(dst, dtype=np.float64).reshape(-1)
    if src.size == 1:
        z = np.zeros(dst.size, dtype=np.int64)
        return z, z, np.ones(dst.size), np.zeros(dst.size)
    hi = np.clip(np.searchsorted(src, dst, side="right"), 1, src.size - 1)
    lo = hi - 1
    lo[dst <= src[0]] = 0
    hi[dst <= src[0]] = 0
    lo[dst >= src[-1]] = src.size - 1
    hi[dst >= src[-1]] = src.size - 1
    den = src[hi] - src[lo]
    al = np.divide(dst - src[lo], den, out=np.zeros_like(dst),
                   where=(hi != lo) & (np.abs(den) > 0))
    wl = 1.0 - al
    wh = al.copy()
    wl[hi == lo] = 1.0
    wh[hi == lo] = 0.0
    return lo, hi, wl, wh


@dataclass
class Monitor:
    """A DFT monitor with one weighted-gather plan per component
    (``monitors/compiler.py:61-102``).  ``plans[c]`` is ``(index[P,K], weight[P,K])``
    into the flattened component array, or ``None`` when inactive."""

    freqs: np.ndarray
    plans: dict
    normal_axis: str
    normal_sign: float
    measure: float
    t_start: float = 0.0
    t_end: float = np.inf
    interval: int = 1
    hann: bool = False

    @property
    def points(self):
        return next(p[0].shape[0] for p in self.plans.values() if p is not None)


def plane_plan_3d(component, shape, grid, dx, normal_axis, plane_index, ranges):
    """Trilinear gather plan onto tangential cell centres of a snapped plane,
    ``lattice.py:252-333`` for a uniform grid.

    ``normal_axis`` in {"x","y","z"}; the plane sits at ``plane_index * dx``;
    ``ranges`` maps each tangential axis to a ``(start, stop)`` cell interval."""
    axes = {"x": ("z", "y"), "y": ("z", "x"), "z": ("y", "x")}[normal_axis]
    pos = {"z": 0, "y": 1, "x": 2}
    coords = {a: (np.arange(shape[pos[a]]) + OFFSETS_3D[component][pos[a]]) * dx
              for a in ("z", "y", "x")}
    t0 = (np.arange(*ranges[axes[0]]) + 0.5) * dx
    t1 = (np.arange(*ranges[axes[1]]) + 0.5) * dx
    g0, g1 = np.meshgrid(t0, t1, indexing="ij")
    plans = (
        linear_interpolation_plan(coords[axes[0]], g0.ravel()),
        linear_interpolation_plan(coords[axes[1]], g1.ravel()),
        linear_interpolation_plan(coords[normal_axis],
                                  np.full(g0.size, plane_index * dx)),
    )
    corners = [(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]
    idx = {}
    w = np.ones((g0.size, 8))
    for d, (lo, hi, wl, wh) in enumerate(plans):
        bits = np.array([c[d] for c in corners])[None, :]
        idx[d] = np.where(bits == 0, lo[:, None], hi[:, None])
        w = w * np.where(bits == 0, wl[:, None], wh[:, None])
    amap = {axes[0]: idx[0], axes[1]: idx[1], normal_axis: idx[2]}
    flat = np.ravel_multi_index((amap["z"], amap["y"], amap["x"]), shape)
    nz = np.abs(w) > 4.0 * np.finfo(np.float32).eps
    w = np.where(nz, w, 0.0)
    rs = w.sum(axis=1, keepdims=True)
    w = np.divide(w, rs, out=np.zeros_like(w), where=np.abs(rs) > 0)
    width = int(np.max(nz.sum(axis=1), initial=0))
    if 0 < width < 8:
        order = np.argsort(~nz, axis=1, kind="stable")[:, :width]
        flat = np.take_along_axis(flat, order, axis=1)
        w = np.take_along_axis(w, order, axis=1)
    return flat.astype(np.int64), w


def line_plan_tm(component, shape, dx, normal_axis, line_index, cell_range):
    """Bilinear gather plan for a 2D line monitor at cell centres of the
    tangential axis, ``monitors/compiler.py:126-197`` and
    ``monitors/monitors.py:176-205`` for a uniform grid."""
    oy, ox = OFFSETS_TM[component]
    ys = (np.arange(shape[0]) + oy) * dx
    xs = (np.arange(shape[1]) + ox) * dx
    tang = (np.arange(*cell_range) + 0.5) * dx
    if normal_axis == "x":
        dst_x, dst_y = np.full(tang.shape, line_index * dx), tang
    else:
        dst_x, dst_y = tang, np.full(tang.shape, line_index * dx)
    x0, x1, wx0, wx1 = linear_interpolation_plan(xs, dst_x)
    y0, y1, wy0, wy1 = linear_interpolation_plan(ys, dst_y)
    bits = np.array([(0, 0), (0, 1), (1, 0), (1, 1)])
    xi = np.where(bits[:, 1][None] == 0, x0[:, None], x1[:, None])
    yi = np.where(bits[:, 0][None] == 0, y0[:, None], y1[:, None])
    wx = np.where(bits[:, 1][None] == 0, wx0[:, None], wx1[:, None])
    wy = np.where(bits[:, 0][None] == 0, wy0[:, None], wy1[:, None])
    flat = np.ravel_multi_index((yi, xi), shape)
    return flat.astype(np.int64), wx * wy


# --------------------------------------------------------------------------
# Problem / state
# --------------------------------------------------------------------------
@dataclass
class Problem:
    """A fully lowered run: what ``compile_simulation`` hands to the step
    (``compile.py:540-793``), restricted to the hot path."""

    grid: tuple
    dx: float
    dt: float
    shapes: dict
    is_3d: bool
    metallic: frozenset
    e_decay: dict
    e_source: dict
    h_decay: dict
    h_source: dict
    hterms: list
    eterms: list
    masks: dict
    sources: list = field(default_factory=list)
    monitors: list = field(default_factory=list)
    t0: float = 0.0
    # derivative metrics of a non-uniform grid (compile.py:209-240): per
    # phase ("h": E -> H forward, "e": H -> E backward) one entry per array
    # axis, None (1 / dx), a scalar or a vector along that axis
    metrics: dict = None

    @property
    def cells(self):
        return int(np.prod(self.grid))


def build_problem(eps_cells, dx, dt, *, cpml=None, metallic=None, sigma_cells=None,
                  sources=(), monitors=(), t0=0.0, metrics=None):
    """Lower a permittivity raster into a ``Problem`` (3D or 2D TMxy).

    ``metallic`` defaults to the reference's boundary resolution
    (``boundaries.py:269-276``, ``_boundary_compile.py:106-115``): all six (four)
    faces PEC when no absorber is given, none when only a CPML is given.
    Material is extruded through the CPML slabs (``_boundary_compile.py:334-385``).
    """
    eps_cells = np.asarray(eps_cells, dtype=np.float64)
    is_3d = eps_cells.ndim == 3
    grid = eps_cells.shape
    axis_edges = AXIS_EDGES_3D if is_3d else AXIS_EDGES_2D
    all_faces = frozenset(e for pair in axis_edges for e in pair)
    if metallic is None:
        metallic = all_faces if cpml is None else frozenset()
    metallic = frozenset(metallic)
    sig = np.zeros(grid) if sigma_cells is None else np.asarray(sigma_cells, np.float64)
    if cpml is not None:
        eps_cells = extrude_into_cpml(eps_cells, cpml.cells, cpml.faces, axis_edges)
        sig = extrude_into_cpml(sig, cpml.cells, cpml.faces, axis_edges)
    shapes = shapes_3d(grid) if is_3d else shapes_tm(grid)
    e_decay, e_source, h_decay, h_source = {}, {}, {}, {}
    e_comps = ("Ex", "Ey", "Ez") if is_3d else ("Ez",)
    h_comps = ("Hx", "Hy", "Hz") if is_3d else ("Hx", "Hy")
    for c in e_comps:
        # kernels.py:712-728: decay=(1-b)/(1+b), source=dt/(eps0 eps)/(1+b), b=sigma dt/(2 eps0 eps)
        eps = e_site_permittivity_3d(eps_cells, c) if is_3d else site_sample_tm(eps_cells, c)
        s = e_site_permittivity_3d(sig, c) if is_3d else site_sample_tm(sig, c)
        beta = s * dt / (2.0 * EPS_0 * eps)
        e_decay[c] = (1.0 - beta) / (1.0 + beta)
        e_source[c] = (dt / (EPS_0 * eps)) / (1.0 + beta)
    for c in h_comps:
        # kernels.py:704-709 with sigma_m = 0 (non-magnetic, lossless H)
        h_decay[c] = np.float64(1.0)
        h_source[c] = np.float64(dt / MU_0)
    if cpml is not None:
        hterms, eterms = build_cpml_terms(grid, shapes, cpml, dx, dt, is_3d)
    else:
        hterms, eterms = [], []
    masks = pec_masks(shapes, metallic, is_3d)
    return Problem(grid, dx, dt, shapes, is_3d, metallic, e_decay, e_source,
                   h_decay, h_source, hterms, eterms, masks, list(sources),
                   list(monitors), t0, metrics)


def grid_metrics(widths, is_3d):
    """The derivative metrics of per-axis cell widths, ``compile.py:209-240``:
    forward 1 / w_i, backward 1 / w_0, 2 / (w_{i-1} + w_i), 1 / w_{n-1};
    ``widths`` maps array axis -> widths (a scalar for one uniform width)."""
    nax = 3 if is_3d else 2
    h, e = [None] * nax, [None] * nax
    for ax in range(nax):
        w = widths.get(ax)
        if w is None:
            continue
        w = np.asarray(w, dtype=np.float64)
        if w.ndim == 0:
            h[ax] = e[ax] = 1.0 / float(w)
            continue
        back = np.empty(w.size + 1)
        back[0], back[-1] = 1.0 / w[0], 1.0 / w[-1]
        back[1:-1] = 2.0 / (w[:-1] + w[1:])
        h[ax], e[ax] = 1.0 / w, back
    return {"h": h, "e": e}


def _scale_axis(d, prob, phase, ax):
    """``kernels.py:164-176``: a difference times its inverse distance (a
    scalar, or a vector broadcast along ``ax``); ``1 / dx`` without metrics."""
    m = None if prob.metrics is None else prob.metrics[phase][ax]
    if m is None:
        return d / prob.dx
    m = np.asarray(m, dtype=np.float64)
    return d * (m if m.ndim == 0 else _bcast(m, ax, d.ndim))


def init_state(prob):
    st = {c.lower(): np.zeros(s) for c, s in prob.shapes.items()}
    st["psi_h"] = [_psi_zeros(prob, t) for t in prob.hterms]
    st["psi_e"] = [_psi_zeros(prob, t) for t in prob.eterms]
    st["dft_re"] = [np.zeros((6, m.freqs.size, m.points)) for m in prob.monitors]
    st["dft_im"] = [np.zeros((6, m.freqs.size, m.points)) for m in prob.monitors]
    st["dft_w"] = [np.zeros(m.freqs.size) for m in prob.monitors]
    st["step"] = 0
    return st


def _psi_zeros(prob, term):
    shape = list(prob.shapes[term.target])
    shape[term.axis] = term.low + term.high
    return np.zeros(shape)


# --------------------------------------------------------------------------
# The step
# --------------------------------------------------------------------------
def _bcast(vec, axis, ndim):
    shape = [1] * ndim
    shape[axis] = vec.size
    return vec.reshape(shape)


def correct_term(d, psi, term):
    """Packed CPML recurrence, ``kernels.py:257-264`` with ``184-222``."""
    if term is None or (term.low == 0 and term.high == 0):
        return (1.0 if term is None else term.sign) * d, psi
    ax, lo, hi, n = term.axis, term.low, term.high, term.n
    slab = np.concatenate([np.take(d, range(lo), axis=ax),
                           np.take(d, range(n - hi, n), axis=ax)], axis=ax)
    nd = d.ndim
    psi = _bcast(term.b, ax, nd) * psi + _bcast(term.a, ax, nd) * slab
    corr = slab * _bcast(term.ik, ax, nd) + psi
    out = d.copy()
    sl_lo = [slice(None)] * nd
    sl_lo[ax] = slice(0, lo)
    sl_hi = [slice(None)] * nd
    sl_hi[ax] = slice(n - hi, n)
    src_lo = [slice(None)] * nd
    src_lo[ax] = slice(0, lo)
    src_hi = [slice(None)] * nd
    src_hi[ax] = slice(lo, lo + hi)
    out[tuple(sl_lo)] = corr[tuple(src_lo)]
    out[tuple(sl_hi)] = corr[tuple(src_hi)]
    return term.sign * out, psi


def ghost_diff(h, axis, metallic, axis_edges, dx):
    """Backward difference of ghost-padded H, ``lattice.py:523-575``: the ghost is
    zero on a metallic face and a copy of the edge sample otherwise (divided
    by ``dx``; ``dx=None`` returns the bare difference)."""
    low_e, high_e = axis_edges[axis]
    first = np.take(h, [0], axis=axis)
    last = np.take(h, [h.shape[axis] - 1], axis=axis)
    lo = np.zeros_like(first) if low_e in metallic else first
    hi = np.zeros_like(last) if high_e in metallic else last
    d = np.diff(np.concatenate([lo, h, hi], axis=axis), axis=axis)
    return d if dx is None else d / dx


def apply_sources(prob, st, timing):
    """``execute.py:436-461``: add every slab of this timing, waveform index
    ``clip(step, 0, T-1)``."""
    for s in prob.sources:
        if s.timing != timing:
            continue
        arr = st[s.component.lower()]
        amp = s.waveform[min(max(st["step"], 0), s.waveform.size - 1)]
        sl = tuple(slice(a, a + n) for a, n in zip(s.start, s.coeff.shape))
        arr[sl] += s.coeff * amp


def curl_e_to_h(prob, st):
    """Uncorrected curls on H sites (no CPML), ``lattice.py:467-473``."""
    out = {}
    table = H_TERMS_3D if prob.is_3d else [(t, s, a) for t, s, a, _ in H_TERMS_TM]
    signs = ([1.0, -1.0] * 3) if prob.is_3d else [1.0, -1.0]
    for i, (tgt, src, ax) in enumerate(table):
        d = signs[i] * np.diff(st[src.lower()], axis=ax) / prob.dx
        out[tgt] = out.get(tgt, 0.0) + d
    return out


def curl_h_to_e(prob, st):
    """Uncorrected ghost-aware curls on E sites, ``lattice.py:578-609``."""
    out = {}
    edges = AXIS_EDGES_3D if prob.is_3d else AXIS_EDGES_2D
    table = E_TERMS_3D if prob.is_3d else [(t, s, a) for t, s, a, _ in E_TERMS_TM]
    signs = ([1.0, -1.0] * 3) if prob.is_3d else [1.0, -1.0]
    for i, (tgt, src, ax) in enumerate(table):
        d = signs[i] * ghost_diff(st[src.lower()], ax, prob.metallic, edges, prob.dx)
        out[tgt] = out.get(tgt, 0.0) + d
    return out


def h_phase(prob, st):
    """``kernels.py:458-503`` (3D) / ``kernels.py:770-786, 1217-1246`` (TMxy);
    with metrics ``kernels.py:506-551`` (3D) / ``880-886, 916-927`` (TMxy)."""
    dx = prob.dx
    terms = prob.hterms or [None] * (6 if prob.is_3d else 2)
    table = H_TERMS_3D if prob.is_3d else [(t, s, a) for t, s, a, _ in H_TERMS_TM]
    signs = ([1.0, -1.0] * 3) if prob.is_3d else [1.0, -1.0]
    corrected = []
    for i, (tgt, src, ax) in enumerate(table):
        d = _scale_axis(np.diff(st[src.lower()], axis=ax), prob, "h", ax)
        if terms[i] is None:
            c, _ = signs[i] * d, None
        else:
            c, st["psi_h"][i] = correct_term(d, st["psi_h"][i], terms[i])
        corrected.append(c)
    comps = ("Hx", "Hy", "Hz") if prob.is_3d else ("Hx", "Hy")
    for j, comp in enumerate(comps):
        curl = corrected[2 * j] + corrected[2 * j + 1] if prob.is_3d else corrected[j]
        key = comp.lower()
        st[key] = prob.h_decay[comp] * st[key] - prob.h_source[comp] * curl


def e_phase(prob, st):
    """``kernels.py:554-626`` (3D) / ``kernels.py:789-811, 1249-1290`` (TMxy);
    with metrics ``kernels.py:629-700`` (3D) / ``888-914, 928-949`` (TMxy)."""
    dx = prob.dx
    edges = AXIS_EDGES_3D if prob.is_3d else AXIS_EDGES_2D
    terms = prob.eterms or [None] * (6 if prob.is_3d else 2)
    table = E_TERMS_3D if prob.is_3d else [(t, s, a) for t, s, a, _ in E_TERMS_TM]
    signs = ([1.0, -1.0] * 3) if prob.is_3d else [1.0, -1.0]
    corrected = []
    for i, (tgt, src, ax) in enumerate(table):
        d = _scale_axis(ghost_diff(st[src.lower()], ax, prob.metallic, edges, None), prob, "e", ax)
        if terms[i] is None:
            c = signs[i] * d
        else:
            c, st["psi_e"][i] = correct_term(d, st["psi_e"][i], terms[i])
        corrected.append(c)
    comps = ("Ex", "Ey", "Ez") if prob.is_3d else ("Ez",)
    for j, comp in enumerate(comps):
        curl = corrected[2 * j] + corrected[2 * j + 1]
        key = comp.lower()
        st[key] = prob.e_decay[comp] * st[key] + prob.e_source[comp] * curl


def apply_masks(prob, st, comps):
    for c in comps:
        m = prob.masks.get(c)
        if m is not None and m.any():
            st[c.lower()][m] = 0.0


def accumulate_monitors(prob, st):
    """``observe.py:529-585`` / ``649-795`` in float64: sample at
    ``t_phys = t0 + (step+1) dt`` with the weighted gather, rect or Hann window."""
    t_phys = prob.t0 + (st["step"] + 1) * prob.dt
    for mi, mon in enumerate(prob.monitors):
        if not (mon.t_start <= t_phys <= mon.t_end and st["step"] % max(mon.interval, 1) == 0):
            continue
        window = 1.0
        if mon.hann and np.isfinite(mon.t_end) and mon.t_end > mon.t_start:
            tau = np.clip((t_phys - mon.t_start) / max(mon.t_end - mon.t_start, 1e-30), 0, 1)
            window = 0.5 * (1.0 - np.cos(2.0 * np.pi * tau))
        theta = 2.0 * np.pi * mon.freqs * t_phys
        cr, ci = np.cos(theta), np.sin(theta)
        for ci_, comp in enumerate(COMPONENTS):
            plan = mon.plans.get(comp)
            if plan is None:
                continue
            idx, w = plan
            v = np.sum(st[comp.lower()].reshape(-1)[idx] * w, axis=1)
            st["dft_re"][mi][ci_] += window * np.outer(cr, v)
            st["dft_im"][mi][ci_] += window * np.outer(ci, v)
        st["dft_w"][mi] += window


def step(prob, st):
    """One leapfrog step in the order of ``execute.py:467-541``."""
    e_comps = ("Ex", "Ey", "Ez") if prob.is_3d else ("Ez",)
    h_comps = ("Hx", "Hy", "Hz") if prob.is_3d else ("Hx", "Hy")
    apply_sources(prob, st, "pre_e")
    h_phase(prob, st)
    apply_sources(prob, st, "h")
    apply_masks(prob, st, h_comps)
    e_phase(prob, st)
    apply_sources(prob, st, "e")
    apply_masks(prob, st, e_comps)
    accumulate_monitors(prob, st)
    st["step"] += 1
    return st


def run(prob, st, nsteps):
    for _ in range(int(nsteps)):
        step(prob, st)
    return st


# --------------------------------------------------------------------------
# Post-processing (host side in the reference too)
# --------------------------------------------------------------------------
def monitor_phasors(prob, st, mi):
    """``observe.py:166-187``: ``2 DFT / wsum`` with the H half-step phase."""
    mon = prob.monitors[mi]
    w = np.maximum(st["dft_w"][mi], 1e-18)[:, None]
    out = {}
    for ci_, comp in enumerate(COMPONENTS):
        if mon.plans.get(comp) is None:
            out[comp] = None
            continue
        v = (2.0 / w) * (st["dft_re"][mi][ci_] + 1j * st["dft_im"][mi][ci_])
        if comp.startswith("H"):
            v = v * np.exp(-1j * np.pi * mon.freqs * prob.dt)[:, None]
        out[comp] = v
    return out


def monitor_flux(prob, st, mi):
    """``observe.py:198-234``: ``1/2 Re sum sign (E x H*)_n * measure``."""
    mon = prob.monitors[mi]
    ph = monitor_phasors(prob, st, mi)
    zero = np.zeros((mon.freqs.size, mon.points), dtype=complex)

    def f(c):
        return ph[c] if ph.get(c) is not None else zero

    if mon.normal_axis == "x":
        comp = f("Ey") * np.conj(f("Hz")) - f("Ez") * np.conj(f("Hy"))
    elif mon.normal_axis == "y":
        comp = f("Ez") * np.conj(f("Hx")) - f("Ex") * np.conj(f("Hz"))
    else:
        comp = f("Ex") * np.conj(f("Hy")) - f("Ey") * np.conj(f("Hx"))
    return 0.5 * np.real(np.sum(mon.normal_sign * comp, axis=1)) * mon.measure
