"""Boundary specifications and their host-side lowering to PEC faces and packed
CPML recurrence terms.  Profiles are evaluated in float64 and cast to float32
only when uploaded.

References: ``beamz/devices/boundaries.py:43-276`` (PEC/PML specs and the
all-PEC default), ``beamz/devices/_boundary_compile.py:44-60`` (term order),
``:106-115`` (edge precedence), ``:244-262`` (automatic sigma_max/alpha_max),
``:334-385`` (material extrusion), ``:425-614`` (staggered profiles);
``beamz/simulation/kernels.py:130-144, 225-254`` (coefficients, slab extents);
``beamz/simulation/compile.py:241-305`` (term plan and signs).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from .constants import EPS_0, ETA_0
from .lattice import EDGES_2D, EDGES_3D, OFFSETS_3D, OFFSETS_TM

_ALL_3D = ("front", "back", "bottom", "top", "left", "right")
_ALL_2D = ("bottom", "top", "left", "right")


def _edges(edges, is_3d):
    if edges == "all":
        return frozenset(_ALL_3D if is_3d else _ALL_2D)
    out = frozenset([edges] if isinstance(edges, str) else edges)
    bad = out - frozenset(_ALL_3D if is_3d else _ALL_2D)
    if bad:
        raise ValueError(f"unsupported boundary edges {sorted(bad)}")
    return out


@dataclass(frozen=True)
class PEC:
    """Perfect electric conductor on the selected faces (``boundaries.py:43-93``)."""

    edges: object = "all"


@dataclass(frozen=True)
class PML:
    """Graded CPML absorber (``boundaries.py:96-186``).

    ``thickness`` is in metres; when omitted the layer is ``cells`` cells thick
    (default 12, ``PML.DEFAULT_CELLS``).  Only the convolutional formulation is
    provided by this backend."""

    edges: object = "all"
    thickness: Optional[float] = None
    sigma_max: Optional[float] = None
    m: int = 3
    formulation: str = "cpml"
    kappa_max: float = 2.0
    alpha_max: Optional[float] = None
    target_reflection: float = 1e-6
    cells: Optional[int] = None

    DEFAULT_CELLS = 12
    DEFAULT_ALPHA_NORMALIZED_2D = 0.1
    DEFAULT_ALPHA_NORMALIZED_3D = 0.05

    def __post_init__(self):
        if str(self.formulation).lower() != "cpml":
            raise NotImplementedError("the HIP backend implements the CPML formulation only")


def resolve_metallic_edges(boundaries, is_3d):
    """PEC precedence (``_boundary_compile.py:106-115``) with the all-PEC default
    for an empty boundary list (``boundaries.py:269-276``)."""
    boundaries = tuple(boundaries) if boundaries else (PEC(),)
    metallic = set()
    for b in boundaries:
        faces = _edges(b.edges, is_3d)
        if isinstance(b, PEC):
            metallic |= faces
        else:
            metallic -= faces
    return frozenset(metallic)


def edge_mask(edges):
    from .lattice import EDGE_BITS
    return sum(EDGE_BITS[e] for e in edges)


def staggered_profile(samples, kind, domain_cells, pml_cells, low, high,
                      sigma_max, kappa_max, alpha_max, m, alpha_order=1.0):
    """sigma/kappa/alpha on E samples (i) or H samples (i + 1/2),
    ``_boundary_compile.py:425-614`` on a uniform grid, in float64."""
    coords = np.arange(samples, dtype=np.float64) + (0.0 if kind == "E" else 0.5)
    sigma = np.zeros(samples)
    kappa = np.ones(samples)
    alpha = np.zeros(samples)
    scale = float(max(pml_cells, 1))
    sides = []
    if low:
        sides.append(np.clip(float(pml_cells) - coords, 0.0, float(pml_cells)))
    if high:
        sides.append(np.clip(coords - float(domain_cells - pml_cells), 0.0, float(pml_cells)))
    for dist in sides:
        u = np.clip(dist / scale, 0.0, 1.0)
        on = dist > 0.0
        sigma = np.where(on, np.maximum(sigma, sigma_max * u ** m), sigma)
        kappa = np.where(on, np.maximum(kappa, 1.0 + (kappa_max - 1.0) * u ** m), kappa)
        alpha = np.where(on, np.maximum(alpha, alpha_max * (1.0 - u) ** alpha_order), alpha)
    return sigma, kappa, alpha


def recurrence_coefficients(sigma, kappa, alpha, dt):
    """``kernels.py:130-144``: ``b = exp(-(sigma/kappa + alpha) dt/eps0)``,
    ``a = (b - 1) sigma / ((sigma + kappa alpha) kappa)``, ``1/kappa``."""
    kappa = np.maximum(kappa, 1.0)
    b = np.exp(-(sigma / kappa + alpha) * dt / EPS_0)
    a = np.nan_to_num(((b - 1.0) * sigma) / np.maximum((sigma + kappa * alpha) * kappa, 1e-30))
    return a, b, 1.0 / kappa


def slab_extents(a, inv_kappa):
    """Active leading/trailing runs of a 1-D profile (``kernels.py:225-234``)."""
    active = (np.abs(a) > 1e-30) | (np.abs(inv_kappa - 1.0) > 1e-7)
    low = int(np.argmin(active)) if not active.all() else active.size
    rest = active[low:][::-1]
    high = int(np.argmin(rest)) if rest.size and not rest.all() else rest.size
    return low, high


@dataclass
class CpmlTerm:
    """One packed derivative (``model.py:306-324``); ``a, b, inv_kappa`` are
    float64 vectors of length ``low + high`` along ``axis`` (storage axis of the
    target array)."""

    component: str
    source: str
    axis: int
    sign: float
    low: int
    high: int
    extent: int
    a: np.ndarray
    b: np.ndarray
    inv_kappa: np.ndarray


# _boundary_compile.py:44-60 and compile.py:287-295: (target, source, axis)
H_TERMS_3D = (("Hx", "Ez", 1), ("Hx", "Ey", 0), ("Hy", "Ex", 0),
              ("Hy", "Ez", 2), ("Hz", "Ey", 2), ("Hz", "Ex", 1))
E_TERMS_3D = (("Ex", "Hz", 1), ("Ex", "Hy", 0), ("Ey", "Hx", 0),
              ("Ey", "Hz", 2), ("Ez", "Hy", 2), ("Ez", "Hx", 1))
H_TERMS_TM = (("Hx", "Ez", 0, 1.0), ("Hy", "Ez", 1, -1.0))
E_TERMS_TM = (("Ez", "Hy", 1, 1.0), ("Ez", "Hx", 0, -1.0))


@dataclass
class CpmlPlan:
    enabled: bool
    cells: int
    faces: frozenset
    h_terms: tuple
    e_terms: tuple


def lower_cpml(pml: PML, grid, shapes, resolution, dt):
    """Compile the H and E term lists for one CPML spec."""
    is_3d = len(grid) == 3
    faces = _edges(pml.edges, is_3d)
    if pml.cells is not None:
        cells = int(pml.cells)
    elif pml.thickness is not None:
        cells = max(int(round(pml.thickness / resolution)), 1)
    else:
        cells = PML.DEFAULT_CELLS
    thickness = pml.thickness if pml.thickness is not None else cells * resolution
    sigma_max = pml.sigma_max
    if sigma_max is None:
        sigma_max = (-(pml.m + 1) * np.log(max(pml.target_reflection, 1e-16))
                     / (2.0 * ETA_0 * max(thickness, resolution)))
    alpha_max = pml.alpha_max
    if alpha_max is None:
        norm = PML.DEFAULT_ALPHA_NORMALIZED_3D if is_3d else PML.DEFAULT_ALPHA_NORMALIZED_2D
        alpha_max = 2.0 * EPS_0 * norm / max(dt, 1e-30)
    edges = EDGES_3D if is_3d else EDGES_2D
    offsets = OFFSETS_3D if is_3d else OFFSETS_TM

    def term(target, source, axis, sign):
        kind = "E" if offsets[target][axis] == 0.0 else "H"
        low_e, high_e = edges[axis]
        n = shapes[target][axis]
        prof = staggered_profile(n, kind, grid[axis], cells, low_e in faces,
                                 high_e in faces, sigma_max, pml.kappa_max, alpha_max, pml.m)
        a, b, ik = recurrence_coefficients(*prof, dt)
        lo, hi = slab_extents(a, ik)

        def pack(v):
            return np.concatenate([v[:lo], v[n - hi:]])

        return CpmlTerm(target, source, axis, sign, lo, hi, n, pack(a), pack(b), pack(ik))

    if is_3d:
        h = tuple(term(t, s, ax, 1.0 if i % 2 == 0 else -1.0) for i, (t, s, ax) in enumerate(H_TERMS_3D))
        e = tuple(term(t, s, ax, 1.0 if i % 2 == 0 else -1.0) for i, (t, s, ax) in enumerate(E_TERMS_3D))
    else:
        h = tuple(term(t, s, ax, sg) for t, s, ax, sg in H_TERMS_TM)
        e = tuple(term(t, s, ax, sg) for t, s, ax, sg in E_TERMS_TM)
    return CpmlPlan(True, cells, faces, h, e)


def extrude_through_cpml(cell, cells, faces, is_3d):
    """Copy the first interior slice through each active slab
    (``_boundary_compile.py:334-385``)."""
    out = np.array(cell, dtype=np.float64, copy=True)
    for axis, (low_e, high_e) in enumerate(EDGES_3D if is_3d else EDGES_2D):
        n = out.shape[axis]
        if cells <= 0 or cells >= n:
            continue
        if low_e in faces:
            sl = [slice(None)] * out.ndim
            sl[axis] = slice(0, cells)
            out[tuple(sl)] = np.take(out, [cells], axis=axis)
        if high_e in faces:
            sl = [slice(None)] * out.ndim
            sl[axis] = slice(n - cells, n)
            out[tuple(sl)] = np.take(out, [n - cells - 1], axis=axis)
    return out
