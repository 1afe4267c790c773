"""Host-side lowering of a simulation into an immutable device program and the
initial state.

References: ``beamz/simulation/compile.py:540-793`` (``compile_simulation``),
``:126-143`` (uniform-coefficient elision), ``:563-733`` (coefficients),
``beamz/simulation/kernels.py:704-728`` (update coefficients),
``beamz/simulation/model.py:127-324`` (state, coefficients, CPML terms),
``beamz/simulation/execute.py:814-905`` (initial state).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import NamedTuple

import numpy as np
import torch

from . import _abi as abi
from .boundaries import (PML, edge_mask, extrude_through_cpml, lower_cpml,
                         resolve_metallic_edges)
from .constants import EPS_0, ETA_0, MU_0
from .lattice import component_shapes, sample_e_site_3d, sample_site_tm
from .monitors import assign_offsets, capture_boxes, lower_monitor, pack_monitors
from .modes import ModeSource, lower_mode_source_2d
from .sources import (CustomSource, GaussianSource, SoftSource, group_sources, lower_gaussian_source,
                      lower_soft_source)

_COMPONENT_CODE = {"Hx": 0, "Hy": 1, "Hz": 2, "Ex": 0, "Ey": 1, "Ez": 2}


class SimulationState(NamedTuple):
    """The step carry (``model.py:127-257``) as device tensors, updated in place
    by the HIP library.  2D TMxy states hold empty ``hz``, ``ex``, ``ey``."""

    hx: torch.Tensor
    hy: torch.Tensor
    hz: torch.Tensor
    ex: torch.Tensor
    ey: torch.Tensor
    ez: torch.Tensor
    psi_h: tuple
    psi_e: tuple
    dft_re: torch.Tensor
    dft_im: torch.Tensor
    dft_w: torch.Tensor
    t: torch.Tensor
    current_step: torch.Tensor
    fields_b: tuple = ()   # second field bank of the two-bank schedule
    psi_b: tuple = ()      # second CPML bank: H terms then E terms

    def fields(self):
        return (self.hx, self.hy, self.hz, self.ex, self.ey, self.ez)


@dataclass
class CompiledProgram:
    grid: tuple
    is_3d: bool
    resolution: float
    dt: float
    t0: float
    num_steps: int
    shapes: dict
    metallic: frozenset
    boundary_code: int
    cpml: object
    coeffs: dict
    h_meta: torch.Tensor
    e_meta: torch.Tensor
    h_term_arrays: tuple
    e_term_arrays: tuple
    source_groups: list
    coincident_mask: int
    monitors: list
    packed_monitors: tuple
    device: torch.device
    field_dtype: torch.dtype
    flags: int
    psi_dtype: torch.dtype
    abi_resolution: float = None   # cell size handed to the kernels (1 when folded)
    cpml_table: torch.Tensor = None  # [terms, 3, st]: each term's a, b, 1/kappa, st its longest slab
    field_scale: dict = None       # storage scale of the E and H fields
    eps_cells: np.ndarray = field(repr=False, default=None)
    eps_site: dict = field(repr=False, default=None)
    # derivative metrics of a non-uniform grid (compile.py:209-240): kind
    # code 0 isotropic, 1 axis-uniform, 2 rectilinear; per phase the (z, y, x)
    # inverse distances (empty tensors on isotropic grids and absent axes)
    metric_kind: int = 0
    h_metrics: tuple = ()
    e_metrics: tuple = ()
    cell_grid: object = None       # the RectilinearGrid, or None for a uniform grid
    # host s32[M, 6, 6] bounding boxes of the cells the monitors gather (3D),
    # handed to the program as its optional trailing argument
    monitor_boxes: torch.Tensor = None

    @property
    def cells(self) -> int:
        """Material cells: the GCUPS numerator (``execute.py:1111-1118``)."""
        return int(np.prod(self.grid))

    @property
    def cpml_enabled(self) -> bool:
        return self.cpml is not None

    @property
    def nterms(self) -> int:
        return 0 if self.cpml is None else len(self.cpml.h_terms)


HALF_DTYPES = (torch.bfloat16, torch.float16)


def storage_scales(customs, field_dtype, field_scale=None):
    """Storage scales ``(s_e, s_h)`` of the E and H fields.  float32 fields are
    stored unscaled.  Half-precision fields are stored as E / s_e and H / s_h
    with s_e the largest per-step source increment (|coefficient| x |waveform|,
    H sources weighted by eta0) -- the order of the peak field a soft source
    builds -- and s_h = s_e / eta0, so both stored fields are O(1) and far from
    the float16 range limits."""
    if field_dtype not in HALF_DTYPES:
        return 1.0, 1.0
    if field_scale is not None:
        s_e = float(field_scale)
        if not (s_e > 0.0 and np.isfinite(s_e)):
            raise ValueError("field_scale must be a positive finite number")
        return s_e, s_e / ETA_0
    peak = 0.0
    for c in customs:
        amp = float(np.max(np.abs(c.coeff), initial=0.0)) * float(np.max(np.abs(c.waveform), initial=0.0))
        peak = max(peak, amp * (ETA_0 if c.component[0] == "H" else 1.0))
    s_e = peak if peak > 0.0 and np.isfinite(peak) else 1.0
    return s_e, s_e / ETA_0


def _dev(x, device, dtype=torch.float32):
    return torch.as_tensor(np.asarray(x), dtype=dtype).to(device).contiguous()


def _elide(arr, device):
    """A uniform coefficient becomes a rank-0 tensor (``compile.py:137-143``)."""
    a = np.asarray(arr, dtype=np.float64)
    a32 = a.astype(np.float32)
    if a32.size and np.all(a32 == a32.flat[0]):
        return torch.tensor(float(a32.flat[0]), dtype=torch.float32, device=device)
    return _dev(a32, device)


# H-timing slabs the fused kernels apply in-kernel per step (kMaxHSlabs of
# the library); programs with more run the per-phase schedule
MAX_H_SLABS = 8


def h_source_slabs(groups):
    """Slab count of the three H-timing source groups."""
    return sum(int(g[0].shape[0]) for g in groups[3:6] if g is not None)


def psi_precision_from_env():
    """CPML memory storage chosen by ``BEAMZ_HIP_CPML_PSI_PRECISION`` (``fp32``
    or ``bf16``), the counterpart of the reference's
    ``BEAMZ_CUDA_CPML_PSI_PRECISION`` (``simulation/backend.py:45-56``)."""
    import os
    value = os.environ.get("BEAMZ_HIP_CPML_PSI_PRECISION", "fp32").strip().lower()
    if value in ("bf16", "bfloat16"):
        return torch.bfloat16
    if value in ("fp32", "float32", ""):
        return torch.float32
    raise ValueError("BEAMZ_HIP_CPML_PSI_PRECISION must be 'fp32' or 'bf16'")


def compile_program(permittivity, resolution, dt, *, boundaries=(), sources=(), monitors=(),
                    num_steps=1, t0=0.0, conductivity=None, device=None,
                    field_dtype=torch.float32, flags=abi.DEFAULT_FLAGS, field_scale=None,
                    psi_dtype=None, grid=None):
    """Lower ``(permittivity raster, dx, dt, boundaries, sources, monitors)``.

    ``psi_dtype`` stores the CPML memories as float32 (default) or bfloat16
    (the reference's BF16 psi option, ``backend.py:45-56``; the recurrence
    itself runs in float32); ``None`` reads ``BEAMZ_HIP_CPML_PSI_PRECISION``.

    ``grid`` (a ``RectilinearGrid``) gives non-uniform cell widths: the curls
    then scale each difference by its inverse physical distance (the
    derivative metrics), ``resolution`` is the grid's smallest width, CPML
    profiles are graded in cell units and monitors weigh each sample by its
    cell's area (or, on 2D lines, width)."""
    device = torch.device(device or ("cuda" if torch.cuda.is_available() else "cpu"))
    psi_dtype = psi_precision_from_env() if psi_dtype is None else psi_dtype
    if psi_dtype not in (torch.float32, torch.bfloat16):
        raise ValueError("CPML memories are stored as float32 or bfloat16")
    if psi_dtype == torch.bfloat16:
        flags = int(flags) | abi.FLAG_BF16_PSI
    eps = np.asarray(permittivity, dtype=np.float64)
    if eps.ndim not in (2, 3):
        raise ValueError("permittivity must be a 2D (y, x) or 3D (z, y, x) raster")
    is_3d = eps.ndim == 3
    cell_grid = grid
    grid = tuple(int(v) for v in eps.shape)
    shapes = component_shapes(grid)
    metric_kind, h_metrics, e_metrics = 0, (), ()
    if cell_grid is not None:
        metric_kind, h_metrics, e_metrics = _lower_metrics(cell_grid, grid, is_3d, device)
        if metric_kind and field_dtype in HALF_DTYPES:
            raise NotImplementedError("half-precision fields run on uniform grids only")
        if metric_kind:
            for s in sources:
                if not isinstance(s, CustomSource):
                    raise NotImplementedError(
                        f"{type(s).__name__} is placed on uniform grids only; use a CustomSource "
                        "on a non-uniform grid")
            for m in monitors:
                if m.placed:
                    raise NotImplementedError(
                        "monitors placed by centre and size need a uniform grid; give a "
                        "non-uniform grid's monitors by index")
            for b in boundaries:
                if isinstance(b, PML) and b.thickness is not None:
                    raise NotImplementedError(
                        "a PML on a non-uniform grid is given in cells (PML(cells=...))")
    sigma = np.zeros(grid) if conductivity is None else np.broadcast_to(
        np.asarray(conductivity, dtype=np.float64), grid).copy()
    pmls = [b for b in boundaries if isinstance(b, PML)]
    if len(pmls) > 1:
        raise NotImplementedError("one PML specification per simulation")
    cpml = lower_cpml(pmls[0], grid, shapes, resolution, dt) if pmls else None
    if cpml is not None:
        eps = extrude_through_cpml(eps, cpml.cells, cpml.faces, is_3d)
        sigma = extrude_through_cpml(sigma, cpml.cells, cpml.faces, is_3d)
    metallic = resolve_metallic_edges(boundaries, is_3d)

    # update coefficients (kernels.py:704-728), evaluated in float64
    coeffs = {}
    eps_site = {}
    sigma_site = {}
    e_source_raw = {}
    e_comps = ("Ex", "Ey", "Ez") if is_3d else ("Ez",)
    for ax in "xyz":
        comp = "E" + ax
        if comp in e_comps:
            es = sample_e_site_3d(eps, comp) if is_3d else sample_site_tm(eps, comp)
            ss = sample_e_site_3d(sigma, comp) if is_3d else sample_site_tm(sigma, comp)
            beta = ss * dt / (2.0 * EPS_0 * es)
            eps_site[comp] = es
            sigma_site[comp] = ss
            coeffs[f"e_decay_{ax}"] = _elide((1.0 - beta) / (1.0 + beta), device)
            e_source_raw[ax] = (dt / (EPS_0 * es)) / (1.0 + beta)
        else:
            coeffs[f"e_decay_{ax}"] = torch.tensor(1.0, device=device)
            e_source_raw[ax] = np.zeros(1)
        coeffs[f"h_decay_{ax}"] = torch.tensor(1.0, device=device)

    # CPML terms (compile.py:241-305) -> metadata and packed coefficient arrays
    thickness = 0
    if cpml is not None:
        allterms = cpml.h_terms + cpml.e_terms
        first = allterms[0].low
        if first > 0 and all(t.low == first and t.high == first for t in allterms):
            thickness = first
    table = None
    if cpml is not None:
        # one contiguous [H terms, E terms] x (a, b, 1/kappa) x st table, st the
        # longest packed slab (2L for one thickness on every face); each term
        # array is the leading low + high entries of its row (the fused kernel
        # reads the table)
        allterms = cpml.h_terms + cpml.e_terms
        st = max(max(t.low + t.high for t in allterms), 1)
        tab = np.zeros((len(allterms), 3, st))
        for i, t in enumerate(allterms):
            tab[i, :, :t.low + t.high] = np.stack([t.a, t.b, t.inv_kappa])
        table = _dev(tab.astype(np.float32), device)

    def term_payload(terms, first):
        if not terms:
            return torch.zeros((0, 5), dtype=torch.int32, device=device), ()
        meta = np.asarray([(_COMPONENT_CODE[t.component], t.axis, t.low, t.high,
                            1 if t.sign > 0 else -1) for t in terms], dtype=np.int32)
        arrays = []
        for i, t in enumerate(terms):
            n = t.low + t.high
            arrays += [table[first + i, 0, :n], table[first + i, 1, :n], table[first + i, 2, :n]]
        while len(arrays) < 18:
            arrays.append(torch.zeros(0, device=device))
        return _dev(meta, device, torch.int32), tuple(arrays)

    h_meta, h_arrays = term_payload(cpml.h_terms if cpml else (), 0)
    e_meta, e_arrays = term_payload(cpml.e_terms if cpml else (), len(cpml.h_terms) if cpml else 0)
    boundary_code = edge_mask(metallic) | (thickness << 8)

    # sources -> nine (timing, component) groups
    customs = []
    for s in sources:
        if isinstance(s, SoftSource):
            customs.append(lower_soft_source(s, eps_site[s.component], dt, t0, num_steps))
        elif isinstance(s, GaussianSource):
            customs.append(lower_gaussian_source(s, eps_site["Ez"], resolution, dt, t0, num_steps))
        elif isinstance(s, CustomSource):
            customs.append(s)
        elif isinstance(s, ModeSource):
            if is_3d and s.launch == "huygens":
                from .modes3d import lower_mode_source_3d
                customs.extend(lower_mode_source_3d(s, eps_site, tuple(eps.shape), resolution, dt,
                                                    t0, num_steps))
            elif is_3d:
                from .tfsf3d import lower_tfsf_mode_source_3d
                customs.extend(lower_tfsf_mode_source_3d(s, eps_site, tuple(eps.shape), resolution,
                                                         dt, t0, num_steps, sigma_site, metallic))
            elif s.launch == "tfsf":
                raise NotImplementedError("2D mode sources lower the line pair (launch='auto')")
            else:
                customs.extend(lower_mode_source_2d(s, eps_site["Ez"], resolution, dt, t0, num_steps))
        else:
            raise TypeError(f"unsupported source {type(s).__name__}")

    # storage scaling of half-precision fields (no counterpart in the
    # reference, whose fields are float32): stored E = E / s_e, stored
    # H = H / s_h.  The scales and 1/dx are folded into the update coefficients
    # (the kernels then run with a unit inverse cell size), the source
    # coefficients are divided by the target's scale and the monitor component
    # masks multiplied by it, so every kernel is unchanged and every monitor
    # output is in physical units.
    s_e, s_h = storage_scales(customs, field_dtype, field_scale)
    half = field_dtype in HALF_DTYPES
    fold = 1.0 / float(resolution) if half else 1.0
    if half:
        customs = [CustomSource(c.component, c.start,
                                np.asarray(c.coeff, dtype=np.float64) / (s_e if c.component[0] == "E" else s_h),
                                c.waveform, c.timing) for c in customs]
    h_source = dt / MU_0 * (s_e / s_h) * fold
    for ax in "xyz":
        coeffs[f"h_source_{ax}"] = torch.tensor(float(np.float32(h_source)), device=device)
    groups = []
    coincident = 0
    for gi, g in enumerate(group_sources(customs, is_3d)):
        if g is None:
            groups.append(None)
            continue
        c, w, st, co = g
        groups.append((_dev(c, device), _dev(w, device), _dev(st, device, torch.int32)))
        coincident |= (1 << gi) if co else 0

    # E source coefficients: float32, or float16 when half-precision fields
    # run the fused step (the folded values c dt / (eps dx) lie in [0.02, 1])
    fused = bool(is_3d and (flags & abi.FLAG_FUSED_STEP) and h_source_slabs(groups) <= MAX_H_SLABS)
    coef_dtype = torch.float16 if (half and fused) else torch.float32
    for ax in "xyz":
        arr = np.asarray(e_source_raw[ax], dtype=np.float64) * (s_h / s_e) * fold
        t = _elide(arr, device)
        coeffs[f"e_source_{ax}"] = t if t.dim() == 0 else t.to(coef_dtype)
    src = [coeffs[f"e_source_{ax}"] for ax in "xyz"]
    if all(t.dim() > 0 for t in src):
        # one allocation, so that the fused step reaches all three from one base
        for ax, view, t in zip("xyz", field_bank([tuple(t.shape) for t in src], coef_dtype, device),
                               src):
            view.copy_(t)
            coeffs[f"e_source_{ax}"] = view

    compiled = [lower_monitor(m, grid, shapes, resolution, i) for i, m in enumerate(monitors)]
    if metric_kind:
        compiled = [_with_cell_areas(cm, cell_grid, is_3d) for cm in compiled]
    assign_offsets(compiled)
    packed = ()
    boxes = None
    if compiled:
        idx, wts, fr, masks, counts, codes, windows = pack_monitors(compiled)
        masks[:, :3] *= s_e    # DFT component order Ex, Ey, Ez, Hx, Hy, Hz
        masks[:, 3:] *= s_h
        packed = (_dev(idx, device, torch.int32), _dev(wts, device), _dev(fr, device, torch.float64),
                  _dev(masks, device), _dev(counts, device, torch.int32),
                  _dev(codes, device, torch.int32), _dev(windows, device))
        if is_3d:
            boxes = torch.from_numpy(capture_boxes(idx, wts, masks, shapes))

    return CompiledProgram(
        grid=grid, is_3d=is_3d, resolution=float(resolution), dt=float(dt), t0=float(t0),
        abi_resolution=1.0 if half else float(resolution), field_scale={"E": s_e, "H": s_h},
        cpml_table=table,
        num_steps=int(num_steps), shapes=shapes, metallic=metallic,
        boundary_code=boundary_code, cpml=cpml, coeffs=coeffs, h_meta=h_meta, e_meta=e_meta,
        h_term_arrays=h_arrays, e_term_arrays=e_arrays, source_groups=groups,
        coincident_mask=coincident, monitors=compiled, packed_monitors=packed,
        monitor_boxes=boxes,
        device=device, field_dtype=field_dtype, psi_dtype=psi_dtype, flags=int(flags), eps_cells=eps,
        eps_site=eps_site, metric_kind=metric_kind, h_metrics=h_metrics, e_metrics=e_metrics,
        cell_grid=cell_grid if metric_kind else None)


def _lower_metrics(cell_grid, grid, is_3d, device):
    """The derivative metrics of ``cell_grid`` as device tensors in the
    kernels' (z, y, x) order per phase: H reads the forward (E -> H) inverse
    distances, E the backward (H -> E) ones (``cuda/runtime.py:89-94``)."""
    from .grid import METRIC_KIND_CODE, derivative_metrics
    axes = ("x", "y", "z") if is_3d else ("x", "y")
    want = tuple(grid) if is_3d else (1,) + tuple(grid)
    have = cell_grid.shape_zyx if is_3d else (1,) + cell_grid.shape_zyx[1:]
    if tuple(have) != want:
        raise ValueError(f"grid cells {cell_grid.shape_zyx} (z, y, x) do not match the "
                         f"permittivity raster {tuple(grid)}")
    kind, fwd, bwd = derivative_metrics(cell_grid, axes)
    code = METRIC_KIND_CODE[kind]
    if code == 0:
        return 0, (), ()
    empty = torch.zeros(0, dtype=torch.float32, device=device)

    def per_phase(table):
        return tuple(_dev(np.asarray(table[a], np.float32), device) if a in table else empty
                     for a in ("z", "y", "x"))
    return code, per_phase(fwd), per_phase(bwd)


def _with_cell_areas(cm, cell_grid, is_3d):
    """An index-space monitor on a non-uniform grid: the gather plans sample
    each component at the physical cell centres of the plane (line) at its
    E-node coordinate, and each sample's measure is the area (3D planes) or
    width (2D lines) of the cell it samples, in place of dx^2 or dx."""
    import dataclasses
    from .lattice import component_coords_3d, line_gather_plan, plane_gather_plan_at
    plans = {}
    if is_3d:
        tangential = {"x": ("z", "y"), "y": ("z", "x"), "z": ("y", "x")}[cm.normal]
        ranges = dict(cm.ranges)
        u = cell_grid.centers(tangential[0])[slice(*ranges[tangential[0]])]
        v = cell_grid.centers(tangential[1])[slice(*ranges[tangential[1]])]
        at = cell_grid.axis_edges(cm.normal)[cm.index]
        for c, plan in cm.plans.items():
            if plan is None:
                plans[c] = None
                continue
            shape = _component_shape(c, cell_grid)
            plans[c] = plane_gather_plan_at(c, shape, None, cm.normal, at, u, v,
                                            coords=component_coords_3d(c, shape, None, cell_grid))
        area = (cell_grid.cell_widths(tangential[0])[slice(*ranges[tangential[0]])][:, None] *
                cell_grid.cell_widths(tangential[1])[slice(*ranges[tangential[1]])][None, :]).reshape(-1)
    else:
        tang = "y" if cm.normal == "x" else "x"
        for c, plan in cm.plans.items():
            if plan is None:
                plans[c] = None
                continue
            shape = _component_shape(c, cell_grid, two_d=True)
            plans[c] = line_gather_plan(c, shape, None, cm.normal, cm.index, cm.cell_range,
                                        cell_grid=cell_grid)
        area = cell_grid.cell_widths(tang)[cm.cell_range[0]:cm.cell_range[1]]
    if area.size != cm.points:
        raise ValueError("monitor samples do not match its cell range")
    return dataclasses.replace(cm, plans=plans, measure=1.0,
                               point_measure=np.asarray(area, np.float64))


def _component_shape(component, cell_grid, two_d=False):
    if two_d:
        _, ny, nx = cell_grid.shape_zyx
        return component_shapes((ny, nx))[component]
    return component_shapes(cell_grid.shape_zyx)[component]


BANK_ALIGN = 256  # bytes between the starts of the fields of one bank


def field_bank(shapes, dtype, device):
    """Zeroed fields of one bank, all views of one allocation, each starting on
    a BANK_ALIGN-byte boundary.  The fused step addresses every field of a
    bank from one base with 32-bit offsets, and two banks built from the same
    shapes have the same offsets."""
    esize = torch.empty(0, dtype=dtype).element_size()
    step = BANK_ALIGN // esize
    offs, total = [], 0
    for sh in shapes:
        offs.append(total)
        n = int(np.prod(sh)) if len(sh) else 0
        total += -(-n // step) * step
    arena = torch.zeros(max(total, 1), dtype=dtype, device=device)
    return tuple(arena[o:o + (int(np.prod(sh)) if len(sh) else 0)].view(sh)
                 for o, sh in zip(offs, shapes))


def initial_state(prog: CompiledProgram, fields=None) -> SimulationState:
    """Zero fields and recurrences (``execute.py:814-905``); ``fields`` may give
    initial values by lower-case component name."""
    dev, dt_ = prog.device, prog.field_dtype
    names = ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez")
    bank = field_bank([tuple(prog.shapes[c]) if c in prog.shapes else (0,) for c in names],
                      dt_, dev)
    out = {c.lower(): t for c, t in zip(names, bank)}
    if fields:
        scale = prog.field_scale or {"E": 1.0, "H": 1.0}
        limit = torch.finfo(dt_).max
        for k, v in fields.items():
            stored = np.asarray(v, dtype=np.float64) / scale[k[0].upper()]
            if stored.size and np.max(np.abs(stored)) >= limit:
                raise ValueError(
                    f"initial {k} exceeds the {dt_} storage range at field scale "
                    f"{scale[k[0].upper()]:.3e}; pass field_scale= to the simulation")
            out[k].copy_(torch.as_tensor(stored, dtype=dt_))

    # both CPML banks live in one arena [2, total] so that every psi of bank B
    # sits at the same element distance from its bank-A twin
    terms = (prog.cpml.h_terms + prog.cpml.e_terms) if prog.cpml else ()
    shapes = []
    for t in terms:
        shape = list(prog.shapes[t.component])
        shape[t.axis] = t.low + t.high
        shapes.append(tuple(shape))
    total = int(sum(int(np.prod(sh)) for sh in shapes))
    arena = torch.zeros((2, max(total, 1)), dtype=prog.psi_dtype, device=dev)
    banks = ([], [])
    off = 0
    for sh in shapes:
        n = int(np.prod(sh))
        for b in (0, 1):
            banks[b].append(arena[b, off:off + n].view(sh))
        off += n
    nh = len(prog.cpml.h_terms) if prog.cpml else 0
    psi_h, psi_e = tuple(banks[0][:nh]), tuple(banks[0][nh:])
    nv = sum(6 * m.freqs.size * m.points for m in prog.monitors)
    nw = sum(m.freqs.size for m in prog.monitors)
    return SimulationState(
        **out, psi_h=psi_h, psi_e=psi_e,
        dft_re=torch.zeros(nv, dtype=torch.float64, device=dev),
        dft_im=torch.zeros(nv, dtype=torch.float64, device=dev),
        dft_w=torch.zeros(nw, dtype=torch.float64, device=dev),
        t=torch.tensor(prog.t0, dtype=torch.float64, device=dev),
        current_step=torch.tensor(0, dtype=torch.int32, device=dev),
        fields_b=(), psi_b=tuple(banks[1]))
