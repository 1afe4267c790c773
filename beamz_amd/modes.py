"""Waveguide mode sources for 2D (TMxy) simulations, lowered on the host into
one timed Hy or Hx slab (timing ``h``) and one timed Ez slab (timing ``e``):
a Huygens pair that launches the mode one way along the guide.

The planning restates the reference's 2D launch (TM branch):
- ``ModeSource`` / ``ModeSpec`` fields: ``beamz/devices/sources/specs.py:79-160``;
- snapping of the plane, its companion plane and the transverse interval:
  ``beamz/devices/_placement.py:151-189``, ``287-371``;
- target effective index ``0.98 * n_max`` by default and the mode index:
  ``beamz/devices/sources/mode_launch.py:315-353``;
- phase alignment on the peak of H, impedance matching of E to
  ``eta0 / n_eff``, Tukey window (alpha 0.3), the (J, M) profile signs, power
  normalization before and after the real projection, and the two entries:
  ``mode_launch.py:266-313``, ``356-475``, ``478-599``;
- coefficients ``profile * dt / (material * mu0|eps0 * spacing)`` and the
  waveform sampled at ``t0 + (i + 1/2) dt``:
  ``beamz/devices/sources/compiler.py:463-490``, ``592-624``.

The reference solves the transverse mode with a full-vector finite-difference
eigen-solver (``beamz/devices/modes/_scipy.py:52-237``).  Here a 2D TMxy mode
is Ez-polarized and the solve is the scalar eigenproblem of the FDTD's own
transverse operator on the Ez nodes,
    (Ez[j+1] - 2 Ez[j] + Ez[j-1]) / dx^2 + k0^2 eps[j] Ez[j] = beta^2 Ez[j],
with Ez = 0 beyond the grid and eps the Ez-site permittivity the update uses,
and the transverse H is -beta Ez / (omega mu0) (Hy for propagation along x;
Hx along y, in the solver's sign convention for y, ``plane.py:196-197``).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .constants import EPS_0, ETA_0, LIGHT_SPEED, MU_0
from .sources import CustomSource, sample_waveform


@dataclass(frozen=True)
class ModeSpec:
    """Which mode to launch (``modes/specs.py``): ``mode_index`` counts modes
    by decreasing effective index; ``target_neff`` defaults to 0.98 of the
    largest index on the source line."""

    mode_index: int = 0
    num_modes: int = 1
    target_neff: Optional[float] = None
    polarization: Optional[str] = None


@dataclass(frozen=True)
class ModeSource:
    """Inject a solved waveguide mode through a finite line (2D).

    ``center`` and ``size`` are (x, y, z) in metres; the zero extent among x
    and y selects the injection axis, ``direction`` is ``"+"`` or ``"-"``
    along it, ``source_time`` provides ``sample(times)`` and ``freq0``.

    ``launch`` selects the 3D injection: ``"auto"`` and ``"tfsf"`` lower the
    discrete total-field / scattered-field residual of the reference's 3D
    launch (``mode_launch.py:602-770``), ``"huygens"`` the two-plane J/M
    current pair.  2D sources always lower the line pair."""

    center: tuple
    size: tuple
    source_time: object
    direction: str = "+"
    mode_spec: ModeSpec = field(default_factory=ModeSpec)
    power: float = 1.0
    launch: str = "auto"

    def __post_init__(self):
        if len(self.center) != 3 or len(self.size) != 3:
            raise ValueError("ModeSource center and size must contain three values.")
        if self.direction not in ("+", "-"):
            raise ValueError("ModeSource direction must be '+' or '-'.")
        if self.launch not in ("auto", "tfsf", "huygens"):
            raise ValueError("ModeSource launch must be 'auto', 'tfsf' or 'huygens'.")
        p = float(self.power)
        if not np.isfinite(p) or p < 0.0:
            raise ValueError("ModeSource power must be non-negative and finite.")
        f0 = getattr(self.source_time, "freq0", None)
        if f0 is None or not np.isfinite(float(f0)) or float(f0) <= 0.0:
            raise ValueError("ModeSource source_time must define a positive freq0.")
        if any(not np.isfinite(float(s)) or float(s) < 0.0 for s in self.size):
            raise ValueError("ModeSource size must be non-negative and finite.")
        if not any(float(s) == 0.0 for s in self.size):
            raise ValueError("ModeSource size needs a zero extent (the injection axis).")

    @property
    def axis3d(self) -> str:
        """The injection axis of a 3D source: its single zero extent."""
        zero = [a for a, s in zip("xyz", self.size) if float(s) == 0.0]
        if len(zero) != 1:
            raise ValueError("a 3D ModeSource needs exactly one zero extent")
        return zero[0]

    @property
    def axis(self) -> str:
        zero = [a for a, s in zip("xy", self.size[:2]) if float(s) == 0.0]
        if len(zero) != 1:
            raise ValueError("a 2D ModeSource needs exactly one zero extent among x and y")
        return zero[0]

    @property
    def frequency(self) -> float:
        return float(self.source_time.freq0)

    @property
    def width(self) -> float:
        return float(self.size[1] if self.axis == "x" else self.size[0])


# ---- snapping (_placement.py:151-189, 287-371) ------------------------------
def snap_cell_center(coord, step, count):
    """Nearest cell centre: index round-half-even(coord/step - 1/2), clamped."""
    if count <= 0:
        return 0
    idx = int(round(float(coord) / float(step) - 0.5))
    return max(0, min(idx, count - 1))


def snap_centered_extent(center, extent, step, count, min_cells=1):
    """Cells [start, stop) covered by [center - extent/2, center + extent/2]."""
    lo, hi = float(center) - 0.5 * float(extent), float(center) + 0.5 * float(extent)
    lo, hi = min(lo, hi), max(lo, hi)
    if count <= 0:
        return 0, 0
    eps = 1e-12 * max(abs(lo), abs(hi), abs(float(step)), 1.0)
    start = max(0, min(int(np.floor((lo + eps) / step)), count))
    stop = max(0, min(int(np.ceil((hi - eps) / step)), count))
    need = max(1, int(min_cells))
    if stop - start < need:
        cell = max(0, min(int(np.floor(0.5 * (lo + hi) / step)), count - 1))
        stop = min(count, cell + need)
        start = max(0, stop - need)
    return start, stop


# ---- the transverse mode ----------------------------------------------------
def solve_line_modes(eps_line, omega, dx, count, target_neff):
    """Ez-polarized modes of a line of Ez nodes with permittivities
    ``eps_line``: the ``count`` eigenpairs whose n_eff^2 lie nearest
    ``target_neff^2`` (the reference's shift-invert selection,
    ``_scipy.py:115-124``), by decreasing effective index.  Returns
    (neff[count], ez[count, N]) with unit-norm real profiles, each signed so
    that its largest-magnitude sample is positive.

    The eigenvalues come first (all of them, no vectors); the ``count``
    nearest the target are a contiguous run of the ascending spectrum, and
    only their vectors are computed (bisection and inverse iteration), which
    on a 600-node line takes a few milliseconds against 0.25 s for the full
    eigendecomposition."""
    from scipy.linalg import eigh_tridiagonal
    eps = np.asarray(eps_line, dtype=np.float64)
    n = eps.size
    k0 = float(omega) / LIGHT_SPEED
    main = -2.0 / dx ** 2 + k0 ** 2 * eps
    off = np.full(n - 1, 1.0 / dx ** 2)
    count = max(1, min(int(count), n))
    w = eigh_tridiagonal(main, off, eigvals_only=True, lapack_driver="stev")   # ascending
    near = np.argsort(np.abs(w / k0 ** 2 - float(target_neff) ** 2))[:count]
    lo, hi = int(near.min()), int(near.max())
    w, v = eigh_tridiagonal(main, off, select="i", select_range=(lo, hi))
    n2 = w / k0 ** 2
    order = np.argsort(n2)[::-1][:count]
    v = v[:, order]
    peak = np.argmax(np.abs(v), axis=0)
    v = v * np.sign(v[peak, np.arange(v.shape[1])])[None, :]
    neff = np.sqrt(np.maximum(n2[order], 0.0))
    return neff, v.T


def _modal_power_2d(e, h, sign, dl):
    """0.5 Re sum(dl * sign * e * conj(h)) over the common length
    (``mode_launch.py:266-277``)."""
    e = np.asarray(e, dtype=np.complex128).reshape(-1)
    h = np.asarray(h, dtype=np.complex128).reshape(-1)
    n = min(e.size, h.size)
    if n <= 0:
        return 0.0
    w = np.asarray(dl, dtype=float)
    w = w if w.ndim == 0 else w.reshape(-1)[:n]
    return float(0.5 * np.real(np.sum(w * float(sign) * e[:n] * np.conjugate(h[:n]))))


def _normalize_pair(h, e, sign, dl):
    """Scale (h, e) to unit modal power (``mode_launch.py:280-288``)."""
    p = _modal_power_2d(e, h, sign, dl)
    if np.isfinite(p) and abs(p) > 1e-30:
        s = np.sqrt(1.0 / abs(p))
        return np.asarray(h) * s, np.asarray(e) * s
    return np.asarray(h), np.asarray(e)


def _impedance_match(e, h, z_target):
    """Scale e so that |sum e h* / sum |h|^2| = |z_target|
    (``mode_launch.py:291-304``)."""
    e = np.asarray(e, dtype=np.complex128).reshape(-1)
    h = np.asarray(h, dtype=np.complex128).reshape(-1)
    n = min(e.size, h.size)
    if n <= 0:
        return e
    denom = np.sum(h[:n] * np.conjugate(h[:n]))
    if abs(denom) <= 1e-12:
        return e
    z = abs(np.sum(e[:n] * np.conjugate(h[:n])) / denom)
    if not np.isfinite(z) or z <= 1e-12:
        return e
    return e * (abs(z_target) / z)


def _window(width, alpha=0.3):
    """Tukey taper over the source line (``mode_launch.py:307-312``)."""
    if width > 2:
        from scipy.signal.windows import tukey
        return tukey(width, alpha=alpha)
    return np.ones(max(1, width))


def _launch_pair(first, second, sign, dl, power):
    """Unit power, real projection, unit power again, then sqrt(power)
    (``mode_launch.py:356-364``, ``205-210``)."""
    first, second = _normalize_pair(first, second, sign, dl)
    first, second = np.real(first), np.real(second)
    first, second = _normalize_pair(first, second, sign, dl)
    s = np.sqrt(float(power))
    return first * s, second * s


@dataclass(frozen=True)
class ModePlan:
    """The planned launch: the solved index, the transverse support and the
    two (component, timing, index, profile) entries before coefficient
    scaling."""

    neff: float
    support: tuple
    entries: tuple


def plan_mode_source_2d(src: ModeSource, eps_ez, resolution):
    """Plan the Huygens pair of a 2D mode source (``mode_launch.py:371-599``).

    ``eps_ez`` is the Ez-site permittivity (ny+1, nx+1) of the simulation."""
    eps_ez = np.asarray(eps_ez, dtype=np.float64)
    ny, nx = eps_ez.shape[0] - 1, eps_ez.shape[1] - 1
    axis = src.axis
    dx = float(resolution)
    pos = {"x": 0, "y": 1}
    counts = {"x": nx, "y": ny}
    center_idx = snap_cell_center(src.center[pos[axis]], dx, counts[axis])
    cmax = max(counts[axis] - 2, 0)
    offset_idx = max(0, center_idx - 1) if src.direction == "+" else min(cmax, center_idx + 1)
    trans = "y" if axis == "x" else "x"
    start, stop = snap_centered_extent(src.center[pos[trans]], src.width, dx, counts[trans])
    line = eps_ez[:, center_idx] if axis == "x" else eps_ez[center_idx, :]

    spec = src.mode_spec
    if spec.polarization not in (None, "tm"):
        raise ValueError("a 2D TMxy simulation launches Ez-polarized (tm) modes only")
    omega = 2.0 * np.pi * src.frequency
    target = spec.target_neff if spec.target_neff is not None else 0.98 * np.sqrt(line.max())
    # as the reference: max(num_modes, mode_index + 1, 3) requested, 2 m + 5
    # solved near the target (plane.py:165-172), sorted by decreasing n_eff
    m = max(int(spec.num_modes), int(spec.mode_index) + 1, 3)
    neffs, ezs = solve_line_modes(line, omega, dx, 2 * m + 5, target)
    mode = min(int(spec.mode_index), neffs.size - 1)
    neff = float(neffs[mode])
    ez = ezs[mode]
    # transverse H of the mode, -beta Ez / (omega mu0) (Faraday's law for
    # propagation along +x; for +y the solver's H carries the flipped sign of
    # plane.py:196-197, which the entry signs below expect), negated for "-"
    beta = neff * omega / LIGHT_SPEED
    h = -beta / (omega * MU_0) * ez
    if src.direction == "-":
        h = -h

    dir_sign = 1.0 if src.direction == "+" else -1.0
    z_target = ETA_0 / max(neff, 1e-6)
    phase_ref = np.angle(h.astype(np.complex128)[np.argmax(np.abs(h))])
    h_al = h * np.exp(-1j * phase_ref)
    e_al = _impedance_match(ez * np.exp(-1j * phase_ref), h_al, z_target)
    win = _window(stop - start)
    h_crop, e_crop = h_al[start:stop], e_al[start:stop]
    if len(h_crop) == len(win):
        h_crop, e_crop = h_crop * win, e_crop * win
    jz = -dir_sign * h_crop
    magnetic = dir_sign * (1.0 if axis == "x" else -1.0) * e_crop
    flux_sign = -1.0 if axis == "x" else 1.0
    jz, magnetic = _launch_pair(jz, magnetic, flux_sign, dx, src.power)
    h_comp = "Hy" if axis == "x" else "Hx"
    if axis == "x":
        h_index, e_index = (start, offset_idx), (start, center_idx)
    else:
        h_index, e_index = (offset_idx, start), (center_idx, start)
    entries = ((h_comp, "h", h_index, -np.asarray(magnetic, dtype=np.float64)),
               ("Ez", "e", e_index, np.asarray(jz, dtype=np.float64)))
    return ModePlan(neff=neff, support=(start, stop), entries=entries)


def lower_mode_source_2d(src: ModeSource, eps_ez, resolution, dt, t0, total_steps):
    """The two timed slab sources of a 2D mode source: coefficient
    ``profile * dt / (mu0 * spacing)`` on H and ``profile * dt / (eps0 *
    eps_Ez * spacing)`` on Ez, waveform at ``t0 + (i + 1/2) dt``
    (``sources/compiler.py:592-624``, ``463-490``)."""
    plan = plan_mode_source_2d(src, eps_ez, resolution)
    wave = sample_waveform(src.source_time, t0, dt, total_steps, half_step=True)
    eps_ez = np.asarray(eps_ez, dtype=np.float64)
    axis = src.axis
    out = []
    for comp, timing, (r0, c0), prof in plan.entries:
        if comp.startswith("H"):
            denom = MU_0 * np.ones_like(prof) * resolution
        else:
            sites = eps_ez[r0:r0 + prof.size, c0] if axis == "x" else eps_ez[r0, c0:c0 + prof.size]
            denom = EPS_0 * sites * resolution
        coeff = prof * dt / denom
        shape = (prof.size, 1) if axis == "x" else (1, prof.size)
        out.append(CustomSource(comp, (r0, c0), coeff.reshape(shape), wave, timing))
    return out
