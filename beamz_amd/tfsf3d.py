"""3D mode sources launched by the discrete total-field / scattered-field
(TF/SF) residual of the Yee step.

The reference launches a 3D ``ModeSource`` through
``sources/mode_launch.py:602-770``.  It builds the incident mode as a complex
carrier phasor F, replicated over the planes along the propagation axis a
(``planar_tfsf.py:97-155``), and masks it to the launched side
(``planar_tfsf.py:189-229``).  It then advances the full and the masked state
through one source-free H and E half-step and keeps the differences

    dH = mask(U_H F) - U_H(mask F)
    dE = mask(U_E(F, U_H F)) - U_E(mask F, mask(U_H F))

cropped to their support (``planar_tfsf.py:253-287``, ``529-768``).  Added
after the H and E updates of every step with the carrier's in-phase and
quadrature waveforms (real part x env cos, -imag part x env sin,
``sources/compiler.py:312-345``, ``463-490``), they keep the grid's field on
the launched side equal to the masked advance of the full incident field:
the mode appears there and nowhere else.  The amplitude is then scaled by
1/sqrt(ratio), where ratio is the launched power over the requested one.
The launched power is that of the reconstructed first step, measured with
the flux monitor's quadrature on the source plane (``mode_launch.py:771-917``).

Mask (``planar_tfsf.py:189-229``).  Components staggered along a (E_a, H_b,
H_c) are launched at a-indices i with d (i - q) >= 0, the others (E_b, E_c,
H_a) at d (i - p) >= 0.  Here d = +-1 is the direction, p the E plane and q
the companion plane, one cell behind p for "+" and one ahead for "-"
(``_placement.py:305-330``).  For "+" the residual sits on H_b, H_c at q and
on E_b, E_c at q.  For "-" it covers two staggered and two integer planes.

Incident field.  The launch is exact when F solves the discrete equations,
and that is where this module differs from the reference.  The reference
takes the wave number from a continuum mode solve and a 1D dispersion
correction (``modes/fields.py:54-69``).  Here the mode is the exact discrete
mode of the lattice, with every sample referenced to its own position and
time.  On the Yee lattice a field F = A exp(i (omega t - k a)) sampled at
the Yee positions and times turns each centred difference into a product:
d/dt -> i w~ with w~ = (2/dt) sin(omega dt/2), and d/da -> -i b~ with
b~ = (2/dx) sin(k dx/2).  The discrete equations are therefore the plane
solver's equations with (omega, beta) replaced by (w~, b~).  The solver run at
w~ returns b~, from which k = (2/dx) arcsin(b~ dx/2).  Its fields, with the
longitudinal ones E_a = -i G/b~ and H_a = i C/(w~ mu0), are the amplitudes:
E_b, E_c and H_a at integer a-indices, E_a, H_b and H_c half a cell ahead.
The backward mode keeps E_b, E_c, H_a and negates E_a, H_b, H_c.  Inside a
PEC window this F is an exact solution.  ``tests/test_tfsf3d.py`` checks
one oracle step of it against exp(i omega dt) to 1e-12, and its residuals
against the textbook TF/SF corrections.

Only isotropic uniform grids and lossless material at the source are lowered,
as in the reference (``mode_launch.py:607-611``).  The residual box must lie
outside the CPML, where the update is the plain curl applied here.
"""

from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np

from .constants import EPS_0, LIGHT_SPEED, MU_0
from .lattice import OFFSETS_3D, plane_gather_plan
from .modes import ModeSource
from .modes3d import _cyclic, plane_setup, solve_source_mode
from .sources import CustomSource, sample_waveform_pair

_STORAGE = ("z", "y", "x")
# residual box half-length along a, in cells around the E plane; the
# reference replicates the incident phasor over +-12 planes
# (mode_launch.py:723-736), of which only the planes next to the mask matter
_HALF_BOX = 5


def _offset(component, axis):
    return OFFSETS_3D[component][_STORAGE.index(axis)]


def staggered_along(component, axis) -> bool:
    """``planar_tfsf.py:33-37``: E_a, H_b and H_c sit half a cell along a."""
    return _offset(component, axis) == 0.5


@dataclass(frozen=True)
class DiscreteMode3d:
    """The exact discrete mode a 3D TF/SF launch injects.

    ``profiles`` maps each of the six components to its complex amplitude on
    the window in (b, c) order, at global b-index ``b0 + j`` and c-index
    ``c0 + k``.  The integer-index components (E_b, E_c, H_a) sit on plane
    ``plane``, the staggered ones on ``companion``."""

    axis: str
    sign: float
    omega: float
    omega_lattice: float
    k_axis: float
    neff: float
    plane: int
    companion: int
    window: tuple
    profiles: dict = field(repr=False)
    resolution: float = 0.0
    dt: float = 0.0

    @property
    def abc(self):
        return _cyclic(self.axis)

    def position(self, component, index) -> float:
        """Lattice coordinate along a (metres) of a-index ``index``."""
        half = 0.5 if staggered_along(component, self.axis) else 0.0
        return (float(index) + half) * self.resolution

    @property
    def ref_position(self) -> float:
        """Phase reference: the staggered companion plane (``discrete.py:304-310``)."""
        return (self.companion + 0.5) * self.resolution

    def launched(self, component, index) -> bool:
        """``planar_tfsf.py:189-214``: staggered components are launched on the
        side of the companion plane, the others on the side of the E plane."""
        edge = self.companion if staggered_along(component, self.axis) else self.plane
        return self.sign * (int(index) - edge) >= 0


def solve_discrete_mode(src: ModeSource, eps_site, grid, resolution, dt) -> DiscreteMode3d:
    """The exact discrete mode of the source's window, scaled to unit power by
    the solver's own quadrature, travelling along ``src.direction``."""
    setup = plane_setup(src, eps_site, grid, resolution)
    dx, dt = float(resolution), float(dt)
    omega = 2.0 * np.pi * src.frequency
    w_lat = 2.0 / dt * np.sin(0.5 * omega * dt)
    modes, m = solve_source_mode(src, setup, w_lat, dx)
    b_lat = float(modes.beta[m])
    if not 0.0 < 0.5 * b_lat * dx < 1.0:
        raise ValueError("the mode's lattice wave number exceeds the grid's Nyquist limit")
    k = 2.0 / dx * np.arcsin(0.5 * b_lat * dx)
    d = 1.0 if src.direction == "+" else -1.0
    a, b, c = _cyclic(setup.axis)
    profiles = {
        "E" + b: modes.eb[m].astype(np.complex128),
        "E" + c: modes.ec[m].astype(np.complex128),
        "E" + a: 1j * d * modes.ea[m],
        "H" + b: d * modes.hb[m].astype(np.complex128),
        "H" + c: d * modes.hc[m].astype(np.complex128),
        "H" + a: 1j * modes.ha[m],
    }
    return DiscreteMode3d(axis=setup.axis, sign=d, omega=omega, omega_lattice=w_lat, k_axis=d * k,
                          neff=k * LIGHT_SPEED / omega, plane=setup.plane,
                          companion=setup.companion, window=setup.window, profiles=profiles,
                          resolution=dx, dt=dt)


# --------------------------------------------------------------------------
# the local box, in (a, b, c) order
# --------------------------------------------------------------------------
@dataclass(frozen=True)
class Box:
    """Cells ``[lo, hi)`` per axis name; component arrays hold the Yee
    samples of those cells (one more along every axis where the component
    sits on nodes)."""

    lo: dict
    hi: dict
    order: tuple            # (a, b, c)

    def shape(self, component):
        return tuple(self.hi[ax] - self.lo[ax] + (1 if _offset(component, ax) == 0.0 else 0)
                     for ax in self.order)

    def to_storage(self, arr):
        perm = [self.order.index(ax) for ax in _STORAGE]
        return np.transpose(arr, perm)

    def from_storage(self, arr):
        perm = [_STORAGE.index(ax) for ax in self.order]
        return np.transpose(arr, perm)

    def storage_slices(self, component):
        """Slices of the box's samples in the global storage-order array."""
        return tuple(slice(self.lo[ax], self.lo[ax] + self.shape(component)[self.order.index(ax)])
                     for ax in _STORAGE)


def source_box(mode: DiscreteMode3d, grid, half=_HALF_BOX) -> Box:
    a, b, c = mode.abc
    counts = dict(zip(_STORAGE, (int(v) for v in grid)))
    (b0, b1), (c0, c1) = mode.window
    lo = {a: max(0, mode.plane - half), b: max(0, b0 - 1), c: max(0, c0 - 1)}
    hi = {a: min(counts[a], mode.plane + half + 1), b: min(counts[b], b1 + 1),
          c: min(counts[c], c1 + 1)}
    return Box(lo, hi, (a, b, c))


def incident_state(mode: DiscreteMode3d, box: Box, *, t_e, t_h, masked, profiles=None):
    """``planar_tfsf.py:97-155``: the carrier phasor of the incident mode on
    the box, ``profile exp(i (omega t - k (pos - ref)))`` on every a-plane of
    the box, E at time ``t_e`` and H at ``t_h``; ``masked`` keeps the
    launched side only.  Returns complex arrays in (a, b, c) order."""
    a, b, c = mode.abc
    (b0, _), (c0, _) = mode.window
    profiles = mode.profiles if profiles is None else profiles
    out = {}
    for comp, prof in profiles.items():
        arr = np.zeros(box.shape(comp), dtype=np.complex128)
        jb, kc = b0 - box.lo[b], c0 - box.lo[c]
        t = float(t_e if comp.startswith("E") else t_h)
        for i in range(arr.shape[0]):
            g = box.lo[a] + i
            if masked and not mode.launched(comp, g):
                continue
            phase = mode.omega * t - mode.k_axis * (mode.position(comp, g) - mode.ref_position)
            arr[i, jb:jb + prof.shape[0], kc:kc + prof.shape[1]] = prof * np.exp(1j * phase)
        out[comp] = arr
    return out


def mask_state(mode: DiscreteMode3d, box: Box, state):
    """``planar_tfsf.py:217-229``: zero the planes off the launched side."""
    a = mode.abc[0]
    out = {}
    for comp, arr in state.items():
        keep = np.array([mode.launched(comp, box.lo[a] + i) for i in range(arr.shape[0])])
        out[comp] = arr * keep[:, None, None]
    return out


def _bdiff(h, axis):
    """Backward difference with a zero sample beyond each end."""
    pad = [(0, 0)] * 3
    pad[axis] = (1, 1)
    return np.diff(np.pad(h, pad), axis=axis)


def advance_h(mode: DiscreteMode3d, state):
    """Source-free H half-step H -= dt / (mu0 dx) curl E on the box
    (``planar_tfsf.py:553-578``)."""
    a, b, c = mode.abc
    ch = mode.dt / (MU_0 * mode.resolution)
    Ea, Eb, Ec = state["E" + a], state["E" + b], state["E" + c]
    return {
        "H" + a: state["H" + a] - ch * (np.diff(Ec, axis=1) - np.diff(Eb, axis=2)),
        "H" + b: state["H" + b] - ch * (np.diff(Ea, axis=2) - np.diff(Ec, axis=0)),
        "H" + c: state["H" + c] - ch * (np.diff(Eb, axis=0) - np.diff(Ea, axis=1)),
    }


def advance_e(mode: DiscreteMode3d, state, h_next, ce, walls=None):
    """Source-free E half-step E += dt / (eps0 eps dx) curl H on the box
    (``planar_tfsf.py:581-626``); ``ce`` maps E components to that factor and
    ``walls`` to a 0/1 mask of the samples on metallic grid faces."""
    a, b, c = mode.abc
    Ha, Hb, Hc = h_next["H" + a], h_next["H" + b], h_next["H" + c]
    out = {
        "E" + a: state["E" + a] + ce["E" + a] * (_bdiff(Hc, 1) - _bdiff(Hb, 2)),
        "E" + b: state["E" + b] + ce["E" + b] * (_bdiff(Ha, 2) - _bdiff(Hc, 0)),
        "E" + c: state["E" + c] + ce["E" + c] * (_bdiff(Hb, 0) - _bdiff(Ha, 1)),
    }
    if walls is not None:
        out = {k: v * walls[k] for k, v in out.items()}
    return out


_FACES = {"z": ("front", "back"), "y": ("bottom", "top"), "x": ("left", "right")}


def wall_masks(box: Box, grid, metallic):
    """0/1 masks (in (a, b, c) order) of the E samples on the metallic grid
    faces the box touches: tangential E vanishes there, as the kernels' PEC
    masks leave it (``execute.py:467-541``).  Beyond the box the H ghost is
    zero, which is the metallic face's ghost (``lattice.py:523-575``) and, on
    interior box faces, the incident field's value outside its window."""
    counts = dict(zip(_STORAGE, (int(v) for v in grid)))
    out = {}
    for comp in ("Ex", "Ey", "Ez"):
        mask = np.ones(box.shape(comp))
        for i, ax in enumerate(box.order):
            if comp[1] == ax or _offset(comp, ax) != 0.0:
                continue
            low, high = _FACES[ax]
            sl = [slice(None)] * 3
            if box.lo[ax] == 0 and low in metallic:
                sl[i] = 0
                mask[tuple(sl)] = 0.0
            if box.hi[ax] == counts[ax] and high in metallic:
                sl[i] = -1
                mask[tuple(sl)] = 0.0
        out[comp] = mask
    return out


@dataclass(frozen=True)
class Residual:
    """One cropped complex residual slab (``planar_tfsf.py:62-75``):
    ``values`` in storage order starting at the global index ``start``."""

    component: str
    timing: str
    start: tuple
    values: np.ndarray


def _crop(box: Box, component, timing, arr, atol=1e-30):
    """``planar_tfsf.py:253-287``: the bounding box of |value| > atol."""
    st = box.to_storage(arr)
    live = np.abs(st) > atol
    if not live.any():
        return None
    where = np.argwhere(live)
    lo, hi = where.min(axis=0), where.max(axis=0) + 1
    region = tuple(slice(int(l), int(h)) for l, h in zip(lo, hi))
    base = [s.start for s in box.storage_slices(component)]
    start = tuple(int(bs + l) for bs, l in zip(base, lo))
    return Residual(component, timing, start, np.ascontiguousarray(st[region]))


@dataclass(frozen=True)
class TfsfPlan:
    """A planned launch: the mode, its residual slabs (already scaled by
    ``amplitude_scale``) and the launch-power diagnostics
    (``mode_launch.py:755-770``)."""

    mode: DiscreteMode3d
    residuals: tuple
    launch_power_ratio: float
    unscaled_launched_power: float
    amplitude_scale: float


def _e_factor(mode, box, eps_site, sigma_site=None):
    ce = {}
    for comp in ("Ex", "Ey", "Ez"):
        eps = box.from_storage(np.asarray(eps_site[comp])[box.storage_slices(comp)])
        if sigma_site is not None:
            sig = np.asarray(sigma_site[comp])[box.storage_slices(comp)]
            if np.any(sig != 0.0):
                raise ValueError("a 3D TF/SF mode launch needs lossless material around the "
                                 "source (mode_launch.py:607-611)")
        ce[comp] = mode.dt / (EPS_0 * eps * mode.resolution)
    return ce


def _residual_states(mode, box, ce, walls=None, profiles=None):
    """The two half-step residuals (``planar_tfsf.py:629-768``) and the
    reconstructed first launched state (``mode_launch.py:780-833``)."""
    kw = dict(t_e=0.0, t_h=-0.5 * mode.dt, profiles=profiles)
    full = incident_state(mode, box, masked=False, **kw)
    masked = incident_state(mode, box, masked=True, **kw)
    h_full = advance_h(mode, full)
    h_target = mask_state(mode, box, h_full)
    h_mask = advance_h(mode, masked)
    e_full = advance_e(mode, full, h_full, ce, walls)
    e_target = mask_state(mode, box, e_full)
    e_mask = advance_e(mode, masked, h_target, ce, walls)
    dh = {k: h_target[k] - h_mask[k] for k in h_target}
    de = {k: e_target[k] - e_mask[k] for k in e_target}
    launched = {**{k: e_mask[k] + de[k] for k in de}, **{k: h_mask[k] + dh[k] for k in dh}}
    return dh, de, launched


def plane_flux(mode: DiscreteMode3d, box: Box, state) -> float:
    """Power of a phasor state (E and H referenced to one time) through the
    source plane, by the flux monitor's quadrature over the window's cells:
    cell-centre interpolation of every component, 1/2 Re sum d (E_b H_c* -
    E_c H_b*) dx^2 (``mode_launch.py:873-917``, ``observe.py:198-234``)."""
    a, b, c = mode.abc
    (b0, b1), (c0, c1) = mode.window
    dx = mode.resolution
    ranges = {b: (b0 - box.lo[b], b1 - box.lo[b]), c: (c0 - box.lo[c], c1 - box.lo[c])}
    sampled = {}
    for comp in ("E" + b, "E" + c, "H" + b, "H" + c):
        arr = box.to_storage(state[comp])
        idx, w = plane_gather_plan(comp, arr.shape, dx, a, mode.plane - box.lo[a], ranges)
        sampled[comp] = (arr.reshape(-1)[idx] * w).sum(axis=1)
    s = sampled["E" + b] * np.conj(sampled["H" + c]) - sampled["E" + c] * np.conj(sampled["H" + b])
    return float(0.5 * np.real(np.sum(mode.sign * s)) * dx * dx)


def deembed(mode: DiscreteMode3d, box: Box, state, *, t_e, t_h):
    """``planar_tfsf.py:158-186``: the profiles read back from a state at
    the profile planes, with the carrier phase at ``t_e`` / ``t_h`` removed."""
    a, b, c = mode.abc
    (b0, _), (c0, _) = mode.window
    jb, kc = b0 - box.lo[b], c0 - box.lo[c]
    out = {}
    for comp, prof in mode.profiles.items():
        g = mode.companion if staggered_along(comp, a) else mode.plane
        t = float(t_e if comp.startswith("E") else t_h)
        phase = mode.omega * t - mode.k_axis * (mode.position(comp, g) - mode.ref_position)
        plane = state[comp][g - box.lo[a], jb:jb + prof.shape[0], kc:kc + prof.shape[1]]
        out[comp] = plane * np.exp(-1j * phase)
    return out


def plan_tfsf_mode_source_3d(src: ModeSource, eps_site, grid, resolution, dt,
                             sigma_site=None, metallic=frozenset()) -> TfsfPlan:
    """Plan the TF/SF launch of a 3D mode source (``mode_launch.py:602-770``);
    ``metallic`` names the grid's PEC faces."""
    mode = solve_discrete_mode(src, eps_site, grid, resolution, dt)
    a = mode.abc[0]
    counts = dict(zip(_STORAGE, (int(v) for v in grid)))
    box = source_box(mode, grid)
    # the residual planes and their stencils stay inside the box
    if box.lo[a] > min(mode.plane, mode.companion) - 3 or \
            box.hi[a] < max(mode.plane, mode.companion) + 4:
        raise ValueError(f"a 3D TF/SF mode source needs 4 cells of grid on each side along {a}")
    if counts[a] < 8:
        raise ValueError("the grid is too short along the propagation axis")
    scale0 = np.sqrt(float(src.power)) if float(src.power) > 0.0 else 0.0
    profiles = {k: v * scale0 for k, v in mode.profiles.items()}
    mode = replace(mode, profiles=profiles)
    ce = _e_factor(mode, box, eps_site, sigma_site)
    dh, de, launched = _residual_states(mode, box, ce, wall_masks(box, grid, frozenset(metallic)))
    # launch-power diagnostics: the first launched state, de-embedded at the
    # profile planes, rebuilt at one time and measured on the source plane
    ratio, unscaled = None, None
    if float(src.power) > 1e-30:
        prof = deembed(mode, box, launched, t_e=mode.dt, t_h=0.5 * mode.dt)
        state = incident_state(mode, box, t_e=0.0, t_h=0.0, masked=False, profiles=prof)
        unscaled = plane_flux(mode, box, state)
        ratio = unscaled / float(src.power)
        if not np.isfinite(ratio) or ratio <= 1e-24:
            ratio = None
    amp = 1.0 if ratio is None else 1.0 / np.sqrt(ratio)
    residuals = []
    for timing, delta in (("h", dh), ("e", de)):
        for comp in sorted(delta):
            r = _crop(box, comp, timing, delta[comp] * amp)
            if r is not None:
                residuals.append(r)
    return TfsfPlan(mode=mode, residuals=tuple(residuals),
                    launch_power_ratio=float("nan") if ratio is None else float(ratio),
                    unscaled_launched_power=float("nan") if unscaled is None else float(unscaled),
                    amplitude_scale=float(amp))


def lower_tfsf_mode_source_3d(src: ModeSource, eps_site, grid, resolution, dt, t0, total_steps,
                              sigma_site=None, metallic=frozenset()):
    """The timed slab sources of a TF/SF launch: each residual's real part
    with the waveform ``env cos`` and its negated imaginary part with the
    quadrature ``env sin``, both sampled at ``t0 + i dt``, timings "h" and "e"
    (``sources/compiler.py:312-345``, ``463-520``)."""
    plan = plan_tfsf_mode_source_3d(src, eps_site, grid, resolution, dt, sigma_site, metallic)
    wave, quad = sample_waveform_pair(src.source_time, t0, dt, total_steps, half_step=False)
    out = []
    for r in plan.residuals:
        out.append(CustomSource(r.component, r.start, np.real(r.values).copy(), wave, r.timing))
        if np.abs(np.imag(r.values)).max() > 1e-30:
            out.append(CustomSource(r.component, r.start, -np.imag(r.values), quad, r.timing))
    return out
