"""Source waveforms and slab sources, lowered on the host into the nine
(timing, component) slab groups the kernels consume.

References: ``beamz/devices/sources/time.py:361-486`` (``GaussianPulse``),
``time.py:67-81`` (waveform sampling), ``sources/compiler.py:418-460``
(Gaussian source: coefficient ``-profile dt / (eps0 eps_site)``, waveform at
``t0 + (i + 1/2) dt``, timing ``pre_e``), ``sources/specs.py:445-499``
(``CustomSource``), ``beamz/simulation/execute.py:48-52`` (phase table).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from .constants import EPS_0

TIMINGS = ("pre_e", "h", "e")
PHASE_COMPONENTS = {"pre_e": ("Ex", "Ey", "Ez"), "h": ("Hx", "Hy", "Hz"), "e": ("Ex", "Ey", "Ez")}


@dataclass(frozen=True)
class GaussianPulse:
    """Gaussian-modulated carrier (``sources/time.py:361-486``).

    ``carrier`` selects ``cos`` (the reference's form) or ``sin``."""

    freq0: float
    fwidth: float
    amplitude: float = 1.0
    offset: float = 4.0
    remove_dc_component: bool = True
    carrier: str = "cos"

    def sample(self, times):
        t = np.asarray(times, dtype=np.float64)
        width = 1.0 / (2.0 * np.pi * max(float(self.fwidth), 1e-30))
        peak = float(self.offset) / max(float(self.fwidth), 1e-30)
        env = float(self.amplitude) * np.exp(-((t - peak) ** 2) / (2.0 * width ** 2))
        phase = 2.0 * np.pi * float(self.freq0) * t
        sig = env * (np.cos(phase) if self.carrier == "cos" else np.sin(phase))
        quad = env * (np.sin(phase) if self.carrier == "cos" else -np.cos(phase))
        if self.remove_dc_component and sig.size:
            sig = sig - np.mean(sig)
        return sig, quad


def sample_waveform_pair(signal, t0, dt, total_steps, half_step=False):
    """In-phase and quadrature samples on ``t0 + (i [+ 1/2]) dt``
    (``time.py:84-135``): a waveform object gives both; a callable has a zero
    quadrature; a sampled array takes the imaginary part of its analytic
    signal."""
    times = float(t0) + (np.arange(int(total_steps), dtype=np.float64) + (0.5 if half_step else 0.0)) * dt
    if hasattr(signal, "sample"):
        sig, quad = signal.sample(times)
        return np.asarray(sig, dtype=np.float64), np.asarray(quad, dtype=np.float64)
    sig = sample_waveform(signal, t0, dt, total_steps, half_step=half_step)
    if callable(signal):
        return sig, np.zeros_like(sig)
    from scipy.signal import hilbert
    return sig, np.imag(hilbert(sig)) if sig.size else sig.copy()


def sample_waveform(signal, t0, dt, total_steps, half_step=True):
    """Sample a waveform on ``t0 + (i + 1/2) dt`` (``time.py:67-81``)."""
    times = float(t0) + (np.arange(int(total_steps), dtype=np.float64) + (0.5 if half_step else 0.0)) * dt
    if hasattr(signal, "sample"):
        return np.asarray(signal.sample(times)[0], dtype=np.float64)
    if callable(signal):
        return np.asarray([float(signal(float(t))) for t in times], dtype=np.float64)
    arr = np.asarray(signal, dtype=np.float64).reshape(-1)
    if arr.size < total_steps:
        arr = np.concatenate([arr, np.full(total_steps - arr.size, arr[-1] if arr.size else 0.0)])
    return arr[:total_steps]


@dataclass(frozen=True)
class CustomSource:
    """An explicit slab injection (``sources/specs.py:445-499``):
    ``field[start:start+coeff.shape] += coeff * waveform[clip(step)]``."""

    component: str
    start: tuple
    coeff: np.ndarray
    waveform: np.ndarray
    timing: str = "pre_e"


@dataclass(frozen=True)
class SoftSource:
    """A soft current source on an index-space slab of one E component with a
    transverse ``profile`` (array of the slab's shape; ones by default).

    Lowered like the reference's ``GaussianSource``
    (``sources/compiler.py:418-460``): coefficient ``-profile dt /
    (eps0 eps_site)``, waveform sampled at ``t0 + (i + 1/2) dt``, timing
    ``pre_e``."""

    component: str
    start: tuple
    shape: tuple
    signal: object
    profile: Optional[np.ndarray] = None


@dataclass(frozen=True)
class GaussianSource:
    """A soft Ez source with a Gaussian spatial profile placed in physical
    coordinates (``sources/specs.py:207-252``): ``position`` (x, y) or
    (x, y, z) in metres, standard width ``width`` in metres, and ``signal``,
    an array of samples on the time grid (the reference's form; a waveform
    object with ``sample`` or a callable of time is also taken)."""

    position: tuple
    width: float
    signal: object

    def __post_init__(self):
        position = tuple(float(v) for v in self.position)
        if len(position) not in (2, 3) or not all(np.isfinite(position)):
            raise ValueError("GaussianSource position must be a finite 2D or 3D coordinate.")
        width = float(self.width)
        if not np.isfinite(width) or width <= 0.0:
            raise ValueError("GaussianSource width must be a positive finite value.")
        object.__setattr__(self, "position", position)
        object.__setattr__(self, "width", width)


def interpolate_signal(signal, time, dt):
    """A sample array read at ``time`` (``time.py:17-32``): linear between
    samples ``floor(time / dt)`` and the next, the last sample at the last
    index, 0 before the first and after the last."""
    arr = np.asarray(signal, dtype=np.float64).reshape(-1)
    if arr.size == 0:
        return 0.0
    pos = float(time / dt)
    low = int(np.floor(pos))
    frac = pos - low
    if 0 <= low < arr.size - 1:
        return float((1.0 - frac) * arr[low] + frac * arr[low + 1])
    if low == arr.size - 1:
        return float(arr[low])
    return 0.0


def gaussian_profile(src: GaussianSource, shape, dx, is_3d):
    """The Ez-site Gaussian of the reference's grid path
    (``sources/compiler.py:649-700``): on each axis the Ez coordinates within
    ``4 width`` of the centre (the nearest one if none is), and
    ``exp(-r^2 / (2 width^2))`` on their product.  Ez sits on nodes in x and
    y, and on cell centres in z.  Returns ``(start, profile)``."""
    centre = dict(zip(("x", "y", "z"), src.position))
    axes = ("z", "y", "x") if is_3d else ("y", "x")
    starts, coords = [], []
    for axis, n in zip(axes, shape):
        values = (np.arange(n) + (0.5 if axis == "z" else 0.0)) * dx
        active = np.flatnonzero(np.abs(values - centre[axis]) <= 4.0 * src.width)
        if active.size:
            lo, hi = int(active[0]), int(active[-1]) + 1
        else:
            lo = int(np.argmin(np.abs(values - centre[axis])))
            hi = lo + 1
        starts.append(lo)
        coords.append(values[lo:hi] - centre[axis])
    mesh = np.meshgrid(*coords, indexing="ij")
    r2 = sum(m * m for m in mesh)
    return tuple(starts), np.exp(-r2 / (2.0 * src.width ** 2))


def lower_gaussian_source(src: GaussianSource, eps_ez, dx, dt, t0, total_steps):
    """``sources/compiler.py:418-460`` for TMxy and 3D: coefficient
    ``-profile dt / (eps0 eps_Ez)``, waveform at ``t0 + (i + 1/2) dt``,
    timing ``pre_e``."""
    eps_ez = np.asarray(eps_ez, dtype=np.float64)
    is_3d = eps_ez.ndim == 3
    if len(src.position) == 3 and not is_3d:
        raise ValueError("a 3D GaussianSource position needs a 3D simulation")
    start, prof = gaussian_profile(src, eps_ez.shape, dx, is_3d)
    region = tuple(slice(a, a + n) for a, n in zip(start, prof.shape))
    coeff = -prof * dt / (EPS_0 * eps_ez[region])
    sig = src.signal
    if hasattr(sig, "sample") or callable(sig):
        wave = sample_waveform(sig, t0, dt, total_steps, half_step=True)
    else:
        times = float(t0) + (np.arange(int(total_steps)) + 0.5) * dt
        wave = np.array([interpolate_signal(sig, t, dt) for t in times])
    return CustomSource("Ez", start, coeff, wave, "pre_e")


def lower_soft_source(src: SoftSource, eps_site, dt, t0, total_steps):
    region = tuple(slice(a, a + n) for a, n in zip(src.start, src.shape))
    eps = np.asarray(eps_site)[region]
    prof = np.ones(src.shape) if src.profile is None else np.asarray(src.profile, dtype=np.float64)
    coeff = -prof * dt / (EPS_0 * eps)
    wave = sample_waveform(src.signal, t0, dt, total_steps, half_step=True)
    return CustomSource(src.component, tuple(src.start), coeff, wave, "pre_e")


def group_sources(customs, is_3d):
    """Stack sources into the nine (timing, component) groups
    (``execute.py:356-371``, ``runtime.py:443-468``).  Slabs of one group are
    zero-padded to a common shape; a padded cell outside the field is skipped
    by the kernel, so padding never moves a slab.  Returns a list of nine
    ``None`` or ``(coeff[n, sz, sy, sx], wave[n, T], starts[n, 3], coincident)``."""
    groups = []
    for timing in TIMINGS:
        for comp in PHASE_COMPONENTS[timing]:
            members = [s for s in customs if s.timing == timing and s.component == comp]
            if not members:
                groups.append(None)
                continue
            shapes = []
            starts = []
            for s in members:
                c = np.asarray(s.coeff, dtype=np.float64)
                st = tuple(int(v) for v in s.start)
                if not is_3d:
                    c = c.reshape((1,) + c.shape[-2:])
                    st = (0,) + st[-2:]
                shapes.append(c.shape)
                starts.append(st)
            mx = tuple(max(sh[d] for sh in shapes) for d in range(3))
            T = max(np.asarray(s.waveform).size for s in members)
            coeff = np.zeros((len(members),) + mx)
            wave = np.zeros((len(members), T))
            for i, s in enumerate(members):
                c = np.asarray(s.coeff, dtype=np.float64).reshape(shapes[i])
                coeff[i, :shapes[i][0], :shapes[i][1], :shapes[i][2]] = c
                w = np.asarray(s.waveform, dtype=np.float64)
                wave[i, :w.size] = w
                wave[i, w.size:] = w[-1] if w.size else 0.0
            coincident = all(st == starts[0] for st in starts)
            groups.append((coeff, wave, np.asarray(starts, dtype=np.int64), coincident))
    return groups
