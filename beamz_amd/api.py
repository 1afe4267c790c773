"""``Simulation``: the public call surface of the HIP time-stepping backend.

Mirrors the parts of ``beamz/simulation/api.py`` that drive the hot path:
``compile`` (``api.py:1071``), ``initial_state``, ``advance`` (``api.py:1339-1458``)
and ``run`` (``api.py:1460-1537``), with backend resolution following
``beamz/simulation/backend.py:223-258``.  Geometry is given as a permittivity
raster (the lowering of ``Design`` objects is outside this backend).
"""

from __future__ import annotations

import time as _time
from dataclasses import dataclass

import numpy as np
import torch

from . import _abi as abi
from . import native, runtime
from .constants import LIGHT_SPEED
from .monitors import flux as _flux, phasors as _phasors
from .program import compile_program, initial_state

BACKENDS = ("auto", "hip")


def resolve_backend(name="auto"):
    """``backend.py:223-258``: ``auto`` and ``hip`` both need the HIP library;
    there is no CPU fallback, so an unavailable library raises."""
    name = str(name or "auto").lower()
    if name not in BACKENDS:
        raise ValueError(f"unknown backend {name!r}; expected one of {BACKENDS}")
    ok, why = native.status()
    if not ok:
        raise native.HipBackendUnavailable(why)
    return "hip"


def courant_dt(resolution, ndim, courant=0.99):
    """``dt = S dx / (c sqrt(ndim))`` (``beamz/_helpers.py:76-164``)."""
    return courant * resolution / (LIGHT_SPEED * np.sqrt(ndim))


@dataclass
class SimulationPerformance:
    """``results.py:59-93``: GCUPS = material cells x steps / runtime / 1e9."""

    cells: int
    steps: int
    runtime_s: float

    @property
    def gcups(self):
        return self.cells * self.steps / max(self.runtime_s, 1e-30) / 1e9


@dataclass(frozen=True)
class MonitorResult:
    """One monitor's frequency-domain acquisition, keyed by name on
    ``SimulationResults`` (the DFT part of ``results.py:537-620``, ``830-880``).

    ``dft_fields``: the raw accumulations sum_t w s m cos/sin(2 pi f t) v as
    complex [F, P] arrays per recorded component (what the reference stores);
    ``phasors``: 2 DFT / wsum with the H half-step phase (``observe.py:166-187``);
    ``flux``: the frequency-domain flux (``observe.py:198-234``)."""

    name: str
    dft_fields: dict
    dft_frequencies: np.ndarray
    dft_weight_sum: np.ndarray
    dft_base_dt: float
    resolution: float
    normal_axis: int
    normal_sign: float
    power_scale: float
    phasors: dict
    flux: np.ndarray


@dataclass
class SimulationResults:
    state: object
    program: object
    performance: SimulationPerformance
    termination: object = None   # RunTermination of an automatically terminated run
    sources: tuple = ()          # the simulation's sources, for source normalization
    t0: float = 0.0              # start of the simulation time grid t0 + i dt, i < steps
    steps: int = 0

    def normalization(self, index=0, source=None):
        """Source normalization at monitor ``index``'s frequencies
        (``observe.py:237-247``), or None when no source has a spectrum."""
        from .normalization import source_normalization
        freqs = np.asarray(self.program.monitors[index].freqs, dtype=float)
        return source_normalization(self.sources, freqs, self.t0, self.program.dt,
                                    self.steps, source=source)

    def normalized_flux(self, index=0, source=None):
        """Flux divided by the source's power norm |S(f)|^2
        (``observe.py:271-312``); the raw flux when there is no normalization."""
        from .normalization import scale_frequency_axis
        flux = self.monitor_flux(index)
        norm = self.normalization(index, source)
        return flux if norm is None else scale_frequency_axis(flux, norm.power_norm)

    def normalized_phasors(self, index=0, source=None):
        """Monitor phasors divided by the source's field-amplitude norm S(f)."""
        from .normalization import scale_frequency_axis
        ph = self.monitor_phasors(index)
        norm = self.normalization(index, source)
        if norm is None:
            return ph
        return {c: (None if v is None else scale_frequency_axis(v, norm.field_amplitude_norm))
                for c, v in ph.items()}

    def monitor_flux(self, index=0):
        m = self.program.monitors[index]
        s = self.state
        return _flux(m, s.dft_re.cpu().numpy(), s.dft_im.cpu().numpy(), s.dft_w.cpu().numpy(),
                     self.program.dt)

    def monitor_phasors(self, index=0):
        m = self.program.monitors[index]
        s = self.state
        return _phasors(m, s.dft_re.cpu().numpy(), s.dft_im.cpu().numpy(), s.dft_w.cpu().numpy(),
                        self.program.dt)

    @property
    def monitors(self):
        """``{name: MonitorResult}`` of every monitor, in program order."""
        s = self.state
        re, im, w = (s.dft_re.cpu().numpy(), s.dft_im.cpu().numpy(), s.dft_w.cpu().numpy())
        out = {}
        for m in self.program.monitors:
            F, P = m.freqs.size, m.points
            raw = {}
            for c, name in enumerate(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")):
                if m.plans.get(name) is None:
                    continue
                sl = slice(m.value_offset + c * F * P, m.value_offset + (c + 1) * F * P)
                raw[name] = (np.asarray(re[sl], np.float64) + 1j * np.asarray(im[sl], np.float64)).reshape(F, P)
            out[m.name] = MonitorResult(
                name=m.name, dft_fields=raw, dft_frequencies=np.asarray(m.freqs, float),
                dft_weight_sum=np.asarray(w[m.weight_offset:m.weight_offset + F], np.float64),
                dft_base_dt=float(self.program.dt), resolution=float(self.program.resolution),
                normal_axis={"x": 0, "y": 1, "z": 2}[m.normal], normal_sign=float(m.normal_sign),
                power_scale=float(m.measure), phasors=_phasors(m, re, im, w, self.program.dt),
                flux=_flux(m, re, im, w, self.program.dt))
        return out

    def __getitem__(self, name):
        """The ``MonitorResult`` of the monitor called ``name``."""
        mons = self.monitors
        if name not in mons:
            raise KeyError(f"no monitor named {name!r}; monitors: {sorted(mons)}")
        return mons[name]

    def field(self, name):
        """Field in physical units (half-precision storage is rescaled)."""
        scale = (self.program.field_scale or {}).get(name[0].upper(), 1.0)
        return getattr(self.state, name.lower()).double().cpu().numpy() * scale


class Simulation:
    """A time-domain run on a uniform Yee grid.

    Parameters
    ----------
    permittivity : array (z, y, x) or (y, x)
        Relative permittivity per material cell.
    resolution : float
        Cell size in metres (with ``grid``: defaults to its smallest width).
    grid : RectilinearGrid, optional
        Non-uniform cell widths (physical x, y, z edges); the curls then scale
        every difference by its inverse physical distance.
    dt : float, optional
        Time step; defaults to ``courant / sqrt(ndim)`` of the Courant limit.
    boundaries, sources, monitors : sequences of ``PEC``/``PML``,
        ``SoftSource``/``CustomSource`` and ``FluxMonitor``.
    steps : int
        Number of time steps of ``run``.
    time : array, optional
        The reference's time grid instead of ``dt``/``steps``/``t0``: finite,
        strictly increasing, uniformly spaced (relative 1e-7), at least two
        entries; then dt = time[1] - time[0], steps = len(time), t0 = time[0]
        (``simulation/api.py:474-489``, ``799-816``).
    psi_dtype : torch.float32 or torch.bfloat16, optional
        CPML memory storage (default: ``BEAMZ_HIP_CPML_PSI_PRECISION``, fp32).
    field_dtype : torch.float32, torch.float16 or torch.bfloat16
        Field storage.  Half-precision fields are stored scaled (E / s_e,
        H / s_h with s_h = s_e / eta0); ``field_scale`` sets s_e, by default
        the largest per-step source increment.
    """

    def __init__(self, permittivity, resolution=None, *, dt=None, courant=0.99, boundaries=(),
                 sources=(), monitors=(), steps=None, t0=None, conductivity=None,
                 backend="auto", field_dtype=torch.float32, device=None,
                 flags=abi.DEFAULT_FLAGS | abi.FLAG_FUSED_STEP, field_scale=None, time=None,
                 psi_dtype=None, grid=None):
        if time is not None:
            values = np.array(time, dtype=float)
            if values.ndim != 1 or values.size < 2:
                raise ValueError("FDTD requires a 1D time array with at least two entries")
            gaps = np.diff(values)
            if not np.all(np.isfinite(values)) or np.any(gaps <= 0.0):
                raise ValueError("Simulation time values must be finite and strictly increasing")
            if not np.allclose(gaps, gaps[0], rtol=1e-7, atol=0.0):
                raise ValueError("Simulation time values must be uniformly spaced")
            if dt is not None and not np.isclose(float(dt), gaps[0], rtol=1e-7, atol=0.0):
                raise ValueError("dt disagrees with the time grid")
            if steps is not None and int(steps) != values.size:
                raise ValueError(f"steps={steps} disagrees with the time grid ({values.size} samples)")
            if t0 is not None and not np.isclose(float(t0), values[0], rtol=1e-12,
                                                 atol=1e-12 * abs(gaps[0])):
                raise ValueError(f"t0={t0} disagrees with the time grid (starts at {values[0]})")
            dt, steps, t0 = float(gaps[0]), int(values.size), float(values[0])
        steps = 1 if steps is None else steps
        t0 = 0.0 if t0 is None else t0
        self.permittivity = np.asarray(permittivity, dtype=np.float64)
        self.grid = grid
        if grid is not None:
            # a realized grid: its smallest width is the resolution and its
            # CFL limit the default step (simulation/api.py:210-217,
            # design/grid.py:358-371)
            axes = ("x", "y", "z")[:self.permittivity.ndim]
            if resolution is None:
                resolution = grid.minimum_spacing(axes)
            if dt is None:
                dt = grid.cfl_time_step(courant, axes=axes)
        if resolution is None:
            raise ValueError("Simulation needs a resolution or a grid")
        self.resolution = float(resolution)
        self.dt = float(dt) if dt is not None else courant_dt(resolution, self.permittivity.ndim, courant)
        self.boundaries = tuple(boundaries)
        self.sources = tuple(sources)
        self.monitors = tuple(monitors)
        self.steps = int(steps)
        self.t0 = float(t0)
        self.conductivity = conductivity
        self.backend = backend
        self.field_dtype = field_dtype
        self.device = device
        self.flags = flags
        self.field_scale = field_scale
        self.psi_dtype = psi_dtype
        self._program = None

    def compile(self):
        resolve_backend(self.backend)
        if self._program is None:
            self._program = compile_program(
                self.permittivity, self.resolution, self.dt, boundaries=self.boundaries,
                sources=self.sources, monitors=self.monitors, num_steps=self.steps,
                t0=self.t0, conductivity=self.conductivity, device=self.device,
                field_dtype=self.field_dtype, flags=self.flags, field_scale=self.field_scale,
                psi_dtype=self.psi_dtype, grid=self.grid)
        return self._program

    def initial_state(self, fields=None):
        return initial_state(self.compile(), fields)

    def advance(self, state=None, steps=None, *, chunk=runtime.GRAPH_MAX_STEPS):
        prog = self.compile()
        state = self.initial_state() if state is None else state
        return runtime.advance(state, prog, self.steps if steps is None else int(steps), chunk=chunk)

    def run(self, state=None, *, chunk=runtime.GRAPH_MAX_STEPS, termination=None):
        """Run the configured steps; with an ``AutoTermination`` policy, run in
        its chunks and stop early once converged (``execute.py:1312-1484``)."""
        prog = self.compile()
        if termination is not None:
            from .termination import run_until_terminated
            state, report, runtime_s = run_until_terminated(self, termination, state)
            steps = report.steps
            perf = SimulationPerformance(prog.cells, steps, runtime_s)
            return SimulationResults(state, prog, perf, termination=report,
                                     sources=tuple(self.sources), t0=self.t0, steps=self.steps)
        state = self.initial_state() if state is None else state
        torch.cuda.synchronize()
        t0 = _time.perf_counter()
        state = runtime.advance(state, prog, self.steps, chunk=chunk)
        torch.cuda.synchronize()
        perf = SimulationPerformance(prog.cells, self.steps, _time.perf_counter() - t0)
        return SimulationResults(state, prog, perf, sources=tuple(self.sources), t0=self.t0,
                                 steps=self.steps)
