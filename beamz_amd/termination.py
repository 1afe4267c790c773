"""Automatic termination: run a simulation in fixed chunks and stop once the
field energy has decayed and the DFT monitors have settled, after the sources
are off.

Restates the reference's chunked convergence loop and its policy:
``beamz/simulation/model.py:336-425`` (``AutoTermination``),
``beamz/simulation/execute.py:1312-1484`` (``run_until_terminated``), with the
diagnostics of ``execute.py:77-125`` (control-volume integration weights),
``127-212`` (integrated energy 1/2 sum eps0 eps E^2 + mu0 H^2, max |field|,
finiteness), ``215-251`` (remaining source activity), ``274-321`` (relative
monitor change) and ``324-356`` (DFT weight completion on convergence).

The chunks run the ordinary compiled program (fused step, sources, DFT); the
diagnostics between chunks are reductions over the device fields.
"""

from __future__ import annotations

import time as _time
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import runtime
from .constants import EPS_0, MU_0
from .monitors import COMPONENTS


@dataclass(frozen=True)
class AutoTermination:
    """Policy of chunked convergence checks (``model.py:336-425``).

    field_decay: largest integrated energy relative to its peak (0 omits it);
    monitor_change: largest relative change of the DFT monitor values between
    checks (None omits it); source_decay: largest remaining source amplitude
    relative to its peak before checks count; chunk_steps: steps per chunk;
    min_steps: first step at which a check may succeed; consecutive_checks:
    successful checks needed; monitor_names: monitors to inspect (empty: all);
    growth_factor, growth_checks: post-source energy growth counted as
    divergence."""

    field_decay: float = 1e-5
    monitor_change: Optional[float] = 1e-3
    source_decay: float = 1e-6
    chunk_steps: int = 256
    min_steps: int = 0
    consecutive_checks: int = 3
    monitor_names: tuple = ()
    growth_factor: float = 1.05
    growth_checks: int = 3

    def __post_init__(self):
        for name in ("field_decay", "source_decay"):
            v = float(getattr(self, name))
            if not np.isfinite(v) or v < 0.0:
                raise ValueError(f"AutoTermination.{name} must be finite and non-negative.")
        if self.monitor_change is not None and (not np.isfinite(float(self.monitor_change))
                                                or float(self.monitor_change) < 0.0):
            raise ValueError("AutoTermination.monitor_change must be finite, non-negative, or None.")
        for name, low in (("chunk_steps", 1), ("consecutive_checks", 1), ("growth_checks", 1),
                          ("min_steps", 0)):
            v = getattr(self, name)
            if int(v) != v or int(v) < low:
                raise ValueError(f"AutoTermination.{name} must be an integer >= {low}.")
        names = tuple(self.monitor_names)
        if len(set(names)) != len(names) or any(not n for n in names):
            raise ValueError("AutoTermination.monitor_names must contain unique non-empty names.")
        if not np.isfinite(float(self.growth_factor)) or float(self.growth_factor) <= 1.0:
            raise ValueError("AutoTermination.growth_factor must be finite and greater than 1.")


@dataclass(frozen=True)
class RunTermination:
    """Why and where the run stopped (``execute.py:1462-1474``)."""

    reason: str            # "converged", "time_limit", "diverged" or "nonfinite"
    steps: int
    time: float
    converged: bool
    field_decay: Optional[float]
    monitor_change: Optional[float]
    source_decay: float
    energy: float
    peak_energy: float
    max_field: float
    consecutive_checks: int


def axis_weights(count, centered, dx):
    """Control-volume widths on a uniform axis (``execute.py:77-93``): a
    centre-aligned support weighs dx per sample, an edge-aligned one dx with
    half weights at the two ends."""
    if centered:
        return np.full(count, dx)
    w = np.full(count, dx)
    w[0] = w[-1] = 0.5 * dx
    return w


def energy_terms(prog):
    """(field name, material, integration weight, constant) per component
    (``execute.py:96-181``), on the device; non-magnetic media (mu = 1)."""
    dev = prog.device
    terms = []
    for comp in COMPONENTS:
        if comp not in prog.shapes:
            continue
        shape = tuple(prog.shapes[comp])
        grid = tuple(prog.grid)
        axes = []
        for n, g in zip(shape, grid):
            axes.append(axis_weights(n, n == g, prog.resolution))
        w = axes[0]
        for a in axes[1:]:
            w = np.multiply.outer(w, a)
        if comp.startswith("E"):
            material, const = np.asarray(prog.eps_site[comp], dtype=np.float64), EPS_0
        else:
            material, const = np.ones(shape), MU_0
        terms.append((comp.lower(), torch.as_tensor(material * w, dtype=torch.float64, device=dev),
                      const))
    return terms


def field_diagnostics(state, prog, terms):
    """Integrated energy 1/2 sum const * material * v^2 * weight, the largest
    |field| and whether every field is finite (``execute.py:184-212``); values
    in physical units (half-precision storage is rescaled)."""
    scale = prog.field_scale or {"E": 1.0, "H": 1.0}
    energy = torch.zeros((), dtype=torch.float64, device=prog.device)
    peak = torch.zeros((), dtype=torch.float64, device=prog.device)
    finite = torch.ones((), dtype=torch.bool, device=prog.device)
    for name, mw, const in terms:
        v = getattr(state, name).double() * scale[name[0].upper()]
        finite &= torch.isfinite(v).all()
        peak = torch.maximum(peak, v.abs().max())
        energy += 0.5 * const * (mw * v * v).sum()
    return float(energy), float(peak), bool(finite)


def remaining_source_activity(prog, total_steps):
    """Largest normalized source drive at or after each step
    (``execute.py:215-243``)."""
    activity = np.zeros(int(total_steps))
    for g in prog.source_groups:
        if g is None:
            continue
        waves = np.abs(g[1].double().cpu().numpy())
        for wave in waves.reshape(waves.shape[0], -1):
            peak = float(wave.max(initial=0.0))
            if peak <= 0.0:
                continue
            sampled = np.full(activity.size, wave[-1] if wave.size else 0.0)
            n = min(activity.size, wave.size)
            sampled[:n] = wave[:n]
            activity = np.maximum(activity, sampled / peak)
    return np.maximum.accumulate(activity[::-1])[::-1] if activity.size else activity


def monitor_vectors(state, prog, names):
    """One raw DFT vector per (monitor, component, frequency)
    (``execute.py:274-289``)."""
    re = state.dft_re.double().cpu().numpy()
    im = state.dft_im.double().cpu().numpy()
    out = {}
    for m in prog.monitors:
        if m.name not in names:
            continue
        F, P = m.freqs.size, m.points
        for c, comp in enumerate(COMPONENTS):
            if m.plans.get(comp) is None:
                continue
            sl = slice(m.value_offset + c * F * P, m.value_offset + (c + 1) * F * P)
            v = (re[sl] + 1j * im[sl]).reshape(F, P)
            for f in range(F):
                out[(m.name, comp, f)] = v[f]
    return out


def relative_monitor_change(current, previous):
    """Largest |current - previous| / max(|current|, |previous|) over the
    vectors (``execute.py:292-321``); None on the first check."""
    if previous is None:
        return None
    if current.keys() != previous.keys():
        raise ValueError("Automatic termination monitor values changed structure.")
    tiny = np.finfo(float).tiny
    changes = []
    for key, cur in current.items():
        prev = previous[key]
        num = float(np.linalg.norm(cur - prev))
        den = max(float(np.linalg.norm(cur)), float(np.linalg.norm(prev)))
        changes.append(0.0 if den <= tiny and num <= tiny else num / den)
    return max(changes, default=0.0)


def configured_weight_sum(m, t0, dt, total_steps):
    """A monitor's window weight over the whole configured run
    (``execute.py:324-336``)."""
    steps = np.arange(int(total_steps), dtype=np.int64)
    times = float(t0) + (steps.astype(np.float64) + 1.0) * float(dt)
    sel = (steps % max(1, int(m.interval)) == 0) & (times >= m.t_start) & (times <= m.t_end)
    if not np.any(sel):
        return 0.0
    if int(m.window_code) != 1 or not np.isfinite(m.t_end):
        return float(np.count_nonzero(sel))
    span = max(float(m.t_end) - float(m.t_start), 1e-30)
    tau = np.clip((times[sel] - float(m.t_start)) / span, 0.0, 1.0)
    return float(np.sum(0.5 * (1.0 - np.cos(2.0 * np.pi * tau))))


def run_until_terminated(sim, policy: AutoTermination, state=None):
    """Run ``sim`` in chunks of ``policy.chunk_steps`` until it converges,
    diverges, turns non-finite or reaches its configured step count.
    Returns (state, RunTermination, runtime seconds)."""
    if not isinstance(policy, AutoTermination):
        raise TypeError("termination must be an AutoTermination instance or None.")
    prog = sim.compile()
    total = int(sim.steps)
    available = {m.name for m in prog.monitors if m.freqs.size > 0 and m.points > 0}
    if policy.monitor_names:
        missing = set(policy.monitor_names) - available
        if missing:
            raise ValueError("Automatic termination monitors must be applicable "
                             f"frequency-domain monitors: {', '.join(sorted(missing))}.")
        names = tuple(policy.monitor_names)
    else:
        names = tuple(m.name for m in prog.monitors if m.name in available)
    tol = None if policy.monitor_change is None else float(policy.monitor_change)
    use_monitor = tol is not None and bool(names)
    if policy.field_decay == 0.0 and not use_monitor:
        raise ValueError("Automatic termination has no applicable field or monitor criterion.")

    activity = remaining_source_activity(prog, total)
    terms = energy_terms(prog)
    state = sim.initial_state() if state is None else state
    tiny = np.finfo(float).tiny
    prev_energy = prev_monitor = None
    energy = peak_energy = max_field = 0.0
    field_decay = monitor_change = None
    source_decay = float(activity[0]) if activity.size else 0.0
    successes = growth = 0
    reason = "time_limit"
    runtime_s = 0.0
    chunk = min(int(policy.chunk_steps), total)
    while int(state.current_step.item()) < total:
        step = int(state.current_step.item())
        n = min(chunk, total - step)
        torch.cuda.synchronize()
        t0 = _time.perf_counter()
        state = runtime.advance(state, prog, n, chunk=n)
        torch.cuda.synchronize()
        runtime_s += _time.perf_counter() - t0
        step = int(state.current_step.item())

        energy, max_field, finite = field_diagnostics(state, prog, terms)
        current = monitor_vectors(state, prog, names)
        if not finite or not np.isfinite(energy) or not all(np.isfinite(v).all()
                                                            for v in current.values()):
            reason = "nonfinite"
            break
        peak_energy = max(peak_energy, energy)
        field_decay = energy / peak_energy if peak_energy > tiny else 0.0
        source_decay = float(activity[step]) if step < activity.size else 0.0
        source_off = source_decay <= policy.source_decay
        monitor_change = relative_monitor_change(current, prev_monitor)
        if (source_off and prev_energy is not None and prev_energy > tiny
                and energy > policy.growth_factor * prev_energy):
            growth += 1
        else:
            growth = 0
        if growth >= policy.growth_checks:
            reason = "diverged"
            break
        eligible = source_off and step >= policy.min_steps
        energy_ok = policy.field_decay == 0.0 or field_decay <= policy.field_decay
        monitor_ok = not use_monitor or (monitor_change is not None and monitor_change <= tol)
        successes = successes + 1 if (eligible and energy_ok and monitor_ok) else 0
        if successes >= policy.consecutive_checks:
            reason = "converged"
            break
        prev_energy, prev_monitor = energy, current

    if reason == "converged":
        # normalize the DFTs as though the negligible tail had been sampled
        w = state.dft_w
        for m in prog.monitors:
            F = m.freqs.size
            w[m.weight_offset:m.weight_offset + F] = configured_weight_sum(
                m, prog.t0, prog.dt, total)
    report = RunTermination(
        reason=reason, steps=int(state.current_step.item()), time=float(state.t.item()),
        converged=reason == "converged", field_decay=field_decay,
        monitor_change=monitor_change if use_monitor else None, source_decay=source_decay,
        energy=energy, peak_energy=peak_energy, max_field=max_field,
        consecutive_checks=successes)
    return state, report, runtime_s
