"""Source normalization of monitor outputs (host post-processing).

Restates ``beamz/simulation/observe.py``:
- ``SourceNormalization`` (``:24-70``): the waveform spectrum S(f), an optional
  launched-power ratio P(f), the field-amplitude norm S sqrt(P) and the power
  norm |S|^2 P;
- ``_signal_dft_spectrum`` (``:88-98``): S(f) = (2/n) sum_i s_i exp(+i 2 pi f t_i);
- ``_sampled_source_spectrum`` (``:101-123``): the source's waveform sampled on
  the simulation time grid t_i = t0 + i dt (not the half-step injection
  times), or its own samples for an explicit waveform array;
- ``source_normalization`` (``:126-147``): the first source (or the chosen
  one) with a spectrum above 1e-12 at some frequency;
- ``_scale_frequency_axis`` (``:257-268``): division along the frequency axis
  where |divisor| > 1e-24, zero elsewhere.

Every source type the reference defines leaves the launched-power ratio unset
(``sources/specs.py:200-203``), so the norms are S and |S|^2.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from .modes import ModeSource
from .sources import CustomSource, GaussianSource, SoftSource


@dataclass(frozen=True)
class SourceNormalization:
    waveform_spectrum: np.ndarray
    launch_power_ratio: Optional[np.ndarray] = None

    def _valid_power(self):
        spectrum = np.asarray(self.waveform_spectrum, dtype=np.complex128).reshape(-1)
        if self.launch_power_ratio is None:
            return spectrum, None
        power = np.asarray(self.launch_power_ratio, dtype=float).reshape(-1)
        if power.shape != spectrum.shape:
            return spectrum, None
        return spectrum, power

    @property
    def field_amplitude_norm(self) -> np.ndarray:
        spectrum, power = self._valid_power()
        if power is None:
            return spectrum
        scale = np.ones_like(power)
        valid = np.isfinite(power) & (power > 1e-24)
        scale[valid] = np.sqrt(power[valid])
        return spectrum * scale

    @property
    def power_norm(self) -> np.ndarray:
        spectrum, power = self._valid_power()
        scale = np.abs(spectrum) ** 2
        if power is not None:
            valid = np.isfinite(power) & (power > 1e-24)
            scale[valid] *= power[valid]
        return scale


def signal_dft_spectrum(signal, sample_times, freqs):
    """S(f) = (2/n) sum_i s_i exp(+i 2 pi f t_i); None when all |S| <= 1e-12."""
    freqs = np.asarray(freqs, dtype=float).reshape(-1)
    times = np.asarray(sample_times, dtype=float).reshape(-1)
    signal = np.asarray(signal, dtype=float).reshape(-1)
    n = min(times.size, signal.size)
    if freqs.size == 0 or n == 0:
        return None
    times, signal = times[:n], signal[:n]
    phase = np.exp(1j * 2.0 * np.pi * times[:, None] * freqs[None, :])
    spectrum = (2.0 / float(n)) * np.sum(signal[:, None] * phase, axis=0)
    return spectrum if np.any(np.abs(spectrum) > 1e-12) else None


def source_signal(source, t0, dt, n):
    """The source's waveform on the time grid t0 + i dt, i < n."""
    times = float(t0) + np.arange(int(n), dtype=np.float64) * float(dt)
    if isinstance(source, CustomSource):
        return np.asarray(source.waveform, dtype=np.float64).reshape(-1)[:n]
    signal = source.source_time if isinstance(source, ModeSource) else (
        source.signal if isinstance(source, (SoftSource, GaussianSource)) else None)
    if signal is None:
        return None
    if hasattr(signal, "sample"):
        return np.asarray(signal.sample(times)[0], dtype=np.float64)
    if callable(signal):
        return np.asarray([float(signal(float(t))) for t in times], dtype=np.float64)
    return np.asarray(signal, dtype=np.float64).reshape(-1)[:n]


def source_normalization(sources, freqs, t0, dt, n, source=None):
    """Normalization from ``sources[source]``, or from the first source with
    a spectrum (``observe.py:126-147``); None when none has one."""
    freqs = np.asarray(freqs, dtype=float).reshape(-1)
    if freqs.size == 0:
        return None
    if source is not None:
        source = int(source)
        if source < 0 or source >= len(sources):
            raise ValueError(f"normalization source {source} is invalid for {len(sources)} sources.")
        candidates = (sources[source],)
    else:
        candidates = tuple(sources)
    times = float(t0) + np.arange(int(n), dtype=np.float64) * float(dt)
    for src in candidates:
        signal = source_signal(src, t0, dt, n)
        if signal is None:
            continue
        spectrum = signal_dft_spectrum(signal, times[:signal.size], freqs)
        if spectrum is None:
            continue
        if spectrum.shape == freqs.shape and np.any(np.abs(spectrum) > 1e-12):
            return SourceNormalization(spectrum)
    return None


def scale_frequency_axis(values, divisor):
    """``values / divisor`` along axis 0 where |divisor| > 1e-24, else 0."""
    array = np.asarray(values)
    divisor = np.asarray(divisor).reshape(-1)
    if not array.size or array.shape[0] != divisor.size:
        return array
    divisor = divisor.reshape((-1,) + (1,) * (array.ndim - 1))
    out = np.zeros_like(array, dtype=np.result_type(array, divisor))
    return np.divide(array, divisor, out=out, where=np.abs(divisor) > 1e-24)
