"""Source normalization (host): the waveform spectrum against closed forms
(a cosine of whole periods has S(f0) = 1, a sine S(f0) = i, which fixes the
sign of the exponent), the norms with a launched-power ratio, source
selection and the division along the frequency axis."""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest

from beamz_amd.normalization import (SourceNormalization, scale_frequency_axis,
                                     signal_dft_spectrum, source_normalization)
from beamz_amd.sources import CustomSource, GaussianPulse, SoftSource

F0 = 2.0e14
N_PER = 40
DT = 1.0 / (F0 * N_PER)


def _times(periods=25):
    return np.arange(periods * N_PER) * DT


def test_cosine_and_sine_spectra_are_the_closed_forms():
    t = _times()
    freqs = np.array([F0, 0.5 * F0])
    cos = signal_dft_spectrum(np.cos(2 * np.pi * F0 * t), t, freqs)
    sin = signal_dft_spectrum(np.sin(2 * np.pi * F0 * t), t, freqs)
    # (2/n) sum cos(th) e^{i th} = 1 + (1/n) sum e^{2 i th} = 1 over whole periods,
    # and (2/n) sum sin(th) e^{i th} = i: a conjugated exponent would give -i
    np.testing.assert_allclose(cos[0], 1.0, atol=1e-12)
    np.testing.assert_allclose(sin[0], 1j, atol=1e-12)
    assert cos.shape == freqs.shape
    assert signal_dft_spectrum(np.zeros(10), t[:10], freqs) is None


def test_norms_with_a_launch_power_ratio():
    s = np.array([2.0 + 1.0j, 0.5j, 1.0])
    p = np.array([4.0, 0.0, np.nan])
    n = SourceNormalization(s, p)
    # sqrt(P) applied where P is finite and positive, 1 elsewhere
    np.testing.assert_allclose(n.field_amplitude_norm, [2.0 * (2.0 + 1.0j), 0.5j, 1.0])
    np.testing.assert_allclose(n.power_norm, [4.0 * 5.0, 0.25, 1.0])
    plain = SourceNormalization(s)
    np.testing.assert_allclose(plain.power_norm, np.abs(s) ** 2)


def test_source_selection_and_sampling_grid():
    freqs = np.array([F0])
    n = 25 * N_PER
    silent = CustomSource("Ez", (0, 0), np.ones((1, 1)), np.zeros(n))
    loud = CustomSource("Ez", (0, 0), np.ones((1, 1)), 3.0 * np.cos(2 * np.pi * F0 * _times()))
    norm = source_normalization([silent, loud], freqs, 0.0, DT, n)
    np.testing.assert_allclose(norm.waveform_spectrum, [3.0], atol=1e-12)
    assert source_normalization([silent], freqs, 0.0, DT, n) is None
    with pytest.raises(ValueError):
        source_normalization([silent], freqs, 0.0, DT, n, source=3)
    # a waveform object is sampled on t0 + i dt (the time grid, not the
    # half-step injection times): shifting t0 by dt turns the phase of S by
    # 2 pi f dt
    pulse = GaussianPulse(F0, 0.3 * F0)
    soft = SoftSource("Ez", (0, 0), (1, 1), pulse)
    a = source_normalization([soft], freqs, 0.0, DT, 4000).waveform_spectrum[0]
    b = source_normalization([soft], freqs, DT, DT, 4000).waveform_spectrum[0]
    t = np.arange(4000) * DT
    want = (2.0 / 4000) * np.sum(pulse.sample(t)[0] * np.exp(2j * np.pi * F0 * t))
    np.testing.assert_allclose(a, want, rtol=1e-12)
    assert abs(b) == pytest.approx(abs(a), rel=1e-3)


def test_scale_frequency_axis():
    v = np.arange(6.0).reshape(3, 2)
    out = scale_frequency_axis(v, np.array([2.0, 0.0, -1.0]))
    np.testing.assert_allclose(out, [[0.0, 0.5], [0.0, 0.0], [-4.0, -5.0]])


def test_results_normalized_flux_divides_by_power_norm():
    from beamz_amd.api import SimulationResults
    n = 25 * N_PER
    src = CustomSource("Ez", (0, 0), np.ones((1, 1)), 2.0 * np.cos(2 * np.pi * F0 * _times()))
    mon = SimpleNamespace(freqs=np.array([F0]))
    prog = SimpleNamespace(monitors=[mon], dt=DT)
    res = SimulationResults(state=None, program=prog, performance=None, sources=(src,), t0=0.0,
                            steps=n)
    res.monitor_flux = lambda index=0: np.array([8.0])
    np.testing.assert_allclose(res.normalized_flux(0), [2.0], rtol=1e-12)
