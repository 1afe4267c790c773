"""Modal S-parameter extraction (host): the projection recovers planted
forward/backward amplitudes, the mode carries unit power, and S-parameters
and power sums follow from synthetic port fields."""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest

from beamz_amd.constants import LIGHT_SPEED
from beamz_amd.sparams import Port, port_projection, s_parameters

DX = 20e-9
F0 = LIGHT_SPEED / 1.55e-6


def _line(n=61, core=(24, 37)):
    eps = np.full(n, 2.085)
    eps[core[0]:core[1]] = 12.25
    return eps


@pytest.mark.parametrize("axis", ["x", "y"])
def test_projection_recovers_planted_amplitudes(axis):
    pts = np.arange(10, 50)
    proj = port_projection(_line(), F0, DX, pts, axis)
    e, h = proj["e"], proj["h"]
    sign = -1.0 if axis == "x" else 1.0
    assert np.isclose(0.5 * np.real(np.sum(DX * sign * e * np.conj(h))), 1.0, rtol=1e-12)
    fwd, bwd = 0.7 - 0.2j, 0.05 + 0.3j
    vec = np.concatenate([fwd * e + bwd * e, fwd * h - bwd * h])
    a = proj["pinv"] @ vec
    np.testing.assert_allclose(a, [fwd, bwd], rtol=1e-10, atol=1e-12)


def _fake_results(fields_by_monitor, freqs, eps_ez):
    monitors = [SimpleNamespace(name=n, normal="x", freqs=freqs, index=i, cell_range=(10, 50))
                for n, (i, _) in fields_by_monitor.items()]
    prog = SimpleNamespace(monitors=monitors, eps_site={"Ez": eps_ez}, resolution=DX)
    order = list(fields_by_monitor)

    def phasors(i):
        return fields_by_monitor[order[i]][1]

    return SimpleNamespace(program=prog, monitor_phasors=phasors)


def test_s_parameters_from_synthetic_port_fields():
    freqs = np.array([0.98 * F0, F0])
    line = _line()
    eps_ez = np.repeat(line[:, None], 200, axis=1)
    pts = np.arange(10, 50)
    fields = {}
    for name, idx, fwd, bwd in (("m_in", 60, 1.0, 0.1), ("m_out", 140, 0.8, 0.0)):
        ez, hy = [], []
        for f in freqs:
            p = port_projection(line, f, DX, pts, "x")
            ez.append(fwd * p["e"] + bwd * p["e"])
            hy.append(fwd * p["h"] - bwd * p["h"])
        fields[name] = (idx, {"Ez": np.array(ez), "Hy": np.array(hy)})
    res = _fake_results(fields, freqs, eps_ez)
    ports = [Port("in", "m_in", "x", "+"), Port("out", "m_out", "x", "-")]
    out = s_parameters(res, "in", ports)
    np.testing.assert_allclose(out.s_matrix[("in", "in")], 0.1, atol=1e-10)
    np.testing.assert_allclose(out.s_matrix[("out", "in")], 0.8, atol=1e-10)
    np.testing.assert_allclose(out.diagnostics["power_sum"], 0.65, atol=1e-9)
    with pytest.raises(ValueError):
        s_parameters(res, "missing", ports)
