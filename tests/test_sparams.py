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


@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_3d_projection_recovers_planted_amplitudes(axis):
    """A 3D port on a strip guide along ``axis``: the unit-power plane mode
    averaged onto the monitor's cell centres, and planted forward/backward
    amplitudes recovered from its fields."""
    from beamz_amd.lattice import sample_e_site_3d
    from beamz_amd.sparams import port_projection_3d
    dx = 40e-9
    n = {"z": 20, "y": 24, "x": 28}
    grid = (n["z"], n["y"], n["x"])
    eps = np.full(grid, 2.085)
    core = {"x": slice(8, 20), "y": slice(6, 18), "z": slice(6, 14)}
    sl = [core["z"], core["y"], core["x"]]
    sl[{"z": 0, "y": 1, "x": 2}[axis]] = slice(None)
    eps[tuple(sl)] = 12.25
    site = {c: sample_e_site_3d(eps, c) for c in ("Ex", "Ey", "Ez")}
    tang = {"x": ("z", "y"), "y": ("z", "x"), "z": ("y", "x")}[axis]
    ranges = tuple((a, (2, n[a] - 2)) for a in tang)
    proj = port_projection_3d(site, grid, dx, F0, axis, n[axis] // 2, ranges)
    e, h = proj["e"], proj["h"]
    P = (n[tang[0]] - 4) * (n[tang[1]] - 4)
    assert e.size == 2 * P and h.size == 2 * P
    # unit power towards +axis: 1/2 sum(E_b H_c - E_c H_b) dA
    assert 0.5 * np.real(np.sum(e[:P] * np.conj(h[P:])) - np.sum(e[P:] * np.conj(h[:P]))) * dx * dx \
        == pytest.approx(1.0, rel=1e-12)
    fwd, bwd = 0.6 + 0.3j, -0.1 + 0.05j
    vec = np.concatenate([fwd * e + bwd * e, fwd * h - bwd * h])
    np.testing.assert_allclose(proj["pinv"] @ vec, [fwd, bwd], rtol=1e-10, atol=1e-12)
    assert np.sqrt(2.085) < proj["mode_neff"] < 3.5   # guided
