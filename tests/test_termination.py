"""Automatic-termination host logic: policy validation, the integration
weights and energy against closed forms, source activity, monitor change and
the configured DFT weight sums (CPU tensors)."""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest
import torch

from beamz_amd.constants import EPS_0, MU_0
from beamz_amd.termination import (AutoTermination, axis_weights, configured_weight_sum,
                                   energy_terms, field_diagnostics, relative_monitor_change,
                                   remaining_source_activity)


@pytest.mark.parametrize("kw", [dict(field_decay=-1.0), dict(monitor_change=float("nan")),
                                dict(chunk_steps=0), dict(consecutive_checks=1.5),
                                dict(growth_factor=1.0), dict(monitor_names=("a", "a")),
                                dict(min_steps=-1)])
def test_policy_rejects_invalid_settings(kw):
    with pytest.raises(ValueError):
        AutoTermination(**kw)


def test_axis_weights_cover_the_domain():
    dx, n = 0.25, 8
    edge = axis_weights(n + 1, False, dx)
    centre = axis_weights(n, True, dx)
    assert edge[0] == edge[-1] == 0.5 * dx and np.all(edge[1:-1] == dx)
    assert np.isclose(edge.sum(), n * dx) and np.isclose(centre.sum(), n * dx)


def _fake_3d(grid=(4, 5, 6), dx=0.5, eps=3.0):
    nz, ny, nx = grid
    shapes = {"Ex": (nz + 1, ny + 1, nx), "Ey": (nz + 1, ny, nx + 1), "Ez": (nz, ny + 1, nx + 1),
              "Hx": (nz, ny, nx + 1), "Hy": (nz, ny + 1, nx), "Hz": (nz + 1, ny, nx)}
    prog = SimpleNamespace(shapes=shapes, grid=grid, resolution=dx, device="cpu", field_scale=None,
                           eps_site={c: np.full(shapes[c], eps) for c in ("Ex", "Ey", "Ez")})
    return prog, shapes


def test_energy_of_uniform_fields_is_the_closed_form():
    """Uniform E and H of unit amplitude: each component integrates to its
    constant times the domain volume, so energy = 1/2 (3 eps0 eps + 3 mu0) V."""
    prog, shapes = _fake_3d()
    state = SimpleNamespace(**{c.lower(): torch.ones(s) for c, s in shapes.items()})
    energy, peak, finite = field_diagnostics(state, prog, energy_terms(prog))
    vol = np.prod(prog.grid) * prog.resolution ** 3
    assert np.isclose(energy, 0.5 * (3 * EPS_0 * 3.0 + 3 * MU_0) * vol, rtol=1e-12)
    assert peak == 1.0 and finite
    getattr(state, "ez")[0, 0, 0] = float("inf")
    assert not field_diagnostics(state, prog, energy_terms(prog))[2]


def test_remaining_source_activity_is_the_suffix_maximum():
    w = torch.tensor([[0.0, 2.0, -4.0, 1.0, 0.5]])
    prog = SimpleNamespace(source_groups=[None, (None, w, None)] + [None] * 7)
    act = remaining_source_activity(prog, 7)
    np.testing.assert_allclose(act, [1.0, 1.0, 1.0, 0.25, 0.125, 0.125, 0.125])


def test_relative_monitor_change():
    a = {("m", "Ez", 0): np.array([1.0, 0.0]), ("m", "Ez", 1): np.array([0.0, 0.0])}
    b = {("m", "Ez", 0): np.array([1.0, 0.1]), ("m", "Ez", 1): np.array([0.0, 0.0])}
    assert relative_monitor_change(a, None) is None
    assert np.isclose(relative_monitor_change(b, a), 0.1 / np.sqrt(1.01))
    with pytest.raises(ValueError):
        relative_monitor_change({("m", "Ez", 0): a[("m", "Ez", 0)]}, a)


def test_configured_weight_sum():
    rect = SimpleNamespace(interval=2, t_start=0.0, t_end=float("inf"), window_code=0)
    assert configured_weight_sum(rect, 0.0, 1.0, 10) == 5.0
    hann = SimpleNamespace(interval=1, t_start=0.0, t_end=10.0, window_code=1)
    t = np.arange(1, 11, dtype=float)
    assert np.isclose(configured_weight_sum(hann, 0.0, 1.0, 10),
                      np.sum(0.5 * (1 - np.cos(2 * np.pi * t / 10.0))))
