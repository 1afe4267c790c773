"""2D mode sources: the transverse solver against closed forms, the snapping
rules, and the launched pulse's direction on the FP64 oracle.

The reference's own mode solver needs ``xarray`` and a package import that
pulls in ``jax``; neither is installed, so its profiles are not reproduced
here.  The solver is pinned by the exact dispersion relation of its discrete
operator and by convergence to the analytic slab-waveguide relation; the
Huygens pair (entry signs, companion plane, power pairing) by the direction
and one-sidedness of the pulse it launches."""

from __future__ import annotations

import numpy as np
import pytest
from scipy.optimize import brentq

import beamz_amd as bz
from beamz_amd.modes import (ModeSource, lower_mode_source_2d, plan_mode_source_2d,
                             snap_cell_center, snap_centered_extent, solve_line_modes)
from oracle import fdtd64 as o

F0 = bz.LIGHT_SPEED / 1.55e-6
OMEGA = 2.0 * np.pi * F0
K0 = OMEGA / bz.LIGHT_SPEED
E_CORE, E_CLAD = 12.25, 2.085


def _lattice_slab(M, dx, margin=1.5e-6):
    n = 2 * M + 1 + 2 * int(margin / dx)
    c = n // 2
    eps = np.full(n, E_CLAD)
    eps[c - M:c + M + 1] = E_CORE
    return eps


def test_solver_matches_discrete_slab_dispersion():
    """Even mode of 2M+1 core nodes: Ez_j = cos(q (j - c)) in the core and
    cos(qM) r^(j-c-M) outside, so the node equations close on
    r = cos(q(M+1)) / cos(qM) and r + 1/r + a2 = 2 cos q + a1
    (a = dx^2 k0^2 eps, beta^2 dx^2 = 2 cos q + a1 - 2)."""
    M, dx = 6, 20e-9
    a1, a2 = dx ** 2 * K0 ** 2 * E_CORE, dx ** 2 * K0 ** 2 * E_CLAD

    def f(q):
        r = np.cos(q * (M + 1)) / np.cos(q * M)
        return r + 1.0 / r + a2 - 2.0 * np.cos(q) - a1

    qs = np.linspace(1e-6, np.pi / (2 * (M + 1)) - 1e-9, 4000)
    v = np.array([f(q) for q in qs])
    i = int(np.nonzero(np.sign(v[:-1]) != np.sign(v[1:]))[0][0])
    q = brentq(f, qs[i], qs[i + 1], xtol=1e-15)
    want = np.sqrt(2.0 * np.cos(q) + a1 - 2.0) / (dx * K0)
    neff, ez = solve_line_modes(_lattice_slab(M, dx), OMEGA, dx, 3, 0.98 * np.sqrt(E_CORE))
    assert abs(neff[0] - want) <= 1e-10 * want, (neff[0], want)
    # the profile is the cosine inside the core
    c = ez.shape[1] // 2
    prof = ez[0] / ez[0][c]
    np.testing.assert_allclose(prof[c - M:c + M + 1], np.cos(q * np.arange(-M, M + 1)), atol=1e-9)


def test_solver_converges_to_the_slab_waveguide():
    """TE0 of a 260 nm slab: x tan x = sqrt(V^2 - x^2), V = k0 d/2 sqrt(n1^2 - n2^2);
    the lattice index approaches it at second order in dx."""
    d = 260e-9
    V = K0 * d / 2.0 * np.sqrt(E_CORE - E_CLAD)
    x = brentq(lambda x: x * np.tan(x) - np.sqrt(V * V - x * x), 1e-9, min(np.pi / 2 - 1e-12, V))
    n_slab = np.sqrt(E_CORE - (2.0 * x / (K0 * d)) ** 2)
    errs = []
    for M in (12, 24, 48):
        dx = d / (2 * M + 1)
        neff, _ = solve_line_modes(_lattice_slab(M, dx), OMEGA, dx, 3, 0.98 * np.sqrt(E_CORE))
        errs.append(abs(neff[0] - n_slab))
    assert errs[-1] < 1e-4
    assert errs[0] / errs[1] > 3.5 and errs[1] / errs[2] > 3.5, errs


def test_snapping_rules():
    dx = 20e-9
    assert snap_cell_center(100.5 * dx, dx, 200) == 100
    assert snap_cell_center(100.0 * dx, dx, 200) == 100      # 99.5 rounds half to even
    assert snap_cell_center(-5 * dx, dx, 200) == 0
    assert snap_cell_center(500 * dx, dx, 200) == 199
    assert snap_centered_extent(30 * dx, 40 * dx, dx, 60) == (10, 50)
    assert snap_centered_extent(30.3 * dx, 40 * dx, dx, 60) == (10, 51)
    assert snap_centered_extent(30.5 * dx, 0.0, dx, 60) == (30, 31)


def _guide(axis):
    ny, nx = (60, 200) if axis == "x" else (200, 60)
    eps = np.full((ny, nx), E_CLAD)
    if axis == "x":
        eps[24:36, :] = E_CORE
    else:
        eps[:, 24:36] = E_CORE
    return eps


def _source(axis, direction, power=1.0, dx=20e-9):
    pulse = bz.GaussianPulse(F0, bz.LIGHT_SPEED * 0.4e-6 / 1.55e-6 ** 2)
    if axis == "x":
        return ModeSource((100.5 * dx, 30 * dx, 0), (0, 40 * dx, 0), pulse, direction, power=power)
    return ModeSource((30 * dx, 100.5 * dx, 0), (40 * dx, 0, 0), pulse, direction, power=power)


def test_power_scales_the_amplitude():
    eps_ez = o.site_sample_tm(_guide("x"), "Ez")
    a = plan_mode_source_2d(_source("x", "+"), eps_ez, 20e-9)
    b = plan_mode_source_2d(_source("x", "+", power=4.0), eps_ez, 20e-9)
    for (_, _, _, pa), (_, _, _, pb) in zip(a.entries, b.entries):
        np.testing.assert_allclose(pb, 2.0 * pa, rtol=1e-12)
    assert a.support == (10, 50)


@pytest.mark.parametrize("axis,direction,leak", [("x", "+", 1e-2), ("x", "-", 5e-2), ("y", "+", 1e-2)])
def test_mode_source_launches_one_way(axis, direction, leak):
    """A 1.55 um pulse launched in a 240 nm guide: the flux at 1.55 um through
    a line 40 cells downstream carries the power, the line 40 cells upstream
    sees at most ``leak`` of it (the "-" launch puts its H plane one cell
    further from the Ez plane, as the reference's companion index does)."""
    dx = 20e-9
    eps = _guide(axis)
    ny, nx = eps.shape
    dt = 0.99 * dx / (o.LIGHT_SPEED * np.sqrt(2))
    steps = 2600
    customs = lower_mode_source_2d(_source(axis, direction), o.site_sample_tm(eps, "Ez"), dx, dt,
                                   0.0, steps)
    assert {(c.component, c.timing) for c in customs} == {("Hy" if axis == "x" else "Hx", "h"),
                                                          ("Ez", "e")}
    srcs = [o.Source(c.component, c.timing, c.start, np.asarray(c.coeff, float),
                     np.asarray(c.waveform, float)) for c in customs]
    shapes = o.shapes_tm((ny, nx))
    h = "Hy" if axis == "x" else "Hx"
    mons = []
    for idx in (60, 140):
        plans = {c: None for c in o.COMPONENTS}
        for c in ("Ez", h):
            plans[c] = o.line_plan_tm(c, shapes[c], dx, axis, idx, (10, 50))
        mons.append(o.Monitor(np.array([F0]), plans, axis, 1.0, dx))
    spec = o.CpmlSpec(cells=10, kappa_max=2.0, faces=frozenset(("bottom", "top", "left", "right")))
    prob = o.build_problem(eps, dx, dt, cpml=spec, sources=srcs, monitors=mons)
    st = o.init_state(prob)
    o.run(prob, st, steps)
    low, high = o.monitor_flux(prob, st, 0)[0], o.monitor_flux(prob, st, 1)[0]
    fwd, back = (high, low) if direction == "+" else (low, high)
    sign = 1.0 if direction == "+" else -1.0
    assert sign * fwd > 0.0, (low, high)
    assert abs(back) <= leak * abs(fwd), (low, high)
