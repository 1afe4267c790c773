"""3D mode sources on the GPU.

- A 1.55 um pulse launched into a 450 x 220 nm silicon strip (25 nm cells,
  10-cell CPML): the flux normalized by the source spectrum |S(f)|^2 through a
  plane 40 cells downstream is the requested power within the launch's
  losses (measured 0.945-0.980 at 0.95-1.05 f0 for "+"), and the plane 40
  cells upstream sees at most 1 % of it ("-": 5 %, its H plane one cell
  further from the E plane, as in 2D).
- The same guide along z gives the same fluxes as along x (the cyclic
  relabelling of the solver and the launch), and power 4 gives 4x power 1.
- The lowered slabs against the FP64 oracle on a small grid: the H-timed
  slabs put the program on the per-phase schedule.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import fdtd64 as o
from tests.test_gpu_parity import _close, _require_gpu

pytestmark = pytest.mark.gpu

DX = 25e-9
F0 = 299792458.0 / 1.55e-6


def _guide(axis, n_long=160):
    nw, nt = 48, 40
    if axis == "x":
        eps = np.full((nt, nw, n_long), 2.085)
        eps[16:25, 15:33, :] = 12.25
    else:  # along z: (z, y, x) with the 9-cell side along y, the 18-cell side along x
        eps = np.full((n_long, nt, nw), 2.085)
        eps[:, 16:25, 15:33] = 12.25
    return eps


def _run(axis, direction, power=1.0, steps=3600):
    import beamz_amd as bz
    eps = _guide(axis)
    pulse = bz.GaussianPulse(F0, bz.LIGHT_SPEED * 0.4e-6 / 1.55e-6 ** 2)
    dt = 0.99 * DX / (bz.LIGHT_SPEED * np.sqrt(3))
    freqs = np.array([0.95 * F0, F0, 1.05 * F0])
    if axis == "x":
        src = bz.ModeSource((80 * DX, 24 * DX, 20.5 * DX), (0, 36 * DX, 30 * DX), pulse, direction,
                            power=power)
        rng = {"z": (10, 30), "y": (10, 38)}
    else:
        src = bz.ModeSource((24 * DX, 20.5 * DX, 80 * DX), (36 * DX, 30 * DX, 0), pulse, direction,
                            power=power)
        rng = {"y": (10, 30), "x": (10, 38)}
    mons = [bz.FluxMonitor(axis, 40, freqs, ranges=rng), bz.FluxMonitor(axis, 120, freqs, ranges=rng)]
    sim = bz.Simulation(eps, DX, dt=dt, boundaries=[bz.PML(cells=10)], sources=[src], monitors=mons,
                        steps=steps)
    res = sim.run()
    torch.cuda.synchronize()
    return res.normalized_flux(0), res.normalized_flux(1)


def test_mode_source_3d_direction_and_power():
    _require_gpu()
    print("case mode3d x+ x- z+ grid=(40, 48, 160) steps=3600", flush=True)
    low, high = _run("x", "+")
    assert np.all(high > 0.9) and np.all(high < 1.05), high
    assert np.all(np.abs(low) <= 1e-2 * high), (low, high)
    low_m, high_m = _run("x", "-")
    assert np.all(-low_m > 0.88) and np.all(-low_m < 1.05), low_m
    assert np.all(np.abs(high_m) <= 5e-2 * -low_m), (low_m, high_m)
    low_z, high_z = _run("z", "+")
    np.testing.assert_allclose(high_z, high, rtol=1e-5)
    np.testing.assert_allclose(low_z, low, rtol=1e-4, atol=1e-6)
    _, high4 = _run("x", "+", power=4.0)
    np.testing.assert_allclose(high4, 4.0 * high, rtol=1e-5)


def test_mode_source_3d_slabs_match_the_oracle():
    _require_gpu()
    import beamz_amd as bz
    from beamz_amd.modes3d import lower_mode_source_3d
    grid = (16, 20, 48)
    nz, ny, nx = grid
    dx = 40e-9
    eps = np.full(grid, 2.085)
    eps[6:10, 6:14, :] = 12.25
    dt = 0.9 * dx / (o.LIGHT_SPEED * np.sqrt(3))
    steps = 40
    pulse = bz.GaussianPulse(F0, 0.5 * F0, offset=1.0)
    src = bz.ModeSource((20.5 * dx, 10 * dx, 8 * dx), (0, 14 * dx, 11 * dx), pulse, "+")
    freqs = np.array([0.9 * F0, F0])
    mon = bz.FluxMonitor("x", 32, freqs, ranges={"z": (3, 13), "y": (3, 17)})
    sim = bz.Simulation(eps, dx, dt=dt, boundaries=[bz.PML(cells=3)], sources=[src], monitors=[mon],
                        steps=steps)
    prog = sim.compile()
    customs = lower_mode_source_3d(src, prog.eps_site, grid, dx, dt, 0.0, steps)
    srcs = [o.Source(c.component, c.timing, c.start,
                     np.asarray(c.coeff, np.float32).astype(np.float64),
                     np.asarray(c.waveform, np.float32).astype(np.float64)) for c in customs]
    shapes = o.shapes_3d(grid)
    plans = {c: None for c in o.COMPONENTS}
    for c in bz.monitors.FLUX_COMPONENTS["x"]:
        plans[c] = o.plane_plan_3d(c, shapes[c], grid, dx, "x", 32, {"z": (3, 13), "y": (3, 17)})
    omon = o.Monitor(freqs, plans, "x", 1.0, dx * dx)
    prob = o.build_problem(eps, dx, dt, cpml=o.CpmlSpec(cells=3, kappa_max=2.0), sources=srcs,
                           monitors=[omon])
    ost = o.init_state(prob)
    print(f"case mode3d oracle grid={grid} steps={steps}", flush=True)
    res = sim.run()
    o.run(prob, ost, steps)
    for c in prob.shapes:
        _close(c, getattr(res.state, c.lower()).cpu().numpy(), ost[c.lower()])
    want = o.monitor_flux(prob, ost, 0)
    np.testing.assert_allclose(res.monitor_flux(0), want, rtol=1e-4, atol=1e-6 * np.abs(want).max())
