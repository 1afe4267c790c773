"""A 2D mode source through ``Simulation`` on the GPU against the FP64
oracle driven by the same lowered Huygens pair (timed H and Ez slabs)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import fdtd64 as o
from tests.test_gpu_parity import _close, _require_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("axis,direction", [("x", "+"), ("y", "-")])
def test_mode_source_parity(axis, direction):
    _require_gpu()
    import beamz_amd as bz
    from beamz_amd.modes import ModeSource, lower_mode_source_2d
    dx, L, steps = 20e-9, 10, 700
    ny, nx = (60, 160) if axis == "x" else (160, 60)
    eps = np.full((ny, nx), 2.085)
    if axis == "x":
        eps[24:36, :] = 12.25
    else:
        eps[:, 24:36] = 12.25
    dt = 0.99 * dx / (o.LIGHT_SPEED * np.sqrt(2))
    f0 = o.LIGHT_SPEED / 1.55e-6
    pulse = bz.GaussianPulse(f0, o.LIGHT_SPEED * 0.4e-6 / 1.55e-6 ** 2, offset=1.5)
    c = (80.5 * dx, 30 * dx, 0) if axis == "x" else (30 * dx, 80.5 * dx, 0)
    size = (0, 40 * dx, 0) if axis == "x" else (40 * dx, 0, 0)
    src = ModeSource(c, size, pulse, direction)
    freqs = np.array([0.95 * f0, f0, 1.05 * f0])
    idx = 120 if direction == "+" else 40
    h = "Hy" if axis == "x" else "Hx"
    mon = bz.FluxMonitor(axis, idx, freqs, ranges={("y" if axis == "x" else "x"): (10, 50)},
                         components=("Ez", h))
    sim = bz.Simulation(eps, dx, dt=dt, boundaries=[bz.PML(cells=L)], sources=[src],
                        monitors=[mon], steps=steps)
    prog = sim.compile()
    eps_ez = o.site_sample_tm(eps, "Ez")
    np.testing.assert_array_equal(prog.eps_site["Ez"], eps_ez)
    customs = lower_mode_source_2d(src, eps_ez, dx, dt, 0.0, steps)
    srcs = [o.Source(s.component, s.timing, s.start,
                     np.asarray(s.coeff, np.float32).astype(np.float64),
                     np.asarray(s.waveform, np.float32).astype(np.float64)) for s in customs]
    shapes = o.shapes_tm((ny, nx))
    plans = {k: None for k in o.COMPONENTS}
    for k in ("Ez", h):
        plans[k] = o.line_plan_tm(k, shapes[k], dx, axis, idx, (10, 50))
    omon = o.Monitor(freqs, plans, axis, 1.0, dx)
    spec = o.CpmlSpec(cells=L, kappa_max=2.0, faces=frozenset(("bottom", "top", "left", "right")))
    prob = o.build_problem(eps, dx, dt, cpml=spec, sources=srcs, monitors=[omon])
    ost = o.init_state(prob)
    state = sim.initial_state()
    print(f"case mode_source axis={axis} direction={direction} grid={(ny, nx)} steps={steps}",
          flush=True)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        state = sim.advance(state, steps, chunk=128)
    torch.cuda.synchronize()
    o.run(prob, ost, steps)
    for k in ("Ez", "Hx", "Hy"):
        _close(k, getattr(state, k.lower()).cpu().numpy(), ost[k.lower()])
    flux = bz.SimulationResults(state, prog, None).monitor_flux(0)
    want = o.monitor_flux(prob, ost, 0)
    assert np.max(np.abs(want)) > 0.0
    np.testing.assert_allclose(flux, want, rtol=1e-4, atol=1e-6 * np.abs(want).max())
