"""Flux monitors placed by physical centre and size, on the GPU, against the
FP64 oracle: the kernels accumulate the DFT of whatever gather plan the host
lowering gives (pinned on the CPU by ``test_monitor_placement``), here
planes between E-node planes with extents off the cell edges (eight-corner
trilinear plans) in the pair and single-step kernels, and 2D lines read at
lattice nodes (one-point plans, weight 0 outside a component's array).
The flux uses the lowered sign and area; tolerances as the other parity
tests (relative L2 2e-5 per field, 5e-5 for DFT arenas, flux 1e-4)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import fdtd64 as o
from tests.test_gpu_parity import _compare_state, _plane_source, _require_gpu, _setup

pytestmark = pytest.mark.gpu

DX = 40e-9


def _flux_close(sim_flux, want):
    np.testing.assert_allclose(sim_flux, want, rtol=1e-4, atol=1e-6 * np.abs(want).max())


@pytest.mark.parametrize("chunk", [12, 5])
def test_placed_planes_3d(chunk):
    _require_gpu()
    import beamz_amd as bz
    grid = (13, 17, 40)
    steps = 12
    dt = 0.9 * DX / (o.LIGHT_SPEED * np.sqrt(3))
    freqs = np.array([1.8e14, 2.0e14, 2.2e14])
    mons = [bz.FluxMonitor(center=(23.3 * DX, 8.6 * DX, 6.2 * DX), size=(0.0, 9.7 * DX, 7.4 * DX),
                           freqs=freqs, name="x_plane"),
            bz.FluxMonitor(center=(20.0 * DX, 8.5 * DX, 9.55 * DX), size=(30.2 * DX, 11.0 * DX, 0.0),
                           freqs=freqs[1:], name="z_plane", interval=2),
            # all six components on a line along y (the reference's FieldMonitor)
            bz.FieldMonitor(center=(17.3 * DX, 8.2 * DX, 5.7 * DX), size=(0.0, 9.4 * DX, 0.0),
                            freqs=freqs, name="y_line")]
    sim, prog, state, prob, ost = _setup(grid, seed=81, cpml_cells=3, steps=steps,
                                         sources=[_plane_source(grid, dt, steps, 7)], monitors=mons)
    print(f"case placed_planes_3d grid={grid} chunk={chunk}", flush=True)
    assert all(m.placed for m in prog.monitors)
    assert all(p is not None for p in prog.monitors[2].plans.values())
    assert max(p[0].shape[1] for m in prog.monitors for p in m.plans.values() if p is not None) == 8
    state = sim.advance(state, steps, chunk=chunk)
    torch.cuda.synchronize()
    o.run(prob, ost, steps)
    _compare_state(prog, state, prob, ost, dft=True)
    res = bz.api.SimulationResults(state=state, program=prog, performance=None)
    for i in range(2):
        _flux_close(res.monitor_flux(i), o.monitor_flux(prob, ost, i))


def test_placed_lines_2d():
    _require_gpu()
    import beamz_amd as bz
    grid = (40, 150)
    steps = 100
    dt = 0.99 * DX / (o.LIGHT_SPEED * np.sqrt(2))
    t = o.sample_times(0.0, dt, steps)
    wave = o.gaussian_pulse(t, 2.0e14, 6e13, offset=2.0, carrier="sin").astype(np.float32).astype(np.float64)
    src = o.Source("Ez", "pre_e", (15, 20), np.full((10, 1), -0.05), wave)
    freqs = np.array([1.9e14, 2.0e14, 2.1e14])
    mons = [bz.FluxMonitor(center=(120.3 * DX, 20.0 * DX), size=(0.0, 30.0 * DX, 1e-6), freqs=freqs,
                           name="x_line"),
            # along the top edge: Hx has no row there, so its samples weigh 0
            bz.FluxMonitor(center=(60.0 * DX, 40.0 * DX), size=(80.0 * DX, 0.0, 1e-6), freqs=freqs,
                           name="top_line"),
            bz.FluxMonitor(center=(70.0 * DX, 31.0 * DX), size=(100.0 * DX, 0.0, 1e-6), freqs=freqs,
                           name="y_line")]
    sim, prog, state, prob, ost = _setup(grid, seed=83, cpml_cells=6, kappa_max=5.0,
                                         sources=[src], monitors=mons, steps=steps)
    print(f"case placed_lines_2d grid={grid} steps={steps}", flush=True)
    assert [m.normal_sign for m in prog.monitors] == [1.0, -1.0, -1.0]
    state = sim.advance(state, steps, chunk=64)
    torch.cuda.synchronize()
    o.run(prob, ost, steps)
    _compare_state(prog, state, prob, ost, dft=True)
    res = bz.api.SimulationResults(state=state, program=prog, performance=None)
    for i in (0, 2):
        _flux_close(res.monitor_flux(i), o.monitor_flux(prob, ost, i))
    assert np.all(res.monitor_flux(1) == 0.0) and np.all(o.monitor_flux(prob, ost, 1) == 0.0)


def test_gaussian_source_and_placed_monitor_2d():
    """A physically placed Gaussian source (the reference's GaussianSource)
    driving placed monitors through ``Simulation.run``: against the oracle run
    on the same lowered slab and plans."""
    _require_gpu()
    import beamz_amd as bz
    from beamz_amd.sources import lower_gaussian_source
    grid = (60, 120)
    steps = 1200
    eps = np.full(grid, 2.085)
    eps[26:34, :] = 12.25
    dt = 0.99 * DX / (o.LIGHT_SPEED * np.sqrt(2))
    f0 = 1.93e14
    # a sample array (the reference's form of the signal): the pulse peaks at
    # 2e-14 s and has passed the x-normal monitor well before 1200 steps (5.6e-14 s)
    signal = bz.GaussianPulse(f0, 1.5e14, offset=3.0, carrier="sin").sample(np.arange(steps) * dt)[0]
    src = bz.GaussianSource((20.3 * DX, 30.0 * DX), 1.7 * DX, signal)
    freqs = np.array([0.95 * f0, f0, 1.05 * f0])
    mons = [bz.FluxMonitor(center=(95.2 * DX, 30.0 * DX), size=(0.0, 40.0 * DX, 1e-6), freqs=freqs),
            bz.FluxMonitor(center=(60.0 * DX, 50.5 * DX), size=(90.0 * DX, 0.0, 1e-6), freqs=freqs)]
    sim = bz.Simulation(eps, DX, dt=dt, boundaries=[bz.PML(cells=8, kappa_max=5.0)], sources=[src],
                        monitors=mons, steps=steps)
    prog = sim.compile()
    cs = lower_gaussian_source(src, prog.eps_site["Ez"], DX, dt, 0.0, steps)
    osrc = o.Source(cs.component, cs.timing, cs.start, np.asarray(cs.coeff, np.float32).astype(np.float64),
                    np.asarray(cs.waveform, np.float32).astype(np.float64))
    shapes = o.shapes_tm(grid)
    omons = []
    for i, m in enumerate(mons):
        cm = bz.monitors.lower_monitor(m, grid, shapes, DX, i)
        omons.append(o.Monitor(cm.freqs, cm.plans, cm.normal, cm.normal_sign, cm.measure))
    prob = o.build_problem(eps, DX, dt, cpml=o.CpmlSpec(cells=8, kappa_max=5.0), sources=[osrc],
                           monitors=omons)
    ost = o.init_state(prob)
    print(f"case gaussian_source_2d grid={grid} steps={steps}", flush=True)
    res = sim.run()
    torch.cuda.synchronize()
    o.run(prob, ost, steps)
    for c in prob.shapes:
        from tests.test_gpu_parity import _close
        _close(c, getattr(res.state, c.lower()).cpu().numpy(), ost[c.lower()])
    for i in range(2):
        _flux_close(res.monitor_flux(i), o.monitor_flux(prob, ost, i))
    assert np.all(res.monitor_flux(0) > 0)
