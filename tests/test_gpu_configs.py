"""Two benchmark configurations at their full size against the FP64 oracle:

- the 2D straight waveguide (200 x 400 cells at 20 nm, 25-cell core of eps
  12.25 in 2.085, 10-cell CPML with kappa_max 5 and alpha 0.05, soft Ez line
  source across the core at x = 40, Gaussian-modulated sine at 1.55 um with
  0.2 um bandwidth, flux line at x = 360 over 21 wavelengths 1.45-1.65 um),
  all 2000 steps;
- the 3D silicon waveguide (128 x 256 x 512 cells at 25 nm, 450 x 220 nm Si
  core along x, 12-cell CPML, Gaussian plane source at x = 30, three flux
  planes at 11 wavelengths), 4 fused steps from random fields (the oracle
  takes seconds per step at 16.8 M cells), and the full 4000-step run with
  two steps per launch against single steps."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import fdtd64 as o
from tests.test_gpu_parity import _close, _require_gpu

pytestmark = pytest.mark.gpu


def _oracle_sources(prog_customs):
    return [o.Source(c.component, c.timing, c.start,
                     np.asarray(c.coeff, np.float32).astype(np.float64),
                     np.asarray(c.waveform, np.float32).astype(np.float64)) for c in prog_customs]


def test_2d_straight_waveguide_full_run():
    _require_gpu()
    import beamz_amd as bz
    from beamz_amd.sources import lower_soft_source
    ny, nx, dx, steps = 200, 400, 20e-9, 2000
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(2))
    eps = np.full((ny, nx), 2.085)
    eps[88:113, :] = 12.25
    f0 = bz.LIGHT_SPEED / 1.55e-6
    fw = bz.LIGHT_SPEED * 0.2e-6 / 1.55e-6 ** 2
    alpha = 2.0 * o.EPS_0 * 0.05 / dt
    src = bz.SoftSource("Ez", (88, 40), (25, 1), bz.GaussianPulse(f0, fw, carrier="sin"))
    freqs = bz.LIGHT_SPEED / np.linspace(1.45e-6, 1.65e-6, 21)
    mon = bz.FluxMonitor("x", 360, freqs, ranges={"y": (10, 190)})
    sim = bz.Simulation(eps, dx, dt=dt, sources=[src], monitors=[mon], steps=steps,
                        boundaries=[bz.PML(cells=10, kappa_max=5.0, alpha_max=alpha)])
    prog = sim.compile()
    custom = lower_soft_source(src, prog.eps_site["Ez"], dt, 0.0, steps)
    shapes = o.shapes_tm((ny, nx))
    plans = {c: None for c in o.COMPONENTS}
    for c in ("Ez", "Hy"):
        plans[c] = o.line_plan_tm(c, shapes[c], dx, "x", 360, (10, 190))
    omon = o.Monitor(freqs, plans, "x", 1.0, dx)
    spec = o.CpmlSpec(cells=10, kappa_max=5.0, alpha_norm=0.05,
                      faces=frozenset(("bottom", "top", "left", "right")))
    prob = o.build_problem(eps, dx, dt, cpml=spec, sources=_oracle_sources([custom]),
                           monitors=[omon])
    ost = o.init_state(prob)
    print(f"case config 2D straight waveguide grid={(ny, nx)} steps={steps}", flush=True)
    res = sim.run()
    o.run(prob, ost, steps)
    for c in ("Ez", "Hx", "Hy"):
        _close(c, getattr(res.state, c.lower()).cpu().numpy(), ost[c.lower()], rel=1e-4, mabs=5e-4)
    flux = res.monitor_flux(0)
    want = o.monitor_flux(prob, ost, 0)
    err = np.max(np.abs(flux - want)) / np.max(np.abs(want))
    print(f"  flux spectrum max relative error {err:.2e}", flush=True)
    assert np.max(np.abs(want)) > 0.0 and err <= 1e-4, err


def test_3d_silicon_waveguide_fused_steps():
    _require_gpu()
    import beamz_amd as bz
    from beamz_amd import native
    from beamz_amd.sources import lower_soft_source
    from tests.cases import random_fields
    nz, ny, nx, dx, L, steps = 128, 256, 512, 25e-9, 12, 4
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(3))
    eps = np.full((nz, ny, nx), 2.085)
    eps[59:68, 119:137, :] = 12.25
    f0 = bz.LIGHT_SPEED / 1.55e-6
    fw = bz.LIGHT_SPEED * 0.2e-6 / 1.55e-6 ** 2
    src = bz.SoftSource("Ez", (L, L, 30), (nz - 2 * L, ny + 1 - 2 * L, 1), bz.GaussianPulse(f0, fw))
    freqs = bz.LIGHT_SPEED / np.linspace(1.5e-6, 1.6e-6, 11)
    rngs = {"z": (L, nz - L), "y": (L, ny - L)}
    mons = [bz.FluxMonitor("x", xi, freqs, ranges=rngs) for xi in (60, 256, 450)]
    sim = bz.Simulation(eps, dx, dt=dt, boundaries=[bz.PML(cells=L)], sources=[src],
                        monitors=mons, steps=4000)
    prog = sim.compile()
    custom = lower_soft_source(src, prog.eps_site["Ez"], dt, 0.0, 4000)
    shapes = o.shapes_3d((nz, ny, nx))
    omons = []
    for m in mons:
        plans = {c: None for c in o.COMPONENTS}
        for c in bz.monitors.FLUX_COMPONENTS["x"]:
            plans[c] = o.plane_plan_3d(c, shapes[c], (nz, ny, nx), dx, "x", m.index, rngs)
        omons.append(o.Monitor(freqs, plans, "x", 1.0, dx * dx))
    spec = o.CpmlSpec(cells=L, kappa_max=2.0)
    prob = o.build_problem(eps, dx, dt, cpml=spec, sources=_oracle_sources([custom]),
                           monitors=omons)
    fields = random_fields(prob.shapes, 31, scale=1e-5)
    fields32 = {k: v.astype(np.float32).astype(np.float64) for k, v in fields.items()}
    state = sim.initial_state(fields32)
    ost = o.init_state(prob)
    for k, v in fields32.items():
        ost[k][...] = v
    print(f"case config 3D silicon waveguide grid={(nz, ny, nx)} steps={steps}", flush=True)
    state = sim.advance(state, steps, chunk=steps)
    torch.cuda.synchronize()
    assert native.last_program_fused()
    o.run(prob, ost, steps)
    for c in prob.shapes:
        _close(c, getattr(state, c.lower()).cpu().numpy(), ost[c.lower()])
    re, im = state.dft_re.cpu().numpy(), state.dft_im.cpu().numpy()
    for mi, m in enumerate(prog.monitors):
        F, P = m.freqs.size, m.points
        for c, name in enumerate(o.COMPONENTS):
            if m.plans.get(name) is None:
                continue
            sl = slice(m.value_offset + c * F * P, m.value_offset + (c + 1) * F * P)
            _close(f"dft_re {name}", re[sl].reshape(F, P), ost["dft_re"][mi][c], rel=5e-5)
            _close(f"dft_im {name}", im[sl].reshape(F, P), ost["dft_im"][mi][c], rel=5e-5)


def test_3d_silicon_waveguide_full_run_pairs_match_single_steps(monkeypatch):
    """The 3D silicon waveguide (128 x 256 x 512, 4000 steps, the plane source
    and three flux planes) run to the end with two steps per launch and with
    single steps: the flux spectra agree to FP32 rounding accumulated over the
    run (relative 1e-5 of the largest)."""
    _require_gpu()
    import beamz_amd as bz
    from beamz_amd import native
    nz, ny, nx, dx, L, steps = 128, 256, 512, 25e-9, 12, 4000
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(3))
    eps = np.full((nz, ny, nx), 2.085)
    eps[59:68, 119:137, :] = 12.25
    f0 = bz.LIGHT_SPEED / 1.55e-6
    fw = bz.LIGHT_SPEED * 0.2e-6 / 1.55e-6 ** 2
    freqs = bz.LIGHT_SPEED / np.linspace(1.5e-6, 1.6e-6, 11)
    rngs = {"z": (L, nz - L), "y": (L, ny - L)}
    fluxes = []
    for pair in ("1", "0"):
        monkeypatch.setenv("BZH_PAIR", pair)
        src = bz.SoftSource("Ez", (L, L, 30), (nz - 2 * L, ny + 1 - 2 * L, 1), bz.GaussianPulse(f0, fw))
        mons = [bz.FluxMonitor("x", xi, freqs, ranges=rngs) for xi in (60, 256, 450)]
        sim = bz.Simulation(eps, dx, dt=dt, boundaries=[bz.PML(cells=L)], sources=[src],
                            monitors=mons, steps=steps)
        print(f"case config 3D silicon waveguide full run pair={pair} steps={steps}", flush=True)
        res = sim.run()
        torch.cuda.synchronize()
        assert native.last_program_fused()
        assert (native.last_program_pairs() > 0) == (pair == "1")
        fluxes.append([res.monitor_flux(i) for i in range(3)])
        del res, sim
        torch.cuda.empty_cache()
    for i, (a, b) in enumerate(zip(*fluxes)):
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        print(f"  monitor {i}: max relative flux difference {err:.2e}", flush=True)
        assert np.max(np.abs(b)) > 0.0 and err <= 1e-5, (i, err)


def test_batched_sweep_64_members_full_run():
    """The batched 2D sweep at its full size: 64 TMxy simulations, (ny, nx)
    ~ U{256..1024}^2 at 20 nm, 2-4 guides each, 10-cell CPML, pre-E Ez line
    sources, 3000 steps in one persistent launch per step.  Every member
    must equal the same member run alone through ``Simulation`` (the
    per-phase kernels) to 5e-5 relative L2 per field, and the two smallest
    members the FP64 oracle to 1e-4 (3000 FP32 steps)."""
    _require_gpu()
    import bench
    from beamz_amd.batch import BatchedSimulation2D
    from beamz_amd.sources import lower_soft_source
    steps = 3000
    sims = bench.build_batched_workload(count=64, steps=steps)
    batch = BatchedSimulation2D(sims)
    cells = batch.cells
    print(f"case config batched sweep: 64 members, {cells} cells, {steps} steps", flush=True)
    state = batch.initial_state()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        state = batch.advance(state, steps)
    torch.cuda.synchronize()
    print(f"batched sweep done (tile_rows={batch.tile_rows}, workgroups={batch.workgroups})",
          flush=True)
    worst = 0.0
    for i, sim in enumerate(sims):
        with torch.cuda.stream(s):
            solo = sim.advance(sim.initial_state(), steps)
        torch.cuda.synchronize()
        for name in ("Ez", "Hx", "Hy"):
            want = getattr(solo, name.lower()).double().cpu().numpy()
            got = batch.field(state, i, name)
            err = np.linalg.norm(got - want) / np.linalg.norm(want)
            worst = max(worst, err)
            assert err <= 5e-5, f"member {i} {sim.permittivity.shape} {name}: {err:.2e}"
    print(f"batched vs solo: worst relative L2 {worst:.2e} over 64 members", flush=True)
    order = sorted(range(len(sims)), key=lambda i: sims[i].permittivity.size)[:2]
    for i in order:
        sim = sims[i]
        prog = sim.compile()
        customs = [lower_soft_source(src, prog.eps_site["Ez"], sim.dt, 0.0, steps)
                   for src in sim.sources]
        spec = o.CpmlSpec(cells=10, kappa_max=sims[i].boundaries[0].kappa_max,
                          faces=frozenset(("bottom", "top", "left", "right")))
        prob = o.build_problem(sim.permittivity, sim.resolution, sim.dt, cpml=spec,
                               sources=_oracle_sources(customs))
        ost = o.init_state(prob)
        o.run(prob, ost, steps)
        for name in ("Ez", "Hx", "Hy"):
            _close(f"member {i} {name}", batch.field(state, i, name), ost[name.lower()], rel=1e-4,
                   mabs=1e-3)
        print(f"member {i} {sim.permittivity.shape} matches the oracle", flush=True)
