"""Batched 2D simulations (one persistent launch per step for the whole
batch) against the FP64 oracle, member by member."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import fdtd64 as o
from tests.cases import eps_raster

pytestmark = pytest.mark.gpu

REL_L2 = 2e-5


def _require_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a GPU")
    from beamz_amd import native
    native.library()


def _member(grid, seed, *, cpml_cells, metallic, steps, dx=20e-9):
    """One 2D member: a random-eps raster, an Ez line source, both the
    beamz_amd Simulation and the oracle problem."""
    import beamz_amd as bz
    ny, nx = grid
    eps = eps_raster(grid, seed, lo=2.085, hi=12.25)
    dt = 0.99 * dx / (o.LIGHT_SPEED * np.sqrt(2))
    t = o.sample_times(0.0, dt, steps)
    wave = np.sin(2 * np.pi * 1.93e14 * t) * np.exp(-(((t - 12 * dt) / (6 * dt)) ** 2))
    wave = wave.astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(seed + 5)
    y0, x0, sy = ny // 4, min(12, nx // 3), ny // 2
    coeff = (-rng.uniform(0.5, 1.0, (sy, 1)) * 1e-1).astype(np.float32).astype(np.float64)
    bounds = []
    spec = None
    if cpml_cells:
        bounds.append(bz.PML(cells=cpml_cells, kappa_max=5.0))
        spec = o.CpmlSpec(cells=cpml_cells, kappa_max=5.0,
                          faces=frozenset(("bottom", "top", "left", "right")))
    if metallic:
        bounds.append(bz.PEC(edges=tuple(metallic)))
    src = o.Source("Ez", "pre_e", (y0, x0), coeff, wave)
    sim = bz.Simulation(eps, dx, dt=dt, boundaries=bounds,
                        sources=[bz.CustomSource("Ez", (y0, x0), coeff, wave, "pre_e")], steps=steps)
    prob = o.build_problem(eps, dx, dt, cpml=spec, metallic=frozenset(metallic) if metallic else None,
                           sources=[src])
    assert sim.compile().metallic == prob.metallic
    return sim, prob


@pytest.mark.parametrize("tile_rows", [8, 16, 64])
def test_batched_members_match_oracle(tile_rows):
    """Uneven grids (CPML on every edge, PEC on every edge, and one with CPML
    next to metal on a grid smaller than a tile) advance together; 41 steps in
    chunks of 16, 16 and 9 (an odd chunk swaps the banks)."""
    _require_gpu()
    from beamz_amd.batch import BatchedSimulation2D
    steps = 41
    specs = [((70, 133), 3, 8, None), ((37, 66), 5, 6, None), ((90, 40), 7, 0, ("bottom", "top", "left", "right")),
             ((21, 30), 9, 4, None)]
    members = [_member(g, s, cpml_cells=L, metallic=m, steps=steps) for g, s, L, m in specs]
    batch = BatchedSimulation2D([sim for sim, _ in members], tile_rows=tile_rows, workgroups=37)
    print(f"case batched tile_rows={tile_rows} grids={[g for g, *_ in specs]}", flush=True)
    state = batch.initial_state()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        state = batch.advance(state, steps, chunk=16)
    torch.cuda.synchronize()
    assert int(state["step"].item()) == steps
    for i, (_, prob) in enumerate(members):
        ost = o.init_state(prob)
        o.run(prob, ost, steps)
        for name in ("Ez", "Hx", "Hy"):
            got = batch.field(state, i, name)
            want = ost[name.lower()]
            err = np.linalg.norm(got - want) / np.linalg.norm(want)
            assert err <= REL_L2, f"member {i} {name}: relative L2 {err:.2e}"


def test_autotune_picks_a_candidate_and_caches_it():
    _require_gpu()
    from beamz_amd.batch import BatchedSimulation2D, autotune
    members = [_member(g, s, cpml_cells=4, metallic=None, steps=8)[0]
               for g, s in (((64, 200), 1), ((48, 100), 2))]
    batch = BatchedSimulation2D(members)
    tr, wg, timings = autotune(batch, steps=2, candidates=[(8, 64), (16, 128)])
    assert (tr, wg) in timings and timings[(tr, wg)] == min(timings.values())
    assert batch.tile_rows == tr and batch.workgroups == wg
    again = autotune(BatchedSimulation2D(members))
    assert again[:2] == (tr, wg)
