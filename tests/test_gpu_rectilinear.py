"""Non-uniform grids on the GPU against the FP64 oracle: rectilinear and
axis-uniform 3D grids with CPML, a source and a flux plane, and a
rectilinear 2D (TMxy) grid.  The HIP program runs the per-phase kernels with
the grid's derivative metrics (``metric_kind`` 1 or 2); the oracle applies
the same metrics (``kernels.py:506-551``, ``629-700``, ``880-949``)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import fdtd64 as o
from tests.test_gpu_parity import _compare_state, _require_gpu

pytestmark = pytest.mark.gpu


def _widths(n, seed, lo=20e-9, hi=35e-9):
    return np.random.default_rng(seed).uniform(lo, hi, n)


def _run(kind, is_3d, steps=40):
    import beamz_amd as bz
    from beamz_amd.grid import RectilinearGrid
    if is_3d:
        nz, ny, nx = 14, 16, 40
        if kind == "rectilinear":
            g = RectilinearGrid.from_widths(_widths(nx, 1), _widths(ny, 2), _widths(nz, 3))
        else:
            g = RectilinearGrid.from_spacing((nx, ny, nz), (20e-9, 26e-9, 31e-9))
        eps = 1.5 + 8.0 * np.random.default_rng(4).random((nz, ny, nx))
        L = 3
        dt = 0.9 * g.cfl_time_step(1.0)
        t = o.sample_times(0.0, dt, steps)
        wave = o.gaussian_pulse(t, 2.0e14, 0.6e14, offset=1.5).astype(np.float32).astype(np.float64)
        coeff = (np.random.default_rng(5).random((nz - 2 * L, ny - 2 * L, 1)) * 0.2)
        coeff = coeff.astype(np.float32).astype(np.float64)
        src = o.Source("Ez", "pre_e", (L, L, 8), coeff, wave)
        mons = [bz.FluxMonitor("x", 30, np.array([1.8e14, 2.0e14, 2.2e14]),
                               ranges={"z": (L, nz - L), "y": (L, ny - L)})]
        faces = frozenset(("front", "back", "bottom", "top", "left", "right"))
        widths = {0: g.cell_widths("z"), 1: g.cell_widths("y"), 2: g.cell_widths("x")}
    else:
        ny, nx = 36, 60
        g = RectilinearGrid.from_widths(_widths(nx, 6), _widths(ny, 7))
        eps = 1.5 + 8.0 * np.random.default_rng(8).random((ny, nx))
        L = 5
        dt = 0.95 * g.cfl_time_step(1.0, axes=("x", "y"))
        t = o.sample_times(0.0, dt, steps)
        wave = o.gaussian_pulse(t, 2.0e14, 0.6e14, offset=1.5).astype(np.float32).astype(np.float64)
        coeff = np.full((ny - 2 * L, 1), -0.1)
        src = o.Source("Ez", "pre_e", (L, 12), coeff, wave)
        mons = [bz.FluxMonitor("x", 48, np.array([1.9e14, 2.1e14]), ranges={"y": (L, ny - L)})]
        faces = frozenset(("bottom", "top", "left", "right"))
        widths = {0: g.cell_widths("y"), 1: g.cell_widths("x")}
    sim = bz.Simulation(eps, grid=g, dt=dt, boundaries=[bz.PML(cells=L)],
                        sources=[bz.CustomSource(src.component, src.start, src.coeff, src.waveform,
                                                 src.timing)],
                        monitors=mons, steps=steps)
    prog = sim.compile()
    assert prog.metric_kind == (2 if kind == "rectilinear" else 1)
    # the oracle: same raster, CPML in cells with the smallest width, the
    # grid's metrics, the lowered (physical-coordinate) monitor plans
    omons = [o.Monitor(m.freqs, m.plans, m.normal, m.normal_sign, m.measure) for m in prog.monitors]
    spec = o.CpmlSpec(cells=L, faces=faces)
    metrics = o.grid_metrics(widths if kind == "rectilinear" else
                             {a: np.float64(w[0]) for a, w in widths.items()}, is_3d)
    prob = o.build_problem(eps, prog.resolution, dt, cpml=spec, sources=[src], monitors=omons,
                           metrics=metrics)
    # random initial fields (as the uniform parity tests start), so that every
    # component and its DFT is well above float32 rounding
    from tests.cases import random_fields
    fields = random_fields(prob.shapes, 9, scale=1e-5)
    fields32 = {k: v.astype(np.float32).astype(np.float64) for k, v in fields.items()}
    state = sim.initial_state(fields32)
    ost = o.init_state(prob)
    for k, v in fields32.items():
        ost[k][...] = v
    print(f"case rectilinear kind={kind} 3d={is_3d} grid={eps.shape} steps={steps}", flush=True)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        state = sim.advance(state, steps, chunk=16)
    torch.cuda.synchronize()
    from beamz_amd import native
    assert not native.last_program_fused()   # metrics run the per-phase kernels
    o.run(prob, ost, steps)
    assert max(np.abs(ost[c.lower()]).max() for c in prob.shapes) > 0.0
    _compare_state(prog, state, prob, ost, dft=True)
    return sim, prog, state


@pytest.mark.parametrize("kind", ["rectilinear", "axis_uniform"])
def test_rectilinear_3d_parity(kind):
    _require_gpu()
    _run(kind, True)


def test_rectilinear_tm2d_parity():
    _require_gpu()
    _run("rectilinear", False)
