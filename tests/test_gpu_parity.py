"""GPU parity of the HIP path against the FP64 oracle on seeded inputs.

Every call goes through ``libbeamz_hip.so`` (``bzh_streamed`` / ``bzh_program``).
Tolerance: the FP32 kernels against the FP64 oracle must agree to a relative
L2 error of ``REL_L2`` per leaf (fields, CPML memories, DFT arenas) and to
``MAX_ABS`` of each leaf's maximum elementwise, after 32-100 steps.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import fdtd64 as o
from tests.cases import eps_raster, random_fields

pytestmark = pytest.mark.gpu

REL_L2 = 2e-5
MAX_ABS = 1e-4


def _close(name, got, want, rel=REL_L2, mabs=MAX_ABS):
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    assert got.shape == want.shape, (name, got.shape, want.shape)
    scale = np.linalg.norm(want)
    if scale == 0.0:
        assert np.abs(got).max(initial=0.0) == 0.0, name
        return
    err = np.linalg.norm(got - want) / scale
    assert err <= rel, f"{name}: relative L2 {err:.3e} > {rel:.1e}"
    assert np.abs(got - want).max() <= mabs * np.abs(want).max(), f"{name}: max abs error"


def _require_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a GPU")
    from beamz_amd import native
    native.library()


def _setup(grid, *, seed, cpml_cells=None, faces=None, metallic=None, kappa_max=2.0,
           sources=(), monitors=(), steps=1, dt_scale=0.9, fused=True):
    import beamz_amd as bz
    eps = eps_raster(grid, seed)
    dx = 40e-9
    dt = dt_scale * dx / (o.LIGHT_SPEED * np.sqrt(len(grid)))
    bounds = []
    spec = None
    if cpml_cells:
        bounds.append(bz.PML(edges=faces or "all", cells=cpml_cells, kappa_max=kappa_max))
        allf = ("front", "back", "bottom", "top", "left", "right") if len(grid) == 3 else \
            ("bottom", "top", "left", "right")
        spec = o.CpmlSpec(cells=cpml_cells, kappa_max=kappa_max, faces=frozenset(faces or allf))
    if metallic:
        bounds.append(bz.PEC(edges=tuple(metallic)))
    customs = [bz.CustomSource(s.component, s.start, s.coeff, s.waveform, s.timing) for s in sources]
    omons = []
    for m in monitors:
        prog_shapes = o.shapes_3d(grid) if len(grid) == 3 else o.shapes_tm(grid)
        plans = {}
        for c in o.COMPONENTS:
            if c not in (m.components or bz.monitors.FLUX_COMPONENTS[m.normal]) or c not in prog_shapes:
                plans[c] = None
                continue
            if len(grid) == 3:
                tang = {"x": ("z", "y"), "y": ("z", "x"), "z": ("y", "x")}[m.normal]
                ax = {"z": 0, "y": 1, "x": 2}
                rng = {a: tuple(m.ranges.get(a, (0, grid[ax[a]]))) for a in tang}
                plans[c] = o.plane_plan_3d(c, prog_shapes[c], grid, dx, m.normal, m.index, rng)
            else:
                tang = "y" if m.normal == "x" else "x"
                rng = tuple(m.ranges.get(tang, (0, grid[{"y": 0, "x": 1}[tang]])))
                plans[c] = o.line_plan_tm(c, prog_shapes[c], dx, m.normal, m.index, rng)
        omons.append(o.Monitor(np.asarray(m.freqs, float), plans, m.normal, 1.0,
                               dx * dx if len(grid) == 3 else dx))
    prob = o.build_problem(eps, dx, dt, cpml=spec,
                           metallic=frozenset(metallic) if metallic else None,
                           sources=sources, monitors=omons)
    flags = bz._abi.DEFAULT_FLAGS | (bz._abi.FLAG_FUSED_STEP if fused else 0)
    sim = bz.Simulation(eps, dx, dt=dt, boundaries=bounds, sources=customs, monitors=monitors,
                        steps=steps, flags=flags)
    prog = sim.compile()
    assert prog.metallic == prob.metallic
    fields = random_fields(prob.shapes, seed + 1)
    fields32 = {k: v.astype(np.float32) for k, v in fields.items()}
    state = sim.initial_state(fields32)
    ost = o.init_state(prob)
    for k, v in fields32.items():
        ost[k][...] = v.astype(np.float64)
    return sim, prog, state, prob, ost


def _compare_state(prog, state, prob, ost, dft=False):
    for c in prob.shapes:
        _close(c, getattr(state, c.lower()).cpu().numpy(), ost[c.lower()])
    for i, (g, w) in enumerate(zip(state.psi_h, ost["psi_h"])):
        _close(f"psi_h{i}", g.cpu().numpy(), w, rel=5e-5)
    for i, (g, w) in enumerate(zip(state.psi_e, ost["psi_e"])):
        _close(f"psi_e{i}", g.cpu().numpy(), w, rel=5e-5)
    if dft:
        re = state.dft_re.cpu().numpy()
        im = state.dft_im.cpu().numpy()
        for mi, m in enumerate(prog.monitors):
            F, P = m.freqs.size, m.points
            for c, name in enumerate(o.COMPONENTS):
                if m.plans.get(name) is None:
                    continue
                sl = slice(m.value_offset + c * F * P, m.value_offset + (c + 1) * F * P)
                _close(f"dft_re {name}", re[sl].reshape(F, P), ost["dft_re"][mi][c], rel=5e-5)
                _close(f"dft_im {name}", im[sl].reshape(F, P), ost["dft_im"][mi][c], rel=5e-5)
        np.testing.assert_allclose(state.dft_w.cpu().numpy(),
                                   np.concatenate(ost["dft_w"]), rtol=1e-12)


@pytest.mark.parametrize("variant", ["cpml", "pec", "mixed"])
def test_single_phase_parity_3d(variant):
    _require_gpu()
    from beamz_amd import runtime
    kw = {"cpml": dict(cpml_cells=3), "pec": {},
          "mixed": dict(cpml_cells=2, faces=("left", "right", "front"),
                        metallic=("bottom", "top", "back"))}[variant]
    sim, prog, state, prob, ost = _setup((9, 11, 70), seed=11, **kw)
    print(f"case single_phase_parity_3d[{variant}] grid=(9,11,70)", flush=True)
    runtime.update_h(state, prog)
    runtime.update_e(state, prog)
    torch.cuda.synchronize()
    o.h_phase(prob, ost)
    o.apply_masks(prob, ost, ("Hx", "Hy", "Hz"))
    o.e_phase(prob, ost)
    o.apply_masks(prob, ost, ("Ex", "Ey", "Ez"))
    _compare_state(prog, state, prob, ost)


def _plane_source(grid, dt, steps, seed):
    rng = np.random.default_rng(seed)
    nz, ny, nx = grid
    coeff = (rng.random((nz - 4, ny - 4, 1)) * 1e-1).astype(np.float32).astype(np.float64)
    t = o.sample_times(0.0, dt, steps)
    f0 = 2.0e14
    wave = o.gaussian_pulse(t, f0, 0.3 * f0, offset=1.5).astype(np.float32).astype(np.float64)
    return o.Source("Ez", "pre_e", (2, 2, 6), coeff, wave)


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per_phase"])
@pytest.mark.parametrize("cpml", [True, False], ids=["cpml", "pec"])
def test_program_parity_3d_with_sources_and_monitors(cpml, fused):
    """Mirrors tests/hardware/test_cuda_backends.py:26-61: an 18x20x35 grid,
    3-cell CPML, a source crossing its peak and DFT monitors, 32 steps."""
    _require_gpu()
    import beamz_amd as bz
    grid = (18, 20, 35)
    steps = 32
    dt = 0.9 * 40e-9 / (o.LIGHT_SPEED * np.sqrt(3))
    src = _plane_source(grid, dt, steps, 5)
    h_src = o.Source("Hy", "h", (3, 4, 20), np.full((5, 6, 2), 3e-3), src.waveform * 0.5)
    e_src = o.Source("Ex", "e", (0, 0, 10), np.full((19, 21, 1), 2e-2), src.waveform)
    freqs = np.array([1.8e14, 2.0e14, 2.2e14])
    mons = [bz.FluxMonitor("x", 28, freqs),
            bz.FluxMonitor("z", 9, freqs[:2], ranges={"y": (3, 17), "x": (4, 30)})]
    srcs = [src, e_src] if fused else [src, h_src, e_src]   # H-timed sources use the per-phase schedule
    sim, prog, state, prob, ost = _setup(grid, seed=21, cpml_cells=3 if cpml else None,
                                         sources=srcs, monitors=mons, steps=steps, fused=fused)
    from beamz_amd import runtime
    assert runtime.uses_two_banks(prog) == fused
    print(f"case program_parity_3d[cpml={cpml},fused={fused}] grid={grid} steps={steps}", flush=True)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        state = sim.advance(state, steps, chunk=12)
    torch.cuda.synchronize()
    o.run(prob, ost, steps)
    _compare_state(prog, state, prob, ost, dft=True)
    flux = bz.SimulationResults(state, prog, None).monitor_flux(0)
    np.testing.assert_allclose(flux, o.monitor_flux(prob, ost, 0), rtol=1e-4,
                               atol=1e-6 * np.abs(o.monitor_flux(prob, ost, 0)).max())


def test_program_parity_tm2d():
    _require_gpu()
    import beamz_amd as bz
    grid = (40, 150)
    steps = 100
    dt = 0.99 * 40e-9 / (o.LIGHT_SPEED * np.sqrt(2))
    t = o.sample_times(0.0, dt, steps)
    wave = o.gaussian_pulse(t, 2.0e14, 6e13, offset=2.0, carrier="sin").astype(np.float32).astype(np.float64)
    src = o.Source("Ez", "pre_e", (15, 20), np.full((10, 1), -0.05), wave)
    mons = [bz.FluxMonitor("x", 120, np.array([1.9e14, 2.0e14, 2.1e14]), ranges={"y": (5, 35)})]
    sim, prog, state, prob, ost = _setup(grid, seed=31, cpml_cells=6, kappa_max=5.0,
                                         sources=[src], monitors=mons, steps=steps)
    print(f"case program_parity_tm2d grid={grid} steps={steps}", flush=True)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        state = sim.advance(state, steps, chunk=64)
    torch.cuda.synchronize()
    o.run(prob, ost, steps)
    _compare_state(prog, state, prob, ost, dft=True)


def test_graph_replay_matches_direct_enqueue_bitwise():
    """The graph cache replays a captured chunk; results equal the direct
    (default-stream, uncaptured) schedule bit for bit."""
    _require_gpu()
    grid = (10, 12, 66)
    dt = 0.9 * 40e-9 / (o.LIGHT_SPEED * np.sqrt(3))
    src = _plane_source(grid, dt, 24, 9)
    runs = []
    for use_stream in (True, False):
        sim, prog, state, prob, ost = _setup(grid, seed=41, cpml_cells=2, sources=[src], steps=24)
        print(f"case graph_replay stream={use_stream}", flush=True)
        if use_stream:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                state = sim.advance(state, 24, chunk=8)   # 3 chunks: capture once, replay twice
        else:
            state = sim.advance(state, 24, chunk=8)
        torch.cuda.synchronize()
        runs.append([t.cpu().numpy().copy() for t in state.fields()])
    for a, b in zip(*runs):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("kernel,rows", [("wg", "8"), ("wg", "15"), ("tile", "4")])
@pytest.mark.parametrize("grid,kz", [((13, 9, 130), "4"), ((40, 11, 65), "7"), ((6, 5, 64), "32")])
@pytest.mark.parametrize("bc", ["cpml", "pec", "mixed"])
def test_fused_step_parity_odd_shapes(grid, kz, bc, kernel, rows, monkeypatch):
    """Fused two-bank kernel on shapes that cut tiles and z-chunks in every
    place (partial x/y tiles, an x-interior box next to x-edge tiles, chunk
    boundaries, odd step counts), for each kernel variant and tile height."""
    _require_gpu()
    monkeypatch.setenv("BZH_FUSED_KZ", kz)
    monkeypatch.setenv("BZH_FUSED_KERNEL", kernel)
    monkeypatch.setenv("BZH_FUSED_ROWS", rows)
    kw = {"cpml": dict(cpml_cells=2), "pec": {},
          "mixed": dict(cpml_cells=2, faces=("left", "back", "top"), metallic=("right", "front"))}[bc]
    steps = 7
    sim, prog, state, prob, ost = _setup(grid, seed=51, steps=steps, **kw)
    print(f"case fused_step_parity grid={grid} kz={kz} bc={bc} kernel={kernel}/{rows}", flush=True)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        state = sim.advance(state, steps, chunk=steps)
    torch.cuda.synchronize()
    o.run(prob, ost, steps)
    _compare_state(prog, state, prob, ost)
