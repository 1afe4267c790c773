"""Two leapfrog steps per launch (the pair kernel) against the FP64 oracle
and against the single-step schedule.

The pair kernel runs wherever the fused step runs; these cases cut its
60 x 13 tiles and z chunks in every place (x-interior and x-edge tiles,
partial y tiles, one-plane chunks), put source slabs of both kinds between
the two steps of every pair (a pre-E Ez plane and an E-timed Ex slab) and
sample x-, y- and z-normal monitors on the intermediate step.  Calls of 5
steps run two pairs and one single step each, so the bank parity across
calls is exercised as well.  Tolerances are the single-step parity tests'
(relative L2 2e-5 per field, 5e-5 for CPML memories and DFT arenas).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import fdtd64 as o
from tests.test_gpu_parity import _close, _compare_state, _plane_source, _require_gpu, _setup

pytestmark = pytest.mark.gpu

DX = 40e-9


def _dt(grid):
    return 0.9 * DX / (o.LIGHT_SPEED * np.sqrt(len(grid)))


def _sources(grid, steps):
    """A pre-E Ez plane and an E-timed Ex slab near the bottom y face."""
    dt = _dt(grid)
    pre = _plane_source(grid, dt, steps, 5)
    nz, ny, nx = grid
    rng = np.random.default_rng(9)
    coeff = (rng.random((3, 4, nx - 6)) * 1e-1).astype(np.float32).astype(np.float64)
    t = o.sample_times(0.0, dt, steps)
    f0 = 2.2e14
    wave = o.gaussian_pulse(t, f0, 0.4 * f0, offset=1.0).astype(np.float32).astype(np.float64)
    post = o.Source("Ex", "e", (nz // 2 - 1, 0, 3), coeff, wave)
    return [pre, post]


def _monitors(grid):
    import beamz_amd as bz
    dt = _dt(grid)
    nz, ny, nx = grid
    freqs = np.array([1.7e14, 1.9e14, 2.1e14])
    return [bz.FluxMonitor("x", nx // 2 + 1, freqs),
            bz.FluxMonitor("y", ny - 4, freqs[1:], interval=2),
            bz.FluxMonitor("z", 3, freqs, window="hann", t_start=2.5 * dt, t_end=9.5 * dt)]


# CPML layouts: every face; none; one face per axis (PEC opposite); x and y
# faces only (no z CPML memories at all)
BOUNDARIES = {"cpml": dict(cpml_cells=3), "pec": {},
              "onesided": dict(cpml_cells=3, faces=("left", "back", "top"),
                               metallic=("right", "front", "bottom")),
              "noz": dict(cpml_cells=2, faces=("left", "right", "front", "back"),
                          metallic=("bottom", "top"))}


@pytest.mark.parametrize("chunk", [12, 5])
@pytest.mark.parametrize("kz", ["1", "3", "16"])
@pytest.mark.parametrize("grid,bc", [((13, 17, 130), "cpml"), ((9, 28, 64), "pec"),
                                     ((11, 14, 197), "cpml"), ((13, 17, 130), "onesided"),
                                     ((11, 14, 197), "noz")])
def test_pair_parity(grid, bc, kz, chunk, monkeypatch):
    _require_gpu()
    from beamz_amd import native
    monkeypatch.setenv("BZH_PAIR_KZ", kz)
    steps = 12
    kw = BOUNDARIES[bc]
    sim, prog, state, prob, ost = _setup(grid, seed=71, steps=steps, sources=_sources(grid, steps),
                                         monitors=_monitors(grid), **kw)
    print(f"case pair_parity grid={grid} bc={bc} kz={kz} chunk={chunk}", flush=True)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        state = sim.advance(state, steps, chunk=chunk)
    torch.cuda.synchronize()
    assert native.last_program_fused()
    assert native.last_program_pairs() == (steps % chunk or chunk) // 2 // 2 * 2
    o.run(prob, ost, steps)
    _compare_state(prog, state, prob, ost, dft=True)


def test_pair_parity_uniform_medium():
    """Scalar E coefficients (uniform permittivity): the kernel variant
    without coefficient arrays."""
    _require_gpu()
    from beamz_amd import native
    grid = (10, 15, 140)
    steps = 8
    sim, prog, state, prob, ost = _setup(grid, seed=73, steps=steps, cpml_cells=4,
                                         sources=_sources(grid, steps), monitors=_monitors(grid),
                                         eps=np.full(grid, 2.25))
    print(f"case pair_parity_uniform grid={grid}", flush=True)
    state = sim.advance(state, steps, chunk=steps)
    torch.cuda.synchronize()
    assert native.last_program_pairs() == 4
    o.run(prob, ost, steps)
    _compare_state(prog, state, prob, ost, dft=True)


def test_pair_matches_single_steps(monkeypatch):
    """The same 16 steps with and without pairs: FP32 results agree to
    rounding (the arithmetic per update is the same; only contraction and
    summation order of sources may differ)."""
    _require_gpu()
    from beamz_amd import native
    grid = (24, 40, 250)
    steps = 16
    runs = []
    for pair in ("1", "0"):
        monkeypatch.setenv("BZH_PAIR", pair)
        sim, prog, state, prob, ost = _setup(grid, seed=75, steps=steps, cpml_cells=5,
                                             sources=_sources(grid, steps), monitors=_monitors(grid))
        print(f"case pair_vs_single grid={grid} pair={pair}", flush=True)
        state = sim.advance(state, steps, chunk=steps)
        torch.cuda.synchronize()
        assert native.last_program_pairs() == (8 if pair == "1" else 0)
        runs.append([t.cpu().numpy().copy() for t in state.fields()] +
                    [state.dft_re.cpu().numpy().copy(), state.dft_im.cpu().numpy().copy()])
    for i, (a, b) in enumerate(zip(*runs)):
        _close(f"leaf {i}", a, b, rel=1e-6, mabs=1e-5)


def _h_sources(grid, steps):
    """H-timing slabs on Hy (touching the bottom y face, where a PEC face
    masks it) and Hz, applied after the H update of both steps of a pair."""
    dt = _dt(grid)
    nz, ny, nx = grid
    rng = np.random.default_rng(13)
    t = o.sample_times(0.0, dt, steps)
    f0 = 1.9e14
    wave = o.gaussian_pulse(t, f0, 0.5 * f0, offset=1.0).astype(np.float32).astype(np.float64)
    hy = (rng.random((4, 3, nx - 10)) * 2e-3).astype(np.float32).astype(np.float64)
    hz = (rng.random((2, 5, 7)) * 2e-3).astype(np.float32).astype(np.float64)
    return [o.Source("Hy", "h", (nz // 2 - 2, 0, 5), hy, wave),
            o.Source("Hz", "h", (2, ny // 2, nx // 2 - 3), hz, wave)]


@pytest.mark.parametrize("chunk", [12, 5])
@pytest.mark.parametrize("grid,bc", [((13, 17, 130), "cpml"), ((9, 28, 64), "pec")])
def test_pair_parity_with_h_sources(grid, bc, chunk):
    """H-timing slabs run inside the pair and single-step kernels (the
    mode-source launch uses them): parity with the oracle as above."""
    _require_gpu()
    from beamz_amd import native
    steps = 12
    kw = {"cpml": dict(cpml_cells=3), "pec": {}}[bc]
    srcs = _sources(grid, steps) + _h_sources(grid, steps)
    sim, prog, state, prob, ost = _setup(grid, seed=77, steps=steps, sources=srcs,
                                         monitors=_monitors(grid), **kw)
    print(f"case pair_parity_h grid={grid} bc={bc} chunk={chunk}", flush=True)
    state = sim.advance(state, steps, chunk=chunk)
    torch.cuda.synchronize()
    assert native.last_program_fused()
    assert native.last_program_pairs() == (steps % chunk or chunk) // 2 // 2 * 2
    o.run(prob, ost, steps)
    _compare_state(prog, state, prob, ost, dft=True)


def _core_sources(grid, steps):
    """Slabs crossing the core tiles (x-interior, rows and planes clear of
    every face and slab): a pre-E Ez plane, an E-timed Ey block and an
    H-timed Hz block, all between the steps of every pair."""
    dt = _dt(grid)
    nz, ny, nx = grid
    rng = np.random.default_rng(21)
    t = o.sample_times(0.0, dt, steps)
    wave = o.gaussian_pulse(t, 2.0e14, 0.5e14, offset=1.0).astype(np.float32).astype(np.float64)
    c1 = (rng.random((nz - 16, ny - 20, 1)) * 1e-1).astype(np.float32).astype(np.float64)
    c2 = (rng.random((6, 9, 30)) * 1e-1).astype(np.float32).astype(np.float64)
    c3 = (rng.random((5, 7, 20)) * 1e-4).astype(np.float32).astype(np.float64)
    return [o.Source("Ez", "pre_e", (8, 10, 75), c1, wave),
            o.Source("Ey", "e", (nz // 2, ny // 2, 64), c2, wave),
            o.Source("Hz", "h", (nz // 2 - 3, ny // 2 + 2, 70), c3, wave)]


@pytest.mark.parametrize("bc", ["cpml", "pec", "onesided", "noz"])
def test_pair_core_parity(bc, monkeypatch):
    """Grids large enough for the core launch (x-interior tiles whose rows
    and chunks lie clear of every face and CPML slab, run by the lean core
    kernel with LDS-DMA prefetch): sources of every timing and an x-normal
    monitor inside the core, against the oracle and against the same run with
    the core folded back into the shell launch (BZH_PAIR_CORE=0)."""
    _require_gpu()
    import beamz_amd as bz
    from beamz_amd import native
    grid = (30, 44, 130) if bc == "pec" else (40, 48, 140)
    steps = 12
    kw = BOUNDARIES[bc]
    freqs = np.array([1.8e14, 2.0e14, 2.2e14])
    runs = []
    for core in ("1", "0"):
        monkeypatch.setenv("BZH_PAIR_CORE", core)
        mons = [bz.FluxMonitor("x", 90, freqs), bz.FluxMonitor("z", grid[0] // 2, freqs[:2])]
        sim, prog, state, prob, ost = _setup(grid, seed=73, steps=steps,
                                             sources=_core_sources(grid, steps), monitors=mons,
                                             **kw)
        print(f"case pair_core grid={grid} bc={bc} core={core}", flush=True)
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            state = sim.advance(state, steps, chunk=steps)
        torch.cuda.synchronize()
        assert native.last_program_fused() and native.last_program_pairs() == 6
        runs.append([t.cpu().numpy().astype(np.float64) for t in
                     (*state.fields(), *state.psi_h, *state.psi_e, state.dft_re, state.dft_im)])
        if core == "1":
            o.run(prob, ost, steps)
            _compare_state(prog, state, prob, ost, dft=True)
    for i, (a, b) in enumerate(zip(*runs)):
        n = np.linalg.norm(b)
        assert n == 0.0 or np.linalg.norm(a - b) / n <= 1e-6, (i, np.linalg.norm(a - b) / n)


def _capture_monitors(grid, layout):
    """Monitor sets whose cells no tile row shares between two monitors, so
    that the pair launches capture them densely: two x-normal planes in
    different 60-column tile columns, one z-normal plane with a placed range,
    one y-normal plane with a Hann window and an interval."""
    import beamz_amd as bz
    dt = _dt(grid)
    nz, ny, nx = grid
    freqs = np.array([1.7e14, 1.9e14, 2.1e14])
    if layout == "x2":
        return [bz.FluxMonitor("x", 20, freqs, ranges={"z": (2, nz - 2), "y": (3, ny - 3)}),
                bz.FluxMonitor("x", 100, freqs[1:])]
    if layout == "z":
        return [bz.FluxMonitor("z", nz // 2, freqs, ranges={"y": (2, ny - 4), "x": (7, nx - 9)})]
    return [bz.FluxMonitor("y", ny // 2, freqs, interval=2, window="hann", t_start=2.5 * dt,
                           t_end=9.5 * dt)]


@pytest.mark.parametrize("layout", ["x2", "z", "y"])
@pytest.mark.parametrize("post", [False, True])
def test_pair_capture_matches_bank_and_oracle(layout, post):
    """The pair launches' dense capture of the monitored cells (the host boxes
    passed as the program's trailing argument) against the FP64 oracle, and
    bit for bit against the same program without the boxes (the intermediate
    bank and a strided gather of the fields).  ``post`` adds an E-timed slab,
    after which the second step is sampled from the fields."""
    _require_gpu()
    from beamz_amd import native
    grid = (13, 17, 130)
    steps = 12
    srcs = _sources(grid, steps)
    if not post:
        srcs = srcs[:1]
    sim, prog, state0, prob, ost = _setup(grid, seed=73, steps=steps, sources=srcs,
                                          monitors=_capture_monitors(grid, layout), cpml_cells=3)
    print(f"case pair_capture layout={layout} post={post} grid={grid}", flush=True)
    assert prog.monitor_boxes is not None
    start = {k: getattr(state0, k).cpu().numpy().copy() for k in ("ex", "ey", "ez", "hx", "hy", "hz")}
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        state = sim.advance(state0, steps, chunk=12)
    torch.cuda.synchronize()
    assert native.last_program_pairs() == 6
    assert native.last_program_captured()
    o.run(prob, ost, steps)
    _compare_state(prog, state, prob, ost, dft=True)
    # the same program through the intermediate bank
    boxes = prog.monitor_boxes
    try:
        prog.monitor_boxes = None
        state_b = sim.initial_state(start)
        with torch.cuda.stream(s):
            state_b = sim.advance(state_b, steps, chunk=12)
        torch.cuda.synchronize()
        assert not native.last_program_captured()
    finally:
        prog.monitor_boxes = boxes
    for name in ("dft_re", "dft_im", "dft_w"):
        np.testing.assert_array_equal(getattr(state, name).cpu().numpy(),
                                      getattr(state_b, name).cpu().numpy(), err_msg=name)
    for c in ("ex", "ey", "ez", "hx", "hy", "hz"):
        np.testing.assert_array_equal(getattr(state, c).cpu().numpy(),
                                      getattr(state_b, c).cpu().numpy(), err_msg=c)
