"""The wide variants of the fused kernels: 64-bit offsets from the group
bases, taken when a field bank (or the E coefficient triple) spans 4 GiB or
more.

- On small grids, with the variants forced (``BZH_WIDE=1``), the single-step
  and pair launches match the FP64 oracle and the 32-bit-offset kernels (the
  same arithmetic: only the address formation differs).
- At 192 x 1024 x 1024 cells a float32 bank spans 4.8 GB, so the program takes
  the wide kernels by itself; a few steps from random fields agree with the
  per-phase kernels (an independent path) to FP32 rounding.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import fdtd64 as o
from tests.test_gpu_fullsize import _rel, _seed_fields
from tests.test_gpu_pair import BOUNDARIES, _monitors, _sources
from tests.test_gpu_parity import _compare_state, _require_gpu, _setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bc", ["cpml", "onesided"])
def test_wide_kernels_match_oracle_and_32bit_offsets(bc, monkeypatch):
    _require_gpu()
    from beamz_amd import native
    grid = (13, 17, 130)
    steps = 12
    runs = []
    for wide in ("0", "1"):
        monkeypatch.setenv("BZH_WIDE", wide)
        sim, prog, state, prob, ost = _setup(grid, seed=81, steps=steps,
                                             sources=_sources(grid, steps),
                                             monitors=_monitors(grid), **BOUNDARIES[bc])
        print(f"case wide grid={grid} bc={bc} wide={wide}", flush=True)
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            state = sim.advance(state, steps, chunk=5)   # pairs and single steps
        torch.cuda.synchronize()
        assert native.last_program_fused() and native.last_program_pairs() == 0
        runs.append([t.clone() for t in (*state.fields(), *state.psi_h, *state.psi_e,
                                          state.dft_re, state.dft_im)])
        if wide == "1":
            o.run(prob, ost, steps)
            _compare_state(prog, state, prob, ost, dft=True)
    for i, (a, b) in enumerate(zip(*runs)):
        if b.numel() == 0 or float(b.abs().max()) == 0.0:
            assert torch.equal(a, b), i
            continue
        assert _rel(a, b) <= 1e-6, (i, _rel(a, b))


def test_wide_bank_over_4gib_matches_per_phase():
    """A float32 bank of 4.8 GB (192 x 1024 x 1024 cells, 8-cell CPML, a
    random-permittivity block, an Ez plane source and a flux plane): six steps
    (two pair launches, two single steps) through the wide kernels against the
    per-phase kernels, fields, CPML memories and DFT arenas."""
    _require_gpu()
    import beamz_amd as bz
    from beamz_amd import _abi, native
    nz, ny, nx = 192, 1024, 1024
    dx = 25e-9
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(3.0))
    eps = np.full((nz, ny, nx), 2.085)
    rng = np.random.default_rng(5)
    blocks = 2.085 + (12.25 - 2.085) * rng.random((8, 32, 32))
    eps[64:128, 384:640, 384:640] = np.repeat(np.repeat(np.repeat(blocks, 8, 0), 8, 1), 8, 2)
    f0 = bz.LIGHT_SPEED / 1.55e-6
    L = 8
    src = bz.SoftSource("Ez", (L, L, 40), (nz - 2 * L, ny + 1 - 2 * L, 1),
                        bz.GaussianPulse(f0, 0.2 * f0, offset=1.5))
    freqs = bz.LIGHT_SPEED / np.linspace(1.45e-6, 1.65e-6, 5)
    mons = [bz.FluxMonitor("x", 700, freqs, ranges={"z": (L, nz - L), "y": (L, ny - L)})]
    steps = 6
    states = []
    for fused in (True, False):
        flags = _abi.DEFAULT_FLAGS | (_abi.FLAG_FUSED_STEP if fused else 0)
        sim = bz.Simulation(eps, dx, dt=dt, boundaries=[bz.PML(cells=L)], sources=[src],
                            monitors=mons, steps=steps, flags=flags)
        prog = sim.compile()
        state = sim.initial_state()
        bank = sum(t.numel() * t.element_size() for t in state.fields())
        print(f"case wide bank {bank / 1e9:.2f} GB fused={fused} grid={(nz, ny, nx)}", flush=True)
        assert bank >= 1 << 32
        _seed_fields(state, prog, 11, {"E": 1e-3, "H": 1e-3 / 376.730313668})
        state = sim.advance(state, steps, chunk=steps)
        torch.cuda.synchronize()
        assert native.last_program_fused() == fused
        if fused:
            assert native.last_program_pairs() == 2
        states.append(state)
    a, b = states
    for name, x, y in zip(("hx", "hy", "hz", "ex", "ey", "ez"), a.fields(), b.fields()):
        err = _rel(x, y)
        print(f"  {name}: wide fused vs per-phase relative L2 {err:.2e}", flush=True)
        assert err <= 1e-6, (name, err)
    for i, (x, y) in enumerate(zip(a.psi_h + a.psi_e, b.psi_h + b.psi_e)):
        assert _rel(x, y) <= 1e-6, (i, _rel(x, y))
    for x, y in ((a.dft_re, b.dft_re), (a.dft_im, b.dft_im)):
        assert _rel(x, y) <= 1e-6, _rel(x, y)
