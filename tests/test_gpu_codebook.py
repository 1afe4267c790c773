"""The reference compiler's packed coefficient-ID form of the E source
coefficients (``compile.py:146-187``): for a lossless program whose source
coefficient arrays take at most 256 values, the E decay slots carry each
component's float32 table and the source slots its cells' 8-bit IDs, four per
int32 word.  The C-ABI accepts that form and runs the per-phase kernels on it.
A few steps of a 3D and a 2D program in that form must give bitwise the fields
of the same program with dense coefficient arrays.  The dense path is pinned
against the FP64 oracle by the parity tests."""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest
import torch

from tests.test_gpu_fullsize import _seed_fields
from tests.test_gpu_parity import _require_gpu

pytestmark = pytest.mark.gpu


def _pack_ids(values):
    """Table and packed IDs of one coefficient array, in the layout of
    ``compile.py:146-168``: sorted unique float32 values, the flat (C-order)
    cell IDs padded to a multiple of four, ID 4w + b in byte b of word w."""
    arr = np.asarray(values, dtype=np.float32)
    table, ids = np.unique(arr, return_inverse=True)
    assert table.size <= 256
    flat = ids.astype(np.uint32).ravel()
    padded = np.zeros((flat.size + 3) & ~3, dtype=np.uint32)
    padded[:flat.size] = flat
    words = padded[0::4] | (padded[1::4] << 8) | (padded[2::4] << 16) | (padded[3::4] << 24)
    return table, words.view(np.int32)


def _codebook_program(prog):
    coeffs = dict(prog.coeffs)
    dev = prog.device
    for ax in ("x", "y", "z"):
        src = coeffs[f"e_source_{ax}"]
        if src.dim() == 0:
            continue
        assert float(coeffs[f"e_decay_{ax}"]) == 1.0   # lossless
        table, words = _pack_ids(src.cpu().numpy())
        coeffs[f"e_decay_{ax}"] = torch.as_tensor(table, device=dev)
        coeffs[f"e_source_{ax}"] = torch.as_tensor(words, device=dev)
    return dataclasses.replace(prog, coeffs=coeffs)


def _materials(shape, seed):
    rng = np.random.default_rng(seed)
    return np.asarray([2.085, 4.0, 12.25])[rng.integers(0, 3, size=shape)]


@pytest.mark.parametrize("dims", [3, 2])
def test_packed_coefficient_ids_match_dense_arrays(dims):
    _require_gpu()
    import beamz_amd as bz
    from beamz_amd import _abi, native, runtime
    dx = 40e-9
    grid = (12, 18, 40) if dims == 3 else (50, 70)
    eps = _materials(grid, 3)
    f0 = 2.0e14
    if dims == 3:
        src = bz.SoftSource("Ez", (3, 3, 10), (6, 12, 1), bz.GaussianPulse(f0, 0.5 * f0))
        mon = bz.FluxMonitor("x", 30, np.array([1.8e14, 2.0e14]))
    else:
        src = bz.SoftSource("Ez", (5, 12), (40, 1), bz.GaussianPulse(f0, 0.5 * f0))
        mon = bz.FluxMonitor("x", 55, np.array([1.8e14, 2.0e14]))
    sim = bz.Simulation(eps, dx, boundaries=[bz.PML(cells=3)], sources=[src], monitors=[mon],
                        steps=24, flags=_abi.DEFAULT_FLAGS)
    prog = sim.compile()
    packed = _codebook_program(prog)
    assert any(packed.coeffs[f"e_source_{a}"].dtype == torch.int32 for a in "xyz")
    runs = []
    for p in (prog, packed):
        print(f"case codebook dims={dims} grid={grid} packed={p is packed}", flush=True)
        state = sim.initial_state()
        _seed_fields(state, p, 5, {"E": 1e-3, "H": 1e-3 / 376.730313668})   # random fields
        state = runtime.advance(state, p, 24, chunk=8)
        torch.cuda.synchronize()
        assert not native.last_program_fused()
        runs.append([t.clone() for t in (*state.fields(), state.dft_re, state.dft_im)])
    for i, (a, b) in enumerate(zip(*runs)):
        assert torch.equal(a, b), i
        if b.numel():   # (2D programs carry empty Ex, Ey, Hz)
            assert float(b.abs().max()) > 0.0, i
