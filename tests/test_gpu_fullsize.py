"""Properties at the benchmark's full size, 1024 x 512 x 256 cells (16-cell
CPML, the random-eps slab, the Ez plane source and two 21-wavelength flux
planes), where the FP64 oracle is too slow to run:
- the fused step and the per-phase kernels (two independent kernel paths)
  agree to FP32 rounding after a few steps from random fields, DFT arenas
  included;
- the fused step is deterministic: the same steps in one chunk or in two
  chunks with the same launch schedule (4 + 2 steps run two pair launches
  and two single steps, like one call of 6) give bitwise the same fields and
  CPML memories (the DFT phases then differ by the float32 rounding of dt);
  the pair launches of the first six steps and the per-phase kernels agree to
  FP32 rounding;
- f16 storage with per-field scaling tracks the FP32 fields within its
  storage precision."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

STEPS = 6


def _require_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a GPU")
    from beamz_amd import native
    native.library()


def _rel(a, b):
    a = a.double()
    b = b.double()
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def _seed_fields(state, prog, seed, amp):
    """Random fields drawn on the device: E with standard deviation amp["E"],
    H with amp["H"] (physical units), stored divided by the program's
    storage scale of half-precision fields."""
    g = torch.Generator(device=state.ex.device)
    g.manual_seed(seed)
    scale = prog.field_scale or {"E": 1.0, "H": 1.0}
    for name, t in zip(("hx", "hy", "hz", "ex", "ey", "ez"), state.fields()):
        key = name[0].upper()
        v = torch.randn(t.shape, generator=g, device=t.device, dtype=torch.float32) * amp[key]
        t.copy_((v / scale[key]).to(t.dtype))


def _run(sim, chunks, seed=7, amp=None):
    state = sim.initial_state()
    _seed_fields(state, sim.compile(), seed, amp or {"E": 1e-3, "H": 1e-3 / 376.730313668})
    for n in chunks:
        state = sim.advance(state, n, chunk=n)
    torch.cuda.synchronize()
    return state


def test_fullsize_fused_matches_per_phase_and_is_deterministic():
    _require_gpu()
    import bench
    from beamz_amd import native
    print(f"case fullsize f32 1024x512x256 steps={STEPS}", flush=True)
    fused, _ = bench.build_workload("f32", monitors=True, fused=True)
    a = _run(fused, [STEPS])
    assert native.last_program_fused() and native.last_program_pairs() == 2
    b = _run(fused, [4, STEPS - 4])
    for x, y in zip(a.fields() + a.psi_h + a.psi_e, b.fields() + b.psi_h + b.psi_e):
        assert torch.equal(x, y), "fused step is not deterministic across chunkings"
    # DFT phases: within a call the sample time is t + (k+1) dt with the
    # call's float32 dt attribute (the ABI's dt:f32), between calls the host
    # advances t in float64, so a split call shifts the later samples' phases
    # by 2 |dt - dt_f32| 2 pi f, about 1e-8 rad here
    for x, y in ((a.dft_re, b.dft_re), (a.dft_im, b.dft_im)):
        err = float((x - y).abs().max() / y.abs().max())
        assert err <= 1e-7, err
    del b
    per_phase, _ = bench.build_workload("f32", monitors=True, fused=False)
    c = _run(per_phase, [STEPS])
    assert not native.last_program_fused()
    for name, x, y in zip(("hx", "hy", "hz", "ex", "ey", "ez"), a.fields(), c.fields()):
        err = _rel(x, y)
        print(f"  {name}: fused vs per-phase relative L2 {err:.2e}", flush=True)
        assert err <= 1e-6, (name, err)
    for i, (x, y) in enumerate(zip(a.psi_h + a.psi_e, c.psi_h + c.psi_e)):
        assert _rel(x, y) <= 1e-6, (i, _rel(x, y))
    for x, y in ((a.dft_re, c.dft_re), (a.dft_im, c.dft_im)):
        assert _rel(x, y) <= 1e-6


def test_fullsize_f16_storage_tracks_f32():
    _require_gpu()
    import bench
    print(f"case fullsize f16 vs f32 1024x512x256 steps={STEPS}", flush=True)
    # fields of order one in f16 storage units: physical amplitude = the
    # storage scale the f16 program chose
    sim, _ = bench.build_workload("f16", monitors=False, fused=True)
    scale = sim.compile().field_scale
    ref_sim, _ = bench.build_workload("f32", monitors=False, fused=True)
    ref = _run(ref_sim, [STEPS], amp=scale)
    want = [t.float() for t in ref.fields()]
    del ref, ref_sim
    torch.cuda.empty_cache()
    got = _run(sim, [STEPS], amp=scale)
    for name, x, y in zip(("hx", "hy", "hz", "ex", "ey", "ez"), got.fields(), want):
        err = _rel(x.float() * scale[name[0].upper()], y)
        print(f"  {name}: f16 vs f32 relative L2 {err:.2e}", flush=True)
        assert err <= 5e-3, (name, err)


_F32_FLUX = {}


def _f32_flux():
    import bench
    if "flux" not in _F32_FLUX:
        _F32_FLUX["flux"], _ = bench.c3_flux_spectra("f32", 3000)
    return _F32_FLUX["flux"]


@pytest.mark.parametrize("dtype", ["f16", "bf16"])
def test_fullsize_half_storage_flux_after_3000_steps(dtype):
    """The mixed-precision configuration in full: 1024 x 512 x 256, the
    random-eps slab, 16-cell CPML, the Gaussian plane source, flux planes at
    x = 64 and x = 320 over 21 wavelengths, 3000 steps, fields stored in f16
    or bf16 (per-field scaling, float32 curl arithmetic, float64 DFT).  The
    flux spectrum of each monitor and the transmission between them must
    match the FP32 run (itself within ~1e-6 of the FP64 oracle on the
    configurations the oracle can run) to 1e-3."""
    _require_gpu()
    import bench
    print(f"case fullsize {dtype} flux 1024x512x256 steps=3000", flush=True)
    ref = _f32_flux()
    assert all(np.max(np.abs(f)) > 0 for f in ref)
    flux, ms = bench.c3_flux_spectra(dtype, 3000)
    errs = bench.flux_errors(flux, ref)
    print(f"  {dtype}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()) +
          f" ({ms:.3f} ms/step)", flush=True)
    bad = {k: v for k, v in errs.items() if not v <= 1e-3}
    assert not bad, f"{dtype} flux errors above 1e-3: {bad}"
