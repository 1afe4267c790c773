"""Adjoint gradient of the 2D mode-overlap objective against central finite
differences of the FP64 oracle."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import fdtd64 as o

pytestmark = pytest.mark.gpu


def _require_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a GPU")
    from beamz_amd import native
    native.library()


def _problem(steps=120, checkpoint=50):
    import beamz_amd as bz
    from beamz_amd.adjoint import AdjointProblem2D
    ny, nx, dx, L = 48, 64, 20e-9, 6
    dt = 0.99 * dx / (o.LIGHT_SPEED * np.sqrt(2))
    base = np.full((ny, nx), 2.085)
    base[20:26, :30] = 12.25                      # input guide
    base[26:, 40:46] = 12.25                      # output guide
    t = o.sample_times(0.0, dt, steps)
    wave = np.sin(2 * np.pi * 1.93e14 * t) * np.exp(-(((t - 30 * dt) / (12 * dt)) ** 2))
    wave = wave.astype(np.float32).astype(np.float64)
    coeff = np.full((10, 1), -0.2)
    src = bz.CustomSource("Ez", (18, 10), coeff, wave, "pre_e")
    freqs = np.array([1.85e14, 1.93e14])
    mon = bz.FluxMonitor("y", 40, freqs, ranges={"x": (34, 52)}, components=("Ez",))
    rng = np.random.default_rng(3)
    design = (16, 30, 24, 38)
    mode = rng.uniform(0.2, 1.0, 18)
    prob = AdjointProblem2D(base, design, eps_min=2.085, eps_max=12.25, resolution=dx, dt=dt,
                            steps=steps, sources=[src], monitor=mon, mode=mode, pml_cells=L,
                            checkpoint=checkpoint)
    return prob, (src, freqs, mode, coeff, wave, L, dx, dt, steps)


def _oracle_J(prob, extra, rho):
    src, freqs, mode, coeff, wave, L, dx, dt, steps = extra
    eps = prob.permittivity(rho)
    ny, nx = eps.shape
    shapes = o.shapes_tm((ny, nx))
    plans = {c: None for c in o.COMPONENTS}
    plans["Ez"] = o.line_plan_tm("Ez", shapes["Ez"], dx, "y", 40, (34, 52))
    mon = o.Monitor(freqs, plans, "y", 1.0, dx)
    spec = o.CpmlSpec(cells=L, kappa_max=2.0, faces=frozenset(("bottom", "top", "left", "right")))
    problem = o.build_problem(eps, dx, dt, cpml=spec,
                              sources=[o.Source("Ez", "pre_e", (18, 10), coeff, wave)],
                              monitors=[mon])
    st = o.init_state(problem)
    o.run(problem, st, steps)
    c = o.COMPONENTS.index("Ez")
    phasor = st["dft_re"][0][c] + 1j * st["dft_im"][0][c]
    O = (phasor * (mode * dx)[None, :]).sum(axis=1)
    return float(np.sum(np.abs(O) ** 2))


def test_adjoint_gradient_matches_oracle_finite_differences():
    _require_gpu()
    from beamz_amd.adjoint import objective
    prob, extra = _problem()
    rng = np.random.default_rng(11)
    rho0 = rng.uniform(0.0, 1.0, (14, 14))
    rho = torch.tensor(rho0, dtype=torch.float64, requires_grad=True)
    print("case adjoint grid=(48, 64) design=14x14 steps=120 checkpoint=50", flush=True)
    J = objective(rho, prob)
    J.backward()
    g = rho.grad.numpy()
    J_ref = _oracle_J(prob, extra, rho0)
    Jv = float(J.detach())
    assert abs(Jv - J_ref) <= 1e-4 * J_ref, (Jv, J_ref)
    h = 1e-3
    pix = [tuple(p) for p in rng.integers(0, 14, size=(6, 2))]
    fd = []
    for (a, b) in pix:
        rp, rm = rho0.copy(), rho0.copy()
        rp[a, b] += h
        rm[a, b] -= h
        fd.append((_oracle_J(prob, extra, rp) - _oracle_J(prob, extra, rm)) / (2 * h))
    fd = np.array(fd)
    got = np.array([g[a, b] for a, b in pix])
    err = np.abs(got - fd) / np.max(np.abs(fd))
    print("adjoint vs FD:", list(zip(got.round(12), fd.round(12))), "max rel", err.max(), flush=True)
    assert err.max() <= 1e-3


def test_adjoint_bend_600x600_matches_fp64_finite_differences():
    """The inverse-design bend at its full size: 600 x 600 cells at 20 nm, a
    100 x 100 design box, the input guide's mode launched by a mode source,
    4000 steps, checkpoints every 200.  The FP32 forward plus hand-written
    adjoint gradient matches central differences (h = 1e-3) of the FP64
    oracle on 32 random pixels to 1e-3 of max |FD|, and J matches the
    oracle's.  The FD values come from tests/golden/make_adjoint_bend_fd.py
    (65 oracle runs on the CPU)."""
    _require_gpu()
    import os

    import bench
    from beamz_amd.adjoint import objective
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "adjoint_bend_600x600_fd.npz"))
    prob, rho0 = bench.build_adjoint_problem()
    assert (prob.steps, prob.checkpoint) == (4000, 200)
    print("case adjoint bend 600x600 design=100x100 steps=4000 checkpoint=200", flush=True)
    rho = torch.tensor(rho0, dtype=torch.float64, requires_grad=True)
    J = objective(rho, prob)
    J.backward()
    g = rho.grad.numpy()
    J_ref = float(gold["J0"])
    rel_J = abs(float(J.detach()) - J_ref) / J_ref
    pix = gold["pixels"]
    fd = gold["fd"]
    got = np.array([g[a, b] for a, b in pix])
    err = np.abs(got - fd) / np.max(np.abs(fd))
    print(f"adjoint bend: J rel err {rel_J:.2e}, gradient vs FP64 FD max {err.max():.2e} "
          f"median {np.median(err):.2e} (h = {float(gold['h'])})", flush=True)
    assert rel_J <= 1e-4
    assert err.max() <= 1e-3


def test_prepared_recompute_gives_the_same_gradient(monkeypatch):
    """The reverse pass recomputes each checkpoint segment one step per call
    through a program call marshalled once (``runtime.prepare_chunk``).  Its
    gradient must equal, bit for bit, the one from a fresh ``run_chunk`` call
    per step (the same native calls, built anew each time), over three
    segments, twice in a row."""
    _require_gpu()
    from beamz_amd import runtime
    prob, _ = _problem(steps=120, checkpoint=50)
    rho = np.random.default_rng(11).uniform(0.0, 1.0, (14, 14))
    _, fwd = prob.forward(rho)
    print("case adjoint prepared recompute 48x64, 120 steps, checkpoint 50", flush=True)
    prepared = [prob.gradient(fwd), prob.gradient(fwd)]
    monkeypatch.setattr(runtime, "prepare_chunk", lambda state, prog, nsteps: None)
    fresh = prob.gradient(fwd)
    assert np.abs(fresh).max() > 0.0
    for g in prepared:
        assert np.array_equal(g, fresh)
