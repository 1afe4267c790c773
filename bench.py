"""Benchmark: GCUPS and HBM-roofline fraction of the 3D Yee+CPML step on one GPU.

Workload (the 3D-at-scale configuration): a 1024 x 512 x 256 (x, y, z) grid at
25 nm (134,217,728 cells); background eps 2.085 with a seeded inner
512 x 256 x 64 slab of 8x8x8-cell blocks, eps iid U[2.085, 12.25] (x from 256
to 768); 16-cell CPML on all faces; a soft Ez plane source (Gaussian pulse at
1.55 um, 0.2 um bandwidth, peak at 1.5 / fwidth so that it has crossed both
monitors within the 3000 steps) at x = 40; two 21-wavelength (1.45-1.65 um)
DFT flux planes with FP64 accumulators, at x = 64 (incident) and x = 320 (64
cells into the random slab).  A "step" is one full leapfrog step (sources, H,
E, DFT) over the whole grid.  Inputs are resident in HBM before timing starts.

The FP32 line also carries a ``half`` object: the f16 and bf16 storage forms
timed the same way, their GCUPS ratio to FP32, and the error of their flux
spectra (both monitors, and the transmission between them) against the FP32
run after the full 3000 steps.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--dtype f32|bf16|f16]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


WORKLOADS = {
    "3d": "3D Yee+CPML 1024x512x256 (x,y,z) at 25 nm, 16-cell CPML, seeded random-eps 512x256x64 "
          "slab, Ez plane source, 2 x 21-wavelength FP64 DFT flux planes (x=64, x=320)",
    "waveguide3d": "3D silicon waveguide 512x256x128 (x,y,z) at 25 nm, 450x220 nm Si core in SiO2 "
                   "along x, 12-cell CPML, 1.55 um Gaussian-pulse Ez plane source at x=30, "
                   "3 x 11-wavelength FP64 DFT flux planes",
}


def build_waveguide_workload(dtype_name, monitors=True, fused=True):
    """The 512x256x128 silicon strip: eps 12.25 core of 9 x 18 cells
    (220 x 450 nm) centred in eps 2.085, running along x."""
    import torch
    import beamz_amd as bz
    from beamz_amd import _abi

    nz, ny, nx, dx, L = 128, 256, 512, 25e-9, 12
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(3.0))
    eps = np.full((nz, ny, nx), 2.085)
    eps[59:68, 119:137, :] = 12.25
    f0 = bz.LIGHT_SPEED / 1.55e-6
    fw = bz.LIGHT_SPEED * 0.2e-6 / 1.55e-6 ** 2
    src = bz.SoftSource("Ez", (L, L, 30), (nz - 2 * L, ny + 1 - 2 * L, 1), bz.GaussianPulse(f0, fw))
    freqs = bz.LIGHT_SPEED / np.linspace(1.5e-6, 1.6e-6, 11)
    rngs = {"z": (L, nz - L), "y": (L, ny - L)}
    mons = [bz.FluxMonitor("x", xi, freqs, ranges=rngs) for xi in (60, 256, 450)]
    field_dtype = {"f32": torch.float32, "bf16": torch.bfloat16, "f16": torch.float16}[dtype_name]
    flags = _abi.DEFAULT_FLAGS | (_abi.FLAG_FUSED_STEP if fused else 0)
    sim = bz.Simulation(eps, dx, dt=dt, boundaries=[bz.PML(cells=L)], sources=[src],
                        monitors=mons if monitors else [], steps=4000, field_dtype=field_dtype,
                        flags=flags)
    return sim, eps


def build_workload(dtype_name, monitors=True, fused=True):
    import torch
    import beamz_amd as bz
    from beamz_amd import _abi

    nz, ny, nx = 256, 512, 1024
    dx = 25e-9
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(3.0))
    eps = np.full((nz, ny, nx), 2.085, dtype=np.float64)
    rng = np.random.default_rng(20260503)
    blocks = 2.085 + (12.25 - 2.085) * rng.random((64 // 8, 256 // 8, 512 // 8))
    z0, y0, x0 = (nz - 64) // 2, (ny - 256) // 2, (nx - 512) // 2
    eps[z0:z0 + 64, y0:y0 + 256, x0:x0 + 512] = np.repeat(np.repeat(np.repeat(
        blocks, 8, axis=0), 8, axis=1), 8, axis=2)
    f0 = bz.LIGHT_SPEED / 1.55e-6
    fw = bz.LIGHT_SPEED * 0.2e-6 / (1.55e-6) ** 2
    L = 16
    src = bz.SoftSource("Ez", (L, L, 40), (nz - 2 * L, ny + 1 - 2 * L, 1),
                        bz.GaussianPulse(f0, fw, offset=1.5))
    lam = np.linspace(1.45e-6, 1.65e-6, 21)
    freqs = bz.LIGHT_SPEED / lam
    rngs = {"z": (L, nz - L), "y": (L, ny - L)}
    mons = [bz.FluxMonitor("x", 64, freqs, ranges=rngs), bz.FluxMonitor("x", 320, freqs, ranges=rngs)]
    field_dtype = {"f32": torch.float32, "bf16": torch.bfloat16, "f16": torch.float16}[dtype_name]
    flags = _abi.DEFAULT_FLAGS | (_abi.FLAG_FUSED_STEP if fused else 0)
    sim = bz.Simulation(eps, dx, dt=dt, boundaries=[bz.PML(cells=L)], sources=[src],
                        monitors=mons if monitors else [], steps=3000, field_dtype=field_dtype,
                        flags=flags)
    return sim, eps


def c3_flux_spectra(dtype_name, steps=3000, chunk=256, sim=None):
    """Run the 1024x512x256 workload (``sim``, or a new one with
    ``dtype_name`` field storage) for ``steps`` steps from zero fields;
    returns (flux spectrum per monitor, ms per step)."""
    import torch
    import beamz_amd as bz
    if sim is None:
        sim, _ = build_workload(dtype_name)
    prog = sim.compile()
    state = sim.initial_state()
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):
        start.record(stream)
        state = sim.advance(state, steps, chunk=chunk)
        stop.record(stream)
    torch.cuda.synchronize()
    res = bz.SimulationResults(state, prog, None)
    flux = [np.asarray(res.monitor_flux(i), dtype=np.float64) for i in range(len(prog.monitors))]
    ms = start.elapsed_time(stop) / steps
    del state, res
    torch.cuda.empty_cache()
    return flux, ms


def flux_errors(flux, ref):
    """Max over the wavelengths of |a - b| / max|b| for each monitor, and of
    the transmission T = flux[1] / flux[0]."""
    out = {f"flux{i}": float(np.max(np.abs(a - b)) / np.max(np.abs(b))) for i, (a, b) in
           enumerate(zip(flux, ref))}
    t, tr = flux[1] / flux[0], ref[1] / ref[0]
    out["transmission"] = float(np.max(np.abs(t - tr)) / np.max(np.abs(tr)))
    return out


def build_batched_workload(seed=4, count=64, steps=3000):
    """The batched 2D sweep: ``count`` TMxy simulations, each (ny, nx) drawn
    uniformly from {256..1024}^2 at 20 nm, with 2-4 x-directed waveguides of
    width U[400, 600] nm (eps 12.25 in 2.085), 10-cell CPML on every edge and
    a soft Ez line source (Gaussian pulse at 1.55 um, 0.2 um bandwidth) across
    every guide at x = 40 cells."""
    import beamz_amd as bz
    rng = np.random.default_rng(seed)
    dx = 20e-9
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(2.0))
    f0 = bz.LIGHT_SPEED / 1.55e-6
    fw = bz.LIGHT_SPEED * 0.2e-6 / (1.55e-6) ** 2
    sims = []
    for _ in range(count):
        ny, nx = (int(v) for v in rng.integers(256, 1025, size=2))
        eps = np.full((ny, nx), 2.085)
        nguides = int(rng.integers(2, 5))
        centres = np.linspace(0, ny, nguides + 2)[1:-1]
        srcs = []
        for c in centres:
            w = int(round(rng.uniform(400e-9, 600e-9) / dx))
            y0 = int(c) - w // 2
            eps[y0:y0 + w, :] = 12.25
            srcs.append(bz.SoftSource("Ez", (y0, 40), (w, 1), bz.GaussianPulse(f0, fw)))
        sims.append(bz.Simulation(eps, dx, dt=dt, boundaries=[bz.PML(cells=10)], sources=srcs[:4],
                                  steps=steps))
    return sims


def run_batched(args):
    import torch
    from beamz_amd.batch import BatchedSimulation2D, autotune
    print("[bench] building the batched 2D workload", file=sys.stderr, flush=True)
    batch = BatchedSimulation2D(build_batched_workload())
    state = batch.initial_state()
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        if args.tile_rows and args.workgroups:
            batch.tile_rows, batch.workgroups = tr, wg = args.tile_rows, args.workgroups
            timings = {}
        else:
            tr, wg, timings = autotune(batch, steps=8, verbose=True)
        batch.advance(state, args.warmup, chunk=args.chunk)
        # the timed window replays captured graphs: its chunk sequence runs
        # once beforehand (a chunk's graph is keyed on its step count)
        batch.advance(state, args.steps, chunk=args.chunk)
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        batch.advance(state, args.steps, chunk=args.chunk)
        stop.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / args.steps
    gcups = batch.cells / (ms * 1e-3) / 1e9
    line = {
        "metric": "GCUPS", "value": round(gcups, 3), "unit": "GCUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched 2D TMxy sweep: 64 simulations, (ny, nx) ~ U{256..1024}^2 at "
                               "20 nm, 2-4 waveguides, 10-cell CPML, Ez line sources; one persistent "
                               "launch per step, hipGraph chunks",
                   "cells": batch.cells, "simulations": len(batch.progs),
                   "tile_rows": tr, "workgroups": wg},
        "autotune_ms_per_step": {f"{k[0]}x{k[1]}": round(v, 4) for k, v in timings.items()},
    }
    print(json.dumps(line), flush=True)


def build_adjoint_problem(steps=4000, checkpoint=200, seed=5):
    """The 2D inverse-design bend: 600 x 600 cells at 20 nm, a 100 x 100
    design box of density rho ~ U[0, 1] (eps = 2.085 + 10.165 rho) at the
    centre, a 500 nm input guide along x from the left edge into the box and a
    500 nm output guide along y from the box to the top edge; 10-cell CPML;
    the fundamental mode of the input guide launched along +x at x = 15 over
    1.2 um (a mode source), Gaussian pulse at 1.55 um with a 0.3 um bandwidth
    peaking at 2 / fwidth, so that the pulse has crossed the output monitor (a
    line across the output guide at y = 400) well within the 4000 steps; the
    objective is the overlap power of Ez on that line with the output guide's
    fundamental mode, at three wavelengths around 1.55 um."""
    import beamz_amd as bz
    from beamz_amd.adjoint import AdjointProblem2D
    from beamz_amd.modes import solve_line_modes
    n, dx, L = 600, 20e-9, 10
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(2.0))
    base = np.full((n, n), 2.085)
    w = 25
    c0 = n // 2 - w // 2                      # guide rows / columns [c0, c0 + w)
    design = (250, 350, 250, 350)
    base[c0:c0 + w, :design[2]] = 12.25       # input guide: y rows, x from 0 to the box
    base[design[1]:, c0:c0 + w] = 12.25       # output guide: x columns, y from the box up
    f0 = bz.LIGHT_SPEED / 1.55e-6
    fw = bz.LIGHT_SPEED * 0.3e-6 / (1.55e-6) ** 2
    src = bz.ModeSource((15.5 * dx, (c0 + w / 2) * dx, 0.0), (0.0, 60 * dx, 0.0),
                        bz.GaussianPulse(f0, fw, offset=2.0), "+")
    freqs = bz.LIGHT_SPEED / np.array([1.5e-6, 1.55e-6, 1.6e-6])
    mon = bz.FluxMonitor("y", 400, freqs, ranges={"x": (c0, c0 + w)}, components=("Ez",))
    # overlap weights: the output guide's fundamental Ez profile across the line
    line = np.concatenate([base[400, :], base[400, -1:]])
    _, ez = solve_line_modes(line, 2 * np.pi * f0, dx, 3, 0.98 * np.sqrt(12.25))
    mode = ez[0][c0:c0 + w] * np.sign(ez[0][c0 + w // 2])
    rho = np.random.default_rng(seed).uniform(0.0, 1.0, (100, 100))
    prob = AdjointProblem2D(base, design, eps_min=2.085, eps_max=12.25, resolution=dx, dt=dt,
                            steps=steps, sources=[src], monitor=mon, mode=mode, pml_cells=L,
                            checkpoint=checkpoint)
    return prob, rho


def run_adjoint(args):
    import torch
    from beamz_amd.adjoint import objective
    print("[bench] building the adjoint bend problem", file=sys.stderr, flush=True)
    prob, rho0 = build_adjoint_problem()
    prob.objective(rho0)                                  # warm: compile, capture graphs
    torch.cuda.synchronize()
    # the stepping alone: one compiled program, 4000 steps in checkpoint-sized calls
    from beamz_amd import runtime
    sim = prob.simulation(rho0)
    prog = sim.compile()
    st = sim.initial_state()
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for _ in range(0, prob.steps, prob.checkpoint):
        st = runtime.advance(st, prog, prob.checkpoint)
    torch.cuda.synchronize()
    stepping_s = time.perf_counter() - ts
    del st
    t0 = time.perf_counter()
    rho = torch.tensor(rho0, dtype=torch.float64, requires_grad=True)
    J = objective(rho, prob)
    t1 = time.perf_counter()
    J.backward()
    t2 = time.perf_counter()
    g = rho.grad.numpy()
    rng = np.random.default_rng(17)
    pix = [tuple(p) for p in rng.integers(0, 100, size=(32, 2))]
    got = np.asarray([g[a, b] for a, b in pix])

    def central_fd(h):
        fd = []
        for a, b in pix:
            rp, rm = rho0.copy(), rho0.copy()
            rp[a, b] += h
            rm[a, b] -= h
            fd.append((prob.objective(rp) - prob.objective(rm)) / (2 * h))
        fd = np.asarray(fd)
        return np.abs(got - fd) / np.max(np.abs(fd))

    # The forward runs in FP32: an h = 1e-3 difference of J carries its
    # rounding (about 1e-3 of max |dJ/drho| at the median pixel here), while
    # h = 1e-2 keeps both that and the O(h^2) truncation below 1e-3.  Both are
    # reported; the value is the h = 1e-2 check.
    rel_fine = central_fd(1e-3)
    rel = central_fd(1e-2)
    cells = prob.base_eps.size
    line = {
        "metric": "adjoint gradient vs central FD", "value": float(rel.max()), "unit": "max relative error",
        "n_gpus": 1, "steps": prob.steps, "warmup": 1, "higher_is_better": False, "data": "synthetic",
        "dtype": "f32",
        "config": {"workload": "2D adjoint bend 600x600 at 20 nm, 100x100 design, mode source "
                               "at 1.55 um, 4000 steps, checkpoint every 200, central FD (FP32 "
                               "forward) h=1e-2 on 32 random pixels; h=1e-3 reported beside it",
                   "J": float(J), "forward_s": round(t1 - t0, 3), "backward_s": round(t2 - t1, 3),
                   "forward_gcups": round(cells * prob.steps / (t1 - t0) / 1e9, 3),
                   "forward_stepping_s": round(stepping_s, 4),
                   "stepping_us_per_step": round(stepping_s / prob.steps * 1e6, 2),
                   "fd_median_rel": float(np.median(rel)),
                   "fd_h1e-3_max_rel": float(rel_fine.max()),
                   "fd_h1e-3_median_rel": float(np.median(rel_fine))},
    }
    print(json.dumps(line), flush=True)


def algorithmic_bytes(prog, state):
    """Compulsory HBM bytes of one full step: every field, coefficient array and
    CPML memory element read once and written once (coefficients only read)."""
    fb = state.ex.element_size()
    fields = sum(t.numel() for t in state.fields())
    coefs = sum(v.numel() * v.element_size() for v in prog.coeffs.values() if v.dim() > 0)
    psi = sum(t.numel() * t.element_size() for t in state.psi_h + state.psi_e)
    return 2 * fields * fb + coefs + 2 * psi


PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")


def measured_traffic(kernel, dtype):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 PMC
    summary (two passes, FETCH_SIZE and WRITE_SIZE, scaled by the counters'
    readings of a plain copy of known size), or None when no summary for this
    kernel and storage type exists.  The summary carries the line
    ``traffic: kernel=<name> dtype=<dtype> bytes_per_launch=<int>``
    (``bytes_per_step`` for a kernel that advances one step per launch)."""
    if not os.path.isdir(PROFILES):
        return None
    for name in sorted(os.listdir(PROFILES)):
        if not name.endswith(".md"):
            continue
        with open(os.path.join(PROFILES, name)) as fh:
            for line in fh:
                line = line.strip().strip("`")
                if not line.startswith("traffic:"):
                    continue
                fields = dict(kv.split("=", 1) for kv in line.split()[1:])
                if fields.get("kernel") == kernel and fields.get("dtype") == dtype:
                    return int(fields.get("bytes_per_launch", fields.get("bytes_per_step")))
    return None


def cpu_baseline(eps, prog, budget_s=12.0):
    """Time the C/OpenMP FP32 restatement (oracle/port.c) on the full grid."""
    from oracle import port
    if not os.path.exists(port.LIB):
        port.build()
    grid = prog.grid
    terms = []
    for t in prog.cpml.h_terms + prog.cpml.e_terms:
        shape = list(prog.shapes[t.component])
        shape[t.axis] = t.low + t.high
        terms.append((t.a, t.b, t.inv_kappa, tuple(shape)))
    e_src = [prog.coeffs[f"e_source_{a}"].cpu().numpy() for a in "xyz"]
    e_src = [np.broadcast_to(v, prog.shapes["E" + a]).copy() if v.ndim == 0 else v
             for v, a in zip(e_src, "xyz")]
    pg = port.PortGrid(grid, 1.0 / prog.resolution, prog.dt / 1.25663706212e-6, e_src,
                       L=prog.cpml.cells, terms=terms, metallic=0, seed=1)
    g1, t1 = port.time_gcups(pg, prog.cells, 1)
    steps = max(2, int(budget_s / max(t1, 1e-3)))
    gcups, tt = port.time_gcups(pg, prog.cells, steps)
    return {"value": round(gcups, 4), "unit": "GCUPS", "cores": pg.threads(), "kind": "port",
            "sample": f"full {grid[2]}x{grid[1]}x{grid[0]} grid with {prog.cpml.cells}-cell CPML, "
                      f"{steps} FP32 steps ({tt:.1f} s) of oracle/port.c (H+E+CPML; no source/DFT)"}


def time_workload(args, build, dtype):
    """Build the workload with ``dtype`` field storage and time ``args.steps``
    steps after ``args.warmup`` warm-up steps; returns (sim, eps, prog,
    state, stream, ms per step)."""
    import torch
    from beamz_amd import runtime

    print(f"[bench] building workload dtype={dtype}", file=sys.stderr, flush=True)
    sim, eps = build(dtype, monitors=not args.no_monitors, fused=not args.per_phase)
    prog = sim.compile()
    state = sim.initial_state()
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    print("[bench] workload resident; warming up", file=sys.stderr, flush=True)

    step = 0

    def run(n):
        nonlocal step, state
        remaining = n
        while remaining > 0:
            k = min(args.chunk, remaining)
            state = runtime.run_chunk(state, prog, k)
            step += k
            remaining -= k
            state.current_step.fill_(step)
            state.t.fill_(prog.t0 + step * prog.dt)

    with torch.cuda.stream(stream):
        run(args.warmup)
        # Replay-only timing: a chunk's graph is keyed on its step count and the
        # bank it starts from, so the exact chunk sequence of the timed window
        # runs once beforehand from the same bank (twice for an odd count,
        # which swaps the banks), and the timed window replays captured graphs.
        for _ in range(2 if args.steps % 2 else 1):
            run(args.steps)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):
        start.record(stream)
        run(args.steps)
        stop.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / args.steps
    print(f"[bench] {dtype}: {ms:.3f} ms/step, {prog.cells / (ms * 1e-3) / 1e9:.2f} GCUPS",
          file=sys.stderr, flush=True)
    return sim, eps, prog, state, stream, ms


def half_leg(args, gcups_f32, sim_f32):
    """The ``half`` object of the FP32 line: f16 and bf16 storage timed like
    FP32, and their 3000-step flux errors against the FP32 run."""
    import torch
    out = {}
    steps = 3000
    print(f"[bench] half-precision accuracy: {steps} steps each of f32, f16, bf16",
          file=sys.stderr, flush=True)
    ref, _ = c3_flux_spectra("f32", steps, sim=sim_f32)
    for dt_name in ("f16", "bf16"):
        sim, eps, prog, state, stream, ms = time_workload(args, build_workload, dt_name)
        g = prog.cells / (ms * 1e-3) / 1e9
        out[dt_name] = {"gcups": round(g, 3), "ms_per_step": round(ms, 4),
                        "vs_f32": round(g / gcups_f32, 4)}
        del state
        torch.cuda.empty_cache()
        flux, _ = c3_flux_spectra(dt_name, steps, sim=sim)
        out[dt_name]["flux_error_vs_f32"] = {k: float(f"{v:.3e}") for k, v in
                                            flux_errors(flux, ref).items()}
        del sim, eps, prog
        torch.cuda.empty_cache()
    out["accuracy_run"] = (f"{steps} steps of this workload; max over 21 wavelengths of "
                           "|flux - flux_f32| / max|flux_f32| per monitor and of the "
                           "transmission flux(x=320) / flux(x=64)")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--dtype", default="f32", choices=("f32", "bf16", "f16"))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-half", action="store_true",
                    help="skip the f16/bf16 timing and accuracy legs of the FP32 line")
    ap.add_argument("--no-monitors", action="store_true")
    ap.add_argument("--per-phase", action="store_true", help="use the per-phase kernels")
    ap.add_argument("--chunk", type=int, default=256)
    ap.add_argument("--tile-rows", type=int, default=0,
                    help="batched2d: item height (with --workgroups, skips the autotuner)")
    ap.add_argument("--workgroups", type=int, default=0, help="batched2d: grid size")
    ap.add_argument("--workload", default="3d", choices=("3d", "waveguide3d", "batched2d", "adjoint2d"),
                    help="3d: the 1024x512x256 Yee+CPML step (default); waveguide3d: the 512x256x128 "
                         "silicon strip; batched2d: the 64-simulation "
                         "2D sweep in one persistent launch per step; adjoint2d: the 600x600 bend "
                         "gradient and its finite-difference check")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device-to-device copy (the byte-counter calibration "
                         "kernel for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 1 << 28  # 2^28 float32 = 1 GiB
        src = torch.ones(n, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(10):
            dst.copy_(src)
        t1.record()
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / 10
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 13,
                          "copy_ms": round(ms, 4), "copy_GBps": round(8 * n / ms / 1e6, 1)}))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path; the value "
              "is one GPU's throughput (the path would shard as z-slabs, one per GPU)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batched2d":
        run_batched(args)
        return
    if args.workload == "adjoint2d":
        run_adjoint(args)
        return

    import torch
    from beamz_amd import native, runtime

    build = build_waveguide_workload if args.workload == "waveguide3d" else build_workload
    sim, eps, prog, state, stream, ms = time_workload(args, build, args.dtype)
    gcups = prog.cells / (ms * 1e-3) / 1e9
    if runtime.uses_two_banks(prog) and not native.last_program_fused():
        raise RuntimeError("the timed steps did not run the fused step kernels")

    pairs = native.last_program_pairs() > 0
    # dominant kernel: the step update launched alone (no sources/DFT), timed
    # with events on its own stream.  One launch advances one full step (the
    # x-interior plus x-edge kernels), or two where the pair kernel runs;
    # algorithmic bytes per launch = that many full steps
    per_launch = 2 if pairs else 1
    step_bytes = algorithmic_bytes(prog, state)
    nb = step_bytes * per_launch
    k_ms = runtime.time_step_kernel(state, prog, launches=20, stream=stream) * per_launch
    achieved = nb / (k_ms * 1e-3) / 1e9
    peak = 8000.0
    line = {
        "metric": "GCUPS",
        "value": round(gcups, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {
            "workload": WORKLOADS[args.workload],
            "cells": prog.cells,
            "field_storage": args.dtype,
            "schedule": "per-phase" if args.per_phase else (
                "fused two-bank, two steps per launch" if pairs else "fused two-bank"),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": peak,
            "unit": "GB/s",
            "frac": round(achieved / peak, 4),
            # the whole step (sources, DFT, every launch) against the same
            # algorithmic bytes per step
            "frac_step": round(step_bytes / (ms * 1e-3) / 1e9 / peak, 4),
            # the committed rocprofv3 byte counts cover the 1024x512x256 workload
            "traffic": (measured_traffic(runtime.step_kernel_name(prog, pairs), args.dtype)
                        if args.workload == "3d" else None),
            "kernel": runtime.step_kernel_name(prog, pairs),
            "kernel_ms": round(k_ms, 4),
            "algorithmic_bytes_per_launch": int(nb),
        },
    }
    if args.workload == "3d" and args.dtype == "f32" and not args.no_half and \
            not args.no_monitors and not args.per_phase:
        del state
        torch.cuda.empty_cache()
        line["half"] = half_leg(args, gcups, sim)
    if not args.no_cpu:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(eps, prog)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
