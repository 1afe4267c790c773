"""Benchmark: GCUPS and HBM-roofline fraction of the 3D Yee+CPML step on one GPU.

Workload (the 3D-at-scale configuration): a 1024 x 512 x 256 (x, y, z) grid at
25 nm (134,217,728 cells); background eps 2.085 with a seeded inner
512 x 256 x 64 slab of 8x8x8-cell blocks, eps iid U[2.085, 12.25]; 16-cell CPML
on all faces; a soft Ez plane source (Gaussian pulse at 1.55 um) at x = 40;
two 21-wavelength (1.45-1.65 um) DFT flux planes at x = 64 and x = 960 with
FP64 accumulators.  A "step" is one full leapfrog step (sources, H, E, DFT)
over the whole grid.  Inputs are resident in HBM before timing starts.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--dtype f32|bf16|f16]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


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
                        bz.GaussianPulse(f0, fw))
    lam = np.linspace(1.45e-6, 1.65e-6, 21)
    freqs = bz.LIGHT_SPEED / lam
    rngs = {"z": (L, nz - L), "y": (L, ny - L)}
    mons = [bz.FluxMonitor("x", 64, freqs, ranges=rngs), bz.FluxMonitor("x", nx - 64, freqs, ranges=rngs)]
    field_dtype = {"f32": torch.float32, "bf16": torch.bfloat16, "f16": torch.float16}[dtype_name]
    flags = _abi.DEFAULT_FLAGS | (_abi.FLAG_FUSED_STEP if fused else 0)
    sim = bz.Simulation(eps, dx, dt=dt, boundaries=[bz.PML(cells=L)], sources=[src],
                        monitors=mons if monitors else [], steps=3000, field_dtype=field_dtype,
                        flags=flags)
    return sim, eps


def algorithmic_bytes(prog, state):
    """Compulsory HBM bytes of one full step: every field, coefficient array and
    CPML memory element read once and written once (coefficients only read)."""
    fb = state.ex.element_size()
    fields = sum(t.numel() for t in state.fields())
    coefs = sum(v.numel() for k, v in prog.coeffs.items() if v.dim() > 0) * 4
    psi = sum(t.numel() * t.element_size() for t in state.psi_h + state.psi_e)
    return 2 * fields * fb + coefs + 2 * psi


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
            "sample": f"full 1024x512x256 grid with 16-cell CPML, {steps} FP32 steps "
                      f"({tt:.1f} s) of oracle/port.c (H+E+CPML; no source/DFT)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--dtype", default="f32", choices=("f32", "bf16", "f16"))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-monitors", action="store_true")
    ap.add_argument("--per-phase", action="store_true", help="use the per-phase kernels")
    ap.add_argument("--chunk", type=int, default=256)
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
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the z-slab sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    from beamz_amd import runtime

    print(f"[bench] building workload dtype={args.dtype}", file=sys.stderr, flush=True)
    sim, eps = build_workload(args.dtype, monitors=not args.no_monitors, fused=not args.per_phase)
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
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):
        start.record(stream)
        run(args.steps)
        stop.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / args.steps
    gcups = prog.cells / (ms * 1e-3) / 1e9
    print(f"[bench] {ms:.3f} ms/step, {gcups:.2f} GCUPS", file=sys.stderr, flush=True)

    # dominant kernel: the full-step update launched alone (no sources/DFT), timed
    # with events on its own stream; algorithmic bytes per launch = one full step
    nb = algorithmic_bytes(prog, state)
    k_ms = runtime.time_step_kernel(state, prog, launches=20, stream=stream)
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
            "workload": "3D Yee+CPML 1024x512x256 (x,y,z) at 25 nm, 16-cell CPML, seeded "
                        "random-eps 512x256x64 slab, Ez plane source, 2 x 21-wavelength FP64 DFT flux planes",
            "cells": prog.cells,
            "field_storage": args.dtype,
            "schedule": "per-phase" if args.per_phase else "fused two-bank",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": peak,
            "unit": "GB/s",
            "frac": round(achieved / peak, 4),
            "traffic": None,
            "kernel": runtime.step_kernel_name(prog),
            "kernel_ms": round(k_ms, 4),
            "algorithmic_bytes_per_launch": int(nb),
        },
    }
    if not args.no_cpu:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(eps, prog)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
