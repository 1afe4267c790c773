"""Golden finite differences of the 2D inverse-design bend, from the FP64 oracle.

The workload is the 600 x 600 bend at 20 nm that ``bench.py --workload
adjoint2d`` runs (``bench.build_adjoint_problem``): a 100 x 100 design box of
seeded density, the input guide's mode launched by a mode source, 4000 steps,
and the objective J = sum_f |sum_p m_p dx DFT_f[Ez](p)|^2 on the output line.

This script runs that problem through the FP64 NumPy oracle on the CPU, with
the mode source lowered by the package's host planner into its two timed slabs
(the design box lies away from the source line, so the slabs do not depend on
rho), and writes

    adjoint_bend_600x600_fd.npz:  J0 (the objective at rho0), pixels (32, 2),
                                   h, J_plus / J_minus (32,), fd (32,)

where fd is the central difference (J(rho0 + h e_p) - J(rho0 - h e_p)) / 2h
with h = 1e-3 on the 32 pixels ``default_rng(17).integers(0, 100, (32, 2))``.
65 oracle runs of 4000 steps; run with ``--workers N`` processes (about a
minute and a quarter per run on one core).

    python tests/golden/make_adjoint_bend_fd.py --workers 6
"""

from __future__ import annotations

import argparse
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "adjoint_bend_600x600_fd.npz")
H = 1e-3
PIXEL_SEED = 17
NPIX = 32


def pixels():
    rng = np.random.default_rng(PIXEL_SEED)
    return rng.integers(0, 100, size=(NPIX, 2))


def oracle_setup():
    """The bend problem, its rho0 and everything the oracle needs that does
    not depend on rho: lowered source slabs, the monitor plan, the CPML."""
    import bench
    from beamz_amd.modes import lower_mode_source_2d
    from oracle import fdtd64 as o
    prob, rho0 = bench.build_adjoint_problem()
    eps0 = prob.permittivity(rho0)
    eps_ez = o.site_sample_tm(eps0, "Ez")
    (src,) = prob.sources
    customs = lower_mode_source_2d(src, eps_ez, prob.resolution, prob.dt, 0.0, prob.steps)
    # the HIP program stores coefficients and waveforms in float32
    srcs = [o.Source(s.component, s.timing, s.start,
                     np.asarray(s.coeff, np.float32).astype(np.float64),
                     np.asarray(s.waveform, np.float32).astype(np.float64)) for s in customs]
    mon = prob.monitor
    ny, nx = eps0.shape
    shapes = o.shapes_tm((ny, nx))
    plans = {c: None for c in o.COMPONENTS}
    (rng_x,) = mon.ranges.values()
    plans["Ez"] = o.line_plan_tm("Ez", shapes["Ez"], prob.resolution, mon.normal, mon.index,
                                 tuple(rng_x))
    omon = o.Monitor(np.asarray(mon.freqs, np.float64), plans, mon.normal, 1.0, prob.resolution)
    spec = o.CpmlSpec(cells=prob.pml_cells, kappa_max=2.0,
                      faces=frozenset(("bottom", "top", "left", "right")))
    return prob, rho0, srcs, omon, spec


def objective(rho):
    from oracle import fdtd64 as o
    prob, _, srcs, omon, spec = oracle_setup()
    eps = prob.permittivity(rho)
    problem = o.build_problem(eps, prob.resolution, prob.dt, cpml=spec, sources=srcs,
                              monitors=[omon])
    st = o.init_state(problem)
    o.run(problem, st, prob.steps)
    c = o.COMPONENTS.index("Ez")
    phasor = st["dft_re"][0][c] + 1j * st["dft_im"][0][c]
    O = (phasor * (prob.mode * prob.resolution)[None, :]).sum(axis=1)
    return float(np.sum(np.abs(O) ** 2))


def _job(arg):
    sign, a, b = arg
    _, rho0, *_ = oracle_setup()
    rho = rho0.copy()
    if sign != 0:
        rho[a, b] += sign * H
    t = time.time()
    J = objective(rho)
    print(f"  J(rho0 {'+-'[sign < 0] if sign else ' '} h e[{a},{b}]) = {J:.12e} "
          f"({time.time() - t:.0f} s)", flush=True)
    return J


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=6)
    args = ap.parse_args()
    pix = pixels()
    jobs = [(0, 0, 0)] + [(s, int(a), int(b)) for a, b in pix for s in (+1, -1)]
    with ProcessPoolExecutor(max_workers=args.workers) as pool:
        vals = list(pool.map(_job, jobs))
    J0 = vals[0]
    Jp = np.asarray(vals[1::2])
    Jm = np.asarray(vals[2::2])
    fd = (Jp - Jm) / (2 * H)
    np.savez(OUT, J0=J0, pixels=pix, h=H, J_plus=Jp, J_minus=Jm, fd=fd)
    print(f"wrote {OUT}: J0 {J0:.6e}, max |fd| {np.max(np.abs(fd)):.3e}")


if __name__ == "__main__":
    main()
