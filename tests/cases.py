"""Shared parity cases: one description lowered both by the oracle and by the
``beamz_amd`` planner, with seeded inputs."""

from __future__ import annotations

import numpy as np

from oracle import fdtd64 as o

C0 = o.LIGHT_SPEED


def eps_raster(grid, seed, lo=1.5, hi=12.25):
    rng = np.random.default_rng(seed)
    return lo + (hi - lo) * rng.random(grid)


def oracle_problem(eps, dx, dt, *, cpml_cells=None, faces=None, kappa_max=2.0,
                   alpha_norm=None, metallic=None, sources=(), monitors=(), t0=0.0):
    spec = None
    if cpml_cells:
        all_faces = ("front", "back", "bottom", "top", "left", "right") if eps.ndim == 3 else \
            ("bottom", "top", "left", "right")
        spec = o.CpmlSpec(cells=cpml_cells, kappa_max=kappa_max, alpha_norm=alpha_norm,
                          faces=frozenset(faces or all_faces))
    return o.build_problem(eps, dx, dt, cpml=spec, metallic=metallic, sources=sources,
                           monitors=monitors, t0=t0)


def amd_boundaries(eps, *, cpml_cells=None, faces=None, kappa_max=2.0, alpha_norm=None,
                   metallic=None):
    from beamz_amd import PEC, PML
    from beamz_amd.constants import EPS_0
    out = []
    if cpml_cells:
        kw = {}
        if alpha_norm is not None:
            kw["alpha_max"] = None  # resolved below
        out.append(PML(edges=faces or "all", cells=cpml_cells, kappa_max=kappa_max))
    if metallic:
        out.append(PEC(edges=tuple(metallic)))
    return out


def random_fields(shapes, seed, scale=1e-5, keys=None):
    rng = np.random.default_rng(seed)
    out = {}
    for c, s in shapes.items():
        if keys and c not in keys:
            continue
        out[c.lower()] = rng.normal(size=s) * scale * (1.0 if c.startswith("E") else 1.0 / o.ETA_0)
    return out
