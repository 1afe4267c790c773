"""The C/OpenMP FP32 CPU baseline (oracle/port.c) agrees with the FP64 oracle."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fdtd64 as o
from oracle import port
from tests.cases import eps_raster


@pytest.mark.parametrize("cpml", [True, False])
def test_port_step_matches_oracle(cpml):
    port.build()
    grid = (8, 9, 12)
    eps = eps_raster(grid, 2)
    dx = 40e-9
    dt = 0.9 * dx / (o.LIGHT_SPEED * np.sqrt(3))
    spec = o.CpmlSpec(cells=2, kappa_max=3.0) if cpml else None
    prob = o.build_problem(eps, dx, dt, cpml=spec)
    terms = []
    if cpml:
        for t in prob.hterms + prob.eterms:
            shape = list(prob.shapes[t.target])
            shape[t.axis] = t.low + t.high
            assert t.low == t.high == 2
            terms.append((t.a, t.b, t.ik, tuple(shape)))
    metal = sum({"front": 1, "back": 2, "bottom": 4, "top": 8, "left": 16, "right": 32}[e]
                for e in prob.metallic)
    pg = port.PortGrid(grid, 1.0 / dx, dt / o.MU_0, [prob.e_source[c] for c in ("Ex", "Ey", "Ez")],
                       L=2 if cpml else 0, terms=terms, metallic=metal, seed=4)
    st = o.init_state(prob)
    for k, v in pg.f.items():
        st[k][...] = v.astype(np.float64)
    pg.step(6)
    o.run(prob, st, 6)
    for k, v in pg.f.items():
        err = np.linalg.norm(v - st[k]) / np.linalg.norm(st[k])
        assert err < 1e-5, (k, err)
    for i, (g, w) in enumerate(zip(pg.psi, st["psi_h"] + st["psi_e"])):
        assert np.linalg.norm(g - w) <= 1e-4 * np.linalg.norm(w) + 1e-30, i
