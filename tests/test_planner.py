"""Host lowering of ``beamz_amd`` agrees with the oracle's restatement:
coefficients, CPML terms and boundary code, source groups and monitor plans.
CPU only (tensors on the CPU device)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from beamz_amd import PEC, PML, FluxMonitor, SoftSource, GaussianPulse, compile_program
from beamz_amd import lattice as L
from oracle import fdtd64 as o
from tests.cases import eps_raster


@pytest.mark.parametrize("grid", [(5, 6, 7), (9, 11)])
def test_coefficients_and_cpml_terms_match_oracle(grid):
    eps = eps_raster(grid, 3)
    dx = 30e-9
    dt = 0.9 * dx / (o.LIGHT_SPEED * np.sqrt(len(grid)))
    prog = compile_program(eps, dx, dt, boundaries=[PML(cells=2, kappa_max=3.0)], device="cpu")
    spec = o.CpmlSpec(cells=2, kappa_max=3.0, faces=frozenset(
        ("front", "back", "bottom", "top", "left", "right") if len(grid) == 3 else
        ("bottom", "top", "left", "right")))
    ref = o.build_problem(eps, dx, dt, cpml=spec)
    for comp, arr in ref.e_source.items():
        got = prog.coeffs[f"e_source_{comp[1].lower()}"].numpy()
        np.testing.assert_allclose(np.broadcast_to(got, arr.shape), arr, rtol=1e-6)
    assert prog.metallic == ref.metallic == frozenset()
    for mine, theirs in zip(prog.cpml.h_terms + prog.cpml.e_terms, ref.hterms + ref.eterms):
        assert (mine.component, mine.axis, mine.sign, mine.low, mine.high) == \
               (theirs.target, theirs.axis, theirs.sign, theirs.low, theirs.high)
        np.testing.assert_allclose(mine.a, theirs.a, rtol=1e-13, atol=1e-300)
        np.testing.assert_allclose(mine.b, theirs.b, rtol=1e-13)
        np.testing.assert_allclose(mine.inv_kappa, theirs.ik, rtol=1e-13)


def test_boundary_code_uniform_and_mixed():
    eps = np.ones((12, 12, 12))
    dx = 20e-9
    dt = 2e-17
    uni = compile_program(eps, dx, dt, boundaries=[PML(cells=3)], device="cpu")
    assert uni.boundary_code == (3 << 8)
    mixed = compile_program(eps, dx, dt, boundaries=[PML(edges=("left", "right"), cells=3),
                                                    PEC(edges=("top", "bottom", "front", "back"))],
                            device="cpu")
    assert mixed.boundary_code & 0xFF == 1 | 2 | 4 | 8
    assert mixed.boundary_code >> 8 == 0
    pec = compile_program(eps, dx, dt, device="cpu")
    assert pec.boundary_code == 63 and pec.cpml is None


def test_plane_and_line_plans_match_oracle():
    grid = (6, 7, 9)
    shapes = L.component_shapes(grid)
    for c in L.COMPONENTS:
        a = L.plane_gather_plan(c, shapes[c], 0.1, "x", 4, {"z": (1, 5), "y": (0, 7)})
        b = o.plane_plan_3d(c, shapes[c], grid, 0.1, "x", 4, {"z": (1, 5), "y": (0, 7)})
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_allclose(a[1], b[1], rtol=1e-15)
    shapes2 = L.component_shapes((8, 10))
    for c in ("Ez", "Hx", "Hy"):
        a = L.line_gather_plan(c, shapes2[c], 0.1, "x", 6, (1, 7))
        b = o.line_plan_tm(c, shapes2[c], 0.1, "x", 6, (1, 7))
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_allclose(a[1], b[1])


def test_soft_source_lowering_matches_gaussian_source_formula():
    grid = (6, 8, 10)
    eps = eps_raster(grid, 5)
    dx, dt = 25e-9, 4e-17
    pulse = GaussianPulse(1.9e14, 2.5e13)
    src = SoftSource("Ez", (1, 2, 3), (3, 4, 1), pulse)
    prog = compile_program(eps, dx, dt, sources=[src], num_steps=50, device="cpu")
    coeff, wave, starts = prog.source_groups[2]   # (pre_e, Ez)
    eps_ez = o.e_site_permittivity_3d(eps, "Ez")[1:4, 2:6, 3:4]
    np.testing.assert_allclose(coeff.numpy()[0], -dt / (o.EPS_0 * eps_ez), rtol=1e-6)
    t = o.sample_times(0.0, dt, 50)
    np.testing.assert_allclose(wave.numpy()[0], o.gaussian_pulse(t, 1.9e14, 2.5e13), rtol=1e-5, atol=1e-7)
    assert starts.numpy().tolist() == [[1, 2, 3]]
    assert all(g is None for i, g in enumerate(prog.source_groups) if i != 2)
