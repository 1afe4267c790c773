"""Pins of the oracle's lattice, curl and boundary rules against closed forms
and the reference's own kernel tests (no reference code runs here)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fdtd64 as o


def test_complete_yee_shapes_match_reference_table():
    # tests/kernels/test_yee_boundary_updates.py:25-38
    expected = {
        "Ex": (5, 6, 6), "Ey": (5, 5, 7), "Ez": (4, 6, 7),
        "Hx": (4, 5, 7), "Hy": (4, 6, 6), "Hz": (5, 5, 6),
    }
    assert o.shapes_3d((4, 5, 6)) == expected
    assert o.shapes_tm((5, 7)) == {"Ez": (6, 8), "Hx": (5, 8), "Hy": (6, 7)}


def _prob(grid, **kw):
    return o.build_problem(np.ones(grid), 1.0, 0.1, **kw)


def test_constant_field_has_zero_curl():
    # tests/kernels/test_maxwell_operators.py:35-48
    prob = _prob((4, 5, 6))
    st = o.init_state(prob)
    for c, v in zip(("ex", "ey", "ez"), (2.0, -3.0, 4.0)):
        st[c][...] = v
    for curl in o.curl_e_to_h(prob, st).values():
        assert np.all(curl == 0.0)


def test_discrete_divergence_of_curl_is_exactly_zero_on_integers():
    # tests/kernels/test_maxwell_operators.py:51-63
    prob = _prob((6, 7, 8))
    rng = np.random.default_rng(20260724)
    st = o.init_state(prob)
    for c in ("ex", "ey", "ez"):
        st[c][...] = rng.integers(-8, 9, size=st[c].shape)
    cur = o.curl_e_to_h(prob, st)
    div = (np.diff(cur["Hx"], axis=2) + np.diff(cur["Hy"], axis=1)
           + np.diff(cur["Hz"], axis=0))
    assert np.all(div == 0.0)


def test_linear_field_has_constant_analytic_curl():
    # tests/kernels/test_curl.py:155-200 (linear field -> constant curl)
    prob = _prob((5, 6, 7))
    st = o.init_state(prob)
    z, y, x = np.meshgrid(*(np.arange(n, dtype=float) for n in st["ez"].shape), indexing="ij")
    st["ez"][...] = 3.0 * y            # dEz/dy = 3 -> curl_x = 3
    z, y, x = np.meshgrid(*(np.arange(n, dtype=float) for n in st["ey"].shape), indexing="ij")
    st["ey"][...] = -2.0 * z + 5.0 * x  # dEy/dz = -2, dEy/dx = 5
    z, y, x = np.meshgrid(*(np.arange(n, dtype=float) for n in st["ex"].shape), indexing="ij")
    st["ex"][...] = 7.0 * z            # dEx/dz = 7
    cur = o.curl_e_to_h(prob, st)
    np.testing.assert_allclose(cur["Hx"], 3.0 - (-2.0))
    np.testing.assert_allclose(cur["Hy"], 7.0 - 0.0)
    np.testing.assert_allclose(cur["Hz"], 5.0 - 0.0)


def test_pec_curl_pair_is_skew_adjoint():
    # tests/validation/invariants/test_3d_curl_adjointness.py:16-72
    prob = _prob((6, 8, 18))
    rng = np.random.default_rng(0)
    st = o.init_state(prob)
    for c in o.COMPONENTS:
        st[c.lower()][...] = rng.standard_normal(st[c.lower()].shape)
        st[c.lower()][prob.masks[c]] = 0.0
    ch = o.curl_e_to_h(prob, st)
    ce = o.curl_h_to_e(prob, st)
    lhs = sum(np.vdot(st[c.lower()], ce[c]) for c in ("Ex", "Ey", "Ez"))
    rhs = sum(np.vdot(st[c.lower()], ch[c]) for c in ("Hx", "Hy", "Hz"))
    assert abs(lhs - rhs) / (abs(lhs) + abs(rhs)) < 1e-12


def test_skew_adjointness_fails_when_a_sign_is_flipped():
    """The adjointness check above must catch a flipped term sign."""
    prob = _prob((4, 5, 6))
    rng = np.random.default_rng(1)
    st = o.init_state(prob)
    for c in o.COMPONENTS:
        st[c.lower()][...] = rng.standard_normal(st[c.lower()].shape)
        st[c.lower()][prob.masks[c]] = 0.0
    ch = o.curl_e_to_h(prob, st)
    ce = o.curl_h_to_e(prob, st)
    ce["Ex"] = -ce["Ex"]
    lhs = sum(np.vdot(st[c.lower()], ce[c]) for c in ("Ex", "Ey", "Ez"))
    rhs = sum(np.vdot(st[c.lower()], ch[c]) for c in ("Hx", "Hy", "Hz"))
    assert abs(lhs - rhs) / (abs(lhs) + abs(rhs)) > 1e-3


def test_pec_masks_cover_both_sides_of_every_axis():
    # tests/kernels/test_yee_boundary_updates.py:55-95
    prob = _prob((4, 5, 6))
    m = prob.masks
    assert m["Ex"][0].all() and m["Ex"][-1].all() and m["Ex"][:, 0].all() and m["Ex"][:, -1].all()
    assert m["Ey"][:, :, 0].all() and m["Ey"][:, :, -1].all()
    assert m["Hz"][0].all() and m["Hz"][-1].all() and not m["Hz"][1:-1].any()
    assert not m["Hx"][:, :, 1:-1].any() and m["Hx"][:, :, 0].all()


def test_e_site_permittivity_is_centered_average():
    # lattice.py:732-786: Ex averages the four cells around its z/y edge
    cell = np.arange(2 * 3 * 4, dtype=float).reshape(2, 3, 4)
    ex = o.e_site_permittivity_3d(cell, "Ex")
    assert ex.shape == (3, 4, 4)
    k, j, i = 1, 1, 2
    expected = 0.25 * (cell[0, 0, i] + cell[0, 1, i] + cell[1, 0, i] + cell[1, 1, i])
    assert ex[k, j, i] == pytest.approx(expected)
    # walls clamp to the edge cell
    assert ex[0, 0, 0] == pytest.approx(cell[0, 0, 0])
    ez = o.e_site_permittivity_3d(cell, "Ez")
    assert ez[1, 2, 3] == pytest.approx(0.25 * (cell[1, 1, 2] + cell[1, 1, 3] + cell[1, 2, 2] + cell[1, 2, 3]))


def test_ghost_rule_gives_edge_value_on_metallic_and_zero_otherwise():
    # lattice.py:523-575 / A.3: metallic low face d=H[0]/dx, high d=-H[last]/dx, open d=0
    h = np.array([[1.0, 2.0, 4.0]])
    d_met = o.ghost_diff(h, 1, frozenset({"left", "right"}), o.AXIS_EDGES_2D, 0.5)
    np.testing.assert_allclose(d_met, [[2.0, 2.0, 4.0, -8.0]])
    d_open = o.ghost_diff(h, 1, frozenset(), o.AXIS_EDGES_2D, 0.5)
    np.testing.assert_allclose(d_open, [[0.0, 2.0, 4.0, 0.0]])
