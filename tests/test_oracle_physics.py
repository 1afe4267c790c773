"""Closed-form and physical pins of the FP64 oracle step: CPML algebra,
dispersion, reversibility, CPML reflection, DFT accumulation and 3D/2D
consistency."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fdtd64 as o


def test_cpml_coefficients_match_independent_derivation():
    # tests/kernels/test_maxwell_operators.py:116-131
    sigma = np.array([0.0, 0.2, 1.3, 2.1])
    kappa = np.array([1.0, 1.0, 2.0, 4.0])
    alpha = np.array([0.0, 0.03, 0.02, 0.0])
    dt = 4e-13
    a, b, ik = o.cpml_coefficients(sigma, kappa, alpha, dt)
    eb = np.exp(-(sigma / kappa + alpha) * dt / o.EPS_0)
    ea = np.nan_to_num(((eb - 1.0) * sigma) / np.maximum((sigma + kappa * alpha) * kappa, 1e-30))
    np.testing.assert_allclose(b, eb, rtol=1e-14)
    np.testing.assert_allclose(a, ea, rtol=1e-14)
    np.testing.assert_allclose(ik, 1.0 / kappa)


def test_packed_recurrence_updates_only_active_slabs():
    # tests/kernels/test_maxwell_operators.py:134-161
    sigma = np.array([0.8, 0.0, 0.0, 1.1])
    kappa = np.array([1.7, 1.0, 1.0, 2.2])
    alpha = np.array([0.04, 0.0, 0.0, 0.02])
    d = np.array([1.0, -2.0, 3.0, -4.0])
    dt = 2e-13
    term = o._make_term("Ez", "Hx", 0, -1.0, d.shape, (sigma, kappa, alpha), dt)
    assert (term.low, term.high) == (1, 1)
    corrected, psi = o.correct_term(d, np.zeros(2), term)
    a, _, ik = o.cpml_coefficients(sigma, kappa, alpha, dt)
    epsi = a * d
    np.testing.assert_allclose(corrected, -(d * ik + epsi))
    np.testing.assert_allclose(psi, epsi[[0, -1]])


def test_cpml_profile_has_identity_interface_and_full_strength_wall():
    # _boundary_compile.py:498-614: E samples at i, H at i+1/2; interface sample inactive
    s, k, a = o.cpml_profile_1d(21, "E", 20, 5, True, True, 7.0, 5.0, 3.0, 3)
    assert s[0] == pytest.approx(7.0) and k[0] == pytest.approx(5.0) and a[0] == 0.0
    assert s[5] == 0.0 and k[5] == 1.0 and a[5] == 0.0
    assert s[1] == pytest.approx(7.0 * (4 / 5) ** 3)
    assert a[1] == pytest.approx(3.0 * (1 - 4 / 5))
    assert s[20] == pytest.approx(7.0) and s[15] == 0.0
    sh, kh, ah = o.cpml_profile_1d(20, "H", 20, 5, True, True, 7.0, 5.0, 3.0, 3)
    assert sh[0] == pytest.approx(7.0 * (4.5 / 5) ** 3)
    assert sh[4] == pytest.approx(7.0 * (0.5 / 5) ** 3) and sh[5] == 0.0
    assert sh[19] == pytest.approx(sh[0])
    ca, cb, cik = o.cpml_coefficients(s, k, a, 1e-17)
    assert o.active_slab_counts(ca, cik) == (5, 5)


def test_sigma_max_formula():
    # _boundary_compile.py:244-253 for the 2D config: R=1e-6, m=3, 10 cells of 20 nm
    smax = o.cpml_sigma_max(3, 1e-6, 10 * 20e-9)
    assert smax == pytest.approx(-4 * np.log(1e-6) / (2 * o.ETA_0 * 200e-9))
    assert smax == pytest.approx(3.667e5, rel=1e-3)


def test_gaussian_pulse_closed_form():
    # sources/time.py:459-486
    fw, f0 = 2.0e13, 1.9e14
    t = np.array([0.0, 4.0 / fw, 5.0 / fw])
    sig = o.gaussian_pulse(t, f0, fw, remove_dc=False)
    tau = 1.0 / (2 * np.pi * fw)
    assert sig[1] == pytest.approx(np.cos(2 * np.pi * f0 * 4.0 / fw))
    assert sig[2] == pytest.approx(np.exp(-(1.0 / fw) ** 2 / (2 * tau ** 2)) * np.cos(2 * np.pi * f0 * 5.0 / fw))
    full = o.gaussian_pulse(np.linspace(0, 8 / fw, 400), f0, fw)
    assert abs(full.mean()) < 1e-12


def _plane_wave_2d(ppw, steps):
    lam = 1e-6
    dx = lam / ppw
    dt = 0.95 * dx / (o.LIGHT_SPEED * np.sqrt(2.0))
    ny, nx = 4 * ppw, 8 * ppw
    prob = o.build_problem(np.ones((ny, nx)), dx, dt)
    st = o.init_state(prob)
    k = 2 * np.pi / lam
    S = o.LIGHT_SPEED * dt / dx
    w = (2.0 / dt) * np.arcsin(S * np.sin(0.5 * k * dx))
    ex = np.arange(st["ez"].shape[1]) * dx
    hx_ = (np.arange(st["hy"].shape[1]) + 0.5) * dx
    st["ez"][...] = np.cos(k * ex)[None, :]
    st["hy"][...] = (-np.cos(k * hx_ + 0.5 * w * dt) / o.ETA_0)[None, :]
    st["ez"][prob.masks["Ez"]] = 0.0
    ez0 = st["ez"].copy()
    o.run(prob, st, steps)
    return prob, st, ez0, ex, hx_, k, w, dt


@pytest.mark.parametrize("ppw", [12, 20])
def test_plane_wave_follows_discrete_dispersion(ppw):
    # tests/validation/analytical/test_plane_wave_dispersion.py:31-150
    steps = 8
    prob, st, ez0, ex, hx_, k, w, dt = _plane_wave_2d(ppw, steps)
    row = st["ez"].shape[0] // 2
    sl = slice(2 * ppw, 6 * ppw)

    def amp(v, x):
        return 2.0 * np.mean(v[sl] * np.exp(-1j * k * x[sl]))

    e0, e1 = amp(ez0[row], ex), amp(st["ez"][row], ex)
    h1 = amp(st["hy"][row], hx_)
    measured_w = -np.angle(e1 / e0) / (steps * dt)
    assert measured_w == pytest.approx(w, rel=1e-6)
    assert abs(e1) / abs(e0) == pytest.approx(1.0, rel=1e-6)
    assert abs(e1 / h1) / o.ETA_0 == pytest.approx(1.0, rel=1e-6)


def test_lossless_step_is_time_reversible():
    # tests/validation/invariants/test_time_reversibility.py:52-84 (no PEC faces)
    ny, nx = 5, 7
    dx = 1e-6
    dt = 0.05 * dx / o.LIGHT_SPEED
    rng = np.random.default_rng(1307)
    fwd = o.build_problem(np.ones((ny, nx)), dx, dt, metallic=frozenset())
    bwd = o.build_problem(np.ones((ny, nx)), dx, -dt, metallic=frozenset())
    st = o.init_state(fwd)
    for c in ("ez", "hx", "hy"):
        st[c][...] = rng.normal(size=st[c].shape)
    init = {c: st[c].copy() for c in ("ez", "hx", "hy")}
    o.h_phase(fwd, st)
    o.e_phase(fwd, st)
    o.e_phase(bwd, st)
    o.h_phase(bwd, st)
    num = np.sqrt(sum(np.linalg.norm(st[c] - init[c]) ** 2 for c in init))
    den = np.sqrt(sum(np.linalg.norm(init[c]) ** 2 for c in init))
    assert num / den < 1e-13


def _packet_reflection_db(n, ppw):
    # tests/validation/analytical/test_cpml_packet_reflection.py:22-98
    lam0 = 1e-6
    lam = lam0 / n
    dx = lam / ppw
    dt = 0.95 * dx / (o.LIGHT_SPEED * np.sqrt(2.0))
    steps = int(np.ceil(12.0 * lam0 / o.LIGHT_SPEED / dt))
    eta = o.ETA_0 / n
    c = o.LIGHT_SPEED / n
    nx, ny = int(round(12.0 * lam / dx)), int(round(26.0 * lam / dx))
    cells = int(round(1.5 * lam / dx))
    spec = o.CpmlSpec(cells=cells, faces=frozenset({"left", "right"}))
    prob = o.build_problem(np.full((ny, nx), n * n), dx, dt, cpml=spec,
                           metallic=frozenset({"top", "bottom"}))
    st = o.init_state(prob)
    xe = np.arange(st["ez"].shape[1]) * dx
    xh = (np.arange(st["hy"].shape[1]) + 0.5) * dx
    x0, sg, k = 3.0 * lam, 0.8 * lam, 2 * np.pi / lam

    def packet(x):
        return np.exp(-0.5 * ((x - x0) / sg) ** 2) * np.cos(k * (x - x0))

    st["ez"][...] = packet(xe)[None, :]
    st["hy"][...] = (-packet(xh + 0.5 * c * dt) / eta)[None, :]
    e0 = 0.5 * (st["ez"][:, 1:] + st["ez"][:, :-1])
    incident = 0.5 * (e0 - eta * st["hy"])
    o.run(prob, st, steps)
    e1 = 0.5 * (st["ez"][:, 1:] + st["ez"][:, :-1])
    reflected = 0.5 * (e1 + eta * st["hy"])
    x = (np.arange(reflected.shape[1]) + 0.5) * dx
    y = np.arange(reflected.shape[0]) * dx
    cy = (y > 11.0 * lam) & (y < 15.0 * lam)
    ix = (x > 1.7 * lam) & (x < 10.3 * lam)
    return 10 * np.log10(np.sum(reflected[np.ix_(cy, ix)] ** 2) / np.sum(incident[np.ix_(cy, ix)] ** 2))


@pytest.mark.parametrize(("n", "ppw"), [(1.0, 10), (1.5, 15)])
def test_cpml_packet_reflection_below_minus_40_db(n, ppw):
    # tests/validation/analytical/test_cpml_packet_reflection.py:101-133
    assert _packet_reflection_db(n, ppw) < -40.0


def test_dft_of_constant_field_is_geometric_sum():
    """observe.py:529-585: t_phys = t0 + (step+1) dt, Re += cos(2 pi f t) v."""
    grid = (3, 4)
    dx, dt, t0 = 1e-8, 1e-17, 3e-17
    prob0 = o.build_problem(np.ones(grid), dx, dt, metallic=frozenset())
    plan = o.line_plan_tm("Ez", prob0.shapes["Ez"], dx, "x", 2, (0, 3))
    freqs = np.array([1e14, 2.5e14])
    mon = o.Monitor(freqs, {"Ez": plan}, "x", 1.0, dx)
    prob = o.build_problem(np.ones(grid), dx, dt, metallic=frozenset(), monitors=[mon], t0=t0)
    st = o.init_state(prob)
    n = 37
    for _ in range(n):
        st["ez"][...] = 1.5          # held constant; accumulate only
        o.accumulate_monitors(prob, st)
        st["step"] += 1
    t = t0 + (np.arange(n) + 1) * dt
    for fi, f in enumerate(freqs):
        np.testing.assert_allclose(st["dft_re"][0][2, fi], 1.5 * np.cos(2 * np.pi * f * t).sum(), rtol=1e-12)
        np.testing.assert_allclose(st["dft_im"][0][2, fi], 1.5 * np.sin(2 * np.pi * f * t).sum(), rtol=1e-12)
    assert np.all(st["dft_w"][0] == n)


def test_line_plan_interpolates_linear_fields_exactly():
    shape = (6, 9)
    dx = 0.5
    for comp in ("Ez", "Hx", "Hy"):
        oy, ox = o.OFFSETS_TM[comp]
        y = (np.arange(shape[0]) + oy) * dx
        x = (np.arange(shape[1]) + ox) * dx
        f = 2.0 * y[:, None] - 3.0 * x[None, :] + 1.0
        idx, w = o.line_plan_tm(comp, shape, dx, "x", 4, (1, 5))
        v = np.sum(f.reshape(-1)[idx] * w, axis=1)
        ys = (np.arange(1, 5) + 0.5) * dx
        np.testing.assert_allclose(v, 2.0 * ys - 3.0 * 4 * dx + 1.0, atol=1e-12)


def test_plane_plan_interpolates_linear_fields_exactly():
    grid = (4, 5, 6)
    shapes = o.shapes_3d(grid)
    dx = 0.25
    for comp in o.COMPONENTS:
        s = shapes[comp]
        oz, oy, ox = o.OFFSETS_3D[comp]
        z = (np.arange(s[0]) + oz) * dx
        y = (np.arange(s[1]) + oy) * dx
        x = (np.arange(s[2]) + ox) * dx
        f = 1.0 + 2 * z[:, None, None] - y[None, :, None] + 0.5 * x[None, None, :]
        idx, w = o.plane_plan_3d(comp, s, grid, dx, "x", 3, {"z": (1, 3), "y": (0, 5)})
        v = np.sum(f.reshape(-1)[idx] * w, axis=1)
        zz, yy = np.meshgrid((np.arange(1, 3) + 0.5) * dx, (np.arange(0, 5) + 0.5) * dx, indexing="ij")
        np.testing.assert_allclose(v, 1.0 + 2 * zz.ravel() - yy.ravel() + 0.5 * 3 * dx, atol=1e-12)


def _embed_tm_in_3d(perm):
    """Run a 2D TMxy problem and the same problem embedded in 3D, uniform along
    the third axis with PEC caps, and compare.  ``perm`` selects which 3D axis
    plays 2D-y/2D-x/the uniform axis, exercising every 3D term and sign."""
    rng = np.random.default_rng(7)
    ny, nx, nu = 9, 11, 3
    dx = 40e-9
    dt = 0.5 * dx / o.LIGHT_SPEED
    # 2D sites own one cell while 3D E sites average four, so keep the
    # permittivity uniform here; site sampling is pinned separately.
    eps2 = np.full((ny, nx), 2.25)
    spec2 = o.CpmlSpec(cells=3, faces=frozenset({"left", "right", "bottom", "top"}),
                       kappa_max=3.0, alpha_norm=0.05)
    p2 = o.build_problem(eps2, dx, dt, cpml=spec2)
    s2 = o.init_state(p2)
    for c in ("ez", "hx", "hy"):
        s2[c][...] = rng.normal(size=s2[c].shape)
    # 3D: axes named by role: (u: uniform, y2: 2D-y, x2: 2D-x)
    ax_u, ax_y, ax_x = perm
    grid3 = [0, 0, 0]
    grid3[ax_u], grid3[ax_y], grid3[ax_x] = nu, ny, nx
    eps3 = np.moveaxis(np.broadcast_to(eps2, (nu, ny, nx)), (0, 1, 2), (ax_u, ax_y, ax_x)).copy()
    names = o.AXIS_EDGES_3D
    faces = frozenset(names[ax_y] + names[ax_x])
    spec3 = o.CpmlSpec(cells=3, faces=faces, kappa_max=3.0, alpha_norm=0.05)
    p3 = o.build_problem(eps3, dx, dt, cpml=spec3, metallic=frozenset(names[ax_u]))
    s3 = o.init_state(p3)
    axname = "zyx"
    comp = {"Ez": "E" + axname[ax_u], "Hx": "H" + axname[ax_x], "Hy": "H" + axname[ax_y]}
    # handedness: (x2, y2, u) must be a right-handed permutation of (x, y, z)
    order = [axname[ax_x], axname[ax_y], axname[ax_u]]
    right = "".join(order) in ("xyz", "yzx", "zxy")
    sgnH = 1.0 if right else -1.0
    for c2, c3 in comp.items():
        src = s2[c2.lower()]
        tgt = s3[c3.lower()]
        # broadcast along the uniform axis, mapping 2D (y, x) onto (ax_y, ax_x)
        full = np.moveaxis(np.broadcast_to(src, (tgt.shape[ax_u],) + src.shape),
                           (0, 1, 2), (ax_u, ax_y, ax_x))
        tgt[...] = full * (sgnH if c2.startswith("H") else 1.0)
    o.run(p2, s2, 12)
    o.run(p3, s3, 12)
    for c2, c3 in comp.items():
        got = s3[c3.lower()]
        got = np.moveaxis(got, (ax_u, ax_y, ax_x), (0, 1, 2))
        want = s2[c2.lower()] * (sgnH if c2.startswith("H") else 1.0)
        for u in range(got.shape[0]):
            np.testing.assert_allclose(got[u], want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    for c in ("ex", "ey", "ez", "hx", "hy", "hz"):
        if c.upper()[0] + c[1] not in comp.values():
            assert np.abs(s3[c]).max() == 0.0


@pytest.mark.parametrize("perm", [(0, 1, 2), (2, 0, 1), (1, 2, 0), (0, 2, 1), (1, 0, 2), (2, 1, 0)])
def test_3d_step_reduces_to_2d_tm_under_every_axis_role(perm):
    _embed_tm_in_3d(perm)
