"""3D mode sources: a full-vector finite-difference mode solver on the source
plane and the Huygens pair it drives.

The reference solves plane modes with its FDFD solver (``devices/modes/*``,
driven from ``sources/solve.py``) and launches them through
``sources/mode_launch.py:371-599``: an E-timed electric current on the source
plane and an H-timed magnetic current on the neighbouring plane
(``sources/compiler.py:592-624``).  This module restates that launch in 3D
with its own solver, written for the Yee lattice the kernels step.

Solver.  For propagation along axis a with transverse axes (b, c) in cyclic
order ((x; y, z), (y; z, x), (z; x, y)) and fields ~ exp(-i beta a), Maxwell's
equations with the longitudinal E and H eliminated give, for the transverse E,

    beta^2 E_b = k0^2 eps_b E_b - d_c C + d_b G
    beta^2 E_c = k0^2 eps_c E_c + d_b C + d_c G

with C = d_b E_c - d_c E_b (at the longitudinal-H sites, cell centres) and
G = eps_a^-1 (d_b(eps_b E_b) + d_c(eps_c E_c)) (at the longitudinal-E sites,
nodes).  On the plane E_b sits at (j + 1/2, k), E_c at (j, k + 1/2), as on the
3D Yee lattice, and every difference is the Yee one between those sites.  The
plane's window is bounded by PEC: tangential E and the longitudinal E vanish
on its edges.  The transverse H follows from Faraday's law:

    H_c = (beta E_b - d_b G / beta) / (omega mu0)     (at the E_b sites)
    H_b = -(beta E_c - d_c G / beta) / (omega mu0)    (at the E_c sites)

so the mode carries power 1/2 sum (E_b H_c - E_c H_b) dx^2 towards +a.
Pinned (``tests/test_modes3d.py``) by the discrete closed form of a PEC
rectangular guide for both polarizations, the analytic TE slab dispersion
relation, and the invariance of the spectrum under a 90-degree rotation of the
plane (which exchanges the two transverse parts of the operator); the launch
is checked end to end on the GPU for direction, purity and power.

Launch.  With n = +-a (the direction), the surface currents J = n x H and
M = -n x E of the mode travelling along n are injected as E increments
-J dt / (eps0 eps dx) on the source plane (timing "e") and H increments
-M dt / (mu0 dx) on the companion plane one cell behind it (timing "h"),
after a separable Tukey(0.3) taper and scaling to the source power
(``mode_launch.py:307-364``).  The mode fields are real, so the reference's
phase alignment and real projection are identities here; its 2D impedance
matching is not applied to vector modes.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .constants import EPS_0, LIGHT_SPEED, MU_0
from .modes import ModeSource, _window, snap_cell_center, snap_centered_extent
from .sources import CustomSource, sample_waveform

# (a, b, c) for each propagation axis, and the field components on b and c
_CYCLIC = {"x": ("x", "y", "z"), "y": ("y", "z", "x"), "z": ("z", "x", "y")}
_AXIS_POS = {"z": 0, "y": 1, "x": 2}   # storage order (z, y, x)


def _cyclic(axis):
    a, b, c = _CYCLIC[axis]
    return a, b, c


@dataclass(frozen=True)
class PlaneModes:
    """Modes of a plane window of Nb x Nc cells, by decreasing n_eff.

    ``eb[m]`` has shape (Nb, Nc + 1) (E_b at (j + 1/2, k)), ``ec[m]`` (Nb + 1,
    Nc) (E_c at (j, k + 1/2)); ``hc[m]`` and ``hb[m]`` sit with ``eb`` and
    ``ec``.  Each mode carries unit power towards +a.  The longitudinal
    fields lie in quadrature with the transverse ones: E_a = i ``ea[m]`` at
    the (Nb + 1, Nc + 1) nodes (E_a = -i G / beta) and H_a = i ``ha[m]`` at the
    (Nb, Nc) cell centres (H_a = i C / (omega mu0)).  ``beta`` holds the
    propagation constants."""

    neff: np.ndarray
    eb: np.ndarray
    ec: np.ndarray
    hb: np.ndarray
    hc: np.ndarray
    ea: np.ndarray = None
    ha: np.ndarray = None
    beta: np.ndarray = None


def _operators(Nb, Nc, h):
    """Sparse Yee differences on the plane's full site sets."""
    import scipy.sparse as sp

    def idx(shape):
        return np.arange(shape[0] * shape[1]).reshape(shape)

    def diff(src_shape, dst_shape, pairs):
        # pairs: list of (dst index array, src index array, sign)
        rows, cols, vals = [], [], []
        for d, s, sign in pairs:
            rows.append(d.ravel())
            cols.append(s.ravel())
            vals.append(np.full(d.size, sign / h))
        n_dst, n_src = dst_shape[0] * dst_shape[1], src_shape[0] * src_shape[1]
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                             shape=(n_dst, n_src))

    sEb, sEc, sN, sQ = (Nb, Nc + 1), (Nb + 1, Nc), (Nb + 1, Nc + 1), (Nb, Nc)
    iEb, iEc, iN, iQ = idx(sEb), idx(sEc), idx(sN), idx(sQ)
    # C = d_b Ec - d_c Eb  (to cell centres Q)
    dbEc = diff(sEc, sQ, [(iQ, iEc[1:, :], 1.0), (iQ, iEc[:-1, :], -1.0)])
    dcEb = diff(sEb, sQ, [(iQ, iEb[:, 1:], 1.0), (iQ, iEb[:, :-1], -1.0)])
    # D = d_b(eps_b Eb) + d_c(eps_c Ec) (to nodes N; sites outside the window are 0)
    dbEb = diff(sEb, sN, [(iN[:-1, :], iEb, 1.0), (iN[1:, :], iEb, -1.0)])
    dcEc = diff(sEc, sN, [(iN[:, :-1], iEc, 1.0), (iN[:, 1:], iEc, -1.0)])
    # d_b G, d_c G (nodes to Eb, Ec sites)
    dbN = diff(sN, sEb, [(iEb, iN[1:, :], 1.0), (iEb, iN[:-1, :], -1.0)])
    dcN = diff(sN, sEc, [(iEc, iN[:, 1:], 1.0), (iEc, iN[:, :-1], -1.0)])
    # d_c C to Eb sites, d_b C to Ec sites (centres outside the window are 0)
    dcQ = diff(sQ, sEb, [(iEb[:, :-1], iQ, 1.0), (iEb[:, 1:], iQ, -1.0)])
    dbQ = diff(sQ, sEc, [(iEc[:-1, :], iQ, 1.0), (iEc[1:, :], iQ, -1.0)])
    return dict(dbEc=dbEc, dcEb=dcEb, dbEb=dbEb, dcEc=dcEc, dbN=dbN, dcN=dcN, dcQ=dcQ, dbQ=dbQ)


def solve_plane_modes(eps_b, eps_c, eps_a, omega, h, count, target_neff):
    """Full-vector modes of a PEC-bounded plane window.

    ``eps_b`` (Nb, Nc + 1), ``eps_c`` (Nb + 1, Nc) and ``eps_a`` (Nb + 1,
    Nc + 1) are the permittivities at the E_b, E_c and longitudinal-E sites.
    Returns the ``count`` modes whose n_eff^2 lie nearest ``target_neff^2``
    (shift-invert), by decreasing n_eff, each scaled to unit power."""
    import scipy.sparse as sp
    from scipy.sparse.linalg import eigs

    eps_b = np.asarray(eps_b, dtype=np.float64)
    eps_c = np.asarray(eps_c, dtype=np.float64)
    eps_a = np.asarray(eps_a, dtype=np.float64)
    Nb, Nc = eps_b.shape[0], eps_c.shape[1]
    if eps_b.shape != (Nb, Nc + 1) or eps_c.shape != (Nb + 1, Nc) or eps_a.shape != (Nb + 1, Nc + 1):
        raise ValueError("plane permittivities do not match the window's Yee sites")
    if Nb < 2 or Nc < 2:
        raise ValueError("a mode plane needs at least 2 x 2 cells")
    k0 = float(omega) / LIGHT_SPEED
    op = _operators(Nb, Nc, float(h))
    # longitudinal E vanishes on the window's edge nodes
    node_in = np.zeros((Nb + 1, Nc + 1))
    node_in[1:-1, 1:-1] = 1.0
    G_of_D = sp.diags(node_in.ravel() / eps_a.ravel())
    Db = op["dbEb"] @ sp.diags(eps_b.ravel())
    Dc = op["dcEc"] @ sp.diags(eps_c.ravel())
    Gb, Gc = G_of_D @ Db, G_of_D @ Dc          # G from Eb, from Ec
    Cb, Cc = -op["dcEb"], op["dbEc"]           # C from Eb, from Ec
    A_bb = k0 ** 2 * sp.diags(eps_b.ravel()) - op["dcQ"] @ Cb + op["dbN"] @ Gb
    A_bc = -op["dcQ"] @ Cc + op["dbN"] @ Gc
    A_cb = op["dbQ"] @ Cb + op["dcN"] @ Gb
    A_cc = k0 ** 2 * sp.diags(eps_c.ravel()) + op["dbQ"] @ Cc + op["dcN"] @ Gc
    # unknowns: E_b off the c-edges, E_c off the b-edges
    mb = np.zeros((Nb, Nc + 1), dtype=bool)
    mb[:, 1:-1] = True
    mc = np.zeros((Nb + 1, Nc), dtype=bool)
    mc[1:-1, :] = True
    ub, uc = np.flatnonzero(mb.ravel()), np.flatnonzero(mc.ravel())
    A = sp.bmat([[A_bb[ub][:, ub], A_bc[ub][:, uc]], [A_cb[uc][:, ub], A_cc[uc][:, uc]]]).tocsc()
    n = A.shape[0]
    count = max(1, min(int(count), n - 2))
    sigma = (k0 * float(target_neff)) ** 2
    # a fixed start vector makes the solve (and every lowering) reproducible
    v0 = np.random.default_rng(0).standard_normal(n)
    w, v = eigs(A, k=count, sigma=sigma, which="LM", v0=v0)
    beta2 = np.real(w)
    order = np.argsort(beta2)[::-1]
    beta2, v = beta2[order], v[:, order]
    # eigenvectors carry an arbitrary complex phase: rotate the largest
    # component to the real axis before keeping the real part
    peak = v[np.argmax(np.abs(v), axis=0), np.arange(v.shape[1])]
    v = np.real(v * np.conj(peak / np.abs(peak))[None, :])
    keep = beta2 > 0.0
    beta2, v = beta2[keep], v[:, keep]
    neff = np.sqrt(beta2) / k0
    modes_eb, modes_ec, modes_hb, modes_hc, modes_ea, modes_ha = [], [], [], [], [], []
    for m in range(beta2.size):
        beta = np.sqrt(beta2[m])
        eb = np.zeros(Nb * (Nc + 1))
        ec = np.zeros((Nb + 1) * Nc)
        eb[ub] = v[:ub.size, m]
        ec[uc] = v[ub.size:, m]
        G = Gb @ eb + Gc @ ec
        C = Cb @ eb + Cc @ ec
        hc = (beta * eb - (op["dbN"] @ G) / beta) / (omega * MU_0)
        hb = -(beta * ec - (op["dcN"] @ G) / beta) / (omega * MU_0)
        ea = -G / beta                      # E_a = i ea
        ha = C / (omega * MU_0)             # H_a = i ha
        p = 0.5 * float(np.sum(eb * hc) - np.sum(ec * hb)) * h * h
        s = 1.0 / np.sqrt(abs(p)) if abs(p) > 1e-300 else 1.0
        if p < 0:
            s = -s
        modes_eb.append((eb * s).reshape(Nb, Nc + 1))
        modes_ec.append((ec * s).reshape(Nb + 1, Nc))
        modes_hb.append((hb * s).reshape(Nb + 1, Nc))
        modes_hc.append((hc * s).reshape(Nb, Nc + 1))
        modes_ea.append((ea * s).reshape(Nb + 1, Nc + 1))
        modes_ha.append((ha * s).reshape(Nb, Nc))
    return PlaneModes(neff, np.array(modes_eb), np.array(modes_ec), np.array(modes_hb),
                      np.array(modes_hc), np.array(modes_ea), np.array(modes_ha),
                      np.sqrt(beta2))


def _plane_index(axis, a_idx, b_sl, c_sl):
    """Storage-order (z, y, x) index tuple of a-index ``a_idx`` and b, c slices."""
    a, b, c = _cyclic(axis)
    out = [None, None, None]
    out[_AXIS_POS[a]] = a_idx
    out[_AXIS_POS[b]] = b_sl
    out[_AXIS_POS[c]] = c_sl
    return tuple(out)


def _to_bc(axis, arr):
    """View of a plane array (storage order with the a axis removed) in (b, c) order."""
    a, b, c = _cyclic(axis)
    rest = [ax for ax in ("z", "y", "x") if ax != a]
    return arr if rest == [b, c] else arr.T


@dataclass(frozen=True)
class ModePlan3d:
    neff: float
    window: tuple           # ((b0, b1), (c0, c1)) cells
    plane: int              # a-index of the E plane
    companion: int          # a-index of the H plane
    entries: tuple          # (component, timing, start (z, y, x), profile in storage order)


@dataclass(frozen=True)
class PlaneSetup:
    """The snapped plane of a 3D mode source: propagation axis, E plane and
    companion (H) plane a-indices, the window's cells and the permittivities
    at its Yee sites, in (b, c) order."""

    axis: str
    plane: int
    companion: int
    window: tuple           # ((b0, b1), (c0, c1)) cells
    eps_b: np.ndarray
    eps_c: np.ndarray
    eps_a: np.ndarray


def plane_setup(src: ModeSource, eps_site, grid, resolution) -> PlaneSetup:
    """Snap a 3D mode source onto the lattice (``_placement.py:305-330``): the
    E plane at the nearest cell centre along a, the companion one cell behind
    it for "+" and one ahead for "-", the window over the cells the source's
    transverse extent covers (at least 2)."""
    axis = src.axis3d
    a, b, c = _cyclic(axis)
    dx = float(resolution)
    counts = {"z": grid[0], "y": grid[1], "x": grid[2]}
    center = dict(zip(("x", "y", "z"), (float(v) for v in src.center)))
    size = dict(zip(("x", "y", "z"), (float(v) for v in src.size)))
    ca = snap_cell_center(center[a], dx, counts[a])
    if src.direction == "+":
        cmp_idx = max(0, ca - 1)
    else:
        cmp_idx = min(max(counts[a] - 2, 0), ca + 1)
    b0, b1 = snap_centered_extent(center[b], size[b], dx, counts[b], min_cells=2)
    c0, c1 = snap_centered_extent(center[c], size[c], dx, counts[c], min_cells=2)
    comp_b, comp_c, comp_a = "E" + b, "E" + c, "E" + a
    # permittivities at the window's sites on the E plane: E_b sites
    # (b0 + j + 1/2, c0 + k) k = 0..Nc, E_c sites (b0 + j, c0 + k + 1/2)
    # j = 0..Nb; the longitudinal E sites are half a cell along a, averaged
    eps_b = _to_bc(axis, np.asarray(eps_site[comp_b])[_plane_index(axis, ca, slice(b0, b1), slice(c0, c1 + 1))])
    eps_c = _to_bc(axis, np.asarray(eps_site[comp_c])[_plane_index(axis, ca, slice(b0, b1 + 1), slice(c0, c1))])
    ea = np.asarray(eps_site[comp_a])
    lo, hi = max(ca - 1, 0), min(ca, ea.shape[_AXIS_POS[a]] - 1)
    eps_a = 0.5 * (_to_bc(axis, ea[_plane_index(axis, lo, slice(b0, b1 + 1), slice(c0, c1 + 1))]) +
                   _to_bc(axis, ea[_plane_index(axis, hi, slice(b0, b1 + 1), slice(c0, c1 + 1))]))
    return PlaneSetup(axis, ca, cmp_idx, ((b0, b1), (c0, c1)), eps_b, eps_c, eps_a)


def solve_source_mode(src: ModeSource, setup: PlaneSetup, omega, resolution):
    """Solve the setup's window at angular frequency ``omega`` and return
    ``(modes, k)``: the plane modes and the index of the one ``src.mode_spec``
    selects (``mode_index`` by decreasing n_eff, near ``target_neff``, by
    default 0.98 of the largest index on the plane)."""
    spec = src.mode_spec
    if spec.polarization not in (None,):
        raise ValueError("3D mode sources select modes by index; polarization filters are not built")
    target = spec.target_neff if spec.target_neff is not None else 0.98 * np.sqrt(
        max(float(setup.eps_b.max()), float(setup.eps_c.max())))
    m = max(int(spec.num_modes), int(spec.mode_index) + 1, 3)
    modes = solve_plane_modes(setup.eps_b, setup.eps_c, setup.eps_a, omega, float(resolution),
                              2 * m + 5, target)
    if modes.neff.size == 0:
        raise ValueError("the plane mode solver found no guided mode")
    return modes, min(int(spec.mode_index), modes.neff.size - 1)


def plan_mode_source_3d(src: ModeSource, eps_site, grid, resolution):
    """Plan the Huygens pair of a 3D mode source.  ``eps_site`` maps Ex, Ey,
    Ez to their Yee-site permittivities; ``grid`` is (nz, ny, nx)."""
    setup = plane_setup(src, eps_site, grid, resolution)
    axis = setup.axis
    a, b, c = _cyclic(axis)
    dx = float(resolution)
    ca, cmp_idx = setup.plane, setup.companion
    (b0, b1), (c0, c1) = setup.window
    Nb, Nc = b1 - b0, c1 - c0
    omega = 2.0 * np.pi * src.frequency
    modes, k = solve_source_mode(src, setup, omega, dx)
    eb, ec, hb, hc = modes.eb[k], modes.ec[k], modes.hb[k], modes.hc[k]
    comp_b, comp_c = "E" + b, "E" + c
    d = 1.0 if src.direction == "+" else -1.0
    # the mode travelling along d a: E unchanged, H times d
    hb, hc = d * hb, d * hc
    # surface currents J = n x H, M = -n x E with n = d a (cyclic components):
    # J_b = -d H_c, J_c = d H_b, M_b = d E_c, M_c = -d E_b; increments -J, -M
    e_inc_b, e_inc_c = d * hc, -d * hb
    h_inc_b, h_inc_c = -d * ec, d * eb
    # taper over the window (separable Tukey), then the source power
    wb = _window(Nb)
    wc = _window(Nc)
    wb_n = np.interp(np.arange(Nb + 1) - 0.5, np.arange(Nb), wb)   # at the E_c nodes in b
    wc_n = np.interp(np.arange(Nc + 1) - 0.5, np.arange(Nc), wc)   # at the E_b nodes in c
    t_b = wb[:, None] * wc_n[None, :]     # E_b sites (Nb, Nc + 1)
    t_c = wb_n[:, None] * wc[None, :]     # E_c sites (Nb + 1, Nc)
    e_inc_b, h_inc_c = e_inc_b * t_b, h_inc_c * t_b
    e_inc_c, h_inc_b = e_inc_c * t_c, h_inc_b * t_c
    # unit power of the tapered pair (E from -M, H from -J), then sqrt(power)
    p = 0.5 * abs(float(np.sum(h_inc_c * e_inc_b) - np.sum(h_inc_b * e_inc_c))) * dx * dx
    s = np.sqrt(float(src.power) / p) if p > 1e-300 else 0.0
    e_inc_b, e_inc_c, h_inc_b, h_inc_c = e_inc_b * s, e_inc_c * s, h_inc_b * s, h_inc_c * s

    def entry(comp, timing, a_idx, prof_bc, b_from, c_from):
        # prof_bc is (nb, nc) in (b, c) order starting at window offsets (b_from, c_from)
        nb_, nc_ = prof_bc.shape
        start = [0, 0, 0]
        start[_AXIS_POS[a]] = a_idx
        start[_AXIS_POS[b]] = b0 + b_from
        start[_AXIS_POS[c]] = c0 + c_from
        shape = [0, 0, 0]
        shape[_AXIS_POS[a]] = 1
        shape[_AXIS_POS[b]] = nb_
        shape[_AXIS_POS[c]] = nc_
        # storage order of the profile
        prof = prof_bc if _AXIS_POS[b] < _AXIS_POS[c] else prof_bc.T
        return (comp, timing, tuple(start), prof.reshape(shape))

    # E_b unknowns k = 1..Nc-1, E_c unknowns j = 1..Nb-1 (edges carry zero)
    entries = (
        entry(comp_b, "e", ca, e_inc_b[:, 1:-1], 0, 1),
        entry(comp_c, "e", ca, e_inc_c[1:-1, :], 1, 0),
        entry("H" + b, "h", cmp_idx, h_inc_b[1:-1, :], 1, 0),
        entry("H" + c, "h", cmp_idx, h_inc_c[:, 1:-1], 0, 1),
    )
    return ModePlan3d(neff=float(modes.neff[k]), window=((b0, b1), (c0, c1)), plane=ca,
                      companion=cmp_idx, entries=entries)


def lower_mode_source_3d(src: ModeSource, eps_site, grid, resolution, dt, t0, total_steps):
    """The four timed slab sources of a 3D mode source: coefficient
    ``profile dt / (eps0 eps_site dx)`` on the E plane and ``profile dt /
    (mu0 dx)`` on the H plane, waveform at ``t0 + (i + 1/2) dt``
    (``sources/compiler.py:592-624``, ``463-490``)."""
    plan = plan_mode_source_3d(src, eps_site, grid, resolution)
    wave = sample_waveform(src.source_time, t0, dt, total_steps, half_step=True)
    out = []
    for comp, timing, start, prof in plan.entries:
        if comp.startswith("H"):
            denom = MU_0 * resolution
        else:
            region = tuple(slice(s, s + n) for s, n in zip(start, prof.shape))
            denom = EPS_0 * np.asarray(eps_site[comp])[region] * resolution
        out.append(CustomSource(comp, start, prof * dt / denom, wave, timing))
    return out
