"""The physically placed Gaussian source (host lowering).

Pinned by properties the reference's grid-path lowering fixes
(``sources/compiler.py:418-460``, ``649-700``; ``time.py:17-32``):
- the profile is a product of 1-D Gaussians (rank one per axis pair), equal
  to 1 on a node at the centre and symmetric about it;
- each axis keeps exactly the Ez coordinates within 4 widths of the centre
  (nodes in x and y, cell centres in z), or the nearest one when none is;
- the coefficient divides by eps0 times the Ez-site permittivity;
- an array signal is read by linear interpolation at the half-step times,
  its last sample at the last index and 0 beyond.
"""

from __future__ import annotations

import numpy as np
import pytest

from beamz_amd import GaussianPulse, GaussianSource, compile_program
from beamz_amd.constants import EPS_0
from beamz_amd.sources import gaussian_profile, interpolate_signal, lower_gaussian_source
from oracle import fdtd64 as o

DX = 20e-9


def test_validation():
    GaussianSource((1e-6, 1.5e-6), 0.2e-6, np.sin(np.linspace(0, 10, 100)))
    with pytest.raises(ValueError, match="width"):
        GaussianSource((1e-6, 1e-6), 0.0, [1.0])
    with pytest.raises(ValueError, match="position"):
        GaussianSource((1e-6,), 1e-7, [1.0])


def test_profile_2d_on_a_node():
    src = GaussianSource((40 * DX, 25 * DX), 1.3 * DX, [1.0])
    start, prof = gaussian_profile(src, (51, 81), DX, False)
    # 4 widths = 5.2 cells: nodes 20..30 in y and 35..45 in x
    assert start == (20, 35) and prof.shape == (11, 11)
    assert prof[5, 5] == 1.0 and prof.max() == 1.0
    np.testing.assert_allclose(prof, prof[::-1, :], rtol=1e-12)
    np.testing.assert_allclose(prof, prof[:, ::-1], rtol=1e-12)
    np.testing.assert_allclose(prof, prof.T, rtol=1e-12)
    # separable: the outer product of its centre row and column
    np.testing.assert_allclose(prof, np.outer(prof[:, 5], prof[5, :]), rtol=1e-12)
    np.testing.assert_allclose(prof[5, 6], np.exp(-1.0 / (2 * 1.3 ** 2)), rtol=1e-12)


def test_profile_2d_between_nodes_and_clipped():
    src = GaussianSource((40.5 * DX, 2.0 * DX), 1.3 * DX, [1.0])
    start, prof = gaussian_profile(src, (51, 81), DX, False)
    # x: |k - 40.5| <= 5.2 -> 36..45; y: clipped at the lower edge -> 0..7
    assert start == (0, 36) and prof.shape == (8, 10)
    np.testing.assert_allclose(prof[:, :5], prof[:, :4:-1], rtol=1e-12)
    assert np.argmax(prof[:, 0]) == 2


def test_profile_3d_z_on_cell_centres_and_nearest_fallback():
    src = GaussianSource((10 * DX, 12 * DX, 10 * DX), 0.8 * DX, [1.0])
    start, prof = gaussian_profile(src, (20, 25, 25), DX, True)  # Ez (nz, ny+1, nx+1)
    # z centres (k + 1/2) dx: |k + 0.5 - 10| <= 3.2 -> 7..12, symmetric pair 9, 10
    assert start[0] == 7 and prof.shape[0] == 6
    np.testing.assert_allclose(prof[2], prof[3], rtol=1e-12)
    far = GaussianSource((-30 * DX, 12 * DX, 10 * DX), DX, [1.0])
    start, prof = gaussian_profile(far, (20, 25, 25), DX, True)
    assert start[2] == 0 and prof.shape[2] == 1


def test_interpolated_array_signal():
    sig = np.arange(5.0)
    got = [interpolate_signal(sig, (i + 0.5) * 1e-17, 1e-17) for i in range(7)]
    np.testing.assert_allclose(got, [0.5, 1.5, 2.5, 3.5, 4.0, 0.0, 0.0], rtol=1e-12)
    assert interpolate_signal(sig, -0.3e-17, 1e-17) == 0.0
    assert interpolate_signal([], 0.0, 1.0) == 0.0


def test_coefficient_and_program_group():
    grid = (30, 50)
    eps = np.full(grid, 2.085)
    eps[12:18, :] = 12.25
    dt = 0.99 * DX / (o.LIGHT_SPEED * np.sqrt(2))
    pulse = GaussianPulse(1.93e14, 2.5e13)
    src = GaussianSource((20 * DX, 15 * DX), 1.5 * DX, pulse)
    prog = compile_program(eps, DX, dt, sources=[src], num_steps=40, device="cpu")
    coeff, wave, starts = prog.source_groups[2]   # (pre_e, Ez)
    eps_ez = o.site_sample_tm(eps, "Ez")
    cs = lower_gaussian_source(src, eps_ez, DX, dt, 0.0, 40)
    s0 = tuple(cs.start)
    region = tuple(slice(a, a + n) for a, n in zip(s0, cs.coeff.shape))
    _, prof = gaussian_profile(src, eps_ez.shape, DX, False)
    np.testing.assert_allclose(cs.coeff, -prof * dt / (EPS_0 * eps_ez[region]), rtol=1e-12)
    # 2D slabs travel as (z, y, x) with z = 0
    assert starts.numpy().tolist() == [[0, *s0]]
    np.testing.assert_allclose(coeff.numpy()[0].reshape(cs.coeff.shape), cs.coeff, rtol=1e-6)
    t = o.sample_times(0.0, dt, 40)
    np.testing.assert_allclose(wave.numpy()[0][:40], o.gaussian_pulse(t, 1.93e14, 2.5e13),
                               rtol=1e-5, atol=1e-7)
    assert all(g is None for i, g in enumerate(prog.source_groups) if i != 2)
