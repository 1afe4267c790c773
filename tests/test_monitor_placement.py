"""Flux monitors placed by physical centre and size (host lowering).

Pinned by:
- the reference's own expectations for snapping: the 3D plane of
  ``tests/contracts/test_device_placement.py:6-49`` (plane cell 4 at 0.9 um,
  z cells [1, 4), y cells [1, 6); shifted by 0.2 um: plane cell 6) and the 2D
  lattice points and normals of
  ``tests/contracts/test_public_input_validation.py:181-195``;
- closed forms: a plane on cell edges and on an E-node plane gives exactly
  the index-space plan, with the reference's area per sample
  dx^2 n/(n-1) m/(m-1); an arbitrary plane's plan reproduces a field linear
  in x, y and z at the sample coordinates (the centres of equal parts of
  the physical extent, on the unsnapped normal coordinate);
- the 2D direct sampling: weight 1 at lattice nodes inside a component's
  array and 0 outside, the sign -1 of a y-normal line.
"""

from __future__ import annotations

import numpy as np
import pytest

from beamz_amd.lattice import OFFSETS_3D, component_shapes, plane_gather_plan
from beamz_amd.monitors import (FluxMonitor, line_grid_points_2d, line_normal_2d, lower_monitor,
                                snap_cell_center, snap_edge_interval)

UM = 1e-6


def test_reference_plane_snap_expectations():
    dx = 0.2 * UM
    mon = FluxMonitor(center=(1.0 * UM, 0.7 * UM, 0.5 * UM), size=(0.0, 0.8 * UM, 0.6 * UM),
                      freqs=(2e14,), name="m3d")
    assert mon.normal == "x"
    idx, coord = snap_cell_center(mon.center[0], dx, 10)
    assert idx == 4 and coord == pytest.approx(0.9 * UM, rel=1e-12)
    assert snap_edge_interval(0.3 * UM, 1.1 * UM, dx, 8) == (1, 6)
    assert snap_edge_interval(0.2 * UM, 0.8 * UM, dx, 5) == (1, 4)
    assert snap_cell_center(1.2 * UM, dx, 10)[0] == 6
    grid = (5, 8, 10)
    cm = lower_monitor(mon, grid, component_shapes(grid), dx, 0)
    assert cm.index == 4 and dict(cm.ranges) == {"z": (1, 4), "y": (1, 6)}
    assert cm.points == 15 and cm.placed and cm.position == pytest.approx(1.0 * UM)


def test_reference_2d_line_points_and_normals():
    horizontal = FluxMonitor(center=(2.0, 1.0, 0.0), size=(4.0, 0.0, 1.0), freqs=[2e14])
    vertical = FluxMonitor(center=(2.0, 1.0, 0.0), size=(0.0, 2.0, 1.0), freqs=[2e14])
    assert (horizontal.normal, vertical.normal) == ("y", "x")
    xs, ys = line_grid_points_2d((0.0, 1.0), (4.0, 1.0), 1.0)
    assert list(zip(xs.tolist(), ys.tolist())) == [(0, 1), (1, 1), (2, 1), (3, 1), (4, 1)]
    xs, ys = line_grid_points_2d((2.0, 0.0), (2.0, 2.0), 1.0)
    assert list(zip(xs.tolist(), ys.tolist())) == [(2, 0), (2, 1), (2, 2)]
    assert line_normal_2d((2.0, 0.0), (2.0, 2.0)) == ("x", 1.0)
    assert line_normal_2d((0.0, 1.0), (4.0, 1.0)) == ("y", -1.0)
    assert line_normal_2d((0.0, 0.0), (1.0, 1.0)) is None


def test_validation():
    with pytest.raises(ValueError, match="exactly one zero"):
        FluxMonitor(center=(0.0, 0.0), size=(0.0, 1.0), freqs=[1e14])  # (0, 1, 0): two zeros
    with pytest.raises(ValueError, match="non-negative"):
        FluxMonitor(center=(0.0, 0.0, 0.0), size=(0.0, -1.0, 1.0), freqs=[1e14])
    with pytest.raises(ValueError, match="finite"):
        FluxMonitor(center=(np.inf, 0.0, 0.0), size=(0.0, 1.0, 1.0), freqs=[1e14])
    with pytest.raises(ValueError, match="both"):
        FluxMonitor(center=(0.0, 0.0, 0.0), freqs=[1e14])
    with pytest.raises(ValueError, match="not both"):
        FluxMonitor(index=3, center=(0.0, 0.0, 0.0), size=(0.0, 1.0, 1.0), freqs=[1e14])
    with pytest.raises(ValueError, match="disagrees"):
        FluxMonitor("y", center=(0.0, 0.0, 0.0), size=(0.0, 1.0, 1.0), freqs=[1e14])
    with pytest.raises(ValueError, match="normal and index"):
        FluxMonitor("x", freqs=[1e14])


def test_edge_interval_clamps_and_widens():
    assert snap_edge_interval(-5.0, 3.0, 1.0, 10) == (0, 3)
    assert snap_edge_interval(8.5, 40.0, 1.0, 10) == (8, 10)
    # a bound within 1e-12 of an edge counts as on it
    assert snap_edge_interval(2.0 + 1e-13, 5.0 - 1e-13, 1.0, 10) == (2, 5)
    # narrower than a cell: the cell around the centre, or min_cells around it
    assert snap_edge_interval(3.2, 3.4, 1.0, 10) == (3, 4)
    assert snap_edge_interval(3.2, 3.4, 1.0, 10, min_cells=3) == (3, 6)
    assert snap_edge_interval(9.2, 9.4, 1.0, 10, min_cells=3) == (7, 10)


def test_aligned_plane_matches_index_space_plan():
    dx = 25e-9
    grid = (12, 14, 20)
    shapes = component_shapes(grid)
    # x-normal plane on the E-node plane 7, y cells [2, 12), z cells [1, 9)
    placed = FluxMonitor(center=(7 * dx, 7 * dx, 5 * dx), size=(0.0, 10 * dx, 8 * dx), freqs=[2e14])
    cm = lower_monitor(placed, grid, shapes, dx, 0)
    assert dict(cm.ranges) == {"z": (1, 9), "y": (2, 12)}
    for c in ("Ey", "Ez", "Hy", "Hz"):
        want_i, want_w = plane_gather_plan(c, shapes[c], dx, "x", 7, {"z": (1, 9), "y": (2, 12)})
        got_i, got_w = cm.plans[c]
        np.testing.assert_array_equal(got_i, want_i)
        np.testing.assert_allclose(got_w, want_w, rtol=0, atol=1e-12)
    assert cm.plans["Ex"] is None and cm.plans["Hx"] is None
    # the reference's area per sample: median spacing x n / (n - 1) on each axis
    assert cm.measure == pytest.approx(dx * dx * (10 / 9) * (8 / 7), rel=1e-12)
    assert cm.normal_sign == 1.0
    index_space = lower_monitor(FluxMonitor("x", 7, [2e14], ranges={"z": (1, 9), "y": (2, 12)}),
                                grid, shapes, dx, 0)
    assert index_space.measure == pytest.approx(dx * dx)


@pytest.mark.parametrize("normal", ["x", "y", "z"])
def test_unaligned_plane_reproduces_a_linear_field(normal):
    dx = 1.0
    grid = (11, 13, 15)  # (z, y, x)
    shapes = component_shapes(grid)
    center = {"x": (6.3, 6.1, 5.45), "y": (7.7, 6.6, 4.9), "z": (7.2, 5.3, 4.35)}[normal]
    size = {"x": (0.0, 5.3, 4.1), "y": (6.2, 0.0, 3.3), "z": (7.1, 4.9, 0.0)}[normal]
    mon = FluxMonitor(center=center, size=size, freqs=[1e14])
    cm = lower_monitor(mon, grid, shapes, dx, 0)
    at = {"x": 0, "y": 1, "z": 2}
    tangential = {"x": ("z", "y"), "y": ("z", "x"), "z": ("y", "x")}[normal]
    spans = {}
    for a in tangential:
        n = dict(cm.ranges)[a][1] - dict(cm.ranges)[a][0]
        lo = center[at[a]] - 0.5 * size[at[a]]
        spans[a] = lo + (np.arange(n) + 0.5) * size[at[a]] / n
    gu, gv = np.meshgrid(spans[tangential[0]], spans[tangential[1]], indexing="ij")
    want_coords = {tangential[0]: gu.ravel(), tangential[1]: gv.ravel(),
                   normal: np.full(gu.size, center[at[normal]])}
    coef = {"x": 0.7, "y": -1.3, "z": 2.1}
    want = sum(coef[a] * want_coords[a] for a in coef) + 0.25
    for c, plan in cm.plans.items():
        if plan is None:
            continue
        shape = shapes[c]
        zz, yy, xx = np.meshgrid(*[(np.arange(shape[d]) + OFFSETS_3D[c][d]) * dx for d in range(3)],
                                 indexing="ij")
        field = coef["x"] * xx + coef["y"] * yy + coef["z"] * zz + 0.25
        idx, w = plan
        np.testing.assert_allclose(w.sum(axis=1), 1.0, atol=1e-12)
        np.testing.assert_allclose((field.reshape(-1)[idx] * w).sum(axis=1), want, atol=1e-9)
    assert cm.index == snap_cell_center(center[at[normal]], dx, grid[2 - at[normal]])[0]
    # area per sample: extent / (n - 1) on each tangential axis
    n = [dict(cm.ranges)[a][1] - dict(cm.ranges)[a][0] for a in tangential]
    want_area = np.prod([size[at[a]] / (k - 1) for a, k in zip(tangential, n)])
    assert cm.measure == pytest.approx(want_area, rel=1e-12)


def test_2d_line_direct_sampling_and_sign():
    dx = 20e-9
    grid = (30, 40)  # (y, x)
    shapes = component_shapes(grid)
    # y-normal line on node row 30 (the top edge) across x nodes 5..35
    top = FluxMonitor(center=(20 * dx, 30 * dx), size=(30 * dx, 0.0, 1e-6), freqs=[2e14])
    cm = lower_monitor(top, grid, shapes, dx, 0)
    assert cm.normal == "y" and cm.normal_sign == -1.0 and cm.measure == pytest.approx(dx)
    assert cm.points == 31 and cm.index == 30 and cm.cell_range == (5, 36)
    xs = np.arange(5, 36)
    np.testing.assert_array_equal(cm.plans["Ez"][0][:, 0], 30 * 41 + xs)
    np.testing.assert_array_equal(cm.plans["Ez"][1][:, 0], 1.0)
    # Hx has 30 rows: row 30 lies outside it, so its samples weigh 0
    np.testing.assert_array_equal(cm.plans["Hx"][1][:, 0], 0.0)
    assert cm.plans["Hy"] is None
    # x-normal line at x = 12.4 cells: node 12, y nodes 3..27, sign +1, element dx
    side = FluxMonitor(center=(12.4 * dx, 15 * dx), size=(0.0, 24 * dx, 1e-6), freqs=[2e14])
    cs = lower_monitor(side, grid, shapes, dx, 1)
    assert cs.normal == "x" and cs.normal_sign == 1.0 and cs.points == 25 and cs.index == 12
    np.testing.assert_array_equal(cs.plans["Hy"][0][:, 0], np.arange(3, 28) * 40 + 12)
    np.testing.assert_array_equal(cs.plans["Hy"][1][:, 0], 1.0)


def test_placed_port_plane_alignment():
    from beamz_amd.sparams import placed_port_plane
    dx = 25e-9
    grid = (56, 64, 170)
    shapes = component_shapes(grid)
    aligned = lower_monitor(FluxMonitor(center=(70 * dx, 32 * dx, 28 * dx), size=(0.0, 44 * dx, 36 * dx),
                                        freqs=[2e14]), grid, shapes, dx, 0)
    assert dict(aligned.ranges) == {"z": (10, 46), "y": (10, 54)}
    assert placed_port_plane(aligned, dx) == 70
    between = lower_monitor(FluxMonitor(center=(70.5 * dx, 32 * dx, 28 * dx),
                                        size=(0.0, 44 * dx, 36 * dx), freqs=[2e14]), grid, shapes, dx, 0)
    off_edges = lower_monitor(FluxMonitor(center=(70 * dx, 32.3 * dx, 28 * dx),
                                          size=(0.0, 44 * dx, 36 * dx), freqs=[2e14]), grid, shapes, dx, 0)
    assert placed_port_plane(between, dx) is None and placed_port_plane(off_edges, dx) is None
    # an extent clamped at the domain edge keeps its physical sample spacing: not a port plane
    clamped = lower_monitor(FluxMonitor(center=(70 * dx, 32 * dx, 28 * dx), size=(0.0, 70 * dx, 36 * dx),
                                        freqs=[2e14]), grid, shapes, dx, 0)
    assert dict(clamped.ranges)["y"] == (0, 64) and placed_port_plane(clamped, dx) is None


def test_placed_monitors_are_refused_as_ports():
    from types import SimpleNamespace
    from beamz_amd.sparams import Port, port_waves
    dx = 20e-9
    grid = (30, 40)
    cm = lower_monitor(FluxMonitor(center=(12 * dx, 15 * dx), size=(0.0, 24 * dx, 1e-6), freqs=[2e14],
                                   name="p"), grid, component_shapes(grid), dx, 0)
    prog = SimpleNamespace(monitors=[cm], eps_site={"Ez": np.ones((31, 41))})
    with pytest.raises(ValueError, match="off the lattice"):
        port_waves(SimpleNamespace(program=prog), Port("in", "p", "x"), [2e14])


def test_field_monitors_and_ragged_arenas():
    """The reference's ragged-arena case (``tests/integration/test_compiled_engine.py:1226-1262``):
    two Ez field monitors on a 1 x 1 domain at resolution 0.1, with 4 and 1
    frequencies, packed back to back."""
    from beamz_amd.monitors import FieldMonitor, assign_offsets
    grid = (10, 10)
    shapes = component_shapes(grid)
    many = FieldMonitor(center=(0.3, 0.5, 0.0), size=(0.0, 0.8, 0.0), freqs=(1e9, 2e9, 3e9, 4e9),
                        fields=("Ez",), name="many_points")
    few = FieldMonitor(center=(0.7, 0.5, 0.0), size=(0.0, 0.2, 0.0), freqs=(2e9,), fields=("Ez",),
                       name="few_points")
    cm = [lower_monitor(m, grid, shapes, 0.1, i) for i, m in enumerate((many, few))]
    assert [m.points for m in cm] == [9, 3]          # nodes y = 1..9 and 4..6
    assert [k for k, v in cm[0].plans.items() if v is not None] == ["Ez"]
    values, weights = assign_offsets(cm)
    assert (cm[0].value_offset, cm[1].value_offset) == (0, 6 * 4 * 9)
    assert (cm[0].weight_offset, cm[1].weight_offset) == (0, 4)
    assert values == 6 * 4 * 9 + 6 * 1 * 3 and weights == 5
    # field monitors need one zero extent at least and known components
    with pytest.raises(ValueError, match="at least one zero"):
        FieldMonitor(center=(0.0, 0.0, 0.0), size=(1.0, 1.0, 1.0), freqs=(1e9,))
    with pytest.raises(ValueError, match="Unsupported"):
        FieldMonitor(center=(0.0, 0.0, 0.0), size=(0.0, 1.0, 1.0), freqs=(1e9,), fields=("Bz",))
    line = lower_monitor(FieldMonitor(center=(5.0, 4.0, 3.0), size=(0.0, 4.0, 0.0), freqs=(1e9,)),
                         (8, 9, 10), component_shapes((8, 9, 10)), 1.0, 0)
    assert dict(line.ranges) == {"z": (3, 4), "y": (2, 6)} and line.points == 4
    assert all(p is not None for p in line.plans.values())



def test_capture_boxes_bound_every_gathered_cell():
    """The host boxes of the pair kernel's monitor capture: for each
    (monitor, component) the inclusive bounding box of the corners the gather
    reads (non-zero weight, index inside the component), -1 where none is read;
    a plane at x = 20 of a 13 x 17 x 130 grid reads E_y, E_z on that plane and
    H_y, H_z on the two half-planes beside it."""
    import beamz_amd as bz
    from beamz_amd.lattice import component_shapes
    from beamz_amd.monitors import capture_boxes, lower_monitor, pack_monitors
    grid = (13, 17, 130)
    shapes = component_shapes(grid)
    freqs = np.array([2e14])
    mons = [lower_monitor(bz.FluxMonitor("x", 20, freqs, ranges={"z": (2, 11), "y": (3, 14)}), grid,
                          shapes, 40e-9, 0),
            lower_monitor(bz.FluxMonitor("z", 6, freqs, ranges={"y": (2, 13), "x": (7, 121)}), grid,
                          shapes, 40e-9, 1)]
    idx, wts, fr, masks, counts, codes, windows = pack_monitors(mons)
    boxes = capture_boxes(idx, wts, masks, shapes)
    assert boxes.shape == (2, 6, 6) and boxes.dtype == np.int32
    names = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
    for m in range(2):
        for c, name in enumerate(names):
            live = (wts[m, c] != 0).reshape(-1)
            if masks[m, c] == 0 or not live.any():
                assert (boxes[m, c] == -1).all(), (m, name)
                continue
            z, y, x = np.unravel_index(idx[m, c].reshape(-1)[live], shapes[name])
            assert tuple(boxes[m, c]) == (z.min(), z.max(), y.min(), y.max(), x.min(), x.max())
    # x-normal flux: Ey and Ez on x = 20, Hy and Hz on the half-planes 19 and 20
    assert boxes[0, 1, 4:].tolist() == [20, 20] and boxes[0, 2, 4:].tolist() == [20, 20]
    assert boxes[0, 4, 4:].tolist() == [19, 20] and boxes[0, 5, 4:].tolist() == [19, 20]
    assert (boxes[0, 0] == -1).all() and (boxes[0, 3] == -1).all()
    # Ey on z nodes 2..11 (cell centres 2.5 .. 10.5 between nodes), y cells 3..13
    assert boxes[0, 1, :4].tolist() == [2, 11, 3, 13]
