"""Host-side checks of the batched 2D path (no GPU: lowering and validation)."""

from __future__ import annotations

import numpy as np
import pytest

import beamz_amd as bz
from beamz_amd.batch import BatchedSimulation2D


def _sim(grid, **kw):
    return bz.Simulation(np.full(grid, 2.085), 20e-9, boundaries=[bz.PML(cells=4)], **kw)


def test_rejects_3d_members():
    with pytest.raises(ValueError, match="2D"):
        BatchedSimulation2D([_sim((8, 20, 20))])


def test_rejects_monitors():
    mon = bz.FluxMonitor("x", 10, np.array([1.9e14]))
    with pytest.raises(NotImplementedError, match="monitors"):
        BatchedSimulation2D([_sim((20, 30), monitors=[mon])])


def test_rejects_non_pre_e_sources():
    src = bz.CustomSource("Ez", (5, 5), np.ones((3, 1)), np.ones(4), "e")
    batch = BatchedSimulation2D([_sim((20, 30), sources=[src])])
    with pytest.raises(NotImplementedError, match="pre-E"):
        batch.initial_state()


def test_signature_and_cells():
    batch = BatchedSimulation2D([_sim((20, 30)), _sim((40, 25))])
    assert batch.cells == 20 * 30 + 40 * 25
    assert batch.signature() == (((20, 30), 4), ((40, 25), 4))


def test_empty_batch():
    with pytest.raises(ValueError):
        BatchedSimulation2D([])
