"""CPU oracle for the Yee/CPML time-stepping path.

TEST INFRASTRUCTURE ONLY.  Nothing in the ``beamz_amd`` package may import,
link or execute anything in this directory; only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker (or, for ``bench.py``, as the timed CPU baseline).

Contents
--------
``fdtd64``
    FP64 NumPy restatement of the reference step (3D Yee + CPML + PEC ghosts,
    2D TMxy, slab sources, weighted-gather DFT monitors, flux post-processing).
``port.c`` / ``Makefile``
    A plain C (OpenMP) FP32 restatement of the 3D step, timed as the CPU
    baseline (``cpu_baseline.kind == "port"``).  The reference's own CPU path
    is JAX, which is not installed in this image, so no ``_ref`` build exists.
"""
