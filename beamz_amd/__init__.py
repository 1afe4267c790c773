"""MI355X-native (gfx950) Yee/CPML time-stepping backend for BeamZ simulations.

The compute path is ``libbeamz_hip.so`` (hand-written HIP kernels behind the
C-ABI in ``include/beamz_hip.h``); this package lowers simulations on the host
and drives that library.  There is no CPU fallback.
"""

from . import _abi
from .api import Simulation, SimulationResults, courant_dt, resolve_backend
from .boundaries import PEC, PML
from .constants import EPS_0, ETA_0, LIGHT_SPEED, MU_0, nm, um
from .grid import RectilinearGrid
from .modes import ModeSource, ModeSpec
from .monitors import FieldMonitor, FluxMonitor
from .native import HipBackendUnavailable, HipLaunchError
from .program import SimulationState, compile_program, initial_state
from .sources import CustomSource, GaussianPulse, GaussianSource, SoftSource
from .termination import AutoTermination, RunTermination

__all__ = [
    "RectilinearGrid",
    "Simulation", "SimulationResults", "SimulationState", "compile_program", "initial_state",
    "courant_dt", "resolve_backend", "PEC", "PML", "FieldMonitor", "FluxMonitor", "CustomSource",
    "GaussianPulse", "GaussianSource", "SoftSource", "ModeSource", "ModeSpec", "AutoTermination",
    "RunTermination", "HipBackendUnavailable",
    "HipLaunchError",
    "EPS_0", "ETA_0", "LIGHT_SPEED", "MU_0", "nm", "um",
]
__version__ = _abi.COMPONENT_VERSION
