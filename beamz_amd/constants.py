"""Physical constants (``beamz/const.py:1-10``)."""

import math

LIGHT_SPEED = 299792458.0
EPS_0 = 8.8541878128e-12
MU_0 = 1.25663706212e-6
ETA_0 = math.sqrt(MU_0 / EPS_0)
um = 1e-6
nm = 1e-9
