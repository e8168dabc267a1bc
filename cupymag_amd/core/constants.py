"""Physical constants (cupymag/core/constants.py:17-30)."""
import math

MU0 = 1.25663706127e-6   # vacuum permeability (H/m)
GAMMA = 1.76e11          # gyromagnetic ratio (rad/(s T))
sqrt2, sqrt3, sqrt6 = math.sqrt(2.0), math.sqrt(3.0), math.sqrt(6.0)
inv_sqrt2, inv_sqrt3, inv_sqrt6 = 1.0 / sqrt2, 1.0 / sqrt3, 1.0 / sqrt6
