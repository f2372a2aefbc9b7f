"""Physical constants and tolerances (utils/constants.py:26-33)."""

DRY_TOL = 1e-4
SED_TOL = 1e-5
VEL_TOL = 1e-6
TIMESTEP_TOL = 5e-7
g = 9.81
