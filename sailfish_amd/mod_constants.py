"""Physical and scheme constants of the hot path (modules/mod_constants.py:5-7, :25-27)."""

# Adams-Moulton 3rd-order coefficients (mod_constants.py:5-7)
AM3_2 = 5.0 / 12.0
AM3_1 = 8.0 / 12.0
AM3_0 = -1.0 / 12.0

rho0 = 1025.0        # mean density, kg/m3 (mod_constants.py:25)
g = 10.0             # gravity, m/s2 (mod_constants.py:26)
gorho0 = g / rho0    # (mod_constants.py:27)

pi = 3.14159265358979323846
