"""Molecular viscosity laws used by the 3D solver (SI units).

Water follows the IAPWS-style four-term power law in reduced temperature
T/300 K; air follows Sutherland's law.  Same laws and coefficients as
``iceflow3d/viscosity.py:5-24`` in the reference (the coefficients are a
published physical table and have to match exactly).
"""

import math

# (coefficient, exponent) pairs of mu/1e-6 Pa s = sum c * (T/300 K)**e.
WATER_VISCOSITY_TERMS = (
    (280.68, -1.9),
    (511.45, -7.7),
    (61.131, -19.6),
    (0.45903, -40.0),
)

_SUTHERLAND_MU0 = 1.716e-5
_SUTHERLAND_T0 = 273.15
_SUTHERLAND_C = 110.33333


def water_dynamic_viscosity_pa_s(temperature_c):
    """Dynamic viscosity of liquid water, valid on [-20, 110] degC."""
    if not math.isfinite(temperature_c) or not -20 <= temperature_c <= 110:
        raise ValueError("IAPWS water viscosity requires [-20, 110] degC")
    reduced = (temperature_c + 273.15) / 300
    return 1e-6 * sum(c * reduced**e for c, e in WATER_VISCOSITY_TERMS)


def air_dynamic_viscosity_pa_s(temperature_c):
    """Sutherland's law for dry air."""
    if not math.isfinite(temperature_c) or temperature_c <= -273.15:
        raise ValueError("air temperature must exceed absolute zero")
    kelvin = temperature_c + 273.15
    return (
        _SUTHERLAND_MU0
        * (kelvin / _SUTHERLAND_T0) ** 1.5
        * (_SUTHERLAND_T0 + _SUTHERLAND_C)
        / (kelvin + _SUTHERLAND_C)
    )
