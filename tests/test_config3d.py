"""Configuration validation (iceflow3d/tests/test_config_and_rigid.py:29-96,
iceflow3d/config.py:10-349): same defaults, same accepted and rejected
values, same exception types."""

from dataclasses import replace

import numpy as np
import pytest

from projecticy_amd.iceflow3d import IceBodyConfig, IceFlowConfig, LatticeScales, ThermalConfig
from projecticy_amd.iceflow3d.config import integer
from projecticy_amd.iceflow3d.rigid import rotation_matrix


def test_defaults_and_units():
    c = IceFlowConfig()
    s = LatticeScales.from_iceflow_config(c)
    assert (c.nx, c.ny, c.nz) == (450, 150, 90)
    assert c.floor_water_depth_cells == 31
    assert (c.water_width, c.water_height) == (198, 138)
    assert len(c.bodies) == 4
    np.testing.assert_allclose([b.center_cells for b in c.bodies],
                               [[255, 21, 45], [309, 21, 45], [363, 21, 45], [417, 21, 45]])
    assert abs(s.dx_m - 1 / 3000) < 1e-15
    assert abs(s.dt_s - 1 / 120000) < 1e-18
    assert abs(s.mass_kg - 1000 * (1 / 3000) ** 3) < 1e-20
    assert c.density_gradient_source_limit == 0.0
    assert not c.gas_state_repair


@pytest.mark.parametrize(
    "kwargs",
    [
        {"nz": 91},
        {"rho_water": float("nan")},
        {"phase_normal_epsilon": -1},
        {"density_gradient_source_limit": 0.1},
        {"cm_acoustic_relaxation_rate": 2},
        {"momentum_collision_model": "cm_hrr"},
        {"gravity": (0, 0, -9.8)},
        {"boundary_cells": True},
        {"water_width": 3},
        {"water_height": 147},
        {"linear_damping": 1.5},
        {"backend": "metal"},
    ],
)
def test_reject_invalid_configuration(kwargs):
    with pytest.raises(ValueError):
        replace(IceFlowConfig(), **kwargs)


def test_reject_overlapping_bodies_and_wall_overlap():
    c = IceFlowConfig()
    with pytest.raises(ValueError):
        replace(c, ice_bodies=(c.bodies[0], c.bodies[0]))
    with pytest.raises(ValueError, match="wall"):
        replace(c, ice_bodies=(replace(c.bodies[0], center_cells=(255, 21, 4)),))


def test_thermal_ranges():
    with pytest.raises(ValueError):
        ThermalConfig(initial_ice_temperature_c=1)
    with pytest.raises(ValueError):
        ThermalConfig(max_fourier_number=0.2)
    with pytest.raises(TypeError):
        ThermalConfig(properties=object())
    with pytest.raises(TypeError):
        replace(IceFlowConfig(), thermal=object())


def test_body_quaternion_is_normalised_and_rotation_orthonormal():
    b = IceBodyConfig((10.0, 10.0, 10.0), (4, 4, 4), (2.0, 0.0, 0.2, 0.0))
    assert abs(np.linalg.norm(b.quaternion_wxyz) - 1) < 1e-15
    r = rotation_matrix(b.quaternion_wxyz)
    np.testing.assert_allclose(r @ r.T, np.eye(3), atol=1e-14)
    with pytest.raises(ValueError):
        IceBodyConfig((1.0, 1.0, 1.0), (4, 4, 4), (0.0, 0.0, 0.0, 0.0))


def test_integer_rejects_bool_fraction_and_negative():
    for bad in (True, 1.5, -1, "x", float("inf")):
        with pytest.raises(ValueError):
            integer("num_steps", bad, 0)
    assert integer("num_steps", 3.0, 0) == 3


def test_config_round_trips_through_dict():
    c = IceFlowConfig()
    d = c.to_dict()
    assert d["nx"] == 450 and d["thermal"]["update_interval_lbm_steps"] == 8
    assert len(d["ice_bodies"]) == 4
