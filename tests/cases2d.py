"""Shared 2D cases: the reference tests' small falling case and the default run."""

import math

from projecticy_amd.iceflow2d.config import ThermalBoundary, ThermalBoundarySet, ThermalConfig, create_iceflow_config
from projecticy_amd.iceflow2d.workloads import default_case


def thermal_config(update_interval=2):
    """iceflow2d/test_iceflow.py:27-44: hot Dirichlet side and bottom walls."""
    hot = ThermalBoundary.dirichlet(90.0)
    return ThermalConfig(
        boundaries=ThermalBoundarySet(left=hot, right=hot, bottom=hot, top=ThermalBoundary.adiabatic()),
        initial_water_temperature_c=90.0, initial_ice_temperature_c=0.0, initial_air_temperature_c=20.0,
        update_interval_lbm_steps=update_interval, water_buoyancy_model="linear",
        buoyancy_reference_temperature_c=90.0, moving_body_scheme="body_ale")


def falling_config(**overrides):
    """iceflow2d/test_iceflow.py:47-77: 32x48, 8x8 ice at 5 degrees above a 24-row pool."""
    nx, ny, boundary = 32, 48, 2
    values = dict(
        resolution=(nx, ny), dx=5.0e-4, reference_velocity=4.0, rho_water=1000.0, rho_air=1.25, rho_ice=917.0,
        viscosity_water=1.0e-6, viscosity_air=1.5e-5, sigma=0.072, gravity=(0.0, -9.8), phase_warmup_steps=2,
        well_balanced_hydrostatics=True, water_width_fraction=(nx - boundary) / nx, water_height_fraction=24.0 / ny,
        ice_width_fraction=8.0 / nx, ice_height_fraction=8.0 / ny, ice_base_y_cells=30.0, boundary_cells=boundary,
        ice_initial_angle=math.radians(5.0), ice_fixed=False, rigid_boundary_scheme="unified",
        air_interface_relaxation_time=0.8, thermal=thermal_config())
    values.update(overrides)
    return create_iceflow_config(**values)


def submerged_config(**overrides):
    """A melting case: the 32x48 tank with the ice lowered into the bath."""
    values = dict(ice_base_y_cells=12.0, ice_initial_angle=math.radians(3.0), phase_warmup_steps=4)
    values.update(overrides)
    return falling_config(**values)


def default_run(warmup=40):
    """The reference's default 100x200 run with a shortened phase warm-up."""
    return default_case(phase_warmup_steps=warmup)
