"""Seeded dam-break workloads built through the reference's ``IceFlowConfig``.

The reference's own dam-break harness (iceflow3d/examples/
dambreak_four_ice_melting_3d.py:142-233) builds a 450x150x90 tank with a
floor pool and four axis-aligned cubes.  The benchmark case here is the
256x128x128 variant: a water column filling x < 0.35 nx up to 0.6 ny at
80 degC, no floor pool, interface width 4, and four 16-cell cubes at seeded
non-overlapping positions in the air region, each with a seeded uniformly
random rotation axis and an angle uniform in [0, 30] degrees.

Draw order (numpy.random.default_rng(seed)), per cube, repeated until the
configuration validates (walls, column and separating-axis overlap):
centre x, centre y, centre z (uniform over the admissible box), rotation
axis (three standard normals, normalised), angle (uniform degrees).
"""

import math

import numpy as np

from .config import IceBodyConfig, IceFlowConfig, ThermalConfig
from .rigid import rotation_matrix, separating_axis

DX_M = 1.0 / 3000.0  # the reference's fine grid spacing (dt = 1/120000 s)


def dam_break_four_ice(nx=256, ny=128, nz=128, seed=2026, dx_m=DX_M, cube_cells=16, interface_width=4.0,
                       water_temperature_c=80.0, max_angle_degrees=30.0, phase_warmup_steps=2000):
    """The D3Q27 HOME dam break with four seeded ice cubes (lattice sizes in cells)."""
    boundary = max(1, round(0.001 / dx_m))
    water_width = int(round(0.35 * nx))
    water_height = int(round(0.6 * ny))
    rng = np.random.default_rng(seed)
    bodies = []
    centres, rotations = [], []
    half = np.full(3, cube_cells / 2.0)
    attempts = 0
    while len(bodies) < 4:
        attempts += 1
        if attempts > 10000:
            raise RuntimeError("could not place four non-overlapping cubes")
        lo = np.array([water_width + 2.0, boundary + 1.0, boundary + 1.0])
        hi = np.array([nx - boundary - 1.0, ny - boundary - 1.0, nz - boundary - 1.0])
        centre = lo + (hi - lo) * rng.random(3)
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        angle = math.radians(rng.uniform(0.0, max_angle_degrees))
        q = (math.cos(angle / 2), *(math.sin(angle / 2) * axis))
        r = rotation_matrix(q)
        extent = np.abs(r) @ half
        if np.any(centre - extent < lo) or np.any(centre + extent > hi):
            continue
        if any(separating_axis(centre, r, half, c, rr, half) is not None for c, rr in zip(centres, rotations)):
            continue
        centres.append(centre)
        rotations.append(r)
        bodies.append(IceBodyConfig(tuple(float(v) for v in centre), (cube_cells,) * 3, tuple(float(v) for v in q)))
    return IceFlowConfig(
        nx=nx, ny=ny, nz=nz,
        domain_size_m=(nx * dx_m, ny * dx_m, nz * dx_m),
        boundary_cells=boundary,
        water_width=water_width,
        water_height=water_height,
        floor_water_depth_cells=0,
        ice_bodies=tuple(bodies),
        interface_width=interface_width,
        phase_warmup_steps=phase_warmup_steps,
        thermal=ThermalConfig(initial_water_temperature_c=water_temperature_c),
    )


def coarse_dam_break_four_ice(seed=2026, **kwargs):
    """The same geometry halved: 128x64x64 with 8-cell cubes (dx doubled)."""
    return dam_break_four_ice(128, 64, 64, seed=seed, dx_m=2 * DX_M, cube_cells=8, **kwargs)
