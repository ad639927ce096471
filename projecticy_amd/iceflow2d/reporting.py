"""Samples and output files of the 2D falling-ice harness (iceflow2d/reporting.py
and iceflow2d/examples/coupled_falling_ice_melting_2d.py:52-923).

A ``Sample`` is one synchronised host copy of the coupled state: the fields
as [ny, nx] arrays (rows are y, as files and images use them), the
phase-change and conservation diagnostics, and the moving body's motion and
melt bookkeeping.  Its attribute names are the column names of
``history.csv`` and the array names of ``fields.npz``, which keep the
reference's names, order and number formatting (``%.17g``): they are the
file format that post-processing reads.  The PNG/GIF sequences (velocity,
vorticity, temperature) and the final plot are drawn with matplotlib in a
layout of this module's own.
"""

from dataclasses import dataclass
import math
import sys
import time
from pathlib import Path

import numpy as np

from .config import LatticeScales

SCENARIO_LABEL = "falling-ice melting"


def _positive(name, value):
    number = float(value)
    if not math.isfinite(number) or number <= 0.0:
        raise ValueError(f"{name} must be finite and positive")
    return number


def _non_negative(name, value):
    number = float(value)
    if not math.isfinite(number) or number < 0.0:
        raise ValueError(f"{name} must be finite and non-negative")
    return number


FIELD_COLUMNS = (
    "physical_time_s", "thermal_time_s", "lbm_steps", "thermal_substeps", "solid_volume_cells",
    "sharp_geometry_volume_cells", "solid_area_m2", "melted_fraction", "equivalent_square_side_m",
    "equivalent_uniform_melt_depth_m", "generated_water_volume_cells", "phase_change_contraction_cells",
    "water_volume_target_cells", "water_volume_current_cells", "water_volume_residual_cells",
    "total_enthalpy_j_m", "boundary_heat_input_j_m", "energy_residual_j_m",
)
MOTION_COLUMNS = (
    "body_active", "body_center_x_m", "body_center_y_m", "body_velocity_x_m_s", "body_velocity_y_m_s",
    "body_angle_rad", "body_angular_velocity_rad_s", "body_mass_lattice", "body_mass_kg_m",
    "body_inertia_lattice", "body_inertia_kg_m", "body_solid_center_local_x_cells",
    "body_solid_center_local_y_cells", "cumulative_melted_mass_lattice", "cumulative_melted_momentum_x_lattice",
    "cumulative_melted_momentum_y_lattice", "cumulative_fluid_melt_momentum_x_lattice",
    "cumulative_fluid_melt_momentum_y_lattice", "melt_momentum_residual_x_lattice",
    "melt_momentum_residual_y_lattice", "cumulative_melted_angular_momentum_lattice",
    "cumulative_fluid_melt_angular_momentum_lattice", "melt_angular_momentum_residual_lattice",
    "ale_water_residual_cells", "ale_energy_residual_j_m", "phase_aperture_water_residual_cells",
    "phase_aperture_energy_residual_j_m", "phase_aperture_capacity_margin_cells",
    "melt_injection_mass_residual_kg_m", "thermal_initial_total_mass_kg_m", "thermal_total_mass_kg_m",
    "thermal_total_mass_residual_kg_m", "water_mean_temperature_c", "aperture_energy_correction_abs_j_m",
)
HISTORY_COLUMNS = FIELD_COLUMNS + MOTION_COLUMNS
_INTEGER_COLUMNS = {"lbm_steps", "thermal_substeps", "body_active"}
FIELD_ARRAYS = ("temperature_c", "liquid_fraction", "enthalpy_j_m3", "water_phase", "solid", "sdf",
                "phase_change_material", "velocity_lattice", "physical_velocity_lattice")


@dataclass(slots=True)
class Sample:
    physical_time_s: float
    thermal_time_s: float
    lbm_steps: int
    thermal_substeps: int
    temperature_c: np.ndarray
    liquid_fraction: np.ndarray
    enthalpy_j_m3: np.ndarray
    water_phase: np.ndarray
    solid: np.ndarray
    sdf: np.ndarray
    phase_change_material: np.ndarray
    velocity_lattice: np.ndarray
    physical_velocity_lattice: np.ndarray
    solid_volume_cells: float
    sharp_geometry_volume_cells: float
    solid_area_m2: float
    melted_fraction: float
    equivalent_square_side_m: float
    equivalent_uniform_melt_depth_m: float
    generated_water_volume_cells: float
    phase_change_contraction_cells: float
    water_volume_target_cells: float
    water_volume_current_cells: float
    water_volume_residual_cells: float
    total_enthalpy_j_m: float
    boundary_heat_input_j_m: float
    energy_residual_j_m: float
    body_active: bool
    body_center_x_m: float
    body_center_y_m: float
    body_velocity_x_m_s: float
    body_velocity_y_m_s: float
    body_angle_rad: float
    body_angular_velocity_rad_s: float
    body_mass_lattice: float
    body_mass_kg_m: float
    body_inertia_lattice: float
    body_inertia_kg_m: float
    body_solid_center_local_x_cells: float
    body_solid_center_local_y_cells: float
    cumulative_melted_mass_lattice: float
    cumulative_melted_momentum_x_lattice: float
    cumulative_melted_momentum_y_lattice: float
    cumulative_fluid_melt_momentum_x_lattice: float
    cumulative_fluid_melt_momentum_y_lattice: float
    melt_momentum_residual_x_lattice: float
    melt_momentum_residual_y_lattice: float
    cumulative_melted_angular_momentum_lattice: float
    cumulative_fluid_melt_angular_momentum_lattice: float
    melt_angular_momentum_residual_lattice: float
    ale_water_residual_cells: float
    ale_energy_residual_j_m: float
    phase_aperture_water_residual_cells: float
    phase_aperture_energy_residual_j_m: float
    phase_aperture_capacity_margin_cells: float
    melt_injection_mass_residual_kg_m: float
    thermal_initial_total_mass_kg_m: float
    thermal_total_mass_kg_m: float
    thermal_total_mass_residual_kg_m: float
    water_mean_temperature_c: float
    aperture_energy_correction_abs_j_m: float


def requested_snapshot_times(end_time_s, interval_s):
    """0, every interval up to the end, and the end itself (reporting.py:158-166)."""
    end, interval = _positive("end_time_s", end_time_s), _positive("output_interval_s", interval_s)
    count = int(math.floor(end / interval + 1.0e-12))
    times = [0.0] + [k * interval for k in range(1, count + 1)]
    if not math.isclose(times[-1], end, rel_tol=0.0, abs_tol=1.0e-12):
        times.append(end)
    return times


def snapshot_step_targets(config, end_time_s, interval_s):
    """LBM steps of the samples: each requested time rounded up to a whole
    thermal update, at least one update, strictly increasing (reporting.py:169-180)."""
    dt = LatticeScales.from_iceflow_config(config).dt_s
    interval = int(config.thermal.update_interval_lbm_steps)
    targets = [0]
    for t in requested_snapshot_times(end_time_s, interval_s)[1:]:
        updates = math.ceil(t / (dt * interval) - 1.0e-12)
        step = max(interval, int(updates) * interval)
        if step > targets[-1]:
            targets.append(step)
    return targets


def _read(source, name, default=None):
    """A 0-d field, derived value, method or attribute of ``source``; the
    default when the solver does not expose it."""
    value = getattr(source, name, None)
    if value is None:
        return default
    if callable(value) and not hasattr(value, "to_numpy"):
        value = value()
    try:
        value = value[None]
    except (KeyError, TypeError, IndexError, AttributeError):
        pass
    return value


def _scalar(source, name, default):
    value = np.asarray(_read(source, name, default), dtype=np.float64).reshape(-1)
    return float(value[0]) if value.size == 1 else float(default)


def _pair(source, name, default):
    value = np.asarray(_read(source, name, default), dtype=np.float64).reshape(-1)
    return (float(value[0]), float(value[1])) if value.size >= 2 else (float(default[0]), float(default[1]))


def _rows_are_y(values, dtype=None):
    array = np.asarray(values, dtype=dtype)
    return np.ascontiguousarray(np.swapaxes(array, 0, 1))


def capture_sample(sim, *, initial_total_enthalpy_j_m, initial_solid_volume_cells, initial_total_mass_kg_m=None):
    """One ``Sample`` of a running ``IceFlow2D`` (reporting.py:183-274 with the
    moving-body terms of coupled_falling_ice_melting_2d.py:490-731)."""
    config, thermal = sim.config, sim.thermal
    totals = thermal.mass_energy_totals()
    solid_volume = float(sim.phase_change_solid_volume_cells())
    geometry_volume = float(sim.phase_change_geometry_volume_cells())
    initial_solid = float(initial_solid_volume_cells)
    melted_solid = max(0.0, initial_solid - solid_volume)
    ratio = float(config.rho_ice / config.rho_water)
    dx = float(config.dx)
    area = solid_volume * dx * dx
    initial_side, side = math.sqrt(initial_solid * dx * dx), math.sqrt(max(0.0, area))
    total_enthalpy = float(totals.total_energy_j_m)
    boundary_heat = float(thermal.boundary_heat_input_j_m[None])
    target, current = float(sim.water_volume_target[None]), float(sim.water_volume_current[None])
    thermal_fields = sim.sample_thermal_fields()
    momentum = np.asarray(sim.momentum_velocity_lattice.to_numpy(), dtype=np.float32)
    force = np.asarray(sim.fluid_acceleration_lattice.to_numpy(), dtype=np.float32)
    physical = momentum + np.float32(0.5) * force  # momentum velocity plus the half-force shift
    scales = LatticeScales.from_iceflow_config(config)
    center = _pair(sim, "body_center", config.ice_initial_center)
    velocity = _pair(sim, "body_velocity", config.ice_initial_velocity)
    local = _pair(sim, "body_local_center_of_mass", (0.0, 0.0))
    melted_momentum = _pair(sim, "cumulative_melted_momentum_lattice", (0.0, 0.0))
    fluid_momentum = _pair(sim, "cumulative_fluid_melt_momentum_lattice", (0.0, 0.0))
    momentum_residual = _pair(sim, "melt_momentum_residual_lattice", (0.0, 0.0))
    mass = _scalar(sim, "body_mass_lattice", solid_volume * ratio)
    initial_mass = float(config.ice_mass_lattice)
    inertia = _scalar(sim, "body_inertia_lattice",
                      0.0 if initial_mass <= 0.0 else config.ice_inertia_lattice * max(0.0, mass) / initial_mass)
    rho_w = float(config.rho_water)
    mass_kg_m = mass * rho_w * dx * dx
    total_mass = float(totals.total_mass_kg_m)
    initial_total = total_mass if initial_total_mass_kg_m is None else float(initial_total_mass_kg_m)
    props = config.thermal.properties
    mean_temperature = props.melting_temperature_c
    if totals.water_mass_kg_m > 0.0:
        mean_temperature += totals.water_sensible_energy_j_m / (totals.water_mass_kg_m * props.specific_heat_water_j_kg_k)
    return Sample(
        physical_time_s=float(sim.physical_time_s), thermal_time_s=float(thermal.time_s),
        lbm_steps=int(sim.steps), thermal_substeps=int(thermal.steps),
        temperature_c=_rows_are_y(thermal_fields["temperature_c"], np.float64),
        liquid_fraction=_rows_are_y(thermal_fields["liquid_fraction"], np.float32),
        enthalpy_j_m3=_rows_are_y(thermal_fields["enthalpy_j_m3"], np.float64),
        water_phase=_rows_are_y(sim.water_phase.to_numpy(), np.float32),
        solid=_rows_are_y(sim.solid_mask.to_numpy(), np.int8),
        sdf=_rows_are_y(sim.body_signed_distance_m.to_numpy(), np.float32),
        phase_change_material=_rows_are_y(thermal_fields["phase_change_material"], np.int8),
        velocity_lattice=_rows_are_y(momentum), physical_velocity_lattice=_rows_are_y(physical),
        solid_volume_cells=solid_volume, sharp_geometry_volume_cells=geometry_volume, solid_area_m2=area,
        melted_fraction=0.0 if initial_solid <= 0.0 else melted_solid / initial_solid,
        equivalent_square_side_m=side, equivalent_uniform_melt_depth_m=0.5 * (initial_side - side),
        generated_water_volume_cells=ratio * melted_solid, phase_change_contraction_cells=(1.0 - ratio) * melted_solid,
        water_volume_target_cells=target, water_volume_current_cells=current,
        water_volume_residual_cells=current - target, total_enthalpy_j_m=total_enthalpy,
        boundary_heat_input_j_m=boundary_heat,
        energy_residual_j_m=total_enthalpy - float(initial_total_enthalpy_j_m) - boundary_heat,
        body_active=bool(_scalar(sim, "body_active", 1.0 if mass > 1.0e-12 else 0.0)),
        body_center_x_m=center[0] * dx, body_center_y_m=center[1] * dx,
        body_velocity_x_m_s=velocity[0] * scales.velocity_scale_m_s,
        body_velocity_y_m_s=velocity[1] * scales.velocity_scale_m_s,
        body_angle_rad=_scalar(sim, "body_angle", config.ice_initial_angle),
        body_angular_velocity_rad_s=_scalar(sim, "body_angular_velocity", config.ice_initial_angular_velocity)
        / scales.dt_s,
        body_mass_lattice=mass, body_mass_kg_m=mass_kg_m, body_inertia_lattice=inertia,
        body_inertia_kg_m=inertia * rho_w * dx**4,
        body_solid_center_local_x_cells=local[0], body_solid_center_local_y_cells=local[1],
        cumulative_melted_mass_lattice=_scalar(sim, "cumulative_melted_mass_lattice", max(0.0, initial_mass - mass)),
        cumulative_melted_momentum_x_lattice=melted_momentum[0],
        cumulative_melted_momentum_y_lattice=melted_momentum[1],
        cumulative_fluid_melt_momentum_x_lattice=fluid_momentum[0],
        cumulative_fluid_melt_momentum_y_lattice=fluid_momentum[1],
        melt_momentum_residual_x_lattice=momentum_residual[0], melt_momentum_residual_y_lattice=momentum_residual[1],
        cumulative_melted_angular_momentum_lattice=_scalar(sim, "cumulative_melted_angular_momentum_lattice", 0.0),
        cumulative_fluid_melt_angular_momentum_lattice=_scalar(
            sim, "cumulative_fluid_melt_angular_momentum_lattice", 0.0),
        melt_angular_momentum_residual_lattice=_scalar(sim, "melt_angular_momentum_residual_lattice", 0.0),
        ale_water_residual_cells=_scalar(sim, "ale_water_residual_cells", 0.0),
        ale_energy_residual_j_m=_scalar(sim, "ale_energy_residual_j_m", 0.0),
        phase_aperture_water_residual_cells=_scalar(sim, "phase_aperture_water_residual_cells", 0.0),
        phase_aperture_energy_residual_j_m=_scalar(thermal, "phase_aperture_energy_residual_j_m", 0.0),
        phase_aperture_capacity_margin_cells=_scalar(sim, "phase_aperture_capacity_margin_cells", 0.0),
        melt_injection_mass_residual_kg_m=_scalar(thermal, "melt_injection_mass_residual_kg_m", 0.0),
        thermal_initial_total_mass_kg_m=initial_total, thermal_total_mass_kg_m=total_mass,
        thermal_total_mass_residual_kg_m=total_mass - initial_total, water_mean_temperature_c=mean_temperature,
        aperture_energy_correction_abs_j_m=_scalar(thermal, "aperture_energy_correction_abs_j_m", 0.0),
    )


def history_row(sample):
    """One ``history.csv`` row: integers as integers, reals as %.17g."""
    row = []
    for name in HISTORY_COLUMNS:
        value = getattr(sample, name)
        row.append(int(value) if name in _INTEGER_COLUMNS else f"{float(value):.17g}")
    return row


def _series(samples, *names, dtype=np.float64):
    if len(names) == 1:
        return np.asarray([getattr(s, names[0]) for s in samples], dtype=dtype)
    return np.asarray([[getattr(s, n) for n in names] for s in samples], dtype=dtype)


def field_arrays(config, samples):
    """The ``fields.npz`` payload (reporting.py:322-384 and
    coupled_falling_ice_melting_2d.py:780-921)."""
    if not samples:
        raise ValueError("at least one snapshot is required")
    centres = lambda n: (np.arange(n, dtype=np.float64) + 0.5) * config.dx  # noqa: E731
    out = {"x_m": centres(config.nx), "y_m": centres(config.ny),
           "physical_time_s": _series(samples, "physical_time_s"), "thermal_time_s": _series(samples, "thermal_time_s"),
           "lbm_steps": _series(samples, "lbm_steps", dtype=np.int64)}
    for name in FIELD_ARRAYS:
        out[name] = np.stack([getattr(s, name) for s in samples])
    for name in ("solid_volume_cells", "sharp_geometry_volume_cells", "solid_area_m2", "melted_fraction",
                 "generated_water_volume_cells", "phase_change_contraction_cells", "water_volume_target_cells",
                 "water_volume_current_cells", "total_enthalpy_j_m", "boundary_heat_input_j_m", "energy_residual_j_m"):
        out[name] = _series(samples, name)
    out["body_active"] = _series(samples, "body_active", dtype=np.int8)
    out["body_center_m"] = _series(samples, "body_center_x_m", "body_center_y_m")
    out["body_velocity_m_s"] = _series(samples, "body_velocity_x_m_s", "body_velocity_y_m_s")
    for name in ("body_angle_rad", "body_angular_velocity_rad_s", "body_mass_lattice", "body_mass_kg_m",
                 "body_inertia_lattice", "body_inertia_kg_m"):
        out[name] = _series(samples, name)
    out["body_solid_center_local_cells"] = _series(samples, "body_solid_center_local_x_cells",
                                                   "body_solid_center_local_y_cells")
    out["water_volume_residual_cells"] = _series(samples, "water_volume_residual_cells")
    out["cumulative_melted_mass_lattice"] = _series(samples, "cumulative_melted_mass_lattice")
    for name in ("cumulative_melted_momentum", "cumulative_fluid_melt_momentum", "melt_momentum_residual"):
        out[name + "_lattice"] = _series(samples, name + "_x_lattice", name + "_y_lattice")
    for name in ("cumulative_melted_angular_momentum_lattice", "cumulative_fluid_melt_angular_momentum_lattice",
                 "melt_angular_momentum_residual_lattice", "ale_water_residual_cells", "ale_energy_residual_j_m",
                 "phase_aperture_water_residual_cells", "phase_aperture_energy_residual_j_m",
                 "phase_aperture_capacity_margin_cells", "melt_injection_mass_residual_kg_m",
                 "thermal_initial_total_mass_kg_m", "thermal_total_mass_kg_m", "thermal_total_mass_residual_kg_m",
                 "water_mean_temperature_c", "aperture_energy_correction_abs_j_m"):
        out[name] = _series(samples, name)
    return out


def write_fields_npz(path, config, samples):
    np.savez_compressed(path, **field_arrays(config, samples))


# ------------------------------------------------------------------ plots
def _wall_mask(config):
    wall = np.zeros((config.ny, config.nx), dtype=bool)
    b = int(config.boundary_cells)
    wall[:b, :] = wall[-b:, :] = True
    wall[:, :b] = wall[:, -b:] = True
    return wall


def _physical_velocity_m_s(config, sample):
    """Water velocity in m/s, zero outside the water (phase >= 0.5)."""
    scale = LatticeScales.from_iceflow_config(config).velocity_scale_m_s
    water = (np.asarray(sample.water_phase) >= 0.5) & (np.asarray(sample.solid) == 0) & ~_wall_mask(config)
    velocity = np.asarray(sample.physical_velocity_lattice, dtype=np.float64) * scale
    return np.where(water[..., None], velocity, 0.0), water


def _vorticity_s_1(config, sample):
    velocity, water = _physical_velocity_m_s(config, sample)
    dvdx = np.gradient(velocity[..., 1], config.dx, axis=1)
    dudy = np.gradient(velocity[..., 0], config.dx, axis=0)
    return np.where(water, dvdx - dudy, np.nan)


def _temperature_c(config, sample):
    air = (np.asarray(sample.water_phase) < 0.5) & (np.asarray(sample.phase_change_material) == 0)
    return np.where(air, np.nan, np.asarray(sample.temperature_c))


def _draw_boundaries(axis, config, sample, extent, colour):
    axis.imshow(np.ma.masked_where(~_wall_mask(config), np.ones((config.ny, config.nx))), origin="lower",
                extent=extent, cmap="gray", vmin=0.0, vmax=2.0, interpolation="nearest")
    phase = np.asarray(sample.water_phase)
    xs = (np.arange(config.nx) + 0.5) * config.dx
    ys = (np.arange(config.ny) + 0.5) * config.dx
    if np.nanmin(phase) < 0.5 < np.nanmax(phase):
        axis.contour(xs, ys, phase, levels=[0.5], colors=colour, linewidths=0.7)
    solid = np.asarray(sample.solid).astype(float)
    if solid.max() > 0.5:
        axis.contour(xs, ys, solid, levels=[0.5], colors="#edae49", linewidths=1.0)


class _Sequence:
    """PNG frames ``<kind>_frames/frame_NNNNN.png`` and ``<kind>_field.gif``
    (reporting.py:774-838, 977-1050, 1156-1224)."""

    def __init__(self, output_dir, config, kind, units, field, cmap, lo, hi):
        from ..iceflow3d.reporting import FrameGif

        self.config, self.kind, self.units, self.field, self.cmap, self.lo, self.hi = (
            config, kind, units, field, cmap, lo, hi)
        self.directory = Path(output_dir) / f"{kind}_frames"
        self.directory.mkdir(parents=True, exist_ok=True)
        for stale in self.directory.glob("frame_*.png"):
            stale.unlink()
        self.animation = Path(output_dir) / f"{kind}_field.gif"
        self.gif = FrameGif(self.animation)
        self.count = 0
        self.observed = [math.inf, -math.inf]

    def append(self, sample):
        import matplotlib

        matplotlib.use("Agg")
        import matplotlib.pyplot as plt

        values = self.field(self.config, sample)
        finite = values[np.isfinite(values)]
        if finite.size:
            self.observed = [min(self.observed[0], float(finite.min())), max(self.observed[1], float(finite.max()))]
        extent = (0.0, self.config.nx * self.config.dx, 0.0, self.config.ny * self.config.dx)
        figure, axis = plt.subplots(figsize=(5.0, 8.0), constrained_layout=True)
        image = axis.imshow(np.ma.masked_invalid(values), origin="lower", extent=extent, cmap=self.cmap,
                            vmin=self.lo, vmax=self.hi, interpolation="nearest", aspect="equal")
        _draw_boundaries(axis, self.config, sample, extent, "white" if self.kind == "temperature" else "black")
        figure.colorbar(image, ax=axis, label=f"{self.kind} ({self.units})")
        axis.set(xlabel="x (m)", ylabel="y (m)",
                 title=f"{SCENARIO_LABEL}: {self.kind}, t = {sample.physical_time_s:.4f} s")
        path = self.directory / f"frame_{self.count:05d}.png"
        figure.savefig(path, dpi=100)
        plt.close(figure)
        self.gif.append(path)
        self.count += 1

    def close(self):
        self.gif.close()
        observed = self.observed if self.count else [None, None]
        return {"directory": self.directory.name, "pattern": "frame_*.png", "frames": self.count,
                "animation": self.animation.name if self.count else None, "units": self.units,
                "observed": observed, "color": [self.lo, self.hi]}


def sequences(output_dir, config, limits):
    """The velocity, vorticity and temperature sequences of one run."""
    speed = lambda c, s: np.where(_physical_velocity_m_s(c, s)[1],  # noqa: E731
                                  np.linalg.norm(_physical_velocity_m_s(c, s)[0], axis=-1), np.nan)
    return {
        "velocity": _Sequence(output_dir, config, "velocity", "m/s", speed, "viridis", 0.0, limits["velocity_max_m_s"]),
        "vorticity": _Sequence(output_dir, config, "vorticity", "s^-1", _vorticity_s_1, "RdBu_r",
                               -limits["vorticity_abs_max_s_1"], limits["vorticity_abs_max_s_1"]),
        "temperature": _Sequence(output_dir, config, "temperature", "degC", _temperature_c, "inferno",
                                 limits["temperature_min_c"], limits["temperature_max_c"]),
    }


def write_final_plot(path, config, sample, *, temperature_max_c):
    """Final temperature and phase state (coupled_falling_ice_melting_2d.py:924-986)."""
    import matplotlib

    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    extent = (0.0, config.nx * config.dx, 0.0, config.ny * config.dx)
    figure, axes = plt.subplots(1, 2, figsize=(10.0, 8.0), constrained_layout=True)
    image = axes[0].imshow(np.ma.masked_invalid(_temperature_c(config, sample)), origin="lower", extent=extent,
                           cmap="inferno", vmin=config.thermal.properties.melting_temperature_c,
                           vmax=temperature_max_c, interpolation="nearest", aspect="equal")
    _draw_boundaries(axes[0], config, sample, extent, "white")
    figure.colorbar(image, ax=axes[0], label="temperature (degC)")
    axes[0].set_title("Final temperature")
    fraction = np.ma.masked_where(np.asarray(sample.phase_change_material) == 0, np.asarray(sample.liquid_fraction))
    image = axes[1].imshow(fraction, origin="lower", extent=extent, cmap="Blues_r", vmin=0.0, vmax=1.0,
                           interpolation="nearest", aspect="equal")
    _draw_boundaries(axes[1], config, sample, extent, "black")
    figure.colorbar(image, ax=axes[1], label="liquid fraction")
    axes[1].set_title("Final phase state")
    for axis in axes:
        axis.set(xlabel="x (m)", ylabel="y (m)")
    figure.suptitle(f"Coupled falling-ice melting, t={sample.physical_time_s:.4f} s, "
                    f"melted={100.0 * sample.melted_fraction:.2f}%, "
                    f"center=({sample.body_center_x_m:.4f}, {sample.body_center_y_m:.4f}) m")
    figure.savefig(path, dpi=120)
    plt.close(figure)


class Progress:
    """A one-line progress report on stderr, at most once a second."""

    def __init__(self, total, *, enabled, label, unit):
        self.total, self.enabled, self.label, self.unit = int(total), bool(enabled), label, unit
        self.done, self.status, self.started, self.shown = 0, "", time.perf_counter(), 0.0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        if self.enabled:
            self._show(force=True)
            print(file=sys.stderr, flush=True)

    def advance(self, count):
        self.done += int(count)
        self._show()

    def set_status(self, text):
        self.status = text
        self._show()

    def _show(self, force=False):
        now = time.perf_counter()
        if not self.enabled or (not force and now - self.shown < 1.0):
            return
        self.shown = now
        share = self.done / self.total if self.total else 1.0
        print(f"\r{self.label}: {self.done}/{self.total} {self.unit} ({100.0 * share:5.1f}%) "
              f"{now - self.started:7.1f} s | {self.status}", end="", file=sys.stderr, flush=True)

