"""Validated configuration of the 2D falling-and-melting ice solver.

Public schema of ``iceflow2d/config.py:1-679``: the same class names, field
names, defaults, derived geometry, exception types and messages, so saved
configurations (``to_dict``) and callers interchange with the reference.
Lattice cells are (i, j) with j vertical.
"""

from __future__ import annotations

import math
from dataclasses import asdict, dataclass, field, fields
from typing import Any, Literal

DEFAULT_REFERENCE_VELOCITY_M_S = 4.0

ThermalBoundaryKind = Literal["adiabatic", "dirichlet"]
WaterBuoyancyModel = Literal["linear"]
MovingBodyThermalScheme = Literal["body_ale"]

ADIABATIC = 0
DIRICHLET = 1
BOUNDARY_CODE = {"adiabatic": ADIABATIC, "dirichlet": DIRICHLET}


def finite(name: str, value: Any) -> float:
    """Float coercion with the reference's messages (config.py:12-22)."""
    if isinstance(value, bool):
        raise ValueError(f"{name} must be a finite real number")
    try:
        number = float(value)
    except (TypeError, ValueError, OverflowError) as exc:
        raise ValueError(f"{name} must be a finite real number") from exc
    if not math.isfinite(number):
        raise ValueError(f"{name} must be finite")
    return number


def positive(name: str, value: Any) -> float:
    number = finite(name, value)
    if number <= 0.0:
        raise ValueError(f"{name} must be positive")
    return number


def non_negative(name: str, value: Any) -> float:
    number = finite(name, value)
    if number < 0.0:
        raise ValueError(f"{name} must be non-negative")
    return number


def _whole(value: Any) -> bool:
    """True for an integral non-bool value (ints and integral floats)."""
    if isinstance(value, bool):
        return False
    try:
        return int(value) == value
    except (TypeError, ValueError, OverflowError):
        return False


def _freeze(obj: Any, name: str, value: Any) -> None:
    object.__setattr__(obj, name, value)


@dataclass(frozen=True, slots=True)
class PhaseChangeProperties:
    """Heat capacities, conductivities and latent heat (config.py:51-84)."""

    melting_temperature_c: float = 0.0
    specific_heat_water_j_kg_k: float = 4186.0
    specific_heat_ice_j_kg_k: float = 2100.0
    specific_heat_air_j_kg_k: float = 1005.0
    conductivity_water_w_m_k: float = 0.60
    conductivity_ice_w_m_k: float = 2.20
    conductivity_air_w_m_k: float = 0.026
    latent_heat_j_kg: float = 334000.0

    def __post_init__(self) -> None:
        _freeze(self, "melting_temperature_c", finite("melting_temperature_c", self.melting_temperature_c))
        for item in fields(self)[1:]:
            _freeze(self, item.name, positive(item.name, getattr(self, item.name)))


@dataclass(frozen=True, slots=True)
class ThermalBoundary:
    """One container side (config.py:87-113)."""

    kind: ThermalBoundaryKind = "adiabatic"
    value: float = 0.0

    def __post_init__(self) -> None:
        if self.kind not in BOUNDARY_CODE:
            raise ValueError("thermal boundary kind must be 'adiabatic' or 'dirichlet'")
        value = finite("thermal boundary value", self.value)
        if self.kind == "adiabatic" and value != 0.0:
            raise ValueError("adiabatic thermal boundary value must be zero")
        _freeze(self, "value", value)

    @classmethod
    def adiabatic(cls) -> "ThermalBoundary":
        return cls("adiabatic", 0.0)

    @classmethod
    def dirichlet(cls, temperature_c: float) -> "ThermalBoundary":
        return cls("dirichlet", temperature_c)


@dataclass(frozen=True, slots=True)
class ThermalBoundarySet:
    left: ThermalBoundary = field(default_factory=ThermalBoundary.adiabatic)
    right: ThermalBoundary = field(default_factory=ThermalBoundary.adiabatic)
    bottom: ThermalBoundary = field(default_factory=ThermalBoundary.adiabatic)
    top: ThermalBoundary = field(default_factory=ThermalBoundary.adiabatic)

    def __post_init__(self) -> None:
        for side in ("left", "right", "bottom", "top"):
            if not isinstance(getattr(self, side), ThermalBoundary):
                raise TypeError(f"{side} must be a ThermalBoundary")


def _positive_int(value: Any) -> bool:
    return _whole(value) and int(value) >= 1


@dataclass(frozen=True, slots=True)
class ThermalConfig:
    """Heat and phase-change controls (config.py:132-247)."""

    properties: PhaseChangeProperties = field(default_factory=PhaseChangeProperties)
    boundaries: ThermalBoundarySet = field(default_factory=ThermalBoundarySet)
    initial_water_temperature_c: float = 20.0
    initial_ice_temperature_c: float = 0.0
    initial_air_temperature_c: float = 20.0
    advection_enabled: bool = True
    water_air_interface_adiabatic: bool = True
    update_interval_lbm_steps: int = 1
    solid_liquid_threshold: float = 0.5
    max_fourier_number: float = 0.15
    max_courant_number: float = 0.50
    max_substeps_per_update: int = 64
    water_buoyancy_model: WaterBuoyancyModel = "linear"
    thermal_expansion_water_1_k: float = 2.1e-4
    buoyancy_reference_temperature_c: float | None = None
    scheme: Literal["enthalpy_fv"] = "enthalpy_fv"
    moving_body_scheme: MovingBodyThermalScheme = "body_ale"

    def __post_init__(self) -> None:
        if not isinstance(self.properties, PhaseChangeProperties):
            raise TypeError("properties must be a PhaseChangeProperties")
        if not isinstance(self.boundaries, ThermalBoundarySet):
            raise TypeError("boundaries must be a ThermalBoundarySet")
        for name, allowed in (("scheme", "enthalpy_fv"), ("moving_body_scheme", "body_ale"),
                              ("water_buoyancy_model", "linear")):
            if getattr(self, name) != allowed:
                raise ValueError(f"{name} must be '{allowed}'")
        for name in ("advection_enabled", "water_air_interface_adiabatic"):
            if not isinstance(getattr(self, name), bool):
                raise ValueError(f"{name} must be a boolean")
        if not _positive_int(self.update_interval_lbm_steps):
            raise ValueError("update_interval_lbm_steps must be a positive integer")
        _freeze(self, "update_interval_lbm_steps", int(self.update_interval_lbm_steps))
        if not self.water_air_interface_adiabatic:
            raise ValueError("body_ale currently requires an adiabatic water/air thermal interface")
        threshold = finite("solid_liquid_threshold", self.solid_liquid_threshold)
        if not 0.0 < threshold < 1.0:
            raise ValueError("solid_liquid_threshold must be in (0, 1)")
        _freeze(self, "solid_liquid_threshold", threshold)
        if self.buoyancy_reference_temperature_c is None:
            _freeze(self, "buoyancy_reference_temperature_c",
                    finite("initial_water_temperature_c", self.initial_water_temperature_c))
        for name in ("initial_water_temperature_c", "initial_ice_temperature_c",
                     "initial_air_temperature_c", "buoyancy_reference_temperature_c"):
            _freeze(self, name, finite(name, getattr(self, name)))
        melting = self.properties.melting_temperature_c
        if self.initial_water_temperature_c < melting:
            raise ValueError("initial_water_temperature_c must not be below melting_temperature_c")
        if self.initial_ice_temperature_c > melting:
            raise ValueError("initial_ice_temperature_c must not exceed melting_temperature_c")
        # A Dirichlet half-cell doubles one face coefficient: a corner cell
        # sees six face contributions, hence the 1/6 bound.
        fourier = positive("max_fourier_number", self.max_fourier_number)
        if fourier > 1.0 / 6.0:
            raise ValueError("max_fourier_number must not exceed 1/6")
        _freeze(self, "max_fourier_number", fourier)
        courant = positive("max_courant_number", self.max_courant_number)
        if courant > 1.0:
            raise ValueError("max_courant_number must not exceed one")
        _freeze(self, "max_courant_number", courant)
        if not _positive_int(self.max_substeps_per_update):
            raise ValueError("max_substeps_per_update must be a positive integer")
        _freeze(self, "max_substeps_per_update", int(self.max_substeps_per_update))
        _freeze(self, "thermal_expansion_water_1_k",
                non_negative("thermal_expansion_water_1_k", self.thermal_expansion_water_1_k))


@dataclass(frozen=True, slots=True)
class LatticeScales:
    """Physical size of one cell and one LBM step (config.py:250-313)."""

    dx_m: float
    dt_s: float
    reference_density_kg_m3: float
    reference_lattice_velocity: float = 0.1

    def __post_init__(self) -> None:
        for item in fields(self):
            _freeze(self, item.name, positive(item.name, getattr(self, item.name)))

    @property
    def velocity_scale_m_s(self) -> float:
        return self.dx_m / self.dt_s

    @property
    def reference_velocity_m_s(self) -> float:
        return self.reference_lattice_velocity * self.velocity_scale_m_s

    @classmethod
    def from_reference_velocity(cls, *, dx_m: float, reference_velocity_m_s: float,
                                reference_density_kg_m3: float,
                                reference_lattice_velocity: float = 0.1) -> "LatticeScales":
        spacing = positive("dx_m", dx_m)
        speed = positive("reference_velocity_m_s", reference_velocity_m_s)
        lattice_speed = positive("reference_lattice_velocity", reference_lattice_velocity)
        return cls(dx_m=spacing, dt_s=lattice_speed * spacing / speed,
                   reference_density_kg_m3=reference_density_kg_m3,
                   reference_lattice_velocity=lattice_speed)

    @classmethod
    def from_iceflow_config(cls, config: Any) -> "LatticeScales":
        try:
            return cls.from_reference_velocity(dx_m=config.dx, reference_velocity_m_s=config.reference_velocity,
                                               reference_density_kg_m3=config.rho_water)
        except (AttributeError, TypeError) as exc:
            raise TypeError("config must provide dx, reference_velocity, and rho_water") from exc


@dataclass(frozen=True, slots=True)
class MovingBodyThermalTotals:
    initial_body_mass_kg_m: float
    solid_body_mass_kg_m: float
    melted_mass_kg_m: float
    water_mass_kg_m: float
    total_mass_kg_m: float
    body_sensible_energy_j_m: float
    water_sensible_energy_j_m: float
    latent_energy_j_m: float
    total_energy_j_m: float


def _default_thermal_config() -> ThermalConfig:
    """The insulated 90 degC bath with an 8-step thermal interval (config.py:330-340)."""
    return ThermalConfig(boundaries=ThermalBoundarySet(), initial_water_temperature_c=90.0,
                         initial_ice_temperature_c=0.0, initial_air_temperature_c=20.0,
                         update_interval_lbm_steps=8, buoyancy_reference_temperature_c=90.0)


# Ranges checked by IceFlowConfig: (name, lower exclusive, upper inclusive).
_FRACTIONS = ("water_width_fraction", "water_height_fraction", "ice_width_fraction", "ice_height_fraction")


@dataclass(slots=True)
class IceFlowConfig:
    """One body-ALE falling-and-melting ice run (config.py:343-658)."""

    resolution: tuple[int, int] = (200, 400)
    dx: float = 1.25e-4
    reference_velocity: float = DEFAULT_REFERENCE_VELOCITY_M_S
    rho_water: float = 1000.0
    rho_air: float = 1.25
    rho_ice: float = 917.0
    viscosity_water: float = 1.0e-6
    viscosity_air: float = 1.5e-5
    sigma: float = 0.072
    gravity: tuple[float, float] = (0.0, -9.8)
    smagorinsky_constant: float = 0.1
    interface_width: float = 5.0
    mobility: float = 0.1
    phase_warmup_steps: int = 500
    thermal: ThermalConfig = field(default_factory=_default_thermal_config)
    well_balanced_hydrostatics: bool = True
    water_width_fraction: float = 0.9875
    water_height_fraction: float = 0.60125
    ice_width_fraction: float = 0.3225
    ice_height_fraction: float = 0.16125
    ice_base_y_cells: float = 254.6672141068609
    boundary_cells: int = 3
    ice_initial_velocity: tuple[float, float] = (0.0, 0.0)
    ice_initial_angle: float = math.radians(5.0)
    ice_initial_angular_velocity: float = 0.0
    ice_fixed: bool = False
    rigid_boundary_scheme: Literal["unified"] = "unified"
    linear_damping: float = 1.0
    angular_damping: float = 0.9990
    volume_projection_tolerance: float = 1.0e-8
    volume_projection_interface_cutoff: float = 1.0e-3
    volume_projection_max_shift: float = 0.50
    volume_projection_max_iterations: int = 64
    air_interface_relaxation_time: float | None = 0.8
    output_dir: str = "iceflow2d/outputs/iceflow2d/coupled_falling_ice_melting_2d"
    show_gui: bool = False
    save_npz: bool = False

    def __post_init__(self) -> None:
        for item in fields(self):
            if item.type == "float":
                setattr(self, item.name, finite(item.name, getattr(self, item.name)))
        if self.rigid_boundary_scheme != "unified":
            raise ValueError("rigid_boundary_scheme must be 'unified'")
        self._check_grid()
        if self.well_balanced_hydrostatics is not True:
            raise ValueError("well_balanced_hydrostatics must be True")
        if not isinstance(self.thermal, ThermalConfig):
            raise TypeError("thermal must be a ThermalConfig")
        self._check_physics()
        self._check_body()
        self._check_projection()
        self._check_layout()
        self._check_relaxation()
        if not isinstance(self.output_dir, str) or not self.output_dir.strip():
            raise ValueError("output_dir must be a non-empty string")

    def _check_grid(self) -> None:
        try:
            nx, ny = self.resolution
        except (TypeError, ValueError) as exc:
            raise ValueError("resolution must contain exactly two integers") from exc
        if not (_whole(nx) and _whole(ny)):
            raise ValueError("resolution must contain exactly two integers")
        nx, ny = int(nx), int(ny)
        self.resolution = (nx, ny)
        if nx <= 0 or ny <= 0:
            raise ValueError("resolution components must be positive")
        if not _whole(self.boundary_cells):
            raise ValueError("boundary_cells must be an integer")
        self.boundary_cells = int(self.boundary_cells)
        if self.boundary_cells < 1:
            raise ValueError("boundary_cells must be at least one")
        if min(nx, ny) <= 2 * self.boundary_cells + 4:
            raise ValueError("resolution is too small for the requested boundary thickness")
        if not _whole(self.phase_warmup_steps):
            raise ValueError("phase_warmup_steps must be an integer")
        self.phase_warmup_steps = int(self.phase_warmup_steps)
        if self.phase_warmup_steps < 0:
            raise ValueError("phase_warmup_steps must be non-negative")

    def _check_physics(self) -> None:
        positive("dx", self.dx)
        self.reference_velocity = positive("reference_velocity", self.reference_velocity)
        for name in ("rho_water", "rho_air", "rho_ice", "viscosity_water", "viscosity_air"):
            positive(name, getattr(self, name))
        for name in ("sigma", "smagorinsky_constant"):
            if finite(name, getattr(self, name)) < 0.0:
                raise ValueError(f"{name} must be non-negative")
        positive("interface_width", self.interface_width)
        positive("mobility", self.mobility)
        if len(self.gravity) != 2:
            raise ValueError("gravity must contain exactly two components")
        self.gravity = tuple(finite(f"gravity[{k}]", g) for k, g in enumerate(self.gravity))
        for name in _FRACTIONS:
            value = finite(name, getattr(self, name))
            if not 0.0 < value <= 1.0:
                raise ValueError(f"{name} must be in (0, 1]")

    def _check_body(self) -> None:
        if finite("ice_base_y_cells", self.ice_base_y_cells) < self.boundary_cells:
            raise ValueError("ice_base_y_cells must not overlap the bottom wall")
        if len(self.ice_initial_velocity) != 2:
            raise ValueError("ice_initial_velocity must contain exactly two components")
        self.ice_initial_velocity = tuple(
            finite(f"ice_initial_velocity[{k}]", v) for k, v in enumerate(self.ice_initial_velocity))
        finite("ice_initial_angle", self.ice_initial_angle)
        finite("ice_initial_angular_velocity", self.ice_initial_angular_velocity)
        if self.ice_fixed is not False:
            raise ValueError("ice_fixed must be False")
        for name in ("linear_damping", "angular_damping"):
            if not 0.0 < finite(name, getattr(self, name)) <= 1.0:
                raise ValueError(f"{name} must be in (0, 1]")

    def _check_projection(self) -> None:
        positive("volume_projection_tolerance", self.volume_projection_tolerance)
        cutoff = finite("volume_projection_interface_cutoff", self.volume_projection_interface_cutoff)
        if not 0.0 < cutoff < 0.5:
            raise ValueError("volume_projection_interface_cutoff must be in (0, 0.5)")
        positive("volume_projection_max_shift", self.volume_projection_max_shift)
        if not _whole(self.volume_projection_max_iterations):
            raise ValueError("volume_projection_max_iterations must be an integer")
        self.volume_projection_max_iterations = int(self.volume_projection_max_iterations)
        if self.volume_projection_max_iterations < 1:
            raise ValueError("volume_projection_max_iterations must be positive")
        tau = self.air_interface_relaxation_time
        if tau is not None:
            if isinstance(tau, bool):
                raise ValueError("air_interface_relaxation_time must be a real number")
            tau = finite("air_interface_relaxation_time", tau)
            if not 0.5 < tau <= 2.0:
                raise ValueError("air_interface_relaxation_time must be in (0.5, 2]")
            self.air_interface_relaxation_time = tau

    def _check_layout(self) -> None:
        nx, ny, wall = self.nx, self.ny, self.boundary_cells
        if self.water_width <= wall or self.water_height <= wall:
            raise ValueError("water block must contain at least one interior cell")
        if self.water_width > nx - wall or self.water_height > ny - wall:
            raise ValueError("water block must fit inside the container walls")
        if self.ice_width <= 1 or self.ice_height <= 1:
            raise ValueError("ice rectangle must be at least two cells in each direction")
        if self.water_width != nx - wall:
            raise ValueError("hydrostatic reference requires a full-width horizontal pool")
        if self.water_height != ny - wall and abs(float(self.gravity[0])) > 1.0e-14:
            raise ValueError("a two-phase hydrostatic reference requires vertical gravity")
        # Axis-aligned extent of the initially rotated rectangle.
        c, s = abs(math.cos(float(self.ice_initial_angle))), abs(math.sin(float(self.ice_initial_angle)))
        half_x = 0.5 * (c * self.ice_width + s * self.ice_height)
        half_y = 0.5 * (s * self.ice_width + c * self.ice_height)
        cx, cy = self.ice_initial_center
        if cx - half_x < wall or cx + half_x > nx - wall:
            raise ValueError("initial ice rectangle overlaps a side wall")
        if cy - half_y < wall - 1.0e-9 or cy + half_y > ny - wall:
            raise ValueError("initial ice rectangle overlaps the bottom or top wall")
        if self.thermal.moving_body_scheme != "body_ale":
            raise ValueError("thermal moving_body_scheme must be 'body_ale'")
        if not self.thermal.water_air_interface_adiabatic:
            raise ValueError("thermal water/air interface must be adiabatic")
        if self.water_height >= ny - wall:
            raise ValueError("thermal phase change requires a water/air free surface for "
                             "continuous phase-volume projection")

    def _check_relaxation(self) -> None:
        for name in ("viscosity_water", "viscosity_air"):
            nu = float(getattr(self, name)) * 0.1 / (float(self.dx) * float(self.reference_velocity))
            tau = 0.5 + 3.0 * nu
            if not math.isfinite(tau) or tau <= 0.5:
                raise ValueError(f"{name} gives an unstable lattice relaxation time ({tau})")

    @property
    def nx(self) -> int:
        return int(self.resolution[0])

    @property
    def ny(self) -> int:
        return int(self.resolution[1])

    @property
    def water_width(self) -> int:
        return max(self.boundary_cells + 2, int(self.nx * self.water_width_fraction))

    @property
    def water_height(self) -> int:
        return max(self.boundary_cells + 2, int(self.ny * self.water_height_fraction))

    @property
    def ice_width(self) -> int:
        return max(2, int(self.nx * self.ice_width_fraction))

    @property
    def ice_height(self) -> int:
        return max(2, int(self.ny * self.ice_height_fraction))

    @property
    def ice_initial_center(self) -> tuple[float, float]:
        """Cell-aligned centre: the rectangle's left edge falls on a cell face."""
        left = (self.nx - self.ice_width) // 2
        return (left + 0.5 * self.ice_width, float(self.ice_base_y_cells) + 0.5 * self.ice_height)

    @property
    def ice_mass_lattice(self) -> float:
        return float(self.rho_ice / self.rho_water * self.ice_width * self.ice_height)

    @property
    def ice_inertia_lattice(self) -> float:
        return self.ice_mass_lattice * (self.ice_width**2 + self.ice_height**2) / 12.0

    def to_dict(self) -> dict:
        return asdict(self)


def create_iceflow_config(**overrides) -> IceFlowConfig:
    names = {item.name for item in fields(IceFlowConfig)}
    for key in overrides:
        if key not in names:
            raise TypeError(f"Unknown IceFlowConfig option: {key}")
    return IceFlowConfig(**overrides)
