"""MI355X-native drop-in for the reference ``iceflow2d`` package."""

from .config import (
    IceFlowConfig,
    LatticeScales,
    MovingBodyThermalScheme,
    MovingBodyThermalTotals,
    PhaseChangeProperties,
    ThermalBoundary,
    ThermalBoundarySet,
    ThermalConfig,
    create_iceflow_config,
)

__all__ = [
    "IceFlowConfig",
    "LatticeScales",
    "MovingBodyThermalScheme",
    "MovingBodyThermalTotals",
    "PhaseChangeProperties",
    "IceFlow2D",
    "ThermalBoundary",
    "ThermalBoundarySet",
    "ThermalConfig",
    "create_iceflow_config",
]


def __getattr__(name):
    if name == "IceFlow2D":
        from .simulator import IceFlow2D

        return IceFlow2D
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
