"""FP16 population storage (FP32 arithmetic) against the FP64 oracle.

The reference has no FP16 mode; the bar here is the diagnostics the ensemble
reports: ice mass and mean bath temperature, plus the melted mass, on the
submerged melting case after 64 steps (32 slow thermal updates: its
thermal interval is 2).  Tolerances (relative): ice mass 1e-6, bath
temperature 1e-6, melted mass 2e-3 for FP16 phase populations and 1e-3 for
FP32 phase populations, with round-to-nearest or stochastic rounding; FP32
storage meets 1e-4 on all three (the parity tests' bar)."""

import numpy as np
import pytest

from cases2d import submerged_config

pytestmark = pytest.mark.gpu

STEPS = 64


def bath_temperature(state):
    volume, energy = state["water_volume_m2"], state["water_sensible_energy"]
    return energy.sum() / (1000.0 * 4186.0 * volume.sum())


@pytest.fixture(scope="module")
def oracle_run():
    from oracle.iceflow2d_oracle import Oracle2D

    ref = Oracle2D(submerged_config())
    ref.step(STEPS)
    return ref.mass_energy_totals(), bath_temperature(ref.state())


@pytest.mark.parametrize("storage,melt_tol", [("fp32", 1e-4), ("fp16-momentum", 1e-3), ("fp16", 2e-3),
                                              ("fp16-momentum-sr", 1e-3), ("fp16-sr", 2e-3)])
def test_storage_against_oracle(storage, melt_tol, oracle_run):
    from projecticy_amd.iceflow2d import IceFlow2D

    totals, bath = oracle_run
    sim = IceFlow2D(submerged_config(), population_storage=storage)
    sim.step(STEPS)
    got = sim.thermal.mass_energy_totals()
    assert got.solid_body_mass_kg_m == pytest.approx(totals["solid_body_mass_kg_m"], rel=1e-6)
    assert bath_temperature(sim.state()) == pytest.approx(bath, rel=1e-6)
    assert totals["melted_mass_kg_m"] > 0.0
    assert got.melted_mass_kg_m == pytest.approx(totals["melted_mass_kg_m"], rel=melt_tol)
    assert np.isfinite(sim.state()["momentum_populations"]).all()
