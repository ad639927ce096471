"""Batched 2D ensembles on the GPU: each case of an ``IceFlowEnsemble2D``
against a standalone ``IceFlow2D`` run of the same configuration (bit for
bit with the default block sizes, FP32 and FP16 storage, with and without
the HIP graph), one case against the FP64 oracle, and the autotuner."""

import math

import numpy as np
import pytest

from cases2d import falling_config, submerged_config

pytestmark = pytest.mark.gpu

FIELDS = ("momentum_populations", "phase_populations", "water_phase", "momentum_velocity_lattice",
          "fluid_acceleration_lattice", "pressure_lattice", "solid_mask", "coverage", "water_volume_m2",
          "water_sensible_energy", "body_solid_mass", "body_sensible_energy", "body_center", "body_velocity",
          "body_angle", "body_mass_lattice")


def sized_falling(nx, ny, **overrides):
    """The 32x48 falling case on another grid: same 2-cell walls, 8-cell ice,
    pool 24 rows deep, ice 6 rows above the pool."""
    values = dict(resolution=(nx, ny), water_width_fraction=(nx - 2) / nx, water_height_fraction=24.0 / ny,
                  ice_width_fraction=8.0 / nx, ice_height_fraction=8.0 / ny, ice_base_y_cells=30.0)
    values.update(overrides)
    return falling_config(**values)


def ensemble_configs():
    return [falling_config(), submerged_config(), sized_falling(40, 56, ice_initial_angle=math.radians(9.0)),
            submerged_config(resolution=(48, 40), water_width_fraction=46.0 / 48, water_height_fraction=24.0 / 40,
                             ice_width_fraction=8.0 / 48, ice_height_fraction=8.0 / 40)]


def run_pair(storage, use_graph, steps):
    from projecticy_amd.iceflow2d import IceFlow2D
    from projecticy_amd.iceflow2d.ensemble import IceFlowEnsemble2D

    configs = ensemble_configs()
    ensemble = IceFlowEnsemble2D(configs, population_storage=storage, use_graph=use_graph)
    ensemble.step(steps)
    singles = []
    for config in configs:
        sim = IceFlow2D(config, population_storage=storage)
        sim.step(steps)
        singles.append(sim)
    return ensemble, singles


@pytest.mark.parametrize("storage,use_graph", [("fp32", True), ("fp16", True), ("fp16-rel-sr", True),
                                               ("fp32", False)])
def test_ensemble_cases_equal_standalone_runs(storage, use_graph):
    steps = 9  # four slow thermal updates (interval 2) and the melt-momentum closures
    ensemble, singles = run_pair(storage, use_graph, steps)
    for k, (case, sim) in enumerate(zip(ensemble.cases, singles)):
        a, b = case.state(), sim.state()
        for name in FIELDS:
            np.testing.assert_array_equal(np.asarray(a[name]), np.asarray(b[name]), err_msg=f"case {k} {name}")
        assert case.steps == sim.steps == steps
        assert case.thermal.steps == sim.thermal.steps
        assert case.thermal.time_s == sim.thermal.time_s
    melted = [t.melted_mass_kg_m for t in ensemble.totals()]
    assert max(melted) > 0.0


def test_ensemble_case_matches_oracle():
    from oracle.iceflow2d_oracle import Oracle2D
    from projecticy_amd.iceflow2d.ensemble import IceFlowEnsemble2D
    from test_parity2d import check

    configs = ensemble_configs()
    ensemble = IceFlowEnsemble2D(configs, population_storage="fp32")
    ref = Oracle2D(configs[1])
    for total in (1, 5):
        ensemble.step(total - ensemble.steps)
        ref.step(total - ref.steps)
        check(ensemble.cases[1], ref, f"ensemble case 1 after {total} steps")


def test_autotune_restores_state_and_keeps_parity():
    from projecticy_amd.iceflow2d.ensemble import THREAD_CHOICES, IceFlowEnsemble2D

    configs = ensemble_configs()
    ensemble = IceFlowEnsemble2D(configs, population_storage="fp32")
    ensemble.step(3)
    before = [case.state() for case in ensemble.cases]
    timings = ensemble.autotune(steps=2)
    assert set(timings) == {(c, s) for c in THREAD_CHOICES for s in THREAD_CHOICES}
    assert all(t > 0.0 for t in timings.values())
    assert ensemble.threads == min(timings, key=timings.get)
    assert ensemble.steps == 3
    for k, (case, old) in enumerate(zip(ensemble.cases, before)):
        new = case.state()
        for name in FIELDS:
            np.testing.assert_array_equal(np.asarray(new[name]), np.asarray(old[name]), err_msg=f"{k} {name}")
    # any block size gives the standalone result up to the impulse's summation order
    ensemble.set_threads(512, 128)
    ensemble.step(4)
    reference, _ = run_pair("fp32", True, 7)
    for case, ref in zip(ensemble.cases, reference.cases):
        a, b = case.state(), ref.state()
        for name in ("momentum_populations", "water_phase", "water_sensible_energy"):
            x, y = np.asarray(a[name], np.float64), np.asarray(b[name], np.float64)
            assert np.abs(x - y).max() <= 1e-5 * max(np.abs(y).max(), 1e-30), name
        np.testing.assert_allclose(a["body_center"], b["body_center"], atol=1e-4)
