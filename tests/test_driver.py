"""CPU: the experiment driver's wire format against the reference's archived run.

The archived manifest of results/quadrupole/full14_uncoupled_tau5_29angles
records configuration_hash = SHA-256 of its configuration record; rebuilding
that record here must give the same hash (every key, value and float format).
The archived raw_sample_yields.csv lists its rows in the report's task order.
"""

import csv
import json
import math
import os
from dataclasses import asdict

import numpy as np

from rpm_amd import core, driver, model

ARCHIVED_UNCOUPLED_HASH = "b0b8037cc352fed6857a4d72bad4bb7107050df712381c3711584550757fea63"
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "archived_full14_yields.csv")


def test_uncoupled_configuration_digest_matches_archived_manifest():
    assert driver.UncoupledExperimentConfig().digest == ARCHIVED_UNCOUPLED_HASH
    # any change of a recorded value changes the digest
    assert driver.UncoupledExperimentConfig(samples=3).digest != ARCHIVED_UNCOUPLED_HASH


def test_published_quadrupole_tensors_match_their_construction():
    for key, vals in model.QUADRUPOLE_PRINCIPAL_VALUES_MT.items():
        built = model.quadrupole_tensor_in_hyperfine_frame(model._BY_KEY[key].tensor_mT, vals,
                                                           model._QUADRUPOLE_FRAME[key])
        np.testing.assert_allclose(built, model.QUADRUPOLE_TENSORS_MT[key], rtol=0, atol=1e-16)
        assert abs(np.trace(model.QUADRUPOLE_TENSORS_MT[key])) < 1e-16


def test_task_order_and_keys_match_archived_report():
    with open(GOLDEN) as f:
        rows = [r for r in csv.DictReader(f) if r["electron_interactions"] == "D=J=0"]
    got = driver.tasks(driver.UncoupledExperimentConfig())
    assert [(float(r["scale"]), float(r["theta_deg"]), int(r["sample_id"])) for r in rows] == \
        [(t.scale, t.theta_deg, t.sample_id) for t in got]
    assert driver.Task(0.0, 0.0, 0).key == "Q0_th000.000_s000"
    assert driver.Task(1.0, 92.0, 3).key == "Q1_th092.000_s003"
    assert str(driver.cell_path("out", driver.Task(1.0, 5.0, 12))).endswith(
        os.path.join("checkpoints", "cells", "Q1_th005.000_s012.npz"))


class _FakeStats:
    def __init__(self):
        self.accepted_steps, self.rejected_steps, self.matvecs = 3, 0, 72


class _FakeResult:
    def __init__(self, times, dims):
        self.times_us = times
        self.observable_values = 0.5 + 0.5 * np.cos(times)
        self.observable_error_bounds = 1e-9 * times
        self.state_error_estimates = 1e-10 * times
        self.final_state = np.zeros(int(np.prod(dims)), dtype=complex)
        self.final_state_device = None
        from rpm_amd.engine import PropagationStats
        self.stats = PropagationStats(3, 0, 72, 24, 0.05, 0.05, 1e-10, 1e-15)


def test_cell_round_trip_passes_the_reference_validation(tmp_path):
    cfg = driver.ExperimentConfig(samples=1, angles_deg=(0., 90.), lifetime_us=.02, tail_bound=math.exp(-2),
                                  state_error_budget=1e-7, smoke=True)
    driver.ensure_manifest(tmp_path, cfg)
    man = json.loads((tmp_path / "manifest.json").read_text())
    assert man["configuration_hash"] == cfg.digest and man["configuration"]["smoke"] is True
    dims = cfg.dimensions()
    task = driver.Task(1.0, 90.0, 0)
    x0 = core.random_phase_initial_state(dims, cfg.seed, 0)
    vals = driver.cell_values(task, cfg, x0, _FakeResult(cfg.sweep.time_grid_us, dims), elapsed=1.0,
                              construction=0.1, initial_singlet=1.0)
    assert sorted(vals) == sorted(["yield", "coarse_yield", "quadrature_error", "yield_propagation_estimate",
                                   "yield_tail_bound", "times_us", "singlet_population",
                                   "population_error_estimate", "state_error_estimates", "final_state",
                                   "initial_singlet", "metadata_json"])
    driver.atomic_npz(driver.cell_path(tmp_path, task), vals)
    values, meta = driver.read_cell(tmp_path, task, cfg)
    assert meta["task"] == asdict(task) and meta["stats"]["matvecs"] == 72
    assert values["yield"] == float(vals["yield"])
    # a cell of another configuration is refused
    other = driver.ExperimentConfig(samples=2, angles_deg=(0., 90.), lifetime_us=.02, tail_bound=math.exp(-2),
                                    state_error_budget=1e-7, smoke=True)
    try:
        driver.read_cell(tmp_path, task, other)
    except ValueError:
        pass
    else:
        raise AssertionError("read_cell accepted a cell of another configuration")
