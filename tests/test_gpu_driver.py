"""GPU: the batched experiment driver writes checkpoint cells the reference's validation accepts,
with yields equal to the oracle's per-trajectory run of the same task (1e-9)."""

import csv
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from oracle import core as ocore  # noqa: E402
from oracle import model as omodel  # noqa: E402
from rpm_amd import driver  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "archived_full14_yields.csv")


def _smoke():
    # R/rpm_full14_quadrupole_experiment.py:101-103
    return driver.ExperimentConfig(samples=1, angles_deg=(0., 90.), lifetime_us=.02, tail_bound=math.exp(-2),
                                   state_error_budget=1e-7, smoke=True)


def test_run_cells_smoke_matches_oracle_and_validates(tmp_path):
    cfg = _smoke()
    written = driver.run_cells(cfg, tmp_path, batch_size=2)
    print("cells written", [t.key for t in written], flush=True)
    assert sorted(t.key for t in written) == sorted(t.key for t in driver.tasks(cfg))
    nuclei = (omodel.FAD_NUCLEI[0], omodel.FAD_NUCLEI[1], omodel.TRP_NUCLEI[0])
    radicals = ("FAD", "FAD", "TrpH")
    ocfg = ocore.SweepConfig(seed=cfg.seed, lifetime_us=cfg.lifetime_us, survival_tail_bound=cfg.tail_bound,
                             state_error_budget=cfg.state_error_budget)
    for task in driver.tasks(cfg):
        values, meta = driver.read_cell(tmp_path, task, cfg)
        model = ocore.ProductionModel(master_seed=cfg.seed, lifetime_us=cfg.lifetime_us,
                                      quadrupole_scale=task.scale, ratio_D_over_J=-3.0)
        h = ocore.build_hamiltonian(task.theta_deg, cfg.field_mT, model=model, nuclei=nuclei, radicals=radicals)
        x0 = ocore.random_phase_initial_state(h.dimensions, cfg.seed, task.sample_id)
        ref = ocore.run_trajectory(h, x0, ocfg)
        print(task.key, values["yield"], ref["yield"], flush=True)
        assert abs(values["yield"] - ref["yield"]) < 1e-9
        assert abs(values["quadrature_error"] - ref["quadrature_error"]) < 1e-9
        assert meta["stats"]["accepted_steps"] == ref["result"].stats.accepted_steps
    rows = driver.raw_sample_rows(tmp_path, cfg, "D/J=-3")
    with open(GOLDEN) as f:
        archived_columns = next(csv.reader(f))
    assert set(archived_columns) <= set(rows[0]) | {"checkpoint"}
    # resume skips existing cells
    assert driver.run_cells(cfg, tmp_path) == []
