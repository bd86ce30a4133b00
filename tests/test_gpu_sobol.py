"""GPU: the Sobol reaction-time estimator (rpm_amd.sobol) against the oracle.

Bars:
- a cell on a 576-level random-tensor pair matches the oracle's cell: yields to 1e-10, propagation
  bounds to 1e-4 relative (error estimates), the same step and matvec counts;
- batched cells (different angles and samples, each with its own reaction times) equal the
  one-at-a-time cells to 1e-12, with the same schedules;
- a full-system cell (dim 221,184) matches the oracle's recorded cell
  (tests/golden/sobol_full_cell.npz) to 1e-10.
"""
import os
from dataclasses import replace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from oracle import core as ocore  # noqa: E402
from oracle import sobol as osob  # noqa: E402
from rpm_amd import sobol as psob  # noqa: E402
from rpm_amd import workloads  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "sobol_full_cell.npz")


def _cfg():
    return replace(psob.RunConfig(), lifetimes_us=(1.0, 3.0), sobol_scrambles=4)


def test_small_cell_matches_oracle():
    print("case: small system cell, Q=32, 4 scrambles", flush=True)
    sys_ = workloads.small_system(0)
    cfg = _cfg()
    task = psob.CellTask(0.0, 2, 32, True, cfg)
    state = psob.cell_initial_state(cfg, 2, sys_.dimensions)
    got = psob.run_cell_full(task, hamiltonian=sys_.hamiltonian(), initial_state=state)
    yo, bo, res = osob.run_cell(cfg, sys_.hamiltonian(builder=ocore.build_hamiltonian), state, 32, 2,
                                trace_times=psob.diagnostic_time_grid_us())
    assert got.yields.shape == (4, 2)
    assert np.abs(got.yields - yo).max() < 1e-10
    # the bounds come from the step error estimates, differences of two nearby Krylov approximations:
    # their relative precision is far below that of the populations
    np.testing.assert_allclose(got.propagation_bounds, bo, rtol=1e-4, atol=1e-16)
    st = res.stats
    assert (got.accepted_steps, got.rejected_steps, got.matvecs) == (st.accepted_steps, st.rejected_steps,
                                                                      st.matvecs)
    tidx = np.searchsorted(res.times_us, got.trace_times_us)
    assert np.abs(got.trace_population - res.observable_values[tidx]).max() < 1e-10


def test_batched_cells_equal_single_cells():
    print("case: 4 batched cells", flush=True)
    sys_ = workloads.small_system(0)
    cfg = _cfg()
    tasks = [psob.CellTask(t, s, 16, False, cfg) for t, s in ((0.0, 0), (35.0, 1), (90.0, 2), (140.0, 3))]
    batched = psob.run_cells(tasks, hamiltonian_for=lambda t: sys_.hamiltonian(theta_deg=t.theta_deg))
    for task, got in zip(tasks, batched):
        one = psob.run_cell_full(task, hamiltonian=sys_.hamiltonian(theta_deg=task.theta_deg),
                                 initial_state=psob.cell_initial_state(cfg, task.sample_id, sys_.dimensions))
        assert (got.theta_deg, got.sample_id) == (task.theta_deg, task.sample_id)
        assert np.abs(got.yields - one.yields).max() < 1e-12
        assert (got.accepted_steps, got.rejected_steps, got.matvecs) == (one.accepted_steps, one.rejected_steps,
                                                                          one.matvecs)


def test_full_system_cell_matches_recorded_oracle():
    print("case: full-system cell theta=90 sample 0", flush=True)
    g = np.load(GOLDEN, allow_pickle=False)
    cfg = replace(psob.RunConfig(), lifetimes_us=tuple(g["lifetimes_us"]), sobol_scrambles=int(g["sobol_scrambles"]))
    q = int(g["q_points"])
    np.testing.assert_array_equal(psob.reaction_time_schedule(cfg, q, int(g["sample_id"])), g["schedule"])
    task = psob.CellTask(float(g["theta_deg"]), int(g["sample_id"]), q, False, cfg)
    got = psob.run_cell(task)
    print("yields", got.yields.ravel(), "oracle", g["yields"].ravel(), flush=True)
    assert np.abs(got.yields - g["yields"]).max() < 1e-10
    assert (got.accepted_steps, got.rejected_steps, got.matvecs) == (int(g["accepted_steps"]),
                                                                      int(g["rejected_steps"]), int(g["matvecs"]))
    # the batched path on the same cell
    again = psob.run_cells([task])[0]
    assert np.abs(again.yields - g["yields"]).max() < 1e-10
