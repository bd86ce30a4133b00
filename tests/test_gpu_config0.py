"""GPU: the CPU-runnable small-system workload end to end (2 electrons + 2 14N + 4 1H, dim 576).

B = 50 uT at theta = 0, J = 0, point-dipole D at 1.5 nm, k = 1/us, horizon 5/k, 0.00125 us grid
(4,001 points), reference Krylov rules, 16 random-phase vectors (seed 0).  All 16 trajectories run
as one batch.  Bars: yields within 1e-9 of the oracle's per-trajectory run (identical schedules) for
two of them; all 16 within their own propagation estimate of the exact (dense eigendecomposition)
Simpson yield.
"""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from oracle import core as ocore  # noqa: E402
from oracle import dense as odense  # noqa: E402
from rpm_amd import batch, core, workloads  # noqa: E402


def test_small_system_sixteen_vectors_batched():
    sys_ = workloads.small_system(0)
    h = sys_.hamiltonian()
    o = sys_.hamiltonian(builder=ocore.build_hamiltonian)
    cfg = sys_.sweep_config()
    grid = cfg.time_grid_us
    assert grid.size == 4001 and abs(grid[-1] - 5.0) < 1e-12
    states = np.stack([core.random_phase_initial_state(h.dimensions, 0, i) for i in range(16)])
    bp = batch.BatchedLanczosPropagator.like(core.make_propagator(h, cfg))
    res = bp.propagate(states, grid)
    yields = np.array([core.integrate_yield(grid, r.observable_values, cfg.lifetime_us) for r in res])
    w = core.yield_weights(grid, cfg.lifetime_us)
    estimates = np.array([float(np.dot(w, r.observable_error_bounds)) for r in res])
    print("config-0 mean yield", yields.mean(), "per-vector spread", yields.std(), "rounds", bp.rounds, flush=True)
    H = o.explicit_matrix()
    for i in range(16):
        exact, _ = odense.dense_grid_yield([(grid[-1], H)], states[i], grid, cfg.lifetime_us)
        assert abs(yields[i] - exact) <= estimates[i] + 1e-9, (i, yields[i], exact, estimates[i])
    ocfg = ocore.SweepConfig(lifetime_us=cfg.lifetime_us, survival_tail_bound=cfg.survival_tail_bound)
    for i in (0, 11):
        ref = ocore.run_trajectory(o, states[i], ocfg)
        assert [(x[1], x[2]) for x in res[i].schedule] == [(x[1], x[2]) for x in ref["result"].schedule]
        assert abs(yields[i] - ref["yield"]) < 1e-9
