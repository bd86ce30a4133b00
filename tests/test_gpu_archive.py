"""End-to-end: GPU trajectories against the reference's archived full-system yields.

results/quadrupole/full14_uncoupled_tau5_29angles/raw_sample_yields.csv of the
reference (fixture copy: tests/golden/archived_full14_yields.csv) holds 464
per-trajectory singlet yields at dim 221,184 (50 uT, tau = 5 us, tail bound
1e-5, horizon 57.5646 us, 46,053-point grid, production Krylov rules, budget
2.4e-5, seed 20260731).  Each archived row took 3.4-7.8 CPU hours.  Two rows
are re-run here from scratch on the GPU, one trajectory at a time; the yield
must agree within the archive's own propagation estimate (~1e-6), and the
fine/coarse quadrature difference within 1e-9.  Four more rows run as one
batched propagation and must agree with the archive to 1e-10.
"""

import csv
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from rpm_amd import core  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "archived_full14_yields.csv")


def _rows():
    with open(GOLDEN) as f:
        return list(csv.DictReader(f))


def _pick(rows, interactions, scale, theta, sample):
    for r in rows:
        if (r["electron_interactions"] == interactions and float(r["scale"]) == scale
                and float(r["theta_deg"]) == theta and int(r["sample_id"]) == sample):
            return r
    raise KeyError((interactions, scale, theta, sample))


@pytest.mark.parametrize("interactions,scale,theta,sample", [
    ("D=J=0", 0.0, 0.0, 0),
    ("D/J=-3", 1.0, 90.0, 1),
])
def test_archived_full_system_yield(interactions, scale, theta, sample):
    row = _pick(_rows(), interactions, scale, theta, sample)
    model = core.ProductionModel(quadrupole_scale=scale,
                                 ratio_D_over_J=None if interactions == "D=J=0" else -3.0)
    h = core.build_hamiltonian(theta, 0.05, model=model)
    state = core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, sample)
    assert core.initial_state_sha256(state) == row["initial_state_sha256"]
    cfg = core.SweepConfig(seed=core.MASTER_SEED, field_microT=50.0, lifetime_us=5.0, survival_tail_bound=1e-5,
                           state_error_budget=2.4e-5, quadrupole_scale=scale)
    out = core.propagate_trajectory(h, state, cfg)
    ref = float(row["yield"])
    print(f"{interactions} Q={scale} theta={theta} s={sample}: GPU {out['yield']:.12f} archive {ref:.12f} "
          f"diff {out['yield'] - ref:.2e}; quad err {out['quadrature_error']:.3e} vs {float(row['quadrature_error']):.3e}; "
          f"prop est {out['yield_propagation_estimate']:.3e} vs {float(row['yield_propagation_estimate']):.3e}; "
          f"steps {out['result'].stats.accepted_steps}/{out['result'].stats.rejected_steps} "
          f"matvecs {out['result'].stats.matvecs}", flush=True)
    assert abs(out["yield"] - ref) < max(2e-6, float(row["yield_propagation_estimate"]))
    assert abs(out["quadrature_error"] - float(row["quadrature_error"])) < 1e-9
    assert out["yield_propagation_estimate"] < 5e-6


def test_archived_yields_batched():
    """Four archived cells (D = J = 0, quadrupoles on, four field angles, sample 2) in one batched
    propagation over the whole horizon: every yield equals the archived one to 1e-10."""
    from rpm_amd import batch
    rows = _rows()
    thetas = (0.0, 60.0, 120.0, 180.0)
    sample = 2
    print(f"case: batched archived cells theta={thetas} sample={sample}", flush=True)
    model = core.ProductionModel(quadrupole_scale=1.0, ratio_D_over_J=None)
    cfg = core.SweepConfig(seed=core.MASTER_SEED, field_microT=50.0, lifetime_us=5.0, survival_tail_bound=1e-5,
                           state_error_budget=2.4e-5, quadrupole_scale=1.0)
    base = core.build_hamiltonian(90.0, 0.05, model=model)
    pairs = batch.pair_matrices(base, [core.build_hamiltonian(t, 0.05, model=model) for t in thetas])
    state = core.random_phase_initial_state(base.dimensions, core.MASTER_SEED, sample)
    bp = batch.BatchedLanczosPropagator.like(core.make_propagator(base, cfg))
    times = cfg.time_grid_us
    res = bp.propagate(np.stack([state] * len(thetas)), times, pairs=pairs)
    for t, r in zip(thetas, res):
        ref = float(_pick(rows, "D=J=0", 1.0, t, sample)["yield"])
        got = core.integrate_yield(times, r.observable_values, cfg.lifetime_us)
        print(f"theta={t}: batched {got:.15f} archive {ref:.15f} diff {got - ref:.2e}", flush=True)
        assert abs(got - ref) < 1e-10


def test_archived_coupled_yields_batched():
    """Four archived cells of the coupled run (D/J = -3, quadrupoles off, four field angles, samples
    0-3) in one batched propagation over the whole horizon, the device-summed yields against the
    archive to 1e-10."""
    from rpm_amd import batch
    rows = _rows()
    cells = ((10.0, 0), (85.0, 1), (91.0, 2), (150.0, 3))
    print(f"case: batched archived coupled cells {cells}", flush=True)
    model = core.ProductionModel(quadrupole_scale=0.0, ratio_D_over_J=-3.0)
    cfg = core.SweepConfig(seed=core.MASTER_SEED, field_microT=50.0, lifetime_us=5.0, survival_tail_bound=1e-5,
                           state_error_budget=2.4e-5, quadrupole_scale=0.0)
    base = core.build_hamiltonian(90.0, 0.05, model=model)
    pairs = batch.pair_matrices(base, [core.build_hamiltonian(t, 0.05, model=model) for t, _ in cells])
    states = np.stack([core.random_phase_initial_state(base.dimensions, core.MASTER_SEED, s) for _, s in cells])
    for (t, s), st in zip(cells, states):
        assert core.initial_state_sha256(st) == _pick(rows, "D/J=-3", 0.0, t, s)["initial_state_sha256"]
    bp = batch.BatchedLanczosPropagator.like(core.make_propagator(base, cfg))
    times = cfg.time_grid_us
    weights = core.yield_weight_table(times, cfg.lifetime_us, core.coarse_grid_indices(cfg))
    res = bp.propagate(states, times, pairs=pairs, yield_weights=weights)
    for (t, s), r in zip(cells, res):
        ref = float(_pick(rows, "D/J=-3", 0.0, t, s)["yield"])
        got = float(r.device_yields[0])
        print(f"theta={t} sample={s}: batched {got:.15f} archive {ref:.15f} diff {got - ref:.2e}", flush=True)
        assert abs(got - ref) < 1e-10
