"""GPU: edge cases and error behaviour of the device propagators (the reference's rules:
R/rpm_statevector_engine.py:717-760 and R/tests/test_statevector_baseline.py:196-205).

- a time grid holding only t = 0 returns the initial population and takes no step;
- decreasing or negative times, the zero vector and a wrong dimension raise ValueError;
- the batched propagator applies the same checks to every trajectory, accepts ragged per-trajectory
  time grids, and rejects a list whose length is not the batch size;
- a ragged batched run equals the separate single runs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from rpm_amd import batch, core, engine, workloads  # noqa: E402


def _small():
    sys_ = workloads.small_system(0)
    h = sys_.hamiltonian()
    s0 = core.random_phase_initial_state(h.dimensions, 0, 0)
    s1 = core.random_phase_initial_state(h.dimensions, 0, 1)
    return h, s0, s1


def _prop(h):
    return engine.AdaptiveLanczosPropagator(h, minimum_dimension=8, maximum_dimension=24, dimension_increment=8)


def test_single_propagator_edges():
    print("case: single propagator edges", flush=True)
    h, s0, _ = _small()
    prop = _prop(h)
    obs = engine.SingletPopulationObservable(h.dimensions)
    res = prop.propagate(s0, [0.0], obs)
    assert res.stats.accepted_steps == 0 and res.stats.matvecs == 0
    assert abs(res.observable_values[0] - engine.singlet_population(s0, h.dimensions)) < 1e-15
    for bad in ([0.2, 0.1], [-0.1, 0.2], [0.1, 0.1], []):
        with pytest.raises(ValueError):
            prop.propagate(s0, bad, obs)
    with pytest.raises(ValueError):
        prop.propagate(np.zeros_like(s0), [0.1], obs)
    with pytest.raises(ValueError):
        prop.propagate(s0[:16], [0.1], obs)


def test_batched_propagator_edges_and_ragged_times():
    print("case: batched propagator edges", flush=True)
    h, s0, s1 = _small()
    bp = batch.BatchedLanczosPropagator.like(_prop(h))
    states = np.stack([s0, s1])
    with pytest.raises(ValueError):
        bp.propagate(states, [[0.0, 0.1], [0.0, 0.05, 0.03]])
    with pytest.raises(ValueError):
        bp.propagate(states, [[0.0, 0.1]])
    with pytest.raises(ValueError):
        bp.propagate(states[:, :16], [0.0, 0.1])
    t0 = np.array([0.0, 0.013, 0.05])
    t1 = np.array([0.0, 0.02, 0.04, 0.07, 0.09])
    got = bp.propagate(states, [t0, t1])
    obs = engine.SingletPopulationObservable(h.dimensions)
    for g, s, t in zip(got, (s0, s1), (t0, t1)):
        one = _prop(h).propagate(s, t, obs)
        np.testing.assert_array_equal(g.times_us, t)
        assert np.abs(g.observable_values - one.observable_values).max() < 1e-12
        assert (g.stats.accepted_steps, g.stats.matvecs) == (one.stats.accepted_steps, one.stats.matvecs)
    only0 = bp.propagate(states, [0.0])
    assert all(r.stats.matvecs == 0 for r in only0)
