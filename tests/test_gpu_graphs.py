"""GPU: Lanczos basis step windows replayed as HIP graphs (use_graphs) give bitwise the eager results.

Each propagation captures every distinct basis window once and replays it many times, with the
propagator's own eager work (candidates, outputs, endpoint updates, host synchronisations) between
the replays; populations, schedules, device yields and final states must equal the eager run's
bitwise. Cases: the production operator (batched and single) and the 576-level random system with a
tight budget and small dimensions (rejections and continued bases).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from rpm_amd import batch, core, engine, quadrature, workloads  # noqa: E402


def _same(a, b):
    assert a.schedule == b.schedule
    assert a.stats == b.stats
    np.testing.assert_array_equal(a.observable_values, b.observable_values)
    np.testing.assert_array_equal(a.state_error_estimates, b.state_error_estimates)
    assert a.device_yields == b.device_yields


def test_batched_graph_replay_equals_eager():
    print("case: production operator, 4 trajectories, 0.5 us, graphs vs eager", flush=True)
    h = core.build_hamiltonian(90.0, 0.05)
    times = np.arange(401) * 0.00125
    states = np.stack([core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, i) for i in range(4)])
    pairs = batch.pair_matrices(h, [core.build_hamiltonian(t, 0.05) for t in (0.0, 0.0, 60.0, 60.0)])
    w = quadrature.yield_weight_table(times, 5.0)
    kw = dict(state_error_budget=1e-9)
    eager = batch.BatchedLanczosPropagator(h, **kw)
    eager.use_graphs = False
    ref = eager.propagate(states, times, pairs=pairs, yield_weights=w)
    ref_final = eager.final_states.clone()
    bp = batch.BatchedLanczosPropagator(h, **kw)
    assert bp.use_graphs          # the default
    got = bp.propagate(states, times, pairs=pairs, yield_weights=w)
    print("captures", bp.graphs.captures, "replays", bp.graphs.replays, flush=True)
    assert bp.graphs.captures >= 1 and bp.graphs.replays > bp.graphs.captures
    for r, g in zip(ref, got):
        _same(r, g)
    assert torch.equal(bp.final_states, ref_final)


def test_single_graph_replay_equals_eager_with_rejections():
    print("case: 576-level system, tight budget, graphs vs eager", flush=True)
    sys_ = workloads.small_system(0)
    h = sys_.hamiltonian()
    times = np.arange(401) * 0.00125
    state = core.random_phase_initial_state(h.dimensions, 0, 1)
    kw = dict(minimum_dimension=8, maximum_dimension=16, dimension_increment=4, embedded_offset=2,
              state_error_budget=1e-10)
    obs = engine.SingletPopulationObservable(h.dimensions)
    w = quadrature.yield_weight_table(times, 1.0)
    ref = engine.AdaptiveLanczosPropagator(h, **kw).propagate(state, times, obs, yield_weights=w)
    prop = engine.AdaptiveLanczosPropagator(h, **kw)
    prop.use_graphs = True
    got = prop.propagate(state, times, obs, yield_weights=w)
    assert ref.stats.rejected_steps > 0
    print("captures", prop._graphs.captures, "replays", prop._graphs.replays, flush=True)
    assert prop._graphs.replays > prop._graphs.captures
    _same(ref, got)
    np.testing.assert_array_equal(got.final_state, ref.final_state)
    # the same propagator, graphs off again, still gives the eager result
    prop.use_graphs = False
    _same(ref, prop.propagate(state, times, obs, yield_weights=w))


def test_single_graph_replay_production_operator():
    print("case: production operator, single trajectory, graphs vs eager", flush=True)
    h = core.build_hamiltonian(90.0, 0.05)
    times = np.arange(401) * 0.00125
    state = core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, 2)
    obs = engine.SingletPopulationObservable(h.dimensions)
    ref = engine.AdaptiveLanczosPropagator(h, state_error_budget=1e-9).propagate(state, times, obs)
    prop = engine.AdaptiveLanczosPropagator(h, state_error_budget=1e-9)
    prop.use_graphs = True
    got = prop.propagate(state, times, obs)
    assert prop._graphs.replays > prop._graphs.captures
    _same(ref, got)
