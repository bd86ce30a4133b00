"""GPU: the directional-map and pulse-optimisation drivers.

- ``batch.directional_yields`` over states x Fibonacci directions x field strengths equals the plain
  one-trajectory propagation of every combination (Simpson yield) to 1e-12;
- ``adjoint.pulse_ascent`` raises the ensemble-mean yield at every step for a small step size
  (first-order ascent), and reports the gradients it used.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from rpm_amd import adjoint, batch, core, engine, quadrature, workloads  # noqa: E402


def test_directional_yields_match_plain_runs():
    print("case: directional yields 2 states x 3 directions x 2 fields", flush=True)
    sys_ = workloads.small_system(0)
    base = sys_.hamiltonian()
    build = lambda t, f, p: sys_.hamiltonian(theta_deg=t, field_mT=f, phi_deg=p)  # noqa: E731
    states = np.stack([core.random_phase_initial_state(base.dimensions, 0, i) for i in range(2)])
    th, ph = workloads.fibonacci_sphere(3)
    fields = [25.0, 50.0]
    times = np.arange(161) * 0.00125
    tau = 1.0
    bp = batch.BatchedLanczosPropagator(base, minimum_dimension=8, maximum_dimension=24, dimension_increment=8)
    got = batch.directional_yields(base, build, bp, states, th, ph, fields, times, tau, batch_size=5)
    assert got.shape == (2, 3, 2)
    obs = engine.SingletPopulationObservable(base.dimensions)
    for s in range(2):
        for d in range(3):
            for f in range(2):
                h = build(th[d], 1e-3 * fields[f], ph[d])
                prop = engine.AdaptiveLanczosPropagator(h, minimum_dimension=8, maximum_dimension=24,
                                                        dimension_increment=8)
                r = prop.propagate(states[s], times, obs)
                want = quadrature.integrate_yield(r.times_us, r.observable_values, tau)
                assert abs(got[s, d, f] - want) < 1e-12, (s, d, f)


def test_pulse_ascent_raises_mean_yield():
    print("case: pulse ascent 3 steps", flush=True)
    sys_ = workloads.small_system(0)
    base = sys_.hamiltonian()
    build = lambda t, f, p: sys_.hamiltonian(theta_deg=t, field_mT=f, phi_deg=p)  # noqa: E731
    states = np.stack([core.random_phase_initial_state(base.dimensions, 0, i) for i in range(3)])
    times = np.arange(241) * 0.00125
    bp = batch.BatchedLanczosPropagator(base, minimum_dimension=24, maximum_dimension=64, state_error_budget=1e-11)
    prob = adjoint.FieldYield(base, build, bp, states, times, 1.0, pulse_window_us=(0.0, 0.2))
    a0 = np.random.default_rng(3).uniform(-20.0, 20.0, size=8)
    out = adjoint.pulse_ascent(prob, 40.0, 0.0, 50.0, a0, steps=3, learning_rate=1.0e5, bound_uT=40.0)
    print("mean yields", out.mean_yields, "max |grad|", np.abs(out.gradients).max(), flush=True)
    assert out.gradients.shape == (3, 8) and out.amplitudes_uT.shape == (8,)
    assert np.all(np.diff(out.mean_yields) > 0.0)
    assert np.all(np.abs(out.amplitudes_uT) <= 40.0)
