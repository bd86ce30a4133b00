"""Oracle Lanczos propagator, initial state and yield quadrature.

Restates R/tests/test_statevector_baseline.py:164-295,371-381,
R/tests/test_poster_reproduction.py:165-238 and R/tests/test_jd_sweeps.py:248-288,376-386,
and pins the oracle against closed forms and the archived initial-state digests.
"""

import csv
import math
import os

import numpy as np
import pytest
import scipy.integrate
import scipy.linalg

from oracle import core, dense, engine, model

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _fixture():
    # R/tests/test_statevector_baseline.py:196-205
    sx, sy, sz = model.spin_operators(0.5)
    dims = (2, 2, 2)
    terms = (engine.LocalTerm((0,), 0.7 * sx + 0.2 * sz),
             engine.LocalTerm((1, 2), 0.4 * np.kron(sy, sx) + 0.3 * np.kron(sz, sz)))
    h = engine.MatrixFreeHamiltonian(dims, terms)
    state = engine.singlet_initial_state(dims, np.random.default_rng(6))
    return h, state, h.explicit_matrix()


def test_lanczos_matches_dense_exponential():
    h, state, d = _fixture()
    times = np.array([0.0, 0.03, 0.2, 0.7])
    prop = engine.AdaptiveLanczosPropagator(h, minimum_dimension=5, maximum_dimension=8, dimension_increment=1,
                                            embedded_offset=2, state_error_budget=1e-11,
                                            initial_step_us=0.4, maximum_step_us=0.4)
    res = prop.propagate(state, times, lambda v: engine.singlet_population(v, (2, 2, 2)), store_states=True)
    exp = [engine.singlet_population(scipy.linalg.expm(-1j * d * t) @ state, (2, 2, 2)) for t in times]
    np.testing.assert_allclose(res.observable_values, exp, atol=2e-12)
    np.testing.assert_allclose(res.final_state, scipy.linalg.expm(-1j * d * times[-1]) @ state, atol=2e-12)
    np.testing.assert_allclose(res.states[-1], res.final_state, atol=2e-12)
    assert res.stats.maximum_norm_error < 2e-12


def test_minimum_step_failure_is_reported():
    h, state, _ = _fixture()
    prop = engine.AdaptiveLanczosPropagator(h, minimum_dimension=3, maximum_dimension=3, dimension_increment=1,
                                            embedded_offset=1, state_error_budget=1e-30, initial_step_us=1.0,
                                            maximum_step_us=1.0, minimum_step_us=0.2)
    with pytest.raises(RuntimeError):
        prop.propagate(state, [1.0], lambda v: engine.singlet_population(v, (2, 2, 2)))


def test_rejected_step_can_recover():
    h, state, _ = _fixture()
    prop = engine.AdaptiveLanczosPropagator(h, minimum_dimension=3, maximum_dimension=3, dimension_increment=1,
                                            embedded_offset=1, state_error_budget=1e-2, initial_step_us=1.0,
                                            maximum_step_us=1.0, minimum_step_us=1e-6)
    res = prop.propagate(state, [1.0], lambda v: engine.singlet_population(v, (2, 2, 2)))
    assert res.stats.rejected_steps > 0 and res.stats.accepted_steps > 0
    assert [s[1] for s in res.schedule].count(False) == res.stats.rejected_steps


def test_maximum_basis_is_reused_after_growth():
    rng = np.random.default_rng(4)
    vals = rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16))
    h = engine.MatrixFreeHamiltonian((2, 2, 2, 2), (engine.LocalTerm((0, 1, 2, 3), 20.0 * (vals + vals.conj().T) / 2),))
    state = rng.normal(size=16) + 1j * rng.normal(size=16)
    state /= np.linalg.norm(state)
    res = engine.AdaptiveLanczosPropagator(h, minimum_dimension=4, maximum_dimension=12, dimension_increment=4,
                                           embedded_offset=2, state_error_budget=1e-4, initial_step_us=0.1,
                                           maximum_step_us=0.1, minimum_step_us=1e-8
                                           ).propagate(state, [0.2], lambda v: float(abs(v[0]) ** 2))
    attempts = res.stats.accepted_steps + res.stats.rejected_steps
    assert res.stats.maximum_krylov_dimension == 12
    assert res.stats.matvecs < attempts * (4 + 12)
    # matvec count equals the sum of all basis sizes recorded in the schedule
    assert res.stats.matvecs == sum(sum(s[3]) for s in res.schedule)


def test_invalid_arguments():
    h, state, _ = _fixture()
    with pytest.raises(ValueError):
        engine.AdaptiveLanczosPropagator(h, minimum_dimension=4, embedded_offset=4)
    with pytest.raises(ValueError):
        engine.AdaptiveLanczosPropagator(h, minimum_dimension=8, maximum_dimension=6)
    prop = engine.AdaptiveLanczosPropagator(h, minimum_dimension=5, maximum_dimension=8, embedded_offset=2)
    with pytest.raises(ValueError):
        prop.propagate(state, [0.2, 0.1], lambda v: 0.0)
    with pytest.raises(ValueError):
        prop.propagate(np.zeros(8), [0.1], lambda v: 0.0)
    with pytest.raises(ValueError):
        prop.propagate(state[:4], [0.1], lambda v: 0.0)


def _small_pair_system(seed=3):
    """Two radicals (eA with a spin-1 and a spin-1/2, eB with a spin-1/2), D != 0."""
    rng = np.random.default_rng(seed)
    nuclei = (model.Nucleus("N5", 1.0, _rand_sym(rng)), model.Nucleus("H6", 0.5, _rand_sym(rng)),
              model.Nucleus("N1", 1.0, _rand_sym(rng)))
    radicals = ("FAD", "FAD", "TrpH")
    q = {"FAD:N5": _traceless(rng), "TrpH:N1": _traceless(rng)}
    h = core.build_hamiltonian(37.0, 0.05, nuclei=nuclei, radicals=radicals, quadrupole_tensors=q)
    return h


def _rand_sym(rng):
    a = rng.normal(0.0, 0.3, size=(3, 3))
    return 0.5 * (a + a.T) + rng.uniform(0.0, 1.0) * np.eye(3)


def _traceless(rng):
    a = rng.normal(size=(3, 3))
    a = 0.5 * (a + a.T)
    return (a - np.trace(a) / 3 * np.eye(3)) * (1.0 / 28.0)


def test_compiled_pair_matches_generic_and_brute_force():
    h = _small_pair_system()
    generic = engine.MatrixFreeHamiltonian(h.dimensions, [engine.LocalTerm(t.axes, t.operator) for t in h.terms])
    brute = dense.dense_from_terms(h.dimensions, h.terms)
    np.testing.assert_allclose(h.explicit_matrix(), brute, atol=1e-12)
    np.testing.assert_allclose(generic.explicit_matrix(), brute, atol=1e-12)
    assert np.linalg.norm(brute - brute.conj().T) < 1e-11


def test_projected_observable_matches_direct_path_and_schedule():
    # R/tests/test_jd_sweeps.py:248-288: identical values, final state and step counts
    h = _small_pair_system(5)
    state = core.random_phase_initial_state(h.dimensions, 11, 2)
    times = np.linspace(0.0, 0.4, 81)
    kw = dict(minimum_dimension=8, maximum_dimension=16, dimension_increment=4, embedded_offset=2,
              state_error_budget=1e-7, initial_step_us=0.02, maximum_step_us=0.1)
    obs = engine.SingletPopulationObservable(h.dimensions)
    a = engine.AdaptiveLanczosPropagator(h, **kw).propagate(state, times, obs)
    b = engine.AdaptiveLanczosPropagator(h, **kw).propagate(state, times, lambda v: obs(v))
    np.testing.assert_allclose(a.observable_values, b.observable_values, atol=1e-12)
    np.testing.assert_allclose(a.final_state, b.final_state, atol=1e-13)
    assert a.stats.accepted_steps == b.stats.accepted_steps
    assert a.stats.rejected_steps == b.stats.rejected_steps
    # and against the exact eigenbasis propagation
    hd = h.explicit_matrix()
    exact = [engine.singlet_population(s, h.dimensions) for s in dense.eigen_propagate(hd, state, times)]
    np.testing.assert_allclose(a.observable_values, exact, atol=1e-6)


def test_yield_against_closed_form_eigenbasis_integral():
    h = _small_pair_system(7)
    state = core.random_phase_initial_state(h.dimensions, 0, 0)
    cfg = core.SweepConfig(lifetime_us=0.05, survival_tail_bound=math.exp(-4.0))
    out = core.run_trajectory(h, state, cfg, propagator=engine.AdaptiveLanczosPropagator(
        h, minimum_dimension=12, maximum_dimension=24, dimension_increment=4, embedded_offset=4,
        state_error_budget=1e-9, initial_step_us=0.01, maximum_step_us=0.05))
    grid = cfg.time_grid_us
    hd = h.explicit_matrix()
    exact = dense.dense_singlet_yield_closed_form(hd, state, cfg.lifetime_us, grid[-1])
    # propagation error alone: Simpson over the exact populations on the same grid
    pops = [engine.singlet_population(s, h.dimensions) for s in dense.eigen_propagate(hd, state, grid)]
    assert abs(out["yield"] - core.integrate_yield(grid, np.array(pops), cfg.lifetime_us)) < 1e-8
    # quadrature error alone: the fine-vs-coarse estimate bounds the distance to the closed form
    assert abs(out["yield"] - exact) < out["quadrature_error"] < 1e-5


def test_random_phase_state_and_archived_digests():
    # bitwise initial_state_sha256 fixtures from the archived full-system run
    with open(os.path.join(GOLDEN, "archived_full14_yields.csv")) as f:
        rows = list(csv.DictReader(f))
    digests = {int(r["sample_id"]): r["initial_state_sha256"] for r in rows}
    assert len(digests) == 4
    for sid, digest in digests.items():
        s = core.random_phase_initial_state(model.SYSTEM_DIMENSIONS, core.MASTER_SEED, sid)
        assert core.initial_state_sha256(s) == digest
    s = core.random_phase_initial_state(model.SYSTEM_DIMENSIONS, core.MASTER_SEED, 3)
    assert abs(np.linalg.norm(s) - 1.0) < 1e-12
    assert abs(engine.singlet_population(s, model.SYSTEM_DIMENSIONS) - 1.0) < 1e-12
    assert core.seed_for(core.MASTER_SEED, 7) != core.seed_for(core.MASTER_SEED, 8)


def test_random_phase_normalization_and_mixed_trace():
    # R/tests/test_statevector_baseline.py:164-177
    st = engine.random_phase_nuclear_state(4, np.random.default_rng(12))
    assert abs(np.vdot(st, st).real - 1.0) < 1e-14
    np.testing.assert_allclose(np.abs(st), 0.5, atol=1e-14)
    obs = np.zeros((4, 4)); obs[0, 1] = obs[1, 0] = 1.0
    rng = np.random.default_rng(22)
    est = [np.vdot(s, obs @ s).real for s in (engine.random_phase_nuclear_state(4, rng) for _ in range(4000))]
    assert abs(np.mean(est)) < 0.02


def test_simpson_exact_on_cubic():
    # R/tests/test_statevector_baseline.py:371-381
    t = np.linspace(0.0, 1.0, 5)
    assert abs(core.integrate_yield(t, np.exp(t) * t ** 3, 1.0) - 0.25) < 1e-14


def test_simpson_constant_population_and_four_level_closed_forms():
    # R/tests/test_poster_reproduction.py:217-238
    cfg = core.PRODUCTION_MODEL.sweep_config
    t = cfg.time_grid_us
    assert t.size == 46053 if cfg.survival_tail_bound == 1e-5 else True
    v = core.integrate_yield(t, np.ones_like(t), cfg.lifetime_us)
    assert abs(v - (1.0 - math.exp(-t[-1] / cfg.lifetime_us))) < 2e-11
    dec = scipy.integrate.simpson(np.exp(-t[::2] / cfg.lifetime_us) / cfg.lifetime_us, x=t[::2])
    assert abs(v - dec) < 2e-10
    he = core.electron_only_hamiltonian(37.0, 2.0, (np.array([1.2, -0.7, 0.3]), np.array([-0.4, 0.1, 0.9])))
    pops = core.dense_singlet_dynamics(he, t)
    assert abs(core.integrate_yield(t, pops, cfg.lifetime_us)
               - core.dense_truncated_singlet_yield(he, cfg.lifetime_us, t[-1])) < 2e-9


def test_time_grid_and_nested_coarse_grid():
    # production grid: tau=5, tail 1e-5 -> 46,053 points, last interval shorter
    cfg = core.SweepConfig(lifetime_us=5.0, survival_tail_bound=1e-5)
    t = cfg.time_grid_us
    assert t.size == 46053
    assert abs(t[-1] - cfg.maximum_time_us) < 1e-12
    idx = core.coarse_grid_indices(cfg)
    assert idx[0] == 0 and idx[-1] == t.size - 1
    np.testing.assert_allclose(t[idx[:-1]], np.arange(idx.size - 1) * 0.0025, atol=1e-12)
    assert core.SweepConfig(lifetime_us=1.0, survival_tail_bound=math.exp(-5.0)).time_grid_us.size == 4001


def test_dense_singlet_closed_form_matches_quadrature():
    h = _small_pair_system(9)
    state = core.random_phase_initial_state(h.dimensions, 3, 1)
    hd = h.explicit_matrix()
    t = np.linspace(0.0, 0.3, 6001)
    pops = [engine.singlet_population(s, h.dimensions) for s in dense.eigen_propagate(hd, state, t)]
    simpson = core.integrate_yield(t, np.array(pops), 0.1)
    assert abs(simpson - dense.dense_singlet_yield_closed_form(hd, state, 0.1, 0.3)) < 1e-8
