"""The Sobol reaction-time estimator: oracle pinned (schedule, uncertainty, convergence to the exact
yield) and the host parts of rpm_amd.sobol (schedule, uncertainty, checkpoint) against it."""
import math
from dataclasses import replace

import numpy as np
import pytest

from oracle import dense, engine as oeng, model as om
from oracle import sobol as osob
from rpm_amd import sobol as psob


def test_schedule_nested_and_independent():
    """R/tests/test_statevector_baseline.py:179-185, for the oracle and the product."""
    cfg = psob.RunConfig()
    for mod in (osob, psob):
        small = mod.reaction_time_schedule(cfg, 32, sample_id=3)
        large = mod.reaction_time_schedule(cfg, 64, sample_id=3)
        np.testing.assert_array_equal(small, large[:, :, :32])
        assert not np.array_equal(small, mod.reaction_time_schedule(cfg, 32, sample_id=4))
    np.testing.assert_array_equal(psob.reaction_time_schedule(cfg, 64, 7), osob.reaction_time_schedule(cfg, 64, 7))
    with pytest.raises(ValueError):
        psob.reaction_time_schedule(cfg, 48)
    # one shared scramble set when samples are not independent
    shared = replace(cfg, independent_sobol_by_nuclear_sample=False)
    np.testing.assert_array_equal(psob.reaction_time_schedule(shared, 16, 1), psob.reaction_time_schedule(shared, 16, 9))


def test_schedule_is_a_stratified_exponential_sample():
    """Each scramble puts exactly one u in every [k/Q, (k+1)/Q) (the (0, m, 1)-net property), and
    T = -tau log(1 - u) inverts the exponential CDF."""
    cfg = psob.RunConfig()
    q = 256
    sched = psob.reaction_time_schedule(cfg, q, sample_id=2)
    for s in range(cfg.sobol_scrambles):
        u = osob.sobol_uniforms(cfg, q, 2, s)
        assert sorted(np.floor(u * q).astype(int)) == list(range(q))
        for i, tau in enumerate(cfg.lifetimes_us):
            np.testing.assert_allclose(np.exp(-sched[i, s] / tau), 1.0 - u, rtol=1e-13, atol=0)
    # the stratified sample's mean is close to tau (exact for the continuous distribution)
    assert abs(sched[2].mean() / cfg.lifetimes_us[2] - 1.0) < 0.02


def test_uncertainty_known_answer():
    """values = arange(32)/100 as 4 x 8 (R/tests/test_statevector_baseline.py:187-191), with the
    variance components worked out by hand and Student-t quantiles from tables."""
    values = np.arange(32, dtype=float).reshape(4, 8) / 100.0
    u = psob.uncertainty_components(values, 200, 99)
    assert u == psob.uncertainty_components(values, 200, 99)
    # row means 0.035, 0.115, 0.195, 0.275: sample variance 0.08^2 * 5/3; over n = 4
    var_nuc = 0.0064 * (5.0 / 3.0) / 4.0
    # each row 0..7 / 100: sample variance 6e-4; over n s = 32
    var_time = 6e-4 / 32.0
    t3, t28 = 3.182446305284263, 2.048407141795244       # t_{0.975} at 3 and 28 degrees of freedom
    assert abs(u.mean - 0.155) < 1e-15
    assert abs(u.nuclear_halfwidth - t3 * math.sqrt(var_nuc)) < 1e-12
    assert abs(u.reaction_time_halfwidth - t28 * math.sqrt(var_time)) < 1e-12
    assert abs(u.joint_halfwidth - t3 * math.sqrt(var_nuc + var_time)) < 1e-12
    assert abs(u.nuclear_sample_std - 0.08 * math.sqrt(5.0 / 3.0)) < 1e-15
    assert np.allclose(osob.uncertainty_components(values), (u.mean, u.nuclear_halfwidth, u.reaction_time_halfwidth,
                                                             u.joint_halfwidth, u.nuclear_sample_std), rtol=0,
                       atol=1e-15)
    with pytest.raises(ValueError):
        psob.uncertainty_components(values[:1], 1, 1)


def test_uncertainty_interval_coverage():
    """Balanced random effects y_ij = mu + a_i + e_ij. When the between-sample effect dominates, the
    joint 95% interval covers mu in about 95% of data sets. When the within-row noise dominates, the
    interval is conservative: that term is also present in the row means, and is counted again."""
    for sa, se, lo, hi in ((0.05, 0.005, 0.93, 0.97), (0.005, 0.05, 0.95, 1.0)):
        rng = np.random.default_rng(5)
        mu, n, s, trials, hits = 0.3, 6, 8, 3000, 0
        for _ in range(trials):
            a = rng.normal(0.0, sa, size=(n, 1))
            e = rng.normal(0.0, se, size=(n, s))
            u = psob.uncertainty_components(mu + a + e, 0, 0)
            hits += abs(u.mean - mu) <= u.joint_halfwidth
        assert lo <= hits / trials <= hi, (sa, se, hits / trials)


def _tiny_pair():
    """2 electrons + one 14N on electron A: dims (2, 2, 3), 12 levels."""
    sx, sy, sz = om.spin_operators(0.5)
    zeeman = om.TWOPI * om.GAMMA_E_MHZ_PER_MT * (0.05 * 0.6 * sx + 0.05 * 0.8 * sz)
    a = np.array([[0.3, 0.1, 0.0], [0.1, -0.2, 0.05], [0.0, 0.05, 1.4]])
    terms = [oeng.LocalTerm((0,), zeeman), oeng.LocalTerm((1,), zeeman),
             oeng.LocalTerm((0, 2), om.TWOPI * om.GAMMA_E_MHZ_PER_MT * oeng.two_spin_coupling_operator(0.5, 1.0, a))]
    dims = (2, 2, 3)
    return oeng.MatrixFreeHamiltonian(dims, terms), dense.dense_from_terms(dims, terms)


def test_oracle_cell_converges_to_exact_yield():
    """Sobol mean over Q = 256 points x 8 scrambles vs the closed-form Haberkorn yield."""
    h, hd = _tiny_pair()
    cfg = replace(psob.RunConfig(), lifetimes_us=(0.5, 2.0), lanczos_minimum_dimension=6,
                  lanczos_maximum_dimension=12, lanczos_dimension_increment=2, lanczos_embedded_offset=2,
                  lanczos_state_error_budget=1e-9)
    rng = np.random.default_rng(osob.seed(cfg.seed, 0xA11CE, 1))
    state = oeng.singlet_initial_state(h.dimensions, rng)
    yields, bounds, res = osob.run_cell(cfg, h, state, 256, 1)
    assert yields.shape == (8, 2) and bounds.shape == (8, 2)
    for i, tau in enumerate(cfg.lifetimes_us):
        exact = dense.dense_singlet_yield_closed_form(hd, state, tau, 60.0 * tau)
        est = yields[:, i]
        half = 4.0 * est.std(ddof=1) / math.sqrt(est.size) + 1e-9
        assert abs(est.mean() - exact) < half, (tau, est.mean(), exact, half)
        assert abs(est.mean() - exact) < 2e-3


def test_checkpoint_round_trip(tmp_path):
    cfg = psob.RunConfig(lifetimes_us=(1.0, 5.0))
    r = psob.CellResult(theta_deg=80.0, sample_id=3, q_points=32, yields=np.arange(16.0).reshape(8, 2) / 7,
                        propagation_bounds=np.full((8, 2), 1e-7), trace_times_us=np.empty(0),
                        trace_population=np.empty(0), trace_bounds=np.empty(0), accepted_steps=11,
                        rejected_steps=2, matvecs=400, maximum_krylov_dimension=40, minimum_step_us=0.05,
                        maximum_step_us=0.25, accumulated_state_error=1e-6, maximum_norm_error=1e-15,
                        elapsed_seconds=1.5)
    path = psob.save_cell(r, tmp_path, cfg.configuration_hash)
    assert path.name == "theta_080p000_sample_0003_q_0032.npz"
    back = psob.load_cell(tmp_path, 80.0, 3, 32, cfg.configuration_hash)
    for k, v in r.__dict__.items():
        np.testing.assert_array_equal(getattr(back, k), v)
    with pytest.raises(RuntimeError):
        psob.load_cell(tmp_path, 80.0, 3, 32, psob.RunConfig().configuration_hash)


def test_configuration_record():
    cfg = psob.RunConfig()
    d = cfg.stable_dict()
    assert d["lifetimes_us"] == [1.0, 2.0, 5.0, 10.0, 20.0, 100.0] and len(d["dimensions"]) == 16
    assert "nominal_quadrupole_tensors_mT" in d and "quadrupole_scale" in d
    q0 = replace(cfg, quadrupole_scale=0.0).stable_dict()
    assert "quadrupole_scale" not in q0 and "nominal_quadrupole_tensors_mT" not in q0
    assert len(cfg.configuration_hash) == 64 and cfg.configuration_hash != replace(cfg, seed=1).configuration_hash
