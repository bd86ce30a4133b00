"""GPU: the singlet yield, coarse yield and propagation estimate summed on the device
(rpm_yield_accumulate) against the same integrals of the returned populations on the host and
against the oracle (R/rpm_quantum_needle_jd_sweeps.py:418-420, R/rpm_full14_quadrupole_experiment.py:179-183).

Bars: device against host quadrature of the same populations 1e-14; against the oracle's yields
1e-9; batched against single trajectories bitwise (the sum order depends only on a trajectory's
own outputs).
"""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from oracle import core as ocore  # noqa: E402
from rpm_amd import _native as nat  # noqa: E402
from rpm_amd import batch, core, engine, quadrature, segments, workloads  # noqa: E402


def _host(times, tau, pops, bounds, coarse=None):
    table = quadrature.yield_weight_table(times, tau, coarse)
    return np.dot(table[:, 0], pops), np.dot(table[:, 1], pops), np.dot(table[:, 0], bounds)


def test_accumulate_kernel_against_numpy():
    print("case: rpm_yield_accumulate on synthetic outputs", flush=True)
    rng = np.random.default_rng(0)
    dev = nat.device()
    nv, nw, B = 700, 900, 5
    vals = rng.uniform(size=nv)
    perr = rng.uniform(0, 1e-3, size=nv)
    weights = rng.normal(size=(nw, 2))
    acc = rng.uniform(0, 1e-3, size=B)
    spans = np.array([[0, 300, 10], [300, 0, 0], [300, 400, 500], [650, 100, 0], [-1, 5, 0]], dtype=np.int64)
    phi0 = rng.normal(size=(B, 3))
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in
         dict(v=vals, e=perr, w=weights, a=acc, s=spans, p=phi0).items()}
    vp = ctypes.c_void_p
    nat.check(nat.load().rpm_yield_accumulate(vp(nat.ptr(d["v"])), vp(nat.ptr(d["e"])), nv, vp(nat.ptr(d["s"])), B,
                                              vp(nat.ptr(d["w"])), nw, vp(nat.ptr(d["a"])), vp(nat.ptr(d["p"])),
                                              vp(nat.stream_ptr())))
    got = d["p"].cpu().numpy()
    want = phi0.copy()
    for b, (r0, cnt, w0) in enumerate(spans):
        if r0 < 0:
            continue
        cnt = min(cnt, nv - r0, nw - w0)       # rows past the values contribute nothing
        if cnt <= 0:
            continue
        p = vals[r0:r0 + cnt]
        e = acc[b] + perr[r0:r0 + cnt]
        bound = np.minimum(1.0, 2 * e + e * e)
        w = weights[w0:w0 + cnt]
        want[b] += [np.dot(w[:, 0], p), np.dot(w[:, 1], p), np.dot(w[:, 0], bound)]
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)
    np.testing.assert_array_equal(got[4], phi0[4])      # a negative row leaves the trajectory alone


def test_single_trajectory_device_yields_match_host_and_oracle():
    print("case: 576-level system, 5 us, single propagator", flush=True)
    sys_ = workloads.small_system(0)
    h = sys_.hamiltonian()
    o = sys_.hamiltonian(builder=ocore.build_hamiltonian)
    cfg = sys_.sweep_config()
    ocfg = ocore.SweepConfig(lifetime_us=cfg.lifetime_us, survival_tail_bound=cfg.survival_tail_bound)
    state = core.random_phase_initial_state(h.dimensions, 0, 2)
    out = core.propagate_trajectory(h, state, cfg)
    times = out["times_us"]
    fine, coarse, est = _host(times, cfg.lifetime_us, out["singlet_population"],
                              out["population_error_estimate"], core.coarse_grid_indices(cfg))
    assert abs(out["yield"] - fine) < 1e-14
    assert abs(out["coarse_yield"] - coarse) < 1e-14
    assert abs(out["yield_propagation_estimate"] - est) <= 1e-14 + 1e-12 * est
    ref = ocore.run_trajectory(o, state, ocfg)
    assert abs(out["yield"] - ref["yield"]) < 1e-9
    assert abs(out["coarse_yield"] - ref["coarse_yield"]) < 1e-9
    np.testing.assert_allclose(out["yield_propagation_estimate"], ref["yield_propagation_estimate"], rtol=1e-4)


def test_batched_device_yields_equal_single_and_ragged_grids():
    print("case: batched device yields, shared and ragged grids", flush=True)
    sys_ = workloads.small_system(0)
    h = sys_.hamiltonian()
    kw = dict(minimum_dimension=16, maximum_dimension=32, dimension_increment=8, state_error_budget=1e-9)
    states = np.stack([core.random_phase_initial_state(h.dimensions, 0, i) for i in range(3)])
    t_a = np.arange(401) * 0.00125
    t_b = np.arange(250) * 0.00125            # an even point count: Simpson's last-interval correction
    tau = 1.0
    w_a = quadrature.yield_weight_table(t_a, tau, np.arange(0, 401, 2))
    w_b = quadrature.yield_weight_table(t_b, tau)
    bp = batch.BatchedLanczosPropagator(h, **kw)
    res = bp.propagate(states, [t_a, t_b, t_a], yield_weights=[w_a, w_b, w_a])
    obs = engine.SingletPopulationObservable(h.dimensions)
    for b, (t, w) in enumerate([(t_a, w_a), (t_b, w_b), (t_a, w_a)]):
        r = res[b]
        host = (np.dot(w[:, 0], r.observable_values), np.dot(w[:, 1], r.observable_values),
                np.dot(w[:, 0], r.observable_error_bounds))
        np.testing.assert_allclose(r.device_yields, host, rtol=1e-13, atol=1e-15)
        one = engine.AdaptiveLanczosPropagator(h, **kw).propagate(states[b], t, obs, yield_weights=w)
        assert one.device_yields == r.device_yields           # bitwise
        assert abs(r.device_yields[0] - quadrature.integrate_yield(t, r.observable_values, tau)) < 1e-14


def test_full_system_device_yield_short_horizon():
    print("case: production operator, 0.1 us, batched device yield", flush=True)
    base = core.build_hamiltonian(90.0, 0.05)
    times = np.arange(81) * 0.00125
    states = np.stack([core.random_phase_initial_state(base.dimensions, core.MASTER_SEED, i) for i in range(2)])
    bp = batch.BatchedLanczosPropagator(base, state_error_budget=1e-9)
    w = quadrature.yield_weight_table(times, 5.0, np.arange(0, 81, 2))
    res = bp.propagate(states, times, yield_weights=w)
    for r in res:
        assert abs(r.device_yields[0] - quadrature.integrate_yield(times, r.observable_values, 5.0)) < 1e-14
        assert abs(r.device_yields[1] - np.dot(w[:, 1], r.observable_values)) < 1e-14


def test_segmented_device_yields_match_host():
    print("case: segment loop with a pulse, device yields", flush=True)
    sys_ = workloads.small_system(0)
    base = sys_.hamiltonian()
    times = np.arange(241) * 0.00125
    states = np.stack([core.random_phase_initial_state(base.dimensions, 0, i) for i in range(2)])
    pairs = batch.pair_matrices(base, [base, base])
    segs = segments.field_pulse_segments(pairs, (1.0, 0.0, 0.0), np.array([10.0, -5.0, 7.0]), 0.0513, 0.2,
                                         times[-1])
    bp = batch.BatchedLanczosPropagator(base, minimum_dimension=16, maximum_dimension=32, dimension_increment=8)
    out = segments.propagate_segments_batched(bp, states, times, segs, 1e-9,
                                              yield_weights=quadrature.yield_weight_table(times, 1.0))
    for b in range(2):
        assert abs(out.yields[b, 0] - quadrature.integrate_yield(times, out.populations[b], 1.0)) < 1e-14
    assert math.isfinite(float(out.yields.sum()))
