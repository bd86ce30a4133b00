"""GPU: the full-size (dim 221,184) workloads against recorded oracle runs (tests/golden/make_full_fixtures.py).

* Random-tensor 14-nucleus system, 32 random-phase vectors batched over the whole 5/k horizon
  (lifetime 1 us, 4,001 grid points), Krylov tolerance 1e-10 taken as the state-error budget, maximum
  dimension 40 (ladder 24, 32, 40): vectors 0 and 1 against the oracle, identical step schedules,
  yields within 1e-9; every yield of the batch summed on the device.
* Production operator, (theta, phi) = (90, 0) and (60, 30) deg, 50 uT, 0.1 us: the adjoint gradient
  with respect to (theta, phi, |B|) against central differences of the oracle's FP64 yield
  (h = 1e-4 rad, 1e-3 uT; the oracle's perturbed runs replay its centre schedule); relative error
  <= 1e-5.
* 64-segment transverse pulse (amplitudes U(-20, 20) uT over [0, 2] us) on 50 uT static along z,
  horizon 2.1 us: per-segment schedules and yields of vectors 0 and 1 against the oracle's segment loop
  (R/rpm_nuclear_spin_rotation_control.py:247-342) within 1e-9, and the adjoint gradient with respect to
  amplitudes 0, 21, 42, 63 against the oracle's central differences (h = 1e-3 uT), relative error <= 1e-5.
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from rpm_amd import adjoint, batch, core, quadrature, workloads  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
TOL_GRAD = 1e-5


def _unpack(rows):
    """(n, 6) rows (segment, trial, accepted, selected, size0, size1 or -1) -> per-segment schedules."""
    out = []
    for seg, trial, ok, sel, s0, s1 in np.asarray(rows, dtype=float):
        while len(out) <= int(seg):
            out.append([])
        sizes = (int(s0),) if s1 < 0 else (int(s0), int(s1))
        out[int(seg)].append((float(trial), bool(ok), int(sel), sizes))
    return out


def _same_schedule(got, want):
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert (g[1], g[2], tuple(g[3])) == (w[1], w[2], w[3])
        assert abs(g[0] - w[0]) <= 1e-15


def _rel(g, fd, scale):
    return abs(g - fd) / max(abs(fd), scale)


def test_random_tensor_system_32_vectors_tol1e10_dim40_matches_oracle():
    print("case: random-tensor 14-nucleus system, 32 vectors, 5 us, tol 1e-10, max dim 40", flush=True)
    d = np.load(os.path.join(GOLDEN, "full_random_tol1e10.npz"))
    sys_ = workloads.full_random_system(1)
    h = sys_.hamiltonian()
    times = d["times"]
    assert times.size == 4001 and abs(times[-1] - 5.0) < 1e-12
    states = np.stack([core.random_phase_initial_state(h.dimensions, 0, i) for i in range(32)])
    bp = batch.BatchedLanczosPropagator(h, minimum_dimension=24, maximum_dimension=40, dimension_increment=8,
                                        embedded_offset=4, safety_factor=2.0, state_error_budget=1e-10,
                                        initial_step_us=0.05, maximum_step_us=0.25)
    coarse = np.arange(0, times.size, 2)
    res = bp.propagate(states, times, yield_weights=quadrature.yield_weight_table(times, 1.0, coarse))
    yields = np.array([r.device_yields[0] for r in res])
    print("yields[:4]", yields[:4], "oracle", d["yields"], flush=True)
    for b in (0, 1):
        _same_schedule(res[b].schedule, _unpack(d[f"schedule{b}"])[0])
        assert res[b].stats.accepted_steps == int(d["accepted"][b])
        assert res[b].stats.matvecs == int(d["matvecs"][b])
        assert abs(yields[b] - d["yields"][b]) < 1e-9
        assert abs(res[b].device_yields[1] - d["coarse_yields"][b]) < 1e-9
        assert np.abs(res[b].observable_values - d["populations"][b]).max() < 1e-9
    assert np.all((yields > 0.0) & (yields < 1.0))
    assert np.all(bp.device_matvecs <= np.array([r.stats.matvecs for r in res]))


def test_full_system_field_gradient_matches_oracle_central_differences():
    print("case: production operator, 2 orientations, 0.1 us, adjoint vs oracle differences", flush=True)
    d = np.load(os.path.join(GOLDEN, "full_field_gradient_fd.npz"))
    base = core.build_hamiltonian(90.0, 0.05)
    times = np.arange(81) * 0.00125
    orient = d["orientations"]
    states = np.stack([core.random_phase_initial_state(base.dimensions, core.MASTER_SEED, int(s))
                       for _, _, s in orient])
    bp = batch.BatchedLanczosPropagator(base, state_error_budget=float(d["budget"]))
    build = lambda t, fm, p: core.build_hamiltonian(t, fm, phi_deg=p)  # noqa: E731
    out = adjoint.yield_and_gradient(base, build, bp, states, times, 5.0, orient[:, 0], orient[:, 1],
                                     [float(d["field_uT"])] * len(orient))
    fd = d["fd_gradient"]
    for b in range(len(orient)):
        assert abs(out.yields[b] - d["center_yields"][b]) < 1e-9
        scale = 1e-3 * np.abs(fd[b]).max()
        errs = [_rel(out.field_gradient[b, k], fd[b, k], scale) for k in range(3)]
        print(f"b={b} adjoint {out.field_gradient[b]} oracle fd {fd[b]} rel {errs}", flush=True)
        assert max(errs) <= TOL_GRAD


def test_pulse64_full_system_matches_oracle_segment_loop_and_differences():
    print("case: 64-segment transverse pulse, production operator, 2.1 us", flush=True)
    d = np.load(os.path.join(GOLDEN, "full_pulse64.npz"))
    base = core.build_hamiltonian(0.0, 0.05)
    times = np.arange(1681) * 0.00125
    amps = d["amplitudes_uT"]
    states = np.stack([core.random_phase_initial_state(base.dimensions, core.MASTER_SEED, i) for i in range(2)])
    bp = batch.BatchedLanczosPropagator(base, state_error_budget=float(d["budget"]))
    build = lambda t, fm, p: core.build_hamiltonian(t, fm, phi_deg=p)  # noqa: E731
    prob = adjoint.FieldYield(base, build, bp, states, times, 5.0, pulse_window_us=tuple(d["window_us"]),
                              pulse_direction=(1.0, 0.0, 0.0))
    y = prob.forward([0.0, 0.0], [0.0, 0.0], [50.0, 50.0], np.tile(amps, (2, 1)))
    sched = prob.saved_state()["schedules"]
    print("yields", y, "oracle", d["yields"], flush=True)
    for b in range(2):
        want = _unpack(d[f"schedule{b}"])
        assert len(sched) == len(want) == 65
        for s in range(65):
            _same_schedule(sched[s][b], want[s])
        assert abs(y[b] - d["yields"][b]) < 1e-9
    _, pgrad, rec = prob.backward()
    print("psi0 recovered to", rec, flush=True)
    assert rec < 1e-8
    fd = d["fd_gradient"]
    got = pgrad[0, d["fd_indices"]]
    scale = 1e-3 * np.abs(fd).max()
    errs = [_rel(got[i], fd[i], scale) for i in range(fd.size)]
    print("adjoint", got, "oracle fd", fd, "rel", errs, flush=True)
    assert max(errs) <= TOL_GRAD
    assert np.all(np.isfinite(pgrad)) and pgrad.shape == (2, 64)
    assert math.isfinite(float(prob.saved_state()["pops"].sum()))
