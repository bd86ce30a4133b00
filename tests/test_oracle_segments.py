"""CPU: independent pins of the oracle's segment loop and its pieces.

The oracle's ``run_segments`` (R/rpm_nuclear_spin_rotation_control.py:247-342),
``local_unitary_apply`` and ``spin_pi_pulse`` (R/poster_reproduction/core.py:285-308)
are checked against what the mathematics fixes, with no shared code path:

* spin_pi_pulse against its closed forms: exp(-i pi n.S) = -i n.sigma = -2i n.S for
  spin 1/2, and 1 - 2 (n.S)^2 for spin 1 (since (n.S)^3 = n.S there);
* local_unitary_apply against the explicit Kronecker embedding 1 (x) U (x) 1;
* the whole segment loop (pi pulses between segments; piecewise-constant fields)
  against exact piecewise dynamics of the dense Hamiltonian built by index arithmetic
  (oracle.dense.dense_from_terms), eigendecomposition per piece, on the 576-level system;
* schedule replay: a replayed run reproduces the recorded run bitwise.
"""

import math

import numpy as np
import pytest

from oracle import core as ocore
from oracle import dense as odense
from oracle import engine as oengine
from oracle.model import spin_operators
from rpm_amd import segments as psegments
from rpm_amd import workloads


def _axes(n, seed):
    rng = np.random.default_rng(seed)
    return [rng.normal(size=3) for _ in range(n)] + [np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])]


@pytest.mark.parametrize("pulse", [ocore.spin_pi_pulse, psegments.spin_pi_pulse])
def test_spin_pi_pulse_closed_forms(pulse):
    for spin in (0.5, 1.0):
        ops = spin_operators(spin)
        for n in _axes(6, 3):
            u = n / np.linalg.norm(n)
            ns = u[0] * ops[0] + u[1] * ops[1] + u[2] * ops[2]
            want = -2.0j * ns if spin == 0.5 else np.eye(3) - 2.0 * ns @ ns
            got = pulse(spin, n)
            assert np.abs(got - want).max() < 1e-13, (spin, n)
            assert np.abs(got.conj().T @ got - np.eye(got.shape[0])).max() < 1e-13


def test_spin_one_operators_cube_identity():
    """(n.S)^3 = n.S for spin 1, the identity the closed form above rests on."""
    ops = spin_operators(1.0)
    for n in _axes(4, 5):
        u = n / np.linalg.norm(n)
        ns = u[0] * ops[0] + u[1] * ops[1] + u[2] * ops[2]
        assert np.abs(ns @ ns @ ns - ns).max() < 1e-14


def test_local_unitary_apply_matches_kron_embedding():
    dims = (2, 2, 3, 2, 3, 2)
    n = math.prod(dims)
    rng = np.random.default_rng(11)
    psi = rng.normal(size=n) + 1j * rng.normal(size=n)
    for axis in range(len(dims)):
        d = dims[axis]
        a = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
        u, _ = np.linalg.qr(a)
        left = math.prod(dims[:axis])
        right = math.prod(dims[axis + 1:])
        full = np.kron(np.kron(np.eye(left), u), np.eye(right))
        got = ocore.local_unitary_apply(psi, dims, axis, u)
        assert np.abs(got - full @ psi).max() < 1e-13, axis


def _dense(h):
    return odense.dense_from_terms(h.dimensions, h.terms)


def _singlet(v):
    n = v.size
    s = (v[n // 4:n // 2] - v[n // 2:3 * n // 4]) / math.sqrt(2.0)
    return float(np.vdot(s, s).real)


def _exact_pieces(pieces, psi0, times, dims, axis=None, unitary=None):
    """P_S(t_k) and the final state of exact piecewise dynamics; ``pieces`` = [(t_end, H)], the unitary
    (embedded by Kronecker products) acting after every piece but the last."""
    psi = np.asarray(psi0, complex).copy()
    pops = np.empty(times.size)
    full_u = None
    if unitary is not None:
        full_u = np.kron(np.kron(np.eye(math.prod(dims[:axis])), unitary), np.eye(math.prod(dims[axis + 1:])))
    start, k = 0.0, 0
    for i, (t_end, h) in enumerate(pieces):
        w, q = np.linalg.eigh(h)
        c = q.conj().T @ psi
        while k < times.size and times[k] <= t_end + 1e-12:
            pops[k] = _singlet(q @ (c * np.exp(-1j * w * (times[k] - start))))
            k += 1
        psi = q @ (c * np.exp(-1j * w * (t_end - start)))
        if full_u is not None and i < len(pieces) - 1:
            psi = full_u @ psi
        start = t_end
    assert k == times.size
    return pops, psi


def _make(h):
    return oengine.AdaptiveLanczosPropagator(h, minimum_dimension=16, maximum_dimension=32, dimension_increment=8,
                                             embedded_offset=4)


def test_run_segments_with_pi_pulses_matches_exact_piecewise_dynamics():
    sys_ = workloads.small_system(0)
    h = sys_.hamiltonian(builder=ocore.build_hamiltonian)
    dims = tuple(h.dimensions)
    times = np.arange(481) * 0.00125
    pulses = np.array([0.13, 0.26, 0.39, 0.5])
    axis = 2
    U = ocore.spin_pi_pulse(1.0, psegments.dominant_principal_axis(sys_.nuclei[0].tensor_mT))
    psi0 = ocore.random_phase_initial_state(dims, 0, 5)
    budget = 1e-10
    pops, bounds, acc, final, sched = ocore.run_segments(h, psi0, times, pulses, _make, budget, axis, U)
    hd = _dense(h)
    ends = list(pulses) + [times[-1]]
    want, want_final = _exact_pieces([(t, hd) for t in ends], psi0, times, dims, axis, U)
    assert len(sched) == len(ends)
    err = np.abs(pops - want)
    print("pi-pulse segment loop: max |P - exact|", err.max(), "accumulated", acc, flush=True)
    assert err.max() < 1e-9
    assert np.all(err <= bounds + 1e-12)          # the loop's own error bounds hold
    assert np.linalg.norm(final - want_final) < 1e-8
    assert acc <= budget * (1 + 1e-12)
    # the pulses matter: without them the populations differ at O(1e-2) or more
    no_pulse, _ = _exact_pieces([(times[-1], hd)], psi0, times, dims)
    assert np.abs(no_pulse - want).max() > 1e-3


def test_run_segments_with_field_pieces_matches_exact_piecewise_dynamics():
    sys_ = workloads.small_system(0)
    dims = sys_.dimensions
    times = np.arange(401) * 0.00125
    fields = [(0.0, 0.0, 0.05), (30.0, 10.0, 0.07), (75.0, -20.0, 0.02), (0.0, 0.0, 0.05)]
    hs = [sys_.hamiltonian(theta_deg=t, field_mT=f, phi_deg=p, builder=ocore.build_hamiltonian)
          for t, p, f in fields]
    bounds_t = np.array([0.1, 0.2375, 0.31])
    psi0 = ocore.random_phase_initial_state(dims, 0, 2)
    pops, bounds, acc, final, sched = ocore.run_segments(hs[0], psi0, times, bounds_t, _make, 1e-10, hamiltonians=hs)
    ends = list(bounds_t) + [times[-1]]
    want, want_final = _exact_pieces([(t, _dense(hh)) for t, hh in zip(ends, hs)], psi0, times, dims)
    err = np.abs(pops - want)
    print("field-piece segment loop: max |P - exact|", err.max(), flush=True)
    assert err.max() < 1e-9
    assert np.all(err <= bounds + 1e-12)
    assert np.linalg.norm(final - want_final) < 1e-8
    # and the same via the piecewise Simpson-yield oracle
    phi, pops_d = odense.dense_grid_yield([(t, _dense(hh)) for t, hh in zip(ends, hs)], psi0, times, 1.0)
    assert np.abs(pops_d - want).max() < 1e-12
    assert abs(ocore.integrate_yield(times, pops, 1.0) - phi) < 1e-9
    with pytest.raises(ValueError):
        ocore.run_segments(hs[0], psi0, times, bounds_t, _make, 1e-10, hamiltonians=hs[:2])


def test_schedule_replay_reproduces_recorded_run():
    sys_ = workloads.small_system(0)
    h = sys_.hamiltonian(builder=ocore.build_hamiltonian)
    times = np.arange(641) * 0.00125
    psi0 = ocore.random_phase_initial_state(h.dimensions, 0, 1)
    obs = oengine.SingletPopulationObservable(h.dimensions)
    prop = oengine.AdaptiveLanczosPropagator(h, minimum_dimension=8, maximum_dimension=16, dimension_increment=4,
                                             embedded_offset=2, state_error_budget=1e-10)
    free = prop.propagate(psi0, times, obs)
    assert free.stats.rejected_steps > 0            # the replay covers rejections too
    again = prop.propagate(psi0, times, obs, schedule=free.schedule)
    assert again.schedule == free.schedule
    assert again.stats == free.stats
    np.testing.assert_array_equal(again.observable_values, free.observable_values)
    np.testing.assert_array_equal(again.final_state, free.final_state)
    # at a nearby field the replay follows the recorded steps and dimensions
    h2 = sys_.hamiltonian(theta_deg=0.01, builder=ocore.build_hamiltonian)
    prop2 = oengine.AdaptiveLanczosPropagator(h2, minimum_dimension=8, maximum_dimension=16, dimension_increment=4,
                                              embedded_offset=2, state_error_budget=1e-10)
    near = prop2.propagate(psi0, times, obs, schedule=free.schedule)
    assert [r[:3] for r in near.schedule] == [r[:3] for r in free.schedule]
    # a schedule that ends early is an error
    with pytest.raises(RuntimeError):
        prop.propagate(psi0, times, obs, schedule=free.schedule[:3])
