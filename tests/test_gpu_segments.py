"""GPU: segmented propagation (pulses, local unitaries) against the oracle's segment loop.

Bars: local unitaries bit-close to the NumPy moveaxis form (1e-14); segmented
populations within 1e-10 of the oracle with identical per-segment schedules;
batched segments reproduce the single-trajectory segment loop.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from oracle import core as ocore  # noqa: E402
from oracle import engine as oengine  # noqa: E402
from rpm_amd import batch, core, engine, segments, workloads  # noqa: E402


def test_local_unitary_matches_oracle():
    sys_ = workloads.small_system(0)
    dims = sys_.dimensions
    rng = np.random.default_rng(1)
    psi = rng.normal(size=int(np.prod(dims))) + 1j * rng.normal(size=int(np.prod(dims)))
    for axis in (2, 3, 5, 7):
        U = segments.spin_pi_pulse((dims[axis] - 1) / 2, rng.normal(size=3))
        got = segments.local_unitary_apply(psi, dims, axis, U)
        want = ocore.local_unitary_apply(psi, dims, axis, U)
        assert np.abs(got - want).max() < 1e-14, axis
    # full system, batched
    full = core.build_hamiltonian(90.0, 0.05).dimensions
    n = int(np.prod(full))
    x = torch.from_numpy(workloads.normal_vectors(n, 2, seed=4)).cuda()
    U = segments.spin_pi_pulse(1.0, [0.3, -0.5, 0.8])
    y = segments.local_unitary_apply(x, full, 3, U).cpu().numpy()
    for b in range(2):
        want = ocore.local_unitary_apply(x[b].cpu().numpy(), full, 3, U)
        assert np.abs(y[b] - want).max() < 1e-14


def _kw():
    return dict(minimum_dimension=8, maximum_dimension=16, dimension_increment=4, embedded_offset=2,
                initial_step_us=0.02, maximum_step_us=0.1)


def test_segment_loop_with_pi_pulses_matches_oracle():
    sys_ = workloads.small_system(0)
    h = sys_.hamiltonian()
    o = sys_.hamiltonian(builder=ocore.build_hamiltonian)
    times = np.arange(321) * 0.00125
    pulses = segments.pulse_schedule(0.13, times[-1])
    axis = 2
    U = segments.spin_pi_pulse(1.0, segments.dominant_principal_axis(sys_.nuclei[0].tensor_mT))
    state = core.random_phase_initial_state(h.dimensions, 0, 5)
    budget = 1e-7
    segs = [segments.Segment(float(t), unitary=(axis, U)) for t in pulses] + [segments.Segment(float(times[-1]))]
    got = segments.propagate_segments(h, state, times, segs, lambda hh: engine.AdaptiveLanczosPropagator(hh, **_kw()),
                                      budget)
    pops, bounds, acc, final, sched = ocore.run_segments(
        o, state, times, pulses, lambda hh: oengine.AdaptiveLanczosPropagator(hh, **_kw()), budget, axis, U)
    assert [[(r[1], r[2], r[3]) for r in s[0]] for s in got.schedules] == [[(r[1], r[2], r[3]) for r in s]
                                                                          for s in sched]
    np.testing.assert_allclose(got.populations[0], pops, rtol=0, atol=1e-10)
    np.testing.assert_allclose(got.bounds[0], bounds, rtol=1e-6, atol=1e-13)   # round-off-level estimates
    assert abs(got.accumulated_state_error[0] - acc) <= 1e-6 * acc + 1e-13
    np.testing.assert_allclose(got.final_states[0].cpu().numpy(), final, rtol=0, atol=1e-10)


def test_batched_segments_reproduce_single_segment_loop():
    sys_ = workloads.small_system(0)
    base = sys_.hamiltonian()
    times = np.arange(241) * 0.00125
    states = np.stack([core.random_phase_initial_state(base.dimensions, 0, i) for i in range(3)])
    orients = [(0.0, 0.0, 0.05), (50.0, 10.0, 0.05), (120.0, 0.0, 0.03)]
    hs = [sys_.hamiltonian(theta_deg=t, field_mT=f, phi_deg=p) for t, p, f in orients]
    pairs = batch.pair_matrices(base, hs)
    U = segments.spin_pi_pulse(0.5, [0.0, 1.0, 0.0])
    amps = np.array([12.0, -7.0, 3.0])
    segs = segments.field_pulse_segments(pairs, (1.0, 0.0, 0.0), amps, 0.05, 0.2, times[-1])
    segs[1].unitary = (3, U)
    budget = 1e-8
    bp = batch.BatchedLanczosPropagator(base, **_kw())
    got = segments.propagate_segments_batched(bp, states, times, segs, budget)
    for b in range(3):
        # a pulse segment's electron matrix is that of the Hamiltonian built with the total field
        th, ph, f = orients[b]
        from rpm_amd.model import field_vector_mT
        for s_idx, a in zip(range(1, 4), amps):
            btot = field_vector_mT(th, 1e3 * f, ph) + 1e-3 * a * np.array([1.0, 0.0, 0.0])
            mag = float(np.linalg.norm(btot))
            th2 = float(np.degrees(np.arccos(btot[2] / mag)))
            ph2 = float(np.degrees(np.arctan2(btot[1], btot[0])))
            want = batch.electron_pair_matrix(sys_.hamiltonian(theta_deg=th2, field_mT=mag, phi_deg=ph2))
            assert np.abs(segs[s_idx].pairs[b] - want).max() < 1e-10
        one_segs = [segments.Segment(sg.end_us, pairs=sg.pairs[b:b + 1], unitary=sg.unitary) for sg in segs]
        bp1 = batch.BatchedLanczosPropagator(base, **_kw())
        ref = segments.propagate_segments_batched(bp1, states[b:b + 1], times, one_segs, budget)
        np.testing.assert_allclose(got.populations[b], ref.populations[0], rtol=0, atol=1e-13)
        assert [sc[b] for sc in got.schedules] == [sc[0] for sc in ref.schedules]
