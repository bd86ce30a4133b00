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


def test_adjoint_sweep_rejects_out_of_range_slots_and_segments():
    """rpm_adjoint_sweep checks every step's slot and segment against nslot / nseg on the host, before any
    launch, so a bad index gives RPM_E_INVALID instead of a write past the accumulators."""
    import ctypes

    from rpm_amd import _native as nat
    print("case: adjoint sweep index checks", flush=True)
    h, s0, _ = _small()
    lib = nat.load()
    op = h._device_op()
    B, kmax = 1, 4
    psi = torch.from_numpy(s0.reshape(1, -1)).cuda()
    ws = nat.workspace(lib.rpm_adjoint_workspace_bytes(ctypes.c_void_p(op.handle), B, kmax))
    acc = torch.zeros((2, B, 16), dtype=torch.complex128, device=psi.device)
    pairs = torch.zeros((2, B, 16), dtype=torch.complex128, device=psi.device)
    G = torch.zeros((1, kmax + 1, kmax + 1), dtype=torch.complex128, device=psi.device)
    nterms = np.array([kmax], dtype=np.int32)
    coef = np.zeros(2 * (kmax + 1))
    cls = np.zeros(1, dtype=np.int32)
    mult = np.ones(1)
    ip = ctypes.POINTER(ctypes.c_int)
    dp = ctypes.POINTER(ctypes.c_double)

    def call(slot, nslot, seg, nseg):
        sl = np.array([slot], dtype=np.int32)
        sg = np.array([seg], dtype=np.int32)
        return lib.rpm_adjoint_sweep(
            ctypes.c_void_p(op.handle), B, ctypes.c_void_p(nat.ptr(psi)), 1, 0.0, 1.0, 1, nterms.ctypes.data_as(ip),
            kmax, coef.ctypes.data_as(dp), ctypes.c_void_p(nat.ptr(G)), cls.ctypes.data_as(ip),
            mult.ctypes.data_as(dp), ctypes.c_void_p(nat.ptr(pairs)), sg.ctypes.data_as(ip), nseg,
            ctypes.c_void_p(nat.ptr(acc)), sl.ctypes.data_as(ip), nslot, ctypes.c_void_p(nat.ptr(ws)), ws.numel(),
            ctypes.c_void_p(nat.stream_ptr()))

    assert call(2, 2, 0, 2) == nat.RPM_E_INVALID      # slot == nslot
    assert call(-1, 2, 0, 2) == nat.RPM_E_INVALID
    assert call(0, 2, 2, 2) == nat.RPM_E_INVALID      # segment == nseg
    assert call(0, 2, 0, 0) == nat.RPM_E_INVALID      # pairs without a segment count
    assert call(0, 0, 0, 2) == nat.RPM_E_INVALID      # no accumulator slot
    torch.cuda.synchronize()
    assert torch.count_nonzero(acc).item() == 0       # nothing ran
