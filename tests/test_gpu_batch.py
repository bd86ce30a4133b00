"""GPU: batched trajectories (many states x field orientations/strengths) against per-trajectory runs.

Bars: per-vector electron matrices give the H·psi of the separately built
Hamiltonian (relative L2 <= 1e-13, checked against the oracle); a batched
basis equals the single-trajectory basis; a batched propagation reproduces the
plain propagation of every trajectory (same schedule, values within 1e-12).
"""

import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from oracle import core as ocore  # noqa: E402
from rpm_amd import _native as nat  # noqa: E402
from rpm_amd import batch, core, engine, workloads  # noqa: E402

ORIENTATIONS = [(90.0, 0.0, 0.05), (0.0, 0.0, 0.05), (37.0, 115.0, 0.025), (141.0, -60.0, 0.05)]


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_per_vector_electron_matrices_match_separate_hamiltonians():
    base = core.build_hamiltonian(90.0, 0.05)
    hs = [core.build_hamiltonian(t, f, phi_deg=p) for t, p, f in ORIENTATIONS]
    pairs = batch.pair_matrices(base, hs)
    v = workloads.normal_vectors(base.dimension, len(hs), seed=21)
    x = torch.from_numpy(v).cuda()
    y = torch.empty_like(x)
    dp = torch.from_numpy(pairs.reshape(len(hs), 16).copy()).cuda()
    op = base._device_op()
    print("per-vector pairs, batch", len(hs), "kernel", op.kernel, flush=True)
    nat.check(nat.load().rpm_operator_apply_pairs(ctypes.c_void_p(op.handle), ctypes.c_void_p(nat.ptr(x)),
                                                  ctypes.c_void_p(nat.ptr(y)), len(hs), base.dimension,
                                                  ctypes.c_void_p(nat.ptr(dp)), ctypes.c_void_p(nat.stream_ptr())))
    yh = y.cpu().numpy()
    for i, (t, p, f) in enumerate(ORIENTATIONS):
        o = ocore.build_hamiltonian(t, f, phi_deg=p)
        assert _rel(yh[i], o.matvec(v[i])) <= 1e-13, (t, p, f)
    # the gather kernel takes the same per-vector matrices
    try:
        op.set_kernel(1)
        y2 = torch.empty_like(x)
        nat.check(nat.load().rpm_operator_apply_pairs(ctypes.c_void_p(op.handle), ctypes.c_void_p(nat.ptr(x)),
                                                      ctypes.c_void_p(nat.ptr(y2)), len(hs), base.dimension,
                                                      ctypes.c_void_p(nat.ptr(dp)),
                                                      ctypes.c_void_p(nat.stream_ptr())))
    finally:
        op.set_kernel(0)
    assert _rel(y2.cpu().numpy(), yh) <= 1e-13


def _basis_batch(h, states, targets, m, pairs=None):
    lib = nat.load()
    op = h._device_op()
    B, n = states.shape
    V = torch.zeros((B, m, n), dtype=torch.complex128, device="cuda")
    alpha = torch.full((B, m), -7.0, dtype=torch.float64, device="cuda")
    beta = torch.full((B, m), -7.0, dtype=torch.float64, device="cuda")
    info = torch.full((B, 4), -1, dtype=torch.int32, device="cuda")
    norm = torch.full((B,), -1.0, dtype=torch.float64, device="cuda")
    ws = nat.workspace(lib.rpm_lanczos_batch_workspace_bytes(ctypes.c_void_p(op.handle), m, B))
    tg = torch.tensor(targets, dtype=torch.int32, device="cuda")
    dp = None if pairs is None else torch.from_numpy(pairs.reshape(B, 16).copy()).cuda()
    vp = ctypes.c_void_p
    nat.check(lib.rpm_lanczos_basis_batch(vp(op.handle), B, vp(nat.ptr(states)), n, vp(nat.ptr(tg)),
                                          max(targets), vp(nat.ptr(dp) if dp is not None else None), vp(nat.ptr(V)),
                                          m, vp(nat.ptr(alpha)), vp(nat.ptr(beta)), vp(nat.ptr(info)),
                                          vp(nat.ptr(norm)), 1e-14, vp(nat.ptr(ws)), ws.numel(),
                                          vp(nat.stream_ptr())))
    return V, alpha, beta, info, norm


def test_batched_basis_matches_single_basis():
    h = core.build_hamiltonian(90.0, 0.05)
    n = h.dimension
    states = torch.from_numpy(workloads.normal_vectors(n, 4, seed=5) * np.array([[1.0], [2.5], [0.3], [1.0]])).cuda()
    targets = [12, 20, 0, 7]
    print("batched basis targets", targets, flush=True)
    V, alpha, beta, info, norm = _basis_batch(h, states, targets, 24)
    info_h = info.cpu().numpy()
    assert info_h[2, 0] == -1 and norm[2].item() == -1.0, "a trajectory with target 0 must stay untouched"
    assert np.all(alpha[2].cpu().numpy() == -7.0)
    prop = engine.AdaptiveLanczosPropagator(h, minimum_dimension=8, maximum_dimension=24)
    prop.paired_steps = False       # the whole-basis entry point takes one step per read of the basis
    for b, t in enumerate(targets):
        if t == 0:
            continue
        size = prop._basis(states[b].contiguous(), t)
        ws = prop._workspace()
        assert info_h[b, 0] == size == t
        np.testing.assert_array_equal(alpha[b, :t].cpu().numpy(), ws.alpha[:t].cpu().numpy())
        np.testing.assert_array_equal(beta[b, :t - 1].cpu().numpy(), ws.beta[:t - 1].cpu().numpy())
        np.testing.assert_array_equal(V[b, :t].cpu().numpy(), ws.V[:t].cpu().numpy())
        assert norm[b].item() == ws.norm.item()


def test_batched_basis_with_orientations_matches_separate_operators():
    base = core.build_hamiltonian(90.0, 0.05)
    hs = [core.build_hamiltonian(t, f, phi_deg=p) for t, p, f in ORIENTATIONS[:3]]
    pairs = batch.pair_matrices(base, hs)
    n = base.dimension
    states = torch.from_numpy(workloads.normal_vectors(n, 3, seed=6)).cuda()
    V, alpha, beta, info, norm = _basis_batch(base, states, [10, 10, 10], 10, pairs)
    for b, h in enumerate(hs):
        prop = engine.AdaptiveLanczosPropagator(h, minimum_dimension=8, maximum_dimension=10)
        prop._basis(states[b].contiguous(), 10)
        ws = prop._workspace()
        np.testing.assert_allclose(alpha[b].cpu().numpy(), ws.alpha.cpu().numpy(), rtol=0, atol=1e-11)
        np.testing.assert_allclose(beta[b, :9].cpu().numpy(), ws.beta[:9].cpu().numpy(), rtol=0, atol=1e-11)
        assert _rel(V[b].cpu().numpy(), ws.V.cpu().numpy()) < 1e-12


def test_batched_breakdown_is_per_trajectory():
    # diagonal pair Hamiltonian: every basis state is an exact eigenvector (breakdown after one vector),
    # a random state spans as many Krylov vectors as H has distinct eigenvalues on its support
    from rpm_amd import model
    sz = model.spin_operators(0.5)[2]
    terms = [engine.LocalTerm((0,), 1.3 * sz), engine.LocalTerm((1,), 0.7 * sz),
             engine.LocalTerm((0, 2), 2.1 * np.kron(sz, sz)), engine.LocalTerm((1, 3), -0.9 * np.kron(sz, sz))]
    h = engine.CompiledFullPairHamiltonian((2, 2, 2, 2), terms, left_axes=(0, 2), right_axes=(1, 3))
    v = workloads.normal_vectors(16, 2, seed=3)
    e3 = np.zeros(16, dtype=complex)
    e3[3] = 1.0
    states = torch.from_numpy(np.stack([v[0], e3, v[1]])).cuda()
    V, alpha, beta, info, norm = _basis_batch(h, states, [4, 4, 3], 4)
    sizes = info[:, 0].cpu().numpy()
    assert list(sizes) == [4, 1, 3]
    assert abs(alpha[1, 0].item() - h.explicit_matrix()[3, 3].real) < 1e-14


def _small_kw():
    return dict(minimum_dimension=8, maximum_dimension=16, dimension_increment=4, embedded_offset=2,
                state_error_budget=1e-7, initial_step_us=0.02, maximum_step_us=0.1)


def test_batched_propagation_reproduces_plain_runs_small_system():
    sys_ = workloads.small_system(0)
    base = sys_.hamiltonian()
    orients = [(0.0, 0.0, 0.05), (70.0, 30.0, 0.05), (0.0, 0.0, 0.05), (120.0, 0.0, 0.025), (45.0, 45.0, 0.1)]
    hs = [sys_.hamiltonian(theta_deg=t, field_mT=f, phi_deg=p) for t, p, f in orients]
    pairs = batch.pair_matrices(base, hs)
    states = np.stack([core.random_phase_initial_state(base.dimensions, 0, i) for i in range(len(hs))])
    times = np.linspace(0.0, 0.5, 401)
    bp = batch.BatchedLanczosPropagator(base, **_small_kw())
    res = bp.propagate(states, times, pairs=pairs)
    print("batched rounds", bp.rounds, flush=True)
    for i, h in enumerate(hs):
        p = engine.AdaptiveLanczosPropagator(h, **_small_kw())
        ref = p.propagate(states[i], times, engine.SingletPopulationObservable(h.dimensions))
        assert res[i].schedule == ref.schedule, i
        np.testing.assert_allclose(res[i].observable_values, ref.observable_values, rtol=0, atol=1e-12)
        np.testing.assert_allclose(res[i].state_error_estimates, ref.state_error_estimates, rtol=1e-9, atol=1e-15)
        assert res[i].stats.matvecs == ref.stats.matvecs
        assert res[i].stats.accepted_steps == ref.stats.accepted_steps
        np.testing.assert_allclose(res[i].final_state_device.cpu().numpy(), ref.final_state, rtol=0, atol=1e-12)
        # replay of the plain schedule
    rep = bp.propagate(states[:2], times, pairs=pairs[:2], schedules=[res[0].schedule, None])
    np.testing.assert_allclose(rep[0].observable_values, res[0].observable_values, rtol=0, atol=1e-13)


def test_batched_propagation_full_system_orientations():
    base = core.build_hamiltonian(90.0, 0.05)
    hs = [core.build_hamiltonian(t, f, phi_deg=p) for t, p, f in ORIENTATIONS[:3]]
    pairs = batch.pair_matrices(base, hs)
    states = np.stack([core.random_phase_initial_state(base.dimensions, core.MASTER_SEED, i) for i in range(3)])
    times = np.arange(17) * 0.00125
    kw = dict(state_error_budget=2.4e-5 * times[-1] / 57.56, initial_step_us=0.05)
    bp = batch.BatchedLanczosPropagator(base, **kw)
    res = bp.propagate(states, times, pairs=pairs)
    for i, h in enumerate(hs):
        o = ocore.build_hamiltonian(ORIENTATIONS[i][0], ORIENTATIONS[i][2], phi_deg=ORIENTATIONS[i][1])
        from oracle import engine as oengine
        op = oengine.AdaptiveLanczosPropagator(o, **kw)
        want = op.propagate(states[i], times, oengine.SingletPopulationObservable(o.dimensions))
        assert [(x[1], x[2], x[3]) for x in res[i].schedule] == [(x[1], x[2], x[3]) for x in want.schedule]
        np.testing.assert_allclose(res[i].observable_values, want.observable_values, rtol=0, atol=1e-10)


def test_batched_yields_match_plain_yield_small_system():
    """Config-0 shape: 576-dim random system, production Krylov rules, 4 random-phase vectors."""
    sys_ = workloads.small_system(0)
    h = sys_.hamiltonian()
    cfg = core.SweepConfig(lifetime_us=1.0, survival_tail_bound=math.exp(-0.5))
    grid = cfg.time_grid_us
    states = np.stack([core.random_phase_initial_state(h.dimensions, 0, i) for i in range(4)])
    bp = batch.BatchedLanczosPropagator.like(core.make_propagator(h, cfg))
    res = bp.propagate(states, grid)
    for i in range(4):
        want = core.propagate_trajectory(h, states[i], cfg)
        got = core.integrate_yield(grid, res[i].observable_values, cfg.lifetime_us)
        assert abs(got - want["yield"]) < 1e-12


def test_singlet_gram_batch_matches_dense_and_clamps_counts():
    """rpm_singlet_gram(_batch) against the dense singlet Gram: P[a][c] = sum conj(A_a) A_c with
    A_a = (V_a[ud] - V_a[du]) / sqrt2; per-trajectory counts above m are clamped to m, 0 skips."""
    nn, m = 1500, 40
    rng = np.random.default_rng(5)
    V = rng.standard_normal((3, m, 4 * nn)) + 1j * rng.standard_normal((3, m, 4 * nn))
    counts = np.array([m, 17, m + 9], dtype=np.int32)
    lib = nat.load()
    dV = torch.from_numpy(V.copy()).cuda()
    dk = torch.from_numpy(counts).cuda()
    P = torch.zeros((3, m * m), dtype=torch.complex128, device="cuda")
    ws = nat.workspace(lib.rpm_singlet_gram_batch_workspace_bytes(nn, 3))
    print("singlet gram batch, nn", nn, "m", m, "counts", counts.tolist(), flush=True)
    nat.check(lib.rpm_singlet_gram_batch(ctypes.c_void_p(nat.ptr(dV)), m, ctypes.c_void_p(nat.ptr(dk)), 3, nn,
                                         ctypes.c_void_p(nat.ptr(P)), ctypes.c_void_p(nat.ptr(ws)), ws.numel(),
                                         ctypes.c_void_p(nat.stream_ptr())))
    got = P.cpu().numpy()
    for b in range(3):
        k = min(int(counts[b]), m)
        A = (V[b, :k, nn:2 * nn] - V[b, :k, 2 * nn:3 * nn]) / math.sqrt(2.0)
        ref = A.conj() @ A.T
        Pb = got[b, :k * k].reshape(k, k)
        assert np.array_equal(Pb, Pb.conj().T)            # Hermitian by construction
        assert np.abs(Pb - ref).max() <= 1e-12 * np.abs(ref).max()
        # the single-trajectory entry point gives the same matrix bitwise
        single = engine.device_singlet_gram(dV[b, :k].contiguous(), k, nn).cpu().numpy()
        assert np.array_equal(single, Pb)
