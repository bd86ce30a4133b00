"""GPU parity: the HIP path through the C-ABI against the CPU oracle.

Bars: H·psi relative L2 error <= 1e-13; Lanczos coefficients / basis to
~1e-12; propagated singlet populations and yields within 1e-9 absolute on the
same seeds with identical step schedules.
"""

import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import core as ocore  # noqa: E402
from oracle import dense as odense  # noqa: E402
from oracle import engine as oengine  # noqa: E402
from rpm_amd import _native as nat  # noqa: E402
from rpm_amd import core, engine, model, workloads  # noqa: E402

HPSI_TOL = 1e-13


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def production_pair():
    h = core.build_hamiltonian(90.0, 0.05)
    o = ocore.build_hamiltonian(90.0, 0.05)
    return h, o


def test_gpu_visible():
    assert torch.cuda.is_available()
    print("device:", torch.cuda.get_device_name(0), flush=True)


def test_pair_stencil_full_system_matches_oracle(production_pair):
    h, o = production_pair
    assert h.uses_pair_stencil
    v = workloads.normal_vectors(h.dimension, 3, seed=0)
    for b in range(3):
        got = h.matvec(v[b])
        want = o.matvec(v[b])
        assert _rel(got, want) <= HPSI_TOL
    # independent dense-block evaluation: H psi = L M + M R^T + pair term
    M = np.transpose(v[0].reshape(o.dimensions), o.permutation).reshape(o.left_dimension, o.right_dimension)
    out = o.left_dense @ M + M @ o.right_dense.T
    four = M.reshape(2, o.left_dimension // 2, 2, o.right_dimension // 2)
    out += np.einsum("abcd,cxdy->axby", o.electron_pair_operator, four).reshape(M.shape)
    want2 = np.transpose(out.reshape(o.left_dimensions + o.right_dimensions), o.inverse_permutation).reshape(-1)
    assert _rel(h.matvec(v[0]), want2) <= HPSI_TOL


def test_pair_stencil_batched_and_torch_inputs(production_pair):
    h, o = production_pair
    v = workloads.normal_vectors(h.dimension, 4, seed=1)
    got = h.matmat(v.T)
    for b in range(4):
        assert _rel(got[:, b], o.matvec(v[b])) <= HPSI_TOL
    t = torch.from_numpy(v[2]).cuda()
    out = h.matvec(t)
    assert out.is_cuda
    assert _rel(out.cpu().numpy(), o.matvec(v[2])) <= HPSI_TOL


@pytest.mark.parametrize("theta", [0.0, 37.0, 90.0])
def test_pair_stencil_orientations_and_random_tensors(theta):
    sysr = workloads.full_random_system(seed=2)
    h = sysr.hamiltonian(theta_deg=theta)
    o = sysr.hamiltonian(theta_deg=theta, builder=ocore.build_hamiltonian)
    v = workloads.normal_vectors(h.dimension, 1, seed=3)[0]
    assert _rel(h.matvec(v), o.matvec(v)) <= HPSI_TOL


def test_generic_local_kernel_full_system():
    h = engine.build_quantum_needle_hamiltonian(90.0, efimova_hore_D_MHz=-9.7405, exchange_J_MHz=3.2468)
    o = oengine.build_quantum_needle_hamiltonian(90.0, efimova_hore_D_MHz=-9.7405, exchange_J_MHz=3.2468)
    v = workloads.normal_vectors(h.dimension, 1, seed=4)[0]
    assert _rel(h.matvec(v), o.matvec(v)) <= HPSI_TOL
    # force the generic gather kernel on the same terms
    g = engine._create_local(h.dimensions, h.terms)
    x = torch.from_numpy(v).cuda()
    y = torch.empty_like(x)
    g.apply(x, y)
    assert _rel(y.cpu().numpy(), o.matvec(v)) <= HPSI_TOL


def test_small_systems_match_brute_force():
    sx, sy, sz = model.spin_operators(0.5)
    sy1 = model.spin_operators(1.0)[1]
    dims = (2, 3, 2)
    terms = (engine.LocalTerm((0, 2), np.kron(sx, sz)), engine.LocalTerm((1,), sy1))
    h = engine.MatrixFreeHamiltonian(dims, terms)
    expected = np.kron(sx, np.kron(np.eye(3), sz)) + np.kron(np.eye(2), np.kron(sy1, np.eye(2)))
    np.testing.assert_allclose(h.explicit_matrix(), expected, atol=1e-13)
    rng = np.random.default_rng(5)
    a = rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16))
    h4 = engine.MatrixFreeHamiltonian((2, 2, 2, 2), [engine.LocalTerm((3, 0, 2, 1), a + a.conj().T)])
    brute = odense.dense_from_terms((2, 2, 2, 2), h4.terms)
    np.testing.assert_allclose(h4.explicit_matrix(), brute, atol=1e-12)
    sm = workloads.small_system(0)
    hs = sm.hamiltonian()
    assert hs.uses_pair_stencil
    np.testing.assert_allclose(hs.explicit_matrix(), sm.hamiltonian(builder=ocore.build_hamiltonian).explicit_matrix(),
                               atol=1e-11)


def test_singlet_population_and_norm():
    h = core.build_hamiltonian(0.0, 0.05)
    s = core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, 3)
    assert abs(engine.singlet_population(s, h.dimensions) - 1.0) < 1e-12
    v = workloads.normal_vectors(h.dimension, 1, seed=6)[0]
    assert abs(engine.singlet_population(v, h.dimensions) - oengine.singlet_population(v, h.dimensions)) < 1e-14
    t = torch.from_numpy(v * 3.0).cuda()
    assert abs(float(engine.device_norm(t).item()) - 3.0) < 1e-13


def _small_pair(seed=3):
    rng = np.random.default_rng(seed)

    def rsym():
        a = rng.normal(0.0, 0.3, size=(3, 3))
        return 0.5 * (a + a.T) + rng.uniform(0.0, 1.0) * np.eye(3)

    def traceless():
        a = rng.normal(size=(3, 3))
        a = 0.5 * (a + a.T)
        return (a - np.trace(a) / 3 * np.eye(3)) / 28.0

    nuclei = (model.Nucleus("N5", 1.0, rsym()), model.Nucleus("H6", 0.5, rsym()), model.Nucleus("N1", 1.0, rsym()))
    radicals = ("FAD", "FAD", "TrpH")
    q = {"FAD:N5": traceless(), "TrpH:N1": traceless()}
    h = core.build_hamiltonian(37.0, 0.05, nuclei=nuclei, radicals=radicals, quadrupole_tensors=q)
    o = ocore.build_hamiltonian(37.0, 0.05, nuclei=nuclei, radicals=radicals, quadrupole_tensors=q)
    return h, o


def test_lanczos_basis_matches_oracle_cgs2():
    h, o = _small_pair()
    state = core.random_phase_initial_state(h.dimensions, 11, 2)
    prop = engine.AdaptiveLanczosPropagator(h, minimum_dimension=8, maximum_dimension=16, embedded_offset=4)
    oprop = oengine.AdaptiveLanczosPropagator(o, minimum_dimension=8, maximum_dimension=16, embedded_offset=4)
    x = torch.from_numpy(state).cuda()
    size = prop._basis(x, 16)
    ov, oa, ob = oprop._basis(state, 16)
    ws = prop._workspace()
    assert size == oa.size == 16
    np.testing.assert_allclose(ws.alpha[:size].cpu().numpy(), oa, rtol=0, atol=1e-10 * np.abs(oa).max())
    np.testing.assert_allclose(ws.beta[: size - 1].cpu().numpy(), ob, rtol=1e-11)
    V = ws.V[:size].cpu().numpy().T
    np.testing.assert_allclose(V, ov, atol=1e-10)
    # orthonormality of the device basis
    assert np.abs(V.conj().T @ V - np.eye(size)).max() < 1e-13


def test_long_production_basis_with_explicit_corrections(production_pair):
    """64 Lanczos vectors of the production operator (dim 221,184) against the oracle's CGS2.

    The device applies the full projection one step late; the pending correction grows by about
    |H| / beta (x2) per step, so a 64-vector basis passes the explicit-correction threshold (1e-6)
    at least once: info[2] counts those corrections, and the basis must still match.
    """
    print("case: production basis m=64", flush=True)
    h, o = production_pair
    state = core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, 3)
    prop = engine.AdaptiveLanczosPropagator(h, minimum_dimension=32, maximum_dimension=64, embedded_offset=4)
    oprop = oengine.AdaptiveLanczosPropagator(o, minimum_dimension=32, maximum_dimension=64, embedded_offset=4)
    size = prop._basis(torch.from_numpy(state).cuda(), 64)
    ws = prop._workspace()
    assert size == 64
    assert int(ws.info[2].item()) >= 1
    ov, oa, ob = oprop._basis(state, 64)
    np.testing.assert_allclose(ws.alpha[:size].cpu().numpy(), oa, rtol=0, atol=1e-10 * np.abs(oa).max())
    np.testing.assert_allclose(ws.beta[: size - 1].cpu().numpy(), ob, rtol=1e-11)
    V = ws.V[:size].cpu().numpy().T
    assert np.abs(V - ov).max() < 1e-10
    gram = V.conj().T @ V
    assert np.abs(gram - np.eye(size)).max() < 1e-13


def test_basis_captured_in_hip_graph_replays_bitwise(production_pair):
    """The basis builder launches only kernels (no host sync, no memset), so a caller can capture it in
    a HIP graph: every replay, with eager work between replays, equals the eager basis bitwise."""
    print("case: basis in a HIP graph, m=24", flush=True)
    h, _ = production_pair
    prop = engine.AdaptiveLanczosPropagator(h, minimum_dimension=16, maximum_dimension=24, embedded_offset=4)
    x = torch.from_numpy(core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, 5)).cuda()
    y = torch.from_numpy(core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, 6)).cuda()
    m = 24
    prop._basis(x, m)
    ws = prop._workspace()
    op = h._device_op()
    lib = nat.load()

    def build(state):
        nat.check(lib.rpm_lanczos_basis(
            ctypes.c_void_p(op.handle), ctypes.c_void_p(nat.ptr(state)), m, ctypes.c_void_p(nat.ptr(ws.V)),
            ctypes.c_void_p(nat.ptr(ws.alpha)), ctypes.c_void_p(nat.ptr(ws.beta)), ctypes.c_void_p(nat.ptr(ws.info)),
            ctypes.c_void_p(nat.ptr(ws.norm)), 1e-14, ctypes.c_void_p(nat.ptr(ws.ws)), ws.ws.numel(),
            ctypes.c_void_p(nat.stream_ptr())))

    build(x)
    torch.cuda.synchronize()
    ref = (ws.alpha[:m].clone(), ws.beta[:m - 1].clone(), ws.V[:m].clone())
    side = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        build(x)
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=side):
            build(x)
    torch.cuda.synchronize()
    for _ in range(3):
        build(y)                      # eager work on the same workspace between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(ws.alpha[:m], ref[0])
        assert torch.equal(ws.beta[:m - 1], ref[1])
        assert torch.equal(ws.V[:m], ref[2])


def test_breakdown_on_invariant_subspace():
    # H with a 3-dimensional invariant subspace containing the start vector -> basis of size 3
    dims = (2, 2, 2)
    rng = np.random.default_rng(2)
    a = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
    h = engine.MatrixFreeHamiltonian(dims, [engine.LocalTerm((2,), a + a.conj().T)])
    state = np.zeros(8, complex)
    state[0] = 1.0
    prop = engine.AdaptiveLanczosPropagator(h, minimum_dimension=5, maximum_dimension=6, embedded_offset=2)
    size = prop._basis(torch.from_numpy(state).cuda(), 5)
    assert size == 2


def test_candidate_errors_match_oracle():
    h, o = _small_pair(4)
    state = core.random_phase_initial_state(h.dimensions, 1, 1)
    prop = engine.AdaptiveLanczosPropagator(h, minimum_dimension=24, maximum_dimension=64)
    oprop = oengine.AdaptiveLanczosPropagator(o, minimum_dimension=24, maximum_dimension=64)
    size = prop._basis(torch.from_numpy(state).cuda(), 64)
    _, oa, ob = oprop._basis(state, 64)
    ladder = [24, 32, 40, 48, 56, 64]
    for t in (0.01, 0.05, 0.2):
        errs = prop._candidates(ladder, 4, t)
        ws = prop._workspace()
        for i, d in enumerate(ladder):
            c, e = oprop._candidate(oa, ob, d, t, 1.0)
            assert abs(errs[i] - e) <= 1e-12 + 1e-6 * e, (t, d, errs[i], e)
            np.testing.assert_allclose(ws.coeffs[i, :d].cpu().numpy(), c, atol=1e-12)


def _propagators(h, o, **kw):
    return engine.AdaptiveLanczosPropagator(h, **kw), oengine.AdaptiveLanczosPropagator(o, **kw)


def test_propagate_matches_oracle_schedule_and_values():
    h, o = _small_pair(5)
    state = core.random_phase_initial_state(h.dimensions, 11, 2)
    times = np.linspace(0.0, 0.4, 321)
    kw = dict(minimum_dimension=8, maximum_dimension=16, dimension_increment=4, embedded_offset=2,
              state_error_budget=1e-7, initial_step_us=0.02, maximum_step_us=0.1)
    p, op = _propagators(h, o, **kw)
    a = p.propagate(state, times, engine.SingletPopulationObservable(h.dimensions))
    b = op.propagate(state, times, oengine.SingletPopulationObservable(o.dimensions))
    assert [(r[1], r[2], r[3]) for r in a.schedule] == [(r[1], r[2], r[3]) for r in b.schedule]
    np.testing.assert_allclose([r[0] for r in a.schedule], [r[0] for r in b.schedule], rtol=1e-12)
    np.testing.assert_allclose(a.observable_values, b.observable_values, atol=1e-10)
    np.testing.assert_allclose(a.final_state, b.final_state, atol=1e-10)
    np.testing.assert_allclose(a.state_error_estimates, b.state_error_estimates, rtol=1e-6, atol=1e-13)
    assert a.stats.accepted_steps == b.stats.accepted_steps and a.stats.rejected_steps == b.stats.rejected_steps
    assert a.stats.matvecs == b.stats.matvecs
    # replaying the oracle's schedule gives the same result
    c = p.propagate(state, times, engine.SingletPopulationObservable(h.dimensions), schedule=b.schedule)
    np.testing.assert_allclose(c.observable_values, b.observable_values, atol=1e-10)


def _ref_fixture():
    sx, sy, sz = model.spin_operators(0.5)
    dims = (2, 2, 2)
    terms = (engine.LocalTerm((0,), 0.7 * sx + 0.2 * sz),
             engine.LocalTerm((1, 2), 0.4 * np.kron(sy, sx) + 0.3 * np.kron(sz, sz)))
    h = engine.MatrixFreeHamiltonian(dims, terms)
    state = engine.singlet_initial_state(dims, np.random.default_rng(6))
    return h, state, odense.dense_from_terms(dims, h.terms)


def test_reference_lanczos_tests_on_gpu():
    import scipy.linalg
    # R/tests/test_statevector_baseline.py:207-234
    h, state, d = _ref_fixture()
    times = np.array([0.0, 0.03, 0.2, 0.7])
    prop = engine.AdaptiveLanczosPropagator(h, minimum_dimension=5, maximum_dimension=8, dimension_increment=1,
                                            embedded_offset=2, state_error_budget=1e-11, initial_step_us=0.4,
                                            maximum_step_us=0.4)
    res = prop.propagate(state, times, lambda v: oengine.singlet_population(v, (2, 2, 2)), store_states=True)
    exp = [oengine.singlet_population(scipy.linalg.expm(-1j * d * t) @ state, (2, 2, 2)) for t in times]
    np.testing.assert_allclose(res.observable_values, exp, atol=2e-12)
    np.testing.assert_allclose(res.final_state, scipy.linalg.expm(-1j * d * times[-1]) @ state, atol=2e-12)
    np.testing.assert_allclose(res.states[-1], res.final_state, atol=2e-12)
    assert res.stats.maximum_norm_error < 2e-12
    # 236-252: step underflow
    prop = engine.AdaptiveLanczosPropagator(h, minimum_dimension=3, maximum_dimension=3, dimension_increment=1,
                                            embedded_offset=1, state_error_budget=1e-30, initial_step_us=1.0,
                                            maximum_step_us=1.0, minimum_step_us=0.2)
    with pytest.raises(RuntimeError):
        prop.propagate(state, [1.0], engine.SingletPopulationObservable((2, 2, 2)))
    # 254-271: rejection recovery
    prop = engine.AdaptiveLanczosPropagator(h, minimum_dimension=3, maximum_dimension=3, dimension_increment=1,
                                            embedded_offset=1, state_error_budget=1e-2, initial_step_us=1.0,
                                            maximum_step_us=1.0, minimum_step_us=1e-6)
    res = prop.propagate(state, [1.0], engine.SingletPopulationObservable((2, 2, 2)))
    assert res.stats.rejected_steps > 0 and res.stats.accepted_steps > 0


def test_maximum_basis_reused_after_growth():
    # R/tests/test_statevector_baseline.py:273-295
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


def test_small_system_yield_matches_oracle():
    """576-dimensional random system, production Krylov rules, short horizon."""
    sm = workloads.small_system(0)
    h = sm.hamiltonian()
    o = sm.hamiltonian(builder=ocore.build_hamiltonian)
    cfg = core.SweepConfig(lifetime_us=1.0, survival_tail_bound=math.exp(-0.5))
    ocfg = ocore.SweepConfig(lifetime_us=1.0, survival_tail_bound=math.exp(-0.5))
    for sid in (0, 1):
        s = core.random_phase_initial_state(h.dimensions, 0, sid)
        a = core.propagate_trajectory(h, s, cfg)
        b = ocore.run_trajectory(o, s, ocfg)
        ra, rb = a["result"], b["result"]
        assert [(x[1], x[2]) for x in ra.schedule] == [(x[1], x[2]) for x in rb.schedule]
        np.testing.assert_allclose(ra.observable_values, rb.observable_values, atol=1e-9)
        assert abs(a["yield"] - b["yield"]) < 1e-9
        assert abs(a["quadrature_error"] - b["quadrature_error"]) < 1e-9
        exact = odense.dense_singlet_yield_closed_form(o.explicit_matrix(), s, 1.0, cfg.time_grid_us[-1])
        assert abs(a["yield"] - exact) < a["quadrature_error"] + 1e-6


def test_full_system_short_horizon_matches_oracle():
    """Production operator, dim 221,184, 0.02 us horizon (a few Lanczos steps)."""
    h = core.build_hamiltonian(90.0, 0.05)
    o = ocore.build_hamiltonian(90.0, 0.05)
    times = np.arange(17) * 0.00125
    s = core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, 0)
    kw = dict(state_error_budget=2.4e-5 * times[-1] / 57.56, initial_step_us=0.05)
    p, op = _propagators(h, o, **kw)
    a = p.propagate(s, times, engine.SingletPopulationObservable(h.dimensions))
    b = op.propagate(s, times, oengine.SingletPopulationObservable(o.dimensions))
    assert [(x[1], x[2], x[3]) for x in a.schedule] == [(x[1], x[2], x[3]) for x in b.schedule]
    np.testing.assert_allclose(a.observable_values, b.observable_values, atol=1e-10)
    assert _rel(a.final_state, b.final_state) < 1e-10


def test_random_tensor_full_system_propagation_matches_oracle():
    """The configuration-1 operator family (14 nuclei with random hyperfine and quadrupole tensors, dim
    221,184) through the adaptive propagator: same schedule and populations as the oracle."""
    print("case: random-tensor full system, 0.02 us", flush=True)
    sysr = workloads.full_random_system(seed=1)
    h = sysr.hamiltonian(theta_deg=35.0, phi_deg=10.0)
    o = sysr.hamiltonian(theta_deg=35.0, phi_deg=10.0, builder=ocore.build_hamiltonian)
    times = np.arange(17) * 0.00125
    s = core.random_phase_initial_state(h.dimensions, 0, 4)
    kw = dict(state_error_budget=1e-10, initial_step_us=0.01, maximum_dimension=40)
    p, op = _propagators(h, o, **kw)
    a = p.propagate(s, times, engine.SingletPopulationObservable(h.dimensions))
    b = op.propagate(s, times, oengine.SingletPopulationObservable(o.dimensions))
    assert [(x[1], x[2], x[3]) for x in a.schedule] == [(x[1], x[2], x[3]) for x in b.schedule]
    np.testing.assert_allclose(a.observable_values, b.observable_values, atol=1e-10)
    assert _rel(a.final_state, b.final_state) < 1e-10


@pytest.mark.parametrize("kernel", [2, 3, 4])
@pytest.mark.parametrize("batch", [1, 3, 8, 16])
def test_tiled_kernel_matches_gather_kernel_and_oracle(production_pair, batch, kernel):
    h, o = production_pair
    dev = h._device_op()
    assert dev.kernel in (2, 3), "production operator should use a tiled pair kernel"
    n = h.dimension
    v = workloads.normal_vectors(n, batch, seed=10 + batch)
    x = torch.from_numpy(v).cuda()
    y_tile = torch.empty_like(x)
    y_gather = torch.empty_like(x)
    try:
        dev.set_kernel(kernel)
        assert dev.kernel == kernel
        dev.apply(x, y_tile, batch=batch, ld=n)
        dev.set_kernel(1)
        assert dev.kernel == 1
        dev.apply(x, y_gather, batch=batch, ld=n)
    finally:
        dev.set_kernel(0)
    yt, yg = y_tile.cpu().numpy(), y_gather.cpu().numpy()
    for b in range(batch):
        want = o.matvec(v[b]) if b < 2 else yg[b]
        assert _rel(yt[b], want) <= HPSI_TOL, b
        assert _rel(yg[b], want) <= HPSI_TOL, b


def test_tiled_kernel_mixed_zero_pattern_matches_oracle():
    """Production tensors with one open methyl proton (H8a) tilted out of the radical plane
    (A_xz, A_yz != 0): its neighbour slot reads both electron halves, so the kernel takes its generic
    form, while every other left coupling keeps its zero entries (skipped per entry)."""
    import dataclasses
    print("case: production with H8a tilted, tiled kernel vs oracle", flush=True)
    base = workloads.production_system()
    nuclei = list(base.nuclei)
    i = [n.name for n in nuclei].index("H8a")
    a = np.array(nuclei[i].tensor_mT, dtype=float)
    a[0, 2] = a[2, 0] = 0.05
    a[1, 2] = a[2, 1] = -0.03
    nuclei[i] = dataclasses.replace(nuclei[i], tensor_mT=a)
    tilted = dataclasses.replace(base, nuclei=tuple(nuclei))
    h = tilted.hamiltonian()
    o = tilted.hamiltonian(builder=ocore.build_hamiltonian)
    assert h._device_op().kernel == 3
    v = workloads.normal_vectors(h.dimension, 2, seed=12)
    for b in range(2):
        assert _rel(h.matvec(v[b]), o.matvec(v[b])) <= HPSI_TOL
    # the tilt changes H·psi: the test is not passing on the unchanged production operator
    p = workloads.production_system().hamiltonian(builder=ocore.build_hamiltonian)
    assert _rel(p.matvec(v[0]), o.matvec(v[0])) > 1e-4


def test_tiled_kernel_random_tensor_system():
    sysr = workloads.full_random_system(seed=7)
    h = sysr.hamiltonian(theta_deg=63.0, phi_deg=20.0)
    o = sysr.hamiltonian(theta_deg=63.0, phi_deg=20.0, builder=ocore.build_hamiltonian)
    assert h._device_op().kernel in (2, 3)
    v = workloads.normal_vectors(h.dimension, 1, seed=8)[0]
    assert _rel(h.matvec(v), o.matvec(v)) <= HPSI_TOL


def test_tiled_kernel_normalising_copy():
    # the fused normalisation used by the Lanczos basis: out = s*H in, copy = s*in
    h = core.build_hamiltonian(10.0, 0.05)
    o = ocore.build_hamiltonian(10.0, 0.05)
    import ctypes
    from rpm_amd import _native as nat
    v = workloads.normal_vectors(h.dimension, 1, seed=9)[0] * 3.7
    x = torch.from_numpy(v).cuda()
    y = torch.empty_like(x)
    c = torch.empty_like(x)
    prop = engine.AdaptiveLanczosPropagator(h, minimum_dimension=5, maximum_dimension=5, embedded_offset=2)
    size = prop._basis(x, 2)
    ws = prop._workspace()
    v0 = ws.V[0].cpu().numpy()
    np.testing.assert_allclose(v0, v / np.linalg.norm(v), atol=1e-15)
    assert abs(ws.alpha[0].item() - np.vdot(v0, o.matvec(v0)).real) < 1e-12
