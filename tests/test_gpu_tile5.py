"""GPU: the closed-first-axis tiled H·psi kernel (k_pair_tile5, kernel 4) against the oracle and the
default LDS-tiled kernel: plain batches, per-vector electron matrices, operators whose couplings have
no structural zeros or a mixed zero pattern, and the Lanczos epilogue inside single and batched
propagations (step schedules and populations against the oracle).

Bars as in test_gpu_parity: H·psi relative L2 error <= 1e-13; populations within 1e-10 on identical
schedules.
"""

import ctypes
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from oracle import core as ocore  # noqa: E402
from oracle import engine as oengine  # noqa: E402
from rpm_amd import _native as nat  # noqa: E402
from rpm_amd import batch, core, engine, workloads  # noqa: E402

HPSI_TOL = 1e-13
ORIENTATIONS = [(0.0, 0.0, 0.05), (60.0, 30.0, 0.025), (90.0, 0.0, 0.05), (135.0, 250.0, 0.05)]


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _apply(op, v, kernel, pairs=None):
    """H·psi of the rows of v with the given kernel; pairs: per-vector electron 4x4 matrices."""
    x = torch.from_numpy(v).cuda()
    y = torch.empty_like(x)
    try:
        op.set_kernel(kernel)
        assert op.kernel == kernel
        if pairs is None:
            op.apply(x, y, v.shape[0])
        else:
            dp = torch.from_numpy(pairs.reshape(v.shape[0], 16).copy()).cuda()
            nat.check(nat.load().rpm_operator_apply_pairs(
                ctypes.c_void_p(op.handle), ctypes.c_void_p(nat.ptr(x)), ctypes.c_void_p(nat.ptr(y)), v.shape[0],
                v.shape[1], ctypes.c_void_p(nat.ptr(dp)), ctypes.c_void_p(nat.stream_ptr())))
    finally:
        op.set_kernel(0)
    return y.cpu().numpy()


@pytest.mark.parametrize("theta", [0.0, 90.0, 137.0])
def test_tile5_production_matches_oracle_and_default_kernel(theta):
    print(f"case: production operator, theta {theta}, 5 vectors, kernel 4", flush=True)
    h = core.build_hamiltonian(theta, 0.05)
    o = ocore.build_hamiltonian(theta, 0.05)
    op = h._device_op()
    v = workloads.normal_vectors(h.dimension, 5, seed=31)
    y5 = _apply(op, v, 4)
    y3 = _apply(op, v, 3)
    assert op.kernel == 3            # automatic choice restored
    for b in range(2):
        assert _rel(y5[b], o.matvec(v[b])) <= HPSI_TOL, b
    for b in range(5):
        assert _rel(y5[b], y3[b]) <= HPSI_TOL, b


def test_tile5_random_tensor_and_tilted_operators():
    print("case: random-tensor system and production with H8a tilted, kernel 4", flush=True)
    sysr = workloads.full_random_system(seed=7)
    h = sysr.hamiltonian(theta_deg=63.0, phi_deg=20.0)
    o = sysr.hamiltonian(theta_deg=63.0, phi_deg=20.0, builder=ocore.build_hamiltonian)
    v = workloads.normal_vectors(h.dimension, 2, seed=8)
    y = _apply(h._device_op(), v, 4)
    for b in range(2):
        assert _rel(y[b], o.matvec(v[b])) <= HPSI_TOL, b
    # one open methyl proton tilted out of the radical plane: its slot reads both electron halves
    base = workloads.production_system()
    nuclei = list(base.nuclei)
    i = [n.name for n in nuclei].index("H8a")
    a = np.array(nuclei[i].tensor_mT, dtype=float)
    a[0, 2] = a[2, 0] = 0.05
    a[1, 2] = a[2, 1] = -0.03
    nuclei[i] = dataclasses.replace(nuclei[i], tensor_mT=a)
    tilted = dataclasses.replace(base, nuclei=tuple(nuclei))
    ht = tilted.hamiltonian()
    ot = tilted.hamiltonian(builder=ocore.build_hamiltonian)
    y = _apply(ht._device_op(), v, 4)
    for b in range(2):
        assert _rel(y[b], ot.matvec(v[b])) <= HPSI_TOL, b


def test_tile5_per_vector_electron_matrices():
    print("case: production operator, 4 per-vector orientations, kernel 4", flush=True)
    base = core.build_hamiltonian(90.0, 0.05)
    hs = [core.build_hamiltonian(t, f, phi_deg=p) for t, p, f in ORIENTATIONS]
    pairs = batch.pair_matrices(base, hs)
    v = workloads.normal_vectors(base.dimension, len(hs), seed=21)
    y = _apply(base._device_op(), v, 4, pairs)
    for i, (t, p, f) in enumerate(ORIENTATIONS):
        o = ocore.build_hamiltonian(t, f, phi_deg=p)
        assert _rel(y[i], o.matvec(v[i])) <= HPSI_TOL, (t, p, f)


def test_tile5_propagations_match_oracle():
    print("case: production operator, kernel 4, single and batched propagations over 0.02 us", flush=True)
    base = core.build_hamiltonian(90.0, 0.05)
    hs = [core.build_hamiltonian(t, f, phi_deg=p) for t, p, f in ORIENTATIONS[:3]]
    pairs = batch.pair_matrices(base, hs)
    states = np.stack([core.random_phase_initial_state(base.dimensions, core.MASTER_SEED, i) for i in range(3)])
    times = np.arange(17) * 0.00125
    kw = dict(state_error_budget=2.4e-5 * times[-1] / 57.56, initial_step_us=0.05)
    op = base._device_op()
    try:
        op.set_kernel(4)
        res = batch.BatchedLanczosPropagator(base, **kw).propagate(states, times, pairs=pairs)
        single = engine.AdaptiveLanczosPropagator(hs[2], **kw)
        hs[2]._device_op().set_kernel(4)
        one = single.propagate(states[2], times, engine.SingletPopulationObservable(base.dimensions))
    finally:
        op.set_kernel(0)
        hs[2]._device_op().set_kernel(0)
    for i in range(3):
        o = ocore.build_hamiltonian(ORIENTATIONS[i][0], ORIENTATIONS[i][2], phi_deg=ORIENTATIONS[i][1])
        want = oengine.AdaptiveLanczosPropagator(o, **kw).propagate(states[i], times,
                                                                    oengine.SingletPopulationObservable(o.dimensions))
        assert [(x[1], x[2], x[3]) for x in res[i].schedule] == [(x[1], x[2], x[3]) for x in want.schedule]
        np.testing.assert_allclose(res[i].observable_values, want.observable_values, rtol=0, atol=1e-10)
        if i == 2:
            assert [(x[1], x[2], x[3]) for x in one.schedule] == [(x[1], x[2], x[3]) for x in want.schedule]
            np.testing.assert_allclose(one.observable_values, want.observable_values, rtol=0, atol=1e-10)
