"""GPU: Lanczos bases built in prefixes (rpm_lanczos_basis_batch_prefix) and the propagators that use them.

* A basis built in pieces (24, then 40, then 64 vectors; with one trajectory starting a new basis
  while another continues) is bitwise the basis built in one call (R/rpm_statevector_engine.py:651-688).
* With paired steps (two steps per read of the basis) the pieces, including odd windows that make a
  trajectory take single steps inside a pair slot, agree with the one-call basis to round-off
  (1e-12), and a trajectory's basis is bitwise the same whatever else the batch holds.
* The propagators keep the reference's schedules and values (the oracle comparisons in
  test_gpu_parity.py run through the same code) while building fewer vectors: the device matvec count
  is below the reference's count whenever a step selects a dimension below the maximum.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from rpm_amd import _native as nat  # noqa: E402
from rpm_amd import batch, core, engine  # noqa: E402


def _bufs(B, m, n, dev):
    return dict(V=torch.zeros((B, m, n), dtype=torch.complex128, device=dev),
                alpha=torch.zeros((B, m), dtype=torch.float64, device=dev),
                beta=torch.zeros((B, m), dtype=torch.float64, device=dev),
                info=torch.zeros((B, 4), dtype=torch.int32, device=dev),
                norm=torch.zeros(B, dtype=torch.float64, device=dev))


def test_prefix_basis_equals_whole_basis_bitwise():
    print("case: production operator, 3 trajectories, prefixes 24/40/64", flush=True)
    h = core.build_hamiltonian(90.0, 0.05)
    lib = nat.load()
    op = ctypes.c_void_p(h._device_op().handle)
    n, B, m = h.dimension, 3, 64
    dev = nat.device()
    states = torch.from_numpy(np.stack([core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, i)
                                        for i in range(B)])).to(dev)
    pairs = torch.from_numpy(np.ascontiguousarray(
        batch.pair_matrices(h, [core.build_hamiltonian(t, 0.05) for t in (0.0, 45.0, 90.0)]).reshape(B, 16))).to(dev)
    vp = ctypes.c_void_p
    sp = vp(nat.stream_ptr())

    whole = _bufs(B, m, n, dev)
    w1 = nat.workspace(lib.rpm_lanczos_batch_workspace_bytes(op, m, B))
    tg = torch.tensor([64, 64, 40], dtype=torch.int32, device=dev)
    nat.check(lib.rpm_lanczos_basis_batch(op, B, vp(nat.ptr(states)), n, vp(nat.ptr(tg)), m, vp(nat.ptr(pairs)),
                                          vp(nat.ptr(whole["V"])), m, vp(nat.ptr(whole["alpha"])),
                                          vp(nat.ptr(whole["beta"])), vp(nat.ptr(whole["info"])),
                                          vp(nat.ptr(whole["norm"])), 1e-14, vp(nat.ptr(w1)), w1.numel(), sp))
    part = _bufs(B, m, n, dev)
    w2 = nat.workspace(lib.rpm_lanczos_prefix_workspace_bytes(op, m, B))

    def call(targets, stops, j0):
        t = torch.tensor(list(targets) + list(stops), dtype=torch.int32, device=dev)
        nat.check(lib.rpm_lanczos_basis_batch_prefix(
            op, B, vp(nat.ptr(states)), n, vp(nat.ptr(t)), m, vp(nat.ptr(t) + 4 * B), j0, max(stops),
            vp(nat.ptr(pairs)), vp(nat.ptr(part["V"])), m, vp(nat.ptr(part["alpha"])), vp(nat.ptr(part["beta"])),
            vp(nat.ptr(part["info"])), vp(nat.ptr(part["norm"])), 1e-14, 0, vp(nat.ptr(w2)), w2.numel(), sp))
        torch.cuda.synchronize()
        return part["info"][:, 0].cpu().numpy().copy()

    # trajectories 0 and 1 start now (prefix 24); trajectory 2 sits out
    assert list(call([64, 64, 0], [24, 24, 0], 0)[:2]) == [24, 24]
    # 0 continues to 40; 1 sits out; 2 starts a basis of 40 whole
    got = call([0, 0, 40], [40, 0, 40], 0)
    assert list(got) == [40, 24, 40]
    # 0 and 1 continue to 64 (1 from 24)
    assert list(call([0, 0, 0], [64, 64, 0], 24)) == [64, 64, 40]
    for b, k in enumerate((64, 64, 40)):
        assert torch.equal(part["V"][b, :k], whole["V"][b, :k]), b
        assert torch.equal(part["alpha"][b, :k], whole["alpha"][b, :k]), b
        assert torch.equal(part["beta"][b, :k - 1], whole["beta"][b, :k - 1]), b
        assert torch.equal(part["norm"][b], whole["norm"][b])
    assert torch.equal(part["info"][:, :3], whole["info"][:, :3])


class _Prefix:
    """Paired prefix calls on B trajectories of the production operator (orientations per trajectory)."""

    def __init__(self, h, thetas, seeds, m=64):
        self.lib = nat.load()
        self.op = ctypes.c_void_p(h._device_op().handle)
        self.n, self.B, self.m = h.dimension, len(thetas), m
        self.dev = nat.device()
        self.states = torch.from_numpy(np.stack([core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, s)
                                                 for s in seeds])).to(self.dev)
        self.pairs = torch.from_numpy(np.ascontiguousarray(
            batch.pair_matrices(h, [core.build_hamiltonian(t, 0.05) for t in thetas]).reshape(self.B, 16))).to(self.dev)
        self.buf = _bufs(self.B, m, self.n, self.dev)
        self.ws = nat.workspace(self.lib.rpm_lanczos_prefix_workspace_bytes(self.op, m, self.B))

    def call(self, targets, stops, j0, flags=nat.RPM_LANCZOS_PAIRED):
        vp = ctypes.c_void_p
        t = torch.tensor(list(targets) + list(stops), dtype=torch.int32, device=self.dev)
        b = self.buf
        nat.check(self.lib.rpm_lanczos_basis_batch_prefix(
            self.op, self.B, vp(nat.ptr(self.states)), self.n, vp(nat.ptr(t)), self.m, vp(nat.ptr(t) + 4 * self.B),
            j0, max(stops), vp(nat.ptr(self.pairs)), vp(nat.ptr(b["V"])), self.m, vp(nat.ptr(b["alpha"])),
            vp(nat.ptr(b["beta"])), vp(nat.ptr(b["info"])), vp(nat.ptr(b["norm"])), 1e-14, flags,
            vp(nat.ptr(self.ws)), self.ws.numel(), vp(nat.stream_ptr())))
        torch.cuda.synchronize()
        return b["info"].cpu().numpy().copy()


def test_paired_prefix_pieces_match_the_whole_basis():
    print("case: production operator, 3 trajectories, paired windows 13/27/40/64 and 24/64 vs one unpaired call",
          flush=True)
    h = core.build_hamiltonian(90.0, 0.05)
    thetas = (0.0, 45.0, 90.0)
    ref = _Prefix(h, thetas, range(3))
    assert list(ref.call([64] * 3, [64] * 3, 0, flags=0)[:, 0]) == [64] * 3
    got = _Prefix(h, thetas, range(3))
    # trajectory 0: odd windows (pairs from step 20: a single step at 26 after two pending vectors,
    # then one at 27 in the next call's first slot); trajectory 1: 24 then 64 (the pair that ends at
    # 24 leaves its last vector materialised); trajectory 2: whole in the second call
    assert list(got.call([64, 64, 0], [13, 24, 0], 0)[:2, 0]) == [13, 24]
    assert list(got.call([0, 0, 64], [27, 0, 64], 0)[:, 0]) == [27, 24, 64]
    assert list(got.call([0, 0, 0], [40, 64, 0], 24)[:2, 0]) == [40, 64]
    info = got.call([0, 0, 0], [64, 0, 0], 40)
    assert list(info[:, 0]) == [64] * 3 and list(info[:, 1]) == [1] * 3 and not info[:, 3].any()
    V0, V1 = ref.buf["V"], got.buf["V"]
    for b in range(3):
        dv = (V1[b] - V0[b]).abs().max().item()
        da = (got.buf["alpha"][b] - ref.buf["alpha"][b]).abs().max().item()
        db = (got.buf["beta"][b, :63] - ref.buf["beta"][b, :63]).abs().max().item()
        scale = ref.buf["alpha"][b].abs().max().item()
        print(f"b={b} |dV|={dv:.2e} |dalpha|/|T|={da / scale:.2e} |dbeta|/|T|={db / scale:.2e}", flush=True)
        assert dv < 1e-12 and da < 1e-12 * scale and db < 1e-12 * scale
        gram = V1[b].conj() @ V1[b].T
        assert (gram - torch.eye(64, dtype=gram.dtype, device=gram.device)).abs().max().item() < 1e-12


def test_paired_basis_of_a_trajectory_does_not_depend_on_the_batch():
    print("case: production operator, paired windows, batch of 3 against each trajectory alone", flush=True)
    h = core.build_hamiltonian(90.0, 0.05)
    thetas, seeds = (0.0, 45.0, 90.0), (3, 4, 5)
    calls = [([64, 64, 0], [13, 24, 0], 0), ([0, 0, 64], [27, 0, 64], 0), ([0, 0, 0], [40, 64, 0], 24),
             ([0, 0, 0], [64, 0, 0], 40)]
    full = _Prefix(h, thetas, seeds)
    for tg, st, j0 in calls:
        full.call(tg, st, j0)
    for b in range(3):
        one = _Prefix(h, thetas[b:b + 1], seeds[b:b + 1])
        for tg, st, j0 in calls:
            if tg[b] > 0 or st[b] > 0:
                one.call([tg[b]], [st[b]], j0 if tg[b] == 0 else 0)
        for key in ("V", "alpha", "beta", "norm"):
            assert torch.equal(one.buf[key][0], full.buf[key][b]), (b, key)


def test_forced_early_stops_continue_to_the_same_basis():
    """RPM_LANCZOS_FORCE_SHORT makes every pair stop its trajectory's call at its first step, as a
    near-breakdown inside a pair does; calls continued until no trajectory is short give the
    one-call basis to round-off."""
    print("case: production operator, 3 trajectories, every pair stopped early, continued", flush=True)
    h = core.build_hamiltonian(90.0, 0.05)
    thetas = (0.0, 45.0, 90.0)
    ref = _Prefix(h, thetas, range(3))
    ref.call([64] * 3, [64] * 3, 0, flags=0)
    got = _Prefix(h, thetas, range(3))
    flags = nat.RPM_LANCZOS_PAIRED | nat.RPM_LANCZOS_FORCE_SHORT
    info = got.call([64, 64, 64], [64, 40, 64], 0, flags=flags)
    calls, shorts = 1, int(info[:, 3].sum())
    while info[:, 3].any():
        stops = [64 if info[b, 3] and b != 1 else (40 if info[b, 3] else 0) for b in range(3)]
        j0 = int(min(info[b, 0] for b in range(3) if info[b, 3]))
        info = got.call([0, 0, 0], stops, j0, flags=flags)
        calls += 1
        shorts += int(info[:, 3].sum())
    info = got.call([0, 0, 0], [0, 64, 0], 40, flags=flags)       # trajectory 1 continues past its prefix
    while info[:, 3].any():
        info = got.call([0, 0, 0], [0, 64, 0], int(info[1, 0]), flags=flags)
        calls += 1
    print("calls", calls, "early stops", shorts, flush=True)
    assert shorts >= 10 and list(got.buf["info"][:, 0].cpu().numpy()) == [64] * 3
    for b in range(3):
        scale = ref.buf["alpha"][b].abs().max().item()
        assert (got.buf["V"][b] - ref.buf["V"][b]).abs().max().item() < 1e-12
        assert (got.buf["alpha"][b] - ref.buf["alpha"][b]).abs().max().item() < 1e-12 * scale
        assert (got.buf["beta"][b, :63] - ref.buf["beta"][b, :63]).abs().max().item() < 1e-12 * scale


def test_propagators_continue_bases_stopped_early():
    """Both propagators continue a basis a call stopped early (forced at every pair) and keep the
    schedule and values of the default run."""
    print("case: production operator, 0.1 us, single and batched, every pair stopped early", flush=True)
    h = core.build_hamiltonian(90.0, 0.05)
    times = np.arange(81) * 0.00125
    states = np.stack([core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, i) for i in range(2)])
    kw = dict(state_error_budget=2.4e-5 * 0.1 / 57.56)
    obs = engine.SingletPopulationObservable(h.dimensions)
    ref = engine.AdaptiveLanczosPropagator(h, **kw).propagate(states[0], times, obs)
    single = engine.AdaptiveLanczosPropagator(h, **kw)
    single._basis_test_flags = nat.RPM_LANCZOS_FORCE_SHORT
    got = single.propagate(states[0], times, obs)
    assert got.schedule == ref.schedule
    np.testing.assert_allclose(got.observable_values, ref.observable_values, rtol=0, atol=1e-12)
    bp = batch.BatchedLanczosPropagator(h, **kw)
    bp._basis_test_flags = nat.RPM_LANCZOS_FORCE_SHORT
    res = bp.propagate(states, times)
    assert res[0].schedule == ref.schedule
    np.testing.assert_allclose(res[0].observable_values, ref.observable_values, rtol=0, atol=1e-12)


def test_propagators_build_fewer_vectors_with_the_same_schedule():
    print("case: production operator, 0.3 us, single and batched", flush=True)
    h = core.build_hamiltonian(90.0, 0.05)
    times = np.arange(241) * 0.00125
    states = np.stack([core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, i) for i in range(2)])
    kw = dict(state_error_budget=2.4e-5 * 0.3 / 57.56)
    obs = engine.SingletPopulationObservable(h.dimensions)
    single = engine.AdaptiveLanczosPropagator(h, **kw)
    one = single.propagate(states[0], times, obs)
    bp = batch.BatchedLanczosPropagator(h, **kw)
    res = bp.propagate(states, times)
    assert res[0].schedule == one.schedule
    assert res[0].stats == one.stats
    np.testing.assert_allclose(res[0].observable_values, one.observable_values, rtol=0, atol=1e-13)
    sel = [r[2] for r in one.schedule if r[1]]
    print("selected", sel, "reference matvecs", one.stats.matvecs, "device", single.device_matvecs, flush=True)
    assert single.device_matvecs == bp.device_matvecs[0]
    assert single.device_matvecs <= one.stats.matvecs
    # the reference's count: every prefer-max step builds the maximum basis
    assert one.stats.matvecs == sum(sum(r[3]) for r in one.schedule)
    if min(sel[1:]) < 64:
        assert single.device_matvecs < one.stats.matvecs
