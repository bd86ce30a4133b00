"""GPU: rpm_lanczos_basis_noalloc (the basis builder on the library's static scratch) against
rpm_lanczos_basis (the same builder on a caller workspace), through the C ABI.

Two calls in a row, the second with a smaller input (the 576-level system after the 221,184-level
one), each bitwise equal to the workspace form: basis, tridiagonal, size/info and norm.  A basis
whose workspace exceeds the scratch is rejected on the host.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from rpm_amd import _native as nat  # noqa: E402
from rpm_amd import core, workloads  # noqa: E402

SCRATCH_BYTES = 16 << 20


def _buffers(n, target, dev):
    return dict(V=torch.empty((target, n), dtype=torch.complex128, device=dev),
                alpha=torch.empty(target, dtype=torch.float64, device=dev),
                beta=torch.empty(target, dtype=torch.float64, device=dev),
                info=torch.empty(4, dtype=torch.int32, device=dev),
                norm=torch.empty(1, dtype=torch.float64, device=dev))


def _vp(t):
    return ctypes.c_void_p(nat.ptr(t))


def _both(h, state, target):
    lib = nat.load()
    op = ctypes.c_void_p(h._device_op().handle)
    n = h.dimension
    dev = nat.device()
    x = torch.from_numpy(np.ascontiguousarray(state)).to(dev)
    ref = _buffers(n, target, dev)
    got = _buffers(n, target, dev)
    for b in (ref, got):      # stale contents must not matter
        b["V"].fill_(complex(7.0, -3.0))
        b["alpha"].fill_(5.0)
        b["beta"].fill_(5.0)
        b["info"].fill_(-9)
    ws = nat.workspace(lib.rpm_lanczos_workspace_bytes(op, target))
    assert lib.rpm_lanczos_workspace_bytes(op, target) <= SCRATCH_BYTES
    stream = ctypes.c_void_p(nat.stream_ptr())
    nat.check(lib.rpm_lanczos_basis(op, _vp(x), target, _vp(ref["V"]), _vp(ref["alpha"]), _vp(ref["beta"]),
                                    _vp(ref["info"]), _vp(ref["norm"]), 1e-14, _vp(ws), ws.numel(), stream))
    nat.check(lib.rpm_lanczos_basis_noalloc(op, _vp(x), target, _vp(got["V"]), _vp(got["alpha"]),
                                            _vp(got["beta"]), _vp(got["info"]), _vp(got["norm"]), 1e-14, stream))
    torch.cuda.synchronize()
    size = int(ref["info"][0].item())
    assert size == target
    assert torch.equal(got["info"][:3], ref["info"][:3])
    assert torch.equal(got["norm"], ref["norm"])
    assert torch.equal(got["alpha"][:size], ref["alpha"][:size])
    assert torch.equal(got["beta"][:size - 1], ref["beta"][:size - 1])
    assert torch.equal(got["V"][:size], ref["V"][:size])
    return got


def test_noalloc_basis_equals_workspace_basis_twice_then_smaller():
    print("case: production operator, target 48", flush=True)
    h = core.build_hamiltonian(90.0, 0.05)
    big = _both(h, core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, 0), 48)
    print("case: 576-level system, target 16 (second call, smaller input)", flush=True)
    sys_ = workloads.small_system(0)
    hs = sys_.hamiltonian()
    small = _both(hs, core.random_phase_initial_state(hs.dimensions, 0, 3), 16)
    # the orthonormality of what the scratch path built
    for out in (big, small):
        V = out["V"]
        g = (V.conj() @ V.T).cpu().numpy()
        assert np.abs(g - np.eye(V.shape[0])).max() < 1e-12


def test_noalloc_basis_rejects_a_workspace_above_the_scratch():
    print("case: N = 2^20 local-term operator, oversize rejection", flush=True)
    lib = nat.load()
    dims = (2,) * 20
    sz = np.array([[0.5, 0.0], [0.0, -0.5]], dtype=complex)
    ops = np.ascontiguousarray(np.stack([sz.real, sz.imag], axis=-1).reshape(-1))
    handle = ctypes.c_void_p()
    nat.check(lib.rpm_operator_create_local(len(dims), (ctypes.c_int * len(dims))(*dims), 1, (ctypes.c_int * 1)(1),
                                            (ctypes.c_int * 1)(0), ops.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                            ctypes.byref(handle)))
    try:
        n = 1 << 20
        assert lib.rpm_lanczos_workspace_bytes(handle, 8) > SCRATCH_BYTES
        dev = nat.device()
        x = torch.zeros(n, dtype=torch.complex128, device=dev)
        out = _buffers(n, 8, dev)
        rc = lib.rpm_lanczos_basis_noalloc(handle, _vp(x), 8, _vp(out["V"]), _vp(out["alpha"]), _vp(out["beta"]),
                                           _vp(out["info"]), _vp(out["norm"]), 1e-14, ctypes.c_void_p(nat.stream_ptr()))
        assert rc == nat.RPM_E_INVALID
        buf = ctypes.create_string_buffer(512)
        lib.rpm_last_error(buf, len(buf))
        assert b"scratch" in buf.value
    finally:
        lib.rpm_operator_destroy(handle)
