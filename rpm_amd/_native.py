"""ctypes binding of librpm_amd.so and the device-memory plumbing (torch/ROCm).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the shared object is missing, and every device entry point raises when
no ROCm GPU is visible.
"""

from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "_lib", "librpm_amd.so")

RPM_E_INVALID = -1
RPM_E_HIP = -2
RPM_E_NOMEM = -3
RPM_E_UNSUPPORTED = -4
RPM_LANCZOS_PAIRED = 1          # rpm_lanczos_basis_batch_prefix: two steps per read of the basis
RPM_LANCZOS_FORCE_SHORT = 2     # test hook: every pair stops its call early, as at a near-breakdown

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_double_p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_ll = ctypes.c_longlong

# name -> (restype, argtypes); the C declarations are in include/rpm_amd.h
SIGNATURES = {
    "rpm_last_error": (ctypes.c_int, [ctypes.c_char_p, _sz]),
    "rpm_version": (ctypes.c_char_p, []),
    "rpm_operator_create_pair": (ctypes.c_int, [ctypes.c_int, _c_int_p, ctypes.c_int, _c_int_p, _c_double_p,
                                                _c_double_p, ctypes.POINTER(_vp)]),
    "rpm_operator_create_local": (ctypes.c_int, [ctypes.c_int, _c_int_p, ctypes.c_int, _c_int_p, _c_int_p,
                                                 _c_double_p, ctypes.POINTER(_vp)]),
    "rpm_operator_destroy": (ctypes.c_int, [_vp]),
    "rpm_operator_dimension": (_ll, [_vp]),
    "rpm_operator_apply": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _ll, _vp]),
    "rpm_operator_set_kernel": (ctypes.c_int, [_vp, ctypes.c_int]),
    "rpm_operator_kernel": (ctypes.c_int, [_vp]),
    "rpm_lanczos_workspace_bytes": (_sz, [_vp, ctypes.c_int]),
    "rpm_lanczos_basis": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, ctypes.c_double, _vp,
                                         _sz, _vp]),
    "rpm_krylov_candidates": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, _c_int_p, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_int, _vp, _vp, _vp, _vp]),
    "rpm_krylov_scratch_bytes": (_sz, [ctypes.c_int, ctypes.c_int]),
    "rpm_singlet_gram": (ctypes.c_int, [_vp, ctypes.c_int, _ll, _vp, _vp, _sz, _vp]),
    "rpm_singlet_gram_workspace_bytes": (_sz, [ctypes.c_int, _ll]),
    "rpm_krylov_outputs": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp,
                                          ctypes.c_double, _vp, _vp, _vp, _vp, _sz, _vp]),
    "rpm_krylov_combine": (ctypes.c_int, [_vp, ctypes.c_int, _ll, _vp, _vp, _vp, _vp]),
    "rpm_norm": (ctypes.c_int, [_vp, _ll, ctypes.c_int, _ll, _vp, _vp, _sz, _vp]),
    "rpm_singlet_population": (ctypes.c_int, [_vp, _ll, ctypes.c_int, _ll, _vp, _vp, _sz, _vp]),
    "rpm_reduce_workspace_bytes": (_sz, [_ll, ctypes.c_int]),
    "rpm_operator_apply_pairs": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _ll, _vp, _vp]),
    "rpm_lanczos_batch_workspace_bytes": (_sz, [_vp, ctypes.c_int, ctypes.c_int]),
    "rpm_lanczos_basis_batch": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _ll, _vp, ctypes.c_int, _vp, _vp,
                                               ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_double, _vp, _sz, _vp]),
    "rpm_krylov_tasks": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, _vp, _ll,
                                        ctypes.c_double, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "rpm_singlet_gram_batch_workspace_bytes": (_sz, [_ll, ctypes.c_int]),
    "rpm_singlet_gram_batch": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int, _ll, _vp, _vp, _sz, _vp]),
    "rpm_krylov_combine_batch": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int, _ll, _vp, ctypes.c_int,
                                                _vp, _vp, _ll, _vp]),
    "rpm_operator_apply_affine": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _ll, _vp, ctypes.c_double,
                                                 ctypes.c_double, _vp, ctypes.c_double, ctypes.c_double, _vp,
                                                 ctypes.c_double, _vp]),
    "rpm_cross_density_workspace_bytes": (_sz, [_ll, ctypes.c_int]),
    "rpm_cross_density": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _ll, _ll, ctypes.c_double, _vp, _vp, _sz, _vp]),
    "rpm_axis_apply": (ctypes.c_int, [ctypes.c_int, _c_int_p, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, _ll,
                                      _vp]),
    "rpm_eigh_workspace_bytes": (_sz, [ctypes.c_int, ctypes.c_int]),
    "rpm_eigh_batched": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _sz, ctypes.c_int, _vp]),
    "rpm_spectral_yields": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, _ll, ctypes.c_int, ctypes.c_int, _vp,
                                           _vp]),
    "rpm_bin_lines": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _ll,
                                     _vp, _ll, ctypes.c_int, _vp, _vp]),
    "rpm_adjoint_workspace_bytes": (_sz, [_vp, ctypes.c_int, ctypes.c_int]),
    "rpm_lanczos_basis_noalloc": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, ctypes.c_double,
                                                 _vp]),
    "rpm_yield_accumulate": (ctypes.c_int, [_vp, _vp, _ll, _vp, ctypes.c_int, _vp, _ll, _vp, _vp, _vp]),
    "rpm_lanczos_basis_batch_prefix": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _ll, _vp, ctypes.c_int, _vp,
                                                      ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, _vp,
                                                      _vp, _vp, ctypes.c_double, ctypes.c_int, _vp, _sz, _vp]),
    "rpm_lanczos_prefix_workspace_bytes": (_sz, [_vp, ctypes.c_int, ctypes.c_int]),
    "rpm_adjoint_sweep": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int, _c_int_p, ctypes.c_int, _c_double_p, _vp, _c_int_p,
                                         _c_double_p, _vp, _c_int_p, ctypes.c_int, _vp, _c_int_p, ctypes.c_int,
                                         _vp, _sz, _vp]),
}

_lib = None
_lock = threading.Lock()


def load():
    """Load librpm_amd.so (built by ``python -m rpm_amd.build``); raises if absent."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIBRARY_PATH):
            raise ImportError(
                f"{LIBRARY_PATH} is missing: build the HIP library first (python -m rpm_amd.build)")
        lib = ctypes.CDLL(LIBRARY_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


class NativeError(RuntimeError):
    pass


def check(rc: int) -> None:
    if rc == 0:
        return
    buf = ctypes.create_string_buffer(1024)
    load().rpm_last_error(buf, len(buf))
    msg = buf.value.decode(errors="replace")
    if rc == RPM_E_INVALID:
        raise ValueError(msg)
    raise NativeError(f"rpm_amd error {rc}: {msg}")


# ---------------------------------------------------------------- device plumbing

def torch_mod():
    import torch
    return torch


def require_gpu():
    torch = torch_mod()
    if not torch.cuda.is_available():
        raise RuntimeError("rpm_amd needs a ROCm GPU (MI355X); none is visible to torch")
    load()
    return torch


def device():
    return require_gpu().device("cuda", torch_mod().cuda.current_device())


def stream_ptr() -> int:
    return require_gpu().cuda.current_stream().cuda_stream


def ptr(t) -> int:
    return t.data_ptr()


def empty_c128(n, *shape):
    torch = require_gpu()
    dims = (n,) + tuple(shape) if shape else (n,)
    return torch.empty(dims, dtype=torch.complex128, device=device())


def to_device_c128(x):
    """numpy / torch -> contiguous complex128 device tensor (1-D view kept)."""
    torch = require_gpu()
    if isinstance(x, torch.Tensor):
        t = x.to(device=device(), dtype=torch.complex128)
        return t.contiguous()
    arr = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
    return torch.from_numpy(arr).to(device())


def workspace(nbytes: int):
    torch = require_gpu()
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device())


def is_torch(x) -> bool:
    try:
        import torch
    except ImportError:  # pragma: no cover
        return False
    return isinstance(x, torch.Tensor)
