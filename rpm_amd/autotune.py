"""Runtime tuning of the shape choices that depend on the device.

Two choices are tuned by timing real launches on the current stream:
- ``hpsi_kernel``: which H·psi kernel an operator uses for a batch size (the per-amplitude gather
  kernel, the register-tiled pair kernel, or one of the two LDS-tiled pair kernels);
- ``propagation_batch``: how many trajectories a batched propagation runs per launch, timed on a
  short prefix of the real workload and bounded by device memory.

Results are cached per process, keyed by device name and shape (a lock guards the cache). Neither
choice changes results beyond rounding: every H·psi kernel forms the same sums, and a batch only
groups independent trajectories.
"""

from __future__ import annotations

import threading
import time

import numpy as np

from . import _native as nat

HPSI_KERNELS = {1: "gather", 2: "register-tiled", 3: "LDS-tiled", 4: "LDS-tiled, closed first axis"}
_CACHE: dict = {}
_LOCK = threading.Lock()


def _bucket(batch: int) -> int:
    """Batch sizes that share a tuning result: powers of two."""
    return 1 << max(0, int(batch) - 1).bit_length()


def clear_cache() -> None:
    with _LOCK:
        _CACHE.clear()


def cached(kind: str) -> list:
    """The cached results of one tuner ("hpsi" or "batch"), oldest first."""
    with _LOCK:
        return [v for k, v in _CACHE.items() if k[0] == kind]


def hpsi_kernel(hamiltonian, batch: int, *, reps: int = 5, apply: bool = True) -> dict:
    """Time every H·psi kernel the operator supports on `batch` resident vectors (HIP events) and pick
    the fastest. Returns {"kernel", "name", "ms_per_launch": {kernel: ms}}; with `apply` the operator
    uses the chosen kernel from then on. A generic local-term operator has one kernel and is returned
    as is."""
    torch = nat.require_gpu()
    op = hamiltonian._device_op()
    if op.kernel not in HPSI_KERNELS:
        return {"kernel": 0, "name": "local-term", "ms_per_launch": {}}
    batch = max(1, int(batch))
    key = ("hpsi", torch.cuda.get_device_name(), tuple(hamiltonian.dimensions), _bucket(batch))
    with _LOCK:
        hit = _CACHE.get(key)
    if hit is None:
        n = hamiltonian.dimension
        gen = torch.Generator(device="cuda").manual_seed(0)
        x = torch.randn((batch, n), dtype=torch.complex128, device="cuda", generator=gen)
        y = torch.empty_like(x)
        before = op.kernel
        times = {}
        for k in HPSI_KERNELS:
            try:
                op.set_kernel(k)
            except (ValueError, nat.NativeError):
                continue          # no such form for this operator shape
            if op.kernel != k:
                continue
            op.apply(x, y, batch)
            start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            start.record(torch.cuda.current_stream())
            for _ in range(reps):
                op.apply(x, y, batch)
            stop.record(torch.cuda.current_stream())
            torch.cuda.synchronize()
            times[k] = start.elapsed_time(stop) / reps
        op.set_kernel(before)
        if not times:
            raise RuntimeError("no H·psi kernel could be timed for this operator")
        best = min(times, key=times.get)
        hit = {"kernel": best, "name": HPSI_KERNELS[best], "ms_per_launch": times}
        with _LOCK:
            _CACHE[key] = hit
    if apply:
        op.set_kernel(hit["kernel"])
    return hit


def propagation_batch(propagator, states, times_us, pairs, *, candidates=(8, 16, 32, 64), prefix_points: int = 9,
                      memory_fraction: float = 0.5) -> dict:
    """Trajectories per batched launch for a BatchedLanczosPropagator: run the first B trajectories
    over the first `prefix_points` output times for each candidate B (after a warm-up) and pick the
    highest trajectory-microseconds per second. Candidates whose Krylov bases (B x m x N complex)
    would take more than `memory_fraction` of the free device memory are skipped. Returns
    {"batch", "trajectory_us_per_s": {B: rate}}."""
    torch = nat.require_gpu()
    times = np.asarray(tuple(float(t) for t in times_us), dtype=float)[:max(2, int(prefix_points))]
    total = states.shape[0]
    n = propagator.hamiltonian.dimension
    m = int(propagator.maximum_dimension)
    free, _ = torch.cuda.mem_get_info()
    fits = [b for b in sorted(set(int(c) for c in candidates))
            if 1 <= b <= total and b * m * n * 16 <= memory_fraction * free]
    if not fits:
        fits = [1]
    key = ("batch", torch.cuda.get_device_name(), n, m, len(times), tuple(fits))
    with _LOCK:
        hit = _CACHE.get(key)
    if hit is not None:
        return hit
    propagator.propagate(states[:fits[0]], times[:3], pairs=pairs[:fits[0]])      # warm-up
    rates = {}
    for b in fits:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        propagator.propagate(states[:b], times, pairs=pairs[:b])
        torch.cuda.synchronize()
        rates[b] = b * float(times[-1]) / (time.perf_counter() - t0)
    best = max(rates, key=rates.get)
    hit = {"batch": best, "trajectory_us_per_s": rates}
    with _LOCK:
        _CACHE[key] = hit
    return hit
