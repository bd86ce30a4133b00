"""Batched trajectories: many random-phase states x field orientations/strengths.

Each trajectory follows exactly the step rules of
``AdaptiveLanczosPropagator.propagate`` (R/rpm_statevector_engine.py:717-933,
singlet-projected path 735-741, 861-902): the rules run as one small host
coroutine per trajectory, and every device operation they ask for (Lanczos
basis, candidate errors, projected outputs, endpoint update) is served for
all trajectories that want it in one batched launch:

* ``rpm_lanczos_basis_batch`` -- fully reorthogonalised Lanczos bases of all
  trajectories that need one,
  each with its own target size, breakdown flag and electron 4x4 (so one
  launch mixes field orientations and strengths);
* ``rpm_krylov_tasks`` -- every candidate and output-time evaluation of the
  round in one launch;
* ``rpm_singlet_gram_batch`` / ``rpm_krylov_combine_batch`` -- projections and
  endpoint updates.

A round serves the phases in the order basis -> candidates -> outputs ->
update, so one host synchronisation per phase covers the whole batch.  The
per-trajectory arithmetic is that of the single-trajectory path (same kernels,
same reduction order), so a batched trajectory reproduces its plain run.
"""

from __future__ import annotations

import ctypes
import math
from typing import Callable, Iterable, Sequence

import numpy as np

from . import _native as nat
from .engine import (
    BasisGraphs, CompiledFullPairHamiltonian, PropagationResult, PropagationStats, YieldAccumulator, _pair_form,
    state_error_to_population_bound,
)

_TASK = np.dtype([("b", "<i4"), ("dim", "<i4"), ("low", "<i4"), ("out", "<i4"), ("t", "<f8"), ("kind", "<i4"),
                  ("pad", "<i4")])
assert _TASK.itemsize == 32

_BASIS, _CAND, _OUT, _ADV, _INIT = "basis", "cand", "out", "adv", "init"
_PHASES = (_INIT, _BASIS, _CAND, _OUT, _ADV)


def electron_pair_matrix(hamiltonian) -> np.ndarray:
    """The 4x4 electron matrix (Zeeman + electron coupling) of a pair-form Hamiltonian, row = eA*2 + eB."""
    form = _pair_form(hamiltonian.dimensions, hamiltonian.terms,
                      getattr(hamiltonian, "left_axes", None), getattr(hamiltonian, "right_axes", None))
    if form is None:
        raise ValueError("The Hamiltonian has no pair-stencil form.")
    return form[3]


def pair_matrices(base, hamiltonians: Sequence) -> np.ndarray:
    """Electron 4x4 of each Hamiltonian, checking that only that part differs from ``base``.

    The batched operator shares the nuclear blocks of ``base``; a Hamiltonian
    whose nuclear part differs cannot ride on it and raises ValueError.
    """
    ref = _pair_form(base.dimensions, base.terms, base.left_axes, base.right_axes)
    out = np.empty((len(hamiltonians), 4, 4), dtype=complex)
    for i, h in enumerate(hamiltonians):
        if tuple(h.dimensions) != tuple(base.dimensions):
            raise ValueError("All Hamiltonians of a batch must share the dimensions.")
        form = _pair_form(h.dimensions, h.terms, h.left_axes, h.right_axes)
        if form is None or len(form[2]) != len(ref[2]) or any(
                not np.array_equal(a, b) for a, b in zip(form[2], ref[2])):
            raise ValueError("Batched Hamiltonians may differ only in their electron terms (Zeeman, coupling).")
        out[i] = form[3]
    return out


def _trajectory_times(times_us, batch: int) -> list:
    """One strictly increasing, non-negative time array per trajectory, from a shared 1-D sequence or
    from `batch` sequences (which may differ in length)."""
    try:
        arr = np.asarray(times_us, dtype=float)
    except (TypeError, ValueError):        # ragged: one sequence per trajectory
        arr = None
    if arr is not None and arr.ndim == 1:
        lists = [arr] * batch
    elif arr is not None and arr.ndim == 2:
        lists = [arr[b] for b in range(arr.shape[0])]
    else:
        lists = [np.asarray(tuple(float(t) for t in x), dtype=float) for x in times_us]
    if len(lists) != batch:
        raise ValueError("times_us must be one sequence or one sequence per trajectory")
    for t in lists:
        if t.ndim != 1 or t.size == 0:
            raise ValueError("At least one observation time is required.")
        if np.any(t < 0.0) or np.any(np.diff(t) <= 0.0):
            raise ValueError("Observation times must be nonnegative and strictly increasing.")
    return lists


class BatchedLanczosPropagator:
    """Adaptive Lanczos propagation of a batch of trajectories with the singlet observable.

    Parameters as ``AdaptiveLanczosPropagator`` (R/rpm_statevector_engine.py:620-649).
    ``hamiltonian`` must have the pair-stencil form; per-trajectory field
    orientations/strengths enter as electron 4x4 matrices (``pairs``).
    """

    def __init__(self, hamiltonian: CompiledFullPairHamiltonian, *, minimum_dimension: int = 24,
                 maximum_dimension: int = 64, dimension_increment: int = 8, embedded_offset: int = 4,
                 safety_factor: float = 2.0, state_error_budget: float = 2.4e-5, initial_step_us: float = 0.05,
                 maximum_step_us: float = 0.25, minimum_step_us: float = 1e-9,
                 breakdown_tolerance: float = 1e-14) -> None:
        if minimum_dimension <= embedded_offset:
            raise ValueError("minimum_dimension must exceed embedded_offset.")
        if maximum_dimension < minimum_dimension:
            raise ValueError("maximum_dimension must not be smaller than minimum_dimension.")
        if not getattr(hamiltonian, "uses_pair_stencil", False):
            raise ValueError("Batched propagation needs a Hamiltonian with the pair-stencil form.")
        self.hamiltonian = hamiltonian
        self.minimum_dimension = minimum_dimension
        self.maximum_dimension = maximum_dimension
        self.dimension_increment = dimension_increment
        self.embedded_offset = embedded_offset
        self.safety_factor = safety_factor
        self.state_error_budget = state_error_budget
        self.initial_step_us = initial_step_us
        self.maximum_step_us = maximum_step_us
        self.minimum_step_us = minimum_step_us
        self.breakdown_tolerance = breakdown_tolerance
        self.rounds = 0
        self.final_states = None     # (B, N) device tensor of the last propagate call
        # matvecs the device made per trajectory in the last call; stats.matvecs counts the reference's
        # whole-basis calls for the same schedule (bases are built only as far as a candidate needs)
        self.device_matvecs = None
        # replay each basis step window as a HIP graph (captured on the window's third run in a
        # propagate call); bitwise the eager launches, and as fast as them once a run has more than a
        # few dozen rounds
        self.use_graphs = True
        self.graphs = None
        # Lanczos steps from the 12th on in pairs that share one read of the basis
        self.paired_steps = True
        # extra flags of the basis builder for tests (RPM_LANCZOS_FORCE_SHORT exercises the
        # continuation of a basis that a call stopped early)
        self._basis_test_flags = 0

    @classmethod
    def like(cls, propagator, hamiltonian=None) -> "BatchedLanczosPropagator":
        """A batched propagator with the settings of a single-trajectory one."""
        return cls(propagator.hamiltonian if hamiltonian is None else hamiltonian,
                   minimum_dimension=propagator.minimum_dimension, maximum_dimension=propagator.maximum_dimension,
                   dimension_increment=propagator.dimension_increment, embedded_offset=propagator.embedded_offset,
                   safety_factor=propagator.safety_factor, state_error_budget=propagator.state_error_budget,
                   initial_step_us=propagator.initial_step_us, maximum_step_us=propagator.maximum_step_us,
                   minimum_step_us=propagator.minimum_step_us, breakdown_tolerance=propagator.breakdown_tolerance)

    def _ladder(self, start):
        c = list(range(start, self.maximum_dimension + 1, self.dimension_increment))
        if not c or c[-1] != self.maximum_dimension:
            c.append(self.maximum_dimension)
        return c

    # -- per-trajectory rules (R/rpm_statevector_engine.py:717-933) ----------

    def _rules(self, times: np.ndarray, replay, callback):
        """Coroutine: yields (phase, args) requests, receives their results, returns the result tuple."""
        nt = times.size
        values = np.empty(nt)
        bounds = np.empty(nt)
        errors = np.empty(nt)
        cursor, now, acc_err = 0, 0.0, 0.0
        accepted = rejected = max_dim = matvecs = 0
        used = []
        record = []
        max_norm_err, pop0 = yield (_INIT, None)
        if times[0] == 0.0:
            values[0] = pop0
            bounds[0] = errors[0] = 0.0
            cursor = 1
        t_final = float(times[-1])
        step = min(self.initial_step_us, t_final) if t_final else 0.0
        prefer_max = False
        mn, mx, inc, off = self.minimum_dimension, self.maximum_dimension, self.dimension_increment, \
            self.embedded_offset
        guess = mx              # basis prefix to build first in prefer-max mode: the last selected dimension
        device_mv = [0]         # matvecs the device made (the reference's count is in the stats)

        while cursor < nt:
            trial = min(step, t_final - now)
            if replay is not None:
                if len(record) >= len(replay):
                    raise RuntimeError("Replayed schedule ended before the final time.")
                rec = replay[len(record)]
                trial = float(rec[0])
            if trial < self.minimum_step_us:
                raise RuntimeError("Lanczos step fell below the configured minimum.")
            if replay is not None:
                def ok(dim, e, rec=rec):
                    return bool(rec[1]) and int(dim) == int(rec[2])
            else:
                budget_now = self.state_error_budget * trial / max(t_final, trial)

                def ok(dim, e, b=budget_now):
                    return e <= b
            budget = self.state_error_budget * trial / max(t_final, trial)

            def search(ladder):
                errs, rows = yield (_CAND, (ladder, off, trial))
                e = math.inf
                for i, dim in enumerate(ladder):
                    e = float(errs[i])
                    if ok(dim, e):
                        return dim, rows[i], e
                return None, None, e

            def exact_step(size):
                _, rows = yield (_CAND, ([size], 0, trial))
                return size, rows[0], 0.0

            def continue_search(built, ladder, done):
                """First passing ladder dimension: those the basis holds above `done`, then the basis is
                continued one ladder point at a time. Returns (selected, row, err, reference basis size)."""
                e = math.inf
                while True:
                    dims = [d for d in ladder if done < d <= built]
                    if dims:
                        sel, r, e = yield from search(dims)
                        if sel is not None:
                            return sel, r, e, mx
                    done = built
                    if built >= mx:
                        return None, None, e, mx
                    stop = min([d for d in ladder if d > built] + [mx])
                    size, _ = yield (_BASIS, (0, stop, True))
                    device_mv[0] += size - built
                    if size < stop:     # breakdown after the prefix: exact, as the whole basis would be
                        sel, r, e = yield from exact_step(size)
                        return sel, r, e, size
                    built = size

            # The reference builds a basis of min (or, prefer-max, max) vectors and searches the ladder
            # (R/rpm_statevector_engine.py:776-846). Here the basis is built in prefixes: up to the
            # dimension the last step selected (min before prefer-max), continued (never rebuilt) one
            # ladder point at a time while no candidate passes. The vectors, candidate errors and decisions
            # are those of the whole-basis calls; only vectors no candidate needs are left unbuilt.
            first = min(max(guess, mn), mx) if prefer_max else mn
            size, snorm = yield (_BASIS, (mx, first, False))
            if snorm == 0.0:
                raise ValueError("Cannot propagate the zero vector.")
            device_mv[0] += size
            selected, row, err = None, None, math.inf
            exact = size < first
            if exact:          # breakdown inside the first prefix: the reference's basis breaks down there too
                sizes = [size]
                selected, row, err = yield from exact_step(size)
            elif not prefer_max:
                sizes = [size]
                errs, rows = yield (_CAND, ([mn], off, trial))
                err = float(errs[0])
                if ok(mn, err):
                    selected, row = mn, rows[0]
                elif mx > mn:
                    # the reference rebuilds a max basis from the same state: the same vectors, continued
                    prefer_max = True
                    selected, row, err, ref = yield from continue_search(size, self._ladder(mn + inc), mn)
                    sizes.append(ref)
                    exact = ref < mx
            else:
                selected, row, err, ref = yield from continue_search(size, self._ladder(mn), 0)
                sizes = [ref]
                exact = ref < mx
            matvecs += sum(sizes)
            if selected is None or (replay is None and err > budget):
                rejected += 1
                record.append((trial, False, 0, tuple(sizes)))
                step = 0.5 * trial
                continue

            record.append((trial, True, int(selected), tuple(sizes)))
            guess = mx if exact else int(selected)
            max_dim = max(max_dim, selected)
            if selected > mn:
                prefer_max = True
            low = max(1, selected - off)
            low_arg = 0 if (exact or low == selected) else low
            step_end = now + trial
            stop = cursor
            while stop < nt and times[stop] <= step_end + 1e-13:
                stop += 1
            if stop > cursor:
                vals, perr = yield (_OUT, (selected, low_arg, times[cursor:stop] - now, cursor, acc_err))
                values[cursor:stop] = vals
                errors[cursor:stop] = acc_err + perr
                bounds[cursor:stop] = [state_error_to_population_bound(e) for e in errors[cursor:stop]]
                cursor = stop
            norm_err = yield (_ADV, (row, selected))
            now = step_end
            acc_err += err
            accepted += 1
            used.append(trial)
            max_norm_err = max(max_norm_err, norm_err)
            if callback is not None:
                callback(now, accepted, rejected, matvecs, acc_err)
            step = min(self.maximum_step_us, 1.5 * trial) if err < 0.1 * budget else trial

        stats = PropagationStats(accepted, rejected, matvecs, max_dim, min(used) if used else 0.0,
                                 max(used) if used else 0.0, acc_err, max_norm_err)
        return values, bounds, errors, stats, tuple(record), device_mv[0]

    # -- driver -------------------------------------------------------------

    def propagate(self, initial_states, times_us: Iterable[float], *, pairs=None, schedules=None,
                  progress_callback: Callable | None = None, yield_weights=None) -> list[PropagationResult]:
        """Propagate B trajectories; returns one ``PropagationResult`` per trajectory.

        ``initial_states``: (B, N) NumPy array or complex128 GPU tensor (copied).
        ``times_us``: one increasing sequence shared by the batch, or B of them (each trajectory
        then follows the step rules for its own final time, as a separate run would).
        ``pairs``: optional (B, 4, 4) electron matrices (see ``pair_matrices``);
        default: the Hamiltonian's own.  ``schedules``: optional per-trajectory
        recorded schedules to replay (entries may be None).
        ``progress_callback(b, now, accepted, rejected, matvecs, acc_err)``
        runs after each accepted step of trajectory b.
        ``yield_weights``: optional (nt, 2) fine and coarse yield weights of the
        grid (``quadrature.yield_weight_table``), shared, or one table per
        trajectory; each trajectory's Phi_S, coarse Phi_S and propagation
        estimate are then summed on the device as its outputs are produced and
        returned in ``PropagationResult.device_yields``.
        """
        torch = nat.require_gpu()
        lib = nat.load()
        dev = nat.device()
        h = self.hamiltonian
        op = h._device_op()
        n = h.dimension
        x0 = initial_states
        if nat.is_torch(x0):
            states = x0.to(device=dev, dtype=torch.complex128).reshape(x0.shape[0], -1).clone()
        else:
            arr = np.asarray(x0, dtype=complex)
            states = torch.from_numpy(np.ascontiguousarray(arr.reshape(arr.shape[0], -1))).to(dev)
        B = states.shape[0]
        if states.shape[1] != n:
            raise ValueError("Initial state dimension does not match the Hamiltonian.")
        if B < 1 or B > 65535:
            raise ValueError("batch size must be in 1..65535")
        traj_times = _trajectory_times(times_us, B)
        m = min(self.maximum_dimension, n)
        if m > 64:
            raise ValueError("maximum_dimension above 64 is not supported by the device kernels.")
        d_pairs = None
        if pairs is not None:
            p = np.asarray(pairs, dtype=complex)
            if p.shape != (B, 4, 4):
                raise ValueError(f"pairs must have shape ({B}, 4, 4)")
            if not np.allclose(p, np.conj(np.transpose(p, (0, 2, 1))), atol=1e-12 * max(1.0, np.abs(p).max())):
                raise ValueError("Electron matrices must be Hermitian.")
            d_pairs = torch.from_numpy(np.ascontiguousarray(p.reshape(B, 16))).to(dev)
        if schedules is not None and len(schedules) != B:
            raise ValueError("schedules must have one entry per trajectory")
        yacc = None
        if yield_weights is not None:
            tables = list(yield_weights) if isinstance(yield_weights, (list, tuple)) else [yield_weights] * B
            if len(tables) != B:
                raise ValueError("yield_weights must be one table or one per trajectory")
            for b, tab in enumerate(tables):
                if np.shape(tab) != (traj_times[b].size, 2):
                    raise ValueError("each yield weight table must have shape (len(times), 2)")
            yacc = YieldAccumulator(tables)

        V = torch.empty((B, m, n), dtype=torch.complex128, device=dev)
        alpha = torch.zeros((B, m), dtype=torch.float64, device=dev)
        beta = torch.zeros((B, m), dtype=torch.float64, device=dev)
        info = torch.zeros((B, 4), dtype=torch.int32, device=dev)
        norm = torch.zeros(B, dtype=torch.float64, device=dev)
        ws = nat.workspace(lib.rpm_lanczos_prefix_workspace_bytes(ctypes.c_void_p(op.handle), m, B))
        flags = (nat.RPM_LANCZOS_PAIRED if self.paired_steps else 0) | self._basis_test_flags
        gram_ws = nat.workspace(lib.rpm_singlet_gram_batch_workspace_bytes(n // 4, B))
        P = torch.zeros((B, m, m), dtype=torch.complex128, device=dev)
        ladder_len = len(self._ladder(self.minimum_dimension))
        coeffs = torch.zeros((B * ladder_len, m), dtype=torch.complex128, device=dev)
        cand_err = torch.zeros(B * ladder_len, dtype=torch.float64, device=dev)
        size_h = np.zeros(B, dtype=np.int64)
        red_ws = nat.workspace(lib.rpm_reduce_workspace_bytes(n, B))
        d_ts = torch.zeros(2 * B, dtype=torch.int32, device=dev)     # basis targets and stops of a call
        # each basis step window as a HIP graph, captured once for this call's buffers and replayed
        graphs = BasisGraphs() if self.use_graphs else None
        self.graphs = graphs
        nrm_out = torch.empty(B, dtype=torch.float64, device=dev)
        pop_out = torch.empty(B, dtype=torch.float64, device=dev)
        sp = ctypes.c_void_p(nat.stream_ptr())
        vp = ctypes.c_void_p

        def norms():
            nat.check(lib.rpm_norm(vp(nat.ptr(states)), n, B, n, vp(nat.ptr(nrm_out)), vp(nat.ptr(red_ws)),
                                   red_ws.numel(), sp))
            return nrm_out.cpu().numpy()

        def run_tasks(tasks, with_p, v0, v1, coeff_buf):
            dt = torch.from_numpy(tasks.view(np.uint8)).to(dev)
            nat.check(lib.rpm_krylov_tasks(
                vp(nat.ptr(dt)), len(tasks), vp(nat.ptr(alpha)), vp(nat.ptr(beta)), m, vp(nat.ptr(norm)),
                vp(nat.ptr(P) if with_p else None), m * m, float(self.safety_factor), vp(nat.ptr(v0)),
                vp(nat.ptr(v1) if v1 is not None else None),
                vp(nat.ptr(coeff_buf) if coeff_buf is not None else None), m, sp))

        cb = progress_callback
        gens, pending, results = {}, {}, {}
        for b in range(B):
            replay = None if schedules is None or schedules[b] is None else list(schedules[b])
            call = None if cb is None else (lambda *a, b=b: cb(b, *a))
            gens[b] = self._rules(traj_times[b], replay, call)
            pending[b] = next(gens[b])

        def resume(b, value):
            try:
                pending[b] = gens[b].send(value)
            except StopIteration as stop:
                results[b] = stop.value
                del pending[b]

        self.rounds = 0
        while pending:
            self.rounds += 1
            for phase in _PHASES:
                group = sorted(b for b, (ph, _) in pending.items() if ph == phase)
                if not group:
                    continue
                if phase == _INIT:
                    nv = norms()
                    nat.check(lib.rpm_singlet_population(vp(nat.ptr(states)), n // 4, B, n, vp(nat.ptr(pop_out)),
                                                         vp(nat.ptr(red_ws)), red_ws.numel(), sp))
                    if yacc is not None:
                        cnt = np.zeros(B, dtype=np.int64)
                        for b in group:
                            cnt[b] = 1 if traj_times[b][0] == 0.0 else 0
                        yacc.add(pop_out, None, np.arange(B), cnt, np.zeros(B, dtype=np.int64))
                    pv = pop_out.cpu().numpy()
                    for b in group:
                        resume(b, (abs(float(nv[b]) - 1.0), float(pv[b])))
                elif phase == _BASIS:
                    # targets (0: continue an earlier prefix) then stops, into the call's own buffer
                    ts = np.zeros(2 * B, dtype=np.int32)
                    j0 = m
                    for b in group:
                        target, stop, cont = pending[b][1]
                        ts[B + b] = min(int(stop), n)
                        if cont:
                            j0 = min(j0, int(size_h[b]))
                        else:
                            ts[b] = min(int(target), n)
                            j0 = 0
                    d_ts.copy_(torch.from_numpy(ts))
                    max_stop = int(ts[B:].max())

                    def launch(stream, j0=j0, max_stop=max_stop):
                        nat.check(lib.rpm_lanczos_basis_batch_prefix(
                            vp(op.handle), B, vp(nat.ptr(states)), n, vp(nat.ptr(d_ts)), m, vp(nat.ptr(d_ts) + 4 * B),
                            j0, max_stop, vp(nat.ptr(d_pairs) if d_pairs is not None else None), vp(nat.ptr(V)), m,
                            vp(nat.ptr(alpha)), vp(nat.ptr(beta)), vp(nat.ptr(info)), vp(nat.ptr(norm)),
                            float(self.breakdown_tolerance), flags, vp(nat.ptr(ws)), ws.numel(), vp(stream)))

                    if graphs is not None:
                        graphs.run((j0, max_stop), launch)
                    else:
                        launch(nat.stream_ptr())
                    inf_h = info.cpu().numpy()
                    while inf_h[group, 3].any():
                        # paired steps stopped these bases early at a near-breakdown: continue them
                        short = [b for b in group if inf_h[b, 3]]
                        ts = np.zeros(2 * B, dtype=np.int32)
                        for b in short:
                            ts[B + b] = min(int(pending[b][1][1]), n)
                        d_ts.copy_(torch.from_numpy(ts))
                        launch(nat.stream_ptr(), j0=int(min(inf_h[b, 0] for b in short)), max_stop=int(ts[B:].max()))
                        inf_new = info.cpu().numpy()
                        for b in short:
                            if inf_new[b, 3] and inf_new[b, 0] <= inf_h[b, 0]:
                                raise RuntimeError(f"Lanczos basis of trajectory {b}: continuing from "
                                                   f"{int(inf_h[b, 0])} vectors made no progress")
                            inf_h[b] = inf_new[b]
                    nrm_h = norm.cpu().numpy()
                    for b in group:
                        size_h[b] = int(inf_h[b, 0])
                        resume(b, (int(inf_h[b, 0]), float(nrm_h[b])))
                elif phase == _CAND:
                    tasks, spans = [], {}
                    for b in group:
                        dims, offset, t = pending[b][1]
                        start = len(tasks)
                        for d in dims:
                            dim = min(int(d), int(size_h[b]))
                            tasks.append((b, dim, max(1, dim - offset), len(tasks), t, 0, 0))
                        spans[b] = (start, len(tasks))
                    ta = np.array(tasks, dtype=_TASK)
                    run_tasks(ta, False, cand_err, None, coeffs)
                    eh = cand_err[:len(ta)].cpu().numpy()
                    for b in group:
                        s0, s1 = spans[b]
                        resume(b, (eh[s0:s1], list(range(s0, s1))))
                elif phase == _OUT:
                    kb = np.zeros(B, dtype=np.int32)
                    tasks, spans = [], {}
                    y_rows = np.zeros(B, dtype=np.int64)
                    y_cnt = np.zeros(B, dtype=np.int64)
                    y_grid = np.zeros(B, dtype=np.int64)
                    y_acc = np.zeros(B)
                    for b in group:
                        sel, low_arg, offs, cursor, acc_b = pending[b][1]
                        kb[b] = sel
                        start = len(tasks)
                        for t in offs:
                            tasks.append((b, sel, low_arg, len(tasks), float(t), 1, 0))
                        spans[b] = (start, len(tasks))
                        y_rows[b], y_cnt[b], y_grid[b], y_acc[b] = start, len(tasks) - start, cursor, acc_b
                    d_kb = torch.from_numpy(kb).to(dev)
                    nat.check(lib.rpm_singlet_gram_batch(vp(nat.ptr(V)), m, vp(nat.ptr(d_kb)), B, n // 4,
                                                         vp(nat.ptr(P)), vp(nat.ptr(gram_ws)), gram_ws.numel(), sp))
                    ta = np.array(tasks, dtype=_TASK)
                    vals = torch.empty(len(ta), dtype=torch.float64, device=dev)
                    perr = torch.empty(len(ta), dtype=torch.float64, device=dev)
                    run_tasks(ta, True, vals, perr, None)
                    if yacc is not None:
                        yacc.add(vals, perr, y_rows, y_cnt, y_grid, y_acc)
                    vh = vals.cpu().numpy()
                    ph_ = perr.cpu().numpy()
                    for b in group:
                        s0, s1 = spans[b]
                        resume(b, (vh[s0:s1], ph_[s0:s1]))
                elif phase == _ADV:
                    kb = np.zeros(B, dtype=np.int32)
                    rows = np.zeros(B, dtype=np.int64)
                    for b in group:
                        row, sel = pending[b][1]
                        kb[b] = sel
                        rows[b] = row
                    d_kb = torch.from_numpy(kb).to(dev)
                    cbuf = coeffs.index_select(0, torch.from_numpy(rows).to(dev)).contiguous()
                    nat.check(lib.rpm_krylov_combine_batch(
                        vp(nat.ptr(V)), m, vp(nat.ptr(d_kb)), B, n, vp(nat.ptr(cbuf)), m, vp(nat.ptr(norm)),
                        vp(nat.ptr(states)), n, sp))
                    nv = norms()
                    for b in group:
                        resume(b, abs(float(nv[b]) - 1.0))

        self.final_states = states
        self.device_matvecs = np.array([results[b][5] for b in range(B)], dtype=np.int64)
        phi = None if yacc is None else yacc.result()
        out = []
        for b in range(B):
            values, bounds, errors, stats, record, _ = results[b]
            dy = None if phi is None else tuple(float(v) for v in phi[b])
            out.append(PropagationResult(traj_times[b], values, bounds, errors, None, None, stats, record, states[b],
                                         dy))
        return out


def directional_yields(base, hamiltonian_for, propagator, states, thetas_deg, phis_deg, fields_uT, times_us,
                       lifetime_us, *, batch_size=32, progress_callback=None) -> np.ndarray:
    """Singlet yields of every (random-phase state, field direction, field strength): (S, D, F).

    The directional-map workload: many random-phase vectors x field directions (e.g.
    ``workloads.fibonacci_sphere``) x field strengths, run ``batch_size`` trajectories per batched
    propagation. ``hamiltonian_for(theta_deg, field_mT, phi_deg)`` builds each trajectory's
    Hamiltonian; only its electron part may differ from ``base``. The yield of each trajectory is the
    reference's Simpson yield on ``times_us`` (R/rpm_quantum_needle_jd_sweeps.py:418-420). The map
    is ``yields.mean(axis=0)``. ``batch_size="auto"`` picks the batch by timing the first
    trajectories over a short prefix (``autotune.propagation_batch``).
    """
    from .quadrature import yield_weight_table

    states = np.asarray(states) if not nat.is_torch(states) else states
    thetas = np.atleast_1d(np.asarray(thetas_deg, dtype=float))
    phis = np.broadcast_to(np.asarray(phis_deg, dtype=float), thetas.shape)
    fields = np.atleast_1d(np.asarray(fields_uT, dtype=float))
    S, D, F = states.shape[0], thetas.size, fields.size
    combos = [(s, d, f) for s in range(S) for d in range(D) for f in range(F)]
    cache: dict = {}

    def electron_matrices(sel):
        """4x4 electron matrices of (direction, strength) pairs, each Hamiltonian built once."""
        todo = [(d, f) for _, d, f in sel if (d, f) not in cache]
        todo = list(dict.fromkeys(todo))
        if todo:
            hs = [hamiltonian_for(float(thetas[d]), 1e-3 * float(fields[f]), float(phis[d])) for d, f in todo]
            for key, mat in zip(todo, pair_matrices(base, hs)):
                cache[key] = mat
        return np.stack([cache[(d, f)] for _, d, f in sel])

    if batch_size == "auto":
        from . import autotune
        lead = combos[:64]
        idx = [s for s, _, _ in lead]
        batch_size = autotune.propagation_batch(propagator, states[idx], times_us, electron_matrices(lead))["batch"]
    out = np.empty((S, D, F))
    weights = yield_weight_table(times_us, lifetime_us)     # Phi_S summed on the device
    for c0 in range(0, len(combos), max(1, int(batch_size))):
        chunk = combos[c0:c0 + max(1, int(batch_size))]
        pairs = electron_matrices(chunk)
        idx = [s for s, _, _ in chunk]
        x0 = states[idx] if not nat.is_torch(states) else states[idx]
        results = propagator.propagate(x0, times_us, pairs=pairs, yield_weights=weights)
        for (s, d, f), r in zip(chunk, results):
            out[s, d, f] = r.device_yields[0]
        if progress_callback is not None:
            progress_callback(min(c0 + len(chunk), len(combos)), len(combos))
    return out
