"""Segmented propagation: piecewise-constant fields and instantaneous local unitaries.

Follows the segment loop of R/rpm_nuclear_spin_rotation_control.py:247-342
(also R/poster_reproduction/runner.py:315-356): the output grid is cut at
the segment boundaries; each segment is propagated by a fresh adaptive Lanczos
propagator whose error budget is the global budget times duration / T_final;
state errors accumulate across segments; after every boundary but the last an
optional local unitary acts on one tensor axis (R/poster_reproduction/core.py:285-292).

Beyond the reference, a segment may carry its own electron 4x4 matrix (a
field pulse: B(t) piecewise constant), and the whole batch of trajectories
runs through ``BatchedLanczosPropagator`` segment by segment.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Callable, Sequence

import numpy as np
import scipy.linalg

from . import _native as nat
from .engine import SingletPopulationObservable, singlet_population, state_error_to_population_bound
from .model import spin_operators


@dataclass
class Segment:
    """One piece of a segmented run.

    ``end_us``: segment end time.  ``pairs``: electron 4x4 per trajectory
    ((B, 4, 4), or (4, 4) for all) during the segment; None keeps the
    Hamiltonian's own.  ``unitary``: (axis, U) applied to the state at the end
    of the segment (skipped at the final time, as the reference does).
    """

    end_us: float
    pairs: object = None
    unitary: tuple | None = None


@dataclass
class SegmentedResult:
    times_us: np.ndarray
    populations: np.ndarray          # (B, nt)
    bounds: np.ndarray               # (B, nt) population error bounds, 2e + e^2
    accumulated_state_error: np.ndarray
    accepted_steps: np.ndarray
    rejected_steps: np.ndarray
    matvecs: np.ndarray
    final_states: object             # device (B, N)
    schedules: list = field(default_factory=list)   # per segment, per trajectory
    yields: np.ndarray | None = None  # (B, 2) Phi_S and coarse Phi_S when yield weights were given


# ---------------------------------------------------------------- reference helpers

def pulse_schedule(interval_us: float, final_time_us: float) -> np.ndarray:
    """R/rpm_nuclear_spin_rotation_control.py:219-223."""
    interval = float(interval_us)
    if interval <= 0.0:
        raise ValueError("Pulse interval must be positive.")
    return np.arange(interval, float(final_time_us), interval)


def dominant_principal_axis(tensor) -> np.ndarray:
    """R/rpm_nuclear_spin_rotation_control.py:201-208."""
    sym = 0.5 * (np.asarray(tensor, float) + np.asarray(tensor, float).T)
    values, vectors = np.linalg.eigh(sym)
    axis = vectors[:, int(np.argmax(np.abs(values)))].copy()
    if axis[int(np.argmax(np.abs(axis)))] < 0.0:
        axis *= -1.0
    return axis


def spin_pi_pulse(spin: float, molecular_axis) -> np.ndarray:
    """exp(-i pi n.I) (R/poster_reproduction/core.py:304-308)."""
    axis = np.asarray(molecular_axis, dtype=float)
    axis = axis / np.linalg.norm(axis)
    gen = sum(v * op for v, op in zip(axis, spin_operators(spin)))
    return scipy.linalg.expm(-1.0j * math.pi * gen)


def local_unitary_apply(state, dimensions, axis: int, unitary):
    """(1 (x) U (x) 1) state on the GPU (R/poster_reproduction/core.py:285-292).

    ``state``: NumPy vector (returns NumPy) or GPU tensor of shape (N,) or (B, N)
    (returns a new GPU tensor).
    """
    torch = nat.require_gpu()
    dims = tuple(int(d) for d in dimensions)
    U = np.asarray(unitary, dtype=complex)
    if not 0 <= int(axis) < len(dims) or U.shape != (dims[axis], dims[axis]):
        raise ValueError("unitary shape does not match the axis dimension")
    host = not nat.is_torch(state)
    x = nat.to_device_c128(np.asarray(state, complex).reshape(-1) if host else state)
    n = math.prod(dims)
    x2 = x.reshape(-1, n).contiguous()
    out = torch.empty_like(x2)
    dU = torch.from_numpy(np.ascontiguousarray(U)).to(x2.device)
    arr = (ctypes.c_int * len(dims))(*dims)
    nat.check(nat.load().rpm_axis_apply(len(dims), arr, int(axis), ctypes.c_void_p(nat.ptr(dU)),
                                        ctypes.c_void_p(nat.ptr(x2)), ctypes.c_void_p(nat.ptr(out)), x2.shape[0], n,
                                        ctypes.c_void_p(nat.stream_ptr())))
    if host:
        return out.reshape(-1).cpu().numpy()
    return out.reshape(x.shape)


# ---------------------------------------------------------------- drivers

def _local_times(times, cursor, boundary, current):
    stop = int(np.searchsorted(times, boundary + 1e-12, side="right"))
    requested = times[cursor:stop] - current
    duration = boundary - current
    if requested.size == 0 or requested[-1] < duration - 1e-12:
        local = np.append(requested, duration)
    else:
        local = requested
    return stop, requested, duration, local


def _check_segments(times, segments):
    if times.ndim != 1 or times.size == 0 or times[0] != 0.0 or np.any(np.diff(times) <= 0.0):
        raise ValueError("The output grid must start at 0 and increase strictly.")
    ends = [float(s.end_us) for s in segments]
    if not ends or np.any(np.diff(ends) <= 0.0) or ends[0] <= 0.0:
        raise ValueError("Segment ends must be positive and increasing.")
    if abs(ends[-1] - times[-1]) > 1e-12:
        raise ValueError("The last segment must end at the final output time.")


def propagate_segments(hamiltonian, initial_state, times_us, segments: Sequence[Segment], make_propagator: Callable,
                       state_error_budget: float, *, hamiltonians=None, schedules=None) -> SegmentedResult:
    """Single trajectory through the reference's segment loop.

    ``make_propagator(h)`` builds a fresh ``AdaptiveLanczosPropagator`` per
    segment (R/poster_reproduction/core.py:232-247); ``hamiltonians``
    optionally gives one Hamiltonian per segment (piecewise-constant fields);
    ``schedules`` optionally gives one recorded schedule per propagated segment
    to replay (as ``SegmentedResult.schedules`` holds them).
    """
    times = np.asarray(times_us, dtype=float)
    _check_segments(times, segments)
    dims = tuple(hamiltonian.dimensions)
    state = nat.to_device_c128(np.asarray(initial_state, complex).reshape(-1)
                               if not nat.is_torch(initial_state) else initial_state).reshape(-1).clone()
    obs = SingletPopulationObservable(dims)
    nt = times.size
    pops = np.empty(nt)
    bounds = np.empty(nt)
    pops[0] = singlet_population(state, dims)
    bounds[0] = 0.0
    cursor, current, acc = 1, 0.0, 0.0
    accepted = rejected = matvecs = 0
    schedule_log = []
    for i, seg in enumerate(segments):
        boundary = float(seg.end_us)
        h = hamiltonian if hamiltonians is None else hamiltonians[i]
        stop, requested, duration, local = _local_times(times, cursor, boundary, current)
        if duration > 0.0:
            prop = make_propagator(h)
            prop.state_error_budget = state_error_budget * duration / times[-1]
            replay = None if schedules is None else schedules[len(schedule_log)][0]
            res = prop.propagate(state, local, obs, schedule=replay)
            count = requested.size
            pops[cursor:stop] = res.observable_values[:count]
            e = acc + res.state_error_estimates[:count]
            bounds[cursor:stop] = 2.0 * e + e ** 2
            state = res.final_state_device
            accepted += res.stats.accepted_steps
            rejected += res.stats.rejected_steps
            matvecs += res.stats.matvecs
            acc += res.stats.accumulated_state_error
            schedule_log.append([res.schedule])
        cursor, current = stop, boundary
        if seg.unitary is not None and boundary < times[-1] - 1e-12:
            axis, U = seg.unitary
            state = local_unitary_apply(state, dims, axis, U)
    return SegmentedResult(times, pops[None], bounds[None], np.array([acc]), np.array([accepted]),
                           np.array([rejected]), np.array([matvecs]), state.reshape(1, -1), schedule_log)


def propagate_segments_batched(propagator, initial_states, times_us, segments: Sequence[Segment],
                               state_error_budget: float, *, schedules=None, yield_weights=None) -> SegmentedResult:
    """A batch of trajectories through the same segment loop with ``BatchedLanczosPropagator``.

    Each segment may give per-trajectory electron matrices (field pulses);
    ``propagator.state_error_budget`` is set per segment as in the reference.
    ``schedules``: optional recorded schedules to replay, one list per
    propagated segment with one entry per trajectory (entries may be None).
    ``yield_weights``: optional (len(times_us), 2) fine/coarse weight table of
    the whole grid (``quadrature.yield_weight_table``); each segment then sums
    its slice of the yields on the device, and ``SegmentedResult.yields``
    holds their totals.
    """
    torch = nat.require_gpu()
    times = np.asarray(times_us, dtype=float)
    _check_segments(times, segments)
    h = propagator.hamiltonian
    dims = tuple(h.dimensions)
    if nat.is_torch(initial_states):
        states = initial_states.to(torch.complex128).reshape(initial_states.shape[0], -1).contiguous().clone()
    else:
        arr = np.asarray(initial_states, dtype=complex)
        states = torch.from_numpy(np.ascontiguousarray(arr.reshape(arr.shape[0], -1))).to(nat.device())
    B = states.shape[0]
    nt = times.size
    pops = np.empty((B, nt))
    bounds = np.empty((B, nt))
    from .engine import device_singlet_population
    pops[:, 0] = device_singlet_population(states.reshape(-1), B).cpu().numpy()
    bounds[:, 0] = 0.0
    table = yields = None
    if yield_weights is not None:
        table = np.asarray(yield_weights, dtype=float)
        if table.shape != (nt, 2):
            raise ValueError("yield_weights must have shape (len(times_us), 2)")
        yields = pops[:, :1] * table[0][None, :]         # the t = 0 term
    cursor, current = 1, 0.0
    acc = np.zeros(B)
    accepted = np.zeros(B, dtype=int)
    rejected = np.zeros(B, dtype=int)
    matvecs = np.zeros(B, dtype=int)
    schedule_log = []
    for seg in segments:
        boundary = float(seg.end_us)
        stop, requested, duration, local = _local_times(times, cursor, boundary, current)
        if duration > 0.0:
            pairs = None
            if seg.pairs is not None:
                pairs = np.asarray(seg.pairs, dtype=complex)
                if pairs.shape == (4, 4):
                    pairs = np.broadcast_to(pairs, (B, 4, 4))
            propagator.state_error_budget = state_error_budget * duration / times[-1]
            replay = None if schedules is None else schedules[len(schedule_log)]
            seg_table = None
            if table is not None:
                # the segment's outputs are grid points cursor .. stop-1; an added end point weighs nothing
                seg_table = np.zeros((local.size, 2))
                seg_table[:requested.size] = table[cursor:stop]
            res = propagator.propagate(states, local, pairs=pairs, schedules=replay, yield_weights=seg_table)
            if yields is not None:
                yields += np.array([r.device_yields[:2] for r in res])
            count = requested.size
            for b, r in enumerate(res):
                pops[b, cursor:stop] = r.observable_values[:count]
                e = acc[b] + r.state_error_estimates[:count]
                bounds[b, cursor:stop] = 2.0 * e + e ** 2
                accepted[b] += r.stats.accepted_steps
                rejected[b] += r.stats.rejected_steps
                matvecs[b] += r.stats.matvecs
                acc[b] += r.stats.accumulated_state_error
            states = propagator.final_states
            schedule_log.append([r.schedule for r in res])
        cursor, current = stop, boundary
        if seg.unitary is not None and boundary < times[-1] - 1e-12:
            axis, U = seg.unitary
            states = local_unitary_apply(states, dims, axis, U)
    return SegmentedResult(times, pops, bounds, acc, accepted, rejected, matvecs, states, schedule_log, yields)


def field_pulse_segments(base_pair: np.ndarray, direction, amplitudes_uT, t_start_us: float, t_end_us: float,
                         t_final_us: float) -> list[Segment]:
    """Segments of a piecewise-constant transverse pulse added to a static field.

    ``base_pair``: (B, 4, 4) static electron matrices; the pulse adds
    amplitudes_uT[s] * 1e-3 mT along ``direction`` to both electrons' Zeeman
    term on equal sub-intervals of [t_start, t_end]; a last segment carries the
    static field to t_final.
    """
    from .adjoint import electron_derivative_operators
    d = np.asarray(direction, float)
    d = d / np.linalg.norm(d)
    M = np.einsum("a,aij->ij", d, electron_derivative_operators())
    amps = np.asarray(amplitudes_uT, dtype=float)
    edges = np.linspace(t_start_us, t_end_us, amps.size + 1)
    segs = []
    if t_start_us > 0.0:
        segs.append(Segment(float(t_start_us), base_pair))
    for s, a in enumerate(amps):
        segs.append(Segment(float(edges[s + 1]), base_pair + 1e-3 * a * M))
    if t_final_us > t_end_us + 1e-12:
        segs.append(Segment(float(t_final_us), base_pair))
    return segs
