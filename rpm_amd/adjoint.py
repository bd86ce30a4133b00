"""Adjoint gradients of singlet yields with respect to field parameters and pulse amplitudes.

The reference computes yields only (R/rpm_quantum_needle_jd_sweeps.py:418-420,
R/rpm_full14_quadrupole_experiment.py:154-194); this module adds their
derivatives.  Only the electron Zeeman term depends on the field
(R/poster_reproduction/core.py:186-194; nuclear Zeeman is omitted), so every
dH/dp is an electron 4x4 operator E_p (x) 1 and

    dPhi/dp = Re sum_{e,f} E_p[e][f] A[e][f],

with A the 4x4 "cross-density" accumulated by the backward (costate) sweep
``rpm_adjoint_sweep``.  The sweep differentiates the yield exactly as it is
defined on the output grid, Phi = sum_k W_k <psi_k|P|psi_k> (SciPy-Simpson
weights times exp(-t/tau)/tau), with the grid-interval propagators
represented by their Chebyshev expansions (Tal-Ezer & Kosloff 1984) to
~1e-16:

* the costate g_k = 2 W_k P psi_k + U^H g_{k+1} receives one kick per output
  time and is carried back interval by interval with Clenshaw's recurrence;
* psi is carried back the same way from the final state of the forward run;
* d/dp of the interval propagator sum_n c_n T_n(X) in direction D = E/hs is
  <nu_1, D u_0> + 2 sum_n <nu_{n+1}, D u_n> with u_n = T_n(X) psi_{k-1} and nu
  the Clenshaw vectors of the costate (the discrete adjoint of the Chebyshev
  recurrence), accumulated as electron cross-densities.

The sweep takes steps of up to MACRO_INTERVALS grid intervals, one Chebyshev
series per step, with the output-time kicks of the step folded into the
Clenshaw seeds. Per grid interval this costs about 3K/64 matvecs (K = 122 for
64 intervals of 0.00125 us on the production operator), independent of the
number of parameters.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _native as nat
from .engine import _pair_form
from .model import GAMMA_E_MHZ_PER_MT, TWOPI, spin_operators
from .quadrature import yield_weights


def _bessel_j(nmax: int, z: float) -> np.ndarray:
    """J_0..J_nmax(z), z >= 0, by Miller's backward recurrence normalised by J_0 + 2 sum J_2k = 1."""
    if z == 0.0:
        out = np.zeros(nmax + 1)
        out[0] = 1.0
        return out
    start = nmax + 40 + int(math.sqrt(40.0 * (nmax + 1))) + int(2 * z)
    jp1, j = 0.0, 1e-300
    vals = np.zeros(nmax + 1)
    norm = 0.0
    for k in range(start, 0, -1):
        jm1 = 2.0 * k / z * j - jp1
        jp1, j = j, jm1
        if k - 1 <= nmax:
            vals[k - 1] = j
        if (k - 1) % 2 == 0 and k - 1 > 0:
            norm += 2.0 * j
        if abs(j) > 1e250:
            j *= 1e-250
            jp1 *= 1e-250
            norm *= 1e-250
            vals *= 1e-250
    norm += j
    return vals / norm


def chebyshev_series(h: float, shift: float, half_width: float, tol: float = 1e-18):
    """Coefficients c_n (n = 0..K) of exp(-i H h) = sum_n c_n T_n((H - shift)/half_width)."""
    z = half_width * h
    nmax = int(1.5 * z) + 60
    J = _bessel_j(nmax, z)
    K = next((n for n in range(2, nmax) if np.all(np.abs(J[n + 1:n + 4]) < tol)), nmax - 4)
    K = max(K, 2)
    n = np.arange(K + 1)
    a = 2.0 * ((-1j) ** n) * J[:K + 1]
    a[0] = J[0]
    return K, np.exp(-1j * shift * h) * a


def electron_derivative_operators() -> np.ndarray:
    """d(pair)/dB_a per mT for a = x, y, z: 2 pi gamma (S_a (x) 1 + 1 (x) S_a), rows eA*2+eB."""
    ops = spin_operators(0.5)
    i2 = np.eye(2, dtype=complex)
    return np.stack([TWOPI * GAMMA_E_MHZ_PER_MT * (np.kron(s, i2) + np.kron(i2, s)) for s in ops])


def field_parameter_derivatives(theta_deg: float, field_uT: float, phi_deg: float = 0.0) -> np.ndarray:
    """d(pair)/d(theta [rad], phi [rad], |B| [uT]) at one orientation: (3, 4, 4)."""
    t, p = math.radians(theta_deg), math.radians(phi_deg)
    b = 1e-3 * field_uT
    dB = np.array([[b * math.cos(t) * math.cos(p), b * math.cos(t) * math.sin(p), -b * math.sin(t)],
                   [-b * math.sin(t) * math.sin(p), b * math.sin(t) * math.cos(p), 0.0],
                   [1e-3 * math.sin(t) * math.cos(p), 1e-3 * math.sin(t) * math.sin(p), 1e-3 * math.cos(t)]])
    M = electron_derivative_operators()
    return np.einsum("pa,aij->pij", dB, M)


def spectral_bounds(hamiltonian, pairs=None):
    """[lo, hi] enclosing the spectrum of H (Weyl: sums of the extreme eigenvalues of its terms).

    ``pairs``: optional array (..., 4, 4) of electron matrices replacing the
    Hamiltonian's own; the bounds then hold for every one of them.
    """
    form = _pair_form(hamiltonian.dimensions, hamiltonian.terms, hamiltonian.left_axes, hamiltonian.right_axes)
    if form is None:
        raise ValueError("Adjoint sweeps need a Hamiltonian with the pair-stencil form.")
    blocks = form[2]
    lo = sum(float(np.linalg.eigvalsh(b).min()) for b in blocks)
    hi = sum(float(np.linalg.eigvalsh(b).max()) for b in blocks)
    ps = np.asarray(form[3] if pairs is None else pairs, dtype=complex).reshape(-1, 4, 4)
    ev = np.linalg.eigvalsh(ps)
    return lo + float(ev.min()), hi + float(ev.max())


# Output-grid intervals per backward Chebyshev step. Longer steps need fewer matvecs per interval
# (K grows more slowly than the step): on 8 production trajectories over 2 us the sweep took 1.05,
# 0.76, 0.67, 0.60, 0.55 and 0.52 s at 8, 16, 24, 32, 48 and 64 intervals per step, with gradients
# equal to 1e-13. A step whose series would pass MAX_TERMS is halved until it fits.
MACRO_INTERVALS = 64
MAX_TERMS = 255           # Chebyshev terms the device tables hold


def chebyshev_series_fixed(h: float, shift: float, half_width: float, K: int) -> np.ndarray:
    """c_0..c_K of exp(-i H h) in the Chebyshev basis of X = (H - shift)/half_width (fixed length K)."""
    J = _bessel_j(K, half_width * h)
    n = np.arange(K + 1)
    a = 2.0 * ((-1j) ** n) * J
    a[0] = J[0]
    return np.exp(-1j * shift * h) * a


def _macro_steps(M: int, seg: np.ndarray, slot: np.ndarray, per: int):
    """Cut M grid intervals into steps (a, b) of at most ``per`` intervals, each within one segment and slot."""
    steps = []
    a = 0
    while a < M:
        b = a + 1
        while b < M and b - a < per and seg[b] == seg[a] and slot[b] == slot[a]:
            b += 1
        steps.append((a, b))
        a = b
    return steps


@dataclass
class _StepClass:
    K: int
    coef: np.ndarray      # (K+1,) complex: c_n of the whole step
    G: np.ndarray         # (K+1, K+1) complex kick matrix


def _step_classes(grid, weights, steps, shift, half, decay_time):
    """Series and kick matrices per step, shared between steps that differ only by the weight factor.

    W_k = mult * w_k with mult = exp(-t_a / decay_time) (1 without a decay time), and
    G[m][j] = sum_k 2 w_k conj(c_m(t_k - t_a)) c_j(t_k - t_a) over the output times of the step.
    """
    classes, index, cls, mult = [], {}, [], []
    for a, b in steps:
        tk = grid[a + 1:b + 1] - grid[a]
        m = math.exp(-grid[a] / decay_time) if decay_time else 1.0
        wk = weights[a + 1:b + 1] / m
        key = (round(float(tk[-1]), 15), tuple(np.round(tk, 15)), tuple(np.round(wk, 14)))
        c = index.get(key)
        if c is None:
            K, cdelta = chebyshev_series(float(tk[-1]), shift, half)
            ck = np.stack([chebyshev_series_fixed(float(t), shift, half, K) for t in tk])
            G = np.einsum("k,km,kj->mj", 2.0 * wk, ck.conj(), ck)
            c = len(classes)
            classes.append(_StepClass(K, cdelta, G))
            index[key] = c
        cls.append(c)
        mult.append(m)
    return classes, np.asarray(cls, dtype=np.int32), np.asarray(mult, dtype=float)


@dataclass
class AdjointResult:
    densities: np.ndarray        # (nslot, B, 4, 4) complex: accumulated electron cross-densities
    initial_states: object       # device (B, N): psi_0 recovered by the backward sweep
    chebyshev_terms: tuple       # K per step class
    bounds: tuple                # spectral enclosure used
    steps: int                   # backward Chebyshev steps


def adjoint_densities(hamiltonian, final_states, times_us, weights, *, pairs=None, segment_of_interval=None,
                      slot_of_interval=None, nslot: int = 1, decay_time=None,
                      intervals_per_step: int = MACRO_INTERVALS) -> AdjointResult:
    """Run the backward sweep from the final states of a forward run.

    ``times_us``: sweep grid t_0 = 0 < ... < t_M (the output grid, plus any
    segment boundaries); ``weights``: W_k of Phi = sum_k W_k <psi_k|P|psi_k>
    on that grid (0 at points that are not output times; W_0 is not used,
    psi_0 does not depend on the parameters); ``decay_time``: tau when
    W_k = s_k exp(-t_k/tau)/tau (lets equal steps share their tables).
    ``final_states``: complex128 GPU tensor (B, N) with psi(t_M) of each
    trajectory (left untouched; a copy is swept).  ``pairs``: electron
    matrices, (B, 4, 4) for one field per trajectory or (S, B, 4, 4) per
    segment with ``segment_of_interval`` (M ints).  ``slot_of_interval``:
    accumulator slot per interval (default all 0).
    """
    torch = nat.require_gpu()
    lib = nat.load()
    h = hamiltonian
    op = h._device_op()
    times = np.asarray(times_us, dtype=float)
    if times.ndim != 1 or times.size < 2 or times[0] != 0.0 or np.any(np.diff(times) <= 0.0):
        raise ValueError("The output grid must start at 0 and increase strictly.")
    w = np.asarray(weights, dtype=float)
    if w.shape != times.shape:
        raise ValueError("weights must match the sweep grid")
    psi = final_states.to(torch.complex128).reshape(final_states.shape[0], -1).contiguous().clone()
    B, n = psi.shape
    if n != h.dimension:
        raise ValueError("State dimension does not match the Hamiltonian.")
    M = times.size - 1
    p = None
    d_pairs = None
    seg = np.zeros(M, dtype=np.int32)
    if pairs is not None:
        p = np.asarray(pairs, dtype=complex)
        if p.ndim == 3:
            p = p[None]
        if p.shape[1:] != (B, 4, 4):
            raise ValueError(f"pairs must have shape (B, 4, 4) or (S, B, 4, 4) with B = {B}")
        if segment_of_interval is not None:
            seg = np.asarray(segment_of_interval, dtype=np.int32)
        if seg.shape != (M,) or seg.min() < 0 or seg.max() >= p.shape[0]:
            raise ValueError("segment_of_interval must give a valid segment for each interval")
        d_pairs = torch.from_numpy(np.ascontiguousarray(p.reshape(p.shape[0], B, 16))).to(psi.device)
    slot = np.zeros(M, dtype=np.int32) if slot_of_interval is None else np.asarray(slot_of_interval, dtype=np.int32)
    if slot.shape != (M,) or slot.min() < 0 or slot.max() >= nslot:
        raise ValueError("slot_of_interval must give a slot in 0..nslot-1 for each interval")
    lo, hi = spectral_bounds(h, p)
    shift = 0.5 * (lo + hi)
    half = 0.5 * (hi - lo) * (1.0 + 1e-9) + 1e-12
    per = max(1, int(intervals_per_step))
    while True:
        steps = _macro_steps(M, seg, slot, per)
        classes, cls, mult = _step_classes(times, w, steps, shift, half, decay_time)
        kmax = max(c.K for c in classes)
        if kmax <= MAX_TERMS or per == 1:
            break
        per = max(1, per // 2)
    if kmax > MAX_TERMS:
        raise ValueError("sweep steps too long for the Chebyshev tables (grid intervals too long for the spectrum)")
    coef = np.zeros((len(classes), kmax + 1), dtype=complex)
    Gs = np.zeros((len(classes), kmax + 1, kmax + 1), dtype=complex)
    nterms = np.zeros(len(classes), dtype=np.int32)
    for c, sc in enumerate(classes):
        nterms[c] = sc.K
        coef[c, :sc.K + 1] = sc.coef
        Gs[c, :sc.K + 1, :sc.K + 1] = sc.G
    step_seg = np.ascontiguousarray(np.array([seg[a] for a, _ in steps], dtype=np.int32))
    step_slot = np.ascontiguousarray(np.array([slot[a] for a, _ in steps], dtype=np.int32))
    acc = torch.zeros((nslot, B, 16), dtype=torch.complex128, device=psi.device)
    dG = torch.from_numpy(np.ascontiguousarray(Gs)).to(psi.device)
    ws = nat.workspace(lib.rpm_adjoint_workspace_bytes(ctypes.c_void_p(op.handle), B, int(kmax)))
    coef_ri = np.ascontiguousarray(np.stack([coef.real, coef.imag], axis=-1).reshape(-1))
    cls_c = np.ascontiguousarray(cls)
    mult_c = np.ascontiguousarray(mult)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    nat.check(lib.rpm_adjoint_sweep(
        ctypes.c_void_p(op.handle), B, ctypes.c_void_p(nat.ptr(psi)), len(steps), shift, half, len(classes),
        nterms.ctypes.data_as(ip), int(kmax), coef_ri.ctypes.data_as(dp), ctypes.c_void_p(nat.ptr(dG)),
        cls_c.ctypes.data_as(ip), mult_c.ctypes.data_as(dp),
        ctypes.c_void_p(nat.ptr(d_pairs) if d_pairs is not None else None),
        step_seg.ctypes.data_as(ip) if d_pairs is not None else None, int(p.shape[0]) if p is not None else 0,
        ctypes.c_void_p(nat.ptr(acc)), step_slot.ctypes.data_as(ip), int(nslot), ctypes.c_void_p(nat.ptr(ws)),
        ws.numel(), ctypes.c_void_p(nat.stream_ptr())))
    dens = acc.cpu().numpy().reshape(nslot, B, 4, 4)
    return AdjointResult(dens, psi, tuple(int(c.K) for c in classes), (lo, hi), len(steps))


def contract(dH: np.ndarray, densities: np.ndarray) -> np.ndarray:
    """dPhi/dp = Re sum_{e,f} dH[b, p, e, f] A[b, e, f]: (B, P, 4, 4) x (B, 4, 4) -> (B, P)."""
    return np.einsum("bpef,bef->bp", np.asarray(dH), np.asarray(densities)).real


# ---------------------------------------------------------------- yields with gradients

def sweep_grid(times_us, lifetime_us, boundaries=()):
    """Union of the output grid and segment boundaries, with the yield weights (0 at added points)."""
    times = np.asarray(times_us, dtype=float)
    w = yield_weights(times, lifetime_us)
    extra = [float(b) for b in boundaries
             if times[0] < b < times[-1] and np.min(np.abs(times - b)) > 1e-12]
    if not extra:
        return times, w
    grid = np.concatenate([times, extra])
    wt = np.concatenate([w, np.zeros(len(extra))])
    order = np.argsort(grid, kind="stable")
    return grid[order], wt[order]


def _segment_index(grid, ends):
    """Segment of each interval [t_{k-1}, t_k]: the first segment whose end is >= t_k."""
    ends = np.asarray(ends, dtype=float)
    return np.searchsorted(ends, grid[1:] - 1e-12, side="left").astype(np.int32)


@dataclass
class YieldGradient:
    yields: np.ndarray               # (B,)
    field_gradient: np.ndarray       # (B, 3): d/d theta [rad], d/d phi [rad], d/d |B| [uT]
    pulse_gradient: np.ndarray | None   # (B, S): d/d amplitude_s [uT]
    populations: np.ndarray          # (B, nt)
    recovered_initial_error: float   # max_b ||psi_0 recovered by the sweep - psi_0|| (consistency check)


class FieldYield:
    """Singlet yields of a batch of trajectories as a function of their fields, with adjoint gradients.

    ``hamiltonian_for(theta_deg, field_mT, phi_deg)`` builds the Hamiltonian of
    one orientation (its electron matrix is taken; the nuclear part must be
    ``base``'s).  ``propagator``: a ``BatchedLanczosPropagator`` on ``base``
    (reference Krylov rules; its ``state_error_budget`` is the global budget).
    With ``pulse_window_us`` = (t0, t1) a piecewise-constant pulse along
    ``pulse_direction`` (amplitudes in uT, one per equal sub-interval) is added
    and the forward run is segmented as in the reference's segment loop.
    """

    def __init__(self, base, hamiltonian_for, propagator, initial_states, times_us, lifetime_us, *,
                 pulse_window_us=None, pulse_direction=(1.0, 0.0, 0.0)):
        self.base = base
        self.hamiltonian_for = hamiltonian_for
        self.propagator = propagator
        self.initial_states = initial_states
        self.times = np.asarray(times_us, dtype=float)
        self.tau = float(lifetime_us)
        self.pulse_window = pulse_window_us
        d = np.asarray(pulse_direction, float)
        self.pulse_direction = d / np.linalg.norm(d)
        self._saved = None

    def forward(self, thetas_deg, phis_deg, fields_uT, pulse_amplitudes_uT=None) -> np.ndarray:
        """Yields (B,); keeps what the backward sweep needs."""
        from .batch import pair_matrices
        from .quadrature import yield_weight_table
        from .segments import Segment, field_pulse_segments, propagate_segments_batched

        times = self.times
        B = len(thetas_deg)
        hs = [self.hamiltonian_for(float(t), 1e-3 * float(f), float(p))
              for t, p, f in zip(thetas_deg, phis_deg, fields_uT)]
        pairs = pair_matrices(self.base, hs)
        prop = self.propagator
        budget = prop.state_error_budget
        if pulse_amplitudes_uT is None:
            res = prop.propagate(self.initial_states, times, pairs=pairs,
                                 yield_weights=yield_weight_table(times, self.tau))
            final = prop.final_states
            pops = np.stack([r.observable_values for r in res])
            yields = np.array([r.device_yields[0] for r in res])
            grid, w = sweep_grid(times, self.tau)
            seg_pairs = pairs[None]
            seg_idx = np.zeros(grid.size - 1, dtype=np.int32)
            pulse_slots = []
            schedules = [[r.schedule for r in res]]
        else:
            if self.pulse_window is None:
                raise ValueError("pulse amplitudes need pulse_window_us")
            amps = np.asarray(pulse_amplitudes_uT, dtype=float)
            amps = np.broadcast_to(amps, (B, amps.shape[-1])) if amps.ndim == 1 else amps
            t0, t1 = self.pulse_window
            per_b = [field_pulse_segments(pairs[b:b + 1], self.pulse_direction, amps[b], t0, t1, times[-1])
                     for b in range(B)]
            segs = [Segment(per_b[0][i].end_us, np.concatenate([per_b[b][i].pairs for b in range(B)]))
                    for i in range(len(per_b[0]))]
            try:
                out = propagate_segments_batched(prop, self.initial_states, times, segs, budget,
                                                 yield_weights=yield_weight_table(times, self.tau))
            finally:
                prop.state_error_budget = budget
            final = out.final_states
            pops = out.populations
            ends = [sg.end_us for sg in segs]
            grid, w = sweep_grid(times, self.tau, ends)
            seg_idx = _segment_index(grid, ends)
            seg_pairs = np.stack([sg.pairs for sg in segs])
            first = 1 if t0 > 0.0 else 0
            pulse_slots = list(range(first, first + amps.shape[1]))
            yields = out.yields[:, 0]
            schedules = out.schedules
        dH = np.stack([field_parameter_derivatives(float(t), float(f), float(p))
                       for t, p, f in zip(thetas_deg, phis_deg, fields_uT)])
        # a new record per call: earlier records stay valid for their own backward()
        self._saved = dict(final=final.clone(), grid=grid, w=w, seg_pairs=seg_pairs, seg_idx=seg_idx,
                           nseg=seg_pairs.shape[0], pulse_slots=pulse_slots, dH=dH, pops=pops,
                           schedules=schedules)
        return yields

    def saved_state(self):
        """The record the last forward() kept for backward(); pass it to backward() to differentiate that call."""
        return self._saved

    def backward(self, saved=None):
        """(field gradient (B, 3) per rad, rad, uT; pulse gradient (B, S) per uT or None; psi_0 recovery error).

        ``saved``: the record of the forward call to differentiate (``saved_state()``
        right after that call); by default the most recent forward call's.
        """
        sv = self._saved if saved is None else saved
        if sv is None:
            raise RuntimeError("backward() needs a forward() first")
        adj = adjoint_densities(self.base, sv["final"], sv["grid"], sv["w"], pairs=sv["seg_pairs"],
                                segment_of_interval=sv["seg_idx"], slot_of_interval=sv["seg_idx"], nslot=sv["nseg"],
                                decay_time=self.tau)
        fgrad = contract(sv["dH"], adj.densities.sum(axis=0))
        pgrad = None
        if sv["pulse_slots"]:
            Mp = 1e-3 * np.einsum("a,aij->ij", self.pulse_direction, electron_derivative_operators())
            pgrad = np.stack([np.einsum("ef,bef->b", Mp, adj.densities[s]).real for s in sv["pulse_slots"]], axis=1)
        x0 = self.initial_states
        if not nat.is_torch(x0):
            torch = nat.require_gpu()
            arr = np.asarray(x0, dtype=complex)
            x0 = torch.from_numpy(np.ascontiguousarray(arr.reshape(arr.shape[0], -1))).to(adj.initial_states.device)
        rec = float((adj.initial_states - x0.reshape(adj.initial_states.shape)).abs().amax().item())
        return fgrad, pgrad, rec


def yield_and_gradient(base, hamiltonian_for, propagator, initial_states, times_us, lifetime_us, thetas_deg,
                       phis_deg, fields_uT, *, pulse_amplitudes_uT=None, pulse_window_us=None,
                       pulse_direction=(1.0, 0.0, 0.0)) -> YieldGradient:
    """Yields of B trajectories and their gradients (one forward run, one backward sweep); see FieldYield."""
    fy = FieldYield(base, hamiltonian_for, propagator, initial_states, times_us, lifetime_us,
                    pulse_window_us=pulse_window_us, pulse_direction=pulse_direction)
    yields = fy.forward(thetas_deg, phis_deg, fields_uT, pulse_amplitudes_uT)
    fgrad, pgrad, rec = fy.backward()
    return YieldGradient(yields, fgrad, pgrad, fy._saved["pops"], rec)


@dataclass
class PulseAscentResult:
    amplitudes_uT: np.ndarray      # (S,) final pulse amplitudes
    mean_yields: np.ndarray        # (steps + 1,) ensemble-mean yield before each step and at the end
    gradients: np.ndarray          # (steps, S) ensemble-mean dPhi/da per uT at each step


def pulse_ascent(problem: FieldYield, theta_deg: float, phi_deg: float, field_uT: float, amplitudes_uT, *,
                 steps: int = 50, learning_rate: float = 1.0, bound_uT=None, callback=None) -> PulseAscentResult:
    """Gradient ascent of the ensemble-mean singlet yield over a piecewise-constant pulse.

    One amplitude vector (uT, one value per sub-interval of ``problem``'s pulse window) is shared by
    every random-phase state of ``problem``. Each step is one batched forward run and one adjoint
    sweep; the amplitudes move by ``learning_rate`` x the ensemble-mean gradient, clipped to
    +-``bound_uT`` when given. ``callback(step, mean_yield, amplitudes, gradient)`` runs after each step.
    """
    if problem.pulse_window is None:
        raise ValueError("pulse ascent needs a FieldYield with pulse_window_us")
    a = np.array(amplitudes_uT, dtype=float).reshape(-1)
    nb = int(problem.initial_states.shape[0])
    th, ph, fb = [float(theta_deg)] * nb, [float(phi_deg)] * nb, [float(field_uT)] * nb
    means, grads = [], []
    for step in range(int(steps)):
        y = problem.forward(th, ph, fb, np.tile(a, (nb, 1)))
        _, pgrad, _ = problem.backward()
        g = pgrad.mean(axis=0)
        means.append(float(y.mean()))
        grads.append(g)
        a = a + learning_rate * g
        if bound_uT is not None:
            a = np.clip(a, -float(bound_uT), float(bound_uT))
        if callback is not None:
            callback(step, means[-1], a, g)
    means.append(float(problem.forward(th, ph, fb, np.tile(a, (nb, 1))).mean()))
    return PulseAscentResult(a, np.array(means), np.array(grads).reshape(len(grads), a.size))


def singlet_yield(problem: FieldYield, field_params, pulse_amplitudes=None):
    """Differentiable yields: ``field_params`` (B, 3) float64 torch tensor of (theta [rad], phi [rad],
    |B| [uT]); ``pulse_amplitudes`` optional (B, S) tensor [uT].  Returns (B,) yields whose
    backward pass runs the adjoint sweep."""
    return _SingletYieldFunction.apply(field_params, pulse_amplitudes, problem)


def _autograd_function():
    import torch

    class SingletYieldFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, field_params, pulse_amplitudes, problem):
            p = field_params.detach().cpu().double().numpy()
            amps = None if pulse_amplitudes is None else pulse_amplitudes.detach().cpu().double().numpy()
            y = problem.forward(np.degrees(p[:, 0]), np.degrees(p[:, 1]), p[:, 2], amps)
            ctx.problem = problem
            # this call's own record: a later forward on the same problem must not change its gradient
            ctx.saved = problem.saved_state()
            ctx.has_pulse = pulse_amplitudes is not None
            return torch.from_numpy(y).to(field_params.device)

        @staticmethod
        def backward(ctx, grad_out):
            fgrad, pgrad, _ = ctx.problem.backward(ctx.saved)
            go = grad_out.detach().cpu().double().numpy()[:, None]
            gf = torch.from_numpy(go * fgrad).to(grad_out.device)
            gp = torch.from_numpy(go * pgrad).to(grad_out.device) if ctx.has_pulse else None
            return gf, gp, None

    return SingletYieldFunction


class _Lazy:
    def __init__(self):
        self._cls = None

    def apply(self, *args):
        if self._cls is None:
            self._cls = _autograd_function()
        return self._cls.apply(*args)


_SingletYieldFunction = _Lazy()
