"""Brute-force dense operators for small systems (independent of the matvec code).

Each LocalTerm is embedded by explicit index arithmetic on the mixed-radix
digits of the basis states, not by ``moveaxis`` / sparse products, so it checks
both matvec restatements (the pattern of R/tests/test_statevector_baseline.py:86-103).
"""

from __future__ import annotations

import math

import numpy as np
import scipy.linalg


def dense_from_terms(dimensions, terms) -> np.ndarray:
    dims = tuple(int(d) for d in dimensions)
    n = math.prod(dims)
    digits = np.array(np.unravel_index(np.arange(n), dims)).T          # (n, naxes)
    h = np.zeros((n, n), dtype=complex)
    for term in terms:
        axes = tuple(int(a) for a in term.axes)
        sub = tuple(dims[a] for a in axes)
        op = np.asarray(term.operator, dtype=complex)
        col_local = np.ravel_multi_index(tuple(digits[:, a] for a in axes), sub)
        for row_local in range(op.shape[0]):
            row_digits = digits.copy()
            parts = np.unravel_index(row_local, sub)
            for a, p in zip(axes, parts):
                row_digits[:, a] = p
            rows = np.ravel_multi_index(tuple(row_digits.T), dims)
            h[rows, np.arange(n)] += op[row_local, col_local]
    return h


def dense_propagate(h, state, times) -> np.ndarray:
    """States exp(-i H t) psi for each t (scipy expm, as R/tests/test_statevector_baseline.py:226-231)."""
    return np.array([scipy.linalg.expm(-1.0j * h * t) @ state for t in times])


def eigen_propagate(h, state, times) -> np.ndarray:
    """Same by eigendecomposition (faster for many times)."""
    w, q = np.linalg.eigh(h)
    c = q.conj().T @ state
    return (q @ (c[:, None] * np.exp(-1.0j * w[:, None] * np.asarray(times)[None, :]))).T


def dense_singlet_yield_closed_form(h, state, tau, t_final) -> float:
    """int_0^T e^{-t/tau}/tau P_S(t) dt in closed form for a pure state.

    P_S(t) = sum_{ab} c_a c_b^* <b|P|a> e^{-i(w_a - w_b)t}, integrated exactly
    (generalises R/poster_reproduction/core.py:455-473 to a many-level H).
    """
    w, q = np.linalg.eigh(h)
    c = q.conj().T @ state
    n = h.shape[0]
    e = q.reshape(2, 2, n // 4, n)
    a = (e[0, 1] - e[1, 0]) / math.sqrt(2.0)              # singlet amplitudes of eigvecs
    pm = a.conj().T @ a                                   # <b|P_S|a>
    k = 1.0 / tau
    decay = k + 1.0j * (w[:, None] - w[None, :])          # index (a, b)
    integral = k * (1.0 - np.exp(-decay * t_final)) / decay
    terms = (c[:, None] * c.conj()[None, :]) * pm.T * integral
    return float(terms.sum().real)


def dense_grid_yield(pieces, state, times, tau):
    """Simpson yield on an output grid with exact piecewise-constant dynamics (test oracle).

    ``pieces``: [(t_end, H_dense), ...] covering [0, times[-1]]; P_S(t_k) =
    ||(psi[01] - psi[10])/sqrt2||^2 of exp(-i H t) psi (eigendecomposition per
    piece); Phi = scipy.integrate.simpson(exp(-t/tau) P_S / tau, x=t)
    (R/rpm_quantum_needle_jd_sweeps.py:418-420).  Returns (Phi, P_S).
    """
    import scipy.integrate
    t = np.asarray(times, dtype=float)
    psi = np.asarray(state, dtype=complex).copy()
    n = psi.size
    pops = np.empty(t.size)
    start = 0.0
    k = 0
    for t_end, h in pieces:
        w, q = np.linalg.eigh(h)
        c = q.conj().T @ psi
        while k < t.size and t[k] <= t_end + 1e-12:
            v = q @ (c * np.exp(-1.0j * w * (t[k] - start)))
            s = (v[n // 4:n // 2] - v[n // 2:3 * n // 4]) / math.sqrt(2.0)
            pops[k] = float(np.vdot(s, s).real)
            k += 1
        psi = q @ (c * np.exp(-1.0j * w * (t_end - start)))
        start = t_end
    if k != t.size:
        raise ValueError("pieces do not cover the output grid")
    return float(scipy.integrate.simpson(np.exp(-t / tau) * pops / tau, x=t)), pops
