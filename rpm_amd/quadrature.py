"""Simpson yield quadrature with SciPy >= 1.11 semantics, as explicit weights.

The reference integrates the singlet yield with ``scipy.integrate.simpson``
(R/rpm_quantum_needle_jd_sweeps.py:418-420).  Writing that rule as a weight
vector w (Phi = sum_i w_i f_i) lets the device accumulate the yield as
populations are produced.  The rule restated here is SciPy's:

* odd point count: composite Simpson on consecutive interval pairs with the
  non-uniform three-point formula
  hsum/6 * (y0 (2 - h1/h0) + y1 hsum^2/(h0 h1) + y2 (2 - h0/h1));
* even point count >= 4: that rule on the first N-1 points, plus the
  last-interval correction alpha*y[-1] + beta*y[-2] - eta*y[-3] with
  alpha = (2 h1^2 + 3 h0 h1) / (6 (h0 + h1)), beta = (h1^2 + 3 h0 h1) / (6 h0),
  eta = h1^3 / (6 h0 (h0 + h1)), h0, h1 the last two spacings;
* two points: the trapezoid.
"""

from __future__ import annotations

import numpy as np


def simpson_weights(x) -> np.ndarray:
    x = np.asarray(x, dtype=float)
    n = x.size
    w = np.zeros(n)
    if n < 2:
        return w
    if n == 2:
        h = x[1] - x[0]
        w[:] = 0.5 * h
        return w
    stop = n - 1 if n % 2 == 1 else n - 2          # last index covered by the pair rule
    h = np.diff(x)
    h0 = h[0:stop:2]
    h1 = h[1:stop:2]
    hsum = h0 + h1
    ratio = np.divide(h0, h1, out=np.zeros_like(h0), where=h1 != 0)
    inv_ratio = np.divide(1.0, ratio, out=np.zeros_like(ratio), where=ratio != 0)
    quot = np.divide(hsum, h0 * h1, out=np.zeros_like(hsum), where=(h0 * h1) != 0)
    np.add.at(w, np.arange(0, stop - 1, 2), hsum / 6.0 * (2.0 - inv_ratio))
    np.add.at(w, np.arange(1, stop, 2), hsum / 6.0 * (hsum * quot))
    np.add.at(w, np.arange(2, stop + 1, 2), hsum / 6.0 * (2.0 - ratio))
    if n % 2 == 0:
        a, b = h[-2], h[-1]
        den_a, den_b, den_e = 6.0 * (b + a), 6.0 * a, 6.0 * a * (a + b)
        w[-1] += (2.0 * b * b + 3.0 * a * b) / den_a if den_a != 0 else 0.0
        w[-2] += (b * b + 3.0 * a * b) / den_b if den_b != 0 else 0.0
        w[-3] -= (b ** 3) / den_e if den_e != 0 else 0.0
    return w


def yield_weights(times_us, lifetime_us) -> np.ndarray:
    """Weights of Phi_S = simpson(exp(-t/tau) P_S(t) / tau, t)."""
    t = np.asarray(times_us, dtype=float)
    return simpson_weights(t) * (np.exp(-t / lifetime_us) / lifetime_us)


def yield_weight_table(times_us, lifetime_us, coarse_indices=None) -> np.ndarray:
    """(n, 2) weights for the device yield accumulation (``rpm_yield_accumulate``).

    Column 0: ``yield_weights`` of the grid (Phi_S, and with the population bounds the propagation
    estimate). Column 1: the same rule on the coarse sub-grid ``times_us[coarse_indices]``, zero at
    the other points (R/rpm_quantum_needle_jd_sweeps.py:423-431 and the coarse yield of
    R/rpm_full14_quadrupole_experiment.py:179-183); all zero without coarse indices.
    """
    t = np.asarray(times_us, dtype=float)
    table = np.zeros((t.size, 2))
    table[:, 0] = yield_weights(t, lifetime_us)
    if coarse_indices is not None:
        idx = np.asarray(coarse_indices, dtype=int)
        table[idx, 1] = yield_weights(t[idx], lifetime_us)
    return table


def integrate_yield(times_us, populations, lifetime_us) -> float:
    """Phi_S for one population trace (R/rpm_quantum_needle_jd_sweeps.py:418-420)."""
    return float(np.dot(yield_weights(times_us, lifetime_us), np.asarray(populations, dtype=float)))


def trapezoid_weights(x) -> np.ndarray:
    """Weights of the composite trapezoid rule on the nodes x: half of each adjacent spacing."""
    x = np.asarray(x, dtype=float)
    w = np.zeros(x.size)
    if x.size < 2:
        return w
    h = np.diff(x)
    w[:-1] += 0.5 * h
    w[1:] += 0.5 * h
    return w


def integrate_singlet_yield(times_us, singlet_population_values, lifetime_us, *, method: str = "simpson") -> float:
    """Phi_S of an already-propagated trajectory (R/rpm_quantum_needle_validation.py:192-210):
    the integral of exp(-t/tau) P_S(t) / tau by Simpson's rule (as ``integrate_yield``) or by the
    trapezoid rule. ValueError when the times and populations are not matching 1-D arrays, or for
    another method."""
    times = np.asarray(times_us, dtype=float)
    populations = np.asarray(singlet_population_values, dtype=float)
    if times.shape != populations.shape or times.ndim != 1:
        raise ValueError("Yield times and singlet populations must be matching 1-D arrays.")
    if method == "simpson":
        weights = simpson_weights(times)
    elif method == "trapezoid":
        weights = trapezoid_weights(times)
    else:
        raise ValueError(f"Unknown yield-integration method: {method}")
    return float(np.dot(weights * (np.exp(-times / lifetime_us) / lifetime_us), populations))
