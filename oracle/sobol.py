"""Sobol reaction-time yield estimator and random-effects uncertainty, a CPU restatement.

Test infrastructure only: the checker for ``rpm_amd.sobol``, imported only by ``tests/``.
Follows R/rpm_quantum_needle_statevector.py:150-248 and 466-500. The schedule uses
scipy.stats.qmc.Sobol (scrambled, seeded), the same third-party routine as the reference. Configs
are duck-typed: any object with the reference's RunConfig fields.

Pinned by (tests/test_oracle_sobol.py):
- the reference's own tests (R/tests/test_statevector_baseline.py:179-191): schedules are nested in
  Q, independent across samples, and the uncertainty is deterministic;
- the net property of scrambled Sobol points: each of the Q intervals [k/Q, (k+1)/Q) holds exactly
  one u per scramble;
- the inverse-CDF map: exp(-T / tau) recovers 1 - u;
- the uncertainty on a hand-computed example (variance components and Student-t quantiles from
  their definitions) and by coverage: on data with known variance components, the 95% joint
  interval covers the true mean about 95% of the time;
- the cell estimator on a small system: the Sobol mean converges to the exact Haberkorn yield
  k int e^{-kt} P(t) dt of the dense propagator.
"""

from __future__ import annotations

import math

import numpy as np
from scipy.stats import qmc
from scipy.stats import t as student_t

from .engine import AdaptiveLanczosPropagator, singlet_population


def seed(base, *keys) -> int:
    """R/rpm_quantum_needle_statevector.py:150-152."""
    return int(np.random.SeedSequence([int(base), *(int(k) for k in keys)]).generate_state(1, dtype=np.uint32)[0])


def sobol_uniforms(config, q_points, sample_id, scramble) -> np.ndarray:
    """Q scrambled Sobol points in [0, 1), clipped below 1 (R/rpm_quantum_needle_statevector.py:165-177)."""
    engine = qmc.Sobol(d=1, scramble=True, seed=seed(
        config.seed, 0x50B01, sample_id if config.independent_sobol_by_nuclear_sample else 0, scramble))
    u = engine.random_base2(int(math.log2(q_points)))[:, 0]
    return np.minimum(u, np.nextafter(1.0, 0.0))


def reaction_time_schedule(config, q_points, sample_id=0) -> np.ndarray:
    """(lifetimes, scrambles, Q): T = -tau log(1 - u) (R/rpm_quantum_needle_statevector.py:155-181)."""
    if q_points < 1 or q_points & (q_points - 1):
        raise ValueError("The Sobol point count Q must be a positive power of two.")
    out = np.empty((len(config.lifetimes_us), config.sobol_scrambles, q_points))
    for s in range(config.sobol_scrambles):
        u = sobol_uniforms(config, q_points, sample_id, s)
        for i, tau in enumerate(config.lifetimes_us):
            out[i, s] = -float(tau) * np.log1p(-u)
    return out


def run_cell(config, hamiltonian, initial_state, q_points, sample_id, trace_times=None):
    """Yields (scrambles, lifetimes) and propagation bounds of one cell, and the propagation result.

    R/rpm_quantum_needle_statevector.py:184-248 with the Hamiltonian and state given.
    """
    schedule = reaction_time_schedule(config, q_points, sample_id)
    reaction = schedule.reshape(-1)
    trace = np.empty(0) if trace_times is None else np.asarray(trace_times, float)
    obs = np.unique(np.concatenate(([0.0], reaction, trace)))
    prop = AdaptiveLanczosPropagator(
        hamiltonian, minimum_dimension=config.lanczos_minimum_dimension,
        maximum_dimension=config.lanczos_maximum_dimension, dimension_increment=config.lanczos_dimension_increment,
        embedded_offset=config.lanczos_embedded_offset, safety_factor=config.lanczos_safety_factor,
        state_error_budget=config.lanczos_state_error_budget, initial_step_us=config.lanczos_initial_step_us,
        maximum_step_us=config.lanczos_maximum_step_us)
    dims = hamiltonian.dimensions
    res = prop.propagate(initial_state, obs, lambda v: singlet_population(v, dims))
    idx = np.searchsorted(obs, reaction)
    pops = res.observable_values[idx].reshape(schedule.shape)
    bounds = res.observable_error_bounds[idx].reshape(schedule.shape)
    return pops.mean(axis=2).T, bounds.mean(axis=2).T, res


def uncertainty_components(values):
    """(mean, nuclear, reaction-time, joint half-widths, nuclear sample std) of an n x s table
    (R/rpm_quantum_needle_statevector.py:466-500)."""
    v = np.asarray(values, dtype=float)
    if v.ndim != 2 or min(v.shape) < 2:
        raise ValueError("Uncertainty analysis needs at least two rows and columns.")
    n, s = v.shape
    rows = v.mean(axis=1)
    var_nuc = rows.var(ddof=1) / n
    var_time = v.var(axis=1, ddof=1).mean() / (n * s)
    dof_t = n * (s - 1)
    return (float(v.mean()),
            float(student_t.ppf(0.975, n - 1)) * math.sqrt(max(0.0, var_nuc)),
            float(student_t.ppf(0.975, dof_t)) * math.sqrt(max(0.0, var_time)),
            float(student_t.ppf(0.975, min(n - 1, dof_t))) * math.sqrt(max(0.0, var_nuc + var_time)),
            float(rows.std(ddof=1)))
