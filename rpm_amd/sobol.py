"""Sobol reaction-time estimator of the singlet yield, and its random-effects uncertainty.

Drop-in for R/rpm_quantum_needle_statevector.py:65-330, 466-500. The yield
Phi_S(tau) = E[P_S(T)] with T ~ Exp(tau) is estimated by evaluating the propagated singlet population
at scrambled-Sobol reaction times T = -tau log(1 - u), which are off the uniform time grid. There are
Q points per scramble and `sobol_scrambles` independent scrambles per nuclear sample. One
propagation per (angle, sample) cell serves every lifetime and scramble.

On the GPU, a cell is one adaptive Lanczos trajectory through the device propagator, with the singlet
population projected in the Krylov basis at every reaction time. ``run_cells`` batches many cells,
each with its own field angle (as a per-trajectory electron matrix) and its own reaction times, through
``BatchedLanczosPropagator``. Each trajectory keeps the step rules of a separate run.
"""

from __future__ import annotations

import hashlib
import json
import math
import os
import time
from dataclasses import asdict, dataclass
from pathlib import Path
from typing import Sequence

import numpy as np
from scipy.stats import qmc
from scipy.stats import t as student_t

from .engine import (AdaptiveLanczosPropagator, SingletPopulationObservable, build_quantum_needle_hamiltonian,
                     singlet_initial_state)
from .model import (DEFAULT_FIELD_MICROT, DEFAULT_LIFETIMES_US, NOMINAL_QUADRUPOLE_SCALE, SPIN_ORDER,
                    SYSTEM_DIMENSIONS, diagnostic_time_grid_us, nominal_quadrupole_metadata)


@dataclass(frozen=True)
class RunConfig:
    """Estimator configuration; fields and defaults of R/rpm_quantum_needle_statevector.py:65-108."""

    backend: str = "full"
    exchange_J_MHz: float = 0.0
    dipolar_tensor_MHz: tuple = ((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0))
    independent_sobol_by_nuclear_sample: bool = True
    field_microT: float = DEFAULT_FIELD_MICROT
    lifetimes_us: tuple = DEFAULT_LIFETIMES_US
    seed: int = 20260713
    sobol_scrambles: int = 8
    bootstrap_replicates: int = 2000
    stochastic_halfwidth_target: float = 2e-4
    propagation_bound_target: float = 5e-5
    pilot_start_m: int = 8
    pilot_start_q: int = 32
    pilot_max_m: int = 2048
    pilot_max_q: int = 256
    lanczos_minimum_dimension: int = 24
    lanczos_maximum_dimension: int = 64
    lanczos_dimension_increment: int = 8
    lanczos_embedded_offset: int = 4
    lanczos_safety_factor: float = 2.0
    lanczos_state_error_budget: float = 2.4e-5
    lanczos_initial_step_us: float = 0.05
    lanczos_maximum_step_us: float = 0.25
    quadrupole_scale: float = NOMINAL_QUADRUPOLE_SCALE

    def stable_dict(self) -> dict:
        result = asdict(self)
        if self.quadrupole_scale == 0:
            result.pop("quadrupole_scale")
        result.update(nominal_quadrupole_metadata(self.quadrupole_scale))
        result["lifetimes_us"] = list(self.lifetimes_us)
        result["spin_order"] = list(SPIN_ORDER)
        result["dimensions"] = list(SYSTEM_DIMENSIONS)
        return result

    @property
    def configuration_hash(self) -> str:
        payload = json.dumps(self.stable_dict(), sort_keys=True, separators=(",", ":"))
        return hashlib.sha256(payload.encode("utf-8")).hexdigest()


@dataclass(frozen=True)
class CellTask:
    theta_deg: float
    sample_id: int
    q_points: int
    include_trace: bool
    config: RunConfig


@dataclass
class CellResult:
    """One (angle, nuclear sample) cell (R/rpm_quantum_needle_statevector.py:120-139).

    yields and propagation_bounds: (sobol_scrambles, lifetimes), each the mean over the Q points.
    """

    theta_deg: float
    sample_id: int
    q_points: int
    yields: np.ndarray
    propagation_bounds: np.ndarray
    trace_times_us: np.ndarray
    trace_population: np.ndarray
    trace_bounds: np.ndarray
    accepted_steps: int
    rejected_steps: int
    matvecs: int
    maximum_krylov_dimension: int
    minimum_step_us: float
    maximum_step_us: float
    accumulated_state_error: float
    maximum_norm_error: float
    elapsed_seconds: float


@dataclass(frozen=True)
class Uncertainty:
    mean: float
    nuclear_halfwidth: float
    reaction_time_halfwidth: float
    joint_halfwidth: float
    nuclear_sample_std: float


def _seed(base: int, *keys: int) -> int:
    """First 32-bit word of SeedSequence([base, *keys]) (R/rpm_quantum_needle_statevector.py:150-152)."""
    sequence = np.random.SeedSequence([int(base), *(int(k) for k in keys)])
    return int(sequence.generate_state(1, dtype=np.uint32)[0])


def reaction_time_schedule(config: RunConfig, q_points: int, sample_id: int = 0) -> np.ndarray:
    """(lifetimes, scrambles, Q) reaction times -tau log(1 - u) from scrambled Sobol points u
    (R/rpm_quantum_needle_statevector.py:155-181). Q must be a power of two; a schedule at Q is the
    first Q points of the schedule at 2Q (nested)."""
    if q_points < 1 or q_points & (q_points - 1):
        raise ValueError("The Sobol point count Q must be a positive power of two.")
    exponent = int(math.log2(q_points))
    schedule = np.empty((len(config.lifetimes_us), config.sobol_scrambles, q_points), dtype=float)
    for scramble in range(config.sobol_scrambles):
        engine = qmc.Sobol(d=1, scramble=True, seed=_seed(
            config.seed, 0x50B01, sample_id if config.independent_sobol_by_nuclear_sample else 0, scramble))
        uniform = engine.random_base2(exponent)[:, 0]
        uniform = np.minimum(uniform, np.nextafter(1.0, 0.0))
        for index, lifetime in enumerate(config.lifetimes_us):
            schedule[index, scramble, :] = -float(lifetime) * np.log1p(-uniform)
    return schedule


def cell_hamiltonian(theta_deg: float, config: RunConfig):
    """R/rpm_quantum_needle_statevector.py:193-199."""
    return build_quantum_needle_hamiltonian(theta_deg, config.field_microT, exchange_J_MHz=config.exchange_J_MHz,
                                            dipolar_tensor_MHz=np.asarray(config.dipolar_tensor_MHz),
                                            quadrupole_scale=config.quadrupole_scale)


def cell_initial_state(config: RunConfig, sample_id: int, dimensions=SYSTEM_DIMENSIONS) -> np.ndarray:
    """Singlet x random-phase nuclear state of the sample (R/rpm_quantum_needle_statevector.py:200-201)."""
    rng = np.random.default_rng(_seed(config.seed, 0xA11CE, sample_id))
    return singlet_initial_state(dimensions, rng)


def _propagator_kwargs(config: RunConfig) -> dict:
    return dict(minimum_dimension=config.lanczos_minimum_dimension,
                maximum_dimension=config.lanczos_maximum_dimension,
                dimension_increment=config.lanczos_dimension_increment,
                embedded_offset=config.lanczos_embedded_offset, safety_factor=config.lanczos_safety_factor,
                state_error_budget=config.lanczos_state_error_budget,
                initial_step_us=config.lanczos_initial_step_us, maximum_step_us=config.lanczos_maximum_step_us)


def _cell_times(task: CellTask):
    schedule = reaction_time_schedule(task.config, task.q_points, task.sample_id)
    reaction = schedule.reshape(-1)
    trace = diagnostic_time_grid_us() if task.include_trace else np.empty(0)
    observation = np.unique(np.concatenate(([0.0], reaction, trace)))
    return schedule, reaction, trace, observation


def _cell_result(task, schedule, reaction, trace, observation, propagated, elapsed) -> CellResult:
    """R/rpm_quantum_needle_statevector.py:223-248."""
    idx = np.searchsorted(observation, reaction)
    populations = np.asarray(propagated.observable_values)[idx].reshape(schedule.shape)
    bounds = np.asarray(propagated.observable_error_bounds)[idx].reshape(schedule.shape)
    if task.include_trace:
        tidx = np.searchsorted(observation, trace)
        trace_population = np.asarray(propagated.observable_values)[tidx]
        trace_bounds = np.asarray(propagated.observable_error_bounds)[tidx]
    else:
        trace_population = np.empty(0)
        trace_bounds = np.empty(0)
    st = propagated.stats
    return CellResult(theta_deg=task.theta_deg, sample_id=task.sample_id, q_points=task.q_points,
                      yields=populations.mean(axis=2).T, propagation_bounds=bounds.mean(axis=2).T,
                      trace_times_us=trace, trace_population=trace_population, trace_bounds=trace_bounds,
                      accepted_steps=st.accepted_steps, rejected_steps=st.rejected_steps, matvecs=st.matvecs,
                      maximum_krylov_dimension=st.maximum_krylov_dimension, minimum_step_us=st.minimum_step_us,
                      maximum_step_us=st.maximum_step_us, accumulated_state_error=st.accumulated_state_error,
                      maximum_norm_error=st.maximum_norm_error, elapsed_seconds=elapsed)


def run_cell_full(task: CellTask, *, hamiltonian=None, initial_state=None) -> CellResult:
    """One cell on the GPU (R/rpm_quantum_needle_statevector.py:184-248).

    `hamiltonian` and `initial_state` (extensions) replace the full-system ones, e.g. for small
    test systems; by default they are built from the task as the reference does.
    """
    started = time.perf_counter()
    schedule, reaction, trace, observation = _cell_times(task)
    h = cell_hamiltonian(task.theta_deg, task.config) if hamiltonian is None else hamiltonian
    state = cell_initial_state(task.config, task.sample_id, h.dimensions) if initial_state is None else initial_state
    prop = AdaptiveLanczosPropagator(h, **_propagator_kwargs(task.config))
    propagated = prop.propagate(state, observation, SingletPopulationObservable(h.dimensions))
    return _cell_result(task, schedule, reaction, trace, observation, propagated, time.perf_counter() - started)


def run_cell(task: CellTask, **kwargs) -> CellResult:
    """R/rpm_quantum_needle_statevector.py:251-254."""
    if task.config.backend != "full":
        raise ValueError("Only the full coupled-space backend is permitted.")
    return run_cell_full(task, **kwargs)


def run_cells(tasks: Sequence[CellTask], *, batch_size: int = 32, hamiltonian_for=None, initial_state_for=None,
              progress_callback=None) -> list:
    """Many cells on the GPU, `batch_size` trajectories per batched propagation.

    Cells of one batch must share the configuration; they may differ in angle (the electron matrix
    of each trajectory) and in sample (initial state and reaction times). Results are in task order,
    each equal to ``run_cell_full`` of the task. `elapsed_seconds` is the batch's wall time divided
    among its cells. `hamiltonian_for(task)` and `initial_state_for(task)` (extensions) replace the
    full-system builders.
    """
    from .batch import BatchedLanczosPropagator, pair_matrices

    tasks = list(tasks)
    for t in tasks:
        if t.config.backend != "full":
            raise ValueError("Only the full coupled-space backend is permitted.")
    out: list = [None] * len(tasks)
    order = sorted(range(len(tasks)), key=lambda i: tasks[i].config.configuration_hash)
    groups: list = []
    for i in order:
        if groups and tasks[groups[-1][0]].config == tasks[i].config and len(groups[-1]) < batch_size:
            groups[-1].append(i)
        else:
            groups.append([i])
    build = hamiltonian_for or (lambda task: cell_hamiltonian(task.theta_deg, task.config))
    for group in groups:
        started = time.perf_counter()
        cfg = tasks[group[0]].config
        hams = [build(tasks[i]) for i in group]
        base = hams[0]
        pairs = pair_matrices(base, hams)
        states = np.stack([
            (initial_state_for(tasks[i]) if initial_state_for is not None
             else cell_initial_state(cfg, tasks[i].sample_id, base.dimensions)) for i in group])
        timing = [_cell_times(tasks[i]) for i in group]
        bp = BatchedLanczosPropagator(base, **_propagator_kwargs(cfg))
        cb = None if progress_callback is None else (lambda b, *a, g=group: progress_callback(tasks[g[b]], *a))
        results = bp.propagate(states, [t[3] for t in timing], pairs=pairs, progress_callback=cb)
        elapsed = (time.perf_counter() - started) / len(group)
        for k, i in enumerate(group):
            out[i] = _cell_result(tasks[i], *timing[k], results[k], elapsed)
    return out


def _angle_token(theta_deg: float) -> str:
    return f"{theta_deg:07.3f}".replace("-", "m").replace(".", "p")


def checkpoint_path(output_dir: Path, theta_deg: float, sample_id: int, q_points: int) -> Path:
    """R/rpm_quantum_needle_statevector.py:261-266."""
    return (Path(output_dir) / "checkpoints"
            / f"theta_{_angle_token(theta_deg)}_sample_{sample_id:04d}_q_{q_points:04d}.npz")


_CELL_SCALARS = (("accepted_steps", int), ("rejected_steps", int), ("matvecs", int),
                 ("maximum_krylov_dimension", int), ("minimum_step_us", float), ("maximum_step_us", float),
                 ("accumulated_state_error", float), ("maximum_norm_error", float), ("elapsed_seconds", float))
_CELL_ARRAYS = ("yields", "propagation_bounds", "trace_times_us", "trace_population", "trace_bounds")


def save_cell(result: CellResult, output_dir: Path, configuration_hash: str) -> Path:
    """Atomic compressed .npz with the reference's keys (R/rpm_quantum_needle_statevector.py:269-296)."""
    path = checkpoint_path(output_dir, result.theta_deg, result.sample_id, result.q_points)
    path.parent.mkdir(parents=True, exist_ok=True)
    temporary = path.with_suffix(path.suffix + f".{os.getpid()}.tmp")
    payload = {"configuration_hash": np.array(configuration_hash), "theta_deg": np.array(result.theta_deg),
               "sample_id": np.array(result.sample_id), "q_points": np.array(result.q_points)}
    for key in _CELL_ARRAYS:
        payload[key] = getattr(result, key)
    for key, _ in _CELL_SCALARS:
        payload[key] = np.array(getattr(result, key))
    with temporary.open("wb") as handle:
        np.savez_compressed(handle, **payload)
    os.replace(temporary, path)
    return path


def load_cell(output_dir: Path, theta_deg: float, sample_id: int, q_points: int,
              configuration_hash: str) -> CellResult:
    """R/rpm_quantum_needle_statevector.py:299-329; a configuration mismatch raises RuntimeError."""
    path = checkpoint_path(output_dir, theta_deg, sample_id, q_points)
    with np.load(path, allow_pickle=False) as data:
        if str(data["configuration_hash"].item()) != configuration_hash:
            raise RuntimeError(f"Checkpoint configuration mismatch: {path}")
        kw = {key: data[key].copy() for key in _CELL_ARRAYS}
        kw.update({key: cast(data[key]) for key, cast in _CELL_SCALARS})
        return CellResult(theta_deg=float(data["theta_deg"]), sample_id=int(data["sample_id"]),
                          q_points=int(data["q_points"]), **kw)


def uncertainty_components(values, replicates: int, seed: int) -> Uncertainty:
    """Random-effects 95% intervals for nuclear samples (rows) x Sobol scrambles (columns)
    (R/rpm_quantum_needle_statevector.py:466-500).

    Between-sample variance of the row means over n, within-row variance over n s, Student-t factors
    with n - 1 and n (s - 1) degrees of freedom (the joint interval uses the smaller). `replicates`
    and `seed` are accepted for the reference's signature and unused, as there.
    """
    values = np.asarray(values, dtype=float)
    if values.ndim != 2 or min(values.shape) < 2:
        raise ValueError("Uncertainty analysis needs at least two rows and columns.")
    del replicates, seed
    n, s = values.shape
    center = float(values.mean())
    row_means = values.mean(axis=1)
    nuclear_variance = float(row_means.var(ddof=1) / n)
    within = float(values.var(axis=1, ddof=1).mean())
    time_variance = within / (n * s)
    nuclear_factor = float(student_t.ppf(0.975, n - 1))
    time_dof = n * (s - 1)
    time_factor = float(student_t.ppf(0.975, time_dof))
    joint_factor = float(student_t.ppf(0.975, min(n - 1, time_dof)))
    return Uncertainty(mean=center,
                       nuclear_halfwidth=nuclear_factor * math.sqrt(max(0.0, nuclear_variance)),
                       reaction_time_halfwidth=time_factor * math.sqrt(max(0.0, time_variance)),
                       joint_halfwidth=joint_factor * math.sqrt(max(0.0, nuclear_variance + time_variance)),
                       nuclear_sample_std=float(row_means.std(ddof=1)))
