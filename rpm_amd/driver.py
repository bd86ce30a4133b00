"""Trajectory driver and checkpoint wire format of the full-system quadrupole experiments.

Restates, with the trajectories propagated in batches on the GPU:

* the configuration record and its digest (R/rpm_full14_quadrupole_experiment.py:47-101):
  ``stable_dict`` has the reference's keys and values, and ``digest`` is the
  SHA-256 of its sorted-key JSON, so manifests and checkpoints written here are
  accepted by the reference's resume and report code;
* the uncoupled-control variant (R/rpm_full14_uncoupled_quadrupole.py:19-44);
* task keys, ordering and checkpoint paths (104-130), the atomic ``.npz`` cell
  with the keys and metadata of ``propagate`` (133-194), and the manifest
  (230-245);
* the checkpoint validation of ``read_cell`` (R/rpm_full14_uncoupled_quadrupole.py:47-59)
  and the ``raw_sample_yields.csv`` row of its report (150-162).
"""

from __future__ import annotations

import csv
import hashlib
import json
import math
import os
import platform
import socket
import time
from dataclasses import asdict, dataclass, field, replace
from pathlib import Path

import numpy as np

from . import core
from .batch import BatchedLanczosPropagator, pair_matrices
from .model import ALL_NUCLEI, FAD_NUCLEI, NUCLEAR_RADICALS, QUADRUPOLE_TENSORS_MT, TRP_NUCLEI

ANGLES = (0., 80., 85., 88., 89., 90., 91., 92., 95., 100., 180.)
PRIORITY = (0., 90., 85., 95., 88., 89., 91., 92., 80., 100., 180.)
SCALES = (0., 1.)
UNCOUPLED_ANGLES = tuple(float(a) for a in sorted(set(range(0, 181, 10)) | set(range(85, 96))))
FIELD_PLANE = "molecular zx plane, B=[sin(theta),0,cos(theta)]"


def _number_token(value: float) -> str:
    return f"{float(value):+.6f}".replace("+", "p").replace("-", "m").replace(".", "p")


def coupling_record(ratio_D_over_J: float | None) -> dict:
    """The reference's EfimovaHoreCondition.stable_dict() minus quadrupole_scale
    (R/rpm_quantum_needle_jd_sweeps.py:167-233)."""
    if ratio_D_over_J is None:
        rec = dict(condition_id="eh_D0_J0_control", family="efimova_hore", role="uncoupled_control",
                   d_paper_uT=0.0, exchange_J_uT=0.0, ratio_D_over_J=None)
    else:
        r = float(ratio_D_over_J)
        rec = dict(condition_id=f"eh_R_{_number_token(r)}", family="efimova_hore", role="paper_cancellation_scan",
                   d_paper_uT=core.R2_D_PAPER_UT, exchange_J_uT=core.R2_D_PAPER_UT / r, ratio_D_over_J=r)
    rec["dipolar_axis"] = (1.0, 0.0, 0.0)
    rec["convention"] = "Efimova-Hore Eqs. 9-10"
    rec["d_paper_MHz"] = core.equivalent_field_uT_to_MHz(rec["d_paper_uT"])
    rec["exchange_J_MHz"] = core.equivalent_field_uT_to_MHz(rec["exchange_J_uT"])
    return rec


@dataclass(frozen=True)
class ExperimentConfig:
    """R/rpm_full14_quadrupole_experiment.py:47-96 (the coupled D/J = -3 experiment)."""

    seed: int = core.MASTER_SEED
    samples: int = 4
    angles_deg: tuple = ANGLES
    field_mT: float = .05
    lifetime_us: float = 5.
    tail_bound: float = 1e-5
    integration_step_us: float = .00125
    coarse_step_us: float = .0025
    state_error_budget: float = 2.4e-5
    bootstrap_replicates: int = 2000
    smoke: bool = False

    ratio_D_over_J = -3.0          # class constants, not part of asdict()
    experiment = "full14_lowfield_quadrupole_off_on"

    @property
    def sweep(self) -> core.SweepConfig:
        return replace(core.SweepConfig(), seed=self.seed, field_microT=self.field_mT * 1000,
                       lifetime_us=self.lifetime_us, survival_tail_bound=self.tail_bound,
                       integration_step_us=self.integration_step_us,
                       coarse_integration_step_us=self.coarse_step_us, state_error_budget=self.state_error_budget)

    def model(self, scale: float) -> core.ProductionModel:
        return core.ProductionModel(master_seed=self.seed, lifetime_us=self.lifetime_us,
                                    quadrupole_scale=float(scale), ratio_D_over_J=self.ratio_D_over_J)

    def system(self):
        if self.smoke:
            return (FAD_NUCLEI[0], FAD_NUCLEI[1], TRP_NUCLEI[0]), ("FAD", "FAD", "TrpH")
        return ALL_NUCLEI, NUCLEAR_RADICALS

    def dimensions(self):
        nuclei, _ = self.system()
        return (2, 2) + tuple(n.dimension for n in nuclei)

    def stable_dict(self) -> dict:
        nuclei, radicals = self.system()
        return {**asdict(self), "schema_version": 1, "dimensions": list(self.dimensions()),
                "ordering": ["eA", "eB"] + [f"{r}:{n.name}" for n, r in zip(nuclei, radicals)],
                "full_coupled_pair": True, "factorized_evolution": False,
                "quadrupole_scales": SCALES,
                "quadrupole_tensors_mT": {k: v.tolist() for k, v in QUADRUPOLE_TENSORS_MT.items()},
                "quadrupole_orientation": "existing molecular-frame tensors; N5 in-plane azimuth provisional, not fitted",
                "hyperfine_tensors_mT": {f"{r}:{n.name}": n.tensor_mT.tolist() for n, r in zip(nuclei, radicals)},
                "coupling": coupling_record(self.ratio_D_over_J),
                "maximum_time_us": self.sweep.maximum_time_us,
                "krylov": {"min_dimension": 24, "max_dimension": 64, "increment": 8, "embedded_offset": 4,
                           "safety_factor": 2., "initial_step_us": .05, "max_step_us": .25,
                           "state_error_budget": self.state_error_budget},
                "initial_state": "electronic singlet times seeded nuclear random phase; common seeds across Q and angle",
                "relaxation": False, "nuclear_Zeeman": False}

    @property
    def digest(self) -> str:
        return hashlib.sha256(json.dumps(self.stable_dict(), sort_keys=True).encode()).hexdigest()


@dataclass(frozen=True)
class UncoupledExperimentConfig(ExperimentConfig):
    """R/rpm_full14_uncoupled_quadrupole.py:19-44: electron J = D = 0, 29 angles."""

    angles_deg: tuple = UNCOUPLED_ANGLES

    ratio_D_over_J = None
    experiment = "full14_lowfield_quadrupole_off_on"

    def stable_dict(self) -> dict:
        return {**super().stable_dict(), "full_coupled_pair": False, "full_pair_state_vector": True,
                "field_plane": FIELD_PLANE}


@dataclass(frozen=True, order=True)
class Task:
    """R/rpm_full14_quadrupole_experiment.py:104-122."""

    scale: float
    theta_deg: float
    sample_id: int

    @property
    def key(self) -> str:
        return f"Q{int(self.scale)}_th{self.theta_deg:07.3f}_s{self.sample_id:03d}"


def tasks(config: ExperimentConfig) -> list[Task]:
    priority = {a: i for i, a in enumerate(PRIORITY)}
    return sorted([Task(q, a, s) for s in range(config.samples) for a in config.angles_deg for q in SCALES],
                  key=lambda t: (t.sample_id, priority.get(t.theta_deg, 999), t.scale))


def cell_path(output: Path, task: Task) -> Path:
    return Path(output) / "checkpoints" / "cells" / (task.key + ".npz")


def atomic_npz(path: Path, values: dict) -> None:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_suffix(path.suffix + f".{os.getpid()}.tmp")
    with tmp.open("wb") as stream:
        np.savez_compressed(stream, **values)
    os.replace(tmp, path)


def atomic_json(path: Path, payload) -> None:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_suffix(path.suffix + f".{os.getpid()}.tmp")
    tmp.write_text(json.dumps(payload, indent=2, sort_keys=True) + "\n", encoding="utf-8")
    os.replace(tmp, path)


def write_csv(path: Path, rows: list[dict]) -> None:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    keys = list(rows[0].keys()) if rows else []
    with path.open("w", newline="", encoding="utf-8") as f:
        w = csv.DictWriter(f, fieldnames=keys)
        w.writeheader()
        w.writerows(rows)


def ensure_manifest(output: Path, config: ExperimentConfig) -> None:
    """R/rpm_full14_quadrupole_experiment.py:230-245 (source digests are of this package's files)."""
    output = Path(output)
    output.mkdir(parents=True, exist_ok=True)
    path = output / "manifest.json"
    if path.exists():
        old = json.loads(path.read_text())
        if old["configuration_hash"] != config.digest:
            raise ValueError("incompatible resume manifest")
        return
    here = Path(__file__).resolve().parent
    files = [here / "driver.py", here / "engine.py", here / "batch.py", here / "model.py", here / "core.py"]
    import scipy
    atomic_json(path, {"experiment": config.experiment, "configuration_hash": config.digest,
                       "configuration": config.stable_dict(),
                       "created_UTC": time.strftime("%Y-%m-%dT%H:%M:%SZ", time.gmtime()),
                       "environment": {"host": socket.gethostname(), "python": platform.python_version(),
                                       "numpy": np.__version__, "scipy": scipy.__version__,
                                       "backend": "rpm_amd (MI355X HIP)"},
                       "source_sha256": {p.name: hashlib.sha256(p.read_bytes()).hexdigest() for p in files}})
    write_csv(output / "task_inventory.csv", [{**asdict(t), "key": t.key} for t in tasks(config)])


def cell_values(task: Task, config: ExperimentConfig, initial: np.ndarray, result, *, elapsed: float,
                construction: float, initial_singlet: float) -> dict:
    """The checkpoint dictionary of R/rpm_full14_quadrupole_experiment.py:171-194 for one trajectory."""
    times = result.times_us
    if getattr(result, "device_yields", None) is not None:
        # summed on the device during the propagation (run_cells passes the fine/coarse weight table)
        fine, coarse, estimate = result.device_yields
    else:
        # a result propagated without yield weights: the same three integrals of its host arrays
        ids = core.coarse_grid_indices(config.sweep)
        table = core.yield_weight_table(times, config.lifetime_us, ids)
        fine = float(np.dot(table[:, 0], result.observable_values))
        coarse = float(np.dot(table[:, 1], result.observable_values))
        estimate = float(np.dot(table[:, 0], result.observable_error_bounds))
    stats = asdict(result.stats)
    metadata = {"task": asdict(task), "configuration_hash": config.digest, "elapsed_seconds": elapsed,
                "construction_seconds": construction, "host": socket.gethostname(), "pid": os.getpid(),
                "initial_state_sha256": hashlib.sha256(initial.tobytes()).hexdigest(), "stats": stats}
    final = result.final_state
    if final is None:
        final = result.final_state_device.cpu().numpy()
    return {"yield": np.array(fine), "coarse_yield": np.array(coarse),
            "quadrature_error": np.array(abs(fine - coarse)), "yield_propagation_estimate": np.array(estimate),
            "yield_tail_bound": np.array(config.tail_bound), "times_us": times,
            "singlet_population": result.observable_values,
            "population_error_estimate": result.observable_error_bounds,
            "state_error_estimates": result.state_error_estimates, "final_state": final,
            "initial_singlet": np.array(float(initial_singlet)),
            "metadata_json": np.array(json.dumps(metadata, sort_keys=True))}


def run_cells(config: ExperimentConfig, output: Path, todo=None, *, batch_size: int = 32, resume: bool = True,
              progress=None) -> list[Task]:
    """Propagate the tasks in batches and write one checkpoint cell per task.

    Tasks sharing a quadrupole scale share the nuclear Hamiltonian, so each
    batch is one scale with many (angle, sample) pairs carried by per-vector
    electron matrices.  Existing cells are skipped with ``resume``.
    Returns the tasks written.
    """
    output = Path(output)
    ensure_manifest(output, config)
    todo = list(tasks(config) if todo is None else todo)
    pending = [t for t in todo if not (resume and cell_path(output, t).exists())]
    nuclei, radicals = config.system()
    dims = config.dimensions()
    cfg = config.sweep
    grid = cfg.time_grid_us
    weights = core.yield_weight_table(grid, config.lifetime_us, core.coarse_grid_indices(cfg))
    written = []
    for scale in SCALES:
        group = [t for t in pending if t.scale == scale]
        if not group:
            continue
        model = config.model(scale)
        start = time.monotonic()
        base = core.build_hamiltonian(90.0, config.field_mT, model=model, nuclei=nuclei, radicals=radicals)
        prop = BatchedLanczosPropagator.like(core.make_propagator(base, cfg))
        for i in range(0, len(group), batch_size):
            chunk = group[i:i + batch_size]
            t0 = time.monotonic()
            hs = [core.build_hamiltonian(t.theta_deg, config.field_mT, model=model, nuclei=nuclei,
                                         radicals=radicals) for t in chunk]
            pairs = pair_matrices(base, hs)
            init = [core.random_phase_initial_state(dims, config.seed, t.sample_id) for t in chunk]
            construction = time.monotonic() - t0 + (time.monotonic() - start if i == 0 else 0.0)
            res = prop.propagate(np.stack(init), grid, pairs=pairs, yield_weights=weights)
            elapsed = time.monotonic() - t0
            for t, x0, r in zip(chunk, init, res):
                # the grid starts at t = 0, so the first population is the initial singlet population
                atomic_npz(cell_path(output, t), cell_values(t, config, x0, r, elapsed=elapsed / len(chunk),
                                                             construction=construction / len(chunk),
                                                             initial_singlet=r.observable_values[0]))
                written.append(t)
            if progress is not None:
                progress(len(written), len(pending))
    return written


def initial_hash(config: ExperimentConfig, sample: int) -> str:
    state = core.random_phase_initial_state(config.dimensions(), config.seed, sample)
    return hashlib.sha256(state.tobytes()).hexdigest()


def read_cell(root: Path, task: Task, config: ExperimentConfig):
    """Checkpoint validation of R/rpm_full14_uncoupled_quadrupole.py:47-59."""
    path = cell_path(root, task)
    with np.load(path, allow_pickle=False) as data:
        metadata = json.loads(str(data["metadata_json"]))
        if (metadata["configuration_hash"] != config.digest or metadata["task"] != asdict(task)
                or metadata["initial_state_sha256"] != initial_hash(config, task.sample_id)):
            raise ValueError(f"Incompatible checkpoint: {path}")
        values = {key: float(data[key]) for key in ("yield", "quadrature_error", "yield_propagation_estimate",
                                                    "yield_tail_bound")}
    if not all(math.isfinite(v) for v in values.values()):
        raise ValueError(f"Nonfinite checkpoint: {path}")
    return values, metadata


def raw_sample_rows(root: Path, config: ExperimentConfig, label: str) -> list[dict]:
    """Rows of raw_sample_yields.csv (R/rpm_full14_uncoupled_quadrupole.py:150-162)."""
    rows = []
    for task in tasks(config):
        values, meta = read_cell(root, task, config)
        rows.append({"electron_interactions": label, **asdict(task), **values,
                     "checkpoint": str(cell_path(root, task)), "initial_state_sha256": meta["initial_state_sha256"]})
    return rows
