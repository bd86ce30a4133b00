"""Production Hamiltonian, seeds, Krylov settings, grids and yields.

Drop-in for the solver-facing parts of R/poster_reproduction/core.py:69-258,
R/rpm_quantum_needle_jd_sweeps.py:81-102,167-235,306-334,373-375,418-431 and
R/rpm_full14_quadrupole_experiment.py:138-194.  Hamiltonians are GPU-backed
``CompiledFullPairHamiltonian`` objects; propagation is the GPU
``AdaptiveLanczosPropagator``.
"""

from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass, replace

import numpy as np

from .engine import (
    SINGLET, AdaptiveLanczosPropagator, CompiledFullPairHamiltonian, LocalTerm, SingletPopulationObservable,
    efimova_hore_electron_coupling_operator, quadrupole_operator, singlet_population, two_spin_coupling_operator,
)
from .model import (
    ALL_NUCLEI, GAMMA_E_MHZ_PER_MT, NUCLEAR_RADICALS, QUADRUPOLE_TENSORS_MT, TWOPI, field_vector_mT,
    spin_operators,
)
from .quadrature import (  # noqa: F401  (re-exported)
    integrate_singlet_yield, integrate_yield, simpson_weights, yield_weight_table, yield_weights,
)

MASTER_SEED = 20260731
R2_D_PAPER_UT = -2783.0 / (2.0 ** 3)


def equivalent_field_uT_to_MHz(value_uT: float) -> float:
    return float(value_uT) * GAMMA_E_MHZ_PER_MT / 1000.0


@dataclass(frozen=True)
class EfimovaHoreCondition:
    """Electron couplings in the Efimova-Hore convention (R/rpm_quantum_needle_jd_sweeps.py:167-199)."""

    d_paper_uT: float
    exchange_J_uT: float
    dipolar_axis: tuple = (1.0, 0.0, 0.0)

    @property
    def d_paper_MHz(self) -> float:
        return equivalent_field_uT_to_MHz(self.d_paper_uT)

    @property
    def exchange_J_MHz(self) -> float:
        return equivalent_field_uT_to_MHz(self.exchange_J_uT)


def efimova_hore_ratio_condition(ratio: float) -> EfimovaHoreCondition:
    r = float(ratio)
    if r == 0.0:
        raise ValueError("Efimova-Hore D/J ratio cannot be zero.")
    return EfimovaHoreCondition(R2_D_PAPER_UT, R2_D_PAPER_UT / r)


def uncoupled_condition() -> EfimovaHoreCondition:
    return EfimovaHoreCondition(0.0, 0.0)


@dataclass(frozen=True)
class SweepConfig:
    """R/rpm_quantum_needle_jd_sweeps.py:306-334."""

    seed: int = 20260715
    field_microT: float = 50.0
    lifetime_us: float = 2.0
    survival_tail_bound: float = 1e-4
    integration_step_us: float = 0.00125
    coarse_integration_step_us: float = 0.0025
    state_error_budget: float = 2.4e-5
    lanczos_minimum_dimension: int = 24
    lanczos_maximum_dimension: int = 64
    lanczos_dimension_increment: int = 8
    lanczos_embedded_offset: int = 4
    lanczos_safety_factor: float = 2.0
    lanczos_initial_step_us: float = 0.05
    lanczos_maximum_step_us: float = 0.25
    quadrupole_scale: float = 1.0

    @property
    def maximum_time_us(self) -> float:
        return -self.lifetime_us * math.log(self.survival_tail_bound)

    @property
    def time_grid_us(self) -> np.ndarray:
        n = int(math.floor(self.maximum_time_us / self.integration_step_us))
        grid = np.arange(n + 1, dtype=float) * self.integration_step_us
        if grid[-1] < self.maximum_time_us - 1e-13:
            grid = np.append(grid, self.maximum_time_us)
        return grid


@dataclass(frozen=True)
class ProductionModel:
    """R/poster_reproduction/core.py:69-101 (solver-facing fields)."""

    master_seed: int = MASTER_SEED
    lifetime_us: float = 5.0
    quadrupole_scale: float = 1.0
    ratio_D_over_J: float | None = -3.0

    @property
    def condition(self) -> EfimovaHoreCondition:
        if self.ratio_D_over_J is None:
            return uncoupled_condition()
        return efimova_hore_ratio_condition(self.ratio_D_over_J)

    @property
    def sweep_config(self) -> SweepConfig:
        return replace(SweepConfig(), seed=self.master_seed, lifetime_us=self.lifetime_us,
                       quadrupole_scale=self.quadrupole_scale)


PRODUCTION_MODEL = ProductionModel()


def sample_seed(master_seed: int, sample_id: int) -> int:
    """uint32 of SeedSequence([seed, 0xA11CE, id]) (R/rpm_quantum_needle_jd_sweeps.py:373-375)."""
    seq = np.random.SeedSequence([int(master_seed), 0xA11CE, int(sample_id)])
    return int(seq.generate_state(1, dtype=np.uint32)[0])


def random_phase_initial_state(dimensions, master_seed: int, sample_id: int) -> np.ndarray:
    """Host-generated, bit-identical to R/poster_reproduction/core.py:250-258."""
    rng = np.random.default_rng(sample_seed(master_seed, sample_id))
    nn = math.prod(dimensions[2:])
    phases = rng.uniform(0.0, TWOPI, nn)
    return np.kron(SINGLET, np.exp(1.0j * phases) / math.sqrt(nn))


def initial_state_sha256(state) -> str:
    return hashlib.sha256(np.asarray(state, dtype=complex).tobytes()).hexdigest()


def electron_coupling_term(condition: EfimovaHoreCondition) -> LocalTerm:
    """R/poster_reproduction/core.py:154-161."""
    op = TWOPI * efimova_hore_electron_coupling_operator(condition.exchange_J_MHz, condition.d_paper_MHz,
                                                         np.asarray(condition.dipolar_axis, float))
    return LocalTerm((0, 1), op, "Efimova-Hore exchange + full nonsecular dipolar")


def zeeman_operator(field_mT_vector) -> np.ndarray:
    ops = spin_operators(0.5)
    b = np.asarray(field_mT_vector, dtype=float)
    return TWOPI * GAMMA_E_MHZ_PER_MT * (b[0] * ops[0] + b[1] * ops[1] + b[2] * ops[2])


def build_hamiltonian(theta_deg: float, field_mT: float, *, model: ProductionModel = PRODUCTION_MODEL,
                      nuclei=ALL_NUCLEI, radicals=NUCLEAR_RADICALS, hyperfine_overrides_mT=None,
                      phi_deg: float = 0.0, condition: EfimovaHoreCondition | None = None,
                      quadrupole_tensors=None) -> CompiledFullPairHamiltonian:
    """Production full-pair H (R/poster_reproduction/core.py:172-229).

    ``phi_deg``, ``condition`` and ``quadrupole_tensors`` are extensions
    (field azimuth, explicit electron coupling, synthetic Q tensors); their
    defaults give the reference Hamiltonian.
    """
    if len(nuclei) != len(radicals):
        raise ValueError("nuclei and radicals must have the same length.")
    dims = (2, 2) + tuple(n.dimension for n in nuclei)
    z = zeeman_operator(field_vector_mT(theta_deg, float(field_mT) * 1000.0, phi_deg))
    terms = [LocalTerm((0,), z, "electron A Zeeman"), LocalTerm((1,), z, "electron B Zeeman")]
    overrides = hyperfine_overrides_mT or {}
    qtab = QUADRUPOLE_TENSORS_MT if quadrupole_tensors is None else quadrupole_tensors
    used: dict[str, int] = {}
    for idx, (nuc, rad) in enumerate(zip(nuclei, radicals)):
        e_axis = 0 if rad == "FAD" else 1
        key = f"{rad}:{nuc.name}"
        tensor = np.asarray(overrides.get(key, nuc.tensor_mT), dtype=float)
        terms.append(LocalTerm((e_axis, idx + 2), TWOPI * GAMMA_E_MHZ_PER_MT *
                               two_spin_coupling_operator(0.5, nuc.spin, tensor), f"{key} hyperfine"))
        q = qtab.get(key)
        if q is not None and model.quadrupole_scale != 0.0:
            terms.append(LocalTerm((idx + 2,), quadrupole_operator(
                nuc.spin, TWOPI * GAMMA_E_MHZ_PER_MT * model.quadrupole_scale * q), f"{key} quadrupole"))
        used[key] = used.get(key, 0) + 1
    if any(v != 1 for v in used.values()):
        raise ValueError("Selected nuclei contain duplicate radical/name keys.")
    terms.append(electron_coupling_term(model.condition if condition is None else condition))
    left = (0,) + tuple(i + 2 for i, r in enumerate(radicals) if r == "FAD")
    right = (1,) + tuple(i + 2 for i, r in enumerate(radicals) if r == "TrpH")
    if len(left) + len(right) != len(dims):
        raise ValueError("Every selected nucleus must belong to FAD or TrpH.")
    return CompiledFullPairHamiltonian(dims, terms, left_axes=left, right_axes=right)


def make_propagator(hamiltonian, config: SweepConfig, *, state_error_budget=None, maximum_dimension=None,
                    probe_horizon=None) -> AdaptiveLanczosPropagator:
    """R/poster_reproduction/core.py:232-247 and R/rpm_full14_quadrupole_experiment.py:138-148."""
    budget = config.state_error_budget if state_error_budget is None else state_error_budget
    if probe_horizon is not None:
        budget *= probe_horizon / config.maximum_time_us
    return AdaptiveLanczosPropagator(
        hamiltonian,
        minimum_dimension=config.lanczos_minimum_dimension,
        maximum_dimension=config.lanczos_maximum_dimension if maximum_dimension is None else maximum_dimension,
        dimension_increment=config.lanczos_dimension_increment,
        embedded_offset=config.lanczos_embedded_offset,
        safety_factor=config.lanczos_safety_factor,
        state_error_budget=budget,
        initial_step_us=config.lanczos_initial_step_us,
        maximum_step_us=config.lanczos_maximum_step_us)


def coarse_grid_indices(config: SweepConfig) -> np.ndarray:
    """R/rpm_quantum_needle_jd_sweeps.py:423-431."""
    ratio = config.coarse_integration_step_us / config.integration_step_us
    stride = int(round(ratio))
    if stride < 1 or not math.isclose(ratio, stride, rel_tol=0.0, abs_tol=1e-12):
        raise ValueError("The coarse integration step must be an integer fine-grid multiple.")
    n = config.time_grid_us.size
    idx = np.arange(0, n, stride, dtype=int)
    if idx[-1] != n - 1:
        idx = np.append(idx, n - 1)
    return idx


def propagate_trajectory(hamiltonian, initial_state, config: SweepConfig, *, propagator=None, times_us=None,
                         lifetime_us=None, schedule=None, progress_callback=None) -> dict:
    """Numerical part of R/rpm_full14_quadrupole_experiment.py:154-194 for one trajectory."""
    tau = config.lifetime_us if lifetime_us is None else lifetime_us
    prop = make_propagator(hamiltonian, config) if propagator is None else propagator
    grid = config.time_grid_us if times_us is None else np.asarray(times_us, dtype=float)
    ids = coarse_grid_indices(config) if times_us is None else None
    # the yield integrals are summed on the device as the populations are produced
    res = prop.propagate(initial_state, grid, SingletPopulationObservable(hamiltonian.dimensions),
                         progress_callback=progress_callback, schedule=schedule,
                         yield_weights=yield_weight_table(grid, tau, ids))
    fine, coarse, estimate = res.device_yields
    out = {"result": res, "yield": fine, "times_us": grid, "singlet_population": res.observable_values,
           "population_error_estimate": res.observable_error_bounds,
           "state_error_estimates": res.state_error_estimates, "final_state": res.final_state}
    if times_us is None:
        out.update(coarse_yield=coarse, quadrature_error=abs(fine - coarse), yield_propagation_estimate=estimate)
    out["initial_singlet"] = singlet_population(initial_state, hamiltonian.dimensions)
    return out
