"""Production builders, seeds, grids and yield quadrature (oracle restatement).

Follows R/poster_reproduction/core.py:69-308,428-473,
R/rpm_quantum_needle_jd_sweeps.py:81-102,167-235,306-334,373-375,418-431 and
R/rpm_full14_quadrupole_experiment.py:138-194.
"""

from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass, replace

import numpy as np
import scipy.integrate
import scipy.linalg

from .engine import (
    AdaptiveLanczosPropagator, CompiledFullPairHamiltonian, LocalTerm, SINGLET,
    SingletPopulationObservable, efimova_hore_electron_coupling_operator, quadrupole_operator,
    singlet_population, two_spin_coupling_operator,
)
from .model import (
    ALL_NUCLEI, GAMMA_E_MHZ_PER_MT, NUCLEAR_RADICALS, QUADRUPOLE_TENSORS_MT, TWOPI,
    field_vector_mT, spin_operators,
)

MASTER_SEED = 20260731                    # R/poster_reproduction/core.py:56
R2_D_PAPER_UT = -2783.0 / (2.0 ** 3)      # R/rpm_quantum_needle_jd_sweeps.py:81


def equivalent_field_uT_to_MHz(value_uT) -> float:
    """R/rpm_quantum_needle_jd_sweeps.py:101-102."""
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


def efimova_hore_ratio_condition(ratio) -> EfimovaHoreCondition:
    """D fixed at the r=2 nm paper value, J = D/ratio (R/rpm_quantum_needle_jd_sweeps.py:224-235)."""
    r = float(ratio)
    if r == 0.0:
        raise ValueError("Efimova-Hore D/J ratio cannot be zero.")
    return EfimovaHoreCondition(R2_D_PAPER_UT, R2_D_PAPER_UT / r)


def uncoupled_condition() -> EfimovaHoreCondition:
    """D = J = 0 (R/rpm_quantum_needle_jd_sweeps.py:202-210)."""
    return EfimovaHoreCondition(0.0, 0.0)


@dataclass(frozen=True)
class SweepConfig:
    """Krylov and grid settings (R/rpm_quantum_needle_jd_sweeps.py:306-334)."""

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
    """R/poster_reproduction/core.py:69-101 (only what the solver path needs)."""

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


def seed_for(base, sample_id) -> int:
    """uint32 from SeedSequence([base, 0xA11CE, id]) (R/rpm_quantum_needle_jd_sweeps.py:373-375)."""
    seq = np.random.SeedSequence([int(base), 0xA11CE, int(sample_id)])
    return int(seq.generate_state(1, dtype=np.uint32)[0])


def random_phase_initial_state(dimensions, master_seed, sample_id) -> np.ndarray:
    """Singlet (x) e^{i phi}/sqrt(Nn) (R/poster_reproduction/core.py:250-258)."""
    rng = np.random.default_rng(seed_for(master_seed, sample_id))
    nn = math.prod(dimensions[2:])
    phases = rng.uniform(0.0, TWOPI, nn)
    return np.kron(SINGLET, np.exp(1.0j * phases) / math.sqrt(nn))


def electron_coupling_term(condition: EfimovaHoreCondition) -> LocalTerm:
    """R/poster_reproduction/core.py:154-161."""
    op = TWOPI * efimova_hore_electron_coupling_operator(
        condition.exchange_J_MHz, condition.d_paper_MHz, np.asarray(condition.dipolar_axis, float))
    return LocalTerm((0, 1), op, "Efimova-Hore exchange + full nonsecular dipolar")


def zeeman_operator(field_mT_vector) -> np.ndarray:
    e = spin_operators(0.5)
    b = np.asarray(field_mT_vector, float)
    return TWOPI * GAMMA_E_MHZ_PER_MT * (b[0] * e[0] + b[1] * e[1] + b[2] * e[2])


def build_hamiltonian(theta_deg, field_mT, *, model=PRODUCTION_MODEL, nuclei=ALL_NUCLEI,
                      radicals=NUCLEAR_RADICALS, phi_deg=0.0, condition=None,
                      quadrupole_tensors=None) -> CompiledFullPairHamiltonian:
    """Production compiled-pair H (R/poster_reproduction/core.py:172-229).

    ``phi_deg`` (field azimuth), ``condition`` and ``quadrupole_tensors``
    extend the reference signature; their defaults reproduce it exactly.
    """
    if len(nuclei) != len(radicals):
        raise ValueError("nuclei and radicals must have the same length.")
    dims = (2, 2) + tuple(n.dimension for n in nuclei)
    z = zeeman_operator(field_vector_mT(theta_deg, float(field_mT) * 1000.0, phi_deg))
    terms = [LocalTerm((0,), z, "electron A Zeeman"), LocalTerm((1,), z, "electron B Zeeman")]
    qtab = QUADRUPOLE_TENSORS_MT if quadrupole_tensors is None else quadrupole_tensors
    seen = set()
    for idx, (nuc, rad) in enumerate(zip(nuclei, radicals)):
        e_axis = 0 if rad == "FAD" else 1
        key = f"{rad}:{nuc.name}"
        if key in seen:
            raise ValueError("Selected nuclei contain duplicate radical/name keys.")
        seen.add(key)
        terms.append(LocalTerm((e_axis, idx + 2), TWOPI * GAMMA_E_MHZ_PER_MT *
                               two_spin_coupling_operator(0.5, nuc.spin, nuc.tensor_mT),
                               f"{key} hyperfine"))
        q = qtab.get(key)
        if q is not None and model.quadrupole_scale != 0.0:
            terms.append(LocalTerm((idx + 2,), quadrupole_operator(
                nuc.spin, TWOPI * GAMMA_E_MHZ_PER_MT * model.quadrupole_scale * q), f"{key} quadrupole"))
    terms.append(electron_coupling_term(model.condition if condition is None else condition))
    left = (0,) + tuple(i + 2 for i, r in enumerate(radicals) if r == "FAD")
    right = (1,) + tuple(i + 2 for i, r in enumerate(radicals) if r == "TrpH")
    if len(left) + len(right) != len(dims):
        raise ValueError("Every selected nucleus must belong to FAD or TrpH.")
    return CompiledFullPairHamiltonian(dims, terms, left_axes=left, right_axes=right)


def make_propagator(hamiltonian, config: SweepConfig, *, state_error_budget=None,
                    maximum_dimension=None) -> AdaptiveLanczosPropagator:
    """R/poster_reproduction/core.py:232-247 / R/rpm_full14_quadrupole_experiment.py:138-148."""
    return AdaptiveLanczosPropagator(
        hamiltonian,
        minimum_dimension=config.lanczos_minimum_dimension,
        maximum_dimension=config.lanczos_maximum_dimension if maximum_dimension is None else maximum_dimension,
        dimension_increment=config.lanczos_dimension_increment,
        embedded_offset=config.lanczos_embedded_offset,
        safety_factor=config.lanczos_safety_factor,
        state_error_budget=config.state_error_budget if state_error_budget is None else state_error_budget,
        initial_step_us=config.lanczos_initial_step_us,
        maximum_step_us=config.lanczos_maximum_step_us)


def integrate_yield(times_us, populations, lifetime_us) -> float:
    """Simpson of exp(-t/tau) P_S / tau (R/rpm_quantum_needle_jd_sweeps.py:418-420)."""
    t = np.asarray(times_us, float)
    return float(scipy.integrate.simpson(np.exp(-t / lifetime_us) * np.asarray(populations) / lifetime_us, x=t))


def coarse_grid_indices(config: SweepConfig) -> np.ndarray:
    """Stride-2 subset plus the last point (R/rpm_quantum_needle_jd_sweeps.py:423-431)."""
    ratio = config.coarse_integration_step_us / config.integration_step_us
    stride = int(round(ratio))
    if stride < 1 or not math.isclose(ratio, stride, rel_tol=0.0, abs_tol=1e-12):
        raise ValueError("The coarse integration step must be an integer fine-grid multiple.")
    n = config.time_grid_us.size
    idx = np.arange(0, n, stride, dtype=int)
    if idx[-1] != n - 1:
        idx = np.append(idx, n - 1)
    return idx


def initial_state_sha256(state) -> str:
    """Digest used by the archived runs (R/rpm_full14_quadrupole_experiment.py:186)."""
    return hashlib.sha256(np.asarray(state, dtype=complex).tobytes()).hexdigest()


def local_unitary_apply(state, dimensions, axis, unitary) -> np.ndarray:
    """One-axis unitary (R/poster_reproduction/core.py:285-292)."""
    psi = np.asarray(state, dtype=complex).reshape(tuple(dimensions))
    moved = np.moveaxis(psi, int(axis), 0)
    out = (np.asarray(unitary, complex) @ moved.reshape(dimensions[axis], -1)).reshape(moved.shape)
    return np.moveaxis(out, 0, axis).reshape(-1)


def spin_pi_pulse(spin, molecular_axis) -> np.ndarray:
    """exp(-i pi n.S) (R/poster_reproduction/core.py:304-308)."""
    n = np.asarray(molecular_axis, float)
    n = n / np.linalg.norm(n)
    ops = spin_operators(spin)
    return scipy.linalg.expm(-1.0j * math.pi * (n[0] * ops[0] + n[1] * ops[1] + n[2] * ops[2]))


def dense_singlet_dynamics(hamiltonian, times_us) -> np.ndarray:
    """Four-level singlet population by eigendecomposition (R/poster_reproduction/core.py:446-452)."""
    w, q = np.linalg.eigh(np.asarray(hamiltonian, complex))
    c = q.conj().T @ SINGLET
    states = q @ (c[:, None] * np.exp(-1.0j * w[:, None] * np.asarray(times_us)[None, :]))
    return np.abs(SINGLET.conj() @ states) ** 2


def dense_truncated_singlet_yield(hamiltonian, lifetime_us, final_time_us) -> float:
    """Closed-form finite-horizon yield of a four-level H (R/poster_reproduction/core.py:455-473)."""
    w, q = np.linalg.eigh(np.asarray(hamiltonian, complex))
    weights = np.abs(q.conj().T @ SINGLET) ** 2
    k = 1.0 / float(lifetime_us)
    decay = k + 1.0j * (w[:, None] - w[None, :])
    integral = k * (1.0 - np.exp(-decay * float(final_time_us))) / decay
    return float(np.sum(weights[:, None] * weights[None, :] * integral).real)


def electron_only_hamiltonian(theta_deg, field_mT, local_fields_mhz=None, model=PRODUCTION_MODEL):
    """R/poster_reproduction/core.py:428-443."""
    ops = spin_operators(0.5)
    i2 = np.eye(2, dtype=complex)
    b = GAMMA_E_MHZ_PER_MT * field_vector_mT(theta_deg, field_mT * 1000.0)
    loc = local_fields_mhz or (np.zeros(3), np.zeros(3))
    h = np.zeros((4, 4), dtype=complex)
    for index, f in enumerate((b + loc[0], b + loc[1])):
        op = TWOPI * (f[0] * ops[0] + f[1] * ops[1] + f[2] * ops[2])
        h += np.kron(op, i2) if index == 0 else np.kron(i2, op)
    h += electron_coupling_term(model.condition).operator
    return 0.5 * (h + h.conj().T)


def run_trajectory(hamiltonian, initial_state, config: SweepConfig, propagator=None,
                   times_us=None, lifetime_us=None):
    """Per-trajectory yield with fine/coarse quadrature and propagation estimate.

    Follows R/rpm_full14_quadrupole_experiment.py:154-194 (the numerical part).
    """
    tau = config.lifetime_us if lifetime_us is None else lifetime_us
    prop = make_propagator(hamiltonian, config) if propagator is None else propagator
    grid = config.time_grid_us if times_us is None else np.asarray(times_us, float)
    res = prop.propagate(initial_state, grid, SingletPopulationObservable(hamiltonian.dimensions))
    fine = integrate_yield(grid, res.observable_values, tau)
    out = {"result": res, "yield": fine}
    if times_us is None:
        ids = coarse_grid_indices(config)
        coarse = integrate_yield(grid[ids], res.observable_values[ids], tau)
        weights = np.exp(-grid / tau) / tau
        out.update(coarse_yield=coarse, quadrature_error=abs(fine - coarse),
                   yield_propagation_estimate=float(scipy.integrate.simpson(
                       weights * res.observable_error_bounds, x=grid)))
    out["initial_singlet"] = singlet_population(initial_state, hamiltonian.dimensions)
    return out


def run_segments(hamiltonian, initial_state, times_us, boundaries, make_prop, state_error_budget, axis=None,
                 unitary=None, *, hamiltonians=None, schedules=None):
    """Segment loop of R/rpm_nuclear_spin_rotation_control.py:247-342 (test oracle).

    ``boundaries``: pulse times (the final time is appended); ``make_prop(h)``
    builds a fresh propagator per segment; the local unitary acts on ``axis``
    after every boundary but the last.  Extensions for piecewise-constant
    fields: ``hamiltonians`` gives one Hamiltonian per segment (len(boundaries)
    + 1), ``schedules`` one recorded schedule per segment to replay.  Returns
    (populations, bounds, accumulated state error, final state, schedules).
    """
    times = np.asarray(times_us, float)
    dims = tuple(hamiltonian.dimensions)
    state = np.asarray(initial_state, complex)
    obs = SingletPopulationObservable(dims)
    pops = np.empty(times.size)
    bounds = np.empty(times.size)
    pops[0] = singlet_population(state, dims)
    bounds[0] = 0.0
    cursor, current, acc = 1, 0.0, 0.0
    schedules_out = []
    ends = np.append(np.asarray(boundaries, float), times[-1])
    if hamiltonians is not None and len(hamiltonians) != ends.size:
        raise ValueError("hamiltonians must give one Hamiltonian per segment")
    for index, boundary in enumerate(ends):
        boundary = float(boundary)
        stop = int(np.searchsorted(times, boundary + 1e-12, side="right"))
        requested = times[cursor:stop] - current
        duration = boundary - current
        local = requested if (requested.size and requested[-1] >= duration - 1e-12) else np.append(requested,
                                                                                                     duration)
        if duration > 0.0:
            prop = make_prop(hamiltonian if hamiltonians is None else hamiltonians[index])
            prop.state_error_budget = state_error_budget * duration / times[-1]
            res = prop.propagate(state, local, obs,
                                 schedule=None if schedules is None else schedules[len(schedules_out)])
            count = requested.size
            pops[cursor:stop] = res.observable_values[:count]
            e = acc + res.state_error_estimates[:count]
            bounds[cursor:stop] = 2.0 * e + e ** 2
            state = res.final_state
            acc += res.stats.accumulated_state_error
            schedules_out.append(res.schedule)
        cursor, current = stop, boundary
        if unitary is not None and boundary < times[-1] - 1e-12:
            state = local_unitary_apply(state, dims, axis, unitary)
    return pops, bounds, acc, state, schedules_out
