"""Matrix-free Hamiltonians, singlet observable and adaptive Lanczos propagator.

Oracle restatement of R/rpm_statevector_engine.py.  NumPy/SciPy FP64, the
same third-party routines the reference calls (``scipy.linalg.eigh_tridiagonal``,
``scipy.sparse`` CSR products), so that decisions taken by the adaptive
propagator follow the reference's arithmetic as closely as a restatement can.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Iterable

import numpy as np
import scipy.linalg
import scipy.sparse

from .model import (
    ALL_NUCLEI, FAD_NUCLEI, GAMMA_E_MHZ_PER_MT, QUADRUPOLE_TENSORS_MT, SPIN_ORDER,
    SYSTEM_DIMENSIONS, TWOPI, field_vector_mT, spin_operators,
)


@dataclass(frozen=True)
class LocalTerm:
    """Operator on a few axes of a tensor-product state (R/rpm_statevector_engine.py:28-34)."""

    axes: tuple
    operator: np.ndarray
    label: str = ""


@dataclass(frozen=True)
class _Term:
    axes: tuple
    active_dimensions: tuple
    remaining_dimensions: tuple
    operator: np.ndarray
    label: str


class MatrixFreeHamiltonian:
    """Generic sum of local operators (R/rpm_statevector_engine.py:46-129)."""

    def __init__(self, dimensions: Iterable[int], terms: Iterable[LocalTerm], *,
                 hermitian_tolerance: float = 1e-12) -> None:
        self.dimensions = tuple(int(d) for d in dimensions)
        if not self.dimensions or min(self.dimensions) < 1:
            raise ValueError("Every tensor dimension must be positive.")
        self.dimension = math.prod(self.dimensions)
        self.terms = tuple(self._check(t, hermitian_tolerance) for t in terms)
        self.shape = (self.dimension, self.dimension)
        self.dtype = np.dtype(np.complex128)
        self.matvec_count = 0

    def _check(self, term: LocalTerm, tol: float) -> _Term:
        # validation rules of R/rpm_statevector_engine.py:67-93
        name = term.label or "Local term"
        axes = tuple(int(a) for a in term.axes)
        if not axes or len(set(axes)) != len(axes):
            raise ValueError(f"{name} must have unique active axes.")
        if min(axes) < 0 or max(axes) >= len(self.dimensions):
            raise ValueError(f"{name} contains an out-of-range axis.")
        active = tuple(self.dimensions[a] for a in axes)
        local = math.prod(active)
        op = np.asarray(term.operator, dtype=complex)
        if op.shape != (local, local):
            raise ValueError(f"{name} has shape {op.shape}; expected {(local, local)}.")
        scale = max(1.0, float(np.linalg.norm(op)))
        if np.linalg.norm(op - op.conj().T) > tol * scale:
            raise ValueError(f"{name} is not Hermitian.")
        rest = tuple(d for i, d in enumerate(self.dimensions) if i not in axes)
        return _Term(axes, active, rest, np.ascontiguousarray(op), term.label)

    def matvec(self, vector) -> np.ndarray:
        # R/rpm_statevector_engine.py:95-108: per term, bring the active axes to
        # the front, multiply, move them back, accumulate.
        flat = np.asarray(vector, dtype=complex).reshape(-1)
        if flat.size != self.dimension:
            raise ValueError(f"Vector length {flat.size} does not match {self.dimension}.")
        psi = flat.reshape(self.dimensions)
        out = np.zeros_like(psi, dtype=complex)
        for t in self.terms:
            k = len(t.axes)
            front = np.moveaxis(psi, t.axes, range(k)).reshape(math.prod(t.active_dimensions), -1)
            applied = (t.operator @ front).reshape(t.active_dimensions + t.remaining_dimensions)
            out += np.moveaxis(applied, range(k), t.axes)
        self.matvec_count += 1
        return out.reshape(-1)

    def matmat(self, matrix) -> np.ndarray:
        m = np.asarray(matrix, dtype=complex)
        if m.ndim != 2 or m.shape[0] != self.dimension:
            raise ValueError(f"Matrix must have shape ({self.dimension}, n_vectors).")
        return np.column_stack([self.matvec(m[:, j]) for j in range(m.shape[1])])

    def explicit_matrix(self) -> np.ndarray:
        return self.matmat(np.eye(self.dimension, dtype=complex))


class CompiledFullPairHamiltonian:
    """Radical-grouped two-sided sparse H (R/rpm_statevector_engine.py:132-266).

    With M the (left x right) regrouped view of psi, H psi = L M + (R M^T)^T
    plus the electron-pair 4x4 term.
    """

    def __init__(self, dimensions, terms, *, left_axes, right_axes,
                 electron_pair_axes=(0, 1)) -> None:
        generic = MatrixFreeHamiltonian(dimensions, terms)
        self.dimensions = generic.dimensions
        self.dimension = generic.dimension
        self.shape = generic.shape
        self.dtype = generic.dtype
        self.terms = generic.terms
        self.matvec_count = 0
        left = tuple(int(a) for a in left_axes)
        right = tuple(int(a) for a in right_axes)
        if not left or not right or set(left) & set(right):
            raise ValueError("Full-pair axis groups must be nonempty and disjoint.")
        if set(left + right) != set(range(len(self.dimensions))):
            raise ValueError("Full-pair axis groups must partition every tensor axis.")
        if left[0] != electron_pair_axes[0] or right[0] != electron_pair_axes[1]:
            raise ValueError("Each radical electron must be first in its axis group.")
        if self.dimensions[left[0]] != 2 or self.dimensions[right[0]] != 2:
            raise ValueError("The electron-pair axes must both have dimension two.")
        lpos = {a: i for i, a in enumerate(left)}
        rpos = {a: i for i, a in enumerate(right)}
        lterms, rterms = [], []
        pair = np.zeros((4, 4), dtype=complex)
        for t in self.terms:
            if all(a in lpos for a in t.axes):
                lterms.append(LocalTerm(tuple(lpos[a] for a in t.axes), t.operator, t.label))
            elif all(a in rpos for a in t.axes):
                rterms.append(LocalTerm(tuple(rpos[a] for a in t.axes), t.operator, t.label))
            elif t.axes == tuple(electron_pair_axes):
                pair += t.operator
            else:
                raise ValueError("Cross-radical terms must act on the configured electron pair.")
        self.left_axes, self.right_axes = left, right
        self.permutation = left + right
        self.inverse_permutation = tuple(int(v) for v in np.argsort(self.permutation))
        self.left_dimensions = tuple(self.dimensions[a] for a in left)
        self.right_dimensions = tuple(self.dimensions[a] for a in right)
        self.left_dimension = math.prod(self.left_dimensions)
        self.right_dimension = math.prod(self.right_dimensions)
        ldense = MatrixFreeHamiltonian(self.left_dimensions, lterms).explicit_matrix()
        rdense = MatrixFreeHamiltonian(self.right_dimensions, rterms).explicit_matrix()
        self.left_dense, self.right_dense = ldense, rdense
        self.left_operator = scipy.sparse.csr_matrix(ldense)
        self.right_operator = scipy.sparse.csr_matrix(rdense)
        self.left_operator.eliminate_zeros()
        self.right_operator.eliminate_zeros()
        self.electron_pair_operator = pair.reshape(2, 2, 2, 2)

    def matvec(self, vector) -> np.ndarray:
        # R/rpm_statevector_engine.py:221-245
        flat = np.asarray(vector, dtype=complex).reshape(-1)
        if flat.size != self.dimension:
            raise ValueError(f"Vector length {flat.size} does not match {self.dimension}.")
        grouped = np.transpose(flat.reshape(self.dimensions), self.permutation)
        m = grouped.reshape(self.left_dimension, self.right_dimension)
        out = self.left_operator @ m + (self.right_operator @ m.T).T
        four = grouped.reshape(2, self.left_dimension // 2, 2, self.right_dimension // 2)
        out += np.einsum("abcd,cxdy->axby", self.electron_pair_operator, four,
                         optimize=True).reshape(self.left_dimension, self.right_dimension)
        back = np.transpose(out.reshape(self.left_dimensions + self.right_dimensions),
                            self.inverse_permutation)
        self.matvec_count += 1
        return back.reshape(-1)

    def matmat(self, matrix) -> np.ndarray:
        m = np.asarray(matrix, dtype=complex)
        if m.ndim != 2 or m.shape[0] != self.dimension:
            raise ValueError(f"Matrix must have shape ({self.dimension}, n_vectors).")
        return np.column_stack([self.matvec(m[:, j]) for j in range(m.shape[1])])

    def explicit_matrix(self) -> np.ndarray:
        return self.matmat(np.eye(self.dimension, dtype=complex))


# ---------------------------------------------------------------- operators

def two_spin_coupling_operator(left_spin, right_spin, tensor) -> np.ndarray:
    """sum_ab T_ab S_a (x) I_b  (R/rpm_statevector_engine.py:269-287)."""
    t = np.asarray(tensor, dtype=float)
    if t.shape != (3, 3):
        raise ValueError("A two-spin coupling tensor must be 3x3.")
    s, i = spin_operators(left_spin), spin_operators(right_spin)
    out = np.zeros((s[0].shape[0] * i[0].shape[0],) * 2, dtype=complex)
    for a in range(3):
        for b in range(3):
            if t[a, b] != 0.0:
                out += t[a, b] * np.kron(s[a], i[b])
    return out


def efimova_hore_electron_coupling_operator(exchange_J_MHz, dipolar_D_MHz, dipolar_axis) -> np.ndarray:
    """-J(S^2 - 1) + D((S.n)^2 - S^2/3) in MHz, Hermitized (R/rpm_statevector_engine.py:290-329)."""
    n = np.asarray(dipolar_axis, dtype=float)
    if n.shape != (3,):
        raise ValueError("dipolar_axis must be a length-3 vector.")
    norm = float(np.linalg.norm(n))
    if norm == 0.0:
        raise ValueError("dipolar_axis must be nonzero.")
    n = n / norm
    one = spin_operators(0.5)
    i2, i4 = np.eye(2, dtype=complex), np.eye(4, dtype=complex)
    total = [np.kron(o, i2) + np.kron(i2, o) for o in one]
    s2 = total[0] @ total[0] + total[1] @ total[1] + total[2] @ total[2]
    sn = n[0] * total[0] + n[1] * total[1] + n[2] * total[2]
    h = -float(exchange_J_MHz) * (s2 - i4) + float(dipolar_D_MHz) * (sn @ sn - s2 / 3.0)
    return 0.5 * (h + h.conj().T)


def quadrupole_operator(spin, tensor) -> np.ndarray:
    """sum_ab Q_ab {I_a, I_b}/2 (R/rpm_statevector_engine.py:332-349)."""
    q = np.asarray(tensor, dtype=float)
    if q.shape != (3, 3):
        raise ValueError("A quadrupole tensor must be 3x3.")
    if not np.allclose(q, q.T, atol=1e-14, rtol=0.0):
        raise ValueError("A quadrupole tensor must be symmetric.")
    ops = spin_operators(spin)
    out = np.zeros_like(ops[0])
    for a in range(3):
        for b in range(3):
            if q[a, b] != 0.0:
                out += 0.5 * q[a, b] * (ops[a] @ ops[b] + ops[b] @ ops[a])
    return out


def build_quantum_needle_hamiltonian(theta_deg, field_microT=50.0, *, exchange_J_MHz=0.0,
                                     dipolar_tensor_MHz=None, efimova_hore_D_MHz=None,
                                     efimova_hore_axis=None, quadrupole_scale=1.0):
    """Generic full-pair H (R/rpm_statevector_engine.py:367-469)."""
    e_ops = spin_operators(0.5)
    b = field_vector_mT(theta_deg, field_microT)
    zeeman = TWOPI * GAMMA_E_MHZ_PER_MT * (b[0] * e_ops[0] + b[1] * e_ops[1] + b[2] * e_ops[2])
    terms = [LocalTerm((0,), zeeman, "electron A Zeeman"),
             LocalTerm((1,), zeeman, "electron B Zeeman")]
    for idx, nuc in enumerate(ALL_NUCLEI):
        e_axis = 0 if idx < len(FAD_NUCLEI) else 1
        terms.append(LocalTerm(
            (e_axis, idx + 2),
            TWOPI * GAMMA_E_MHZ_PER_MT * two_spin_coupling_operator(0.5, nuc.spin, nuc.tensor_mT),
            f"{nuc.name} hyperfine"))
        q = QUADRUPOLE_TENSORS_MT.get(SPIN_ORDER[idx + 2])
        if q is not None and quadrupole_scale != 0.0:
            terms.append(LocalTerm(
                (idx + 2,),
                quadrupole_operator(nuc.spin, TWOPI * GAMMA_E_MHZ_PER_MT * float(quadrupole_scale) * q),
                f"{nuc.name} nuclear quadrupole"))
    if efimova_hore_D_MHz is not None and dipolar_tensor_MHz is not None:
        raise ValueError("Choose either dipolar_tensor_MHz or efimova_hore_D_MHz, not both.")
    if efimova_hore_D_MHz is not None:
        axis = np.array([1.0, 0.0, 0.0]) if efimova_hore_axis is None else np.asarray(efimova_hore_axis, float)
        op = TWOPI * efimova_hore_electron_coupling_operator(exchange_J_MHz, efimova_hore_D_MHz, axis)
        if np.any(op != 0.0):
            terms.append(LocalTerm((0, 1), op, "Efimova-Hore electron exchange + dipolar coupling"))
        return MatrixFreeHamiltonian(SYSTEM_DIMENSIONS, terms)
    dip = np.zeros((3, 3)) if dipolar_tensor_MHz is None else np.asarray(dipolar_tensor_MHz, float)
    if dip.shape != (3, 3):
        raise ValueError("dipolar_tensor_MHz must be a 3x3 tensor.")
    tensor = TWOPI * (dip - 2.0 * float(exchange_J_MHz) * np.eye(3))
    if np.any(tensor != 0.0):
        terms.append(LocalTerm((0, 1), two_spin_coupling_operator(0.5, 0.5, tensor),
                               "electron exchange + dipolar coupling"))
    return MatrixFreeHamiltonian(SYSTEM_DIMENSIONS, terms)


# ---------------------------------------------------------------- states and observable

SINGLET = np.array([0.0, 1.0, -1.0, 0.0], dtype=complex) / math.sqrt(2.0)


def random_phase_nuclear_state(nuclear_dimension, rng) -> np.ndarray:
    """R/rpm_statevector_engine.py:472-477."""
    phases = rng.uniform(0.0, TWOPI, int(nuclear_dimension))
    return np.exp(1.0j * phases) / math.sqrt(nuclear_dimension)


def singlet_initial_state(dimensions, rng) -> np.ndarray:
    """R/rpm_statevector_engine.py:480-489."""
    dims = tuple(int(d) for d in dimensions)
    if dims[:2] != (2, 2):
        raise ValueError("The first two axes must be the two electron spins.")
    return np.kron(SINGLET, random_phase_nuclear_state(math.prod(dims[2:]), rng))


def singlet_population(vector, dimensions) -> float:
    """|| (psi[0,1,:] - psi[1,0,:]) / sqrt2 ||^2 (R/rpm_statevector_engine.py:492-497)."""
    e = np.asarray(vector, dtype=complex).reshape(2, 2, -1)
    amp = (e[0, 1] - e[1, 0]) / math.sqrt(2.0)
    return float(np.vdot(amp, amp).real)


class SingletPopulationObservable:
    """Singlet population with the Krylov projection shortcut (R/rpm_statevector_engine.py:500-538)."""

    def __init__(self, dimensions) -> None:
        dims = tuple(int(d) for d in dimensions)
        if dims[:2] != (2, 2):
            raise ValueError("The first two axes must be the two electron spins.")
        self.dimensions = dims

    def __call__(self, vector) -> float:
        return singlet_population(vector, self.dimensions)

    def project_basis(self, basis) -> np.ndarray:
        v = np.asarray(basis, dtype=complex)
        n = math.prod(self.dimensions)
        if v.ndim != 2 or v.shape[0] != n:
            raise ValueError(f"Krylov basis must have shape ({n}, m).")
        e = v.reshape(2, 2, -1, v.shape[1])
        a = (e[0, 1] - e[1, 0]) / math.sqrt(2.0)
        p = a.conj().T @ a
        return 0.5 * (p + p.conj().T)

    @staticmethod
    def evaluate_projected(projected, coefficients, state_norm) -> float:
        return float(state_norm * state_norm * np.vdot(coefficients, projected @ coefficients).real)


def state_error_to_population_bound(state_error) -> float:
    """min(1, 2e + e^2) (R/rpm_statevector_engine.py:541-543)."""
    e = max(0.0, float(state_error))
    return min(1.0, 2.0 * e + e * e)


@dataclass(frozen=True)
class PropagationStats:
    """R/rpm_statevector_engine.py:546-555."""

    accepted_steps: int
    rejected_steps: int
    matvecs: int
    maximum_krylov_dimension: int
    minimum_step_us: float
    maximum_step_us: float
    accumulated_state_error: float
    maximum_norm_error: float


@dataclass(frozen=True)
class PropagationResult:
    """R/rpm_statevector_engine.py:558-566, plus the recorded step schedule."""

    times_us: np.ndarray
    observable_values: np.ndarray
    observable_error_bounds: np.ndarray
    state_error_estimates: np.ndarray
    states: np.ndarray | None
    final_state: np.ndarray
    stats: PropagationStats
    schedule: tuple = field(default=())


def _tridiagonal_eigh(diagonal, off_diagonal):
    if diagonal.size == 1:
        return np.asarray(diagonal, float), np.ones((1, 1)), np.ones(1)
    w, q = scipy.linalg.eigh_tridiagonal(diagonal, off_diagonal, check_finite=False)
    return w, q, q[0, :].copy()


def projected_coefficients(diagonal, off_diagonal, time_us) -> np.ndarray:
    """c(t) = Q (exp(-i w t) * Q[0,:]) (R/rpm_statevector_engine.py:569-582)."""
    if diagonal.size == 1:
        return np.array([np.exp(-1.0j * diagonal[0] * time_us)])
    w, q, first = _tridiagonal_eigh(diagonal, off_diagonal)
    return q @ (np.exp(-1.0j * w * time_us) * first)


class _Evolution:
    """R/rpm_statevector_engine.py:585-609."""

    def __init__(self, diagonal, off_diagonal):
        self.w, self.q, self.first = _tridiagonal_eigh(diagonal, off_diagonal)

    def coefficients(self, t):
        return self.q @ (np.exp(-1.0j * self.w * t) * self.first)


class AdaptiveLanczosPropagator:
    """Adaptive Hermitian Lanczos propagation (R/rpm_statevector_engine.py:612-933).

    ``propagate`` also records the step schedule: one tuple per attempt,
    ``(trial_step_us, accepted, selected_dimension, basis_sizes)``.
    """

    def __init__(self, hamiltonian, *, minimum_dimension=24, maximum_dimension=64,
                 dimension_increment=8, embedded_offset=4, safety_factor=2.0,
                 state_error_budget=2.4e-5, initial_step_us=0.05, maximum_step_us=0.25,
                 minimum_step_us=1e-9, breakdown_tolerance=1e-14) -> None:
        if minimum_dimension <= embedded_offset:
            raise ValueError("minimum_dimension must exceed embedded_offset.")
        if maximum_dimension < minimum_dimension:
            raise ValueError("maximum_dimension must not be smaller than minimum_dimension.")
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

    def _basis(self, state, target):
        """CGS2 Lanczos (R/rpm_statevector_engine.py:651-688)."""
        norm = float(np.linalg.norm(state))
        if norm == 0.0:
            raise ValueError("Cannot propagate the zero vector.")
        count = min(int(target), state.size)
        v = np.empty((state.size, count), dtype=complex)
        vc = np.empty_like(v)
        alpha = np.empty(count)
        beta = np.empty(max(count - 1, 0))
        v[:, 0] = state / norm
        np.conjugate(v[:, 0], out=vc[:, 0])
        size = count
        for j in range(count):
            w = self.hamiltonian.matvec(v[:, j])
            alpha[j] = float(np.vdot(v[:, j], w).real)
            w -= alpha[j] * v[:, j]
            if j:
                w -= beta[j - 1] * v[:, j - 1]
            active, active_h = v[:, : j + 1], vc[:, : j + 1].T
            for _ in range(2):
                w -= active @ (active_h @ w)
            if j == count - 1:
                break
            b = float(np.linalg.norm(w))
            beta[j] = b
            if b <= self.breakdown_tolerance:
                size = j + 1
                break
            v[:, j + 1] = w / b
            np.conjugate(v[:, j + 1], out=vc[:, j + 1])
        return v[:, :size], alpha[:size], beta[: max(0, size - 1)]

    def _candidate(self, diagonal, off_diagonal, dimension, time_us, state_norm):
        """R/rpm_statevector_engine.py:690-715."""
        dimension = min(dimension, diagonal.size)
        high = projected_coefficients(diagonal[:dimension], off_diagonal[: max(0, dimension - 1)], time_us)
        low_dim = max(1, dimension - self.embedded_offset)
        if low_dim == dimension:
            return high, 0.0
        low = projected_coefficients(diagonal[:low_dim], off_diagonal[: max(0, low_dim - 1)], time_us)
        diff = high.copy()
        diff[:low_dim] -= low
        return high, self.safety_factor * state_norm * float(np.linalg.norm(diff))

    def _ladder(self, start):
        c = list(range(start, self.maximum_dimension + 1, self.dimension_increment))
        if not c or c[-1] != self.maximum_dimension:
            c.append(self.maximum_dimension)
        return c

    def propagate(self, initial_state, times_us, observable, *, store_states=False,
                  progress_callback=None, schedule=None) -> PropagationResult:
        """R/rpm_statevector_engine.py:717-933 (rule set restated one for one).

        ``schedule`` (test extension, no reference counterpart): a recorded
        schedule as returned in ``PropagationResult.schedule``.  Its trial steps,
        accept/reject decisions and selected dimensions are replayed instead of
        being decided from this run's error estimates, so that runs at nearby
        parameters share one step sequence (central finite differences).
        """
        times = np.asarray(tuple(float(t) for t in times_us), dtype=float)
        if times.ndim != 1 or times.size == 0:
            raise ValueError("At least one observation time is required.")
        if np.any(times < 0.0) or np.any(np.diff(times) <= 0.0):
            raise ValueError("Observation times must be nonnegative and strictly increasing.")
        state = np.asarray(initial_state, dtype=complex).reshape(-1).copy()
        if state.size != self.hamiltonian.dimension:
            raise ValueError("Initial state dimension does not match the Hamiltonian.")
        project = getattr(observable, "project_basis", None)
        evaluate = getattr(observable, "evaluate_projected", None)
        projected_path = (not store_states) and callable(project) and callable(evaluate)

        values = np.empty(times.size)
        bounds = np.empty(times.size)
        errors = np.empty(times.size)
        stored = np.empty((times.size, state.size), dtype=complex) if store_states else None
        cursor, now, acc_err = 0, 0.0, 0.0
        accepted = rejected = max_dim = 0
        used = []
        replay = None if schedule is None else list(schedule)
        record = []
        max_norm_err = abs(float(np.linalg.norm(state)) - 1.0)
        mv0 = self.hamiltonian.matvec_count
        if times[0] == 0.0:
            values[0] = observable(state)
            bounds[0] = errors[0] = 0.0
            if stored is not None:
                stored[0] = state
            cursor = 1
        t_final = float(times[-1])
        step = min(self.initial_step_us, t_final) if t_final else 0.0
        prefer_max = False

        while cursor < times.size:
            trial = min(step, t_final - now)
            rec = None
            if replay is not None:
                if len(record) >= len(replay):
                    raise RuntimeError("Replayed schedule ended before the final time.")
                rec = replay[len(record)]
                trial = float(rec[0])
            if trial < self.minimum_step_us:
                raise RuntimeError("Lanczos step fell below the configured minimum.")
            snorm = float(np.linalg.norm(state))
            target = self.maximum_dimension if prefer_max else self.minimum_dimension
            basis, d, e = self._basis(state, target)
            sizes = [d.size]
            budget = self.state_error_budget * trial / max(t_final, trial)

            def ok(dim, err, rec=rec, budget=budget):
                if rec is None:
                    return err <= budget
                return bool(rec[1]) and int(dim) == int(rec[2])

            selected, c_end, err = None, None, math.inf
            exact = d.size < target
            if exact:
                selected, err = d.size, 0.0
                c_end = projected_coefficients(d, e, trial)
            elif not prefer_max:
                c_end, err = self._candidate(d, e, self.minimum_dimension, trial, snorm)
                if ok(self.minimum_dimension, err):
                    selected = self.minimum_dimension
                elif self.maximum_dimension > self.minimum_dimension:
                    basis, d, e = self._basis(state, self.maximum_dimension)
                    sizes.append(d.size)
                    prefer_max = True
                    exact = d.size < self.maximum_dimension
                    if exact:
                        selected, err = d.size, 0.0
                        c_end = projected_coefficients(d, e, trial)
                    else:
                        for dim in self._ladder(self.minimum_dimension + self.dimension_increment):
                            c_end, err = self._candidate(d, e, dim, trial, snorm)
                            if ok(dim, err):
                                selected = dim
                                break
            else:
                for dim in self._ladder(self.minimum_dimension):
                    c_end, err = self._candidate(d, e, dim, trial, snorm)
                    if ok(dim, err):
                        selected = dim
                        break
            if exact and rec is not None and not rec[1]:
                selected = None
            if selected is None or (rec is None and err > budget):
                rejected += 1
                record.append((trial, False, 0, tuple(sizes)))
                step = 0.5 * trial
                continue

            record.append((trial, True, int(selected), tuple(sizes)))
            max_dim = max(max_dim, selected)
            if selected > self.minimum_dimension:
                prefer_max = True
            vb = basis[:, :selected]
            p = project(vb) if projected_path else None
            high = _Evolution(d[:selected], e[: max(0, selected - 1)])
            low_dim = max(1, selected - self.embedded_offset)
            low = None if (exact or low_dim == selected) else _Evolution(
                d[:low_dim], e[: max(0, low_dim - 1)])
            step_end = now + trial
            while cursor < times.size and times[cursor] <= step_end + 1e-13:
                offset = float(times[cursor] - now)
                c = high.coefficients(offset)
                if low is None:
                    perr = 0.0
                else:
                    diff = c.copy()
                    diff[:low_dim] -= low.coefficients(offset)
                    perr = self.safety_factor * snorm * float(np.linalg.norm(diff))
                if p is not None:
                    values[cursor] = evaluate(p, c, snorm)
                    pstate = None
                else:
                    pstate = snorm * (vb @ c)
                    values[cursor] = observable(pstate)
                errors[cursor] = acc_err + perr
                bounds[cursor] = state_error_to_population_bound(errors[cursor])
                if stored is not None:
                    stored[cursor] = pstate
                cursor += 1
            state = snorm * (vb @ c_end)
            now = step_end
            acc_err += err
            accepted += 1
            used.append(trial)
            max_norm_err = max(max_norm_err, abs(float(np.linalg.norm(state)) - 1.0))
            if progress_callback is not None:
                progress_callback(now, accepted, rejected, self.hamiltonian.matvec_count - mv0, acc_err)
            step = min(self.maximum_step_us, 1.5 * trial) if err < 0.1 * budget else trial

        stats = PropagationStats(accepted, rejected, self.hamiltonian.matvec_count - mv0, max_dim,
                                 min(used) if used else 0.0, max(used) if used else 0.0,
                                 acc_err, max_norm_err)
        return PropagationResult(times, values, bounds, errors, stored, state, stats, tuple(record))
