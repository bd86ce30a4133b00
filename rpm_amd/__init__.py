"""rpm_amd -- MI355X-native radical-pair solver path (H·psi, adaptive Lanczos, singlet yield).

Drop-in for the solver entry points of sam-not-smart/RPM
(R/rpm_statevector_engine.py, R/poster_reproduction/core.py,
R/rpm_quantum_needle_jd_sweeps.py).  The numerics run in librpm_amd.so
(hand-written HIP for gfx950), loaded through ``rpm_amd._native``.
"""

from .engine import (  # noqa: F401
    AdaptiveLanczosPropagator, CompiledFullPairHamiltonian, LocalTerm, MatrixFreeHamiltonian, PropagationResult,
    PropagationStats, SingletPopulationObservable, build_quantum_needle_hamiltonian,
    efimova_hore_electron_coupling_operator, quadrupole_operator, random_phase_nuclear_state,
    singlet_initial_state, singlet_population, state_error_to_population_bound, two_spin_coupling_operator,
)
from .model import (  # noqa: F401
    ALL_NUCLEI, FAD_NUCLEI, GAMMA_E_MHZ_PER_MT, NUCLEAR_DIMENSION, NUCLEAR_RADICALS, QUADRUPOLE_TENSORS_MT,
    SPIN_ORDER, STATE_DIMENSION, SYSTEM_DIMENSIONS, TRP_NUCLEI, TWOPI, Nucleus, field_vector_mT, spin_operators,
)

__version__ = "0.1.0"
