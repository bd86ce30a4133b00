"""Oracle against the reference's own outputs (tests/golden/reference_vectors.npz).

The fixture was written by tests/golden/make_golden.py from the reference
solver: propagation schedules (through its progress callback), singlet
populations, state-error estimates and final states on two reduced pairs, and
the compiled full-pair matvec on one seeded 221,184-vector.
"""

import os

import numpy as np
import pytest

from oracle import core, engine, model

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz")
CASES = {
    "six": ((model.FAD_NUCLEI[0], model.FAD_NUCLEI[1], model.FAD_NUCLEI[2], model.TRP_NUCLEI[0],
             model.TRP_NUCLEI[1], model.TRP_NUCLEI[2]), ("FAD", "FAD", "FAD", "TrpH", "TrpH", "TrpH")),
    "smoke": ((model.FAD_NUCLEI[0], model.FAD_NUCLEI[1], model.TRP_NUCLEI[0]), ("FAD", "FAD", "TrpH")),
}


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_reference_propagation(golden, name):
    nuclei, radicals = CASES[name]
    theta = float(golden[f"{name}_theta"])
    h = core.build_hamiltonian(theta, 0.05, nuclei=nuclei, radicals=radicals)
    state = core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, 0)
    times = golden[f"{name}_times"]
    events = []
    prop = core.make_propagator(h, core.PRODUCTION_MODEL.sweep_config)
    res = prop.propagate(state, times, engine.SingletPopulationObservable(h.dimensions),
                         progress_callback=lambda t, a, r, m, e: events.append((t, a, r, m, e)))
    ref_events = golden[f"{name}_events"]
    ev = np.array(events)
    assert ev.shape == ref_events.shape
    # identical schedule: step end times, acceptance/rejection counts, matvec counts
    np.testing.assert_allclose(ev[:, 0], ref_events[:, 0], rtol=0, atol=1e-14)
    np.testing.assert_array_equal(ev[:, 1:4], ref_events[:, 1:4])
    np.testing.assert_allclose(ev[:, 4], ref_events[:, 4], rtol=1e-8)
    np.testing.assert_allclose(res.observable_values, golden[f"{name}_values"], atol=1e-12)
    # the estimates are norms of differences of two nearly equal Krylov coefficient vectors: their
    # round-off floor (~1e-14 absolute) moves with the BLAS summation order (thread count)
    np.testing.assert_allclose(res.state_error_estimates, golden[f"{name}_state_errors"], rtol=1e-6, atol=2e-13)
    np.testing.assert_allclose(res.final_state, golden[f"{name}_final_state"], atol=1e-12)
    s = res.stats
    ref = golden[f"{name}_stats"]
    assert [s.accepted_steps, s.rejected_steps, s.matvecs, s.maximum_krylov_dimension] == list(ref[:4].astype(int))


def test_oracle_full_pair_matvec_matches_reference(golden):
    h = core.build_hamiltonian(37.0, 0.05)
    rng = np.random.default_rng(20260731)
    v = rng.normal(size=h.dimension) + 1j * rng.normal(size=h.dimension)
    hv = h.matvec(v)
    idx = golden["full_matvec_index"]
    scale = float(golden["full_matvec_norm"])
    assert abs(np.linalg.norm(hv) - scale) < 1e-12 * scale
    assert abs(np.vdot(v, hv) - golden["full_matvec_vhv"]) < 1e-11 * abs(golden["full_matvec_vhv"]) + 1e-9
    np.testing.assert_allclose(hv[idx], golden["full_matvec_samples"], atol=1e-13 * scale)


def test_generic_and_compiled_full_system_agree():
    # R/tests/test_poster_reproduction.py:165-183 on the oracle
    compiled = core.build_hamiltonian(37.0, 2.0)
    generic = engine.MatrixFreeHamiltonian(compiled.dimensions,
                                           [engine.LocalTerm(t.axes, t.operator, t.label) for t in compiled.terms])
    rng = np.random.default_rng(20260731)
    v = rng.normal(size=compiled.dimension) + 1j * rng.normal(size=compiled.dimension)
    want = generic.matvec(v)
    assert np.linalg.norm(compiled.matvec(v) - want) / np.linalg.norm(want) < 1e-12
