"""Oracle model tables and operators, pinned by the reference's own known answers.

Restates R/tests/test_statevector_baseline.py:55-160, R/tests/test_jd_sweeps.py:80-121,186-194
and R/tests/test_poster_reproduction.py:204-215 against the oracle, plus brute-force
identities that a transposed operand or a sign slip would fail.
"""

import math

import numpy as np
import pytest

from oracle import core, dense, engine, model


def test_dimensions_and_order():
    assert model.SYSTEM_DIMENSIONS == (2, 2, 3, 3, 2, 2, 2, 2, 2, 3, 2, 2, 2, 2, 2, 2)
    assert model.NUCLEAR_DIMENSION == 55_296
    assert model.STATE_DIMENSION == 221_184
    assert len(model.SPIN_ORDER) == 16
    assert model.SPIN_ORDER[2] == "FAD:N5" and model.SPIN_ORDER[9] == "TrpH:N1"


def test_quadrupole_tensors_traceless_with_literature_eigenvalues():
    reported = {"FAD:N5": (0.031, 0.041, -0.072), "FAD:N10": (0.032, 0.041, -0.072),
                "TrpH:N1": (0.016, 0.020, -0.036)}
    assert set(model.QUADRUPOLE_TENSORS_MT) == set(reported)
    for key, vals in reported.items():
        q = model.QUADRUPOLE_TENSORS_MT[key]
        np.testing.assert_allclose(q, q.T, atol=1e-14)
        assert abs(np.trace(q)) < 1e-14
        np.testing.assert_allclose(np.linalg.eigvalsh(q), np.sort(np.asarray(vals) - np.mean(vals)), atol=1e-14)


def test_quadrupole_tensors_match_methods_json_values():
    # methods/quadrupole_tensors.json "tensor_mT" for FAD:N5 (provisional frame)
    expected_n5 = np.array([[-0.062249980262593874, 0.030152763522984453, 0.0],
                            [0.030152763522984453, 0.02124998026259388, 0.0],
                            [0.0, 0.0, 0.040999999999999995]])
    np.testing.assert_allclose(model.QUADRUPOLE_TENSORS_MT["FAD:N5"], expected_n5, atol=1e-15)


def test_spin_operators_algebra():
    for s in (0.5, 1.0, 1.5):
        sx, sy, sz = model.spin_operators(s)
        np.testing.assert_allclose(np.diag(sz).real, s - np.arange(int(2 * s + 1)))
        # [Sx, Sy] = i Sz and S^2 = s(s+1)
        np.testing.assert_allclose(sx @ sy - sy @ sx, 1j * sz, atol=1e-14)
        np.testing.assert_allclose(sx @ sx + sy @ sy + sz @ sz, s * (s + 1) * np.eye(sx.shape[0]), atol=1e-14)
    _, _, sz1 = model.spin_operators(1.0)
    np.testing.assert_allclose(np.diag(sz1), [1.0, 0.0, -1.0])


def test_field_vector_plane_and_azimuth():
    np.testing.assert_allclose(model.field_vector_mT(90.0, 50.0), [0.05, 0.0, 0.05 * math.cos(math.pi / 2)])
    np.testing.assert_allclose(model.field_vector_mT(30.0, 50.0, phi_deg=0.0),
                               model.field_vector_mT(30.0, 50.0))
    b = model.field_vector_mT(40.0, 25.0, phi_deg=90.0)
    np.testing.assert_allclose(b, 0.025 * np.array([0.0, math.sin(math.radians(40)), math.cos(math.radians(40))]),
                               atol=1e-18)


def test_local_terms_match_dense_kronecker():
    # R/tests/test_statevector_baseline.py:86-103
    sx, sz = model.spin_operators(0.5)[0], model.spin_operators(0.5)[2]
    sy1 = model.spin_operators(1.0)[1]
    dims = (2, 3, 2)
    terms = (engine.LocalTerm((0, 2), np.kron(sx, sz)), engine.LocalTerm((1,), sy1))
    h = engine.MatrixFreeHamiltonian(dims, terms)
    expected = np.kron(sx, np.kron(np.eye(3), sz)) + np.kron(np.eye(2), np.kron(sy1, np.eye(2)))
    np.testing.assert_allclose(h.explicit_matrix(), expected, atol=1e-13)
    np.testing.assert_allclose(dense.dense_from_terms(dims, terms), expected, atol=1e-15)


def test_reversed_axis_order_term_is_not_symmetric_in_axes():
    # a term listed as (2, 0) must act as op on (axis2, axis0), i.e. a transposed Kronecker order
    sx, sy, sz = model.spin_operators(0.5)
    dims = (2, 3, 2)
    op = np.kron(sz, sx) + 0.3 * np.kron(sx, sy)
    h1 = engine.MatrixFreeHamiltonian(dims, (engine.LocalTerm((2, 0), op),)).explicit_matrix()
    brute = dense.dense_from_terms(dims, (engine.LocalTerm((2, 0), op),))
    np.testing.assert_allclose(h1, brute, atol=1e-14)
    swapped = dense.dense_from_terms(dims, (engine.LocalTerm((0, 2), op),))
    assert np.linalg.norm(h1 - swapped) > 0.1


def test_singlet_projection_matches_dense_projector():
    # R/tests/test_statevector_baseline.py:105-113
    rng = np.random.default_rng(4)
    state = rng.normal(size=12) + 1j * rng.normal(size=12)
    state /= np.linalg.norm(state)
    s = np.array([0.0, 1.0, -1.0, 0.0]) / np.sqrt(2.0)
    proj = np.kron(np.outer(s, s), np.eye(3))
    assert abs(engine.singlet_population(state, (2, 2, 3)) - np.vdot(state, proj @ state).real) < 1e-13


def test_quadrupole_operator_hermitian_and_rejects_asymmetric():
    t = np.array([[-0.2, 0.03, 0.0], [0.03, -0.1, 0.02], [0.0, 0.02, 0.3]])
    op = engine.quadrupole_operator(1.0, t)
    np.testing.assert_allclose(op, op.conj().T, atol=1e-14)
    with pytest.raises(ValueError):
        engine.quadrupole_operator(1.0, np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]]))
    # traceless Q on spin-1: trace of the operator is (I(I+1)) tr(Q) ... = 0 only if traceless
    assert abs(np.trace(op).real - 2.0 * np.trace(t)) < 1e-14


def test_non_hermitian_term_rejected():
    with pytest.raises(ValueError):
        engine.MatrixFreeHamiltonian((2,), (engine.LocalTerm((0,), np.array([[0, 1], [0, 0]])),))
    with pytest.raises(ValueError):
        engine.MatrixFreeHamiltonian((2, 2), (engine.LocalTerm((0, 0), np.eye(4)),))


def test_two_spin_coupling_isotropic_is_s_dot_i():
    sx, sy, sz = model.spin_operators(0.5)
    ix, iy, iz = model.spin_operators(1.0)
    op = engine.two_spin_coupling_operator(0.5, 1.0, 0.7 * np.eye(3))
    np.testing.assert_allclose(op, 0.7 * (np.kron(sx, ix) + np.kron(sy, iy) + np.kron(sz, iz)), atol=1e-15)
    # asymmetric tensor: T_xz S_x I_z, not S_z I_x
    t = np.zeros((3, 3)); t[0, 2] = 1.0
    np.testing.assert_allclose(engine.two_spin_coupling_operator(0.5, 1.0, t), np.kron(sx, iz), atol=1e-15)


def test_efimova_hore_operator_identities():
    # R/tests/test_jd_sweeps.py:80-106: J only -> -J(S^2-1) has eigenvalues J (singlet, S=0) and -J (triplet)
    op = engine.efimova_hore_electron_coupling_operator(1.3, 0.0, [1.0, 0.0, 0.0])
    s = np.array([0.0, 1.0, -1.0, 0.0]) / np.sqrt(2.0)
    assert abs(np.vdot(s, op @ s).real - 1.3) < 1e-14
    np.testing.assert_allclose(np.sort(np.linalg.eigvalsh(op)), [-1.3, -1.3, -1.3, 1.3], atol=1e-14)
    # D only: singlet is annihilated, triplet splitting D(1 - 2/3), D(0 - 2/3) along the axis
    d = engine.efimova_hore_electron_coupling_operator(0.0, 2.0, [0.0, 0.0, 3.0])
    assert np.linalg.norm(d @ s) < 1e-14
    np.testing.assert_allclose(np.sort(np.linalg.eigvalsh(d)), np.sort([0.0, 2.0 / 3, 2.0 / 3, -4.0 / 3]), atol=1e-14)
    with pytest.raises(ValueError):
        engine.efimova_hore_electron_coupling_operator(1.0, 1.0, [0.0, 0.0, 0.0])


def test_production_parameters_and_conversions():
    # R/tests/test_poster_reproduction.py:209-215; R/tests/test_jd_sweeps.py:186-194
    c = core.PRODUCTION_MODEL.condition
    assert c.d_paper_uT == pytest.approx(-347.875)
    assert c.exchange_J_uT == pytest.approx(115.95833333333333)
    assert core.equivalent_field_uT_to_MHz(-347.875) == pytest.approx(-9.7405)
    assert c.d_paper_MHz == pytest.approx(-9.7405)
    assert core.PRODUCTION_MODEL.quadrupole_scale == 1.0 and core.PRODUCTION_MODEL.lifetime_us == 5.0


def test_eh_crossing_at_ratio_minus_three():
    # R/tests/test_jd_sweeps.py:108-121: D/J=-3 makes the singlet degenerate with a triplet level
    c = core.efimova_hore_ratio_condition(-3.0)
    op = engine.efimova_hore_electron_coupling_operator(c.exchange_J_MHz, c.d_paper_MHz, [1, 0, 0])
    w = np.sort(np.linalg.eigvalsh(op))
    gaps = np.diff(w)
    assert gaps.min() < 1e-12


def test_generic_builder_quadrupole_axes_and_electron_term():
    # R/tests/test_statevector_baseline.py:115-153
    h = engine.build_quantum_needle_hamiltonian(90.0)
    q_axes = {t.axes for t in h.terms if "nuclear quadrupole" in t.label}
    assert q_axes == {(2,), (3,), (9,)}
    assert not any("nuclear quadrupole" in t.label for t in
                   engine.build_quantum_needle_hamiltonian(90.0, quadrupole_scale=0.0).terms)
    zero = engine.build_quantum_needle_hamiltonian(90.0, exchange_J_MHz=0.0, dipolar_tensor_MHz=np.zeros((3, 3)),
                                                   quadrupole_scale=0.0)
    assert not any(t.axes == (0, 1) for t in zero.terms)
    dip = np.diag([-0.5, -0.5, 1.0])
    coupled = engine.build_quantum_needle_hamiltonian(90.0, exchange_J_MHz=0.25, dipolar_tensor_MHz=dip,
                                                      quadrupole_scale=0.0)
    et = [t for t in coupled.terms if t.axes == (0, 1)]
    sx, sy, sz = model.spin_operators(0.5)
    expected = 2 * np.pi * sum(c * np.kron(o, o) for c, o in zip(np.diag(dip) - 0.5, (sx, sy, sz)))
    np.testing.assert_allclose(et[0].operator, expected, atol=1e-14)


def test_production_terms_hermitian_and_partitioned():
    h = core.build_hamiltonian(0.0, 0.05)
    labels = [t.label for t in h.terms]
    assert labels[:2] == ["electron A Zeeman", "electron B Zeeman"]
    assert sum("quadrupole" in s for s in labels) == 3
    assert sum("hyperfine" in s for s in labels) == 14
    assert h.left_axes == (0, 2, 3, 4, 5, 6, 7, 8) and h.right_axes == (1, 9, 10, 11, 12, 13, 14, 15)
    for t in h.terms:
        assert np.linalg.norm(t.operator - t.operator.conj().T) < 1e-11
