"""Host-side logic of rpm_amd that runs without a GPU.

* the C-ABI library loads and exports every symbol include/rpm_amd.h declares;
* the pair-stencil table compiler (``_pair_form``) reproduces the brute-force
  dense H when its tables are expanded back on the CPU;
* Simpson weights reproduce scipy.integrate.simpson;
* the product's model tables and validation rules match the oracle's.
"""

import ctypes
import math
import os
import re

import numpy as np
import pytest
import scipy.integrate

from oracle import core as ocore
from oracle import dense as odense
from oracle import engine as oengine
from oracle import model as omodel
from rpm_amd import _native, engine, model, quadrature
from rpm_amd import core as pcore

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "rpm_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rpm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    declared = _declared_symbols()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), name
        assert name in _native.SIGNATURES, f"{name} has no ctypes signature"
    assert lib.rpm_version().decode().startswith("rpm_amd")


def _declared_arg_counts():
    text = open(os.path.join(ROOT, "include", "rpm_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(rpm_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_ctypes_signatures_match_header_argument_counts():
    counts = _declared_arg_counts()
    assert set(counts) == set(_native.SIGNATURES)
    for name, n in counts.items():
        assert len(_native.SIGNATURES[name][1]) == n, name


def test_last_error_and_invalid_arguments_without_gpu():
    lib = _native.load()
    h = ctypes.c_void_p()
    # invalid: zero-length axis list is rejected before any device call
    rc = lib.rpm_operator_create_local(0, (ctypes.c_int * 1)(2), 0, None, None, None, ctypes.byref(h))
    assert rc == _native.RPM_E_INVALID
    with pytest.raises(ValueError):
        _native.check(rc)
    assert lib.rpm_operator_dimension(None) == -1
    assert lib.rpm_krylov_scratch_bytes(6, 64) >= 0
    assert lib.rpm_lanczos_workspace_bytes(None, 8) == 0
    assert lib.rpm_lanczos_prefix_workspace_bytes(None, 8, 1) == 0
    # unknown flag bits are rejected before the operator is touched
    one = ctypes.c_void_p(1)
    bad = 1 << 5
    assert not bad & (_native.RPM_LANCZOS_PAIRED | _native.RPM_LANCZOS_FORCE_SHORT)
    rc = lib.rpm_lanczos_basis_batch_prefix(one, 1, one, 4, one, 8, one, 0, 8, None, one, 8, one, one, one, one, 1e-14,
                                            bad, one, 1 << 20, None)
    assert rc == _native.RPM_E_INVALID


def test_spectral_entry_points_reject_bad_arguments_without_gpu():
    """Argument checks run before any device call (null buffers, empty or oversized shapes)."""
    lib = _native.load()
    vp = ctypes.c_void_p
    one = vp(1)
    # rpm_bin_lines: null buffers, a stride shorter than the grid, a non-positive bin width
    assert lib.rpm_bin_lines(None, one, 4, 1, 2.0, 0.1, 10, one, 21, 0, one, None) == _native.RPM_E_INVALID
    assert lib.rpm_bin_lines(one, one, 4, 1, 2.0, 0.1, 10, one, 20, 0, one, None) == _native.RPM_E_INVALID
    assert lib.rpm_bin_lines(one, one, 4, 1, 2.0, 0.0, 10, one, 21, 0, one, None) == _native.RPM_E_INVALID
    assert lib.rpm_bin_lines(one, one, 0, 1, 2.0, 0.1, 10, one, 21, 0, one, None) == _native.RPM_E_INVALID
    assert lib.rpm_bin_lines(one, one, 4, 1, 0.0, 0.1, 10, one, 21, 0, one, None) == _native.RPM_E_INVALID
    assert lib.rpm_bin_lines(one, one, 4, 1, 2.0, 0.1, 10, one, 21, 2, one, None) == _native.RPM_E_INVALID
    # rpm_spectral_yields: more than 8 lifetimes per pass, empty batch
    assert lib.rpm_spectral_yields(one, one, one, one, 2, 100, 9, 4, one, None) == _native.RPM_E_INVALID
    assert lib.rpm_spectral_yields(one, one, one, one, 0, 100, 6, 4, one, None) == _native.RPM_E_INVALID
    # rpm_eigh_batched: unknown method, workspace smaller than rpm_eigh_workspace_bytes
    need = lib.rpm_eigh_workspace_bytes(8, 3)
    assert need >= 3 * 4 + 8 * 3 * 8 and lib.rpm_eigh_workspace_bytes(0, 3) == 0
    assert lib.rpm_eigh_batched(one, 8, 3, one, one, need, 2, None) == _native.RPM_E_INVALID
    assert lib.rpm_eigh_batched(one, 8, 3, one, one, need - 1, 0, None) == _native.RPM_E_INVALID
    with pytest.raises(ValueError):
        _native.check(lib.rpm_eigh_batched(None, 8, 3, one, one, need, 0, None))


def _expand_pair(dims, form):
    """Dense H from the pair-stencil tables (the kernel's formula, written with Kronecker products)."""
    ldims, rdims, blocks, pair = form
    nl, nr = len(ldims), len(rdims)
    full = (2, 2) + tuple(ldims) + tuple(rdims)
    assert full == tuple(dims)
    terms = [oengine.LocalTerm((0, 1), pair)]
    for i in range(nl):
        terms.append(oengine.LocalTerm((0, 2 + i), blocks[i]))
    for i in range(nr):
        terms.append(oengine.LocalTerm((1, 2 + nl + i), blocks[nl + i]))
    return odense.dense_from_terms(dims, terms)


def _random_pair_terms(rng, dims_l, dims_r):
    dims = (2, 2) + tuple(dims_l) + tuple(dims_r)
    terms = []

    def herm(d):
        a = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
        return a + a.conj().T

    terms.append(engine.LocalTerm((0,), herm(2)))
    terms.append(engine.LocalTerm((1,), herm(2)))
    terms.append(engine.LocalTerm((1, 0), herm(4)))          # reversed pair order
    for i, d in enumerate(dims_l):
        ax = 2 + i
        terms.append(engine.LocalTerm((0, ax) if i % 2 == 0 else (ax, 0), herm(2 * d)))
        terms.append(engine.LocalTerm((ax,), herm(d)))
    for i, d in enumerate(dims_r):
        ax = 2 + len(dims_l) + i
        terms.append(engine.LocalTerm((1, ax) if i % 2 else (ax, 1), herm(2 * d)))
    return dims, terms


def test_pair_form_tables_reproduce_brute_force():
    rng = np.random.default_rng(3)
    dims, terms = _random_pair_terms(rng, (3, 2), (2, 3))
    compiled = [engine._validate_term(dims, t, 1e-12) for t in terms]
    form = engine._pair_form(dims, compiled)
    assert form is not None
    brute = odense.dense_from_terms(dims, compiled)
    np.testing.assert_allclose(_expand_pair(dims, form), brute, atol=1e-12)


def test_pair_form_rejects_unsupported_layouts():
    sx = model.spin_operators(0.5)[0]
    dims = (2, 2, 2, 2)
    # nucleus-nucleus coupling is not an (electron, nucleus) term
    bad = [engine._validate_term(dims, engine.LocalTerm((2, 3), np.kron(sx, sx)), 1e-12)]
    assert engine._pair_form(dims, bad) is None
    # interleaved radicals (right nucleus before a left one) do not fit the canonical layout
    terms = [engine._validate_term(dims, engine.LocalTerm((1, 2), np.kron(sx, sx)), 1e-12),
             engine._validate_term(dims, engine.LocalTerm((0, 3), np.kron(sx, sx)), 1e-12)]
    assert engine._pair_form(dims, terms) is None


def test_production_hamiltonian_uses_pair_stencil_and_matches_oracle_terms():
    h = pcore.build_hamiltonian(37.0, 0.05)
    o = ocore.build_hamiltonian(37.0, 0.05)
    assert h.uses_pair_stencil
    assert [t.label for t in h.terms] == [t.label for t in o.terms]
    for a, b in zip(h.terms, o.terms):
        assert a.axes == b.axes
        if "quadrupole" in a.label:
            # the product uses the published tensors; the oracle rebuilds them with the local
            # LAPACK, which may differ in the last bit
            np.testing.assert_allclose(a.operator, b.operator, rtol=0, atol=1e-13)
        else:
            np.testing.assert_array_equal(a.operator, b.operator)
    form = engine._pair_form(h.dimensions, h.terms, h.left_axes, h.right_axes)
    ldims, rdims, blocks, pair = form
    assert ldims == [3, 3, 2, 2, 2, 2, 2] and rdims == [3, 2, 2, 2, 2, 2, 2]
    # the expanded stencil equals L (x) I + I (x) R + pair of the oracle's compiled form on a
    # reduced (smoke) system that is small enough to expand densely
    nuclei = (omodel.FAD_NUCLEI[0], omodel.FAD_NUCLEI[1], omodel.TRP_NUCLEI[0])
    small = pcore.build_hamiltonian(90.0, 0.05, nuclei=nuclei, radicals=("FAD", "FAD", "TrpH"))
    small_o = ocore.build_hamiltonian(90.0, 0.05, nuclei=nuclei, radicals=("FAD", "FAD", "TrpH"))
    f = engine._pair_form(small.dimensions, small.terms, small.left_axes, small.right_axes)
    np.testing.assert_allclose(_expand_pair(small.dimensions, f), small_o.explicit_matrix(), atol=1e-11)


def test_model_tables_match_oracle():
    assert model.SYSTEM_DIMENSIONS == omodel.SYSTEM_DIMENSIONS
    assert model.SPIN_ORDER == omodel.SPIN_ORDER
    for a, b in zip(model.ALL_NUCLEI, omodel.ALL_NUCLEI):
        assert a.name == b.name and a.spin == b.spin
        np.testing.assert_array_equal(a.tensor_mT, b.tensor_mT)
    for k in omodel.QUADRUPOLE_TENSORS_MT:
        np.testing.assert_allclose(model.QUADRUPOLE_TENSORS_MT[k], omodel.QUADRUPOLE_TENSORS_MT[k], atol=1e-16)
    for s in (0.5, 1.0):
        for a, b in zip(model.spin_operators(s), omodel.spin_operators(s)):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("x", [
    np.linspace(0.0, 1.0, 5), np.linspace(0.0, 1.0, 6), np.array([0.0, 0.3]),
    np.array([0.0, 0.1, 0.25, 0.3, 0.7, 0.71, 1.0]), np.array([0.0, 0.1, 0.25, 0.3, 0.7, 1.0]),
    np.append(np.arange(101) * 0.00125, 0.1259), np.append(np.arange(100) * 0.00125, 0.1241),
])
def test_simpson_weights_match_scipy(x):
    rng = np.random.default_rng(1)
    y = rng.normal(size=x.size)
    assert abs(np.dot(quadrature.simpson_weights(x), y) - scipy.integrate.simpson(y, x=x)) < 1e-14


@pytest.mark.parametrize("x", [
    np.linspace(0.0, 1.0, 5), np.linspace(0.0, 1.0, 6), np.array([0.0, 0.3]),
    np.array([0.0, 0.1, 0.25, 0.3, 0.7, 0.71, 1.0]),
])
def test_integrate_singlet_yield_rules(x):
    """R/rpm_quantum_needle_validation.py:192-210: both rules against the library routines."""
    rng = np.random.default_rng(2)
    p = rng.uniform(size=x.size)
    f = np.exp(-x / 0.7) * p / 0.7
    assert abs(pcore.integrate_singlet_yield(x, p, 0.7) - scipy.integrate.simpson(f, x=x)) < 1e-14
    assert abs(pcore.integrate_singlet_yield(x, p, 0.7, method="trapezoid") - np.trapezoid(f, x)) < 1e-14
    assert pcore.integrate_singlet_yield(x, p, 0.7) == pcore.integrate_yield(x, p, 0.7)


def test_integrate_singlet_yield_errors():
    t = np.linspace(0.0, 1.0, 5)
    with pytest.raises(ValueError, match="matching 1-D"):
        pcore.integrate_singlet_yield(t, np.ones(4), 1.0)
    with pytest.raises(ValueError, match="matching 1-D"):
        pcore.integrate_singlet_yield(t[None, :], np.ones((1, 5)), 1.0)
    with pytest.raises(ValueError, match="Unknown yield-integration method"):
        pcore.integrate_singlet_yield(t, np.ones(5), 1.0, method="midpoint")


def test_production_grid_yield_weights():
    cfg = pcore.SweepConfig(lifetime_us=5.0, survival_tail_bound=1e-5)
    t = cfg.time_grid_us
    assert t.size == 46053
    pops = 0.25 + 0.1 * np.cos(3.0 * t)
    assert abs(pcore.integrate_yield(t, pops, 5.0) - ocore.integrate_yield(t, pops, 5.0)) < 1e-14
    np.testing.assert_array_equal(pcore.coarse_grid_indices(cfg), ocore.coarse_grid_indices(ocore.SweepConfig(
        lifetime_us=5.0, survival_tail_bound=1e-5)))


def test_validation_errors_match_reference_rules():
    sx = model.spin_operators(0.5)[0]
    with pytest.raises(ValueError, match="unique active axes"):
        engine.MatrixFreeHamiltonian((2, 2), [engine.LocalTerm((0, 0), np.eye(4))])
    with pytest.raises(ValueError, match="out-of-range"):
        engine.MatrixFreeHamiltonian((2, 2), [engine.LocalTerm((2,), np.eye(2))])
    with pytest.raises(ValueError, match="not Hermitian"):
        engine.MatrixFreeHamiltonian((2,), [engine.LocalTerm((0,), np.array([[0, 1], [0, 0]]))])
    with pytest.raises(ValueError, match="shape"):
        engine.MatrixFreeHamiltonian((2,), [engine.LocalTerm((0,), np.eye(3))])
    with pytest.raises(ValueError, match="Cross-radical"):
        engine.CompiledFullPairHamiltonian((2, 2, 2), [engine.LocalTerm((1, 2), np.kron(sx, sx))],
                                           left_axes=(0, 2), right_axes=(1,)) if False else \
            engine.CompiledFullPairHamiltonian((2, 2, 2, 2), [engine.LocalTerm((1, 2), np.kron(sx, sx))],
                                               left_axes=(0, 2), right_axes=(1, 3))
    with pytest.raises(ValueError, match="partition"):
        engine.CompiledFullPairHamiltonian((2, 2, 2), [], left_axes=(0,), right_axes=(1,))
    with pytest.raises(ValueError):
        engine.AdaptiveLanczosPropagator(None, minimum_dimension=4, embedded_offset=4)
    with pytest.raises(ValueError):
        engine.quadrupole_operator(1.0, np.array([[0.0, 1.0, 0.0], [0.0] * 3, [0.0] * 3]))


def test_seeds_and_initial_state_digests_match_archive():
    import csv
    with open(os.path.join(ROOT, "tests", "golden", "archived_full14_yields.csv")) as f:
        rows = list(csv.DictReader(f))
    digests = {int(r["sample_id"]): r["initial_state_sha256"] for r in rows}
    for sid, digest in digests.items():
        s = pcore.random_phase_initial_state(model.SYSTEM_DIMENSIONS, pcore.MASTER_SEED, sid)
        assert pcore.initial_state_sha256(s) == digest
    assert pcore.sample_seed(5, 7) == ocore.seed_for(5, 7)


def test_fibonacci_sphere_directions():
    from rpm_amd import workloads
    for n in (1, 7, 64):
        th, ph = workloads.fibonacci_sphere(n)
        assert th.shape == ph.shape == (n,)
        z = np.cos(np.radians(th))
        np.testing.assert_allclose(z, 1.0 - (2.0 * np.arange(n) + 1.0) / n, atol=1e-12)   # equal-area bands
        assert np.all((ph >= 0.0) & (ph < 360.0))
    th, ph = workloads.fibonacci_sphere(64)
    v = np.stack([np.sin(np.radians(th)) * np.cos(np.radians(ph)), np.sin(np.radians(th)) * np.sin(np.radians(ph)),
                  np.cos(np.radians(th))], axis=1)
    assert np.abs(v.mean(axis=0)).max() < 0.01          # no preferred direction
    with pytest.raises(ValueError):
        workloads.fibonacci_sphere(0)


def test_yield_weight_table_matches_scipy_on_fine_and_coarse_grids():
    """Column 0 reproduces simpson(exp(-t/tau) f / tau, t); column 1 the same on the coarse sub-grid."""
    cfg = pcore.SweepConfig(lifetime_us=1.0, survival_tail_bound=math.exp(-0.6))
    t = cfg.time_grid_us
    ids = pcore.coarse_grid_indices(cfg)
    f = 0.5 + 0.3 * np.cos(7.0 * t) * np.exp(-t)
    tab = quadrature.yield_weight_table(t, 1.0, ids)
    g = np.exp(-t) * f
    assert abs(tab[:, 0] @ f - scipy.integrate.simpson(g, x=t)) < 1e-15
    assert abs(tab[:, 1] @ f - scipy.integrate.simpson(g[ids], x=t[ids])) < 1e-15
    assert np.all(tab[np.setdiff1d(np.arange(t.size), ids), 1] == 0.0)
    assert np.all(quadrature.yield_weight_table(t, 1.0)[:, 1] == 0.0)
