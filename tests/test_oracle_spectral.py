"""The factorized spectral oracle (oracle/spectral.py) pinned against the archived curve, the
reference's own toy self-test and a dense brute force; host parts of rpm_amd.spectral."""
import csv
import json
import os

import numpy as np
import pytest

from oracle import model as om
from oracle import spectral as osp
from rpm_amd import model as pm
from rpm_amd import spectral as ps

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "archived_fig1b_q0.csv")
TAUS = (1.0, 2.0, 5.0, 10.0, 20.0, 100.0)


def _archive():
    with open(GOLDEN, newline="") as f:
        return {float(r["theta_deg"]): {k: float(v) for k, v in r.items()} for r in csv.DictReader(f)}


def test_theta_grid_matches_archive():
    thetas = np.array(sorted(_archive()))
    assert len(thetas) == 199
    np.testing.assert_array_equal(osp.theta_grid(), thetas)
    np.testing.assert_array_equal(pm.theta_grid(), thetas)


@pytest.mark.parametrize("theta", [0.0, 88.75])
def test_oracle_reproduces_archived_curve(theta):
    """R/rpm_quantum_needle_reproduction.py:217-237 at Q = 0: the archived row to rounding."""
    fad = osp.build_radical("FAD", om.FAD_NUCLEI)
    trp = osp.build_radical("TrpH", om.TRP_NUCLEI)
    row = osp.solve_orientation(fad, trp, theta, 50.0, TAUS, 8.0)
    ref = _archive()[theta]
    assert row["delta_omega_rad_per_us"] == ref["delta_omega_rad_per_us"] == 0.00125
    assert abs(row["initial_singlet"] - 1.0) < 1e-12
    for tau in TAUS:
        key = f"phi_s_tau_{tau:g}_us"
        assert abs(row[key] - ref[key]) < 1e-13, key


def _toy():
    a = osp.build_radical("X", (om.Nucleus("N", 1.0, np.diag([-0.0989, -0.0989, 1.7569])),))
    b = osp.build_radical("Y", (om.Nucleus("N", 1.0, np.diag([0.0, 0.0, 1.0812])),))
    return a, b


def _brute_force_pair_yield(a, b, theta, field, taus):
    """Full-pair dense Hamiltonian H_A (x) 1 + 1 (x) H_B, rho0 = P_S / M, eigenbasis yield
    Phi = (1/M) sum_ij |P_S,ij|^2 k^2 / (k^2 + w_ij^2). Independent of the line formula."""
    fv = om.field_vector_mT(theta, field)

    def full(r):
        h = r.h0.copy()
        for c, op in zip(fv, r.electron_ops):
            h = h + om.TWOPI * om.GAMMA_E_MHZ_PER_MT * c * op
        return h

    ha, hb = full(a), full(b)
    ia, ib = np.eye(ha.shape[0]), np.eye(hb.shape[0])
    h = np.kron(ha, ib) + np.kron(ia, hb)
    sdot = sum(np.kron(sa, sb) for sa, sb in zip(a.electron_ops, b.electron_ops))
    ps_ = 0.25 * np.eye(h.shape[0]) - sdot
    w, v = np.linalg.eigh(h)
    p = v.conj().T @ ps_ @ v
    m = a.nuclear_dimension * b.nuclear_dimension
    dw = w[:, None] - w[None, :]
    out = {}
    for tau in taus:
        k = 1.0 / tau
        out[tau] = float(np.sum(np.abs(p) ** 2 * k * k / (k * k + dw * dw)) / m)
    return out


def test_toy_direct_lines_equal_brute_force():
    a, b = _toy()
    taus = (1.0, 10.0, 100.0)
    sa = osp.diagonalize_radical(a, 89.0, 50.0)
    sb = osp.diagonalize_radical(b, 89.0, 50.0)
    direct = osp.direct_line_yields(sa, sb, taus)
    brute = _brute_force_pair_yield(a, b, 89.0, 50.0, taus)
    for tau in taus:
        assert abs(direct[tau] - brute[tau]) < 1e-12


def test_toy_self_test_fft_route():
    """R/rpm_quantum_needle_reproduction.py:259-286: P_S(0) = 1 and FFT within 2e-4 of exact."""
    a, b = _toy()
    taus = (1.0, 10.0, 100.0)
    sa = osp.diagonalize_radical(a, 89.0, 50.0)
    sb = osp.diagonalize_radical(b, 89.0, 50.0)
    exact = osp.direct_line_yields(sa, sb, taus)
    omega, comb, initial = osp.convolved_correlation_spectrum(sa, sb, (1.0 / 100.0) / 16.0)
    approx = osp.yields_from_spectrum(omega, comb, taus)
    assert abs(initial - 1.0) < 2e-12
    assert max(abs(exact[t] - approx[t]) for t in taus) < 2e-4


def test_binning_conserves_weight_and_first_moment():
    """Linear interpolation keeps sum(c) and sum(c * omega) of the lines exactly (up to rounding)."""
    rng = np.random.default_rng(3)
    freq = rng.uniform(-2.0, 2.0, 500)
    coeff = rng.normal(size=500) + 1j * rng.normal(size=500)
    delta = 0.01
    g = osp.binned_spectrum(freq, coeff, delta, 205)
    grid = (np.arange(g.size) - 205) * delta
    assert abs(g.sum() - coeff.sum()) < 1e-12
    assert abs((g * grid).sum() - (coeff * freq).sum()) < 1e-11
    with pytest.raises(ValueError):
        osp.binned_spectrum(freq, coeff, delta, 150)


@pytest.mark.parametrize("scale", [0.0, 1.0])
def test_product_radical_matches_oracle(scale):
    for name, nuc_p, nuc_o in (("FAD", pm.FAD_NUCLEI, om.FAD_NUCLEI), ("TrpH", pm.TRP_NUCLEI, om.TRP_NUCLEI)):
        p = ps.build_radical(name, nuc_p, scale)
        o = osp.build_radical(name, nuc_o, om.QUADRUPOLE_TENSORS_MT, scale)
        assert p.dimension == o.h0.shape[0] and p.nuclear_dimension == o.nuclear_dimension
        np.testing.assert_allclose(p.h0, o.h0, rtol=0, atol=1e-13)
    assert ps.reference_models(0.0)[0].dimension == 576 and ps.reference_models(0.0)[1].dimension == 384


def test_product_field_components_match_model():
    th = np.array([0.0, 33.0, 90.0, 180.0])
    got = ps._field_components_mT(th, 50.0, 0.0)
    for i, t in enumerate(th):
        np.testing.assert_array_equal(got[i], om.field_vector_mT(t, 50.0))
    got = ps._field_components_mT(th, 50.0, 40.0)
    for i, t in enumerate(th):
        np.testing.assert_allclose(got[i], om.field_vector_mT(t, 50.0, 40.0), rtol=0, atol=1e-17)


def test_fast_length_is_minimal_smooth():
    def smooth(n):
        for p in (2, 3, 5, 7):
            while n % p == 0:
                n //= p
        return n == 1

    for n in list(range(1, 400)) + [1955659, 1934555]:
        m = ps._fast_length(n)
        assert m >= n and smooth(m)
        assert all(not smooth(k) for k in range(n, m))
