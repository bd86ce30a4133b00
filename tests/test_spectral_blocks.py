"""Host side of the multiplet-block route of rpm_amd.spectral (no GPU).

Equivalent spin-1/2 nuclei (identical hyperfine tensors) enter H only through their total spin, so
the radical splits into multiplet blocks. Pinned by what the mathematics fixes:
- the multiplet count: sum over I of multiplicity x (2I + 1) = 2^n, and the known n = 2, 3 cases;
- each multiplet basis is orthonormal and carries the textbook spin-I matrices;
- FAD's blocks reproduce the full 576-state radical: eigenvalues (with multiplicity) and the binned
  transition-line grids of the oracle, at several fields and with the quadrupole terms on;
- a group that h0 does not treat symmetrically is left whole.
"""
import math

import numpy as np
import pytest

from oracle import model as om
from oracle import spectral as osp
from rpm_amd import model as pm
from rpm_amd import spectral as ps


def _dense(model, theta, phi=0.0):
    """Eigenvalues and electron operators in the eigenbasis (numpy, electron index slowest)."""
    fv = om.field_vector_mT(theta, 50.0, phi)
    m = model.dimension // 2
    s = [np.kron(o, np.eye(m)) for o in om.spin_operators(0.5)]
    h = model.h0 + om.TWOPI * om.GAMMA_E_MHZ_PER_MT * sum(c * o for c, o in zip(fv, s))
    w, v = np.linalg.eigh(h)
    return w, [v.conj().T @ o @ v for o in s]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_multiplet_counts(n):
    mult = ps._multiplets(n)
    assert sum(c * (2 * spin + 1) for spin, c in mult) == 2 ** n
    assert all(c >= 1 for _, c in mult)
    if n == 2:
        assert mult == [(1.0, 1), (0.0, 1)]
    if n == 3:
        assert mult == [(1.5, 1), (0.5, 2)]


@pytest.mark.parametrize("n,spin", [(2, 1.0), (2, 0.0), (3, 1.5), (3, 0.5), (4, 1.0)])
def test_multiplet_basis_carries_spin_matrices(n, spin):
    b = ps._multiplet_basis(n, spin)
    assert b.shape == (2 ** n, int(round(2 * spin + 1)))
    assert np.abs(b.conj().T @ b - np.eye(b.shape[1])).max() < 1e-13
    dims = (2,) * n
    for o, want in zip(pm.spin_operators(0.5), pm.spin_operators(spin)):
        total = sum(ps._embed(o, k, dims) for k in range(n))
        assert np.abs(b.conj().T @ total @ b - want).max() < 1e-13
        # the span is invariant: total-spin operators do not leave the multiplet
        assert np.abs(total @ b - b @ (b.conj().T @ total @ b)).max() < 1e-13


@pytest.mark.parametrize("scale", [0.0, 1.0])
def test_fad_blocks_reproduce_the_full_radical(scale):
    fad, trp = ps.build_radical("FAD", pm.FAD_NUCLEI, scale), ps.build_radical("TrpH", pm.TRP_NUCLEI, scale)
    blocks = ps.symmetry_blocks(fad)
    assert [(b.dimension, b.multiplicity) for b in blocks] == [(288, 1), (144, 2)]
    assert all(b.nuclear_dimension == 288 for b in blocks)
    assert sum(b.dimension * b.multiplicity for b in blocks) == fad.dimension
    assert ps.symmetry_blocks(trp) == (trp,)
    for theta, phi in ((37.0, 0.0), (90.0, 0.0), (123.0, 40.0)):
        w, ops = _dense(fad, theta, phi)
        parts = [(_dense(b, theta, phi), b.multiplicity) for b in blocks]
        merged = np.sort(np.concatenate([np.repeat(wb, mult) for (wb, _), mult in parts]))
        assert np.abs(merged - w).max() < 1e-10 * np.abs(w).max()
        delta = 0.00125
        half = int(np.ceil(np.ptp(w) / delta)) + osp.MARGIN_BINS
        for p, q in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)):
            full = osp.binned_spectrum((w[:, None] - w[None, :]).ravel(), (ops[p] * ops[q].T / 288).ravel(),
                                       delta, half)
            acc = np.zeros_like(full)
            for (wb, ob), mult in parts:
                acc += osp.binned_spectrum((wb[:, None] - wb[None, :]).ravel(),
                                           (ob[p] * ob[q].T * mult / 288).ravel(), delta, half)
            # eigenvalue rounding (~1e-13 rad/us) moves a line by ~1e-10 of a bin
            assert np.abs(acc - full).max() < 1e-8 * np.abs(full).max(), (theta, p, q)
            assert abs(acc.sum() - full.sum()) < 1e-13, (theta, p, q)


def test_asymmetric_group_is_left_whole():
    """Tensors that are equal in the nuclei list but not in h0: the projection check refuses the split."""
    iso = pm.Nucleus("Ha", 0.5, np.eye(3) * 0.44)
    other = pm.Nucleus("Hb", 0.5, np.eye(3) * 0.41)
    model = ps.build_radical("X", (iso, other), 0.0)
    claimed = ps.RadicalModel("X", (iso, pm.Nucleus("Hb", 0.5, np.eye(3) * 0.44)), model.h0, model.nuclear_dimension)
    assert ps.symmetry_blocks(claimed) == (claimed,)
    # the same pair with truly equal tensors splits into a triplet and a singlet
    pair = ps.build_radical("X", (iso, pm.Nucleus("Hb", 0.5, np.eye(3) * 0.44)), 0.0)
    assert [(b.dimension, b.multiplicity) for b in ps.symmetry_blocks(pair)] == [(6, 1), (2, 1)]
    assert math.isclose(sum(b.dimension for b in ps.symmetry_blocks(pair)), pair.dimension)
