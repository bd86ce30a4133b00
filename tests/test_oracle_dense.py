"""CPU: the piecewise dense yield oracle used to pin adjoint gradients."""

import math

import numpy as np
import scipy.linalg

from oracle import dense


def _random_h(n, seed, scale=30.0):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return scale * (a + a.conj().T) / 2


def _state(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=n) + 1j * rng.normal(size=n)
    return v / np.linalg.norm(v)


def test_dense_grid_yield_single_piece_matches_expm_and_closed_form():
    n = 16
    h = _random_h(n, 1)
    psi = _state(n, 2)
    t = np.linspace(0.0, 0.5, 401)
    phi, pops = dense.dense_grid_yield([(t[-1], h)], psi, t, 1.0)
    for k in (0, 7, 200, 400):
        v = scipy.linalg.expm(-1j * h * t[k]) @ psi
        s = (v[4:8] - v[8:12]) / math.sqrt(2.0)
        assert abs(pops[k] - np.vdot(s, s).real) < 1e-12
    exact = dense.dense_singlet_yield_closed_form(h, psi, 1.0, t[-1])
    assert abs(phi - exact) < 1e-6


def test_dense_grid_yield_pieces_compose():
    n = 16
    h1, h2 = _random_h(n, 3), _random_h(n, 4)
    psi = _state(n, 5)
    t = np.linspace(0.0, 0.2, 161)
    whole, _ = dense.dense_grid_yield([(0.2, h1)], psi, t, 1.0)
    split, _ = dense.dense_grid_yield([(0.05, h1), (0.2, h1)], psi, t, 1.0)
    assert abs(whole - split) < 1e-13
    # two different pieces: the second piece starts from the first piece's end state
    _, pops = dense.dense_grid_yield([(0.1, h1), (0.2, h2)], psi, t, 1.0)
    mid = scipy.linalg.expm(-1j * h1 * 0.1) @ psi
    end = scipy.linalg.expm(-1j * h2 * 0.1) @ mid
    s = (end[4:8] - end[8:12]) / math.sqrt(2.0)
    assert abs(pops[-1] - np.vdot(s, s).real) < 1e-12
