"""Factorized spectral yield oracle (J = D = 0), a CPU restatement.

Test infrastructure only: it is the checker for ``rpm_amd.spectral`` and is imported only by
``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py``.

With no electron-electron coupling the two radicals evolve independently, and the singlet yield of
the random-phase ensemble follows from each radical's electron-spin correlation spectrum:
- diagonalise each radical's Hamiltonian (electron + its own nuclei);
- list its transition lines (frequency, weight) for every pair (p, q) of electron axes;
- deposit the lines on a uniform frequency grid by linear interpolation;
- convolve the two radicals' grids (FFT) and sum over (p, q);
- Phi_S = 1/4 + Re sum_n c_n k / (k - i omega_n).

Follows R/rpm_quantum_needle_reproduction.py:93-256. Pinned by:
- the archived 199-angle curve results/quadrupole/zeeman_hyperfine_reference/quantum_needle_fig1b.csv
  (Q = 0, 50 uT, six lifetimes), checked at several angles against the committed fixture;
- the reference's own self-test (R/rpm_quantum_needle_reproduction.py:259-286): on the two-nitrogen
  toy pair the FFT route agrees with the exact line-by-line sum to 2e-4 and P_S(0) = 1 to 2e-12;
- brute force: on the toy pair, the exact line sum equals the yield of the dense full-pair
  Hamiltonian, computed independently from its eigenbasis.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import scipy.linalg
from scipy.fft import fft, ifft, next_fast_len

from .model import GAMMA_E_MHZ_PER_MT, TWOPI, field_vector_mT, spin_operators

MARGIN_BINS = 3  # R/rpm_quantum_needle_reproduction.py:180


@dataclass
class Radical:
    name: str
    h0: np.ndarray                      # dense, electron x nuclei, rad/us
    electron_ops: tuple                 # (Sx, Sy, Sz) embedded, dense
    nuclear_dimension: int


@dataclass
class Spectrum:
    eigenvalues: np.ndarray
    electron_ops_eigenbasis: tuple
    nuclear_dimension: int

    @property
    def max_gap(self) -> float:
        return float(np.ptp(self.eigenvalues))


def _embed(op, position, dims):
    out = np.ones((1, 1), dtype=complex)
    for i, d in enumerate(dims):
        out = np.kron(out, op if i == position else np.eye(d))
    return out


def build_radical(name, nuclei, quadrupole_tensors=None, quadrupole_scale=0.0) -> Radical:
    """Electron + its nuclei: sum_n 2 pi gamma S.A_n.I_n (+ Q terms).

    R/rpm_quantum_needle_reproduction.py:93-124.
    """
    nuclei = tuple(nuclei)
    dims = (2,) + tuple(int(round(2 * n.spin + 1)) for n in nuclei)
    s = tuple(_embed(o, 0, dims) for o in spin_operators(0.5))
    dim = math.prod(dims)
    h0 = np.zeros((dim, dim), dtype=complex)
    for index, nucleus in enumerate(nuclei, start=1):
        i_ops = tuple(_embed(o, index, dims) for o in spin_operators(nucleus.spin))
        for a in range(3):
            for b in range(3):
                c = nucleus.tensor_mT[a, b]
                if abs(c) > 1e-15:
                    h0 += TWOPI * GAMMA_E_MHZ_PER_MT * c * (s[a] @ i_ops[b])
        tensor = (quadrupole_tensors or {}).get(f"{name}:{nucleus.name}")
        if tensor is not None and quadrupole_scale:
            for a in range(3):
                for b in range(3):
                    h0 += TWOPI * GAMMA_E_MHZ_PER_MT * quadrupole_scale * tensor[a, b] * (i_ops[a] @ i_ops[b])
    return Radical(name, h0, s, math.prod(dims[1:]))


def diagonalize_radical(model: Radical, theta_deg, field_microT) -> Spectrum:
    """Add the Zeeman term, eigh, electron operators in the eigenbasis.

    R/rpm_quantum_needle_reproduction.py:127-142.
    """
    field = field_vector_mT(theta_deg, field_microT)
    h = model.h0.copy()
    for comp, op in zip(field, model.electron_ops):
        h = h + TWOPI * GAMMA_E_MHZ_PER_MT * comp * op
    w, v = scipy.linalg.eigh(h, check_finite=False, driver="evd")
    ops = tuple(v.conj().T @ (op @ v) for op in model.electron_ops)
    return Spectrum(w, ops, model.nuclear_dimension)


def transition_lines(spectrum: Spectrum, p: int, q: int):
    """omega_ij = lambda_i - lambda_j, weight T_p[i, j] T_q[j, i] / M.

    R/rpm_quantum_needle_reproduction.py:145-152.
    """
    lam = spectrum.eigenvalues
    freq = (lam[:, None] - lam[None, :]).reshape(-1)
    ops = spectrum.electron_ops_eigenbasis
    coeff = (ops[p] * ops[q].T / spectrum.nuclear_dimension).reshape(-1)
    return freq, coeff


def binned_spectrum(freq, coeff, delta_omega, half_width_bins):
    """Linear-interpolation deposit on 2 half + 1 bins centred at index half.

    R/rpm_quantum_needle_reproduction.py:155-172.
    """
    pos = freq / delta_omega + half_width_bins
    lower = np.floor(pos).astype(np.int64)
    frac = pos - lower
    size = 2 * half_width_bins + 1
    if lower.min() < 0 or lower.max() + 1 >= size:
        raise ValueError("Frequency grid is too small for transition spectrum.")
    out = np.zeros(size, dtype=complex)
    np.add.at(out, lower, coeff * (1.0 - frac))
    np.add.at(out, lower + 1, coeff * frac)
    return out


def half_width_bins(spectrum: Spectrum, delta_omega: float) -> int:
    return int(np.ceil(spectrum.max_gap / delta_omega)) + MARGIN_BINS


def convolved_correlation_spectrum(a: Spectrum, b: Spectrum, delta_omega: float):
    """Sum over (p, q) of the two radicals' binned spectra, convolved by FFT.

    Returns (omega, combined, initial_singlet). R/rpm_quantum_needle_reproduction.py:175-201.
    """
    ha, hb = half_width_bins(a, delta_omega), half_width_bins(b, delta_omega)
    la, lb = 2 * ha + 1, 2 * hb + 1
    out_len = la + lb - 1
    nfft = next_fast_len(out_len)
    acc = np.zeros(nfft, dtype=complex)
    for p in range(3):
        for q in range(3):
            ga = binned_spectrum(*transition_lines(a, p, q), delta_omega, ha)
            gb = binned_spectrum(*transition_lines(b, p, q), delta_omega, hb)
            acc += fft(ga, nfft) * fft(gb, nfft)
    combined = ifft(acc)[:out_len]
    omega = (np.arange(out_len, dtype=float) - (ha + hb)) * delta_omega
    return omega, combined, float((0.25 + np.sum(combined)).real)


def yields_from_spectrum(omega, coefficients, lifetimes_us):
    """Phi_S(tau) = 1/4 + Re sum c k / (k - i omega), k = 1/tau.

    R/rpm_quantum_needle_reproduction.py:204-214.
    """
    out = {}
    for tau in lifetimes_us:
        k = 1.0 / float(tau)
        out[float(tau)] = float((0.25 + np.sum(coefficients * (k / (k - 1j * omega)))).real)
    return out


def solve_orientation(fad: Radical, trp: Radical, theta_deg, field_microT, lifetimes_us, bins_per_linewidth=8.0):
    """One row of the Fig. 1B curve. R/rpm_quantum_needle_reproduction.py:217-237."""
    sa = diagonalize_radical(fad, theta_deg, field_microT)
    sb = diagonalize_radical(trp, theta_deg, field_microT)
    delta = (1.0 / max(lifetimes_us)) / bins_per_linewidth
    omega, comb, p0 = convolved_correlation_spectrum(sa, sb, delta)
    row = {"theta_deg": float(theta_deg), "initial_singlet": p0, "delta_omega_rad_per_us": delta}
    for tau, value in yields_from_spectrum(omega, comb, lifetimes_us).items():
        row[f"phi_s_tau_{tau:g}_us"] = value
    return row


def direct_line_yields(a: Spectrum, b: Spectrum, lifetimes_us):
    """Exact O(d_A^2 d_B^2) line-by-line sum, for small radicals.

    R/rpm_quantum_needle_reproduction.py:240-256.
    """
    out = {float(t): 0.25 for t in lifetimes_us}
    for p in range(3):
        for q in range(3):
            wa, ca = transition_lines(a, p, q)
            wb, cb = transition_lines(b, p, q)
            w = wa[:, None] + wb[None, :]
            c = ca[:, None] * cb[None, :]
            for tau in out:
                k = 1.0 / tau
                out[tau] += float(np.sum(c * k / (k - 1j * w)).real)
    return out


def theta_grid(step_deg=1.0, fine_step_deg=0.25, fine_half_width_deg=3.0):
    """0..180 in step_deg, plus fine_step_deg within +-fine_half_width_deg of 90.

    R/rpm_quantum_needle_model.py:202-213.
    """
    coarse = np.arange(0.0, 180.0 + 0.5 * step_deg, step_deg)
    fine = np.arange(90.0 - fine_half_width_deg, 90.0 + fine_half_width_deg + 0.5 * fine_step_deg, fine_step_deg)
    return np.unique(np.round(np.concatenate([coarse, fine]), 10))
