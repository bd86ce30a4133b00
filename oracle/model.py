"""Physical model of the FAD-TrpH radical pair (oracle restatement).

Follows R/rpm_quantum_needle_model.py:16-199.  Frequencies are angular, in
rad/us.  Hyperfine and quadrupole tensors are field-equivalent mT and are
converted with 2*pi*28 MHz/mT.  Local spin bases are ordered m = I, I-1, ..., -I.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

TWOPI = 2.0 * np.pi                      # R/rpm_quantum_needle_model.py:16
GAMMA_E_MHZ_PER_MT = 28.0                # R/rpm_quantum_needle_model.py:17
DEFAULT_FIELD_MICROT = 50.0              # R/rpm_quantum_needle_model.py:18


@dataclass(frozen=True)
class Nucleus:
    """One magnetic nucleus (R/rpm_quantum_needle_model.py:24-34)."""

    name: str
    spin: float
    tensor_mT: np.ndarray

    @property
    def dimension(self) -> int:
        return int(round(2.0 * self.spin + 1.0))


def _sym(xx, xy, xz, yy, yz, zz) -> np.ndarray:
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]], dtype=float)


def _iso(value: float) -> np.ndarray:
    return value * np.eye(3)


# Hyperfine tensors (mT), Hiscock et al. 2016 SI Tables S1-S2, as tabulated in
# R/rpm_quantum_needle_model.py:37-56 and methods/hyperfine_tensors.json.
FAD_NUCLEI = (
    Nucleus("N5", 1.0, _sym(-0.0989, 0.0039, 0.0, -0.0881, 0.0, 1.7569)),
    Nucleus("N10", 1.0, _sym(-0.0190, -0.0048, 0.0, -0.0196, 0.0, 0.6046)),
    Nucleus("H6", 0.5, _sym(-0.2569, -0.1273, 0.0, -0.4711, 0.0, -0.4336)),
    Nucleus("H8a", 0.5, _iso(0.4399)),
    Nucleus("H8b", 0.5, _iso(0.4399)),
    Nucleus("H8c", 0.5, _iso(0.4399)),
    Nucleus("Hbeta", 0.5, _iso(0.4070)),
)
TRP_NUCLEI = (
    Nucleus("N1", 1.0, _sym(-0.0336, 0.0924, -0.1354, 0.3303, -0.5318, 0.6680)),
    Nucleus("H1", 0.5, _sym(-0.9920, -0.2091, -0.2003, -0.2631, 0.2803, -0.5398)),
    Nucleus("H2", 0.5, _sym(-0.2843, 0.1757, 0.1525, -0.2798, 0.0975, -0.2699)),
    Nucleus("H4", 0.5, _sym(-0.5596, -0.1956, -0.1657, -0.4020, 0.0762, -0.5021)),
    Nucleus("H5", 0.5, _sym(0.0051, 0.0616, 0.0694, -0.0665, 0.0391, -0.0586)),
    Nucleus("H7", 0.5, _sym(-0.4355, -0.1541, -0.1239, -0.2777, 0.0864, -0.3770)),
    Nucleus("Hbeta1", 0.5, _sym(1.5808, -0.0453, -0.0506, 1.5575, 0.0988, 1.6752)),
)

ALL_NUCLEI = FAD_NUCLEI + TRP_NUCLEI
NUCLEAR_RADICALS = ("FAD",) * len(FAD_NUCLEI) + ("TrpH",) * len(TRP_NUCLEI)
SPIN_ORDER = ("eA", "eB") + tuple(
    f"{radical}:{nucleus.name}" for radical, nucleus in zip(NUCLEAR_RADICALS, ALL_NUCLEI)
)
SYSTEM_DIMENSIONS = (2, 2) + tuple(n.dimension for n in ALL_NUCLEI)
NUCLEAR_DIMENSION = math.prod(SYSTEM_DIMENSIONS[2:])
STATE_DIMENSION = math.prod(SYSTEM_DIMENSIONS)


def _quadrupole_in_hyperfine_frame(hyperfine_mT, principal_mT, ring_normal: bool) -> np.ndarray:
    """Traceless Q assigned to hyperfine principal axes.

    Follows R/rpm_quantum_needle_model.py:69-102: subtract the triplet mean,
    find the hyperfine axis of largest |eigenvalue| (the ring normal), and
    either put Q_zz along it (``ring_normal``) or put Q_zz on the first in-plane
    axis, Q_xx on the second and Q_yy along the normal (provisional N5 frame).
    """
    values = np.asarray(principal_mT, dtype=float)
    values = values - np.mean(values)
    hf_values, hf_axes = np.linalg.eigh(hyperfine_mT)
    normal = int(np.argmax(np.abs(hf_values)))
    plane = [i for i in range(3) if i != normal]
    diag = np.empty(3)
    if ring_normal:
        diag[normal], diag[plane[0]], diag[plane[1]] = values[2], values[0], values[1]
    else:
        diag[plane[0]], diag[plane[1]], diag[normal] = values[2], values[0], values[1]
    tensor = hf_axes @ np.diag(diag) @ hf_axes.T
    return 0.5 * (tensor + tensor.T)


# Efimova & Hore 2009 Table 2 principal values (R/rpm_quantum_needle_model.py:109-113).
QUADRUPOLE_PRINCIPAL_VALUES_MT = {
    "FAD:N5": (0.031, 0.041, -0.072),
    "FAD:N10": (0.032, 0.041, -0.072),
    "TrpH:N1": (0.016, 0.020, -0.036),
}
QUADRUPOLE_TENSORS_MT = {
    "FAD:N5": _quadrupole_in_hyperfine_frame(
        FAD_NUCLEI[0].tensor_mT, QUADRUPOLE_PRINCIPAL_VALUES_MT["FAD:N5"], ring_normal=False),
    "FAD:N10": _quadrupole_in_hyperfine_frame(
        FAD_NUCLEI[1].tensor_mT, QUADRUPOLE_PRINCIPAL_VALUES_MT["FAD:N10"], ring_normal=True),
    "TrpH:N1": _quadrupole_in_hyperfine_frame(
        TRP_NUCLEI[0].tensor_mT, QUADRUPOLE_PRINCIPAL_VALUES_MT["TrpH:N1"], ring_normal=True),
}


def spin_operators(spin: float):
    """(Sx, Sy, Sz) in the descending-m basis (R/rpm_quantum_needle_model.py:178-192)."""
    d = int(round(2.0 * spin + 1.0))
    m = spin - np.arange(d, dtype=float)
    splus = np.zeros((d, d), dtype=complex)
    for row in range(d - 1):
        # <m+1| S+ |m> with m = m[row + 1]
        mm = m[row + 1]
        splus[row, row + 1] = math.sqrt(spin * (spin + 1.0) - mm * (mm + 1.0))
    sminus = splus.conj().T
    return 0.5 * (splus + sminus), (splus - sminus) / 2.0j, np.diag(m.astype(complex))


def field_vector_mT(theta_deg: float, field_microT: float = DEFAULT_FIELD_MICROT,
                    phi_deg: float = 0.0) -> np.ndarray:
    """Field in mT.  At phi=0 this is R/rpm_quantum_needle_model.py:195-199
    (B in the molecular zx plane); phi rotates it about z."""
    theta = np.deg2rad(theta_deg)
    if phi_deg == 0.0:
        return 1e-3 * field_microT * np.array([np.sin(theta), 0.0, np.cos(theta)])
    phi = np.deg2rad(phi_deg)
    return 1e-3 * field_microT * np.array(
        [np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
