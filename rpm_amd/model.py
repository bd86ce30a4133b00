"""FAD-TrpH radical-pair model: spin order, dimensions, tensors, spin operators.

Same names and values as R/rpm_quantum_needle_model.py:16-199.  Angular
frequencies in rad/us; tensors in field-equivalent mT; local bases ordered
m = I, I-1, ..., -I.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

TWOPI = 2.0 * np.pi
GAMMA_E_MHZ_PER_MT = 28.0
DEFAULT_FIELD_MICROT = 50.0
DEFAULT_LIFETIMES_US = (1.0, 2.0, 5.0, 10.0, 20.0, 100.0)
NOMINAL_QUADRUPOLE_SCALE = 1.0
PILOT_ANGLES_DEG = (0.0, 80.0, 90.0, 100.0, 180.0)
DIAGNOSTIC_ANGLES_DEG = PILOT_ANGLES_DEG
MT_TO_RAD_PER_US = TWOPI * GAMMA_E_MHZ_PER_MT


@dataclass(frozen=True)
class Nucleus:
    name: str
    spin: float
    tensor_mT: np.ndarray

    @property
    def dimension(self) -> int:
        return int(round(2.0 * self.spin + 1.0))


# (name, spin, (xx, xy, xz, yy, yz, zz)) in mT -- Hiscock et al. 2016 SI Tables S1-S2
_FAD_TABLE = (
    ("N5", 1.0, (-0.0989, 0.0039, 0.0, -0.0881, 0.0, 1.7569)),
    ("N10", 1.0, (-0.0190, -0.0048, 0.0, -0.0196, 0.0, 0.6046)),
    ("H6", 0.5, (-0.2569, -0.1273, 0.0, -0.4711, 0.0, -0.4336)),
    ("H8a", 0.5, (0.4399, 0.0, 0.0, 0.4399, 0.0, 0.4399)),
    ("H8b", 0.5, (0.4399, 0.0, 0.0, 0.4399, 0.0, 0.4399)),
    ("H8c", 0.5, (0.4399, 0.0, 0.0, 0.4399, 0.0, 0.4399)),
    ("Hbeta", 0.5, (0.4070, 0.0, 0.0, 0.4070, 0.0, 0.4070)),
)
_TRP_TABLE = (
    ("N1", 1.0, (-0.0336, 0.0924, -0.1354, 0.3303, -0.5318, 0.6680)),
    ("H1", 0.5, (-0.9920, -0.2091, -0.2003, -0.2631, 0.2803, -0.5398)),
    ("H2", 0.5, (-0.2843, 0.1757, 0.1525, -0.2798, 0.0975, -0.2699)),
    ("H4", 0.5, (-0.5596, -0.1956, -0.1657, -0.4020, 0.0762, -0.5021)),
    ("H5", 0.5, (0.0051, 0.0616, 0.0694, -0.0665, 0.0391, -0.0586)),
    ("H7", 0.5, (-0.4355, -0.1541, -0.1239, -0.2777, 0.0864, -0.3770)),
    ("Hbeta1", 0.5, (1.5808, -0.0453, -0.0506, 1.5575, 0.0988, 1.6752)),
)


def symmetric_tensor(xx, xy, xz, yy, yz, zz) -> np.ndarray:
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]], dtype=float)


def _nuclei(table):
    return tuple(Nucleus(name, spin, symmetric_tensor(*vals)) for name, spin, vals in table)


FAD_NUCLEI = _nuclei(_FAD_TABLE)
TRP_NUCLEI = _nuclei(_TRP_TABLE)
ALL_NUCLEI = FAD_NUCLEI + TRP_NUCLEI
NUCLEAR_RADICALS = ("FAD",) * len(FAD_NUCLEI) + ("TrpH",) * len(TRP_NUCLEI)
SPIN_ORDER = ("eA", "eB") + tuple(f"{r}:{n.name}" for r, n in zip(NUCLEAR_RADICALS, ALL_NUCLEI))
SYSTEM_DIMENSIONS = (2, 2) + tuple(n.dimension for n in ALL_NUCLEI)
NUCLEAR_DIMENSION = math.prod(SYSTEM_DIMENSIONS[2:])
STATE_DIMENSION = math.prod(SYSTEM_DIMENSIONS)

# Efimova & Hore 2009, Table 2: principal values (mT) and frame choice
QUADRUPOLE_PRINCIPAL_VALUES_MT = {
    "FAD:N5": (0.031, 0.041, -0.072),
    "FAD:N10": (0.032, 0.041, -0.072),
    "TrpH:N1": (0.016, 0.020, -0.036),
}
_QUADRUPOLE_FRAME = {"FAD:N5": "ring_plane_provisional", "FAD:N10": "ring_normal", "TrpH:N1": "ring_normal"}


def quadrupole_tensor_in_hyperfine_frame(hyperfine_mT, principal_mT, orientation: str) -> np.ndarray:
    """Traceless Q on the hyperfine principal axes (R/rpm_quantum_needle_model.py:69-102).

    The hyperfine axis with the largest |eigenvalue| is the ring normal.
    ``ring_normal``: Q_zz on the normal, Q_xx, Q_yy on the in-plane axes in
    order.  ``ring_plane_provisional``: Q_zz on the first in-plane axis, Q_xx
    on the second, Q_yy on the normal.
    """
    q = np.asarray(principal_mT, dtype=float)
    q = q - q.mean()
    values, axes = np.linalg.eigh(hyperfine_mT)
    normal = int(np.argmax(np.abs(values)))
    in_plane = [i for i in range(3) if i != normal]
    diag = np.zeros(3)
    if orientation == "ring_normal":
        diag[normal], diag[in_plane[0]], diag[in_plane[1]] = q[2], q[0], q[1]
    elif orientation == "ring_plane_provisional":
        diag[in_plane[0]], diag[in_plane[1]], diag[normal] = q[2], q[0], q[1]
    else:
        raise ValueError(f"Unknown quadrupole orientation: {orientation}")
    t = axes @ np.diag(diag) @ axes.T
    return 0.5 * (t + t.T)


_BY_KEY = {f"{r}:{n.name}": n for r, n in zip(NUCLEAR_RADICALS, ALL_NUCLEI)}

# The tensors as published with the reference (methods/quadrupole_tensors.json "tensor_mT", the values
# its archived runs used).  quadrupole_tensor_in_hyperfine_frame reproduces them to the last bit or
# two; the eigenvectors of the local LAPACK decide that bit, so the published values are used as
# data and the construction is checked against them.
QUADRUPOLE_TENSORS_MT = {
    "FAD:N5": np.array([[-0.062249980262593874, 0.030152763522984453, 0.0],
                        [0.030152763522984453, 0.02124998026259388, 0.0],
                        [0.0, 0.0, 0.040999999999999995]]),
    "FAD:N10": np.array([[0.03644736895436497, -0.0044912366031729985, 0.0],
                         [-0.0044912366031729985, 0.03588596437896834, 0.0],
                         [0.0, 0.0, -0.07233333333333333]]),
    "TrpH:N1": np.array([[0.01706521016250016, -0.0061896630223540324, 0.005136399168541131],
                         [-0.0061896630223540324, -0.00025698660522767674, 0.02498806846622416],
                         [0.005136399168541131, 0.02498806846622416, -0.016808223557272517]]),
}


def spin_operators(spin: float):
    """(Sx, Sy, Sz) in the descending-m basis (R/rpm_quantum_needle_model.py:178-192)."""
    d = int(round(2.0 * spin + 1.0))
    m = spin - np.arange(d)
    raise_ = np.zeros((d, d), dtype=complex)
    idx = np.arange(d - 1)
    raise_[idx, idx + 1] = np.sqrt(spin * (spin + 1.0) - m[idx + 1] * (m[idx + 1] + 1.0))
    lower = raise_.conj().T
    return 0.5 * (raise_ + lower), (raise_ - lower) / 2.0j, np.diag(m.astype(complex))


def field_vector_mT(theta_deg: float, field_microT: float = DEFAULT_FIELD_MICROT, phi_deg: float = 0.0):
    """B in mT: [sin t cos p, sin t sin p, cos t] * |B| (zx plane at phi = 0,
    R/rpm_quantum_needle_model.py:195-199)."""
    theta = np.deg2rad(theta_deg)
    if phi_deg == 0.0:
        return 1e-3 * field_microT * np.array([np.sin(theta), 0.0, np.cos(theta)])
    phi = np.deg2rad(phi_deg)
    return 1e-3 * field_microT * np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi),
                                           np.cos(theta)])


def theta_grid(step_deg: float = 1.0, fine_step_deg: float = 0.25, fine_half_width_deg: float = 3.0) -> np.ndarray:
    """Field angles (deg): 0..180 in step_deg, plus fine_step_deg within +-fine_half_width_deg of 90
    (R/rpm_quantum_needle_model.py:202-213; 199 angles by default)."""
    coarse = np.arange(0.0, 180.0 + 0.5 * step_deg, step_deg)
    fine = np.arange(90.0 - fine_half_width_deg, 90.0 + fine_half_width_deg + 0.5 * fine_step_deg, fine_step_deg)
    return np.unique(np.round(np.concatenate([coarse, fine]), 10))


def diagnostic_time_grid_us() -> np.ndarray:
    """0..5 us every 0.01 us, then 257 log-spaced points to 500 us (R/rpm_quantum_needle_model.py:216-219)."""
    dense = np.arange(0.0, 5.0 + 0.005, 0.01)
    logarithmic = np.geomspace(5.0, 500.0, 257)
    return np.unique(np.concatenate([dense, logarithmic]))


def nominal_quadrupole_metadata(scale: float) -> dict:
    """Configuration-record entries for active Q tensors; empty at scale 0 so that Q-off records keep
    their fingerprints (R/rpm_quantum_needle_model.py:153-161)."""
    scale = float(scale)
    if not math.isfinite(scale) or scale < 0.0:
        raise ValueError("quadrupole scale must be finite and non-negative")
    if scale == 0:
        return {}
    return {
        "nominal_quadrupole_tensors_mT": {k: q.tolist() for k, q in QUADRUPOLE_TENSORS_MT.items()},
        "quadrupole_frame": "shared flavin molecular frame; N5 in-plane azimuth provisional",
    }
