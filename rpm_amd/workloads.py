"""Synthetic workloads: random-tensor spin systems and seeded vectors.

* ``small_system`` -- 2 electrons + 2 14N (spin 1, random traceless symmetric Q
  with Frobenius norm 1 MHz) + 4 1H, dimension 4*9*16 = 576; hyperfine tensors
  random symmetric with isotropic part U(0,1) mT and anisotropic part
  N(0, 0.3 mT); B = 50 uT at theta = 0; J = 0; Efimova-Hore D of a point dipole
  at r = 1.5 nm (D = -2783/r^3 uT nm^3); lifetime 1 us; horizon 5 us.
* ``full_random_system`` -- the 14 FAD-TrpH nuclei (species, spins, order of
  the production model) with tensors drawn as above, Q on the three 14N.
* ``production_system`` -- the reference's production operator.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import core
from .model import ALL_NUCLEI, NUCLEAR_RADICALS, Nucleus

POINT_DIPOLE_UT_NM3 = -2783.0


def _random_hyperfine(rng) -> np.ndarray:
    a = rng.normal(0.0, 0.3, size=(3, 3))
    aniso = 0.5 * (a + a.T)
    aniso -= np.trace(aniso) / 3.0 * np.eye(3)
    return aniso + rng.uniform(0.0, 1.0) * np.eye(3)


def _random_quadrupole_mT(rng, magnitude_MHz=1.0) -> np.ndarray:
    a = rng.normal(size=(3, 3))
    q = 0.5 * (a + a.T)
    q -= np.trace(q) / 3.0 * np.eye(3)
    q *= magnitude_MHz / np.linalg.norm(q)
    return q / core.GAMMA_E_MHZ_PER_MT


@dataclass(frozen=True)
class SpinSystem:
    nuclei: tuple
    radicals: tuple
    quadrupole_tensors: dict
    condition: core.EfimovaHoreCondition
    theta_deg: float
    field_mT: float
    lifetime_us: float
    horizon_us: float

    @property
    def dimensions(self):
        return (2, 2) + tuple(n.dimension for n in self.nuclei)

    def hamiltonian(self, theta_deg=None, field_mT=None, phi_deg=0.0, builder=None):
        build = core.build_hamiltonian if builder is None else builder
        return build(self.theta_deg if theta_deg is None else theta_deg,
                     self.field_mT if field_mT is None else field_mT,
                     nuclei=self.nuclei, radicals=self.radicals, condition=self.condition,
                     quadrupole_tensors=self.quadrupole_tensors, phi_deg=phi_deg)

    def sweep_config(self, **kw) -> core.SweepConfig:
        return core.SweepConfig(lifetime_us=self.lifetime_us,
                                survival_tail_bound=math.exp(-self.horizon_us / self.lifetime_us), **kw)


def small_system(seed: int = 0) -> SpinSystem:
    rng = np.random.default_rng(seed)
    nuclei = (Nucleus("N5", 1.0, _random_hyperfine(rng)), Nucleus("H6", 0.5, _random_hyperfine(rng)),
              Nucleus("H8a", 0.5, _random_hyperfine(rng)), Nucleus("N1", 1.0, _random_hyperfine(rng)),
              Nucleus("H1", 0.5, _random_hyperfine(rng)), Nucleus("H2", 0.5, _random_hyperfine(rng)))
    radicals = ("FAD", "FAD", "FAD", "TrpH", "TrpH", "TrpH")
    q = {"FAD:N5": _random_quadrupole_mT(rng), "TrpH:N1": _random_quadrupole_mT(rng)}
    d_uT = POINT_DIPOLE_UT_NM3 / 1.5 ** 3
    return SpinSystem(nuclei, radicals, q, core.EfimovaHoreCondition(d_uT, 0.0), 0.0, 0.05, 1.0, 5.0)


def full_random_system(seed: int = 1) -> SpinSystem:
    rng = np.random.default_rng(seed)
    nuclei = tuple(Nucleus(n.name, n.spin, _random_hyperfine(rng)) for n in ALL_NUCLEI)
    q = {f"{r}:{n.name}": _random_quadrupole_mT(rng) for n, r in zip(nuclei, NUCLEAR_RADICALS) if n.spin == 1.0}
    return SpinSystem(nuclei, NUCLEAR_RADICALS, q, core.PRODUCTION_MODEL.condition, 0.0, 0.05, 1.0, 5.0)


def production_system(theta_deg: float = 90.0) -> SpinSystem:
    from .model import QUADRUPOLE_TENSORS_MT
    return SpinSystem(ALL_NUCLEI, NUCLEAR_RADICALS, QUADRUPOLE_TENSORS_MT, core.PRODUCTION_MODEL.condition,
                      theta_deg, 0.05, 5.0, -5.0 * math.log(1e-5))


def normal_vectors(n: int, batch: int, seed: int = 0) -> np.ndarray:
    """i.i.d. complex normal vectors from default_rng(seed), each normalised: (batch, n)."""
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(batch, n)) + 1j * rng.normal(size=(batch, n))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v


def fibonacci_sphere(n: int):
    """n field directions spread evenly over the sphere (golden-angle spiral): (theta_deg, phi_deg).

    Point i sits at z = 1 - (2i + 1) / n and azimuth i times the golden angle, so the points cover equal
    areas of the sphere. These are the orientations of the directional-map workload (64 of them).
    """
    if n < 1:
        raise ValueError("need at least one direction")
    i = np.arange(n, dtype=float) + 0.5
    z = 1.0 - 2.0 * i / n
    theta = np.degrees(np.arccos(z))
    phi = np.degrees(np.mod(np.pi * (1.0 + math.sqrt(5.0)) * i, 2.0 * np.pi))
    return theta, phi
