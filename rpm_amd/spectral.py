"""Factorized spectral yields on the GPU (no electron-electron coupling: J = D = 0).

Drop-in for R/rpm_quantum_needle_reproduction.py:93-237. The reference computes one orientation at
a time on the CPU. Here a batch of orientations is processed per pass:
- The radical Hamiltonians are built once on the host (dense, 576 and 384 for FAD and TrpH). The
  Zeeman term of every orientation is added on the device.
- Batched ``eigh``: the whole batch in one rocSOLVER call (rpm_eigh_batched). The electron operators
  go to the eigenbasis with three batched products that use the structure S_c = sigma_c / 2 (x) 1.
- The HIP kernel ``rpm_bin_lines`` deposits every transition line on the frequency grid by linear
  interpolation (R/rpm_quantum_needle_reproduction.py:145-172). Only the six pairs p <= q are
  binned, because the (q, p) grid is the conjugate of the (p, q) grid.
- One batched FFT per radical. The yields then come from Parseval (the HIP kernel
  ``rpm_spectral_yields``, one pass over both transforms):
      sum_n c[n] K[n] = sum_f C(f) ifft(K)[f],
  with K[n] = k / (k - i omega_n) on the reference's output grid and 0 beyond it. The pairs
  p < q enter through x + conj(x), i.e. with the kernel 2 Re K. This replaces the reference's 9 FFT
  products plus an inverse FFT per orientation by 6 products and one small matrix product per batch.

Every orientation in a batch shares one grid: the half width (in bins) is the largest over the
batch. Zero padding leaves the convolution unchanged, so the result is the reference's up to
rounding.

Equivalent nuclei. When several spin-1/2 nuclei of a radical have the same hyperfine tensor (FAD's
methyl protons H8a, H8b and H8c, 0.4399 mT each), the Hamiltonian depends on them only through their
total spin. The radical's space then splits into multiplet blocks: for three protons, one block with
total spin 3/2 (FAD: 288 states) and two identical blocks with spin 1/2 (144 states each), with no
matrix element of H or S between blocks. Each distinct block is diagonalised and binned on its own,
its lines counted with the block's multiplicity. The lines, and with them the spectrum, are those of
the full radical; eigh runs on 288 + 144 states instead of 576 (about 7x fewer flops) and 3.2x fewer
lines are binned.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Iterable, Sequence

import numpy as np

from . import _native
from .model import (DEFAULT_LIFETIMES_US, FAD_NUCLEI, GAMMA_E_MHZ_PER_MT, NOMINAL_QUADRUPOLE_SCALE,
                    QUADRUPOLE_TENSORS_MT, TRP_NUCLEI, TWOPI, Nucleus, spin_operators, theta_grid)

MARGIN_BINS = 3          # R/rpm_quantum_needle_reproduction.py:180
PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))   # the grid order of rpm_bin_lines
DEFAULT_CHUNK = 16       # orientations per grid pass; about 1 GB of device memory each at the Fig. 1B grid

__all__ = ["RadicalModel", "Spectra", "build_radical", "reference_models", "diagonalize_radical",
           "symmetry_blocks", "eigh_batched", "binned_spectra", "correlation_yields", "solve_orientations", "solve_orientation",
           "theta_grid", "PAIRS"]


@dataclass
class RadicalModel:
    """One radical: electron plus its own nuclei (R/rpm_quantum_needle_reproduction.py:60-70).

    A multiplet block of a radical (symmetry_blocks) is a RadicalModel too: its h0 is the block, its
    nuclei list holds the equivalent group as one nucleus of the block's total spin, nuclear_dimension
    stays the whole radical's M (the 1/M of the random-phase ensemble), and multiplicity counts the
    identical copies of the block.
    """

    name: str
    nuclei: tuple
    h0: np.ndarray               # dense hyperfine (+ quadrupole) Hamiltonian, rad/us
    nuclear_dimension: int
    multiplicity: int = 1
    _blocks: tuple | None = field(default=None, repr=False, compare=False)

    @property
    def dimension(self) -> int:
        return self.h0.shape[0]


def _embed(op: np.ndarray, position: int, dims: Sequence[int]) -> np.ndarray:
    left = int(np.prod(dims[:position]))
    right = int(np.prod(dims[position + 1:]))
    return np.kron(np.kron(np.eye(left), op), np.eye(right))


def build_radical(name: str, nuclei: Iterable[Nucleus], quadrupole_scale: float = NOMINAL_QUADRUPOLE_SCALE,
                  quadrupole_tensors=None) -> RadicalModel:
    """h0 = sum_n 2 pi gamma S.A_n.I_n + scale * 2 pi gamma I_n.Q_n.I_n.

    Basis: electron (m = +1/2, -1/2) slowest, then the nuclei in order, each with m descending.
    Same terms and order as R/rpm_quantum_needle_reproduction.py:93-124.
    """
    quadrupole_scale = float(quadrupole_scale)
    if not math.isfinite(quadrupole_scale) or quadrupole_scale < 0.0:
        raise ValueError("quadrupole scale must be finite and non-negative")
    tensors = QUADRUPOLE_TENSORS_MT if quadrupole_tensors is None else quadrupole_tensors
    nuclei = tuple(nuclei)
    dims = (2,) + tuple(n.dimension for n in nuclei)
    dim = math.prod(dims)
    s = tuple(_embed(o, 0, dims) for o in spin_operators(0.5))
    h0 = np.zeros((dim, dim), dtype=complex)
    for index, nucleus in enumerate(nuclei, start=1):
        i_ops = tuple(_embed(o, index, dims) for o in spin_operators(nucleus.spin))
        for a in range(3):
            for b in range(3):
                c = nucleus.tensor_mT[a, b]
                if abs(c) > 1e-15:
                    h0 += TWOPI * GAMMA_E_MHZ_PER_MT * c * (s[a] @ i_ops[b])
        q = tensors.get(f"{name}:{nucleus.name}")
        if q is not None and quadrupole_scale:
            for a in range(3):
                for b in range(3):
                    h0 += TWOPI * GAMMA_E_MHZ_PER_MT * quadrupole_scale * q[a, b] * (i_ops[a] @ i_ops[b])
    return RadicalModel(name, nuclei, h0, math.prod(dims[1:]))


def _equivalent_groups(nuclei) -> list:
    """Index tuples (size >= 2) of spin-1/2 nuclei with identical hyperfine tensors."""
    groups: dict = {}
    for i, n in enumerate(nuclei):
        if n.spin == 0.5:
            groups.setdefault(tuple(np.asarray(n.tensor_mT, dtype=float).ravel()), []).append(i)
    return [tuple(g) for g in groups.values() if len(g) >= 2]


def _multiplets(n: int) -> list:
    """(total spin I, multiplicity) of n coupled spin-1/2: I = n/2 - k with multiplicity
    C(n, k) - C(n, k - 1), k = 0 .. n // 2."""
    return [(n / 2 - k, math.comb(n, k) - (math.comb(n, k - 1) if k else 0)) for k in range(n // 2 + 1)]


def _multiplet_basis(n: int, spin: float) -> np.ndarray:
    """(2^n, 2I + 1): one irreducible multiplet |I, m>, m = I .. -I, of n spin-1/2.

    The top state is a J^2 = I(I+1) eigenvector inside Jz = I (Jz is diagonal in the product basis);
    the rest follow by J-, so the block's spin matrices are spin_operators(I).
    """
    dims = (2,) * n
    jx, jy, jz = (sum(_embed(o, k, dims) for k in range(n)) for o in spin_operators(0.5))
    top_space = np.flatnonzero(np.abs(np.diag(jz).real - spin) < 1e-9)
    j2 = (jx @ jx + jy @ jy + jz @ jz)[np.ix_(top_space, top_space)]
    w, v = np.linalg.eigh(j2)
    top = np.zeros(2 ** n, dtype=complex)
    top[top_space] = v[:, int(np.flatnonzero(np.abs(w - spin * (spin + 1.0)) < 1e-9)[0])]
    cols = [top]
    lower = jx - 1j * jy
    m = spin
    for _ in range(int(round(2 * spin))):
        cols.append(lower @ cols[-1] / math.sqrt(spin * (spin + 1.0) - m * (m - 1.0)))
        m -= 1.0
    return np.stack(cols, axis=1)


def _group_isometry(dims: Sequence[int], group: tuple, basis: np.ndarray):
    """Isometry from the reduced nuclear space (the group replaced, at its first position, by the
    multiplet spanned by `basis`) into the product space: (prod dims, prod reduced dims)."""
    keep = [i for i in range(len(dims)) if i == group[0] or i not in group]
    reduced = [basis.shape[1] if i == group[0] else dims[i] for i in keep]
    gdims = [dims[i] for i in group]
    u = np.zeros((math.prod(dims), math.prod(reduced)), dtype=complex)
    for a in np.ndindex(*dims):
        row = np.ravel_multi_index(a, dims)
        g = np.ravel_multi_index([a[i] for i in group], gdims)
        for m in range(basis.shape[1]):
            col = np.ravel_multi_index([m if i == group[0] else a[i] for i in keep], reduced)
            u[row, col] = basis[g, m]
    return u


def _decompose(model: RadicalModel) -> tuple:
    groups = _equivalent_groups(model.nuclei)
    if not groups:
        return (model,)
    group = groups[0]
    dims = [n.dimension for n in model.nuclei]
    scale = max(1.0, float(np.abs(model.h0).max()))
    blocks = []
    for spin, count in _multiplets(len(group)):
        u = np.kron(np.eye(2), _group_isometry(dims, group, _multiplet_basis(len(group), spin)))
        h = u.conj().T @ model.h0 @ u
        if np.abs(model.h0 @ u - u @ h).max() > 1e-12 * scale:
            return (model,)      # a term of h0 tells the group's nuclei apart: no block structure
        first = model.nuclei[group[0]]
        merged = Nucleus("+".join(model.nuclei[i].name for i in group), spin, first.tensor_mT)
        nuclei = tuple(merged if i == group[0] else n for i, n in enumerate(model.nuclei)
                       if i == group[0] or i not in group)
        blocks.extend(_decompose(RadicalModel(model.name, nuclei, 0.5 * (h + h.conj().T), model.nuclear_dimension,
                                              model.multiplicity * count)))
    return tuple(blocks)


def symmetry_blocks(model: RadicalModel) -> tuple:
    """The distinct multiplet blocks of a radical with groups of equivalent spin-1/2 nuclei (identical
    hyperfine tensors), each a RadicalModel with its multiplicity; (model,) when there are none.

    The blocks' transition lines, each counted multiplicity times, are the full radical's lines
    (R/rpm_quantum_needle_reproduction.py:145-152): H and S_c have no element between blocks. A group
    that h0 does not treat symmetrically (checked on the projected matrices) is left whole.
    """
    if model._blocks is None:
        model._blocks = _decompose(model)
    return model._blocks


_MODELS: dict = {}


def reference_models(quadrupole_scale: float = NOMINAL_QUADRUPOLE_SCALE):
    """(FAD, TrpH) radical models (R/rpm_quantum_needle_reproduction.py:465-473)."""
    key = float(quadrupole_scale)
    if key not in _MODELS:
        _MODELS[key] = (build_radical("FAD", FAD_NUCLEI, key), build_radical("TrpH", TRP_NUCLEI, key))
    return _MODELS[key]


@dataclass
class Spectra:
    """Eigen-decompositions of one radical (or one multiplet block of it) at a batch of orientations
    (device tensors).

    eigenvalues: (B, d) float64, ascending. ops: (B, 3, d, d) complex128, V^H S_c V. Each line weighs
    multiplicity / nuclear_dimension.
    """

    eigenvalues: object
    ops: object
    nuclear_dimension: int
    multiplicity: int = 1

    @property
    def max_gap(self):
        """Per orientation, lambda_max - lambda_min (host array)."""
        w = self.eigenvalues
        return (w[:, -1] - w[:, 0]).cpu().numpy()


def _field_components_mT(thetas_deg, field_microT, phis_deg):
    theta = np.deg2rad(np.asarray(thetas_deg, dtype=float))
    phi = np.deg2rad(np.broadcast_to(np.asarray(phis_deg, dtype=float), theta.shape))
    b = 1e-3 * float(field_microT)
    # R/rpm_quantum_needle_model.py:195-199 at phi = 0: (sin t, 0, cos t)
    return b * np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], axis=1)


EIGH_METHODS = {"evd": 0, "jacobi": 1, "torch": None}


def eigh_batched(h, method: str = "evd"):
    """Eigen-decomposition of a batch of Hermitian matrices h (B, d, d), overwritten.

    Returns (w (B, d) ascending, vh (B, d, d) = V^H, i.e. row k is the conjugated eigenvector k).
    "evd" and "jacobi" hand the whole batch to rocSOLVER in one call (rpm_eigh_batched); "torch" is
    torch.linalg.eigh.
    """
    torch = _native.require_gpu()
    if method not in EIGH_METHODS:
        raise ValueError(f"eigh method must be one of {sorted(EIGH_METHODS)}")
    nb, d, _ = h.shape
    if method == "torch":
        w, v = torch.linalg.eigh(h)
        return w, v.transpose(1, 2).conj().resolve_conj().contiguous()
    if not h.is_contiguous() or h.dtype != torch.complex128:
        raise ValueError("h: contiguous complex128 (B, d, d)")
    lib = _native.load()
    w = torch.empty((nb, d), dtype=torch.float64, device=h.device)
    nbytes = lib.rpm_eigh_workspace_bytes(d, nb)
    ws = _native.workspace(nbytes)
    _native.check(lib.rpm_eigh_batched(_native.ptr(h), d, nb, _native.ptr(w), _native.ptr(ws), nbytes,
                                       EIGH_METHODS[method], _native.stream_ptr()))
    info = ws[:4 * nb].view(torch.int32)
    if bool((info != 0).any()):
        raise RuntimeError(f"eigh did not converge for {int((info != 0).sum())} of {nb} matrices")
    return w, h


def diagonalize_radical(model: RadicalModel, thetas_deg, field_microT: float, phis_deg=0.0,
                        eigh: str = "evd") -> Spectra:
    """Batched R/rpm_quantum_needle_reproduction.py:127-142: add 2 pi gamma B.S, eigh, and transform
    the electron operators into each eigenbasis."""
    torch = _native.require_gpu()
    dev = _native.device()
    thetas = np.atleast_1d(np.asarray(thetas_deg, dtype=float))
    b = torch.from_numpy(_field_components_mT(thetas, field_microT, phis_deg) * (TWOPI * GAMMA_E_MHZ_PER_MT)).to(dev)
    d = model.dimension
    m = d // 2
    h0 = torch.from_numpy(np.ascontiguousarray(model.h0)).to(dev)
    h = h0.unsqueeze(0).repeat(len(thetas), 1, 1)
    # 2 pi gamma B.S with S_c = sigma_c / 2 on the electron (the slowest index: rows [0, m) are m_S = +1/2).
    bx, by, bz = (b[:, c].to(torch.complex128).view(-1, 1) for c in range(3))
    diag = torch.arange(m, device=dev)
    h[:, diag, diag] += 0.5 * bz
    h[:, diag + m, diag + m] -= 0.5 * bz
    h[:, diag, diag + m] += 0.5 * (bx - 1j * by)
    h[:, diag + m, diag] += 0.5 * (bx + 1j * by)
    w, vh = eigh_batched(h, eigh)
    # V = [V_t; V_b] (electron up / down rows); V^H S_c V from V_t^H = vh[:, :, :m], V_b^H = vh[:, :, m:]
    vth = vh[:, :, :m].contiguous()
    vbh = vh[:, :, m:].contiguous()
    del vh, h
    vt = vth.transpose(1, 2).conj().resolve_conj().contiguous()
    vb = vbh.transpose(1, 2).conj().resolve_conj().contiguous()
    x = torch.matmul(vth, vb)                       # V_t^H V_b
    xh = x.transpose(1, 2).conj().resolve_conj()
    ops = torch.empty((len(thetas), 3, d, d), dtype=torch.complex128, device=dev)
    ops[:, 0] = 0.5 * (x + xh)                      # V^H S_x V
    ops[:, 1] = 0.5j * (xh - x)                     # V^H S_y V
    ops[:, 2] = 0.5 * (torch.matmul(vth, vt) - torch.matmul(vbh, vb))   # V^H S_z V
    return Spectra(w.contiguous(), ops, model.nuclear_dimension, model.multiplicity)


def half_width_bins(max_gap: float, delta_omega: float) -> int:
    """R/rpm_quantum_needle_reproduction.py:180-182."""
    return int(np.ceil(max_gap / delta_omega)) + MARGIN_BINS


def binned_spectra(spectra: Spectra, delta_omega: float, half: int, stride: int | None = None, out=None,
                   overflow=None, accumulate: bool = False):
    """Transition-line grids of the six pairs p <= q: (B, 6, stride) complex128, bin `half` = 0 rad/us.

    The HIP kernel rpm_bin_lines (R/rpm_quantum_needle_reproduction.py:145-172). Every element of the
    result is written, bins 2 half + 1 .. stride - 1 with zeros; with `accumulate` the lines are added
    to `out` instead. Raises ValueError when a line falls outside the 2 half + 1 bins, as the
    reference does (with `overflow` given, the flag is only accumulated there and the caller checks it).
    """
    torch = _native.require_gpu()
    lib = _native.load()
    w, ops = spectra.eigenvalues, spectra.ops
    nb, d = w.shape
    if ops.shape != (nb, 3, d, d) or not ops.is_contiguous() or not w.is_contiguous():
        raise ValueError("spectra: eigenvalues (B, d) and ops (B, 3, d, d), contiguous")
    stride = 2 * half + 1 if stride is None else int(stride)
    if out is None:
        if accumulate:
            raise ValueError("accumulate needs the grid to add to (out)")
        out = torch.empty((nb, len(PAIRS), stride), dtype=torch.complex128, device=w.device)
    elif out.shape != (nb, len(PAIRS), stride) or not out.is_contiguous():
        raise ValueError("out: (B, 6, stride) contiguous")
    own = overflow is None
    if own:
        overflow = torch.zeros(1, dtype=torch.int32, device=w.device)
    divisor = spectra.nuclear_dimension / spectra.multiplicity
    _native.check(lib.rpm_bin_lines(_native.ptr(w), _native.ptr(ops), d, nb, float(divisor), float(delta_omega),
                                    int(half), _native.ptr(out), stride, int(bool(accumulate)),
                                    _native.ptr(overflow), _native.stream_ptr()))
    if own and int(overflow.item()) != 0:
        raise ValueError("Frequency grid is too small for transition spectrum.")
    return out


def _fast_length(n: int) -> int:
    """Smallest 2^a 3^b 5^c 7^d >= n (lengths rocFFT factors into its fastest radices)."""
    best = 1 << (int(n) - 1).bit_length()
    p7 = 1
    while p7 < 2 * n:
        p5 = p7
        while p5 < 2 * n:
            p3 = p5
            while p3 < 2 * n:
                m = p3
                while m < n:
                    m *= 2
                best = min(best, m)
                p3 *= 3
            p5 *= 5
        p7 *= 7
    return best


def _kernel_tables(torch, dev, lifetimes, out_len, origin, delta, nfft):
    """ifft of K_tau on the output grid (zero beyond it), and of 2 Re K_tau: (L, nfft) each."""
    omega = (torch.arange(nfft, dtype=torch.float64, device=dev) - origin) * delta
    live = torch.arange(nfft, device=dev) < out_len
    ks = torch.tensor([1.0 / float(t) for t in lifetimes], dtype=torch.float64, device=dev).view(-1, 1)
    kern = ks / (ks - 1j * omega.view(1, -1))                    # k / (k - i omega)
    kern = torch.where(live.view(1, -1), kern, torch.zeros((), dtype=kern.dtype, device=dev))
    return torch.fft.ifft(kern, dim=-1), torch.fft.ifft((2.0 * kern.real).to(torch.complex128), dim=-1)


@dataclass
class _Grid:
    """The shared frequency grid of a set of orientations and the yield kernels on it."""

    delta: float
    ha: int
    hb: int
    out_len: int
    nfft: int
    ld: object        # (L, nfft): ifft of K_tau on the output grid
    lo: object        # (L, nfft): ifft of 2 Re K_tau


def _grid(torch, dev, gap_a: float, gap_b: float, delta: float, lifetimes) -> _Grid:
    ha, hb = half_width_bins(gap_a, delta), half_width_bins(gap_b, delta)
    out_len = (2 * ha + 1) + (2 * hb + 1) - 1
    nfft = _fast_length(out_len)
    ld, lo = _kernel_tables(torch, dev, lifetimes, out_len, ha + hb, delta, nfft)
    return _Grid(delta, ha, hb, out_len, nfft, ld, lo)


YIELD_BLOCKS = 1024   # frequency ranges of rpm_spectral_yields (x the orientation groups of 8)


def _as_blocks(spectra) -> tuple:
    return (spectra,) if isinstance(spectra, Spectra) else tuple(spectra)


def _max_gap(blocks) -> np.ndarray:
    """Per orientation, the whole radical's lambda_max - lambda_min over its blocks (host array)."""
    torch = _native.require_gpu()
    hi = torch.stack([b.eigenvalues[:, -1] for b in blocks]).amax(dim=0)
    lo = torch.stack([b.eigenvalues[:, 0] for b in blocks]).amin(dim=0)
    return (hi - lo).cpu().numpy()


def _binned_blocks(blocks, delta: float, half: int, stride: int, overflow):
    """One radical's grids, its blocks' lines deposited into one (B, 6, stride) array."""
    out = binned_spectra(blocks[0], delta, half, stride=stride, overflow=overflow)
    for b in blocks[1:]:
        binned_spectra(b, delta, half, stride=stride, out=out, overflow=overflow, accumulate=True)
    return out


def _yields_on_grid(torch, spec_a, spec_b, g: _Grid):
    lib = _native.load()
    blocks_a, blocks_b = _as_blocks(spec_a), _as_blocks(spec_b)
    nb = blocks_a[0].eigenvalues.shape[0]
    overflow = torch.zeros(1, dtype=torch.int32, device=blocks_a[0].eigenvalues.device)
    fa = torch.fft.fft(_binned_blocks(blocks_a, g.delta, g.ha, g.nfft, overflow), dim=-1)
    fb = torch.fft.fft(_binned_blocks(blocks_b, g.delta, g.hb, g.nfft, overflow), dim=-1)
    sums = []
    for t0 in range(0, g.ld.shape[0], 8):                      # the kernel takes up to 8 lifetimes per pass
        ld, lo = g.ld[t0:t0 + 8], g.lo[t0:t0 + 8]
        partial = torch.empty((YIELD_BLOCKS, nb, ld.shape[0]), dtype=torch.float64, device=fa.device)
        _native.check(lib.rpm_spectral_yields(_native.ptr(fa), _native.ptr(fb), _native.ptr(ld), _native.ptr(lo), nb,
                                              g.nfft, ld.shape[0], YIELD_BLOCKS, _native.ptr(partial),
                                              _native.stream_ptr()))
        sums.append(partial.sum(dim=0))
    yields = 0.25 + torch.cat(sums, dim=1)
    dc = fa[:, :, 0] * fb[:, :, 0]                             # sum_n c[n] = C(0)
    initial = 0.25 + dc[:, :3].sum(dim=1).real + 2.0 * dc[:, 3:].sum(dim=1).real
    del fa, fb
    if int(overflow.item()) != 0:
        raise ValueError("Frequency grid is too small for transition spectrum.")
    return initial.cpu().numpy(), yields.cpu().numpy()


def correlation_yields(spec_a, spec_b, delta_omega: float, lifetimes_us: Sequence[float]):
    """Batched convolved_correlation_spectrum + yields_from_spectrum
    (R/rpm_quantum_needle_reproduction.py:175-214).

    spec_a, spec_b: one radical's Spectra, or the Spectra of its multiplet blocks. Returns
    (initial_singlet (B,), yields (B, L)) as host float64 arrays.
    """
    torch = _native.require_gpu()
    blocks_a, blocks_b = _as_blocks(spec_a), _as_blocks(spec_b)
    g = _grid(torch, blocks_a[0].eigenvalues.device, float(_max_gap(blocks_a).max()),
              float(_max_gap(blocks_b).max()), delta_omega, lifetimes_us)
    return _yields_on_grid(torch, spec_a, spec_b, g)


SUPER_CHUNK = 512   # orientations diagonalised at once and sharing one grid (about 12 MB each for FAD + TrpH)


def _rows(spec: Spectra, part: slice) -> Spectra:
    """The orientations `part` of a batch of spectra (contiguous views)."""
    return Spectra(spec.eigenvalues[part], spec.ops[part], spec.nuclear_dimension, spec.multiplicity)


def solve_orientations(thetas_deg, field_microT: float = 50.0,
                       lifetimes_us: Sequence[float] = DEFAULT_LIFETIMES_US, bins_per_linewidth: float = 8.0,
                       quadrupole_scale: float = NOMINAL_QUADRUPOLE_SCALE, phis_deg=0.0,
                       chunk: int = DEFAULT_CHUNK, models=None, eigh: str = "evd",
                       symmetry: bool = True) -> list:
    """Rows of R/rpm_quantum_needle_reproduction.py:217-237 (solve_orientation) for many angles.

    Each row holds theta_deg, initial_singlet, delta_omega_rad_per_us and phi_s_tau_<tau>_us, with the
    reference's keys, in the order of `thetas_deg`. All orientations of a call (up to SUPER_CHUNK of
    them) share one frequency grid, so the FFT length, and with it the FFT plan, is the same for every
    chunk. With `symmetry`, radicals with equivalent nuclei are solved block by block (symmetry_blocks).
    """
    if bins_per_linewidth < 2.0:
        raise ValueError("--bins-per-linewidth must be at least 2.")
    lifetimes = tuple(float(t) for t in lifetimes_us)
    if not lifetimes or any(t <= 0.0 for t in lifetimes):
        raise ValueError("Lifetimes must be positive.")
    torch = _native.require_gpu()
    fad, trp = reference_models(quadrupole_scale) if models is None else models
    parts_a = symmetry_blocks(fad) if symmetry else (fad,)
    parts_b = symmetry_blocks(trp) if symmetry else (trp,)
    thetas = np.atleast_1d(np.asarray(thetas_deg, dtype=float))
    phis = np.broadcast_to(np.asarray(phis_deg, dtype=float), thetas.shape)
    delta = (1.0 / max(lifetimes)) / bins_per_linewidth
    step = max(1, int(chunk))
    rows = []
    for s0 in range(0, len(thetas), SUPER_CHUNK):
        s1 = min(len(thetas), s0 + SUPER_CHUNK)
        # eigen-decompositions of the whole super-chunk at once (about 12 MB per FAD + TrpH orientation):
        # rocSOLVER's batched kernels are launch-bound at small batches
        sa = [diagonalize_radical(m, thetas[s0:s1], field_microT, phis[s0:s1], eigh=eigh) for m in parts_a]
        sb = [diagonalize_radical(m, thetas[s0:s1], field_microT, phis[s0:s1], eigh=eigh) for m in parts_b]
        g = _grid(torch, sa[0].eigenvalues.device, float(_max_gap(sa).max()), float(_max_gap(sb).max()), delta,
                  lifetimes)
        # the frequency grids (about 0.4 GB per orientation at the Fig. 1B grid) are formed `chunk` at a time
        for c0 in range(0, s1 - s0, step):
            part = slice(c0, min(s1 - s0, c0 + step))
            initial, yields = _yields_on_grid(torch, [_rows(x, part) for x in sa], [_rows(x, part) for x in sb], g)
            for i, theta in enumerate(thetas[s0 + part.start:s0 + part.stop]):
                row = {"theta_deg": float(theta), "initial_singlet": float(initial[i]),
                       "delta_omega_rad_per_us": delta}
                for j, tau in enumerate(lifetimes):
                    row[f"phi_s_tau_{tau:g}_us"] = float(yields[i, j])
                rows.append(row)
        del sa, sb
    return rows


def solve_orientation(theta_deg: float, field_microT: float, lifetimes_us: Sequence[float],
                      bins_per_linewidth: float, quadrupole_scale: float = NOMINAL_QUADRUPOLE_SCALE) -> dict:
    """One orientation, the reference's signature (R/rpm_quantum_needle_reproduction.py:217-237)."""
    return solve_orientations([theta_deg], field_microT, lifetimes_us, bins_per_linewidth, quadrupole_scale)[0]
