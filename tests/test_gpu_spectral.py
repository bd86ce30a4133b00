"""GPU: factorized spectral yields (rpm_amd.spectral) against the oracle and the archived curve.

Bars:
- the HIP line-binning kernel, fed the oracle's eigensystem, reproduces the oracle's grids bin by
  bin to 1e-12 of the bin's absolute weight (FP64 atomics sum each bin in another order);
- the 199-angle Fig. 1B curve (Q = 0, 50 uT, six lifetimes) is within 1e-10 of the archived one,
  with P_S(0) = 1 to 1e-12;
- the toy pair's FFT-route yields agree with the oracle's to 1e-12.
"""
import csv
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from oracle import model as om  # noqa: E402
from oracle import spectral as osp  # noqa: E402
from rpm_amd import model as pm  # noqa: E402
from rpm_amd import spectral as ps  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "archived_fig1b_q0.csv")
TAUS = (1.0, 2.0, 5.0, 10.0, 20.0, 100.0)


def _device_spectra(spec):
    ev = torch.from_numpy(np.ascontiguousarray(spec.eigenvalues[None])).cuda()
    ops = torch.from_numpy(np.ascontiguousarray(np.stack(spec.electron_ops_eigenbasis)[None])).cuda()
    return ps.Spectra(ev, ops, spec.nuclear_dimension)


def test_binning_kernel_matches_oracle_grids():
    """k_bin_lines on the oracle's own eigen-decomposition: every (p, q) grid to rounding."""
    print("case: bin FAD theta=37", flush=True)
    fad = osp.build_radical("FAD", om.FAD_NUCLEI)
    spec = osp.diagonalize_radical(fad, 37.0, 50.0)
    delta = 0.00125
    half = osp.half_width_bins(spec, delta)
    got = ps.binned_spectra(_device_spectra(spec), delta, half).cpu().numpy()[0]
    ops = spec.electron_ops_eigenbasis
    for k, (p, q) in enumerate(ps.PAIRS):
        freq, coeff = osp.transition_lines(spec, p, q)
        weight = osp.binned_spectrum(freq, np.abs(coeff), delta, half).real
        # the kernel's weight T_p[i][j] conj(T_q[i][j]) / M (T_q Hermitian): bin by bin to rounding
        herm = (ops[p] * ops[q].conj()).reshape(-1) / spec.nuclear_dimension
        want = osp.binned_spectrum(freq, herm, delta, half)
        excess = np.abs(got[k] - want) - 1e-12 * weight
        worst = int(np.argmax(excess))
        assert excess[worst] <= 1e-300, ((p, q), worst, got[k][worst], want[worst], weight[worst])
        # the reference's weight T_p[i][j] T_q[j][i] / M differs only where T is rounding noise
        want = osp.binned_spectrum(freq, coeff, delta, half)
        assert np.all(np.abs(got[k] - want) <= 1e-12 * weight + 1e-17), (p, q)
        assert abs(got[k].sum() - coeff.sum()) < 1e-13, (p, q)
    # a grid too narrow for the lines is reported, as the reference does
    with pytest.raises(ValueError):
        ps.binned_spectra(_device_spectra(spec), delta, half // 2)


@pytest.mark.parametrize("method", ["evd", "jacobi", "torch"])
def test_eigensystem_matches_oracle(method):
    print(f"case: diagonalize TrpH, eigh {method}", flush=True)
    trp_o = osp.build_radical("TrpH", om.TRP_NUCLEI)
    trp_p = ps.reference_models(0.0)[1]
    got = ps.diagonalize_radical(trp_p, [12.0, 90.0], 50.0, eigh=method)
    w = got.eigenvalues.cpu().numpy()
    ops = got.ops.cpu().numpy()
    for i, th in enumerate((12.0, 90.0)):
        spec = osp.diagonalize_radical(trp_o, th, 50.0)
        assert np.abs(w[i] - spec.eigenvalues).max() < 1e-11
        for c in range(3):
            t = ops[i, c]
            assert np.abs(t - t.conj().T).max() < 1e-15
            # basis-independent: spectrum of T_c is +-1/2, and T_c^2 = 1/4
            assert abs(np.trace(t @ t).real - t.shape[0] / 4) < 1e-10
            # diagonal of T_c in the eigenbasis (unique up to degeneracy) matches the oracle
            assert np.abs(np.diag(t).real - np.diag(spec.electron_ops_eigenbasis[c]).real).max() < 1e-8


def test_fig1b_curve_matches_archive():
    print("case: fig1b 199 angles", flush=True)
    with open(GOLDEN, newline="") as f:
        ref = [{k: float(v) for k, v in r.items()} for r in csv.DictReader(f)]
    thetas = pm.theta_grid()
    t0 = time.perf_counter()
    rows = ps.solve_orientations(thetas, 50.0, TAUS, 8.0, quadrupole_scale=0.0)
    print(f"199 angles: {time.perf_counter() - t0:.2f} s", flush=True)
    worst = 0.0
    for row, r in zip(rows, ref):
        assert row["theta_deg"] == r["theta_deg"]
        assert row["delta_omega_rad_per_us"] == r["delta_omega_rad_per_us"]
        assert abs(row["initial_singlet"] - 1.0) < 1e-12
        for tau in TAUS:
            key = f"phi_s_tau_{tau:g}_us"
            worst = max(worst, abs(row[key] - r[key]))
    print(f"worst |Phi_gpu - Phi_archive| = {worst:.3e}")
    assert worst < 1e-10


def test_toy_pair_fft_route():
    """The reference's toy self-test pair through the GPU route, against the oracle's FFT route."""
    print("case: toy pair", flush=True)
    taus = (1.0, 10.0, 100.0)
    na = pm.Nucleus("N", 1.0, np.diag([-0.0989, -0.0989, 1.7569]))
    nb = pm.Nucleus("N", 1.0, np.diag([0.0, 0.0, 1.0812]))
    models = (ps.build_radical("X", (na,), 0.0), ps.build_radical("Y", (nb,), 0.0))
    rows = ps.solve_orientations([89.0, 30.0], 50.0, taus, 16.0, models=models)
    oa = osp.build_radical("X", (om.Nucleus("N", 1.0, np.diag([-0.0989, -0.0989, 1.7569])),))
    ob = osp.build_radical("Y", (om.Nucleus("N", 1.0, np.diag([0.0, 0.0, 1.0812])),))
    for row, th in zip(rows, (89.0, 30.0)):
        sa = osp.diagonalize_radical(oa, th, 50.0)
        sb = osp.diagonalize_radical(ob, th, 50.0)
        omega, comb, initial = osp.convolved_correlation_spectrum(sa, sb, (1.0 / 100.0) / 16.0)
        want = osp.yields_from_spectrum(omega, comb, taus)
        exact = osp.direct_line_yields(sa, sb, taus)
        assert abs(row["initial_singlet"] - initial) < 1e-12
        for tau in taus:
            assert abs(row[f"phi_s_tau_{tau:g}_us"] - want[tau]) < 1e-12
            assert abs(row[f"phi_s_tau_{tau:g}_us"] - exact[tau]) < 2e-4


def test_many_lifetimes_and_phi():
    """More than 8 lifetimes (two yield passes) and an out-of-plane field, against the oracle."""
    print("case: 10 lifetimes, phi=30", flush=True)
    taus = (0.5, 1.0, 2.0, 3.0, 5.0, 7.0, 10.0, 20.0, 50.0, 100.0)
    na = pm.Nucleus("N", 1.0, np.diag([-0.0989, -0.0989, 1.7569]))
    nb = pm.Nucleus("N", 1.0, np.diag([0.0, 0.0, 1.0812]))
    models = (ps.build_radical("X", (na,), 0.0), ps.build_radical("Y", (nb,), 0.0))
    row = ps.solve_orientations([61.0], 50.0, taus, 8.0, phis_deg=30.0, models=models)[0]
    oa = osp.build_radical("X", (om.Nucleus("N", 1.0, np.diag([-0.0989, -0.0989, 1.7569])),))
    ob = osp.build_radical("Y", (om.Nucleus("N", 1.0, np.diag([0.0, 0.0, 1.0812])),))

    def diag(r):
        fv = om.field_vector_mT(61.0, 50.0, 30.0)
        h = r.h0 + sum(om.TWOPI * om.GAMMA_E_MHZ_PER_MT * c * op for c, op in zip(fv, r.electron_ops))
        w, v = np.linalg.eigh(h)
        return osp.Spectrum(w, tuple(v.conj().T @ op @ v for op in r.electron_ops), r.nuclear_dimension)

    omega, comb, initial = osp.convolved_correlation_spectrum(diag(oa), diag(ob), (1.0 / 100.0) / 8.0)
    want = osp.yields_from_spectrum(omega, comb, taus)
    assert abs(row["initial_singlet"] - initial) < 1e-12
    for tau in taus:
        assert abs(row[f"phi_s_tau_{tau:g}_us"] - want[tau]) < 1e-12, tau


@pytest.mark.parametrize("scale", [0.0, 1.0])
def test_multiplet_blocks_match_the_whole_radical(scale):
    """The block route (FAD as 288 + 2 x 144) against the whole 576-state radical, Q off and on."""
    print(f"case: symmetry blocks, Q scale {scale}", flush=True)
    thetas = [0.0, 37.0, 89.5, 90.0, 151.0]
    blocks = ps.solve_orientations(thetas, 50.0, TAUS, 8.0, quadrupole_scale=scale, phis_deg=20.0)
    whole = ps.solve_orientations(thetas, 50.0, TAUS, 8.0, quadrupole_scale=scale, phis_deg=20.0, symmetry=False)
    for a, b in zip(blocks, whole):
        assert a.keys() == b.keys()
        for k in a:
            assert abs(a[k] - b[k]) < 1e-12, (a["theta_deg"], k, a[k], b[k])


def test_eigh_from_two_host_threads_on_two_streams():
    """rpm_eigh_batched keeps one rocBLAS handle per stream and holds it for the whole solve, so host
    threads on different streams get the same eigensystems as serial calls."""
    import threading
    print("case: eigh from two threads, two streams", flush=True)
    rng = np.random.default_rng(3)
    mats = []
    for _ in range(2):
        a = rng.normal(size=(8, 96, 96)) + 1j * rng.normal(size=(8, 96, 96))
        mats.append(np.ascontiguousarray(a + np.conj(np.transpose(a, (0, 2, 1)))))
    ref = [ps.eigh_batched(torch.from_numpy(m.copy()).cuda())[0].cpu().numpy() for m in mats]
    out, errors = [None, None], []

    def work(i):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for _ in range(6):
                    w, vh = ps.eigh_batched(torch.from_numpy(mats[i].copy()).to("cuda", non_blocking=False))
                s.synchronize()
                out[i] = (w.cpu().numpy(), vh.cpu().numpy())
        except Exception as exc:   # reported after the join
            errors.append(exc)

    threads = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    for i in range(2):
        w, vh = out[i]
        np.testing.assert_allclose(w, ref[i], rtol=0, atol=1e-11)
        v = np.conj(np.transpose(vh, (0, 2, 1)))
        rec = v @ (w[:, :, None] * vh)
        assert np.abs(rec - mats[i]).max() < 1e-10
