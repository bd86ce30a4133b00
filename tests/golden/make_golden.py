"""Write golden vectors from the reference solver (run only where /root/reference exists).

The reference is imported read-only from its own tree to record input/output
vectors that no closed form or known-answer test pins:

* the adaptive Lanczos step schedule (accepted-step end times, cumulative
  rejections and matvec counts, reported through the reference's own
  progress_callback), the singlet populations, state-error estimates, final
  state and statistics of ``AdaptiveLanczosPropagator.propagate`` on a reduced
  6-nucleus pair (dim 864) and on the 3-nitrogen smoke subsystem (dim 108);
* ``CompiledFullPairHamiltonian.matvec`` on one seeded full-size vector
  (dim 221,184): norms, <v|Hv> and 512 sampled entries.

Output: tests/golden/reference_vectors.npz (small).  Usage:
    python tests/golden/make_golden.py
"""

from __future__ import annotations

import os
import sys

import numpy as np

REF = "/root/reference/Realistic system size/code"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.npz")


def main() -> None:
    sys.path.insert(0, REF)
    import matplotlib
    matplotlib.use("Agg")
    from poster_reproduction import core as rcore  # noqa: E402
    from rpm_quantum_needle_model import FAD_NUCLEI, TRP_NUCLEI  # noqa: E402
    from rpm_statevector_engine import SingletPopulationObservable  # noqa: E402

    data = {}
    cases = {
        "six": ((FAD_NUCLEI[0], FAD_NUCLEI[1], FAD_NUCLEI[2], TRP_NUCLEI[0], TRP_NUCLEI[1], TRP_NUCLEI[2]),
                ("FAD", "FAD", "FAD", "TrpH", "TrpH", "TrpH"), 37.0, 0.5),
        "smoke": ((FAD_NUCLEI[0], FAD_NUCLEI[1], TRP_NUCLEI[0]), ("FAD", "FAD", "TrpH"), 90.0, 0.5),
    }
    for name, (nuclei, radicals, theta, horizon) in cases.items():
        h = rcore.build_hamiltonian(theta, 0.05, nuclei=nuclei, radicals=radicals)
        dims = rcore.dimensions_for(nuclei)
        state = rcore.random_phase_initial_state(dims, rcore.MASTER_SEED, 0)
        times = np.arange(int(round(horizon / 0.00125)) + 1) * 0.00125
        events = []
        prop = rcore.make_propagator(h)
        res = prop.propagate(state, times, SingletPopulationObservable(dims),
                             progress_callback=lambda t, a, r, m, e: events.append((t, a, r, m, e)))
        data[f"{name}_theta"] = np.array(theta)
        data[f"{name}_times"] = times
        data[f"{name}_values"] = res.observable_values
        data[f"{name}_state_errors"] = res.state_error_estimates
        data[f"{name}_final_state"] = res.final_state
        data[f"{name}_events"] = np.array(events, dtype=float)
        s = res.stats
        data[f"{name}_stats"] = np.array([s.accepted_steps, s.rejected_steps, s.matvecs, s.maximum_krylov_dimension,
                                          s.minimum_step_us, s.maximum_step_us, s.accumulated_state_error,
                                          s.maximum_norm_error], dtype=float)
        print(name, s, flush=True)

    h = rcore.build_hamiltonian(37.0, 0.05)
    rng = np.random.default_rng(20260731)
    v = rng.normal(size=h.dimension) + 1j * rng.normal(size=h.dimension)
    hv = h.matvec(v)
    idx = np.linspace(0, h.dimension - 1, 512).astype(np.int64)
    data["full_matvec_index"] = idx
    data["full_matvec_samples"] = hv[idx]
    data["full_matvec_norm"] = np.array(np.linalg.norm(hv))
    data["full_matvec_vhv"] = np.array(np.vdot(v, hv))
    np.savez_compressed(OUT, **data)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
