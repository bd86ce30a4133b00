"""Record one full-system Sobol cell of the oracle as a fixture (sobol_full_cell.npz).

The cell: theta = 90 deg, 50 uT, nominal 14N quadrupoles, J = D = 0 (the estimator's default
Hamiltonian, R/rpm_quantum_needle_statevector.py:193-199), sample 0, tau = 1 us only, 2 scrambles,
Q = 4, with the estimator's default Lanczos settings. The oracle needs about 7 minutes on one core,
too long for a test, so its output is stored here and the GPU cell is checked against it.

    python tests/golden/make_sobol_golden.py
"""
import os
import sys
from dataclasses import replace

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import engine as oeng  # noqa: E402
from oracle import model as om  # noqa: E402
from oracle import sobol as osob  # noqa: E402
from rpm_amd import sobol as psob  # noqa: E402  (the configuration record only)

THETA, SAMPLE, Q = 90.0, 0, 4
cfg = replace(psob.RunConfig(), lifetimes_us=(1.0,), sobol_scrambles=2)
h = oeng.build_quantum_needle_hamiltonian(THETA, cfg.field_microT, quadrupole_scale=cfg.quadrupole_scale)
state = oeng.singlet_initial_state(om.SYSTEM_DIMENSIONS, np.random.default_rng(osob.seed(cfg.seed, 0xA11CE, SAMPLE)))
yields, bounds, res = osob.run_cell(cfg, h, state, Q, SAMPLE)
st = res.stats
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sobol_full_cell.npz")
np.savez(out, theta_deg=THETA, sample_id=SAMPLE, q_points=Q, lifetimes_us=np.array(cfg.lifetimes_us),
         sobol_scrambles=cfg.sobol_scrambles, schedule=osob.reaction_time_schedule(cfg, Q, SAMPLE),
         yields=yields, propagation_bounds=bounds, observation_times=res.times_us,
         populations=res.observable_values, accepted_steps=st.accepted_steps, rejected_steps=st.rejected_steps,
         matvecs=st.matvecs, maximum_krylov_dimension=st.maximum_krylov_dimension)
print("wrote", out, yields.ravel(), st)
