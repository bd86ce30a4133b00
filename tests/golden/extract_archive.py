"""Copy archived reference results into small fixtures.

1. The archived full-system per-trajectory yields (archived_full14_yields.csv).

Source: results/quadrupole/full14_uncoupled_tau5_29angles/raw_sample_yields.csv of
the reference (464 rows: {D=J=0, D/J=-3} x Q in {0,1} x 29 angles x 4 samples,
B = 50 uT, tau = 5 us, tail bound 1e-5, seed 20260731).  Only data columns are
kept; host-specific checkpoint paths are dropped.  Run in a container that has
the reference mounted:  python tests/golden/extract_archive.py

2. The factorized Fig. 1B curve (archived_fig1b_q0.csv), copied unchanged from
results/quadrupole/zeeman_hyperfine_reference/quantum_needle_fig1b.csv: 199 angles
(theta_grid: 1 deg, plus 0.25 deg within 3 deg of 90), Q = 0, J = D = 0, B = 50 uT,
bins_per_linewidth 8, lifetimes 1, 2, 5, 10, 20, 100 us.
"""
import csv
import os

SRC = ("/root/reference/results/quadrupole/full14_uncoupled_tau5_29angles/"
       "raw_sample_yields.csv")
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "archived_full14_yields.csv")
KEEP = ("electron_interactions", "scale", "theta_deg", "sample_id", "yield", "quadrature_error",
        "yield_propagation_estimate", "yield_tail_bound", "initial_state_sha256")

with open(SRC, newline="") as fin, open(DST, "w", newline="") as fout:
    writer = csv.DictWriter(fout, fieldnames=KEEP)
    writer.writeheader()
    for row in csv.DictReader(fin):
        writer.writerow({k: row[k] for k in KEEP})
print("wrote", DST)

SRC2 = "/root/reference/results/quadrupole/zeeman_hyperfine_reference/quantum_needle_fig1b.csv"
DST2 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "archived_fig1b_q0.csv")
with open(SRC2, newline="") as fin, open(DST2, "w", newline="") as fout:
    rows = list(csv.DictReader(fin))
    writer = csv.DictWriter(fout, fieldnames=list(rows[0]))
    writer.writeheader()
    writer.writerows(rows)
print("wrote", DST2)
