"""Copy the archived full-system per-trajectory yields into a small fixture.

Source: results/quadrupole/full14_uncoupled_tau5_29angles/raw_sample_yields.csv of
the reference (464 rows: {D=J=0, D/J=-3} x Q in {0,1} x 29 angles x 4 samples,
B = 50 uT, tau = 5 us, tail bound 1e-5, seed 20260731).  Only data columns are
kept; host-specific checkpoint paths are dropped.  Run in a container that has
the reference mounted:  python tests/golden/extract_archive.py
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
