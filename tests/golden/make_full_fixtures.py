"""Record full-system (dim 221,184) oracle runs as fixtures for the GPU parity tests.

The oracle needs minutes to tens of minutes per run at this size on one core, too
long for a test, so its outputs are stored here (a few tens of KB) and the GPU runs
are checked against them:

* ``full_random_tol1e10.npz`` -- the random-tensor 14-nucleus system (tensors drawn
  as for the 576-level system: isotropic U(0,1) mT + anisotropic N(0, 0.3 mT),
  random traceless Q of 1 MHz on the three 14N), 50 uT at theta = 0, production
  J and D, lifetime 1 us, horizon 5 us (4,001 grid points), Krylov tolerance
  (state-error budget) 1e-10 and maximum dimension 40 (ladder 24, 32, 40);
  random-phase vectors 0 and 1 of seed 0.  Step schedules, P_S(t), yields.
* ``full_field_gradient_fd.npz`` -- the production operator at (theta, phi) =
  (90, 0) and (60, 30) deg, 50 uT, lifetime 5 us, horizon 0.1 us (81 points),
  budget 1e-10; the centre run and six runs at theta +- 1e-4 rad, phi +- 1e-4 rad,
  |B| +- 1e-3 uT replaying the centre's step schedule; their central differences.
* ``full_pulse64.npz`` -- the production operator with 50 uT static along z plus a
  transverse (x) piecewise-constant pulse of 64 amplitudes U(-20, 20) uT
  (default_rng(11)) over [0, 2] us, horizon 2.1 us (1,681 points), lifetime 5 us,
  budget 1e-10, segment loop with a fresh propagator per segment; random-phase
  vectors 0 and 1 of the production seed; for vector 0 the central differences
  (h = 1e-3 uT) of the yield with respect to amplitudes 0, 21, 42 and 63, the
  perturbed runs replaying the centre's per-segment schedules.

Schedules are stored as float rows (segment, trial_us, accepted, selected, size0,
size1 or -1), one per attempt.

    python tests/golden/make_full_fixtures.py [--workers 7] [--cache DIR] [config1|grad|pulse ...]

With ``--cache DIR`` every finished run is also stored there and reused by a later
invocation, so an interrupted generation resumes where it stopped.
"""
from __future__ import annotations

import argparse
import math
import pickle
import os
import sys
import time
from concurrent.futures import FIRST_COMPLETED, ProcessPoolExecutor, wait

for _v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ[_v] = "1"

import numpy as np  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import core as ocore  # noqa: E402
from oracle import engine as oengine  # noqa: E402
from oracle.model import field_vector_mT  # noqa: E402

STEP = 0.00125
H_ANGLE = 1e-4       # rad
H_FIELD = 1e-3       # uT
PULSE_SEGMENTS = 64
PULSE_WINDOW = (0.0, 2.0)
PULSE_HORIZON = 2.1
PULSE_FD_INDICES = (0, 21, 42, 63)
GRAD_ORIENTATIONS = ((90.0, 0.0, 0), (60.0, 30.0, 1))   # theta, phi, random-phase vector
GRAD_HORIZON = 0.1
BUDGET = 1e-10


def grid(horizon):
    return np.arange(int(round(horizon / STEP)) + 1) * STEP


def pack(schedules):
    """Per-segment schedules -> (n, 6) float rows."""
    rows = []
    for s, sched in enumerate(schedules):
        for trial, ok, sel, sizes in sched:
            sizes = tuple(sizes)
            rows.append((s, trial, float(ok), sel, sizes[0], sizes[1] if len(sizes) > 1 else -1))
    return np.array(rows, dtype=float).reshape(-1, 6)


def unpack(rows):
    """(n, 6) rows -> list per segment of (trial, accepted, selected, sizes) tuples."""
    out = []
    for seg, trial, ok, sel, s0, s1 in np.asarray(rows, dtype=float):
        while len(out) <= int(seg):
            out.append([])
        sizes = (int(s0),) if s1 < 0 else (int(s0), int(s1))
        out[int(seg)].append((float(trial), bool(ok), int(sel), sizes))
    return out


# ---------------------------------------------------------------- config 1: random tensors, tol 1e-10, dim <= 40

def random_system():
    from rpm_amd import workloads   # the seeded tensor draws only (host NumPy)
    return workloads.full_random_system(1)


def config1_run(sample):
    sys_ = random_system()
    h = sys_.hamiltonian(builder=ocore.build_hamiltonian)
    times = grid(5.0)
    prop = oengine.AdaptiveLanczosPropagator(h, minimum_dimension=24, maximum_dimension=40, dimension_increment=8,
                                             embedded_offset=4, safety_factor=2.0, state_error_budget=BUDGET,
                                             initial_step_us=0.05, maximum_step_us=0.25)
    state = ocore.random_phase_initial_state(h.dimensions, 0, sample)
    t0 = time.time()
    res = prop.propagate(state, times, oengine.SingletPopulationObservable(h.dimensions))
    st = res.stats
    fine = ocore.integrate_yield(times, res.observable_values, sys_.lifetime_us)
    coarse_idx = np.append(np.arange(0, times.size, 2), [] if (times.size - 1) % 2 == 0 else [times.size - 1])
    coarse = ocore.integrate_yield(times[coarse_idx.astype(int)], res.observable_values[coarse_idx.astype(int)],
                                   sys_.lifetime_us)
    return dict(sample=sample, seconds=time.time() - t0, schedule=pack([res.schedule]),
                populations=res.observable_values, errors=res.state_error_estimates, yield_=fine, coarse=coarse,
                accepted=st.accepted_steps, rejected=st.rejected_steps, matvecs=st.matvecs,
                acc_err=st.accumulated_state_error, final_norm=float(np.linalg.norm(res.final_state)),
                final_singlet=oengine.singlet_population(res.final_state, h.dimensions))


# ---------------------------------------------------------------- config 3: (theta, phi, |B|) central differences

def grad_run(theta, phi, field_uT, sample, schedule_rows=None):
    h = ocore.build_hamiltonian(theta, 1e-3 * field_uT, phi_deg=phi)
    times = grid(GRAD_HORIZON)
    prop = oengine.AdaptiveLanczosPropagator(h, state_error_budget=BUDGET)
    state = ocore.random_phase_initial_state(h.dimensions, ocore.MASTER_SEED, sample)
    sched = None if schedule_rows is None else unpack(schedule_rows)[0]
    res = prop.propagate(state, times, oengine.SingletPopulationObservable(h.dimensions), schedule=sched)
    y = ocore.integrate_yield(times, res.observable_values, 5.0)
    return dict(yield_=y, schedule=pack([res.schedule]), populations=res.observable_values)


# ---------------------------------------------------------------- config 4: 64-segment transverse pulse

def pulse_amplitudes():
    return np.random.default_rng(11).uniform(-20.0, 20.0, size=PULSE_SEGMENTS)


def _field_h(bvec_mT):
    mag = float(np.linalg.norm(bvec_mT))
    th = math.degrees(math.acos(bvec_mT[2] / mag))
    ph = math.degrees(math.atan2(bvec_mT[1], bvec_mT[0]))
    return ocore.build_hamiltonian(th, mag, phi_deg=ph)


def pulse_run(sample, amps, schedule_rows=None):
    static = field_vector_mT(0.0, 50.0)
    edges = np.linspace(PULSE_WINDOW[0], PULSE_WINDOW[1], PULSE_SEGMENTS + 1)
    hs = [_field_h(static + 1e-3 * a * np.array([1.0, 0.0, 0.0])) for a in amps]
    hs.append(ocore.build_hamiltonian(0.0, 0.05))
    times = grid(PULSE_HORIZON)
    state = ocore.random_phase_initial_state(hs[0].dimensions, ocore.MASTER_SEED, sample)
    sched = None if schedule_rows is None else unpack(schedule_rows)
    t0 = time.time()
    pops, bounds, acc, final, schedules = ocore.run_segments(
        hs[0], state, times, edges[1:], lambda h: oengine.AdaptiveLanczosPropagator(h), BUDGET,
        hamiltonians=hs, schedules=sched)
    y = ocore.integrate_yield(times, pops, 5.0)
    return dict(sample=sample, seconds=time.time() - t0, yield_=y, schedule=pack(schedules), populations=pops,
                bounds=bounds, acc_err=acc, final_singlet=oengine.singlet_population(final, hs[0].dimensions))


# ---------------------------------------------------------------- driver

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", nargs="*", default=["config1", "grad", "pulse"])
    ap.add_argument("--workers", type=int, default=7)
    ap.add_argument("--cache", default=None, help="directory keeping finished runs between invocations")
    args = ap.parse_args()
    pool = ProcessPoolExecutor(args.workers)
    pending = {}
    results = {}
    cached = []
    if args.cache:
        os.makedirs(args.cache, exist_ok=True)

    def cache_file(key):
        return os.path.join(args.cache, "_".join(str(k) for k in key) + ".pkl")

    def submit(key, fn, *a):
        if args.cache and os.path.exists(cache_file(key)):
            with open(cache_file(key), "rb") as fh:    # this script's own earlier output
                cached.append((key, pickle.load(fh)))
            return
        pending[pool.submit(fn, *a)] = key

    def finish(key, res):
        results[key] = res
        print(f"[{time.time() - t0:7.0f} s] done {key}", flush=True)
        if args.cache and not os.path.exists(cache_file(key)):
            with open(cache_file(key), "wb") as fh:
                pickle.dump(res, fh)
        if key[0] == "g" and key[2] == "c":
            th, ph, s = GRAD_ORIENTATIONS[key[1]]
            dd = math.degrees(H_ANGLE)
            for k, (a, b, c) in enumerate(((dd, 0, 0), (0, dd, 0), (0, 0, H_FIELD))):
                for sign in (1, -1):
                    submit(("g", key[1], k, sign), grad_run, th + sign * a, ph + sign * b, 50.0 + sign * c, s,
                           res["schedule"])
        if key[0] == "p" and key[1] == 0 and key[2] == "c":
            for idx in PULSE_FD_INDICES:
                for sign in (1, -1):
                    e = np.zeros_like(amps)
                    e[idx] = sign * H_FIELD
                    submit(("p", 0, idx, sign), pulse_run, 0, amps + e, res["schedule"])

    amps = pulse_amplitudes()
    t0 = time.time()
    if "config1" in args.which:
        for s in (0, 1):
            submit(("c1", s), config1_run, s)
    if "grad" in args.which:
        for i, (th, ph, s) in enumerate(GRAD_ORIENTATIONS):
            submit(("g", i, "c"), grad_run, th, ph, 50.0, s)
    if "pulse" in args.which:
        for s in (0, 1):
            submit(("p", s, "c"), pulse_run, s, amps)
    while pending or cached:
        while cached:
            finish(*cached.pop(0))
        if not pending:
            break
        done, _ = wait(list(pending), return_when=FIRST_COMPLETED)
        for fut in done:
            key = pending.pop(fut)
            finish(key, fut.result())

    if "config1" in args.which:
        r = [results[("c1", s)] for s in (0, 1)]
        out = os.path.join(HERE, "full_random_tol1e10.npz")
        np.savez_compressed(
            out, samples=np.array([0, 1]), horizon_us=5.0, lifetime_us=1.0, budget=BUDGET, maximum_dimension=40,
            times=grid(5.0), populations=np.stack([x["populations"] for x in r]),
            errors=np.stack([x["errors"] for x in r]), yields=np.array([x["yield_"] for x in r]),
            coarse_yields=np.array([x["coarse"] for x in r]),
            schedule0=r[0]["schedule"], schedule1=r[1]["schedule"],
            accepted=np.array([x["accepted"] for x in r]), rejected=np.array([x["rejected"] for x in r]),
            matvecs=np.array([x["matvecs"] for x in r]), accumulated_error=np.array([x["acc_err"] for x in r]),
            final_norm=np.array([x["final_norm"] for x in r]), final_singlet=np.array([x["final_singlet"] for x in r]),
            oracle_seconds=np.array([x["seconds"] for x in r]))
        print("wrote", out, [x["yield_"] for x in r], [x["seconds"] for x in r], flush=True)
    if "grad" in args.which:
        n = len(GRAD_ORIENTATIONS)
        fd = np.zeros((n, 3))
        center = np.zeros(n)
        plus = np.zeros((n, 3))
        minus = np.zeros((n, 3))
        for i in range(n):
            center[i] = results[("g", i, "c")]["yield_"]
            for k in range(3):
                plus[i, k] = results[("g", i, k, 1)]["yield_"]
                minus[i, k] = results[("g", i, k, -1)]["yield_"]
                fd[i, k] = (plus[i, k] - minus[i, k]) / (2.0 * (H_ANGLE if k < 2 else H_FIELD))
        out = os.path.join(HERE, "full_field_gradient_fd.npz")
        np.savez_compressed(
            out, orientations=np.array(GRAD_ORIENTATIONS, dtype=float), field_uT=50.0, horizon_us=GRAD_HORIZON,
            lifetime_us=5.0, budget=BUDGET, h_angle=H_ANGLE, h_field=H_FIELD, center_yields=center,
            plus_yields=plus, minus_yields=minus, fd_gradient=fd,
            center_populations=np.stack([results[("g", i, "c")]["populations"] for i in range(n)]),
            **{f"schedule{i}": results[("g", i, "c")]["schedule"] for i in range(n)})
        print("wrote", out, fd, flush=True)
    if "pulse" in args.which:
        c = [results[("p", s, "c")] for s in (0, 1)]
        fd = np.array([(results[("p", 0, i, 1)]["yield_"] - results[("p", 0, i, -1)]["yield_"]) / (2 * H_FIELD)
                       for i in PULSE_FD_INDICES])
        out = os.path.join(HERE, "full_pulse64.npz")
        np.savez_compressed(
            out, amplitudes_uT=amps, window_us=np.array(PULSE_WINDOW), horizon_us=PULSE_HORIZON, lifetime_us=5.0,
            budget=BUDGET, samples=np.array([0, 1]), yields=np.array([x["yield_"] for x in c]),
            populations=np.stack([x["populations"] for x in c]), bounds=np.stack([x["bounds"] for x in c]),
            accumulated_error=np.array([x["acc_err"] for x in c]),
            final_singlet=np.array([x["final_singlet"] for x in c]),
            schedule0=c[0]["schedule"], schedule1=c[1]["schedule"], fd_indices=np.array(PULSE_FD_INDICES),
            fd_gradient=fd, h_field=H_FIELD,
            fd_plus=np.array([results[("p", 0, i, 1)]["yield_"] for i in PULSE_FD_INDICES]),
            fd_minus=np.array([results[("p", 0, i, -1)]["yield_"] for i in PULSE_FD_INDICES]),
            oracle_seconds=np.array([x["seconds"] for x in c]))
        print("wrote", out, fd, flush=True)
    pool.shutdown()


if __name__ == "__main__":
    main()
