"""Benchmark: batched H·psi on the 221,184-dimensional FAD-TrpH pair state (one MI355X).

A step is one batched H·psi launch over B resident complex128 vectors of the
production operator (theta = 90 deg, 50 uT, Efimova-Hore D/J = -3, 14N
quadrupoles on).  value = H·psi per second over the timed steps; the roofline
object prices the kernel's algorithmic bytes (read psi + write H psi,
32 B per amplitude = 7,077,888 B per vector) against the 8 TB/s HBM peak.
The same JSON line carries the other two parts of the metric over the whole
57.56 us horizon (the reference's unit of work is a whole trajectory):
"batched" (32 whole trajectories of the reference's archived run, 8 field
angles x 4 random-phase vectors, in one batched propagation, against
"trajectory", one whole trajectory of the same run on the plain path; yields
per second and their ratio; yields checked against the archive) and "gradient"
(8 whole production trajectories: yield plus adjoint (theta, phi, |B|) gradient
over yield alone), plus "spectral" (the factorized J = D = 0 oracle's 199-angle
curve on the GPU against the oracle's per-orientation CPU time), "sobol" (16
batched Sobol reaction-time cells on the full system), "directional" (the
directional-map composition over a short horizon) and "propagation" (the
reference's 0.1 us speed probe, priced on the bytes this implementation moves).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--no-cpu] [--no-extras]
    python bench.py --mode propagate|batch|directional|grad|spectral|sobol|trajectory|copy ...
    python bench.py --mode archive_batch   # 32 whole archived trajectories in one batch (about a minute)
    python bench.py --mode archive_all     # every archived trajectory (464, about 10 minutes); --archive-groups
    python bench.py --mode gradcheck       # whole-horizon adjoint gradient against central differences
    python bench.py --mode pulse           # one whole-horizon ascent step of the 64-segment pulse
    python bench.py --mode ceiling         # copy and two-pass streaming bounds beside the H·psi kernel
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = ("H·ψ/s at dim 221184 (% HBM peak); batched yields/s; yield+grad vs yield time")
HBM_PEAK_GBS = 8000.0
BYTES_PER_AMPLITUDE = 32


def cpu_baseline(seconds: float = 12.0) -> dict:
    """Single-core H·psi/s of the oracle (NumPy/SciPy restatement of the reference matvec)."""
    code = r"""
import json, sys, time
sys.path.insert(0, %r)
import numpy as np
from oracle import core
h = core.build_hamiltonian(90.0, 0.05)
rng = np.random.default_rng(0)
v = rng.normal(size=h.dimension) + 1j * rng.normal(size=h.dimension)
v /= np.linalg.norm(v)
h.matvec(v)
n, t0 = 0, time.perf_counter()
while time.perf_counter() - t0 < %f:
    v = h.matvec(v); v /= np.linalg.norm(v); n += 1
dt = time.perf_counter() - t0
print(json.dumps({"n": n, "dt": dt}))
""" % (ROOT, seconds)
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1",
               NUMEXPR_NUM_THREADS="1")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, check=True)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return {"value": r["n"] / r["dt"], "unit": "H·ψ/s", "cores": 1, "kind": "port",
            "sample": f"{r['n']} sequential oracle CompiledFullPairHamiltonian.matvec calls on one vector "
                      f"(production operator, dim 221184) over {r['dt']:.1f} s, 1 thread"}


def cpu_extra_baselines(prefix_us: float = 0.1, all_core_seconds: float = 6.0) -> dict:
    """The oracle on the host: (ii) single-core trajectory times, config 0 (576 levels, 5 us) and a
    0.1 us full-system prefix with the horizon-scaled budget, extrapolated to the full horizon;
    (iii) all-core compiled-pair H.psi throughput, one worker per core up to 16."""
    code = r"""
import json, math, sys, time
sys.path.insert(0, %r)
import numpy as np
from oracle import core as ocore
from rpm_amd import workloads
out = {}
sys0 = workloads.small_system(0)
h0 = sys0.hamiltonian(builder=ocore.build_hamiltonian)
cfg0 = ocore.SweepConfig(lifetime_us=1.0, survival_tail_bound=math.exp(-5.0))
t0 = time.perf_counter()
ocore.run_trajectory(h0, ocore.random_phase_initial_state(h0.dimensions, 0, 0), cfg0)
out["config0_seconds"] = time.perf_counter() - t0
hp = ocore.build_hamiltonian(90.0, 0.05)
cfg = ocore.SweepConfig(lifetime_us=5.0, survival_tail_bound=1e-5)
times = np.arange(int(round(%f / 0.00125)) + 1) * 0.00125
prop = ocore.make_propagator(hp, cfg, state_error_budget=cfg.state_error_budget * float(times[-1]) / cfg.maximum_time_us)
s = ocore.random_phase_initial_state(hp.dimensions, ocore.MASTER_SEED, 0)
t0 = time.perf_counter()
prop.propagate(s, times, ocore.SingletPopulationObservable(hp.dimensions))
dt = time.perf_counter() - t0
out["prefix_us"] = float(times[-1])
out["prefix_seconds"] = dt
out["full_horizon_us"] = cfg.maximum_time_us
out["extrapolated_trajectory_seconds"] = dt * cfg.maximum_time_us / float(times[-1])
print(json.dumps(out))
""" % (ROOT, prefix_us)
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1",
               NUMEXPR_NUM_THREADS="1")
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, check=True)
    out = json.loads(res.stdout.strip().splitlines()[-1])
    # one worker per CPU this process may run on (its affinity set), at most 16: the GPU box gives one
    # GPU's run a share of 16 CPUs whatever os.cpu_count() reports
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    workers = max(1, min(16, affinity))
    worker = r"""
import json, sys, time
sys.path.insert(0, %r)
import numpy as np
from oracle import core
h = core.build_hamiltonian(90.0, 0.05)
v = np.random.default_rng(int(sys.argv[1])).normal(size=h.dimension) + 0j
v /= np.linalg.norm(v)
h.matvec(v)
n, t0 = 0, time.perf_counter()
while time.perf_counter() - t0 < %f:
    v = h.matvec(v); v /= np.linalg.norm(v); n += 1
print(json.dumps({"n": n, "dt": time.perf_counter() - t0}))
""" % (ROOT, all_core_seconds)
    procs = [subprocess.Popen([sys.executable, "-c", worker, str(i)], stdout=subprocess.PIPE, text=True, env=env)
             for i in range(workers)]
    rate = 0.0
    for pr in procs:
        o, _ = pr.communicate()
        r = json.loads(o.strip().splitlines()[-1])
        rate += r["n"] / r["dt"]
    out["all_core_hpsi_per_s"] = rate
    out["all_core_workers"] = workers
    out["host_cpu_count"] = os.cpu_count()
    out["host_affinity_cpus"] = affinity
    out["kind"] = "port"
    return out


def measured_traffic(batch: int):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary, if present."""
    path = os.path.join(ROOT, "profiles", "hpsi_pair_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    per_vector = data.get("hbm_bytes_per_vector")
    return None if per_vector is None else per_vector * batch


def lanczos_algorithmic_bytes(schedule, n: int) -> int:
    """Algorithmic HBM bytes of a propagation under the reference's CGS2 semantics: per Krylov basis of
    size m, 16 N (2m [matvec] + 2m(m+1) [two projections over the basis] + 7m [alpha, three-term step,
    norm, scale]); per accepted step, 16 N (m+1) for the endpoint V c and 8 N m for the singlet
    projection, m the selected dimension. Summed over the recorded schedule (accepted and rejected)."""
    total = 0
    for _, accepted, selected, sizes in schedule:
        for m in sizes:
            total += 16 * n * (2 * m + 2 * m * (m + 1) + 7 * m)
        if accepted:
            total += 16 * n * (selected + 1) + 8 * n * selected
    return total


PAIR_FROM = 12   # first paired Lanczos step of the device basis builder


def basis_device_vectors(m: int) -> int:
    """Vector reads and writes (units of 16 N bytes) of one device basis of m vectors built in one call:
    a single step j moves j + 6 (matvec in and out; the pass over V[0..j), x_j and H x_j read, v_j and
    y written); a pair slot (j, j+1) from step 12 on moves j + 17 (two matvecs, 4; the prep of z, 4; the
    pass over the j - 1 materialised columns with x_j, z, H z and the pending z' staged and v_{j-1},
    v_j written, j + 5; the close forming x_{j+2}, 4), j + 16 for the first pair (one pending vector);
    the basis's last pair materialises its last vector instead of forming x_{j+2} (j + 3)."""
    vecs, j, first = 0, 0, True
    while j < m:
        if j < PAIR_FROM or j + 1 >= m:
            vecs += j + 6
            j += 1
            continue
        vecs += j + (16 if first else 17)
        if j + 2 >= m:
            vecs += (j + 3) - 4
        first = False
        j += 2
    return vecs


def lanczos_device_bytes(schedule, n: int, built=None) -> int:
    """HBM bytes this implementation moves for the same schedule (one read of the basis per step, or per
    pair of steps from step 12 on; basis_device_vectors); per accepted step the same endpoint and
    singlet terms as lanczos_algorithmic_bytes. ``built``: the vectors the device built per attempt (bases
    are built only as far as a candidate needs), counted as one call each; default, the reference's
    basis sizes. The rare explicit corrections are not counted."""
    total = 0
    for i, (_, accepted, selected, sizes) in enumerate(schedule):
        for m in ([built[i]] if built is not None else sizes):
            total += 16 * n * basis_device_vectors(m)
        if accepted:
            total += 16 * n * (selected + 1) + 8 * n * selected
    return total


def propagate_probe(h, args, horizon=None) -> dict:
    """One production trajectory over a short horizon with the horizon-scaled error budget
    of the reference's speed probe (R/rpm_full14_quadrupole_experiment.py:138-148)."""
    import numpy as np
    import torch
    from rpm_amd import core
    horizon = args.horizon if horizon is None else horizon
    cfg = core.SweepConfig(lifetime_us=5.0, survival_tail_bound=1e-5)
    times = np.arange(int(round(horizon / 0.00125)) + 1) * 0.00125
    prop = core.make_propagator(h, cfg, probe_horizon=float(times[-1]))
    # the initial state is resident in HBM before the timed region
    s = torch.from_numpy(core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, 0)).cuda()
    obs = core.SingletPopulationObservable(h.dimensions)
    prop.propagate(s, times[:3], obs)          # warm-up (allocations, first launches)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = prop.propagate(s, times, obs)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = res.stats
    cgs2_bytes = lanczos_algorithmic_bytes(res.schedule, h.dimension)
    dbytes = lanczos_device_bytes(res.schedule, h.dimension, prop.device_basis_sizes)
    return {"mode": "propagate", "horizon_us": float(times[-1]), "seconds": dt,
            "simulated_us_per_s": float(times[-1]) / dt, "accepted": st.accepted_steps,
            "rejected": st.rejected_steps, "matvecs_reference_count": st.matvecs,
            "matvecs_device": prop.device_matvecs, "max_krylov": st.maximum_krylov_dimension,
            "ms_per_device_matvec_incl_orthogonalisation": dt * 1e3 / prop.device_matvecs,
            # priced on the bytes this implementation must move (one read of the basis per Lanczos step);
            # the reference's CGS2 would move cgs2_bytes for the same schedule
            "roofline": {"bound": "hbm", "achieved": dbytes / dt / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": dbytes / dt / 1e9 / HBM_PEAK_GBS, "traffic": None, "algorithmic_bytes": dbytes,
                         "definition": "device bytes of the recorded schedule: per single Lanczos step j, 16 N "
                                       "(j + 6); per pair of steps (j, j+1) from step 12 on, 16 N (j + 17); per "
                                       "accepted step the endpoint V c and the singlet projection",
                         "reference_cgs2_bytes": cgs2_bytes}}


# Wall time per full-system trajectory of the reference's archived uncoupled run (n = 232, one
# numerical thread per process): results/quadrupole/full14_uncoupled_tau5_29angles/
# execution_result.json, durations_seconds.
REFERENCE_TRAJECTORY_SECONDS = {"median": 25978.29, "min": 12351.87, "max": 28132.45, "n": 232, "threads": 1}
ARCHIVED_YIELDS = os.path.join(ROOT, "tests", "golden", "archived_full14_yields.csv")


def trajectory_probe(theta: float = 90.0, quadrupole_scale: float = 1.0, sample: int = 0) -> dict:
    """One whole trajectory of the reference's archived uncoupled run, end to end: D = J = 0,
    14N quadrupoles on, 50 uT at theta, tau = 5 us, tail bound 1e-5 (horizon 57.5646 us on the
    46,053-point grid), budget 2.4e-5, seed 20260731, sample `sample`
    (R/rpm_full14_uncoupled_quadrupole.py, R/rpm_full14_quadrupole_experiment.py:154-194).
    Timed from the host initial state to the host yield, as the reference's per-trajectory wall
    time is; the archived yield of the same cell is read from the repository's fixture."""
    import csv

    import numpy as np
    import torch
    from rpm_amd import core
    model = core.ProductionModel(quadrupole_scale=quadrupole_scale, ratio_D_over_J=None)
    cfg = core.SweepConfig(seed=core.MASTER_SEED, field_microT=50.0, lifetime_us=5.0, survival_tail_bound=1e-5,
                           state_error_budget=2.4e-5, quadrupole_scale=quadrupole_scale)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    h = core.build_hamiltonian(theta, 0.05, model=model)
    state = core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, sample)
    out = core.propagate_trajectory(h, state, cfg)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = out["result"].stats
    archived = None
    if os.path.exists(ARCHIVED_YIELDS):
        with open(ARCHIVED_YIELDS) as f:
            for r in csv.DictReader(f):
                if (r["electron_interactions"] == "D=J=0" and float(r["scale"]) == quadrupole_scale
                        and float(r["theta_deg"]) == theta and int(r["sample_id"]) == sample):
                    archived = float(r["yield"])
    horizon = float(out["times_us"][-1])
    ref = REFERENCE_TRAJECTORY_SECONDS
    return {"mode": "trajectory", "cell": {"electron_interactions": "D=J=0", "quadrupole_scale": quadrupole_scale,
                                            "theta_deg": theta, "field_uT": 50.0, "tau_us": 5.0, "sample_id": sample},
            "horizon_us": horizon, "grid_points": int(out["times_us"].size), "seconds": dt,
            "simulated_us_per_s": horizon / dt, "accepted": st.accepted_steps, "rejected": st.rejected_steps,
            "matvecs": st.matvecs, "yield": out["yield"], "archived_yield": archived,
            "yield_minus_archived": None if archived is None else out["yield"] - archived,
            "yield_propagation_estimate": out["yield_propagation_estimate"],
            "reference_seconds": ref, "speedup_vs_reference_median": ref["median"] / dt}


def archive_batch_probe(thetas=(0.0, 20.0, 40.0, 60.0, 80.0, 90.0, 120.0, 160.0), samples: int = 4,
                        quadrupole_scale: float = 1.0, interactions: str = "D=J=0", batch_size: int | None = None,
                        label: str = "archive_batch") -> dict:
    """Whole trajectories of the reference's archived full-system run in batched propagations, each
    yield compared with its archived value. Default: 32 cells of the uncoupled run (8 field angles x
    samples 0-3, D = J = 0, quadrupoles on, the settings of trajectory_probe) in one batch. thetas=None
    takes every archived angle of the (interactions, quadrupole_scale) group; batch_size splits the
    cells into consecutive batches. Timed from the host states to the host yields."""
    import csv

    import numpy as np
    import torch
    from rpm_amd import batch, core
    model = core.ProductionModel(quadrupole_scale=quadrupole_scale,
                                 ratio_D_over_J=None if interactions == "D=J=0" else -3.0)
    if interactions not in ("D=J=0", "D/J=-3"):
        raise ValueError(f"no archived group {interactions!r}")
    cfg = core.SweepConfig(seed=core.MASTER_SEED, field_microT=50.0, lifetime_us=5.0, survival_tail_bound=1e-5,
                           state_error_budget=2.4e-5, quadrupole_scale=quadrupole_scale)
    archived = {}
    with open(ARCHIVED_YIELDS) as f:
        for r in csv.DictReader(f):
            if r["electron_interactions"] == interactions and float(r["scale"]) == quadrupole_scale:
                archived[(float(r["theta_deg"]), int(r["sample_id"]))] = float(r["yield"])
    if thetas is None:
        thetas = sorted({t for t, _ in archived})
    cells = [(float(t), s) for t in thetas for s in range(samples)]
    missing = [c for c in cells if c not in archived]
    if missing:
        raise ValueError(f"cells not in the archived run: {missing}")
    times = cfg.time_grid_us
    size = len(cells) if batch_size is None else max(1, int(batch_size))

    def progress(b, now, accepted, rejected, matvecs, acc_err):
        if b == 0 and accepted % 200 == 0:
            print(f"[bench] {label}: trajectory 0 of the batch at {now:.2f} us, {accepted} steps", file=sys.stderr,
                  flush=True)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    base = core.build_hamiltonian(90.0, 0.05, model=model)
    hs = {t: core.build_hamiltonian(t, 0.05, model=model) for t in thetas}
    weights = core.yield_weight_table(times, cfg.lifetime_us, core.coarse_grid_indices(cfg))
    yields, host, matvecs_ref, matvecs_dev, rounds = [], [], 0, 0, 0
    for k in range(0, len(cells), size):
        part = cells[k:k + size]
        pairs = batch.pair_matrices(base, [hs[t] for t, _ in part])
        states = np.stack([core.random_phase_initial_state(base.dimensions, core.MASTER_SEED, smp) for _, smp in part])
        bp = batch.BatchedLanczosPropagator.like(core.make_propagator(base, cfg))
        res = bp.propagate(states, times, pairs=pairs, progress_callback=progress, yield_weights=weights)
        yields += [r.device_yields[0] for r in res]       # summed on the device
        host += [core.integrate_yield(times, r.observable_values, cfg.lifetime_us) for r in res]
        matvecs_ref += int(sum(r.stats.matvecs for r in res))
        matvecs_dev += int(bp.device_matvecs.sum())
        rounds += bp.rounds
        print(f"[bench] {label}: {min(k + size, len(cells))} of {len(cells)} cells after "
              f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    yields, host = np.array(yields), np.array(host)
    diff = yields - np.array([archived[c] for c in cells])
    horizon = float(times[-1])
    per_traj = dt / len(cells)
    ref = REFERENCE_TRAJECTORY_SECONDS
    return {"mode": label, "electron_interactions": interactions, "quadrupole_scale": quadrupole_scale,
            "trajectories": len(cells), "thetas_deg": list(thetas), "samples": samples, "batch_size": size,
            "horizon_us": horizon, "seconds": dt,
            "trajectory_us_per_s": len(cells) * horizon / dt, "seconds_per_trajectory": per_traj,
            "matvecs_reference_count": matvecs_ref, "matvecs_device": matvecs_dev, "rounds": rounds,
            "max_abs_yield_minus_archived": float(np.abs(diff).max()),
            "max_abs_device_minus_host_yield": float(np.abs(yields - host).max()),
            "reference_seconds": ref, "speedup_vs_reference_median": ref["median"] / per_traj}


def archive_all_probe(groups, batch_size: int) -> dict:
    """Every archived trajectory of the given (interactions, quadrupole scale) groups (29 angles x 4
    samples each; the archive holds four groups, 464 trajectories), batched, against its archived yield."""
    out = {"mode": "archive_all", "groups": []}
    for interactions, scale in groups:
        out["groups"].append(archive_batch_probe(thetas=None, quadrupole_scale=scale, interactions=interactions,
                                                 batch_size=batch_size, label="archive_all"))
    g = out["groups"]
    out["trajectories"] = sum(x["trajectories"] for x in g)
    out["seconds"] = sum(x["seconds"] for x in g)
    out["seconds_per_trajectory"] = out["seconds"] / out["trajectories"]
    out["max_abs_yield_minus_archived"] = max(x["max_abs_yield_minus_archived"] for x in g)
    out["max_abs_device_minus_host_yield"] = max(x["max_abs_device_minus_host_yield"] for x in g)
    return out


def batch_probe(h, args, B=None, horizon=None, plain_sample=None) -> dict:
    """Batched trajectories (random-phase vectors x field orientations) against the plain per-vector path.

    Both run the same trajectories over the same short horizon with the
    horizon-scaled budget of propagate_probe; prints trajectory-microseconds
    simulated per second for each and their ratio.
    """
    import numpy as np
    import torch
    from rpm_amd import batch, core
    cfg = core.SweepConfig(lifetime_us=5.0, survival_tail_bound=1e-5)
    horizon_req = args.horizon if horizon is None else horizon
    times = np.arange(int(round(horizon_req / 0.00125)) + 1) * 0.00125
    B = args.batch if B is None else B
    orients = [(float(t), 0.0, 0.05) for t in np.linspace(0.0, 180.0, max(1, B // 4))]
    hs, states = [], []
    for i in range(B):
        th, ph, f = orients[i % len(orients)]
        hs.append(core.build_hamiltonian(th, f, phi_deg=ph))
        states.append(core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, i // len(orients)))
    # the initial states are resident in HBM before the timed region
    states = torch.from_numpy(np.stack(states)).cuda()
    pairs = batch.pair_matrices(h, hs)
    bp = batch.BatchedLanczosPropagator.like(core.make_propagator(h, cfg, probe_horizon=float(times[-1])))
    bp.use_graphs = not bool(getattr(args, "eager", False))
    bp.propagate(states, times[:3], pairs=pairs)      # warm-up at the full batch (allocations, first launches)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = bp.propagate(states, times, pairs=pairs)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    print(f"[bench] batched {B} trajectories: {t_batch:.3f} s, {bp.rounds} rounds", file=sys.stderr, flush=True)
    nplain = min(B, args.plain_sample if plain_sample is None else plain_sample)
    t_plain = 0.0
    for i in range(nplain):
        prop = core.make_propagator(hs[i], cfg, probe_horizon=float(times[-1]))
        obs = core.SingletPopulationObservable(h.dimensions)
        if i == 0:
            prop.propagate(states[i], times[:3], obs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = prop.propagate(states[i], times, obs)
        torch.cuda.synchronize()
        t_plain += time.perf_counter() - t0
        assert r.schedule == res[i].schedule
        assert float(np.abs(r.observable_values - res[i].observable_values).max()) < 1e-12
    horizon = float(times[-1])
    batched_rate = B * horizon / t_batch
    plain_rate = nplain * horizon / t_plain if nplain else float("nan")
    matvecs = sum(r.stats.matvecs for r in res)
    return {"mode": "batch", "trajectories": B, "orientations": len(orients), "horizon_us": horizon,
            "seconds_batched": t_batch, "trajectory_us_per_s_batched": batched_rate, "plain_sample": nplain,
            "trajectory_us_per_s_plain": plain_rate, "speedup_vs_plain": batched_rate / plain_rate,
            "yields_per_s_batched_equiv": batched_rate / cfg.maximum_time_us,
            "yields_per_s_plain_equiv": plain_rate / cfg.maximum_time_us,
            "full_horizon_us": cfg.maximum_time_us, "matvecs": matvecs, "rounds": bp.rounds,
            "graphs": bp.use_graphs}


def directional_probe(h, args, n_dirs=64, fields_uT=(25.0, 50.0), n_states=2, horizon=0.25) -> dict:
    """The directional-map composition of the batched workload: random-phase states x Fibonacci-sphere
    field directions x field strengths, through batch.directional_yields with the batch size picked by
    the run-time tuner, over a short horizon with the speed-probe budget."""
    import numpy as np
    import torch
    from rpm_amd import batch, core, workloads
    cfg = core.SweepConfig(lifetime_us=5.0, survival_tail_bound=1e-5)
    times = np.arange(int(round(horizon / 0.00125)) + 1) * 0.00125
    th, ph = workloads.fibonacci_sphere(n_dirs)
    states = torch.from_numpy(np.stack([core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, i)
                                        for i in range(n_states)])).cuda()
    bp = batch.BatchedLanczosPropagator.like(core.make_propagator(h, cfg, probe_horizon=float(times[-1])))
    build = lambda t, f_mT, p: core.build_hamiltonian(t, f_mT, phi_deg=p)  # noqa: E731
    batch.directional_yields(h, build, bp, states[:1], th[:2], ph[:2], fields_uT[:1], times[:3], cfg.lifetime_us,
                             batch_size=2)                                   # warm-up
    t0 = time.perf_counter()
    hs = [build(float(t), 1e-3 * f, float(p)) for t, p in zip(th, ph) for f in fields_uT]
    batch.pair_matrices(h, hs)
    t_host = time.perf_counter() - t0                                          # host-side Hamiltonians alone
    del hs
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Y = batch.directional_yields(h, build, bp, states, th, ph, fields_uT, times, cfg.lifetime_us, batch_size="auto")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    from rpm_amd import autotune
    chosen = [v["batch"] for v in autotune.cached("batch")]
    n = Y.size
    return {"mode": "directional", "trajectories": n, "states": n_states, "directions": n_dirs,
            "fields_uT": list(fields_uT), "horizon_us": float(times[-1]), "seconds": dt,
            "host_hamiltonian_seconds": t_host, "batch": chosen[-1] if chosen else None,
            "trajectory_us_per_s": n * float(times[-1]) / dt,
            "yields_per_s_equiv": n * float(times[-1]) / dt / cfg.maximum_time_us,
            "map_mean": float(Y.mean()), "map_range": [float(Y.mean(axis=0).min()), float(Y.mean(axis=0).max())]}


def grad_probe(h, args, B=None, horizon=None, whole=False) -> dict:
    """Yield-only vs yield + adjoint gradient for a batch of production trajectories; ``whole``: over the
    whole 57.56 us horizon (46,053-point grid, production budget), else over a short horizon with the
    horizon-scaled budget."""
    import numpy as np
    import torch
    from rpm_amd import adjoint, batch, core
    cfg = core.SweepConfig(lifetime_us=5.0, survival_tail_bound=1e-5)
    if whole:
        times = cfg.time_grid_us
    else:
        horizon_req = args.horizon if horizon is None else horizon
        times = np.arange(int(round(horizon_req / 0.00125)) + 1) * 0.00125
    B = args.batch if B is None else B
    thetas = list(np.linspace(10.0, 170.0, B))
    phis = [0.0] * B
    fields = [50.0] * B
    states = np.stack([core.random_phase_initial_state(h.dimensions, core.MASTER_SEED, i) for i in range(B)])
    bp = batch.BatchedLanczosPropagator.like(core.make_propagator(h, cfg, probe_horizon=float(times[-1])))
    build = lambda t, f, p: core.build_hamiltonian(t, f, phi_deg=p)  # noqa: E731
    warm = adjoint.FieldYield(h, build, bp, states[:1], times[:9], cfg.lifetime_us)
    warm.forward(thetas[:1], phis[:1], fields[:1])
    warm.backward()
    prob = adjoint.FieldYield(h, build, bp, states, times, cfg.lifetime_us)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    y = prob.forward(thetas, phis, fields)
    torch.cuda.synchronize()
    t_yield = time.perf_counter() - t0
    t0 = time.perf_counter()
    fgrad, _, rec = prob.backward()
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    return {"mode": "grad", "trajectories": B, "horizon_us": float(times[-1]), "intervals": times.size - 1,
            "parameters": ["theta", "phi", "|B|"], "seconds_yield": t_yield, "seconds_adjoint": t_grad,
            "yield_plus_grad_over_yield": (t_yield + t_grad) / t_yield, "psi0_recovery_error": rec}


def gradcheck_probe(theta: float = 60.0, phi: float = 30.0, field_uT: float = 50.0, sample: int = 0,
                    budget: float = 1e-12) -> dict:
    """Adjoint gradient of one production trajectory over the whole 57.56 us horizon (46,052 output
    intervals) against central differences (h = 1e-4 rad, 1e-3 uT) of the GPU Krylov yield, with a
    tight error budget and the center run's step schedule replayed for the six perturbed runs (one
    batched propagation). Returns the gradient, the differences and the relative errors."""
    import math

    import numpy as np
    from rpm_amd import adjoint, batch, core
    cfg = core.SweepConfig(lifetime_us=5.0, survival_tail_bound=1e-5)
    times = cfg.time_grid_us
    base = core.build_hamiltonian(90.0, 0.05)
    state = core.random_phase_initial_state(base.dimensions, core.MASTER_SEED, sample)[None, :]
    bp = batch.BatchedLanczosPropagator(base, state_error_budget=budget, initial_step_us=0.05)
    build = lambda t, f_mT, p: core.build_hamiltonian(t, f_mT, phi_deg=p)  # noqa: E731
    t0 = time.perf_counter()
    out = adjoint.yield_and_gradient(base, build, bp, state, times, cfg.lifetime_us, [theta], [phi], [field_uT])
    t_grad = time.perf_counter() - t0
    center = bp.propagate(state, times, pairs=batch.pair_matrices(base, [build(theta, 1e-3 * field_uT, phi)]))
    h_angle, h_field = 1e-4, 1e-3
    shifts = [(math.degrees(h_angle), 0.0, 0.0), (0.0, math.degrees(h_angle), 0.0), (0.0, 0.0, h_field)]
    params = []
    for dt_, dp, df in shifts:
        params += [(theta + dt_, phi + dp, field_uT + df), (theta - dt_, phi - dp, field_uT - df)]
    pairs = batch.pair_matrices(base, [build(t, 1e-3 * f, p) for t, p, f in params])
    res = bp.propagate(np.repeat(state, len(params), axis=0), times, pairs=pairs,
                       schedules=[center[0].schedule] * len(params),
                       yield_weights=core.yield_weight_table(times, cfg.lifetime_us))
    y = np.array([r.device_yields[0] for r in res])
    fd = np.array([(y[2 * k] - y[2 * k + 1]) / (2 * h) for k, h in enumerate((h_angle, h_angle, h_field))])
    g = np.asarray(out.field_gradient[0], dtype=float)
    scale = 1e-3 * np.abs(fd).max()
    rel = [float(abs(g[k] - fd[k]) / max(abs(fd[k]), scale)) for k in range(3)]
    return {"mode": "gradcheck", "theta_deg": theta, "phi_deg": phi, "field_uT": field_uT, "sample": sample,
            "horizon_us": float(times[-1]), "intervals": int(times.size - 1), "budget": budget,
            "steps": center[0].stats.accepted_steps, "seconds_yield_and_gradient": t_grad,
            "parameters": ["theta (per rad)", "phi (per rad)", "|B| (per uT)"],
            "adjoint": g.tolist(), "central_differences": fd.tolist(), "relative_error": rel,
            "psi0_recovery_error": out.recovered_initial_error}


def pulse_probe(vectors: int = 8, segments: int = 64, window_us: float = 2.0, max_move_uT: float = 1.0) -> dict:
    """One ascent step of the field-pulse optimisation over the whole 57.56 us horizon: 50 uT static
    along z plus a transverse (x) piecewise-constant pulse of `segments` amplitudes U(-20, 20) uT over
    [0, window_us]; `vectors` random-phase states. Times the forward run and the adjoint sweep
    (gradient with respect to every amplitude), then moves the amplitudes along the ensemble-mean
    gradient (largest move `max_move_uT`) and runs the forward again."""
    import numpy as np
    import torch
    from rpm_amd import adjoint, batch, core
    cfg = core.SweepConfig(lifetime_us=5.0, survival_tail_bound=1e-5)
    times = cfg.time_grid_us
    base = core.build_hamiltonian(0.0, 0.05)
    states = np.stack([core.random_phase_initial_state(base.dimensions, core.MASTER_SEED, i) for i in range(vectors)])
    bp = batch.BatchedLanczosPropagator.like(core.make_propagator(base, cfg))
    build = lambda t, f_mT, p: core.build_hamiltonian(t, f_mT, phi_deg=p)  # noqa: E731
    prob = adjoint.FieldYield(base, build, bp, states, times, cfg.lifetime_us, pulse_window_us=(0.0, window_us),
                              pulse_direction=(1.0, 0.0, 0.0))
    amps = np.random.default_rng(11).uniform(-20.0, 20.0, size=segments)
    fixed = ([0.0] * vectors, [0.0] * vectors, [50.0] * vectors)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    y0 = prob.forward(*fixed, np.tile(amps, (vectors, 1)))
    torch.cuda.synchronize()
    t_forward = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, pgrad, rec = prob.backward()
    torch.cuda.synchronize()
    t_sweep = time.perf_counter() - t0
    g = pgrad.mean(axis=0)
    gmax = float(np.abs(g).max())
    step = max_move_uT / gmax if gmax > 0.0 else 0.0
    y1 = prob.forward(*fixed, np.tile(amps + step * g, (vectors, 1)))
    predicted = step * float(np.dot(g, g))      # first-order change of the mean yield along the step
    return {"mode": "pulse", "vectors": vectors, "segments": segments, "window_us": window_us,
            "horizon_us": float(times[-1]), "seconds_forward": t_forward, "seconds_sweep": t_sweep,
            "seconds_per_vector_step": (t_forward + t_sweep) / vectors,
            "yield_plus_grad_over_yield": (t_forward + t_sweep) / t_forward,
            "mean_yield": float(y0.mean()), "mean_yield_after_step": float(y1.mean()),
            "observed_change": float(y1.mean() - y0.mean()), "predicted_change_first_order": predicted,
            "gradient_max_abs_per_uT": float(np.abs(g).max()), "psi0_recovery_error": rec}


def sobol_probe(args, cells: int = 16, q_points: int = 16) -> dict:
    """Batched Sobol reaction-time cells on the full system: tau = 1 us, Q points x 8 scrambles each,
    angles spread over 0..180 deg, one batched propagation."""
    from dataclasses import replace
    import numpy as np
    import torch
    from rpm_amd import sobol
    cfg = replace(sobol.RunConfig(), lifetimes_us=(1.0,))
    tasks = [sobol.CellTask(float(t), i, q_points, False, cfg) for i, t in enumerate(np.linspace(0.0, 180.0, cells))]
    sobol.run_cells([replace(tasks[0], q_points=1)])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = sobol.run_cells(tasks, batch_size=cells)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    horizon = sum(float(sobol.reaction_time_schedule(cfg, q_points, t.sample_id).max()) for t in tasks)
    return {"mode": "sobol", "cells": cells, "q_points": q_points, "scrambles": cfg.sobol_scrambles,
            "lifetimes_us": list(cfg.lifetimes_us), "seconds": dt, "cells_per_s": cells / dt,
            "reaction_times_per_s": cells * q_points * cfg.sobol_scrambles / dt,
            "trajectory_us_per_s": horizon / dt, "matvecs": int(sum(r.matvecs for r in res)),
            "mean_yield": float(np.mean([r.yields.mean() for r in res]))}


def spectral_cpu_sample(n_angles: int = 3) -> dict:
    """Seconds per orientation of the oracle's factorized route (one thread), on a few angles."""
    code = r"""
import json, sys, time
sys.path.insert(0, %r)
from oracle import model, spectral
fad = spectral.build_radical("FAD", model.FAD_NUCLEI)
trp = spectral.build_radical("TrpH", model.TRP_NUCLEI)
thetas = [0.0, 45.0, 90.0, 135.0, 180.0][:%d]
t0 = time.perf_counter()
for t in thetas:
    spectral.solve_orientation(fad, trp, t, 50.0, (1.0, 2.0, 5.0, 10.0, 20.0, 100.0), 8.0)
print(json.dumps({"n": len(thetas), "dt": time.perf_counter() - t0}))
""" % (ROOT, n_angles)
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1",
               NUMEXPR_NUM_THREADS="1")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, check=True)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return {"seconds_per_orientation": r["dt"] / r["n"], "cores": 1, "kind": "port",
            "sample": f"{r['n']} orientations of the oracle's factorized route (Q = 0, 50 uT, six lifetimes), 1 thread"}


def spectral_probe(args, with_cpu=True) -> dict:
    """The 199-angle factorized Fig. 1B curve (J = D = 0, Q = 0, 50 uT, 6 lifetimes) on the GPU."""
    import torch
    from rpm_amd import model, spectral
    thetas = model.theta_grid()
    spectral.solve_orientations(thetas[:4], 50.0, model.DEFAULT_LIFETIMES_US, 8.0, quadrupole_scale=0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rows = spectral.solve_orientations(thetas, 50.0, model.DEFAULT_LIFETIMES_US, 8.0, quadrupole_scale=0.0)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out = {"mode": "spectral", "orientations": len(rows), "lifetimes": len(model.DEFAULT_LIFETIMES_US),
           "seconds": dt, "orientations_per_s": len(rows) / dt,
           "max_initial_singlet_error": max(abs(r["initial_singlet"] - 1.0) for r in rows)}
    # the line-binning kernel alone: one launch = 32 TrpH orientations x 6 pairs (TrpH, d = 384, is the
    # largest block the curve bins: FAD goes as 288 + 2 x 144), grid stride of the curve
    fad, trp = spectral.reference_models(0.0)
    sa = [spectral.diagonalize_radical(b, thetas[:32], 50.0) for b in spectral.symmetry_blocks(fad)]
    sb = spectral.diagonalize_radical(trp, thetas[:32], 50.0)
    delta = (1.0 / max(model.DEFAULT_LIFETIMES_US)) / 8.0
    ha = spectral.half_width_bins(float(spectral._max_gap(sa).max()), delta)
    hb = spectral.half_width_bins(float(sb.max_gap.max()), delta)
    stride = spectral._fast_length(2 * ha + 2 * hb + 1)
    buf = torch.empty((32, len(spectral.PAIRS), stride), dtype=torch.complex128, device="cuda")
    flag = torch.zeros(1, dtype=torch.int32, device="cuda")
    spectral.binned_spectra(sb, delta, hb, stride, out=buf, overflow=flag)
    reps = 10
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(torch.cuda.current_stream())
    for _ in range(reps):
        spectral.binned_spectra(sb, delta, hb, stride, out=buf, overflow=flag)
    stop.record(torch.cuda.current_stream())
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / reps
    d = trp.dimension
    nbytes = buf.numel() * 16 + 32 * 3 * d * d * 16 + 32 * d * 8   # grid written once + T_x,y,z + eigenvalues read
    out["bin_kernel"] = {"kernel": "k_bin_tiles", "orientations": 32, "radical": "TrpH (d=384)", "grid_stride": stride,
                         "ms_per_launch": ms, "bytes_per_launch": nbytes, "achieved_GBps": nbytes / ms / 1e6,
                         "frac_of_hbm_peak": nbytes / ms / 1e6 / HBM_PEAK_GBS}
    del buf, sa, sb
    if with_cpu:
        cpu = spectral_cpu_sample()
        out["cpu_baseline"] = cpu
        out["gpu_over_cpu_core"] = cpu["seconds_per_orientation"] * len(rows) / dt
    return out


def ceiling_probe(h, B: int = 256, reps: int = 20) -> dict:
    """The streaming bounds an H.psi design over B resident vectors cannot beat, measured on this box
    with plain device copies (no stencil work), next to the fused kernel itself:
    - "copy": read B vectors and write B vectors once (the bytes of a single fused pass);
    - "two_pass_chunked": per chunk of c vectors, a copy (the first pass writing an intermediate) and
      an add-to (the second pass reading the input and the intermediate, writing the result), with
      chunks small enough (c <= 32: at most 113 MB per array) for the 256 MB Infinity Cache to hold
      the intermediate: the bound of a design that splits H into a FAD pass and a TrpH pass;
    - "kernel": the H.psi kernel at the same batch.
    Times per B vectors from HIP events on the current stream."""
    import torch
    from rpm_amd import workloads
    n = h.dimension
    x = torch.from_numpy(workloads.normal_vectors(n, B, seed=0)).cuda()
    y = torch.empty_like(x)

    def timed(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    out = {"mode": "ceiling", "batch": B, "dimension": n, "bytes_per_vector": BYTES_PER_AMPLITUDE * n}
    algo = BYTES_PER_AMPLITUDE * n * B
    ms = timed(lambda: y.copy_(x))
    out["copy"] = {"ms": ms, "GBps": algo / (ms / 1e3) / 1e9, "frac": algo / (ms / 1e3) / 1e9 / HBM_PEAK_GBS}
    chunks = {}
    for c in (8, 16, 32):
        def two_pass(c=c):
            for i in range(0, B, c):
                y[i:i + c].copy_(x[i:i + c])
                y[i:i + c].add_(x[i:i + c])
        ms = timed(two_pass)
        chunks[str(c)] = {"ms": ms, "GBps_algorithmic": algo / (ms / 1e3) / 1e9,
                          "frac": algo / (ms / 1e3) / 1e9 / HBM_PEAK_GBS}
    out["two_pass_chunked"] = chunks
    ms = timed(lambda: h.apply_device(x, y, batch=B, ld=n))
    out["kernel"] = {"ms": ms, "GBps": algo / (ms / 1e3) / 1e9, "frac": algo / (ms / 1e3) / 1e9 / HBM_PEAK_GBS}
    return out


KERNEL_NAMES = {1: "k_pair_apply", 2: "k_pair_tile", 3: "k_pair_tile3", 4: "k_pair_tile5", 9: "k_local_apply"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the batched-yield and gradient probes")
    ap.add_argument("--kernel", type=int, default=0,
                    help="0 tuned at run time (rpm_amd.autotune), 1 gather, 2 register-tiled, 3 LDS-tiled")
    ap.add_argument("--operator", default="production", choices=("production", "random"),
                    help="hpsi mode: the production operator, or the 14-nucleus random-tensor system")
    ap.add_argument("--horizon", type=float, default=0.05, help="propagate mode: simulated horizon in us")
    ap.add_argument("--plain-sample", type=int, default=8, help="batch mode: plain trajectories timed for the ratio")
    ap.add_argument("--archive-groups", default="D=J=0:0,D=J=0:1,D/J=-3:0,D/J=-3:1",
                    help="archive_all mode: comma-separated electron_interactions:quadrupole_scale groups")
    ap.add_argument("--archive-batch", type=int, default=58,
                    help="archive_all mode: trajectories per batched propagation (two batches per group of 116)")
    ap.add_argument("--eager", action="store_true",
                    help="batch mode: launch each window of basis steps eagerly instead of replaying it as a "
                         "captured HIP graph (the default)")
    ap.add_argument("--mode", default="hpsi",
                    choices=("hpsi", "copy", "propagate", "batch", "grad", "spectral", "sobol", "directional",
                             "trajectory", "archive_batch", "archive_all", "gradcheck", "pulse", "ceiling"),
                    help="copy: time a plain device copy of the same bytes (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(trajectories are independent; a multi-GPU run gives each GPU its own block of them).", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    sys.path.insert(0, ROOT)
    from rpm_amd import core, workloads

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        return 2
    h = core.build_hamiltonian(90.0, 0.05)
    assert h.uses_pair_stencil
    if args.mode == "propagate":
        print(json.dumps(propagate_probe(h, args)), flush=True)
        return 0
    if args.mode == "batch":
        print(json.dumps(batch_probe(h, args)), flush=True)
        return 0
    if args.mode == "grad":
        print(json.dumps(grad_probe(h, args)), flush=True)
        return 0
    if args.mode == "spectral":
        print(json.dumps(spectral_probe(args, with_cpu=not args.no_cpu)), flush=True)
        return 0
    if args.mode == "sobol":
        print(json.dumps(sobol_probe(args)), flush=True)
        return 0
    if args.mode == "directional":
        print(json.dumps(directional_probe(h, args)), flush=True)
        return 0
    if args.mode == "trajectory":
        print(json.dumps(trajectory_probe()), flush=True)
        return 0
    if args.mode == "archive_batch":
        print(json.dumps(archive_batch_probe()), flush=True)
        return 0
    if args.mode == "archive_all":
        groups = []
        for item in args.archive_groups.split(","):
            name, scale = item.rsplit(":", 1)
            groups.append((name, float(scale)))
        print(json.dumps(archive_all_probe(groups, args.archive_batch)), flush=True)
        return 0
    if args.mode == "gradcheck":
        print(json.dumps(gradcheck_probe()), flush=True)
        return 0
    if args.mode == "pulse":
        print(json.dumps(pulse_probe()), flush=True)
        return 0
    if args.mode == "ceiling":
        print(json.dumps(ceiling_probe(h)), flush=True)
        return 0
    if args.operator == "random":
        h = workloads.full_random_system(seed=1).hamiltonian()      # the H·ψ measurement only
    n = h.dimension
    B = args.batch
    tuned = None
    if args.kernel == 0 and args.mode == "hpsi":
        from rpm_amd import autotune
        tuned = autotune.hpsi_kernel(h, B)      # times the H·ψ kernels at this batch and keeps the fastest
    else:
        h._device_op().set_kernel(args.kernel)
    print(f"[bench] building {B} synthetic vectors of dim {n}", file=sys.stderr, flush=True)
    x = torch.from_numpy(workloads.normal_vectors(n, B, seed=0)).cuda()
    y = torch.empty_like(x)
    if args.mode == "copy":
        for _ in range(args.warmup):
            y.copy_(x)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            y.copy_(x)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        print(json.dumps({"mode": "copy", "bytes_per_step": 2 * 16 * n * B, "ms_per_step": dt * 1e3 / args.steps,
                          "GBps": 2 * 16 * n * B * args.steps / dt / 1e9}), flush=True)
        return 0
    # Clock settle (untimed, before the W warm-up steps): after an idle gap (host-side set-up) the first
    # 50-80 launches run 10-30% slow while the engine clock ramps up (0.70 -> 0.535 ms per launch at
    # B = 256), so a short warm-up would time the ramp, not the kernel.  Launch the same step until
    # 0.2 s of it has run back to back; the count and time are reported.
    settle_n, settle_ms, settle_groups = 0, 0.0, []
    while settle_ms < 200.0 and settle_n < 4000:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            h.apply_device(x, y, batch=B, ld=n)
        e1.record()
        torch.cuda.synchronize()
        settle_n += 20
        settle_groups.append(e0.elapsed_time(e1) / 20)
        settle_ms += 20 * settle_groups[-1]
    for _ in range(args.warmup):
        h.apply_device(x, y, batch=B, ld=n)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        h.apply_device(x, y, batch=B, ld=n)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = start.elapsed_time(stop)
    ms_per_step = ms_total / args.steps
    value = B * args.steps / (ms_total / 1e3)
    algo_bytes = BYTES_PER_AMPLITUDE * n * B
    achieved = algo_bytes / (ms_per_step / 1e3) / 1e9
    print(f"[bench] {args.steps} steps in {ms_total:.3f} ms (wall {wall:.3f} s)", file=sys.stderr, flush=True)
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "H·ψ/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "clock_settle": {"launches": settle_n, "ms": settle_ms,
                         "ms_per_launch_first_20": settle_groups[0], "ms_per_launch_last_20": settle_groups[-1],
                         "note": "untimed launches of the same step before the warm-up, until 0.2 s has run"},
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {"workload": ("batched H·ψ, production FAD-TrpH operator (theta=90 deg, 50 uT, "
                                "EH D/J=-3, Q on), dim 221184, complex128") if args.operator == "production" else
                               ("batched H·ψ, 14-nucleus random-tensor system (seed 1), dim 221184, complex128"),
                   "batch_vectors": B, "dimension": n,
                   "kernel": KERNEL_NAMES.get(h._device_op().kernel, f"kernel {h._device_op().kernel}")},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(B) if args.operator == "production" else None},
    }
    if tuned is not None:
        from rpm_amd.autotune import HPSI_KERNELS
        result["autotune"] = {"hpsi_kernel": tuned["name"],
                              "ms_per_launch": {HPSI_KERNELS[k]: ms for k, ms in tuned["ms_per_launch"].items()}}
    if not args.no_extras:
        # the same kernel at other batch sizes: B <= 32 fits the 256 MB Infinity Cache, B >= 128 is HBM-bound
        del y
        sweep = {}
        for bs in (1, 16, 64, 1024):
            if bs <= B:
                xs = x[:bs]
            else:   # synthetic data drawn on the device (the values do not affect the timing)
                gen = torch.Generator(device="cuda")
                gen.manual_seed(1)
                xs = torch.randn((bs, n), dtype=torch.complex128, device="cuda", generator=gen)
            ys = torch.empty_like(xs)
            for _ in range(3):
                h.apply_device(xs, ys, batch=bs, ld=n)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            reps = 20
            e0.record()
            for _ in range(reps):
                h.apply_device(xs, ys, batch=bs, ld=n)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            gbs = BYTES_PER_AMPLITUDE * n * bs / (ms / 1e3) / 1e9
            sweep[str(bs)] = {"ms_per_launch": ms, "hpsi_per_s": bs / (ms / 1e3), "achieved_GBps": gbs,
                              "frac": gbs / HBM_PEAK_GBS}
            del xs, ys
        result["hpsi_sweep"] = sweep
        if args.operator == "production" and B <= x.shape[0]:
            # the configuration-1 operator family (random, fully anisotropic tensors on all 14 nuclei, no
            # structurally zero couplings) on the same resident vectors
            hr = workloads.full_random_system(seed=1).hamiltonian()
            ys = torch.empty_like(x)
            for _ in range(3):
                hr.apply_device(x, ys, batch=B, ld=n)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            reps = 20
            e0.record()
            for _ in range(reps):
                hr.apply_device(x, ys, batch=B, ld=n)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            gbs = BYTES_PER_AMPLITUDE * n * B / (ms / 1e3) / 1e9
            result["hpsi_random_operator"] = {"batch_vectors": B, "ms_per_launch": ms, "hpsi_per_s": B / (ms / 1e3),
                                              "achieved_GBps": gbs, "frac": gbs / HBM_PEAK_GBS,
                                              "kernel": KERNEL_NAMES.get(hr._device_op().kernel, "?"),
                                              "operator": "14 nuclei, random hyperfine and quadrupole tensors, seed 1"}
            del ys, hr
        # the drop-in API with host (NumPy) buffers: host layout copies, PCIe both ways and the kernel
        # (reported beside the resident-vector value, never as it)
        host = {}
        rng = np.random.default_rng(3)
        for name, cols in (("matvec", 0), ("matmat_16", 16)):
            shape = (n,) if cols == 0 else (n, cols)
            hv = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
            call = h.matvec if cols == 0 else h.matmat
            call(hv)
            torch.cuda.synchronize()
            reps = 5
            t0 = time.perf_counter()
            for _ in range(reps):
                call(hv)
            dt = (time.perf_counter() - t0) / reps
            k = max(cols, 1)
            host[name] = {"vectors": k, "ms_per_call": dt * 1e3, "hpsi_per_s": k / dt,
                          "GBps_algorithmic": BYTES_PER_AMPLITUDE * n * k / dt / 1e9}
            del hv
        result["hpsi_host_buffers"] = host
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline()
        result["gpu_over_cpu"] = value / result["cpu_baseline"]["value"]
    if not args.no_extras:
        # the other two parts of the metric over the whole 57.56 us horizon (the reference's unit of work
        # is a whole trajectory): batched yields/s against the plain one-trajectory path, and the
        # yield + gradient time over the yield time
        h = core.build_hamiltonian(90.0, 0.05)
        del x
        torch.cuda.empty_cache()
        print("[bench] whole-trajectory probe (plain path)", file=sys.stderr, flush=True)
        result["trajectory"] = trajectory_probe()
        print("[bench] 32 whole archived trajectories batched", file=sys.stderr, flush=True)
        ab = archive_batch_probe()
        plain = result["trajectory"]["seconds"]
        ab.update({"yields_per_s": ab["trajectories"] / ab["seconds"], "plain_seconds_per_trajectory": plain,
                   "yields_per_s_plain": 1.0 / plain, "speedup_vs_plain": plain / ab["seconds_per_trajectory"],
                   "plain_cell": result["trajectory"]["cell"]})
        result["batched"] = ab
        print("[bench] whole-horizon gradient probe", file=sys.stderr, flush=True)
        result["gradient"] = grad_probe(h, args, B=8, whole=True)
        print("[bench] directional-map probe", file=sys.stderr, flush=True)
        result["directional"] = directional_probe(h, args)
        print("[bench] factorized spectral probe", file=sys.stderr, flush=True)
        result["spectral"] = spectral_probe(args, with_cpu=not args.no_cpu)
        print("[bench] Sobol reaction-time probe", file=sys.stderr, flush=True)
        result["sobol"] = sobol_probe(args)
        print("[bench] propagation probe", file=sys.stderr, flush=True)
        result["propagation"] = propagate_probe(h, args, horizon=0.1)
    if not args.no_cpu and not args.no_extras:
        print("[bench] CPU trajectory and all-core baselines", file=sys.stderr, flush=True)
        extra = cpu_extra_baselines()
        result["cpu_baseline_extra"] = extra
        if "propagation" in result:
            result["propagation"]["gpu_over_cpu_core"] = extra["prefix_seconds"] / result["propagation"]["seconds"]
        result["gpu_over_cpu_all_core"] = value / extra["all_core_hpsi_per_s"]
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
