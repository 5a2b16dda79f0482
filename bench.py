"""Throughput of the DEM step on one MI355X.

Workload: the plain GPU DEM at scale -- 1,000,000 polydisperse spheres
(log-normal diameters, median 2 mm, sigma 0.25, clipped to [1,4] mm, 32
types), FP64, periodic in x and y (0.2 x 0.2 m), flat bottom wall, open top,
Hertz-Mindlin with tangential history (E=5e7 Pa, nu=0.3, e=0.8, mu=0.4),
seeded random loose pack at solid fraction 0.3 falling under gravity,
dt=5e-6 s, AB2.  The state is uploaded once; the timed region runs K steps
with everything resident in HBM.

Prints one JSON line: particle-steps/s, the roofline of the dominant kernel
(the fused step kernel) and of the whole step, and a CPU baseline (the C
restatement of the reference step on every host core at once, one thread
per bounded sample of the same workload).  Under "contact_bed": the same bed
settled by the GPU path into a resting pack (about 2.3 contact records per
sphere) and stepped with the same options, with its own roofline and CPU
baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes (FP64 SoA, AB2 translation + AB2 rotation, int32 types):
#   per particle and step 428 B = 220 read + 192 written + 16 for the cell sort
#   per contact record 80 B (old key + history 40 B, new key + history 40 B)
# the fused step kernel owns all of it except the 16 B of the sort
BYTES_PER_PARTICLE_STEP = 428
BYTES_PER_PARTICLE_KERNEL = 412
BYTES_PER_CONTACT = 80
# search tuning (does not change results): cell edges relative to the
# reference's Prtcl_cs_ratio * d_max, candidate-list skin as a fraction of
# d_max; the per-contact force arrays are not written (they are computed and
# integrated, but only sdem_get_forces reads them)
CELL_FACTOR = (0.5, 0.5, 0.5)
SKIN = 0.05
# The three-size pack's hierarchical search (cells halve from level to
# level): at 0.5 its finest level has 0.625 mm cells, ~456M keys in all,
# and settled on the incline it steps in 2.70 ms; at 1.0 (1.25 mm, ~57M
# keys) in 2.16 ms (0.75: 2.28 ms)
POLY_CELL_FACTOR = (1.0, 1.0, 1.0)
# HBM bytes per k_force launch on the default 1M bed and on that bed
# settled into contact, measured with rocprofv3 --pmc FETCH_SIZE /
# WRITE_SIZE and scaled by what the same counters read for a plain 512 MiB
# copy (x2.00 and x1.00): 84,370 KiB x 2 read + 176,270 KiB written
# (profiles/k_force_settling_1M.md); 434,902 KiB x 2 + 408,823 KiB for the
# bed in contact, where the force kernel runs at three waves per SIMD and
# its spills add to the traffic (profiles/k_force_contact_bed_1M.md)
KFORCE_TRAFFIC_1M = (84_370 * 2 + 176_270) * 1024
KFORCE_TRAFFIC_CONTACT_1M = (434_902 * 2 + 408_823) * 1024
# steps per library call in a timed region (a few seconds of work at most,
# so that a long run still prints progress)
CHUNK = 20_000


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cpu-sample", type=int, default=20000, help="particles in the CPU baseline sample")
    ap.add_argument("--cpu-steps", type=int, default=0, help="0 = pick ~15 s of CPU work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-contact", action="store_true", help="leave out the settled bed in contact")
    ap.add_argument("--contact-steps", type=int, default=200, help="timed steps of the bed in contact")
    ap.add_argument("--contact-state", default=None,
                    help=".npz of the settled bed: read if it exists, else written after the settle")
    ap.add_argument("--workload", choices=("settling", "contact", "sandbox", "polydisperse", "ensemble",
                                           "surrogate"),
                    default="settling",
                    help="settling: the 1M-sphere bed (the headline line); contact: the same bed settled "
                         "into contact, alone; sandbox: the 10k-sphere sand "
                         "box; polydisperse: 4M spheres of three sizes on an incline; ensemble: 256 "
                         "independent sand boxes stepped together; surrogate: the learned surrogate "
                         "trained on DEM trajectories, then against the DEM on the 1M-sphere bed")
    ap.add_argument("--train-snapshots", type=int, default=240)
    ap.add_argument("--train-spacing", type=int, default=10,
                    help="DEM steps between training snapshots (surrogate workload)")
    ap.add_argument("--dem-every", type=int, default=500,
                    help="surrogate workload: calls between the plain DEM's 10-step takeovers")
    ap.add_argument("--surrogate-window", choices=("contact", "free"), default="contact",
                    help="surrogate workload: train and evaluate on the bed collapsed into contact by the DEM "
                         "first (contact) or on the loose pack in free fall (free)")
    ap.add_argument("--poly-settle", action="store_true",
                    help="polydisperse workload: settle the pack onto the incline first (POLY_SETTLE), then "
                         "time it in contact with the configuration's own options")
    ap.add_argument("--poly-state", default=None,
                    help=".npz of the settled polydisperse pack: read if it exists, else written after the settle")
    ap.add_argument("--boxes", type=int, default=256)
    ap.add_argument("--ensemble-settle", type=int, default=0,
                    help="ensemble workload: settle the boxes this many steps first (dt 1e-6 s)")
    ap.add_argument("--ensemble-state", default=None,
                    help=".npz of the settled boxes: read if it exists, else written after the settle")
    return ap.parse_args()


def host_threads() -> int:
    """Host threads the CPU baselines use: the cores this process may run on,
    at most 16 (one GPU's share of a GPU box's cores)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_rate(scenarios_, steps_for, target_s=15.0):
    """The oracle (C restatement of the reference step) on one thread per
    scenario, all at once: each thread steps an independent oracle context
    (the library has no global state and ctypes releases the interpreter
    lock during a call).  Returns (particle-steps/s summed over the threads,
    steps per thread)."""
    import threading

    import oracle
    from sdem_amd import native

    lib = oracle.load()
    engs = [native.build_engine(lib, sc, prefix=oracle.PREFIX) for sc in scenarios_]
    for e in engs:
        e.iterate(2)
    t0 = time.perf_counter()
    engs[0].iterate(3)
    per_step = (time.perf_counter() - t0) / 3
    steps = steps_for or max(5, int(target_s / max(per_step, 1e-6)))
    ns = [e.num_particles() for e in engs]
    barrier = threading.Barrier(len(engs) + 1)

    def run(e):
        barrier.wait()
        e.iterate(steps)

    th = [threading.Thread(target=run, args=(e,)) for e in engs]
    for t in th:
        t.start()
    barrier.wait()
    t0 = time.perf_counter()
    for t in th:
        t.join()
    wall = time.perf_counter() - t0
    for e in engs:
        e.close()
    return sum(ns) * steps / wall, steps


def cpu_baseline(args):
    """The oracle on every host core at once (one thread each, at most 16),
    each on its own seeded 20k-sphere box of the headline workload (same
    pack fraction, size distribution and physics; xy area scaled to the
    particle count)."""
    from sdem_amd import scenarios

    nth = host_threads()
    n = args.cpu_sample
    scs = [scenarios.settling(n=n, seed=args.seed + 101 * k) for k in range(nth)]
    rate, steps = oracle_rate(scs, args.cpu_steps)
    return {"value": rate, "unit": "particle-steps/s", "cores": nth, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"{nth} independent settling beds of {n} spheres (seeds {args.seed} + 101 k; same "
                      f"distribution, phi=0.3, xy area scaled), {steps} steps each after 5 warm-up steps, "
                      f"oracle, one thread per bed"}


# ---------------------------------------------------------------------------
# the settling bed in contact
# ---------------------------------------------------------------------------
# the settle: 60,000 steps at twice gravity (the fall from the loose pack),
# then 12,000 at gravity, all with the configuration's own physics and dt
CONTACT_SETTLE = ((60_000, 2.0), (12_000, 1.0))
# The 4M three-size pack settled onto its incline: steps and gravity scale
# at dt 5e-6 s (the 1M bed's time step; the configuration's own 2e-6 s
# would take 2.5 times the steps to let the 0.56 m column fall)
POLY_SETTLE = ((60_000, 2.0), (12_000, 1.0))
POLY_SETTLE_DT = 5.0e-6


def settled_poly(sc, log, state_path=None):
    """The polydisperse pack settled onto the incline by the GPU path
    (`scenarios.settle`, POLY_SETTLE at POLY_SETTLE_DT); with `state_path`
    the settled state and contact records are read from / written to that
    .npz file."""
    from sdem_amd import scenarios

    if state_path and os.path.exists(state_path):
        z = np.load(state_path)
        sc["state"] = {k[3:]: z[k] for k in z.files if k.startswith("st_")}
        sc["contacts"] = {k[3:]: z[k] for k in z.files if k.startswith("ct_")}
        return sc, 0.0
    t0 = time.perf_counter()
    for steps, g in POLY_SETTLE:
        sc = scenarios.settle(sc, steps, dt=POLY_SETTLE_DT, gravity_scale=g, chunk=4000, log=log,
                              tuning=dict(cell_factor=POLY_CELL_FACTOR, skin=SKIN))
    t_settle = time.perf_counter() - t0
    if state_path:
        np.savez(state_path, **{"st_" + k: v for k, v in sc["state"].items()},
                 **{"ct_" + k: v for k, v in sc["contacts"].items()})
    return sc, t_settle


def settled_bed(n, seed, log, state_path=None):
    """The 1M-sphere settling bed (configuration 2's physics, time step and
    walls) settled into a resting pack by the GPU path itself
    (`scenarios.settle`, CONTACT_SETTLE); with `state_path` the settled
    state and contact records are read from / written to that .npz file."""
    from sdem_amd import scenarios

    sc = scenarios.settling(n=n, seed=seed)
    if state_path and os.path.exists(state_path):
        z = np.load(state_path)
        sc["state"] = {k[3:]: z[k] for k in z.files if k.startswith("st_")}
        sc["contacts"] = {k[3:]: z[k] for k in z.files if k.startswith("ct_")}
        return sc, 0.0
    t0 = time.perf_counter()
    for steps, g in CONTACT_SETTLE:
        sc = scenarios.settle(sc, steps, gravity_scale=g, chunk=steps // 4, log=log,
                              tuning=dict(cell_factor=CELL_FACTOR, skin=SKIN))
    t_settle = time.perf_counter() - t0
    if state_path:
        np.savez(state_path, **{"st_" + k: v for k, v in sc["state"].items()},
                 **{"ct_" + k: v for k, v in sc["contacts"].items()})
    return sc, t_settle


def cpu_baseline_contact(args, log):
    """The oracle on every host core at once, each thread on its own copy of
    a 20k-sphere settling bed in contact (configuration 2's physics, settled
    on the GPU path as the 1M bed is, scaled to 20k spheres)."""
    from sdem_amd import scenarios

    n = args.cpu_sample
    sc = scenarios.settling(n=n, seed=args.seed)
    for steps, g in CONTACT_SETTLE:
        sc = scenarios.settle(sc, steps, gravity_scale=g, chunk=steps, log=log)
    nth = host_threads()
    rate, steps = oracle_rate([sc] * nth, args.cpu_steps)
    P = len(sc["contacts"]["kind"])
    return {"value": rate, "unit": "particle-steps/s", "cores": nth, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"{nth} copies of a {n}-sphere settling bed in contact ({P} contact records), "
                      f"{steps} steps each after 5 warm-up steps, oracle, one thread per copy"}


def timed_region(eng, steps, log):
    """K steps bracketed by stream synchronisation (Engine.iterate returns
    after synchronising the library's stream); one library call per CHUNK
    steps, with a progress line and a contact-count sample after each.
    Returns (wall seconds, mean contact records per step sampled)."""
    t0 = time.perf_counter()
    done, counts_sum, nsamp = 0, np.zeros(4, np.int64), 0
    while done < steps:
        k = min(CHUNK, steps - done)
        eng.iterate(k)
        counts_sum += eng.get_counts()
        nsamp += 1
        done += k
        log(f"# {done}/{steps} steps")
    return time.perf_counter() - t0, float(counts_sum.sum()) / max(1, nsamp), counts_sum // max(1, nsamp)


def kernel_roofline(eng, n, P, steps, ms_step):
    """Per-kernel durations (HIP events around every kernel on the library's
    stream over `steps` more steps, queued one by one) and the roofline of
    the force kernel and of the whole step on the algorithmic bytes."""
    eng.set_profiling(True)
    eng.iterate(steps)
    prof = eng.get_profile()
    eng.set_profiling(False)
    k_ms, k_n = prof["force"]
    k_avg_s = k_ms / max(1, k_n) / 1e3
    kernel_bytes = BYTES_PER_PARTICLE_KERNEL * n + BYTES_PER_CONTACT * P
    achieved = kernel_bytes / k_avg_s / 1e9
    step_bytes = BYTES_PER_PARTICLE_STEP * n + BYTES_PER_CONTACT * P
    step_achieved = step_bytes / (ms_step / 1e3) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "kernel": "k_force", "kernel_ms_avg": k_avg_s * 1e3, "algorithmic_bytes_per_launch": kernel_bytes,
            "whole_step": {"achieved": step_achieved, "frac": step_achieved / HBM_PEAK_GBS,
                           "algorithmic_bytes_per_step": step_bytes,
                           "definition": "B_step = 428 N + 80 P per step over the timed wall time"}}
    return roof, {k: v[0] / max(1, v[1]) for k, v in prof.items()}


def contact_bed_bench(args, log):
    """The settled 1M bed stepped with configuration 2's options: the
    contact regime of the headline workload (P ~ 2.3 contact records per
    sphere)."""
    from sdem_amd import native

    n = args.n
    sc, t_settle = settled_bed(n, args.seed, log, args.contact_state)
    lib = native.load_hip_library()
    eng = native.build_engine(lib, sc, tuning=dict(cell_factor=CELL_FACTOR, skin=SKIN, keep_forces=0))
    eng.iterate(args.warmup)
    steps = args.contact_steps
    wall, P, by_kind = timed_region(eng, steps, log)
    ms_step = wall / steps * 1e3
    roof, breakdown = kernel_roofline(eng, n, P, min(steps, 200), ms_step)
    if n == 1_000_000 and args.seed == 20261203:
        roof["traffic"] = KFORCE_TRAFFIC_CONTACT_1M
        roof["traffic_source"] = "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this bed (constant of that run)"
    out = {"value": n * steps / wall, "unit": "particle-steps/s", "steps": steps, "warmup": args.warmup,
           "ms_per_step": ms_step, "contact_records_per_step": P,
           "contact_records_by_kind": {"PP": int(by_kind[0]), "PPG": int(by_kind[1]), "PPFix": int(by_kind[2]),
                                       "PW": int(by_kind[3])},
           "settle": {"schedule_steps_gravity_scale": [list(x) for x in CONTACT_SETTLE], "seconds": t_settle},
           "roofline": roof, "step_breakdown_ms": breakdown, "particles_at_end": eng.num_particles()}
    eng.close()
    return out


def box_alive(eng, nbox):
    """Live spheres summed over the boxes."""
    n = 0
    for b in range(nbox):
        eng.select(b)
        n += eng.num_particles()
    eng.select(0)
    return n


def box_contacts(eng, nbox):
    """Contact records per sphere of every box after the last step."""
    per = []
    for b in range(nbox):
        eng.select(b)
        per.append(float(eng.get_counts().sum()) / max(1, eng.num_particles()))
    eng.select(0)
    return np.asarray(per)


def ensemble_bench(args):
    """256 closed sand boxes (log-uniform 2,000-50,000 spheres each, mixed
    linear / Hertz-Mindlin, per-box friction and restitution) stepped
    together by one context; block and tile shapes from the autotuner.
    With --ensemble-settle S the boxes are first settled into resting packs
    (S steps, every box at dt 1e-6 s and twice gravity, on the GPU path) and
    the timed steps run in contact; every box is then stepped at dt 1e-6 s,
    since at the configuration's 1e-5 s a contact of the linear boxes (kn up
    to 2e5 N/m between spheres of 1.5 mm) lasts one to six steps and their
    packs blow up under AB2."""
    from sdem_amd import native, scenarios

    t0 = time.perf_counter()
    scs = scenarios.ensemble(nbox=args.boxes, seed=args.seed)
    n_total = sum(len(sc["state"]["id"]) for sc in scs)
    print(f"# generated {args.boxes} boxes, {n_total} spheres in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    settle_s = None
    if args.ensemble_settle:
        path = args.ensemble_state
        if path and os.path.exists(path):
            # the settled states and contact records of every box, by box
            z = np.load(path)
            for b, sc in enumerate(scs):
                sc["state"] = {k.split("_", 2)[2]: z[k] for k in z.files if k.startswith(f"st_{b}_")}
                sc["contacts"] = {k.split("_", 2)[2]: z[k] for k in z.files if k.startswith(f"ct_{b}_")}
            settle_s = 0.0
        else:
            t0 = time.perf_counter()
            log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
            scs = scenarios.settle(scs, args.ensemble_settle, dt=1.0e-6, gravity_scale=2.0,
                                   chunk=max(1, args.ensemble_settle // 10), log=log)
            settle_s = time.perf_counter() - t0
            if path:
                np.savez(path, **{f"st_{b}_{k}": v for b, sc in enumerate(scs) for k, v in sc["state"].items()},
                         **{f"ct_{b}_{k}": v for b, sc in enumerate(scs) for k, v in sc["contacts"].items()})
        for sc in scs:
            sc["options"]["dt"] = 1.0e-6
    lib = native.load_hip_library()
    eng = native.build_ensemble(lib, scs, tuning=dict(keep_forces=0))
    # warm up, then let the autotuner pick the shapes on the moving packs
    # (its trials are ordinary steps of the run)
    eng.iterate(args.warmup)
    t0 = time.perf_counter()
    tuned_steps, tuned_ms = eng.autotune(steps_per_trial=20)
    print(f"# autotune: {tuned_steps} steps in {time.perf_counter() - t0:.1f}s, best {tuned_ms:.3f} ms/step, "
          f"{eng.get_tuning()}", file=sys.stderr, flush=True)
    # Spheres that leave their box are removed (the linear boxes at the
    # configuration's dt=1e-5 s blow up under AB2: a contact lasts one to
    # six steps).  The timed steps run in chunks of 250; the live spheres
    # are counted between the chunks, outside the timed steps, and the
    # value counts the particle-steps done (each chunk by the mean of its
    # first and last count).
    alive0 = box_alive(eng, args.boxes)
    alive = alive0
    wall, psteps, done = 0.0, 0.0, 0
    while done < args.steps:
        k = min(250, args.steps - done)
        t0 = time.perf_counter()
        eng.iterate(k)
        wall += time.perf_counter() - t0
        done += k
        now = box_alive(eng, args.boxes)
        psteps += 0.5 * (alive + now) * k
        alive = now
        print(f"# {done}/{args.steps} steps, {alive} spheres", file=sys.stderr, flush=True)
    per_box = box_contacts(eng, args.boxes)
    eng.set_profiling(True)
    eng.iterate(min(args.steps, 200))
    prof = eng.get_profile()
    dt_label = (f"dt=1e-6 s after {args.ensemble_settle} settling steps at dt 1e-6 s and 2 g"
                if args.ensemble_settle else "dt=1e-5 s")
    out = {
        "metric": "particle-steps/s, batched ensemble of independent boxes" +
                  (", settled packs in contact" if args.ensemble_settle else ""),
        "value": psteps / wall,
        "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded random loose packs, phi=0.3)",
        "config": {"workload": f"{args.boxes} closed sand boxes, 2,000-50,000 spheres each (log-uniform), "
                               "d U[1.5,3] mm, mu U[0.2,0.6], e U[0.5,0.95], half linear kn U[5e4,2e5], "
                               f"half Hertz-Mindlin E=5e7, {dt_label}, AB2",
                   "boxes": args.boxes, "spheres": n_total, "seed": args.seed,
                   "spheres_timed": {"first_step": alive0, "last_step": alive,
                                     "removed_before_timing": n_total - alive0,
                                     "note": "spheres that leave their box are removed; value counts the "
                                             "particle-steps done"},
                   "tuning": eng.get_tuning(), "autotune_steps": tuned_steps,
                   "settle_steps": args.ensemble_settle, "settle_s": settle_s},
        "contact_records_per_sphere_by_box": {"min": float(per_box.min()), "median": float(np.median(per_box)),
                                              "max": float(per_box.max()), "mean_weighted": None},
        "step_breakdown_ms": {k: v[0] / max(1, v[1]) for k, v in prof.items()},
    }
    tot = 0.0
    for b in range(args.boxes):
        eng.select(b)
        tot += float(eng.get_counts().sum())
    eng.select(0)
    out["contact_records_per_sphere_by_box"]["mean_weighted"] = tot / max(1, alive)
    eng.close()
    print(json.dumps(out), flush=True)


def surrogate_bench(args):
    """Learned surrogate: trained online on three 50k-sphere settling boxes
    (same physics as the 1M bed).  Speed: 20 surrogate calls (200 DEM
    steps) on the 1M bed against 200 DEM steps.  Error: a 200k-sphere bed
    of the same physics over args.steps DEM steps, by the DEM, by the
    surrogate (with 10 DEM steps after every 500 calls) and by the
    ballistic step without the network."""
    from sdem_amd import native, surrogate

    lib = native.load_hip_library()
    model = surrogate.make_model(hidden=128, layers=8, seed=1, lib=lib)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    settle = surrogate.CONTACT_WINDOW if args.surrogate_window == "contact" else ()
    t0 = time.perf_counter()
    losses = surrogate.train(model, n=50_000, seeds=(101, 102, 103), warm=200, snapshots=args.train_snapshots,
                             spacing=args.train_spacing, log=log, lib=lib, settle=settle)
    t_train = time.perf_counter() - t0
    speed = surrogate.rollout(model, n=args.n, seed=args.seed, warm=args.warmup, steps=200, log=log, lib=lib,
                              settle=settle)
    err = surrogate.rollout(model, n=200_000, seed=args.seed, warm=args.warmup, steps=args.steps, log=log, lib=lib,
                            settle=settle, dem_every=args.dem_every)
    H, L = 128, 8
    calls = speed["steps"] // speed["nsub"]
    ne = speed["edges"]
    # matrix-core work per call: 2 flop per multiply-add of every GEMM
    flop_edge = 2 * (3 * H * H + H * H + H * H) * L + 2 * (16 * H + 2 * H * H)
    flop_node = 2 * (2 * H * H + 2 * H * H) * L + 2 * (16 * H + 2 * H * H) + 2 * (H * H * 2 + H * 16)
    tflops = (flop_edge * ne + flop_node * args.n) / (speed["surrogate_s"] / calls) / 1e12
    out = {
        "metric": "surrogate speed-up over the plain GPU DEM at equal simulated time",
        "value": speed["speedup"], "unit": "x", "n_gpus": 1, "steps": 200, "warmup": args.warmup,
        "ms_per_step": speed["surrogate_s"] / 200 * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 MFMA, fp32 accumulation, fp64 positions",
        "data": "synthetic (seeded random loose packs, phi=0.3)",
        "config": {"workload": f"settling bed, {args.n} spheres; graph network 8 layers x 2x128 SiLU MLPs, "
                               f"10 DEM steps per call, trained on {args.train_snapshots} snapshots of "
                               "3 x 50k-sphere boxes", "train_s": t_train,
                   "window": (f"contact: the packs first collapsed by the DEM ({list(settle)} steps x gravity "
                              "scale), then stepped at gravity" if settle else "free: the loose pack falling"),
                   "contact_records_per_sphere": {"speed_bed_start": speed["contact_records_per_sphere_start"],
                                                  "error_bed_start": err["contact_records_per_sphere_start"],
                                                  "error_bed_end": err["contact_records_per_sphere_end"]},
                   "train_loss_ratio_last": float(np.mean(losses[-20:])),
                   "train_snapshot_spacing": args.train_spacing},
        "timing_s_1M_200_steps": {"dem": speed["dem_s"], "surrogate": speed["surrogate_s"],
                                  "ballistic": speed["ballistic_s"]},
        "errors_vs_dem_200k": {k: v for k, v in err.items()
                               if "err" in k or "rms" in k or "lost" in k or k.endswith("_dem") or "fallback" in k or "blew" in k},
        "error_window_steps": args.steps, "dem_every_calls": args.dem_every,
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": tflops / 2500.0, "traffic": None, "edges": ne},
    }
    model.close()
    print(json.dumps(out), flush=True)


WORKLOADS = {
    "settling": ("particle-steps/s at 1M polydisperse spheres",
                 "settling bed: 1M log-normal spheres, periodic x/y 0.2x0.2 m, bottom wall, "
                 "Hertz-Mindlin with history, dt=5e-6 s, AB2"),
    "sandbox": ("particle-steps/s, 10k-sphere sand box",
                "sand box: 10k spheres d U[2,3] mm (11 types), 6 plane walls 0.05x0.05x0.2 m, linear "
                "kn=1e5 N/m, kt=0.8 kn, e=0.9, mu=0.3, jittered lattice, dt=1e-5 s, AB2"),
    "polydisperse": ("particle-steps/s, 4M spheres of three size classes",
                     "wide polydispersity: 4M spheres of 1/2.5/5 mm (0.85/0.12/0.03), closed 1:1:2 box with a "
                     "30-degree plane, Hertz-Mindlin mu=0.5, random pack phi=0.2, hierarchical search, "
                     "dt=2e-6 s, AB2"),
}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the multi-GPU decomposition is design only)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        ensemble_bench(args)
        return
    if args.workload == "surrogate":
        surrogate_bench(args)
        return
    from sdem_amd import native, scenarios

    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if args.workload == "contact":
        out = contact_bed_bench(args, log)
        out.update({"metric": "particle-steps/s at 1M polydisperse spheres, settled bed in contact", "n_gpus": 1,
                    "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                    "data": "synthetic (seeded loose pack settled by the GPU path)",
                    "config": {"workload": WORKLOADS["settling"][1] + ", settled into contact", "n": args.n,
                               "seed": args.seed}})
        print(json.dumps(out), flush=True)
        return
    cpu = None
    if not args.no_cpu and args.workload == "settling":
        log("# cpu baseline (before any GPU work)")
        cpu = cpu_baseline(args)
    t0 = time.perf_counter()
    if args.workload == "sandbox":
        n = 10_000
        sc = scenarios.sandbox(n=n, seed=args.seed)
    elif args.workload == "polydisperse":
        n = 4_000_000
        sc = scenarios.polydisperse(n=n, seed=args.seed)
        if args.poly_settle:
            sc, poly_settle_s = settled_poly(sc, log, args.poly_state)
    else:
        n = args.n
        sc = scenarios.settling(n=n, seed=args.seed)
    log(f"# generated {n} spheres in {time.perf_counter() - t0:.1f}s")
    lib = native.load_hip_library()
    cell_factor = POLY_CELL_FACTOR if args.workload == "polydisperse" else CELL_FACTOR
    eng = native.build_engine(lib, sc, tuning=dict(cell_factor=cell_factor, skin=SKIN, keep_forces=0))
    eng.iterate(args.warmup)
    log(f"# warm-up {args.warmup} steps done")
    n_start = eng.num_particles()
    # Timed region: steps queued as HIP graphs, bracketed by device
    # synchronisation on both sides (timed_region).
    wall, P, _ = timed_region(eng, args.steps, log)
    ms_step = wall / args.steps * 1e3
    n_end = eng.num_particles()
    # spheres that leave through an open side are removed: the particle-steps
    # done, by the mean of the live counts before and after the timed steps
    value = 0.5 * (n_start + n_end) * args.steps / wall
    roof, breakdown = kernel_roofline(eng, n, P, min(args.steps, 1000), ms_step)
    n_now = n_end
    eng.close()
    metric, workload = WORKLOADS[args.workload]
    if args.workload == "polydisperse" and args.poly_settle:
        metric += ", settled onto the incline"
        workload += (f", settled onto the incline first ({', '.join(f'{k} steps at {g:g} g' for k, g in POLY_SETTLE)}"
                     f", dt {POLY_SETTLE_DT:g} s, {poly_settle_s:.0f} s), then timed with the configuration's options")
    if args.workload == "settling" and n != 1_000_000:
        side = 0.2 * (n / 1.0e6) ** 0.5  # scenarios.settling scales the xy area with n
        metric = f"particle-steps/s, settling bed of {n} spheres"
        workload = workload.replace("1M", f"{n}").replace("0.2x0.2 m", f"{side:.4g}x{side:.4g} m")
    headline = args.workload == "settling" and n == 1_000_000 and args.seed == 20261203
    if headline:
        # HBM bytes of one k_force launch on this bed, from rocprofv3 --pmc
        # FETCH_SIZE and WRITE_SIZE scaled by a plain copy's counts: a
        # constant of that measurement, not measured in this run
        roof["traffic"] = KFORCE_TRAFFIC_1M
        roof["traffic_source"] = "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this bed (constant of that run)"
    out = {
        "metric": metric,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded initial packing)",
        "config": {"workload": workload, "n": n, "seed": args.seed, "contact_records_per_step": P,
                   "tuning": {"cell_factor": list(cell_factor), "skin_d_max": SKIN},
                   "particles_timed": {"first_step": n_start, "last_step": n_now}, "particles_at_end": n_now},
        "roofline": roof,
        "step_breakdown_ms": breakdown,
        "whole_step_hbm_frac": roof["whole_step"]["frac"],
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
    if args.workload == "settling" and not args.no_contact:
        log("# the same bed settled into contact")
        out["contact_bed"] = contact_bed_bench(args, log)
        if not args.no_cpu:
            log("# cpu baseline in contact")
            out["contact_bed"]["cpu_baseline"] = cpu_baseline_contact(args, log)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
