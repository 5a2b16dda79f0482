"""Training and rollout of the learned surrogate (include/sdem_gnn.h) on
trajectories of the plain GPU DEM.

Targets: over T = nsub * dt, the DEM's deviation from the ballistic step,
    x(t+T) - (x + v T + g T^2 / 2)   and   v(t+T) - (v + g T),
in units of dx_scale = |g| T^2 / 2 and dv_scale = |g| T (a particle at
rest has the target (0, 0, 1, 0, 0, 1) * sign: it must cancel gravity).
The ballistic step itself (a network that outputs zero) is the method
without the network; rollout() measures both against the DEM.
"""
from __future__ import annotations

import time
from typing import Dict, List, Optional

import numpy as np

from . import gnn, native, scenarios

G_EARTH = 9.81


def _index(state: Dict, ids: np.ndarray) -> np.ndarray:
    """rows of an ascending-id state for the given particle ids"""
    return np.searchsorted(state["id"], ids)


def residuals(s0: Dict, s1: Dict, ids: np.ndarray, T: float, g, L, per):
    """DEM minus ballistic over T for the particles `ids` (rows in that order)."""
    i0, i1 = _index(s0, ids), _index(s1, ids)
    x0, v0 = s0["posr"][i0, :3], s0["lin_vel"][i0, 0, :]
    x1, v1 = s1["posr"][i1, :3], s1["lin_vel"][i1, 0, :]
    g = np.asarray(g, float)
    dx = x1 - (x0 + v0 * T + 0.5 * g * T * T)
    for d in range(3):
        if per[d]:
            dx[:, d] -= L[d] * np.round(dx[:, d] / L[d])
    dv = v1 - (v0 + g * T)
    return dx, dv


def make_model(hidden=128, layers=8, nsub=10, dt=5.0e-6, dmax=4.0e-3, vel_scale=0.1, seed=1,
               lib=None) -> gnn.Model:
    T = nsub * dt
    cfg = gnn.default_config(lib, hidden=hidden, layers=layers, nsub=nsub, len_scale=2.0e-3, acc_scale=G_EARTH,
                             vel_scale=vel_scale, radius=1.2 * dmax * 1.0001, cutoff_factor=1.2, floor_z=0.0,
                             dx_scale=0.5 * G_EARTH * T * T, dv_scale=G_EARTH * T)
    return gnn.Model(cfg, seed=seed, lib=lib)


# A contact-bearing window of the settling bed: the loose pack is first
# collapsed by the DEM itself for these (steps, gravity scale) stages, which
# leave it in contact with the floor and with itself (about one contact
# record per sphere, still moving), and the window starts there at the
# configuration's own gravity.
CONTACT_WINDOW = ((14_000, 2.0),)


def start_state(n: int, seed: int, settle=(), log=None, lib=None):
    """The settling bed of n spheres (configuration-2 physics, xy area scaled
    to n), advanced by the DEM through the `settle` stages ((steps, gravity
    scale) pairs) when given."""
    sc = scenarios.settling(n=n, seed=seed)
    for steps, gscale in settle:
        sc = scenarios.settle(sc, steps, gravity_scale=gscale, chunk=max(1, steps // 2), log=log, lib=lib,
                              tuning=dict(cell_factor=(0.5, 0.5, 0.5), skin=0.05))
    return sc


def train(model: gnn.Model, n: int = 50_000, seeds=(101, 102, 103), warm: int = 200, snapshots: int = 300,
          lr: float = 1e-3, lr_end: float = 2e-4, spacing: int = 10, log=print, lib=None, settle=()) -> List[float]:
    """Online training on settling boxes of n spheres (configuration-2
    physics, xy area scaled to n), each first advanced through the `settle`
    stages (start_state): after `warm` steps, every `spacing` DEM steps one
    snapshot -- graph and features of the current state, the DEM's next
    nsub steps as target -- and one optimisation step.  Returns per
    snapshot the loss divided by the loss of a zero prediction (the
    ballistic step without the network)."""
    lib = lib or native.load_hip_library()
    nsub = model.cfg.nsub
    boxes = []
    for sd in seeds:
        sc = start_state(n, sd, settle, log=None, lib=lib)
        eng = native.build_engine(lib, sc, tuning=dict(cell_factor=(0.5, 0.5, 0.5), skin=0.05, keep_forces=0))
        eng.iterate(warm)
        o = sc["options"]
        L = np.array(o["domain_max"]) - np.array(o["domain_min"])
        boxes.append((eng, o["dt"] * nsub, o["gravity"], L, o["is_periodic"]))
    losses = []
    t0 = time.perf_counter()
    for k in range(snapshots):
        eng, T, g, L, per = boxes[k % len(boxes)]
        model.graph_from_dem(eng, for_training=True)
        ids = model.node_ids()
        s0 = eng.get_state()
        eng.iterate(nsub)
        s1 = eng.get_state()
        dx, dv = residuals(s0, s1, ids, T, g, L, per)
        target = np.concatenate([dx / model.cfg.dx_scale, dv / model.cfg.dv_scale], 1).astype(np.float32)
        base = float(np.mean(target.astype(np.float64) ** 2)) + 1e-30
        frac = k / max(1, snapshots - 1)
        rate = lr * (lr_end / lr) ** frac
        losses.append(model.train_step(target, lr=rate, clip=1.0) / base)
        if spacing > nsub:
            eng.iterate(spacing - nsub)
        if log and (k % 25 == 0 or k == snapshots - 1):
            log(f"# train {k + 1}/{snapshots}: loss / no-network loss {losses[-1]:.4f} "
                f"(target rms {np.sqrt(base):.3g}, {time.perf_counter() - t0:.1f}s)")
    for b in boxes:
        b[0].close()
    return losses


def bed_metrics(state: Dict, area: float, slab: float = 0.02) -> Dict[str, float]:
    """Ensemble statistics of a bed (one density): height 2 <z> (the height
    of a uniform column with the same centroid), the solid fraction of the
    bottom slab [0, slab], and the translational kinetic energy per unit
    mass 0.5 <|v|^2> (mass-weighted)."""
    z = state["posr"][:, 2]
    r = state["posr"][:, 3]
    vol = 4.0 / 3.0 * np.pi * r ** 3
    low = z < slab
    v2 = (np.asarray(state["lin_vel"])[:, 0, :] ** 2).sum(1)
    return dict(height=float(2.0 * np.average(z, weights=vol)),
                phi_bottom=float(vol[low].sum() / (area * slab)),
                ke=float(0.5 * np.average(v2, weights=vol)))


def rollout(model: gnn.Model, n: int = 1_000_000, seed: int = 20261203, warm: int = 200, steps: int = 2000,
            dem_every: int = 500, dem_steps: int = 10, log=print, lib=None, settle=()) -> Dict:
    """The same settling bed (advanced through the `settle` stages first,
    start_state) three ways from one state after `warm` DEM steps: plain
    DEM for `steps` steps (the reference trajectory), the surrogate for steps / nsub calls
    (with `dem_steps` plain DEM steps after every `dem_every` calls, as
    check and fallback), and the ballistic step without the network.
    Returns timings and errors against the DEM."""
    lib = lib or native.load_hip_library()
    nsub = model.cfg.nsub
    sc = start_state(n, seed, settle, log=None, lib=lib)
    tune = dict(cell_factor=(0.5, 0.5, 0.5), skin=0.05, keep_forces=0)
    o = sc["options"]
    area = (o["domain_max"][0] - o["domain_min"][0]) * (o["domain_max"][1] - o["domain_min"][1])
    out = {}
    if log:
        log(f"# rollout: {n} spheres, start state ready")
    ref = native.build_engine(lib, sc, tuning=tune)
    ref.iterate(warm)
    out["contact_records_per_sphere_start"] = float(ref.get_counts().sum()) / max(1, ref.num_particles())
    t = time.perf_counter()
    ref.iterate(steps)
    out["dem_s"] = time.perf_counter() - t
    out["contact_records_per_sphere_end"] = float(ref.get_counts().sum()) / max(1, ref.num_particles())
    s_ref = ref.get_state()
    ref.close()
    if log:
        log(f"# rollout: DEM {steps} steps in {out['dem_s']:.2f}s")

    def run(m: gnn.Model, name: str):
        eng = native.build_engine(lib, sc, tuning=tune)
        eng.iterate(warm)
        calls = steps // nsub
        done = 0
        t0 = time.perf_counter()
        fell_back = None
        while done < calls:
            k = min(100, dem_every - done % dem_every, calls - done)
            try:
                for c in range(k):
                    if fell_back is None:
                        try:
                            m.advance(eng, 1)
                        except native.SdemError as err:
                            # the state collapsed under the surrogate: the
                            # plain DEM takes over for the rest of the window
                            fell_back = done + c
                            if log:
                                log(f"# rollout {name}: DEM fallback after {fell_back} calls: {err}")
                    if fell_back is not None:
                        eng.iterate(nsub)
                done += k
                if done < calls and dem_steps > 0 and done % dem_every == 0:
                    eng.iterate(dem_steps)  # the DEM takes over briefly
            except native.SdemError as err:
                # the DEM refuses the state the surrogate left (spheres
                # driven into each other blow the bed apart): no result
                out[name + "_blew_up_after_calls"] = done
                if log:
                    log(f"# rollout {name}: the DEM blew up after {done} calls: {err}")
                eng.close()
                out[name + "_fallback_after_calls"] = fell_back
                return None, time.perf_counter() - t0, 0
            if log:
                log(f"# rollout {name}: {done}/{calls} calls, {time.perf_counter() - t0:.2f}s, "
                    f"{eng.num_particles()} spheres")
        el = time.perf_counter() - t0
        extra = (calls - 1) // dem_every * dem_steps if dem_steps > 0 else 0
        st = eng.get_state()
        eng.close()
        out[name + "_fallback_after_calls"] = fell_back
        return st, el, extra

    s_sur, out["surrogate_s"], extra = run(model, "surrogate")
    out["edges"] = int(model.ne) if hasattr(model, "ne") else 0
    zero = gnn.Model(model.cfg, seed=1, lib=lib)
    zero.set_params(np.zeros(zero.num_params(), np.float32))
    s_bal, out["ballistic_s"], _ = run(zero, "ballistic")
    zero.close()
    mr = bed_metrics(s_ref, area)
    ms = bed_metrics(s_sur, area) if s_sur is not None else None
    mb = bed_metrics(s_bal, area) if s_bal is not None else None
    for k in ("height", "phi_bottom", "ke"):
        out[k + "_dem"] = mr[k]
        out[k + "_err_surrogate"] = abs(ms[k] - mr[k]) / mr[k] if ms else None
        out[k + "_err_ballistic"] = abs(mb[k] - mr[k]) / mr[k] if mb else None
    L = np.array(o["domain_max"]) - np.array(o["domain_min"])
    per = np.array(o["is_periodic"], bool)

    def rms(s):
        if s is None:
            return None, None, None
        # particles present in both (a particle thrown out of the domain is deleted)
        ids, ia, ib = np.intersect1d(s["id"], s_ref["id"], return_indices=True)
        d = s["posr"][ia, :3] - s_ref["posr"][ib, :3]
        d[:, per] -= L[per] * np.round(d[:, per] / L[per])
        low = s_ref["posr"][ib, 2] < 0.02
        return (float(np.sqrt((d ** 2).sum(1).mean())), float(np.sqrt((d[low] ** 2).sum(1).mean())),
                int(len(s_ref["id"]) - len(ids)))
    out["rms_surrogate"], out["rms_bottom_surrogate"], out["lost_surrogate"] = rms(s_sur)
    out["rms_ballistic"], out["rms_bottom_ballistic"], out["lost_ballistic"] = rms(s_bal)
    out["speedup"] = out["dem_s"] / out["surrogate_s"]
    out.update(n=n, steps=steps, nsub=nsub, extra_dem_steps=extra)
    if log:
        log(f"# rollout: {out}")
    return out


def time_calls(n: int = 1_000_000, calls: int = 10, seed: int = 20261203, warm: int = 100, lib=None) -> Dict:
    """Wall time of surrogate calls on a settling bed of n spheres (an
    untrained network: the cost does not depend on the weights).  Returns
    ms per call, the edges of the last graph and ms per DEM step of the
    same bed for comparison."""
    lib = lib or native.load_hip_library()
    model = make_model(lib=lib)
    sc = scenarios.settling(n=n, seed=seed)
    eng = native.build_engine(lib, sc, tuning=dict(cell_factor=(0.5, 0.5, 0.5), skin=0.05, keep_forces=0))
    eng.iterate(warm)
    t = time.perf_counter()
    eng.iterate(10 * model.cfg.nsub)
    dem_ms = (time.perf_counter() - t) / (10 * model.cfg.nsub) * 1e3
    model.advance(eng, 2)  # buffers sized, kernels loaded
    t = time.perf_counter()
    model.advance(eng, calls)
    call_ms = (time.perf_counter() - t) / calls * 1e3
    out = dict(n=n, calls=calls, ms_per_call=call_ms, edges=model.ne, dem_ms_per_step=dem_ms,
               dem_ms_per_call_equivalent=dem_ms * model.cfg.nsub)
    eng.close()
    model.close()
    return out


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="time surrogate calls on a settling bed")
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--calls", type=int, default=10)
    ap.add_argument("--seed", type=int, default=20261203)
    a = ap.parse_args()
    print(json.dumps(time_calls(n=a.n, calls=a.calls, seed=a.seed)), flush=True)
