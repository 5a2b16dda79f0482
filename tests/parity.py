"""Comparison helpers for HIP-vs-oracle parity (used by tests and smoke).

Bar: particle ids, contact keys (kind, owner, partner) and positions after
one step from an identical state are bit-exact; FP64 forces, torques,
velocities, accelerations and contact history agree within 1e-12 relative,
with the relative error of a per-particle vector taken against its own size
plus the largest size of that quantity in the system (summation order of
contact forces differs between the two implementations, and a resting
particle's net force is a small difference of large contact forces).
"""
import math

import numpy as np

from sdem_amd import native

RTOL = 1e-12


def close(a, b, rtol=RTOL, axis_scale=None):
    a, b = np.asarray(a, float), np.asarray(b, float)
    if a.shape != b.shape:
        return False, f"shape {a.shape} vs {b.shape}"
    if a.size == 0:
        return True, ""
    scale = np.abs(b).max() if axis_scale is None else axis_scale
    tol = rtol * (np.abs(b) + scale) + 1e-300
    bad = np.abs(a - b) > tol
    if bad.any():
        i = np.argwhere(bad)[0]
        return False, f"max rel err {np.max(np.abs(a - b) / (np.abs(b) + scale + 1e-300)):.3e} at {tuple(i)}: {a[tuple(i)]!r} vs {b[tuple(i)]!r}"
    return True, ""


def compare(gpu: native.Engine, ref: native.Engine, exact_positions=True, rtol=RTOL, wall_cadence=True):
    """Return a list of failure strings (empty = parity).  wall_cadence=False
    compares iterNumber only: a context started from a restart file begins
    a new near-wall cadence, as the reference's restart does."""
    errs = []
    g, r = gpu.get_state(), ref.get_state()
    if not np.array_equal(g["id"], r["id"]):
        return [f"particle ids differ: {len(g['id'])} vs {len(r['id'])}"]
    for k in ("ptype", "usr_mark"):
        if not np.array_equal(g[k], r[k]):
            errs.append(f"{k} differs")
    if exact_positions:
        if not np.array_equal(g["posr"], r["posr"]):
            d = np.abs(g["posr"] - r["posr"]).max()
            errs.append(f"positions not bit-exact (max diff {d:.3e})")
    else:
        ok, m = close(g["posr"][:, :3], r["posr"][:, :3], rtol)
        if not ok:
            errs.append("positions: " + m)
    for k in ("lin_vel", "lin_acc", "theta", "rot_vel", "rot_acc"):
        ok, m = close(g[k], r[k], rtol)
        if not ok:
            errs.append(f"{k}: {m}")
    gc, rc = gpu.get_contacts(), ref.get_contacts()
    for k in ("kind", "owner", "partner"):
        if not np.array_equal(gc[k], rc[k]):
            gs = set(zip(gc["owner"].tolist(), gc["partner"].tolist(), gc["kind"].tolist()))
            rs = set(zip(rc["owner"].tolist(), rc["partner"].tolist(), rc["kind"].tolist()))
            errs.append(f"contact {k} differs ({len(gc[k])} vs {len(rc[k])} records); "
                        f"only hip: {sorted(gs - rs)[:5]}, only oracle: {sorted(rs - gs)[:5]}")
            return errs
    ok, m = close(gc["tan_delta"], rc["tan_delta"], rtol)
    if not ok:
        errs.append("tan_delta: " + m)
    ok, m = close(gc["vrel_init"], rc["vrel_init"], rtol)
    if not ok:
        errs.append("vrel_init: " + m)
    if not np.array_equal(gpu.get_counts(), ref.get_counts()):
        errs.append(f"counts {gpu.get_counts()} vs {ref.get_counts()}")
    gi, gf, gt = gpu.get_forces()
    ri, rf, rt = ref.get_forces()
    ok, m = close(gf, rf, rtol)
    if not ok:
        errs.append("force: " + m)
    ok, m = close(gt, rt, rtol)
    if not ok:
        errs.append("torque: " + m)
    gi, ri = gpu.get_iter(), ref.get_iter()
    if (gi if wall_cadence else gi[0]) != (ri if wall_cadence else ri[0]):
        errs.append(f"iter/next wall update {gi} vs {ri}")
    return errs


def close_rows(a, b, rtol=RTOL):
    """Row by row: each row (a record's or a particle's vector) against the
    size of its own row in b, not the system's largest."""
    a, b = np.atleast_2d(np.asarray(a, float)), np.atleast_2d(np.asarray(b, float))
    if a.shape != b.shape:
        return False, f"shape {a.shape} vs {b.shape}"
    if a.size == 0:
        return True, ""
    scale = np.linalg.norm(b, axis=1, keepdims=True)
    bad = np.abs(a - b) > rtol * scale + 1e-300
    if bad.any():
        i = np.argwhere(bad)[0]
        rel = np.abs(a - b).max(axis=1) / (scale[:, 0] + 1e-300)
        return False, f"max row-relative err {rel.max():.3e} at {tuple(i)}: {a[tuple(i)]!r} vs {b[tuple(i)]!r}"
    return True, ""


def compare_per_record(gpu: native.Engine, ref: native.Engine, rtol=RTOL):
    """After one step from an identical state, the quantities that no
    summation order enters, each against its own size: every contact
    record's TanDelta (vector) and VelRel_Init, and the force and torque of
    every particle that takes part in exactly one contact record (its sum
    has one term).  Returns a list of failure strings."""
    errs = []
    gc, rc = gpu.get_contacts(), ref.get_contacts()
    for k in ("kind", "owner", "partner"):
        if not np.array_equal(gc[k], rc[k]):
            return [f"contact {k} differs"]
    ok, m = close_rows(gc["tan_delta"], rc["tan_delta"], rtol)
    if not ok:
        errs.append("tan_delta per record: " + m)
    ok, m = close_rows(gc["vrel_init"][:, None], rc["vrel_init"][:, None], rtol)
    if not ok:
        errs.append("vrel_init per record: " + m)
    # the records a particle's sum takes: those it owns (each side of a
    # periodic pair owns its own), and the PP records of a lower-id partner
    ids, cnt = np.unique(np.concatenate([rc["owner"], rc["partner"][(rc["kind"] == native.KIND_PP)]]),
                         return_counts=True)
    single = set(ids[cnt == 1].tolist())
    gi, gf, gt = gpu.get_forces()
    ri, rf, rt = ref.get_forces()
    if not np.array_equal(gi, ri):
        return errs + ["force ids differ"]
    sel = np.isin(ri, np.fromiter(single, dtype=ri.dtype, count=len(single)))
    for name, a, b in (("force", gf, rf), ("torque", gt, rt)):
        ok, m = close_rows(a[sel], b[sel], rtol)
        if not ok:
            errs.append(f"{name} of single-contact particles: " + m)
    return errs


def aging_lists_scene(seed=31, speed=3.0):
    """400 sparse spheres (no contact at the start), every one at the same
    speed in a random direction, periodic in x, y and z, dt 1e-5 s: pairs
    close head-on at up to twice the largest single move, so a candidate
    list ages over several steps and the gap bound of the contact test is
    exercised (tests/test_list_bounds.py, tests/test_gpu_parity.py)."""
    from sdem_amd import scenarios

    sc = scenarios.random_cloud(n=400, seed=seed, spacing=1.8, vel=1.0, omega=0.0, dt=1.0e-5,
                                periodic=(1, 1, 1), walls="")
    st = sc["state"]
    rng = np.random.default_rng(seed + 1)
    u = rng.normal(size=(len(st["id"]), 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    st["lin_vel"][:, :, :] = (speed * u)[:, None, :]
    st["lin_acc"][:, :, :] = 0.0
    return sc



def drift_past_fixed_scene(seed=9, drift=(2.0, 0.6, 0.0)):
    """A sparse cloud (no contact at the start) moving as one body at
    ~2 m/s through 40 fixed spheres, periodic in x, y and z, dt 1e-5 s: the
    spheres do not move relative to each other until they hit a fixed one,
    so only their moves relative to the fixed spheres age the candidate
    lists (tests/test_gpu_parity.py)."""
    from sdem_amd import scenarios

    sc = scenarios.random_cloud(n=300, seed=seed, spacing=2.0, vel=0.0, omega=0.0, dt=1.0e-5,
                                periodic=(1, 1, 1), walls="", nfixed=40)
    st, fx = sc["state"], sc["fixed"]
    L = np.array(sc["options"]["domain_max"]) - np.array(sc["options"]["domain_min"])
    p, f = st["posr"], fx["posr"]
    d = p[:, None, :3] - f[None, :, :3]
    d -= L * np.round(d / L)
    gap = np.sqrt((d * d).sum(2)) - p[:, None, 3] - f[None, :, 3]
    keep = gap.min(1) > 0.2 * f[0, 3]  # no sphere touches a fixed one at the start
    n = len(st["id"])
    for k, v in st.items():
        if isinstance(v, np.ndarray) and v.shape[:1] == (n,):
            st[k] = np.ascontiguousarray(v[keep])
    st["lin_vel"][:, :, :] += np.asarray(drift)[None, None, :]
    st["lin_acc"][:, :, :] = 0.0
    return sc


# finite plane quads for the wall-geometry tests
def _rect(origin, eu, ev):
    o, eu, ev = (np.asarray(a, float) for a in (origin, eu, ev))
    return [o, o + eu, o + eu + ev, o + ev]


def _rotation(seed):
    q = np.random.default_rng(seed).normal(size=4)
    a, b, c, d = q / np.linalg.norm(q)
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def quads():
    """(corners, both_side): an axis-aligned quad, a 30-degree incline, the
    incline two-sided, and a rectangle turned by a random rotation."""
    c30, s30 = math.cos(math.radians(30)), math.sin(math.radians(30))
    R = _rotation(7)
    out = [
        (_rect((0.010, 0.020, 0.010), (0.030, 0.0, 0.0), (0.0, 0.015, 0.0)), False),
        (_rect((0.010, 0.010, 0.010), (0.028 * c30, 0.0, 0.028 * s30), (0.0, 0.020, 0.0)), False),
        (_rect((0.010, 0.010, 0.010), (0.028 * c30, 0.0, 0.028 * s30), (0.0, 0.020, 0.0)), True),
    ]
    ctr = np.array([0.025, 0.025, 0.025])
    eu, ev = R @ np.array([0.022, 0.0, 0.0]), R @ np.array([0.0, 0.016, 0.0])
    out.append((_rect(ctr - 0.5 * eu - 0.5 * ev, eu, ev), False))
    return out
