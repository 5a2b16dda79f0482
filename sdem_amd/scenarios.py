"""Seeded initial conditions for the DEM workloads.

The reference builds its initial state with a clock-seeded lattice and
shuffle (Prtcl_Variables.f90:191-307) that cannot be reproduced; these
generators replace it with seeded ones and return a *scenario*: options,
particle and wall properties, plane walls and the particle state, in the
shapes `sdem_amd.native.build_engine` pushes through the C ABI.

Workloads (sizes are arguments; defaults are the full configurations):
  * ``sandbox``  : closed box closed by 6 plane walls, jittered cubic lattice,
                   diameters uniform in [2,3] mm discretised to 11 types.
  * ``settling`` : periodic in x and y, flat bottom wall, open top, log-normal
                   diameters (median 2 mm, sigma 0.25, clipped to [1,4] mm) in
                   32 log-spaced types, random loose pack at solid fraction 0.3.
  * ``polydisperse`` : closed box with an extra 30-degree plane, three size
                   classes 1 / 2.5 / 5 mm at number fractions 0.85/0.12/0.03.
  * ``two_spheres``, ``random_cloud`` : small cases for tests.
"""
from __future__ import annotations

import math
from typing import Dict, List, Optional, Sequence

import numpy as np

from .native import (CFT_LINEAR, CFT_LSD, CFT_NLIN, CSM_NBS_MUNJIZA, CSM_NBS_MUNJIZA_HRCHL, CTM_CONSTANT_TORQUE,
                     PIM_AB2, level_size)

DEFAULT_SEED = 20261203


# --------------------------------------------------------------------------
# helpers
# --------------------------------------------------------------------------
def box_walls(lo: Sequence[float], hi: Sequence[float], infinite: bool = False,
              faces: str = "xXyYzZ", wall_type: int = 1) -> List[dict]:
    """Plane quads on the faces of an axis-aligned box, normals pointing in.
    Corner order gives n = (p2-p1)x(p3-p1) (G_add_PlaneWall,
    Prtcl_Geometry.f90:492)."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    quads = {
        "z": [(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0)],
        "Z": [(x0, y0, z1), (x0, y1, z1), (x1, y1, z1), (x1, y0, z1)],
        "x": [(x0, y0, z0), (x0, y1, z0), (x0, y1, z1), (x0, y0, z1)],
        "X": [(x1, y0, z0), (x1, y0, z1), (x1, y1, z1), (x1, y1, z0)],
        "y": [(x0, y0, z0), (x0, y0, z1), (x1, y0, z1), (x1, y0, z0)],
        "Y": [(x0, y1, z0), (x1, y1, z0), (x1, y1, z1), (x0, y1, z1)],
    }
    out = []
    for k, f in enumerate(faces):
        p = quads[f]
        out.append(dict(p1=p[0], p2=p[1], p3=p[2], p4=p[3], user_id=k + 1, wall_type=wall_type,
                        both_side=False, is_infinite=infinite, trans_vel=(0.0, 0.0, 0.0)))
    return out


def zero_state(ids, ptype, posr, pi_method=PIM_AB2, pri_method=PIM_AB2, lin_vel1=None) -> Dict:
    n = len(ids)
    ts, rs = level_size(pi_method), level_size(pri_method)
    lv = np.zeros((n, ts, 3))
    if lin_vel1 is not None:
        lv[:, 0, :] = lin_vel1
    return dict(id=np.asarray(ids, np.int32), ptype=np.asarray(ptype, np.int32),
                usr_mark=np.ones(n, np.int32), posr=np.asarray(posr, np.float64),
                lin_vel=lv, lin_acc=np.zeros((n, ts, 3)), theta=np.zeros((n, 3)),
                rot_vel=np.zeros((n, rs, 3)), rot_acc=np.zeros((n, rs, 3)))


def discretise(d: np.ndarray, edges: np.ndarray, centers: np.ndarray):
    """Map diameters to type ids (1-based) by bins."""
    t = np.clip(np.searchsorted(edges, d, side="right") - 1, 0, len(centers) - 1)
    return t.astype(np.int32) + 1, centers[t]


def rsa_pack(radii: np.ndarray, lo, hi, seed: int, periodic=(0, 0, 0), max_attempts: int = 100000):
    """Random sequential addition of non-overlapping spheres in a box,
    largest first (sdem_util_random_pack in libsdem_hip.so, host code)."""
    import ctypes as C

    from .native import load_hip_library

    lib = load_hip_library()
    f = lib.sdem_util_random_pack
    DP = C.POINTER(C.c_double)
    f.argtypes = [C.c_int64, DP, DP, DP, C.POINTER(C.c_int32), C.c_uint64, C.c_int32, DP]
    f.restype = C.c_int
    r = np.ascontiguousarray(radii, np.float64)
    lo_, hi_ = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    per = np.asarray(periodic, np.int32)
    pos = np.zeros((len(r), 3))
    rc = f(len(r), r.ctypes.data_as(DP), lo_.ctypes.data_as(DP), hi_.ctypes.data_as(DP),
           per.ctypes.data_as(C.POINTER(C.c_int32)), int(seed), int(max_attempts), pos.ctypes.data_as(DP))
    if rc != 0:
        raise RuntimeError(f"random packing failed (code {rc})")
    return pos


# --------------------------------------------------------------------------
# small cases
# --------------------------------------------------------------------------
def base_options(**kw) -> dict:
    o = dict(dt=1.0e-5, gravity=(0.0, 0.0, -9.81), domain_min=(0.0, 0.0, 0.0),
             domain_max=(0.05, 0.05, 0.05), prtcl_cs_ratio=1.0, wall_neighbor_ratio=2.0,
             is_periodic=(0, 0, 0), cs_method=CSM_NBS_MUNJIZA, cf_type=CFT_NLIN,
             ct_model=CTM_CONSTANT_TORQUE, pi_method=PIM_AB2, pri_method=PIM_AB2, cs_numlvls=0,
             cntct_list_size=6, base_wall_id=10000000, wall_max_update_iter=100, num_prtcl=100,
             ifirst=1)
    o.update(kw)
    return o


def two_spheres(d=2.0e-3, v=0.5, gap=0.0, cf_type=CFT_NLIN, rest=0.8, youngs=5.0e7, poisson=0.3,
                density=2500.0, fric=0.4, roll=0.0, offset=(0.0, 0.0), kn=1.0e5, kt_over_kn=0.8,
                **okw) -> Dict:
    """Two equal spheres approaching head-on along x with speeds +v, -v
    (kn, kt_over_kn: the constant-stiffness linear model's parameters)."""
    L = 0.05
    c = np.array([L / 2, L / 2, L / 2])
    s = d + gap
    posr = np.array([[c[0] - s / 2, c[1], c[2], d / 2],
                     [c[0] + s / 2 + offset[0], c[1] + offset[1], c[2], d / 2]])
    st = zero_state([1, 2], [1, 1], posr, okw.get("pi_method", PIM_AB2), okw.get("pri_method", PIM_AB2),
                    lin_vel1=np.array([[v, 0, 0], [-v, 0, 0]]))
    opts = base_options(cf_type=cf_type, gravity=(0.0, 0.0, 0.0), num_prtcl=2,
                        domain_max=(L, L, L), **okw)
    sc = dict(options=opts,
              ptypes=dict(diameter=[d], density=[density], youngs=[youngs], poisson=[poisson]),
              pp=(fric, fric, roll, rest, rest),
              wtypes=dict(youngs=[youngs], poisson=[poisson]), pw=(fric, fric, roll, rest, rest),
              walls=[], state=st)
    if cf_type == CFT_LINEAR:
        sc["linear"] = (kn, kt_over_kn)
    return sc


def random_cloud(n=300, seed=DEFAULT_SEED, box=None, diameters=(1.0e-3, 1.5e-3, 2.0e-3),
                 spacing=0.9, periodic=(0, 0, 0), walls="xXyYzZ", infinite_walls=False,
                 vel=0.3, omega=50.0, cf_type=CFT_NLIN, ct_model=CTM_CONSTANT_TORQUE, pi=PIM_AB2,
                 pri=PIM_AB2, cs_method=CSM_NBS_MUNJIZA, roll=0.05, nfixed=0, **okw) -> Dict:
    """A dense random cloud with many initial contacts: spheres on a jittered
    cubic lattice whose spacing is `spacing` x the mean diameter, random
    velocities, spins and AB histories, optional periodicity, walls and fixed
    spheres.  Used by parity tests."""
    rng = np.random.default_rng(seed)
    diam = np.asarray(diameters, float)
    t = rng.integers(0, len(diam), n)
    r = diam[t] / 2
    a = spacing * float(np.mean(diam[t]))
    if box is None:
        k = int(math.ceil(n ** (1 / 3)))
        box = (a * k, a * k, a * math.ceil(n / (k * k)))
    box = np.asarray(box, float)
    nx = np.maximum(1, np.floor(box / a).astype(int))
    while np.prod(nx) < n:
        nx[np.argmax(box / nx)] += 1
    grid = np.stack(np.meshgrid(*[(np.arange(k) + 0.5) / k for k in nx], indexing="ij"), -1).reshape(-1, 3)
    sel = rng.permutation(len(grid))[:n]
    pos = grid[sel] * box
    pos += (rng.random((n, 3)) - 0.5) * (box / nx) * 0.3
    for d in range(3):
        if not periodic[d]:
            pos[:, d] = np.clip(pos[:, d], 0.6 * r, box[d] - 0.6 * r)
        else:
            pos[:, d] = np.mod(pos[:, d], box[d])
            pos[pos[:, d] >= box[d], d] = 0.0
    ids = rng.permutation(np.arange(1, n + 1)).astype(np.int32)
    posr = np.concatenate([pos, r[:, None]], 1)
    st = zero_state(ids, t + 1, posr, pi, pri)
    st["lin_vel"][:, 0, :] = (rng.random((n, 3)) - 0.5) * 2 * vel
    st["rot_vel"][:, 0, :] = (rng.random((n, 3)) - 0.5) * 2 * omega
    ts, rs = level_size(pi), level_size(pri)
    for l in range(1, ts):
        st["lin_vel"][:, l, :] = (rng.random((n, 3)) - 0.5) * vel
    for l in range(ts):
        st["lin_acc"][:, l, :] = (rng.random((n, 3)) - 0.5) * 20.0
    for l in range(1, rs):
        st["rot_vel"][:, l, :] = (rng.random((n, 3)) - 0.5) * omega
    for l in range(rs):
        st["rot_acc"][:, l, :] = (rng.random((n, 3)) - 0.5) * 1000.0
    if ts > 1:  # the reference's own restarts satisfy linAcc(1)=linAcc(2)
        st["lin_acc"][:, 0, :] = st["lin_acc"][:, 1, :]
    if rs > 1:
        st["rot_acc"][:, 0, :] = st["rot_acc"][:, 1, :]
    st["theta"][:] = rng.random((n, 3))
    wl = box_walls((0, 0, 0), box, infinite=infinite_walls, faces=walls) if walls else []
    okw.setdefault("dt", 2.0e-6)
    opts = base_options(domain_max=tuple(box), is_periodic=tuple(periodic), num_prtcl=n, cf_type=cf_type,
                        ct_model=ct_model, pi_method=pi, pri_method=pri, cs_method=cs_method, **okw)
    sc = dict(options=opts,
              ptypes=dict(diameter=list(diam), density=[2500.0] * len(diam),
                          youngs=[5.0e7] * len(diam), poisson=[0.3] * len(diam)),
              pp=(0.4, 0.35, roll, 0.8, 0.8),
              wtypes=dict(youngs=[5.0e7], poisson=[0.3]), pw=(0.45, 0.4, roll, 0.7, 0.7),
              walls=wl, state=st)
    if cf_type == CFT_LINEAR:
        sc["linear"] = (5.0e4, 0.8)
    if nfixed:
        fr = diam.max() / 2
        fpos = rng.random((nfixed, 3)) * (box - 2 * fr) + fr
        sc["fixed"] = dict(id=np.arange(n + 1, n + nfixed + 1, dtype=np.int32),
                           ptype=np.full(nfixed, len(diam), np.int32),
                           posr=np.concatenate([fpos, np.full((nfixed, 1), fr)], 1))
    return sc


# --------------------------------------------------------------------------
# full workloads
# --------------------------------------------------------------------------
def settling(n=1_000_000, seed=DEFAULT_SEED, phi=0.3, ntypes=32, area=None, cs_method=CSM_NBS_MUNJIZA,
             zmax=None) -> Dict:
    """Periodic x,y bed with a flat bottom wall and an open top; Hertz-Mindlin
    with E=5e7 Pa, nu=0.3, e=0.8, mu=0.4; dt=5e-6 s; AB2 translation and
    rotation; log-normal diameters (median 2 mm, sigma 0.25, clipped to
    [1,4] mm) in `ntypes` log-spaced types; density 2500 kg/m^3 (assumed).
    The xy area scales with n so that n=1e6 gives 0.2 x 0.2 m."""
    rng = np.random.default_rng(seed)
    edges = np.geomspace(1.0e-3, 4.0e-3, ntypes + 1)
    centers = np.sqrt(edges[:-1] * edges[1:])
    d = np.exp(np.log(2.0e-3) + 0.25 * rng.standard_normal(n))
    d = np.clip(d, 1.0e-3, 4.0e-3)
    t, dq = discretise(d, edges, centers)
    side = area if area is not None else 0.2 * math.sqrt(n / 1.0e6)
    vol_solid = float(np.sum(np.pi / 6 * dq ** 3))
    height = vol_solid / phi / (side * side)
    pos = rsa_pack(dq / 2, (0.0, 0.0, 0.0), (side, side, height), seed, periodic=(1, 1, 0))
    order = np.lexsort((pos[:, 0], pos[:, 1], pos[:, 2]))  # any fixed order
    pos, t, dq = pos[order], t[order], dq[order]
    ids = np.arange(1, n + 1, dtype=np.int32)
    posr = np.concatenate([pos, dq[:, None] / 2], 1)
    st = zero_state(ids, t, posr)
    zmax = zmax if zmax is not None else max(1.3 * height, height + 0.02)
    opts = base_options(dt=5.0e-6, domain_max=(side, side, zmax), is_periodic=(1, 1, 0), num_prtcl=n,
                        cf_type=CFT_NLIN, cs_method=cs_method)
    walls = box_walls((0, 0, 0), (side, side, zmax), infinite=True, faces="z")
    return dict(options=opts,
                ptypes=dict(diameter=list(centers), density=[2500.0] * ntypes, youngs=[5.0e7] * ntypes,
                            poisson=[0.3] * ntypes),
                pp=(0.4, 0.4, 0.0, 0.8, 0.8),
                wtypes=dict(youngs=[5.0e7], poisson=[0.3]), pw=(0.4, 0.4, 0.0, 0.8, 0.8),
                walls=walls, state=st, meta=dict(height=height, phi=phi, side=side))


def sandbox(n=10_000, seed=DEFAULT_SEED, box=(0.05, 0.05, 0.2), ntypes=11, jitter=0.1,
            model="linear") -> Dict:
    """Closed box of 6 finite plane walls, diameters U[2,3] mm in `ntypes`
    types, density 2500, e=0.9, mu=0.3, dt=1e-5 s, AB2.  model="linear": the
    constant-stiffness spring-dashpot kn=1e5 N/m, kt=0.8 kn; model="lsd": the
    reference's velocity-dependent linear model (LSD) with E=1e7 Pa and
    nu=1/3, so that its kt/kn is also 0.8.  Particles sit on a jittered cubic
    lattice filling the box (the upper half cannot hold 10,000
    non-overlapping spheres of up to 3 mm)."""
    rng = np.random.default_rng(seed)
    edges = np.linspace(2.0e-3, 3.0e-3, ntypes + 1)
    centers = np.linspace(2.0e-3, 3.0e-3, ntypes)
    d = rng.uniform(2.0e-3, 3.0e-3, n)
    t = np.clip(np.rint((d - 2.0e-3) / 1.0e-3 * (ntypes - 1)).astype(int), 0, ntypes - 1)
    dq = centers[t]
    box = np.asarray(box, float)
    a = (np.prod(box) / n) ** (1 / 3)
    nx = np.maximum(1, np.floor(box / a).astype(int))
    while np.prod(nx) < n:
        nx[np.argmax(box / nx)] += 1
    h = box / nx
    grid = np.stack(np.meshgrid(*[(np.arange(k) + 0.5) for k in nx], indexing="ij"), -1).reshape(-1, 3)
    grid = grid[np.lexsort((grid[:, 0], grid[:, 1], grid[:, 2]))][:n] * h
    pos = grid + (rng.random((n, 3)) - 0.5) * 2 * jitter * dq[:, None]
    for k in range(3):
        pos[:, k] = np.clip(pos[:, k], dq / 2, box[k] - dq / 2)
    ids = np.arange(1, n + 1, dtype=np.int32)
    posr = np.concatenate([pos, dq[:, None] / 2], 1)
    st = zero_state(ids, t + 1, posr)
    if model not in ("linear", "lsd"):
        raise ValueError("model must be 'linear' or 'lsd'")
    opts = base_options(dt=1.0e-5, domain_max=tuple(box), num_prtcl=n,
                        cf_type=CFT_LINEAR if model == "linear" else CFT_LSD)
    sc = dict(options=opts,
              ptypes=dict(diameter=list(centers), density=[2500.0] * ntypes, youngs=[1.0e7] * ntypes,
                          poisson=[1.0 / 3.0] * ntypes),
              pp=(0.3, 0.3, 0.0, 0.9, 0.9),
              wtypes=dict(youngs=[1.0e7], poisson=[1.0 / 3.0]), pw=(0.3, 0.3, 0.0, 0.9, 0.9),
              walls=box_walls((0, 0, 0), box), state=st)
    if model == "linear":
        sc["linear"] = (1.0e5, 0.8)
    return sc


def incline_quad(L: float, W: float, angle_deg: float = 30.0):
    """The inclined plane of the three-size box: a quad rising from the
    floor edge x = 0 along +x at `angle_deg` over half the box length L,
    across the full width W, facing up.  Its upper edge (x = L/2,
    z = L/2 tan angle) is free, so spheres resting on that edge are
    touched through the edge test of PW_isInContact
    (Prtcl_Geometry.f90:89-104), not the plane test."""
    xr = 0.5 * L
    zr = xr * math.tan(math.radians(angle_deg))
    return dict(p1=(0.0, 0.0, 0.0), p2=(xr, 0.0, zr), p3=(xr, W, zr), p4=(0.0, W, 0.0), user_id=7,
                wall_type=1, both_side=False, is_infinite=False, trans_vel=(0.0, 0.0, 0.0))


def polydisperse(n=4_000_000, seed=DEFAULT_SEED, phi=0.2, box=None) -> Dict:
    """Three size classes 1 / 2.5 / 5 mm (number fractions 0.85/0.12/0.03) in a
    closed box with a 30-degree plane; Hertz-Mindlin, mu=0.5, dt=2e-6 s,
    hierarchical search.  The box keeps the 1:1:2 aspect of 0.3x0.3x0.6 m,
    sized so the pack fits at fraction `phi` in its lower 80 %, and is
    raised by the incline's height: the pack starts above the incline (the
    plane rises at 30 degrees along +x over the first half of the floor,
    `incline_quad`) and collapses onto it."""
    rng = np.random.default_rng(seed)
    centers = np.array([1.0e-3, 2.5e-3, 5.0e-3])
    t = rng.choice(3, size=n, p=[0.85, 0.12, 0.03])
    dq = centers[t]
    vol_solid = float(np.sum(np.pi / 6 * dq ** 3))
    tan30 = math.tan(math.radians(30.0))
    if box is None:
        s = (vol_solid / phi / 0.8 / 2.0) ** (1 / 3)
        box = (s, s, 2 * s + 0.5 * s * tan30)
    box = np.asarray(box, float)
    L = box[0]
    ramp = incline_quad(L, box[1])
    zr = ramp["p2"][2]
    height = vol_solid / phi / (box[0] * box[1])
    z0 = zr + centers.max()
    pos = rsa_pack(dq / 2, (0, 0, z0), (box[0], box[1], min(z0 + height, box[2] * 0.999)), seed)
    order = np.lexsort((pos[:, 0], pos[:, 1], pos[:, 2]))
    pos, t, dq = pos[order], t[order], dq[order]
    ids = np.arange(1, n + 1, dtype=np.int32)
    st = zero_state(ids, t + 1, np.concatenate([pos, dq[:, None] / 2], 1))
    walls = box_walls((0, 0, 0), box)
    walls.append(ramp)
    opts = base_options(dt=2.0e-6, domain_max=tuple(box), num_prtcl=n, cf_type=CFT_NLIN,
                        cs_method=CSM_NBS_MUNJIZA_HRCHL)
    return dict(options=opts,
                ptypes=dict(diameter=list(centers), density=[2500.0] * 3, youngs=[5.0e7] * 3,
                            poisson=[0.3] * 3),
                pp=(0.5, 0.5, 0.0, 0.8, 0.8), wtypes=dict(youngs=[5.0e7], poisson=[0.3]),
                pw=(0.5, 0.5, 0.0, 0.8, 0.8), walls=walls, state=st)


# --------------------------------------------------------------------------
# batched ensemble of independent sand boxes
# --------------------------------------------------------------------------
def ensemble_box(n: int, seed: int, model: str = "hm", fric: float = 0.4, rest: float = 0.8,
                 kn: float = 1.0e5, phi: float = 0.3, ntypes: int = 8, dt: float = 1.0e-5) -> Dict:
    """One closed sand box of an ensemble: n spheres with diameters U[1.5,3]
    mm in `ntypes` types, density 2500, a random loose pack at solid fraction
    `phi` filling a 1:1:2 box closed by 6 planes, gravity -z, AB2, dt 1e-5 s.
    model "hm": Hertz-Mindlin, E=5e7 Pa, nu=0.3; model "linear": constant kn,
    kt = 0.8 kn.  Friction and restitution as given (particles and walls)."""
    rng = np.random.default_rng(seed)
    edges = np.linspace(1.5e-3, 3.0e-3, ntypes + 1)
    centers = 0.5 * (edges[:-1] + edges[1:])
    d = rng.uniform(1.5e-3, 3.0e-3, n)
    t, dq = discretise(d, edges, centers)
    vol_solid = float(np.sum(np.pi / 6 * dq ** 3))
    s = (vol_solid / phi / 2.0) ** (1 / 3)
    s = max(s, 6 * 3.0e-3)
    box = (s, s, 2 * s)
    height = min(vol_solid / phi / (s * s), box[2] * 0.999)
    pos = rsa_pack(dq / 2, (0.0, 0.0, 0.0), (s, s, height), seed)
    order = np.lexsort((pos[:, 0], pos[:, 1], pos[:, 2]))
    pos, t, dq = pos[order], t[order], dq[order]
    ids = np.arange(1, n + 1, dtype=np.int32)
    st = zero_state(ids, t, np.concatenate([pos, dq[:, None] / 2], 1))
    cf = CFT_LINEAR if model == "linear" else CFT_NLIN
    opts = base_options(dt=dt, domain_max=box, num_prtcl=n, cf_type=cf)
    sc = dict(options=opts,
              ptypes=dict(diameter=list(centers), density=[2500.0] * ntypes, youngs=[5.0e7] * ntypes,
                          poisson=[0.3] * ntypes),
              pp=(fric, fric, 0.0, rest, rest),
              wtypes=dict(youngs=[5.0e7], poisson=[0.3]), pw=(fric, fric, 0.0, rest, rest),
              walls=box_walls((0, 0, 0), box), state=st, meta=dict(model=model, box=box))
    if cf == CFT_LINEAR:
        sc["linear"] = (kn, 0.8)
    return sc


def ensemble(nbox: int = 256, seed: int = DEFAULT_SEED, nmin: int = 2000, nmax: int = 50000) -> List[Dict]:
    """The batched-ensemble workload: `nbox` boxes with log-uniform particle
    counts in [nmin, nmax], friction U[0.2,0.6], restitution U[0.5,0.95];
    even boxes use the linear model with kn U[5e4,2e5] N/m, odd boxes
    Hertz-Mindlin with E=5e7 Pa."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(nbox):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        fric = float(rng.uniform(0.2, 0.6))
        rest = float(rng.uniform(0.5, 0.95))
        kn = float(rng.uniform(5.0e4, 2.0e5))
        out.append(ensemble_box(n, seed + 7919 * (b + 1), model="linear" if b % 2 == 0 else "hm",
                                fric=fric, rest=rest, kn=kn))
    return out


# --------------------------------------------------------------------------
# contact-bearing starts
# --------------------------------------------------------------------------
def _with_start(sc: Dict, state: Dict, contacts: Dict) -> Dict:
    out = dict(sc)
    out["state"] = state
    out["contacts"] = contacts
    out["options"] = dict(sc["options"])
    return out


def settle(sc, steps: int, lib=None, dt: Optional[float] = None, gravity_scale: float = 1.0,
           chunk: int = 5000, log=None, tuning: Optional[dict] = None):
    """A contact-bearing start for a scenario (or a list of scenarios, stepped
    together as one ensemble): the pack is advanced `steps` steps on the
    product's GPU path with the scenario's own particle types, materials,
    walls and contact model, and the state and contact records (tangential
    history included) at the end replace the scenario's start.  `dt`
    (default: the scenario's own) and `gravity_scale` apply to these
    steps only -- a configuration whose own time step resolves a contact in
    two or three steps cannot come to rest under Adams-Bashforth, and a
    stronger pull only shortens the fall.  The returned scenario keeps its
    own options, so it is stepped with its own dt and gravity from there."""
    from . import native

    lib = lib or native.load_hip_library()
    many = isinstance(sc, (list, tuple))
    scs = list(sc) if many else [sc]
    prep = []
    for s in scs:
        p = dict(s)
        o = dict(s["options"])
        if dt is not None:
            o["dt"] = dt
        o["gravity"] = tuple(gravity_scale * np.asarray(o.get("gravity", (0.0, 0.0, -9.81)), float))
        p["options"] = o
        prep.append(p)
    tune = dict(keep_forces=0)
    tune.update(tuning or {})
    eng = native.build_ensemble(lib, prep, tuning=tune) if many else native.build_engine(lib, prep[0], tuning=tune)
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        eng.iterate(k)
        done += k
        if log:
            log(f"settle: {done}/{steps} steps, contacts {eng.get_counts().tolist()}")
    out = []
    for b, s in enumerate(scs):
        if many:
            eng.select(b)
        st, ct = eng.get_state(), eng.get_contacts()
        out.append(_with_start(s, st, ct))
    eng.close()
    return out if many else out[0]
