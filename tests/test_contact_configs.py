"""The benchmark configurations in contact: HIP path vs the oracle.

Each configuration (shrunk so that the oracle finishes in seconds) is first
settled into a resting pack on the HIP path with its own particle types,
materials, walls and contact model (`scenarios.settle`; a smaller time step
or a stronger pull where only that shortens the fall), so that its contact
physics -- the 32-type Hertz-Mindlin table with the row-type rule, the
linear and LSD models against six finite walls, the three-size hierarchical
search against the 30-degree incline and its free upper edge, per-box
materials of an ensemble -- meets the oracle.  From the settled state and
contact records both sides step on with the configuration's own options;
the contact set is compared exactly, the state and history within the
stated tolerance, and the counts of each contact kind are asserted so that
a contact-free window cannot pass.
"""
import functools

import numpy as np
import pytest

import oracle
from parity import compare
from sdem_amd import native, scenarios

pytestmark = pytest.mark.gpu

WALL_BASE = 10000000  # base_wall_id of the scenarios; wall k (1-based) has id base + k
INCLINE_ID = WALL_BASE + 7


def _log(m):
    print(m, flush=True)


def settled_sandbox(model):
    sc = scenarios.sandbox(n=1500, box=(0.04, 0.04, 0.08), model=model)
    # dt 1e-5 resolves a contact of kn = 1e5 N/m in ~3 steps, too few for
    # the pack to come to rest under AB2: settle with 1e-6
    return scenarios.settle(sc, 250_000, dt=1.0e-6, chunk=50_000, log=_log)


def settled_settling():
    sc = scenarios.settling(n=20000)
    s = scenarios.settle(sc, 40_000, gravity_scale=4.0, chunk=20_000, log=_log)
    s = scenarios.settle(s, 6_000, gravity_scale=2.0, chunk=6_000, log=_log)
    return scenarios.settle(s, 12_000, chunk=6_000, log=_log)


def settled_polydisperse():
    sc = scenarios.polydisperse(n=3000)
    s = scenarios.settle(sc, 40_000, gravity_scale=4.0, chunk=20_000, log=_log)
    return scenarios.settle(s, 20_000, chunk=10_000, log=_log)


SETTLED = {
    "sandbox_linear": functools.partial(settled_sandbox, "linear"),
    "sandbox_lsd": functools.partial(settled_sandbox, "lsd"),
    "settling": settled_settling,
    "polydisperse_hierarchical": settled_polydisperse,
}


@functools.lru_cache(maxsize=None)
def settled(name):
    _log(f"settling {name} on the HIP path")
    return SETTLED[name]()


def incline_branch_counts(sc, eng):
    """PW records on the incline, split by the branch of PW_isInContact that
    found them (Prtcl_Geometry.f90:84-104): the projection of the centre on
    the plane inside the quad (plane test), or outside it (edge and vertex
    tests, Line_point_check)."""
    w = sc["walls"][-1]
    p1, p2, p4 = (np.asarray(w[k], float) for k in ("p1", "p2", "p4"))
    n = np.cross(p2 - p1, np.asarray(w["p3"], float) - p1)
    n /= np.linalg.norm(n)
    ct, st = eng.get_contacts(), eng.get_state()
    sel = (ct["kind"] == native.KIND_PW) & (ct["partner"] == INCLINE_ID)
    pos = st["posr"][np.searchsorted(st["id"], ct["owner"][sel]), :3]
    cp = pos - ((pos - p1) @ n)[:, None] * n
    u, v = p2 - p1, p4 - p1
    su, sv = (cp - p1) @ u / (u @ u), (cp - p1) @ v / (v @ v)
    inside = (su >= 0) & (su <= 1) & (sv >= 0) & (sv <= 1)
    return int(inside.sum()), int((~inside).sum())


@pytest.mark.parametrize("name", sorted(SETTLED))
def test_settled_config_parity(name, hip_lib, oracle_lib):
    sc = settled(name)
    n = len(sc["state"]["id"])
    _log(f"config {name} in contact: {n} spheres, {len(sc['contacts']['kind'])} records at the start; 1 + 29 steps")
    gpu = native.build_engine(hip_lib, sc)
    ref = native.build_engine(oracle_lib, sc, prefix=oracle.PREFIX)
    for k, exact, rtol in ((1, True, 1e-12), (9, False, 1e-11), (20, False, 1e-10)):
        gpu.iterate(k)
        ref.iterate(k)
        errs = compare(gpu, ref, exact_positions=exact, rtol=rtol)
        assert not errs, (k, errs)
        pp, ppg, ppfix, pw = ref.get_counts().tolist()
        _log(f"  after {k}: counts PP {pp} PPG {ppg} PPFix {ppfix} PW {pw}")
        assert pp > n // 2, (pp, n)
        assert pw > 0
        if name == "settling":
            assert ppg > 0
        if name == "polydisperse_hierarchical":
            plane, edge = incline_branch_counts(sc, ref)
            _log(f"  incline records: {plane} by the plane test, {edge} by the edge/vertex tests")
            assert plane > 0 and edge > 0, (plane, edge)


def test_settled_ensemble_boxes_match_oracle(hip_lib, oracle_lib):
    """Four boxes of an ensemble settled together (per-box kn, friction and
    restitution; boxes 0 and 2 linear, 1 and 3 Hertz-Mindlin), then stepped
    together with their own options: every box equals a context of its own
    bit for bit, and a linear and a Hertz-Mindlin box equal the oracle in
    contact.  The linear boxes' own time step (1e-5 s) resolves a contact of
    kn in U[5e4, 2e5] N/m between 1.5 and 3 mm spheres in one to six steps,
    so under AB2 their packs blow up within tens of steps (the oracle's
    arithmetic does the same) and the growth amplifies the summation-order
    roundings: the linear box is compared over its first 10 steps, the
    Hertz-Mindlin box over 30."""
    _log("ensemble of 4 boxes in contact: settle, then 30 steps")
    scs = scenarios.settle(scenarios.ensemble(nbox=4, seed=9, nmin=400, nmax=1500), 250_000, dt=1.0e-6,
                           chunk=50_000, log=_log)
    ens = native.build_ensemble(hip_lib, scs)
    refs = [native.build_engine(oracle_lib, scs[b], prefix=oracle.PREFIX) for b in (0, 1)]
    for k, rtol in ((1, 1e-12), (9, 1e-11), (20, 1e-10)):
        ens.iterate(k)
        for b, ref in enumerate(refs):
            if b == 0 and k == 20:
                continue  # the linear box after 10 steps: see above
            ref.iterate(k)
            ens.select(b)
            errs = compare(ens, ref, exact_positions=(k == 1), rtol=rtol)
            assert not errs, (b, k, errs)
            n = len(scs[b]["state"]["id"])
            pp, _, _, pw = ref.get_counts().tolist()
            _log(f"  box {b} ({scs[b]['meta']['model']}) after {k}: {n} spheres, PP {pp} PW {pw}")
            assert pp > n // 2 and pw > 0, (b, pp, pw)
    for b in range(4):
        ens.select(b)
        one = native.build_engine(hip_lib, scs[b])
        one.iterate(30)
        sa, sb = ens.get_state(), one.get_state()
        for key in sa:
            assert np.array_equal(sa[key], sb[key]), (b, key)
        ca, cb = ens.get_contacts(), one.get_contacts()
        for key in ca:
            assert np.array_equal(ca[key], cb[key]), (b, key)
        one.close()
