"""HIP path vs the CPU oracle on identical seeded states (needs the GPU)."""
import numpy as np
import pytest

import oracle
from parity import aging_lists_scene, compare, compare_per_record, drift_past_fixed_scene, quads
from sdem_amd import native, scenarios

pytestmark = pytest.mark.gpu

CASES = {
    "box_nlin_ab2": dict(),
    "box_lsd_ab2": dict(cf_type=native.CFT_LSD),
    "box_linear_ab2": dict(cf_type=native.CFT_LINEAR),
    "periodic_fixed_linear": dict(cf_type=native.CFT_LINEAR, periodic=(1, 1, 0), walls="zZ", nfixed=30),
    "periodic_x_walls_yz": dict(periodic=(1, 0, 0), walls="yYzZ"),
    "periodic_xy_bottom_infinite": dict(periodic=(1, 1, 0), walls="z", infinite_walls=True),
    "periodic_xyz": dict(periodic=(1, 1, 1), walls=""),
    "ab3_ab3": dict(pi=native.PIM_AB3, pri=native.PIM_AB3),
    "fe_fe_variable_torque": dict(pi=native.PIM_FE, pri=native.PIM_FE,
                                  ct_model=native.CTM_VARIABLE_TORQUE),
    "fixed_particles": dict(nfixed=25),
    "hierarchical": dict(diameters=(0.6e-3, 1.2e-3, 3.0e-3), cs_method=native.CSM_NBS_MUNJIZA_HRCHL,
                         nfixed=10),
    # no walls, fast particles: the near-wall cadence still follows max|v|
    "no_walls_fast": dict(periodic=(1, 1, 1), walls="", vel=3.0, dt=1.0e-5),
}


def engines(sc, hip_lib, oracle_lib):
    gpu = native.build_engine(hip_lib, sc)
    ref = native.build_engine(oracle_lib, sc, prefix=oracle.PREFIX)
    return gpu, ref


@pytest.mark.parametrize("name", sorted(CASES))
def test_one_step_parity(name, hip_lib, oracle_lib):
    print(f"case {name}: one step", flush=True)
    sc = scenarios.random_cloud(n=400, seed=11, **CASES[name])
    gpu, ref = engines(sc, hip_lib, oracle_lib)
    gpu.iterate(1)
    ref.iterate(1)
    errs = compare(gpu, ref)
    assert not errs, errs
    errs = compare_per_record(gpu, ref)
    assert not errs, errs
    assert gpu.get_counts().sum() > 50


@pytest.mark.parametrize("name", ["box_nlin_ab2", "periodic_xy_bottom_infinite", "hierarchical"])
def test_many_steps_parity(name, hip_lib, oracle_lib):
    print(f"case {name}: 40 steps", flush=True)
    sc = scenarios.random_cloud(n=300, seed=5, **CASES[name])
    gpu, ref = engines(sc, hip_lib, oracle_lib)
    for k in range(4):
        gpu.iterate(10)
        ref.iterate(10)
        errs = compare(gpu, ref, exact_positions=False, rtol=1e-10)
        assert not errs, (k, errs)


@pytest.mark.parametrize("skin", [0.0, 0.02, 0.3])
def test_candidate_list_skin_does_not_change_results(skin, hip_lib, oracle_lib):
    """Lists rebuilt every step (skin 0), rarely (0.02) or seldom (0.3 of
    d_max): the contact set and the state stay the oracle's over 30 steps."""
    print(f"skin {skin}", flush=True)
    sc = scenarios.random_cloud(n=300, seed=21, periodic=(1, 1, 0), walls="z", infinite_walls=True)
    gpu = native.build_engine(hip_lib, sc, tuning=dict(skin=skin, cell_factor=(0.5, 0.75, 1.0)))
    ref = native.build_engine(oracle_lib, sc, prefix=oracle.PREFIX)
    for k in range(3):
        gpu.iterate(10)
        ref.iterate(10)
        errs = compare(gpu, ref, exact_positions=False, rtol=1e-10)
        assert not errs, (k, errs)


def test_contacts_found_while_lists_age(hip_lib, oracle_lib):
    """Sparse, fast spheres and a skin of 0.3 d_max: the candidate lists
    stay in use for several steps, and a particle skips its contact test
    while its closest listed neighbour cannot have closed the gap (twice the
    box's summed largest moves since the lists were built).  Pairs close in
    head-on from every gap (tests/test_list_bounds.py checks on the oracle's
    trajectory that this scene separates twice the summed moves from once);
    the contact set and the state are the oracle's after every step."""
    print("aging lists: 400 sparse fast spheres, skin 0.3, 80 single steps", flush=True)
    sc = aging_lists_scene()
    gpu = native.build_engine(hip_lib, sc, tuning=dict(skin=0.3))
    ref = native.build_engine(oracle_lib, sc, prefix=oracle.PREFIX)
    first = None
    for step in range(80):
        gpu.iterate(1)
        ref.iterate(1)
        errs = compare(gpu, ref, exact_positions=False, rtol=1e-10)
        assert not errs, (step, errs)
        if first is None and gpu.get_counts().sum() > 0:
            first = step
    assert first is not None and first > 0, first  # no contact at the start; contacts form later
    assert gpu.get_counts().sum() > 5


def test_drift_past_fixed_spheres(hip_lib, oracle_lib):
    """A sparse cloud moving as one body at ~2 m/s through 40 fixed spheres
    (parity.drift_past_fixed_scene), skin 0.05 d_max: the lists age with the
    spheres' moves, which are moves relative to the fixed spheres, and are
    rebuilt every few steps; every contact with a fixed sphere is found (the
    first after 14 steps, some 0.3 mm of drift: three skins).  The contact
    set is the oracle's after every step; the hits rebound at up to ~4 m/s,
    which amplifies the rounding differences of the two implementations'
    force sums (other orders of summation), so the state is compared within
    1e-8."""
    print("drift past fixed spheres: 262 spheres + 40 fixed, skin 0.05, 80 single steps", flush=True)
    sc = drift_past_fixed_scene()
    gpu = native.build_engine(hip_lib, sc, tuning=dict(skin=0.05))
    ref = native.build_engine(oracle_lib, sc, prefix=oracle.PREFIX)
    fixed_records = 0
    for step in range(80):
        gpu.iterate(1)
        ref.iterate(1)
        gc, rc = gpu.get_contacts(), ref.get_contacts()
        for key in ("kind", "owner", "partner"):
            assert np.array_equal(gc[key], rc[key]), (step, key)
        errs = compare(gpu, ref, exact_positions=False, rtol=1e-8)
        assert not errs, (step, errs)
        fixed_records += int(gpu.get_counts()[2])
    assert fixed_records > 50, fixed_records
    assert gpu.get_rebuilds() > 5, gpu.get_rebuilds()


def test_capacity_overflow_is_replayed(hip_lib, oracle_lib):
    """Tiny history and list capacities force overflows; the library grows
    them and replays the failed step from its intact input, so the result is
    the oracle's."""
    print("capacity replay: hist_slots=1, list_slots=2", flush=True)
    sc = scenarios.random_cloud(n=300, seed=4)
    gpu = native.build_engine(hip_lib, sc, capacity=dict(hist_slots=1, list_slots=2))
    ref = native.build_engine(oracle_lib, sc, prefix=oracle.PREFIX)
    bytes0 = gpu.device_bytes()
    # 5 + 7 steps: the halted remainder of a call differs in length
    for k in (5, 7):
        gpu.iterate(k)
        ref.iterate(k)
        assert gpu.get_iter() == ref.get_iter()
    assert gpu.device_bytes() > bytes0, "no capacity was regrown"
    errs = compare(gpu, ref, exact_positions=False, rtol=1e-10)
    assert not errs, errs


def test_history_overflow_on_a_fast_step(hip_lib, oracle_lib):
    """A third sphere closes on a pair in contact while the candidate lists
    stay valid (skin 0.4 d, moves of 5e-8 m per step), so the middle
    sphere's second contact exceeds its one history slot on a step queued
    without the rebuild kernels, whose control runs in the contact test:
    the step halts, the slots grow and it is replayed from its input.  The
    state is the oracle's every 10 steps."""
    print("history overflow on a fast step: 3 spheres, hist_slots=1", flush=True)
    d = 2.0e-3
    sc = scenarios.two_spheres(d=d, v=0.01, gap=-0.5e-6, dt=1.0e-6)
    st = sc["state"]
    x2 = st["posr"][1, 0]
    posr = np.vstack([st["posr"], [[x2 + d + 3.0e-6, st["posr"][1, 1], st["posr"][1, 2], d / 2]]])
    lv1 = np.array([[0.01, 0, 0], [-0.01, 0, 0], [-0.05, 0, 0]])
    sc["state"] = scenarios.zero_state([1, 2, 3], [1, 1, 1], posr, lin_vel1=lv1)
    sc["options"]["num_prtcl"] = 3
    gpu = native.build_engine(hip_lib, sc, tuning=dict(skin=0.4), capacity=dict(hist_slots=1))
    ref = native.build_engine(oracle_lib, sc, prefix=oracle.PREFIX)
    bytes0 = gpu.device_bytes()
    for _ in range(10):
        gpu.iterate(10)
        ref.iterate(10)
        errs = compare(gpu, ref, exact_positions=False, rtol=1e-10)
        assert not errs, (gpu.get_iter(), errs)
    ct = ref.get_contacts()
    assert np.sum(ct["kind"] != native.KIND_PW) >= 2, "the third sphere never touched"
    assert gpu.device_bytes() > bytes0, "no history slot was added"


def test_wall_cadence_without_acceleration(hip_lib, oracle_lib):
    """Two spheres apart and at rest, no gravity: the cadence formula is
    0/0 and the near-wall lists are refreshed at every step, as the
    oracle does (tests/test_oracle_pinning.py pins why)."""
    print("near-wall cadence with nothing accelerating: 2 spheres, 12 steps", flush=True)
    sc = scenarios.two_spheres(v=0.0, gap=1.0e-3)
    gpu, ref = engines(sc, hip_lib, oracle_lib)
    for k in (1, 3, 8):
        gpu.iterate(k)
        ref.iterate(k)
        assert gpu.get_iter() == ref.get_iter(), (gpu.get_iter(), ref.get_iter())
        errs = compare(gpu, ref)
        assert not errs, errs


def test_two_runs_are_bitwise_identical(hip_lib):
    print("determinism: n=2000, 25 steps twice", flush=True)
    sc = scenarios.random_cloud(n=2000, seed=8, periodic=(1, 0, 0), walls="yYzZ")
    out = []
    for _ in range(2):
        g = native.build_engine(hip_lib, sc)
        g.iterate(25)
        st, ct = g.get_state(), g.get_contacts()
        out.append((st["posr"].tobytes(), st["lin_vel"].tobytes(), ct["tan_delta"].tobytes()))
        g.close()
    assert out[0] == out[1]


@pytest.mark.parametrize("skin", [0.0, 0.02, 0.3])
def test_speculative_steps_match_full_steps(hip_lib, skin):
    """A single-box context queues its steps without the sort and list
    kernels and replays a step on the full path when its box must rebuild
    (skin 0: every step; 0.02: often; 0.3: seldom).  The bits are those of
    steps queued with every kernel (profiling on), over calls of several
    lengths."""
    print(f"speculative steps, skin {skin}", flush=True)
    sc = scenarios.random_cloud(n=600, seed=7, periodic=(1, 1, 0), walls="z", vel=2.0, dt=1.0e-5)
    a = native.build_engine(hip_lib, sc, tuning=dict(skin=skin))
    b = native.build_engine(hip_lib, sc, tuning=dict(skin=skin))
    b.set_profiling(True)
    for k in (1, 7, 30, 2, 64):
        a.iterate(k)
        b.iterate(k)
        sa, sb = a.get_state(), b.get_state()
        for key in sa:
            assert np.array_equal(sa[key], sb[key]), (k, key)
        ca, cb = a.get_contacts(), b.get_contacts()
        for key in ca:
            assert np.array_equal(ca[key], cb[key]), (k, key)
        assert a.get_iter() == b.get_iter()
        assert np.array_equal(a.get_counts(), b.get_counts())


def test_tuning_does_not_change_bits(hip_lib):
    """Contact forces are summed in ascending partner id, so the cell shape,
    the candidate-list skin and the block sizes leave every bit of the result
    unchanged."""
    print("tuning independence: 3 settings x 20 steps", flush=True)
    sc = scenarios.random_cloud(n=1500, seed=13, periodic=(1, 1, 0), walls="z", infinite_walls=True)
    out = []
    for (cf, skin), bs in zip([((1, 1, 1), 0.1), ((0.5, 0.75, 0.6), 0.0), ((0.3, 0.5, 1.0), 0.3)],
                              [(256, 256), (128, 64), (64, 128)]):
        g = native.build_engine(hip_lib, sc, tuning=dict(cell_factor=cf, skin=skin))
        g.set_block_sizes(*bs)
        g.iterate(20)
        st, ct = g.get_state(), g.get_contacts()
        out.append((st["posr"].tobytes(), st["lin_vel"].tobytes(), st["rot_vel"].tobytes(),
                    ct["tan_delta"].tobytes()))
        g.close()
    assert out[0] == out[1] == out[2]


def test_force_builds_give_the_same_bits(hip_lib):
    """The force kernel's builds for two and three waves per SIMD (the host
    picks one per batch by contact density) run the same arithmetic: on a
    pack with ~3 contact records per sphere each gives every bit of the
    other's result, and the automatic choice gives them too."""
    print("force builds: 2 vs 3 waves per SIMD vs chosen, 40 steps in contact", flush=True)
    sc = scenarios.random_cloud(n=2000, seed=8, periodic=(1, 0, 0), walls="yYzZ")
    out = []
    for build in (2, 3, 0):
        g = native.build_engine(hip_lib, sc)
        g.set_force_build(build)
        g.iterate(40)
        st, ct = g.get_state(), g.get_contacts()
        out.append(([st[k].tobytes() for k in sorted(st)], [ct[k].tobytes() for k in sorted(ct)],
                    g.get_counts().tolist()))
        g.close()
    assert out[0] == out[1] == out[2]
    assert sum(out[0][2]) > 2000, out[0][2]
    with pytest.raises(native.SdemError):
        g = native.build_engine(hip_lib, sc)
        g.set_force_build(4)


# the benchmark configurations, shrunk so the oracle finishes in seconds
CONFIGS = {
    "sandbox_linear": lambda: scenarios.sandbox(n=1500, box=(0.04, 0.04, 0.08)),
    "sandbox_lsd": lambda: scenarios.sandbox(n=1500, box=(0.04, 0.04, 0.08), model="lsd"),
    "settling": lambda: scenarios.settling(n=20000),
    "polydisperse_hierarchical": lambda: scenarios.polydisperse(n=3000),
}


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_config_parity(name, hip_lib, oracle_lib):
    print(f"config {name}: 30 steps", flush=True)
    sc = CONFIGS[name]()
    gpu, ref = engines(sc, hip_lib, oracle_lib)
    for k in (1, 29):
        gpu.iterate(k)
        ref.iterate(k)
        errs = compare(gpu, ref, exact_positions=False, rtol=1e-11)
        assert not errs, errs


def quad_edge_scene(seed=23, n=400):
    """Spheres around the edges and corners of three finite quads (a
    30-degree incline, the same incline two-sided, a rectangle turned by a
    random rotation), on the front side within a radius of the plane, with
    random velocities and spins: many of them touch a quad through the edge
    and corner tests of PW_isInContact (Line_point_check,
    Prtcl_Geometry.f90:226-261)."""
    rng = np.random.default_rng(seed)
    qs = quads()[1:]
    pos, rad = [], []
    for i in range(n):
        corners, both = qs[i % len(qs)]
        p1, p2, p3, p4 = corners
        nrm = np.cross(p2 - p1, p3 - p1)
        nrm /= np.linalg.norm(nrm)
        eu, ev = p2 - p1, p4 - p1
        r = rng.choice([0.6e-3, 0.9e-3])
        side = rng.integers(4)
        s = rng.uniform(-0.05, 1.05)
        ou = rng.uniform(-1.2, 1.2) * r / np.linalg.norm(eu)  # off edge 2 / 4, in units of the side
        ov = rng.uniform(-1.2, 1.2) * r / np.linalg.norm(ev)  # off edge 1 / 3
        u, v = [(s, ov), (1 + ou, s), (s, 1 + ov), (ou, s)][side]
        h = rng.uniform(0.05, 0.95) * r * (rng.choice([-1, 1]) if both else 1)
        pos.append(p1 + u * eu + v * ev + h * nrm)
        rad.append(r)
    posr = np.concatenate([np.array(pos), np.array(rad)[:, None]], 1)
    types = np.where(posr[:, 3] < 0.7e-3, 1, 2)
    st = scenarios.zero_state(np.arange(1, n + 1), types, posr)
    st["lin_vel"][:, 0, :] = rng.uniform(-0.2, 0.2, (n, 3))
    st["rot_vel"][:, 0, :] = rng.uniform(-100, 100, (n, 3))
    walls = [dict(p1=tuple(c[0]), p2=tuple(c[1]), p3=tuple(c[2]), p4=tuple(c[3]), user_id=k + 1, wall_type=1,
                  both_side=both, is_infinite=False, trans_vel=(0.0, 0.0, 0.0)) for k, (c, both) in enumerate(qs)]
    return dict(options=scenarios.base_options(dt=1.0e-6, domain_max=(0.05, 0.05, 0.05), num_prtcl=n),
                ptypes=dict(diameter=[1.2e-3, 1.8e-3], density=[2500.0] * 2, youngs=[5.0e7] * 2, poisson=[0.3] * 2),
                pp=(0.4, 0.3, 0.05, 0.8, 0.8), wtypes=dict(youngs=[5.0e7], poisson=[0.3]),
                pw=(0.45, 0.35, 0.05, 0.7, 0.7), walls=walls, state=st), qs


def test_finite_quad_edges_and_corners(hip_lib, oracle_lib):
    """Wall contacts through the edge and corner tests of finite quads
    (tilted, two-sided, rotated): the contact set and the state equal the
    oracle's after 1 and 10 steps, and most wall records are off the quad's
    face (their centre projects outside the quad)."""
    print("finite quads: 400 spheres around edges and corners, 1 + 9 steps", flush=True)
    sc, qs = quad_edge_scene()
    gpu, ref = engines(sc, hip_lib, oracle_lib)
    for k, exact, rtol in ((1, True, 1e-12), (9, False, 1e-11)):
        gpu.iterate(k)
        ref.iterate(k)
        errs = compare(gpu, ref, exact_positions=exact, rtol=rtol)
        assert not errs, (k, errs)
    ct, st = ref.get_contacts(), ref.get_state()
    off_face = 0
    for (kind, owner, partner) in zip(ct["kind"], ct["owner"], ct["partner"]):
        if kind != native.KIND_PW:
            continue
        p1, p2, p3, p4 = qs[partner - 10000001][0]
        p = st["posr"][np.searchsorted(st["id"], owner), :3]
        eu, ev = p2 - p1, p4 - p1
        u, v = (p - p1) @ eu / (eu @ eu), (p - p1) @ ev / (ev @ ev)
        off_face += not (0 <= u <= 1 and 0 <= v <= 1)
    print(f"  wall records {int(np.sum(ct['kind'] == native.KIND_PW))}, off the face {off_face}", flush=True)
    assert off_face > 50
