"""Pin the oracle's wall geometry, near-wall cadence, pair table, friction,
rolling resistance, torque split and FE rotation against what the
mathematics fixes (closed forms, an ODE integrated by scipy, independent
geometry, brute force), so that a plausible mistake in the restatement --
a wrong edge order, the true pair mass where the reference takes the row
type's, the static friction coefficient where it caps with the kinetic one,
a dropped factor of the cadence -- fails one of them.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

import oracle
from parity import quads
from sdem_amd import native, scenarios

RHO, E, NU = 2500.0, 5.0e7, 0.3


def make(lib, sc):
    return native.build_engine(lib, sc, prefix=oracle.PREFIX)


def one_sphere_on_floor(d=2.0e-3, L=0.2, H=0.01, dt=1.0e-6, fric=(0.4, 0.4), roll=0.0, rest=0.3,
                        ct_model=native.CTM_CONSTANT_TORQUE, pi=native.PIM_AB3, pri=native.PIM_AB3):
    """One Hertz-Mindlin sphere resting on a floor wall at the static overlap
    delta = (m g / K_h)^(2/3), K_h = 4/3 E* sqrt(R) (flat wall: R* = R,
    Prtcl_Property.f90:211-216), so that the normal force starts at m g."""
    r = d / 2
    m = RHO * math.pi / 6 * d ** 3
    estar = 1.0 / (2 * (1 - NU * NU) / E)
    kh = 4.0 / 3.0 * estar * math.sqrt(r)
    delta = (m * 9.81 / kh) ** (2.0 / 3.0)
    sc = scenarios.two_spheres(d=d, rest=rest, dt=dt, pi_method=pi, pri_method=pri, ct_model=ct_model)
    sc["pp"] = (fric[0], fric[1], roll, rest, rest)
    sc["pw"] = (fric[0], fric[1], roll, rest, rest)
    st = sc["state"]
    for k in ("id", "ptype", "usr_mark"):
        st[k] = st[k][:1]
    for k in ("lin_vel", "lin_acc", "rot_vel", "rot_acc", "theta"):
        st[k] = np.zeros_like(st[k][:1])
    st["posr"] = np.array([[0.01, L / 2, r - delta, r]])
    sc["options"].update(gravity=(0.0, 0.0, -9.81), domain_max=(L, L, H), num_prtcl=1)
    sc["walls"] = scenarios.box_walls((0, 0, 0), (L, L, H), faces="z")
    return sc, m


# ---------------------------------------------------------------------------
# (i) finite plane walls: PW_isInContact against an independent distance
# ---------------------------------------------------------------------------
def expected_wall_contact(corners, both, p, r):
    """Sphere (p, r) against the quad p1..p4: the plane distance gates as in
    PW_isInContact (one side unless both-sided, |dist| <= r); the contact
    point is the quad point nearest to p.  One deviation is the reference's
    own and is stated here: its edge tests run in the order p1p2, p2p3,
    p3p4, p4p1, and each one also takes its *start* corner for a centre
    that projects up to r/length before the edge (Line_point_check,
    Prtcl_Geometry.f90:244-251), so a centre outside the last edge p4p1
    but within r of p1 and projecting within r/length before p1 on p1p2 is
    given corner p1, not its nearest point on p4p1.  (The end-corner branch
    `t >= 1 + r/length` never overlaps, and between the edges the next
    edge's start corner covers the corners.)"""
    p1, p2, p3, p4 = corners
    n = np.cross(p2 - p1, p3 - p1)
    n /= np.linalg.norm(n)
    dist = float((p - p1) @ n)
    if not both and dist < 0:
        return None
    ov_plane = r - abs(dist)
    if ov_plane < 0:
        return None
    eu, ev = p2 - p1, p4 - p1
    u, v = float((p - p1) @ eu / (eu @ eu)), float((p - p1) @ ev / (ev @ ev))
    if 0 <= u <= 1 and 0 <= v <= 1:
        return ov_plane, np.sign(dist) * n
    len1 = float(np.linalg.norm(eu))
    if -r / len1 <= u < 0 and np.linalg.norm(p - p1) <= r:
        q = p1
    else:
        best = None
        for a, b in ((p1, p2), (p2, p3), (p3, p4), (p4, p1)):
            t = np.clip((p - a) @ (b - a) / ((b - a) @ (b - a)), 0.0, 1.0)
            qq = a + t * (b - a)
            dd = np.linalg.norm(p - qq)
            if best is None or dd < best[0]:
                best = (dd, qq)
        q = best[1]
    dq = float(np.linalg.norm(p - q))
    if r - dq < 0:
        return None
    return r - dq, (p - q) / dq


def wall_probe_context(lib):
    sc = scenarios.two_spheres()
    sc["options"]["domain_max"] = (0.05, 0.05, 0.05)
    sc["walls"] = [dict(p1=tuple(c[0]), p2=tuple(c[1]), p3=tuple(c[2]), p4=tuple(c[3]), user_id=k + 1,
                        wall_type=1, both_side=both, is_infinite=False, trans_vel=(0.0, 0.0, 0.0))
                   for k, (c, both) in enumerate(quads())]
    return make(lib, sc)


def test_finite_wall_contact_matches_point_to_quad_distance(oracle_lib):
    e = wall_probe_context(oracle_lib)
    nw = C.c_int(0)
    oracle_lib.oracle_get_num_walls(e.h, C.byref(nw))
    assert nw.value == len(quads())
    rng = np.random.default_rng(42)
    seen = dict(plane=0, edge=0, corner_quirk=0, miss=0)
    for k, (corners, both) in enumerate(quads(), start=1):
        p1, p2, p3, p4 = corners
        n = np.cross(p2 - p1, p3 - p1)
        n /= np.linalg.norm(n)
        eu, ev = p2 - p1, p4 - p1
        for i in range(4000):
            r = rng.uniform(0.5e-3, 2.5e-3)
            if i % 4 == 0:   # near the corners
                u = rng.choice([0.0, 1.0]) + rng.uniform(-1.5, 1.5) * r / np.linalg.norm(eu)
                v = rng.choice([0.0, 1.0]) + rng.uniform(-1.5, 1.5) * r / np.linalg.norm(ev)
            else:
                u, v = rng.uniform(-0.15, 1.15, 2)
            h = rng.uniform(-1.2, 1.2) * r
            p = p1 + u * eu + v * ev + h * n
            ov = C.c_double(0)
            nv = (C.c_double * 3)()
            pr = (C.c_double * 4)(*p, r)
            hit = oracle_lib.oracle_wall_contact(e.h, k, pr, C.byref(ov), nv)
            want = expected_wall_contact(corners, both, p, r)
            if want is None:
                assert hit == 0, (k, u, v, h, r)
                seen["miss"] += 1
                continue
            assert hit == 1, (k, u, v, h, r, want)
            assert abs(ov.value - want[0]) <= 1e-12 * r, (k, u, v, ov.value, want[0])
            assert np.allclose(np.array(nv[:]), want[1], rtol=0, atol=1e-9), (k, u, v, nv[:], want[1])
            inside = 0 <= u <= 1 and 0 <= v <= 1
            quirk = (not inside) and -r / np.linalg.norm(eu) <= u < 0 and 0 < v <= 1
            seen["plane" if inside else ("corner_quirk" if quirk else "edge")] += 1
    # every branch is exercised
    assert min(seen.values()) > 20, seen


def test_edge_test_order_is_the_reference_one(oracle_lib):
    """The start-corner window of p1p2 takes precedence over the edge p4p1
    (checked last): a centre just outside p4p1, near p1, gets corner p1's
    smaller overlap.  A restatement that checked the edges in another order,
    or took the true nearest point, fails here."""
    e = wall_probe_context(oracle_lib)
    corners, _ = quads()[0]
    p1, p2, p3, p4 = corners
    r = 2.0e-3
    eu, ev = (p2 - p1) / np.linalg.norm(p2 - p1), (p4 - p1) / np.linalg.norm(p4 - p1)
    p = p1 - 0.6 * r * eu + 0.3 * r * ev + 0.5 * r * np.array([0, 0, 1.0])
    ov = C.c_double(0)
    nv = (C.c_double * 3)()
    assert oracle_lib.oracle_wall_contact(e.h, 1, (C.c_double * 4)(*p, r), C.byref(ov), nv) == 1
    d_corner = np.linalg.norm(p - p1)
    d_edge = np.linalg.norm((p - p1) - ((p - p1) @ ev) * ev)
    assert d_edge < d_corner
    assert abs(ov.value - (r - d_corner)) < 1e-15
    assert np.allclose(nv[:], (p - p1) / d_corner, atol=1e-12)


# ---------------------------------------------------------------------------
# (ii) near-wall list: cadence kinematics and no missed wall contact
# ---------------------------------------------------------------------------
def cadence_scene(seed=3):
    """Spheres above an infinite floor and below an infinite lid, all moving
    vertically at up to 3 m/s (the fastest ones toward a wall from outside
    its near-wall box), gravity on, periodic in x and y."""
    rng = np.random.default_rng(seed)
    n, L, H, d = 120, 0.03, 0.03, 2.0e-3
    pos = np.stack([rng.uniform(0, L, n), rng.uniform(0, L, n), rng.uniform(0.2 * H, 0.8 * H, n)], 1)
    ang = rng.permutation(n)
    pos[:, 0] = (ang % 12 + 0.5) * L / 12
    pos[:, 1] = (ang // 12 + 0.5) * L / 10
    posr = np.concatenate([pos, np.full((n, 1), d / 2)], 1)
    st = scenarios.zero_state(np.arange(1, n + 1), np.ones(n, int), posr)
    vz = rng.uniform(-3.0, 3.0, n)
    vz[:4] = [-3.0, 3.0, -2.9, 2.9]
    st["lin_vel"][:, :, 2] = vz[:, None]
    sc = dict(options=scenarios.base_options(dt=1.0e-5, domain_max=(L, L, H), is_periodic=(1, 1, 0),
                                             num_prtcl=n, wall_max_update_iter=1000),
              ptypes=dict(diameter=[d], density=[RHO], youngs=[E], poisson=[NU]),
              pp=(0.4, 0.4, 0.0, 0.8, 0.8), wtypes=dict(youngs=[E], poisson=[NU]), pw=(0.4, 0.4, 0.0, 0.8, 0.8),
              walls=scenarios.box_walls((0, 0, 0), (L, L, H), infinite=True, faces="zZ"), state=st)
    return sc


def test_near_wall_cadence_kinematics_and_no_missed_contact(oracle_lib):
    """numIter is the time a particle needs to cover DeltaXmax1 =
    Wall_neighbor_ratio * r_max starting at v = 1.5 max(max|v| + v_wall,
    0.3) under a = 1.5 max(max|a|, |g|) (Prtcl_ContactSearchPW.f90:201-226),
    here from v t + a t^2 / 2 = DeltaXmax1 solved as t = 2 dX / (v +
    sqrt(v^2 + 2 a dX)).  And at every step the wall contacts are those of
    an all-walls brute-force check at the step's start positions: the
    near-wall box gate never hides a contact."""
    sc = cadence_scene()
    e = make(oracle_lib, sc)
    o = sc["options"]
    rmax = max(sc["ptypes"]["diameter"]) / 2
    dx1 = o["wall_neighbor_ratio"] * rmax
    H = o["domain_max"][2]
    refreshes, skipped, wall_steps = 0, 0, 0
    for step in range(400):
        s0 = e.get_state()
        it0, nx0 = e.get_iter()
        e.iterate(1)
        it1, nx1 = e.get_iter()
        if it0 >= nx0:  # the near-wall lists were refreshed in this step
            refreshes += 1
            v = 1.5 * max(np.linalg.norm(s0["lin_vel"][:, 0], axis=1).max(), 0.3)
            a = 1.5 * max(np.linalg.norm(s0["lin_acc"][:, 0], axis=1).max(), 9.81)
            t = 2 * dx1 / (v + math.sqrt(v * v + 2 * a * dx1))
            q = t / o["dt"]
            if abs(q - round(q)) > 1e-9:
                num = min(max(int(q), 1), o["wall_max_update_iter"])
                assert nx1 == it0 + num, (step, nx1, it0, num)
            skipped = max(skipped, nx1 - it0)
        else:
            assert nx1 == nx0
        # brute force: every sphere against both walls at the step's start
        z, r, ids = s0["posr"][:, 2], s0["posr"][:, 3], s0["id"]
        want = {(int(i), 10000001) for i in ids[z <= r]} | {(int(i), 10000002) for i in ids[H - z <= r]}
        c = e.get_contacts()
        got = {(int(a), int(b)) for a, b, k in zip(c["owner"], c["partner"], c["kind"]) if k == native.KIND_PW}
        assert got == want, (step, sorted(got ^ want)[:5])
        wall_steps += bool(want)
    assert refreshes > 5 and skipped > 5 and wall_steps > 100, (refreshes, skipped, wall_steps)


# ---------------------------------------------------------------------------
# (iii) the pair table's row-type rule (A1) for unequal spheres
# ---------------------------------------------------------------------------
def hertz_collision_ode(mu_true, rstar, mstar, rest, v0):
    """mu_true x'' = -K x^1.5 - D x^0.25 x' from x=0, x'=v0 until x returns to
    0, with K = 4/3 E* sqrt(R*) and D = -2.2664 ln(e) sqrt(m* K) /
    sqrt(ln^2 e + 10.1354) from the given R*, m* (Prtcl_Property.f90:239-244).
    Returns (rebound ratio, contact time)."""
    estar = E / (2 * (1 - NU * NU))
    K = 4.0 / 3.0 * estar * math.sqrt(rstar)
    le = math.log(rest)
    D = -2.2664 * le * math.sqrt(mstar * K) / math.sqrt(le * le + 10.1354)

    def f(t, y):
        x, xd = y
        xp = max(x, 0.0)
        return [xd, (-K * xp ** 1.5 - D * xp ** 0.25 * xd) / mu_true]

    def ends(t, y):
        return y[0]
    ends.terminal, ends.direction = True, -1
    sol = solve_ivp(f, (0, 1e-2), [0.0, v0], events=ends, rtol=1e-11, atol=1e-18, max_step=1e-7,
                    first_step=1e-10)
    t_end = sol.t_events[0][-1]
    return -sol.y_events[0][-1][1] / v0, t_end


@pytest.mark.parametrize("small_first", [True, False])
def test_unequal_spheres_take_the_row_type(oracle_lib, small_first):
    """Head-on collision of a 1.5 mm and a 3 mm sphere: the force law takes
    R* = R/2 and m* = m/2 of the lower-id sphere's type for both partners
    (Prtcl_Property.f90:137-138, the row = the lower id's type,
    Prtcl_CL_and_CF.f90:743), while the motion has the true masses.  The
    rebound ratio and contact time follow the 1-D ODE with those
    coefficients, and swapping the ids changes both."""
    d1, d2, v, rest, dt = 1.5e-3, 3.0e-3, 0.5, 0.8, 2.0e-8
    sc = scenarios.two_spheres(d=d1, v=v, rest=rest, dt=dt)
    L = 0.05
    c = L / 2
    posr = np.array([[c - d1 / 2, c, c, d1 / 2], [c + d2 / 2, c, c, d2 / 2]])
    ids = [1, 2] if small_first else [2, 1]
    st = scenarios.zero_state(ids, [1, 2], posr, lin_vel1=np.array([[v, 0, 0], [-v, 0, 0]]))
    sc["state"] = st
    sc["ptypes"] = dict(diameter=[d1, d2], density=[RHO] * 2, youngs=[E] * 2, poisson=[NU] * 2)
    e = make(oracle_lib, sc)
    steps_in = 0
    for _ in range(60000):
        e.iterate(1)
        if e.get_counts().sum():
            steps_in += 1
        elif steps_in:
            break
    s = e.get_state()  # ascending id
    v_small = s["lin_vel"][list(s["id"]).index(ids[0]), 0, 0]
    v_big = s["lin_vel"][list(s["id"]).index(ids[1]), 0, 0]
    e_meas = (v_big - v_small) / (2 * v)
    m1, m2 = RHO * math.pi / 6 * d1 ** 3, RHO * math.pi / 6 * d2 ** 3
    row_d, row_m = (d1, m1) if small_first else (d2, m2)
    e_ode, t_ode = hertz_collision_ode(m1 * m2 / (m1 + m2), row_d / 4, row_m / 2, rest, 2 * v)
    assert abs(e_meas - e_ode) < 2e-4, (e_meas, e_ode)
    assert abs(steps_in * dt - t_ode) < 3 * dt, (steps_in * dt, t_ode)
    # the true pair (R* of both, m* of both) would give another result
    e_true, t_true = hertz_collision_ode(m1 * m2 / (m1 + m2), (d1 * d2 / 4) / (d1 + d2), m1 * m2 / (m1 + m2),
                                         rest, 2 * v)
    assert abs(e_true - e_ode) > 20 * 2e-4 or abs(t_true - t_ode) > 20 * dt


# ---------------------------------------------------------------------------
# (iv) sliding friction with mu_s != mu_k, (v) Hertz tangential stiffness
# ---------------------------------------------------------------------------
def test_sliding_decelerates_with_the_kinetic_coefficient(oracle_lib):
    """A sphere launched sliding (v0, no spin) on a floor with mu_s = 0.4,
    mu_k = 0.25: each step slips (the elastic trial force exceeds mu_s |fn|
    by far) and is capped at mu_k |fn| (Prtcl_ContactForce_inc.f90:44-48),
    so the sphere decelerates at mu_k g and its spin grows at 5 mu_k g /
    (2 R) until it rolls, at t = 2 v0 / (7 mu_k g), with v = 5/7 v0."""
    mus, muk, v0, dt, g = 0.4, 0.25, 0.2, 2.0e-6, 9.81
    sc, m = one_sphere_on_floor(fric=(mus, muk), dt=dt)
    sc["state"]["lin_vel"][0, :, 0] = v0
    e = make(oracle_lib, sc)
    t_roll = 2 * v0 / (7 * muk * g)
    n1 = int(0.1 * t_roll / dt)
    n2 = int(0.5 * t_roll / dt)
    e.iterate(n1)
    va = e.get_state()["lin_vel"][0, 0, 0]
    e.iterate(n2 - n1)
    vb = e.get_state()["lin_vel"][0, 0, 0]
    decel = (va - vb) / ((n2 - n1) * dt)
    assert abs(decel - muk * g) < 2e-3 * muk * g, (decel, muk * g, mus * g)
    # rolling sets in (contact point at rest) at t_roll
    r = sc["state"]["posr"][0, 3]
    k = n2
    while True:
        e.iterate(50)
        k += 50
        s = e.get_state()
        if s["lin_vel"][0, 0, 0] - s["rot_vel"][0, 0, 1] * r < 1e-3 * v0:
            break
        assert k * dt < 2 * t_roll
    assert abs(k * dt - t_roll) < 0.02 * t_roll, (k * dt, t_roll)
    assert abs(s["lin_vel"][0, 0, 0] - 5.0 / 7.0 * v0) < 2e-3 * v0


def test_hertz_tangential_stiffness(oracle_lib):
    """First step of a sheared Hertz-Mindlin contact: the history is u dt and
    the tangential force is kt u dt with kt = (16/3) G* sqrt(R* delta),
    G* = 1 / ((2-nu_i)/G_i + (2-nu_j)/G_j), G = E / (2 (1 + nu))
    (Prtcl_Property.f90:243, _inc:35; not the common 8 G* sqrt(R* delta));
    the normal force is K delta^1.5 (no normal motion)."""
    d, u, dt, delta = 2.0e-3, 0.05, 1.0e-8, 2.0e-6
    sc = scenarios.two_spheres(d=d, v=0.0, gap=-delta, rest=0.8, dt=dt, fric=0.5)
    sc["state"]["lin_vel"][1, 0, :] = (0.0, u, 0.0)
    e = make(oracle_lib, sc)
    e.iterate(1)
    _, f, _ = e.get_forces()
    G = E / (2 * (1 + NU))
    gstar = 1.0 / (2 * (2 - NU) / G)
    estar = 1.0 / (2 * (1 - NU * NU) / E)
    rstar = d / 4
    kt = 16.0 / 3.0 * gstar * math.sqrt(rstar * delta)
    fn = 4.0 / 3.0 * estar * math.sqrt(rstar) * delta ** 1.5
    assert abs(f[0, 1] - kt * u * dt) < 1e-9 * kt * u * dt, (f[0, 1], kt * u * dt)
    assert abs(-f[0, 0] - fn) < 1e-9 * fn
    # the history is kept from the lower id's side: its relative velocity
    # Vel_i - Vel_j is -u along y
    ct = e.get_contacts()
    assert abs(ct["tan_delta"][0, 1] + u * dt) < 1e-9 * u * dt


# ---------------------------------------------------------------------------
# (vi) rolling resistance, both torque models
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("ct_model", [native.CTM_CONSTANT_TORQUE, native.CTM_VARIABLE_TORQUE])
def test_rolling_resistance_decelerates_a_rolling_sphere(oracle_lib, ct_model):
    """A sphere rolling on a floor (v = omega R) under the rolling torque
    M_r = -mu_r |fn| R* w_hat (constant torque, _inc:57) or -mu_r |fn| R*
    |v_w| w_hat (variable torque, _inc:59; |v_w| = omega R = v): with I =
    2/5 m R^2 and no slip, m a (1 + I/(m R^2)) = -M_r / R, i.e. a = -5/7 mu_r
    g (linear decay) or a = -5/7 mu_r g v (exponential decay)."""
    # the variable-torque coefficient multiplies a speed (its unit is s/m)
    mur = 0.05 if ct_model == native.CTM_CONSTANT_TORQUE else 2.0
    v0, dt, g = 0.1, 2.0e-6, 9.81
    sc, m = one_sphere_on_floor(fric=(0.5, 0.5), roll=mur, ct_model=ct_model, dt=dt)
    r = sc["state"]["posr"][0, 3]
    sc["state"]["lin_vel"][0, :, 0] = v0
    sc["state"]["rot_vel"][0, :, 1] = v0 / r
    e = make(oracle_lib, sc)
    T = 0.05
    nst = int(T / dt)
    e.iterate(nst)
    s = e.get_state()
    v = s["lin_vel"][0, 0, 0]
    # still rolling (the tangential spring's elastic creep leaves a slip of
    # under 1e-3 v0)
    assert abs(v - s["rot_vel"][0, 0, 1] * r) < 2e-3 * v0
    if ct_model == native.CTM_CONSTANT_TORQUE:
        want = v0 - 5.0 / 7.0 * mur * g * T
    else:
        want = v0 * math.exp(-5.0 / 7.0 * mur * g * T)
    assert abs(v - want) < 1e-3 * v0, (v, want)
    assert abs(v - v0) > 0.1 * v0  # the resistance did act


# ---------------------------------------------------------------------------
# (vii) the PP torque split
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("roll", [0.0, 0.1])
def test_pair_torque_split(oracle_lib, roll):
    """Isolated touching pairs with random velocities and spins: the lower
    id gets F and r_lo (n x F_t) + M_r, the higher id -F and r_hi (n x F_t)
    - M_r (ContactForce_PP, Prtcl_CL_and_CF.f90:732-756), where n is the
    unit vector from the lower to the higher id's centre and F_t = F - (F.n)
    n; M_r = -mu_r |fn| R* unit(w_lo - w_hi) (constant torque)."""
    rng = np.random.default_rng(17)
    npair, L = 40, 0.06
    d = np.array([1.5e-3, 2.0e-3, 3.0e-3])
    rows = []
    for p in range(npair):
        c = np.array([(p % 8 + 0.5) * L / 8, (p // 8 + 0.5) * L / 5, L / 2])
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        ta, tb = rng.integers(0, 3, 2)
        ra, rb = d[ta] / 2, d[tb] / 2
        s = (ra + rb) * (1 - 2e-3)
        rows.append((c - 0.5 * s * u, ra, ta + 1))
        rows.append((c + 0.5 * s * u, rb, tb + 1))
    n = len(rows)
    posr = np.array([[*x, r] for x, r, _ in rows])
    ids = rng.permutation(np.arange(1, n + 1))
    st = scenarios.zero_state(ids, [t for _, _, t in rows], posr, native.PIM_AB3, native.PIM_AB3)
    st["lin_vel"][:, 0, :] = rng.uniform(-0.2, 0.2, (n, 3))
    st["rot_vel"][:, 0, :] = rng.uniform(-200, 200, (n, 3))
    sc = dict(options=scenarios.base_options(dt=1.0e-7, gravity=(0, 0, 0), domain_max=(L, L, L), num_prtcl=n,
                                             pi_method=native.PIM_AB3, pri_method=native.PIM_AB3),
              ptypes=dict(diameter=list(d), density=[RHO] * 3, youngs=[E] * 3, poisson=[NU] * 3),
              pp=(0.3, 0.3, roll, 0.8, 0.8), wtypes=dict(youngs=[E], poisson=[NU]), pw=(0.3, 0.3, roll, 0.8, 0.8),
              walls=[], state=st)
    e = make(oracle_lib, sc)
    e.iterate(1)
    fid, F, T = e.get_forces()
    assert e.get_counts()[0] == npair
    at = {int(i): k for k, i in enumerate(fid)}
    for p in range(npair):
        ia, ib = ids[2 * p], ids[2 * p + 1]
        lo, hi = (2 * p, 2 * p + 1) if ia < ib else (2 * p + 1, 2 * p)
        xlo, xhi = posr[lo, :3], posr[hi, :3]
        rlo, rhi = posr[lo, 3], posr[hi, 3]
        nvec = (xhi - xlo) / np.linalg.norm(xhi - xlo)
        Flo, Fhi = F[at[int(ids[lo])]], F[at[int(ids[hi])]]
        Tlo, Thi = T[at[int(ids[lo])]], T[at[int(ids[hi])]]
        scale = np.linalg.norm(Flo)
        assert np.allclose(Flo, -Fhi, rtol=0, atol=1e-13 * scale)
        fn = Flo @ nvec
        ft = Flo - fn * nvec
        m_t = np.cross(nvec, ft)
        w = st["rot_vel"][lo, 0] - st["rot_vel"][hi, 0]
        row = rows[lo][2] - 1  # the lower id's type for both (Prtcl_Property.f90:137-138)
        rstar = d[row] / 4
        mr = -roll * abs(fn) * rstar * w / np.linalg.norm(w)
        tol = 1e-9 * scale * max(rlo, rhi)
        assert np.allclose(Tlo, rlo * m_t + mr, rtol=0, atol=tol), (p, Tlo, rlo * m_t + mr)
        assert np.allclose(Thi, rhi * m_t - mr, rtol=0, atol=tol), (p, Thi, rhi * m_t - mr)
        assert np.linalg.norm(ft) > 1e-3 * scale  # the tangential part is not trivial


# ---------------------------------------------------------------------------
# (viii) forward-Euler rotation
# ---------------------------------------------------------------------------
def test_fe_rotation_under_a_constant_torque(oracle_lib):
    """A sphere resting on a floor and spinning about the vertical axis:
    no contact-point velocity, so the only torque is the constant rolling
    torque -mu_r m g R (its normal force stays m g), alpha = -5 mu_r g /
    (2 R).  Forward Euler (Prtcl_Integration.f90:56-60) gives omega_n =
    omega_0 + n alpha dt and theta_n = theta_0 + n omega_0 dt + alpha dt^2
    n (n-1) / 2; the AB2 rotation of the reference would instead take the
    linear velocity (zero) as the spin."""
    mur, w0, dt, g, nst = 0.02, 30.0, 1.0e-6, 9.81, 2000
    sc, m = one_sphere_on_floor(roll=mur, dt=dt, pi=native.PIM_FE, pri=native.PIM_FE, rest=0.5)
    r = sc["state"]["posr"][0, 3]
    sc["state"]["rot_vel"][0, 0, 2] = w0
    e = make(oracle_lib, sc)
    e.iterate(nst)
    s = e.get_state()
    alpha = -5.0 * mur * g / (2 * r)
    w_want = w0 + nst * alpha * dt
    th_want = nst * w0 * dt + alpha * dt * dt * nst * (nst - 1) / 2
    assert abs(s["rot_vel"][0, 0, 2] - w_want) < 1e-6 * w0, (s["rot_vel"][0, 0, 2], w_want)
    assert abs(s["theta"][0, 2] - th_want) < 1e-6 * abs(th_want), (s["theta"][0, 2], th_want)
    assert np.abs(s["rot_vel"][0, 0, :2]).max() < 1e-9 * w0
