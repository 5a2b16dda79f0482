"""Edge cases of the step against the oracle: an empty box, a single sphere,
particles leaving through open sides until the box is empty, particle ids
at the top of their range, and a long chaotic rollout judged by ensemble
statistics (needs the GPU)."""
import numpy as np
import pytest

import oracle
from parity import compare
from sdem_amd import native, scenarios

pytestmark = pytest.mark.gpu


def pair(sc, hip_lib, oracle_lib):
    return native.build_engine(hip_lib, sc), native.build_engine(oracle_lib, sc, prefix=oracle.PREFIX)


def subset(sc, keep):
    """The scenario with only the particles `keep` (indices)."""
    s2 = dict(sc)
    s2["state"] = {k: np.asarray(v)[keep].copy() for k, v in sc["state"].items()}
    return s2


def test_empty_box(hip_lib, oracle_lib):
    print("empty box: 10 steps", flush=True)
    sc = subset(scenarios.random_cloud(n=50, seed=3), np.zeros(0, int))
    gpu, ref = pair(sc, hip_lib, oracle_lib)
    for k in (1, 9):
        gpu.iterate(k)
        ref.iterate(k)
        assert gpu.num_particles() == 0 == ref.num_particles()
        assert gpu.get_iter() == ref.get_iter()
        assert gpu.get_counts().tolist() == [0, 0, 0, 0]
        assert len(gpu.get_contacts()["kind"]) == 0


def test_single_sphere_on_a_wall(hip_lib, oracle_lib):
    """One sphere falls onto the bottom wall and bounces (PW history only)."""
    print("single sphere: 400 steps", flush=True)
    sc = subset(scenarios.random_cloud(n=50, seed=5, vel=0.0, omega=0.0), np.array([0]))
    sc["options"] = dict(sc["options"], gravity=(0.0, 0.0, -9.81))
    sc["state"]["posr"][0, 2] = sc["state"]["posr"][0, 3] * 1.02
    gpu, ref = pair(sc, hip_lib, oracle_lib)
    seen_contact = False
    for _ in range(8):
        gpu.iterate(50)
        ref.iterate(50)
        errs = compare(gpu, ref, exact_positions=False, rtol=1e-10)
        assert not errs, errs
        seen_contact |= gpu.get_counts()[3] > 0
    assert seen_contact


def test_particles_leave_through_open_sides(hip_lib, oracle_lib):
    """Only a bottom wall: fast particles leave through the open sides and
    top and are removed, down to an empty box; the survivors and the step
    counter stay the oracle's throughout."""
    print("open sides: deletions down to an empty box", flush=True)
    sc = scenarios.random_cloud(n=400, seed=9, walls="z", vel=0.5, dt=1.0e-5)
    st = sc["state"]
    rng = np.random.default_rng(1)
    st["lin_vel"][:, 0, 2] += rng.uniform(20.0, 60.0, len(st["id"]))  # upwards, out of the top
    st["lin_vel"][:, 0, 1] += rng.uniform(-30.0, 30.0, len(st["id"]))
    st["lin_vel"][:, 0, 0] += rng.uniform(-30.0, 30.0, len(st["id"]))
    gpu, ref = pair(sc, hip_lib, oracle_lib)
    counts = []
    for k in range(12):
        gpu.iterate(10)
        ref.iterate(10)
        errs = compare(gpu, ref, exact_positions=False, rtol=1e-10)
        assert not errs, (k, errs)
        counts.append(gpu.num_particles())
    assert counts[0] < 400 and counts[-1] == 0, counts
    gpu.iterate(5)
    ref.iterate(5)
    assert gpu.num_particles() == 0 and gpu.get_iter() == ref.get_iter()


def test_largest_ids(hip_lib, oracle_lib):
    """Particle ids up to 2^29 - 1, the largest the contact keys hold."""
    print("ids near 2^29", flush=True)
    sc = scenarios.random_cloud(n=300, seed=12, periodic=(1, 1, 0), walls="z")
    top = (1 << 29) - 1
    sc["state"]["id"] = (top - np.asarray(sc["state"]["id"]) + 1).astype(np.int32)
    sc["options"] = dict(sc["options"], base_wall_id=1000)
    gpu, ref = pair(sc, hip_lib, oracle_lib)
    for _ in range(3):
        gpu.iterate(10)
        ref.iterate(10)
        errs = compare(gpu, ref, exact_positions=False, rtol=1e-10)
        assert not errs, errs
    assert gpu.get_contacts()["owner"].max() > top - 400


def test_long_rollout_statistics(hip_lib, oracle_lib):
    """3,000 steps of a dense, colliding, dissipating cloud in a closed box:
    single trajectories part ways (chaos amplifies the last-ulp differences
    of the summation order), so the run is judged as the reference's long
    runs are, by ensemble statistics.  Kinetic energy (its decay), bed
    height and contacts per particle agree with the oracle's at every 500
    steps: height within 0.5 %, kinetic energy within 5 %, contacts within
    3 %."""
    print("long rollout: 1200 spheres, 3000 steps, ensemble statistics", flush=True)
    from sdem_amd import surrogate
    sc = scenarios.random_cloud(n=1200, seed=17, vel=0.5, omega=200.0)
    sc["options"] = dict(sc["options"], gravity=(0.0, 0.0, -9.81))
    gpu, ref = pair(sc, hip_lib, oracle_lib)
    o = sc["options"]
    area = (o["domain_max"][0] - o["domain_min"][0]) * (o["domain_max"][1] - o["domain_min"][1])
    ke0 = None
    for k in range(6):
        gpu.iterate(500)
        ref.iterate(500)
        g, r = surrogate.bed_metrics(gpu.get_state(), area), surrogate.bed_metrics(ref.get_state(), area)
        cg, cr = gpu.get_counts().sum(), ref.get_counts().sum()
        n = gpu.num_particles()
        print(f"  step {500 * (k + 1)}: height {g['height']:.6e} / {r['height']:.6e}, "
              f"ke {g['ke']:.4e} / {r['ke']:.4e}, contacts {cg} / {cr}", flush=True)
        assert n == ref.num_particles()
        assert abs(g["height"] - r["height"]) <= 5e-3 * r["height"]
        assert abs(g["ke"] - r["ke"]) <= 5e-2 * r["ke"]
        assert abs(int(cg) - int(cr)) <= 0.03 * cr + 2
        ke0 = r["ke"] if ke0 is None else ke0
    assert r["ke"] < 0.5 * ke0  # the energy decays


def test_runaway_particle_is_an_error(hip_lib):
    """A sphere that crosses a periodic dimension many times in one step
    (1e4 m/s, dt 1e-4 s, a 0.05 m period) leaves a coordinate that one
    wrap does not bring back into the domain; the next rebuild's cell walk
    refuses it and sdem_step returns SDEM_E_STATE (a blown-up state) instead
    of reading past the cell grid.  The reference has no such check (its
    cell index would fall outside its arrays)."""
    print("runaway particle: 2 steps, expecting SDEM_E_STATE", flush=True)
    sc = scenarios.random_cloud(n=60, seed=12, spacing=1.5, periodic=(1, 1, 1), walls="", dt=1.0e-4)
    sc["state"]["lin_vel"][5, :, 0] = 1.0e4
    g = native.build_engine(hip_lib, sc)
    with pytest.raises(native.SdemError) as err:
        g.iterate(3)
    assert err.value.code == native.SDEM_E_STATE
    assert "blown up" in str(err.value)
    g.close()


def test_too_fine_a_grid_is_an_error(hip_lib):
    """A cell factor so small that the cell count of an axis would overflow
    is refused when the grid is made (before any kernel runs), with
    SDEM_E_ARG, instead of a grid built from a wrapped count."""
    print("grid of 1e-12 cell factor: expecting SDEM_E_ARG", flush=True)
    sc = scenarios.random_cloud(n=60, seed=12)
    with pytest.raises(native.SdemError) as err:
        g = native.build_engine(hip_lib, sc, tuning=dict(cell_factor=(1.0e-12, 1.0, 1.0)))
        g.iterate(1)
    assert err.value.code == native.SDEM_E_ARG
    assert "search grid too large" in str(err.value)
