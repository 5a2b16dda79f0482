"""A step phase by phase: the eight entry points that stand for the calls of
DEMS_iterate (Prtcl_DEMSystem.f90:213-248) give, in the reference's order,
the bits of sdem_step, enforce that order, recover from a capacity overflow,
and fill the reference's step timers (needs the GPU)."""
import numpy as np
import pytest

import oracle
from parity import compare
from sdem_amd import native, scenarios

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    ca, cb = a.get_contacts(), b.get_contacts()
    for k in ca:
        assert np.array_equal(ca[k], cb[k]), k
    assert a.get_iter() == b.get_iter()
    assert np.array_equal(a.get_counts(), b.get_counts())
    fa, fb = a.get_forces(), b.get_forces()
    for x, y in zip(fa, fb):
        assert np.array_equal(x, y)


CASES = {
    "walls_ab2": dict(),
    "periodic_fixed_ab3": dict(periodic=(1, 1, 0), walls="zZ", nfixed=20, pi=native.PIM_AB3, pri=native.PIM_AB3),
    "fe_variable_torque": dict(pi=native.PIM_FE, pri=native.PIM_FE, ct_model=native.CTM_VARIABLE_TORQUE,
                               cf_type=native.CFT_LSD),
    "hierarchical_fast": dict(diameters=(0.6e-3, 1.2e-3, 3.0e-3), cs_method=native.CSM_NBS_MUNJIZA_HRCHL,
                              vel=2.0, dt=1.0e-5, walls="z"),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_phases_give_the_step_bits(hip_lib, name):
    """Stepping phase by phase and with sdem_step give the same bits over
    steps with and without rebuilds (and deletions through open sides)."""
    print(f"phases vs sdem_step: {name}, 30 steps", flush=True)
    sc = scenarios.random_cloud(n=800, seed=17, **CASES[name])
    a = native.build_engine(hip_lib, sc)
    b = native.build_engine(hip_lib, sc)
    for k in (1, 4, 25):
        a.iterate(k)
        b.iterate_phases(k)
        same_bits(a, b)
        if k == 1:
            assert a.get_counts().sum() > 100


def test_phases_match_the_oracle(hip_lib, oracle_lib):
    print("phases vs the oracle: periodic cloud, 20 steps", flush=True)
    sc = scenarios.random_cloud(n=400, seed=3, periodic=(1, 0, 0), walls="yYzZ")
    g = native.build_engine(hip_lib, sc)
    r = native.build_engine(oracle_lib, sc, prefix=oracle.PREFIX)
    for k in (1, 19):
        g.iterate_phases(k)
        r.iterate(k)
        errs = compare(g, r, exact_positions=(k == 1), rtol=1e-11)
        assert not errs, errs


def test_phase_order_is_enforced(hip_lib):
    print("phase order", flush=True)
    sc = scenarios.random_cloud(n=100, seed=2)
    g = native.build_engine(hip_lib, sc)
    with pytest.raises(native.SdemError) as err:
        g.phase("find_contacts")  # no step open
    assert err.value.code == native.SDEM_E_STATE
    g.phase("pre_iteration")
    g.phase("comm_for_cntct")
    with pytest.raises(native.SdemError) as err:
        g.phase("integrate")  # skips the force phases
    assert err.value.code == native.SDEM_E_STATE
    with pytest.raises(native.SdemError) as err:
        g.iterate(1)  # a phased step is open
    assert err.value.code == native.SDEM_E_STATE
    for name in native.STEP_PHASES[2:]:
        g.phase(name)
    g.iterate(2)  # closed again


def test_phases_recover_from_a_capacity_overflow(hip_lib, oracle_lib):
    """One history slot and two list entries per particle: the search or
    force phase runs out, grows the capacity and puts the step back to its
    start; the step run again gives the oracle's result."""
    print("phases: capacity overflow, 6 steps", flush=True)
    sc = scenarios.random_cloud(n=300, seed=4)
    g = native.build_engine(hip_lib, sc, capacity=dict(hist_slots=1, list_slots=2))
    r = native.build_engine(oracle_lib, sc, prefix=oracle.PREFIX)
    with pytest.raises(native.SdemError) as err:
        for name in native.STEP_PHASES:
            g.phase(name)
    assert err.value.code == native.SDEM_E_CAPACITY
    g.iterate_phases(6)
    r.iterate(6)
    errs = compare(g, r, exact_positions=False, rtol=1e-11)
    assert not errs, errs


def test_timers_have_the_reference_names(hip_lib):
    print("timers: 5 phased steps, then 5 fused", flush=True)
    sc = scenarios.random_cloud(n=500, seed=6, walls="z")
    g = native.build_engine(hip_lib, sc)
    g.set_profiling(True)
    g.iterate_phases(5)
    t = g.get_timers()
    for name in native.TIMER_NAMES[1:8]:
        assert t[name][1] == 5, (name, t[name])
        assert t[name][0] > 0.0, name
    assert t["total"][1] == 5 and t["write_to_file"] == (0.0, 0)
    assert t["total"][0] >= sum(t[n][0] for n in native.TIMER_NAMES[1:8]) * (1 - 1e-9)
    g.set_profiling(True)
    g.iterate(5)
    t = g.get_timers()
    assert t["total"][1] == 5 and t["pre_iteration"][1] == 5 and t["cs_and_cf_pp"][1] == 5
    assert t["integration"] == (0.0, 0)  # fused into the force kernel, booked under CS and CF P-P
