"""The benchmark configurations at their full sizes against the oracle.

The other parity tests shrink the configurations so that the oracle
finishes in seconds.  Here configuration 2 (1,000,000 spheres: periodic x
and y, bottom wall, 32-type Hertz-Mindlin with history, AB2) and
configuration 3 (4,000,000 spheres of three sizes, hierarchical search,
closed box with the 30-degree incline) run at the sizes the benchmark
times, for the few steps the oracle can take in seconds at that size:

* configuration 2 from its own start (a loose pack falling, no contacts
  yet: the integration, the periodic images and the near-wall lists over
  the whole grid);
* configuration 2 settled into a resting pack on the HIP path first (about
  2.3 contact records per sphere: the contact test, the history and the
  forces at the full grid), then stepped by both sides; and on the HIP path
  alone, the same pack under two search tunings, which must give the same
  bits (the contact set is decided exactly every step);
* configuration 3 from its own start, over the three-level grid;
* configuration 4 (256 independent sand boxes, 3.76M spheres, linear and
  Hertz-Mindlin boxes with their own materials) settled into contact on
  the HIP path first, every box against an oracle context of its own.

The contact set is compared exactly, the state and history within the
stated tolerance (the first step with exact positions).
"""
import functools

import numpy as np
import pytest

import oracle
from parity import compare
from sdem_amd import native, scenarios

pytestmark = pytest.mark.gpu


def _log(m):
    print(m, flush=True)


def _same_bits(a, b):
    sa, sb = a.get_state(), b.get_state()
    ca, cb = a.get_contacts(), b.get_contacts()
    return all(np.array_equal(sa[k], sb[k]) for k in sa) and all(np.array_equal(ca[k], cb[k]) for k in ca)


@functools.lru_cache(maxsize=None)
def settling_1m():
    _log("configuration 2 at 1M spheres: building the scenario")
    return scenarios.settling()


@functools.lru_cache(maxsize=None)
def settled_1m():
    _log("configuration 2 at 1M spheres: settling on the HIP path (40k steps at 4 g, 6k at 2 g, 12k at g)")
    s = scenarios.settle(settling_1m(), 40_000, gravity_scale=4.0, chunk=10_000, log=_log)
    s = scenarios.settle(s, 6_000, gravity_scale=2.0, chunk=6_000, log=_log)
    return scenarios.settle(s, 12_000, chunk=6_000, log=_log)


def _step_and_compare(sc, hip_lib, oracle_lib, schedule):
    gpu = native.build_engine(hip_lib, sc)
    ref = native.build_engine(oracle_lib, sc, prefix=oracle.PREFIX)
    for k, exact, rtol in schedule:
        gpu.iterate(k)
        ref.iterate(k)
        errs = compare(gpu, ref, exact_positions=exact, rtol=rtol)
        assert not errs, (k, errs)
        assert gpu.num_particles() == ref.num_particles()
        assert gpu.get_counts().tolist() == ref.get_counts().tolist()
        _log(f"  after {k} more: {gpu.num_particles()} spheres, counts {ref.get_counts().tolist()}")
    return gpu, ref


def test_settling_bed_1m_from_its_start(hip_lib, oracle_lib):
    sc = settling_1m()
    _log(f"configuration 2 from its start: {len(sc['state']['id'])} spheres, 1 + 4 steps")
    _step_and_compare(sc, hip_lib, oracle_lib, ((1, True, 1e-12), (4, False, 1e-11)))


def test_settled_bed_1m_in_contact(hip_lib, oracle_lib):
    sc = settled_1m()
    n = len(sc["state"]["id"])
    _log(f"configuration 2 in contact: {n} spheres, {len(sc['contacts']['kind'])} records; 1 + 3 steps")
    _, ref = _step_and_compare(sc, hip_lib, oracle_lib, ((1, True, 1e-12), (3, False, 1e-11)))
    pp, ppg, ppfix, pw = ref.get_counts().tolist()
    assert pp > n, (pp, n)  # ~2.2 records per sphere: each PP pair from both sides
    assert ppg > 0 and pw > 0


def test_settled_bed_1m_search_tuning_keeps_the_bits(hip_lib):
    sc = settled_1m()
    _log("configuration 2 in contact: 40 steps under two search tunings")
    a = native.build_engine(hip_lib, sc, tuning=dict(cell_factor=(0.5, 0.5, 0.5), skin=0.05))
    b = native.build_engine(hip_lib, sc, tuning=dict(cell_factor=(1.0, 1.0, 1.0), skin=0.15))
    for _ in range(2):
        a.iterate(20)
        b.iterate(20)
        assert _same_bits(a, b)
    assert a.get_counts()[0] > len(sc["state"]["id"])


def test_polydisperse_4m_from_its_start(hip_lib, oracle_lib):
    _log("configuration 3 at 4M spheres: building the scenario")
    sc = scenarios.polydisperse()
    _log(f"configuration 3 from its start: {len(sc['state']['id'])} spheres, 1 + 1 steps")
    _step_and_compare(sc, hip_lib, oracle_lib, ((1, True, 1e-12), (1, False, 1e-11)))


def test_ensemble_256_boxes_in_contact(hip_lib, oracle_lib):
    _log("configuration 4, 256 boxes: building the scenarios")
    scs = scenarios.ensemble()
    _log("configuration 4: settling on the HIP path (60k steps at dt 1e-6 s and 2 g)")
    scs = scenarios.settle(scs, 60_000, dt=1.0e-6, gravity_scale=2.0, chunk=10_000, log=_log)
    for sc in scs:
        sc["options"]["dt"] = 1.0e-6  # at 1e-5 s the linear boxes' contacts last one to six steps
    n = sum(len(sc["state"]["id"]) for sc in scs)
    _log(f"configuration 4 in contact: {len(scs)} boxes, {n} spheres; 1 + 2 steps, each box against the oracle")
    ens = native.build_ensemble(hip_lib, scs)
    refs = [native.build_engine(oracle_lib, sc, prefix=oracle.PREFIX) for sc in scs]
    for k, exact, rtol in ((1, True, 1e-12), (2, False, 1e-11)):
        ens.iterate(k)
        tot = np.zeros(4, np.int64)
        for b, ref in enumerate(refs):
            ref.iterate(k)
            ens.select(b)
            errs = compare(ens, ref, exact_positions=exact, rtol=rtol)
            assert not errs, (b, k, errs)
            c = ref.get_counts()
            assert ens.get_counts().tolist() == c.tolist(), b
            tot += c
        _log(f"  after {k} more: counts over the boxes {tot.tolist()}")
    assert tot[0] > n // 2 and tot[3] > 0, tot
