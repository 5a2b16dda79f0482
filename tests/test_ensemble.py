"""Batched ensembles: every box of an ensemble context gives the same bits as
a context of its own (the single-box GPU path, itself checked against the
oracle), for uneven box sizes, mixed contact models and after autotuning."""
import numpy as np
import pytest

import oracle
from parity import compare
from sdem_amd import native, scenarios

pytestmark = pytest.mark.gpu


def boxes(nbox, nmin=300, nmax=3000, seed=5):
    return scenarios.ensemble(nbox=nbox, seed=seed, nmin=nmin, nmax=nmax)


def snapshot(eng):
    st, ct = eng.get_state(), eng.get_contacts()
    return dict(st=st, ct=ct, counts=eng.get_counts().tolist(), it=eng.get_iter())


def same_bits(a, b):
    for k in ("id", "ptype", "posr", "lin_vel", "lin_acc", "theta", "rot_vel", "rot_acc"):
        assert np.array_equal(a["st"][k], b["st"][k]), k
    for k in ("kind", "owner", "partner", "tan_delta", "vrel_init"):
        assert np.array_equal(a["ct"][k], b["ct"][k]), k
    assert a["counts"] == b["counts"]
    assert a["it"] == b["it"]


def test_every_box_matches_its_own_context(hip_lib):
    print("ensemble: 9 boxes, 300-3000 spheres, 60 steps", flush=True)
    scs = boxes(9)
    ens = native.build_ensemble(hip_lib, scs)
    ens.iterate(60)
    for b, sc in enumerate(scs):
        ens.select(b)
        one = native.build_engine(hip_lib, sc)
        one.iterate(60)
        same_bits(snapshot(ens), snapshot(one))
        one.close()
    ens.close()


def test_ensemble_box_matches_oracle(hip_lib, oracle_lib):
    print("ensemble: box 1 of 4 against the oracle, 40 steps", flush=True)
    scs = boxes(4, nmin=400, nmax=1500, seed=9)
    ens = native.build_ensemble(hip_lib, scs)
    ens.iterate(40)
    for b in (0, 1):
        ens.select(b)
        ref = native.build_engine(oracle_lib, scs[b], prefix=oracle.PREFIX)
        ref.iterate(40)
        errs = compare(ens, ref, exact_positions=False, rtol=1e-11)
        assert not errs, (b, errs)


def test_autotune_keeps_the_bits(hip_lib):
    """The autotuner runs real steps under several tunings; the result is the
    same as a plain run of as many steps."""
    print("ensemble: autotune 5 boxes, then compare", flush=True)
    scs = boxes(5, seed=17)
    ens = native.build_ensemble(hip_lib, scs)
    ran, ms = ens.autotune(steps_per_trial=2)
    assert ran > 0 and ms > 0
    ens.iterate(6)
    for b, sc in enumerate(scs):
        ens.select(b)
        one = native.build_engine(hip_lib, sc)
        one.iterate(ran + 6)
        same_bits(snapshot(ens), snapshot(one))
        one.close()
