"""Restart and visualisation files in the reference's formats.

CPU tests pin the restatement in oracle/restart_format.py (its writer lays
the file out sequentially, its reader finds the blocks by the reference
reader's own offset arithmetic) on states and contact records of the C
oracle, and the library's XDMF text against it and against lines derived
by hand from the Fortran edit descriptors.  GPU tests check the library's
readers and writers against the restatement and resume runs from restart
files, including files the restatement wrote from the oracle's state.
"""
import os

import numpy as np
import pytest

import oracle
from oracle import restart_format as R
from parity import compare
from sdem_amd import native, scenarios

# a periodic box with walls and fixed particles: PP, PPG, PPFix and PW records
CASE = dict(cf_type=native.CFT_NLIN, periodic=(1, 1, 0), walls="zZ", nfixed=30)


def case(n=300, seed=13, **kw):
    sc = scenarios.random_cloud(n=n, seed=seed, **{**CASE, **kw})
    return sc


def ids_of(sc):
    base = sc["options"]["base_wall_id"]
    walls = list(range(base + 1, base + 1 + len(sc.get("walls", []))))
    fixed = list(sc["fixed"]["id"]) if "fixed" in sc else []
    return walls, fixed


def record_set(ct):
    return {(int(k), int(o), int(p)): (tuple(np.asarray(t).tolist()), float(v))
            for k, o, p, t, v in zip(ct["kind"], ct["owner"], ct["partner"], ct["tan_delta"], ct["vrel_init"])}


def restarted(sc, k):
    """The same scenario continued from step k (ifirst = k + 1)."""
    s2 = dict(sc)
    s2["options"] = dict(sc["options"], ifirst=sc["options"].get("ifirst", 1) + k)
    return s2


# ----------------------------------------------------------------- CPU
@pytest.mark.parametrize("pi", [native.PIM_FE, native.PIM_AB2, native.PIM_AB3])
def test_restatement_round_trip_on_oracle_state(tmp_path, oracle_lib, pi):
    """Oracle state and records after 6 steps -> restated writer -> restated
    reader (offsets of PIO_Read_Restart / PIO_RestartCL) ->
    CL_Add_RestartCntctlink: the same state and the same records, every
    kind present."""
    sc = case(pi=pi, pri=pi)
    eng = native.build_engine(oracle_lib, sc, prefix=oracle.PREFIX)
    eng.iterate(6)
    st, ct = eng.get_state(), eng.get_contacts()
    assert set(ct["kind"].tolist()) == {1, 2, 3, 4}
    ts = rs = native.level_size(pi)
    path = str(tmp_path / "RestartForCase0000000006")
    R.write_restart(path, st, ct, ts, rs)
    st2, links = R.read_restart(path, ts, rs, np.asarray(sc["ptypes"]["diameter"]) / 2)
    for k in st:
        assert np.array_equal(np.asarray(st[k]).reshape(st2[k].shape), st2[k]), k
    walls, fixed = ids_of(sc)
    o = sc["options"]
    ct2 = R.add_restart_links(st2, links, fixed, walls, o["is_periodic"], o["domain_min"], o["domain_max"])
    assert record_set(ct2) == record_set(ct)
    # file size: header, per particle 3 + 2ts + 2rs + 1 real3 + 3 int + ncv,
    # two ints per link, one real4 per record
    nlinks = sum(len(l) for l in links)
    n, m = len(st["id"]), len(ct["kind"])
    assert os.path.getsize(path) == 8 + n * (24 * (2 * ts + 2 * rs + 2) + 12 + 4) + 8 * nlinks + 32 * m
    assert nlinks == m + int(np.sum(ct["kind"] == 1))


def test_oracle_resumes_from_restated_file(tmp_path, oracle_lib):
    """Oracle run of 5 + 7 steps against 5 steps, the restated file, a new
    oracle context from it (set_state / set_contacts) and 7 more steps: the
    restated reader loses nothing the step needs."""
    sc = case()
    a = native.build_engine(oracle_lib, sc, prefix=oracle.PREFIX)
    a.iterate(5)
    path = str(tmp_path / "RestartForCase0000000005")
    R.write_restart(path, a.get_state(), a.get_contacts(), 2, 2)
    a.iterate(7)
    st, links = R.read_restart(path, 2, 2, np.asarray(sc["ptypes"]["diameter"]) / 2)
    walls, fixed = ids_of(sc)
    o = sc["options"]
    ct = R.add_restart_links(st, links, fixed, walls, o["is_periodic"], o["domain_min"], o["domain_max"])
    s2 = restarted(sc, 5)
    s2["state"] = st
    s2["contacts"] = ct
    b = native.build_engine(oracle_lib, s2, prefix=oracle.PREFIX)
    b.iterate(7)
    errs = compare(b, a, exact_positions=False, rtol=1e-11, wall_cadence=False)
    assert not errs, errs
    assert a.get_iter()[0] == b.get_iter()[0]


def test_xdmf_text(tmp_path, hip_lib):
    """The library's XDMF index equals the restatement, whose key lines are
    checked against the Fortran formats written out by hand."""
    xmf = str(tmp_path / "PartVisuForRun.xmf")
    fields = {"ID", "Diameter", "LinVel", "Torque"}
    mask = sum(native.VISU[f] for f in fields)
    native.xdmf_begin(hip_lib, xmf, 1, 300, 100)
    native.xdmf_grid(hip_lib, xmf, "PartVisuForRun", 0, 1000, mask)
    native.xdmf_grid(hip_lib, xmf, "PartVisuForRun", 100, 1000, mask)
    native.xdmf_end(hip_lib, xmf)
    text = open(xmf).read()
    want = (R.xdmf_begin(1, 300, 100) + R.xdmf_grid("PartVisuForRun", 0, 1000, fields)
            + R.xdmf_grid("PartVisuForRun", 100, 1000, fields) + R.xdmf_end())
    assert text == want
    lines = text.splitlines()
    # '(A,I6,A)' and four I9 times on one line with the closing tag
    assert lines[6] == '            <DataItem Format="XML" NumberType="Int" Dimensions="     4">'
    assert lines[7] == " " * 12 + "        0      100      200      300" + " " * 12 + "</DataItem>"
    # '(A,I10.10,A)', '(A,I9,A)', '(A,I1,A,I2,I9,A,I15,A)'
    assert '        <Grid Name="T0000000100" GridType="Uniform">' in lines
    assert '            <Topology TopologyType="Polyvertex" NodesPerElement="     1000"/>' in lines
    assert ('                <DataItem Format="Binary" DataType="Float" Precision="8" Endian="Native"'
            ' Dimensions=" 3     1000" Seek="              0">') in lines
    # ID (int32) then Diameter (f64): the Diameter block starts after 24000 + 4000 bytes
    assert ('                <DataItem Format="Binary" DataType="Float" Precision="8" Endian="Native"'
            ' Dimensions=" 1     1000" Seek="          28000">') in lines
    assert "                    PartVisuForRun0000000100" in lines
    assert lines[-3:] == ["    </Grid>", "</Domain>", "</Xdmf>"]


# ----------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_library_restart_matches_restatement(tmp_path, hip_lib):
    print("restart: library writer vs restated reader", flush=True)
    sc = case()
    eng = native.build_engine(hip_lib, sc)
    eng.iterate(6)
    path = str(tmp_path / "RestartForCase0000000006")
    eng.write_restart(path)
    st, ct = eng.get_state(), eng.get_contacts()
    st2, links = R.read_restart(path, 2, 2, np.asarray(sc["ptypes"]["diameter"]) / 2)
    for k in st:
        assert np.array_equal(np.asarray(st[k]).reshape(st2[k].shape), st2[k]), k
    walls, fixed = ids_of(sc)
    o = sc["options"]
    ct2 = R.add_restart_links(st2, links, fixed, walls, o["is_periodic"], o["domain_min"], o["domain_max"])
    assert record_set(ct2) == record_set(ct)
    assert set(ct["kind"].tolist()) == {1, 2, 3, 4}


@pytest.mark.gpu
def test_library_reads_oracle_restart(tmp_path, hip_lib, oracle_lib):
    """The oracle runs 5 steps and the restated writer saves them; the
    library reads that file and runs 7 more; the oracle runs 12 straight."""
    print("restart: oracle file -> library -> 7 steps vs oracle 12", flush=True)
    sc = case()
    ref = native.build_engine(oracle_lib, sc, prefix=oracle.PREFIX)
    ref.iterate(5)
    path = str(tmp_path / "RestartForCase0000000005")
    R.write_restart(path, ref.get_state(), ref.get_contacts(), 2, 2)
    ref.iterate(7)
    gpu = native.build_engine(hip_lib, restarted(sc, 5))
    gpu.read_restart(path)
    gpu.iterate(7)
    errs = compare(gpu, ref, exact_positions=False, rtol=1e-10, wall_cadence=False)
    assert not errs, errs


@pytest.mark.gpu
def test_resume_from_library_restart_is_bitwise(tmp_path, hip_lib):
    print("restart: 8 + 9 steps vs 17 straight, bitwise", flush=True)
    sc = case(seed=17)
    a = native.build_engine(hip_lib, sc)
    a.iterate(8)
    path = str(tmp_path / "RestartForCase0000000008")
    a.write_restart(path)
    a.iterate(9)
    b = native.build_engine(hip_lib, restarted(sc, 8))
    b.read_restart(path)
    b.iterate(9)
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    assert record_set(a.get_contacts()) == record_set(b.get_contacts())
    assert a.get_iter()[0] == b.get_iter()[0]


@pytest.mark.gpu
def test_mark_links(tmp_path, hip_lib):
    """With mark_links the context's marks become the link counts after the
    write; the file keeps the marks from before (PIO_Write_Restart :979-984)."""
    print("restart: mark_links", flush=True)
    sc = case(seed=19)
    eng = native.build_engine(hip_lib, sc)
    eng.iterate(4)
    before = eng.get_state()["usr_mark"].copy()
    path = str(tmp_path / "RestartForCase0000000004")
    eng.write_restart(path, mark_links=True)
    st2, links = R.read_restart(path, 2, 2, np.asarray(sc["ptypes"]["diameter"]) / 2)
    assert np.array_equal(st2["usr_mark"], before)
    st = eng.get_state()
    after = st["usr_mark"]
    assert np.array_equal(after, np.array([len(l) for l in links]))
    want = dict(zip(st["id"].tolist(), after.tolist()))
    # the steps carry the new marks on, in either state buffer (odd and
    # even numbers of steps, most of them without a re-sort; a sphere that
    # leaves through a wall takes its mark with it)
    for k in (1, 2, 1, 3):
        eng.iterate(k)
        st = eng.get_state()
        assert [want[i] for i in st["id"].tolist()] == st["usr_mark"].tolist(), k


@pytest.mark.gpu
def test_more_than_20_links_is_refused(tmp_path, hip_lib):
    print("restart: 21 links", flush=True)
    sc = case(seed=23)
    eng = native.build_engine(hip_lib, sc)
    st = sc["state"]
    owner = int(st["id"][0])
    partners = [int(p) for p in st["id"] if p > owner][:21]
    assert len(partners) == 21
    eng.set_contacts([1] * 21, [owner] * 21, partners, np.zeros((21, 3)), np.full(21, -2.0))
    with pytest.raises(native.SdemError) as e:
        eng.write_restart(str(tmp_path / "r"))
    assert e.value.code == native.SDEM_E_CAPACITY
    assert not os.path.exists(str(tmp_path / "r"))


@pytest.mark.gpu
def test_visu_dump(tmp_path, hip_lib):
    print("visu dump: all fields", flush=True)
    sc = case(seed=29)
    eng = native.build_engine(hip_lib, sc, tuning=dict(keep_forces=1))
    eng.iterate(3)
    path = str(tmp_path / "PartVisuForCase0000000003")
    eng.write_visu(path, sum(native.VISU.values()))
    st = eng.get_state()
    ids, f, t = eng.get_forces()
    d = R.read_visu(path, len(st["id"]), set(R.VISU_ORDER))
    assert np.array_equal(d["Pos"], st["posr"][:, :3])
    assert np.array_equal(d["ID"], st["id"])
    assert np.array_equal(d["Diameter"], 2 * st["posr"][:, 3])
    assert np.array_equal(d["Type"], st["ptype"])
    assert np.array_equal(d["UsrMark"], st["usr_mark"])
    for name, key in (("LinVel", "lin_vel"), ("LinAcc", "lin_acc"), ("RotVel", "rot_vel"), ("RotAcc", "rot_acc")):
        assert np.array_equal(d[name], st[key][:, 0, :]), name
    assert np.array_equal(d["Theta"], st["theta"])
    assert np.array_equal(ids, st["id"])
    assert np.array_equal(d["CntctForce"], f) and np.array_equal(d["Torque"], t)
    # only positions and the type
    eng.write_visu(path, native.VISU["Type"])
    d = R.read_visu(path, len(st["id"]), {"Type"})
    assert np.array_equal(d["Type"], st["ptype"])


def write_fixed_coord(path, fx):
    """FixedSpheresCoord.dat as PIO_ReadFixedCoord reads it
    (Prtcl_IOAndVisu.f90:164-291): packed {f64 x, y, z, f64 diameter, int32 type}."""
    rec = np.dtype([("p", "<f8", 3), ("d", "<f8"), ("t", "<i4")])
    a = np.zeros(len(fx["id"]), rec)
    a["p"] = fx["posr"][:, :3]
    a["d"] = 2 * fx["posr"][:, 3]
    a["t"] = fx["ptype"]
    assert a.itemsize == 36
    a.tofile(path)


@pytest.mark.gpu
def test_fixed_spheres_file(tmp_path, hip_lib):
    """Fixed spheres read from FixedSpheresCoord.dat (ids numPrtcl + k) run
    bit for bit as the same spheres passed by sdem_set_fixed_particles; a
    diameter that does not match its type, or a file of the wrong size, is
    refused."""
    print("fixed spheres file", flush=True)
    sc = case(seed=31)
    fx = sc["fixed"]
    assert np.array_equal(fx["id"], sc["options"]["num_prtcl"] + 1 + np.arange(len(fx["id"])))
    path = str(tmp_path / "FixedSpheresCoord.dat")
    write_fixed_coord(path, fx)
    a = native.build_engine(hip_lib, sc)
    s2 = {k: v for k, v in sc.items() if k != "fixed"}
    s2["fixed_file"] = (path, len(fx["id"]))
    b = native.build_engine(hip_lib, s2)
    a.iterate(10)
    b.iterate(10)
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    assert record_set(a.get_contacts()) == record_set(b.get_contacts())
    assert (a.get_contacts()["kind"] == 3).any()
    bad = dict(fx, posr=fx["posr"].copy())
    bad["posr"][3, 3] *= 1.01
    write_fixed_coord(path, bad)
    s2["fixed_file"] = (path, len(fx["id"]))
    with pytest.raises(native.SdemError):
        native.build_engine(hip_lib, s2)
    write_fixed_coord(path, fx)
    s2["fixed_file"] = (path, len(fx["id"]) + 1)
    with pytest.raises(native.SdemError):
        native.build_engine(hip_lib, s2)
