"""The Fortran host (sdem_amd/fortran): the reference's .prm namelists and
call sequence driving libsdem_hip.so through ISO_C_BINDING."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from fortran_case import hourglass_walls, prm_text, read_dump, sandbox_prm_text, sandbox_scenario, scenario
from oracle import restart_format as R
from sdem_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "sdem_amd", "fortran")
EXE = os.path.join(FDIR, "sdem_main")


def driver():
    if not os.path.exists(EXE):
        if not shutil.which("amdflang"):
            pytest.skip("amdflang is not available and the driver is not built")
        subprocess.run(["make", "-s", "-C", FDIR], check=True)
    return EXE


def test_options_struct_matches_c():
    out = subprocess.run([driver(), "--abi-check"], capture_output=True, text=True, check=True).stdout
    assert "sizeof(sdem_options) = 152" in out


def test_prm_namelists_are_read(tmp_path):
    prm = tmp_path / "case.prm"
    prm.write_text(prm_text())
    out = subprocess.run([driver(), "--dry-run", str(prm)], capture_output=True, text=True, check=True).stdout
    assert "numPrtcl 1500 types 2 walls 2 steps 40" in out
    assert "periodic 1 0 1" in out and "methods 1 2 1 2 2" in out
    bins = [int(x) for x in out.split("bins")[1].split("\n")[0].split()]
    assert sum(bins) == 1500 and bins[0] == 1000 and bins[1] == 500


def dry_run_walls(prm):
    out = subprocess.run([driver(), "--dry-run", str(prm)], capture_output=True, text=True, check=True).stdout
    walls = [np.array([float(x) for x in line.split()[2:]]).reshape(4, 3)
             for line in out.splitlines() if line.startswith("wall ")]
    return out, walls


def test_program_geometry_is_the_reference_hourglass(tmp_path):
    """GeometrySource = 0: the driver builds the reference program's own 22
    walls (the fixture read out of Prtcl_Geometry.f90:557-615), corner for
    corner and in the reference's order."""
    prm = tmp_path / "sandbox.prm"
    prm.write_text(sandbox_prm_text())
    out, walls = dry_run_walls(prm)
    want = hourglass_walls()
    assert len(walls) == len(want) == 22
    for got, w in zip(walls, want):
        assert np.array_equal(got, np.array([w["p1"], w["p2"], w["p3"], w["p4"]]))


def test_reference_sandbox_input_is_read(tmp_path):
    """The reference's own Input/SandBox_DEM.prm (GeometrySource = 0), where
    the reference tree is at hand: 8000 spheres, the hourglass walls."""
    ref = "/root/reference/Input/SandBox_DEM.prm"
    if not os.path.exists(ref):
        pytest.skip("the reference tree is not here")
    prm = tmp_path / "SandBox_DEM.prm"
    shutil.copy(ref, prm)
    out, walls = dry_run_walls(prm)
    assert "numPrtcl 8000 types 1 walls 22" in out
    assert len(walls) == 22


@pytest.mark.gpu
@pytest.mark.parametrize("cf_type", [native.CFT_NLIN, native.CFT_LINEAR])
def test_fortran_run_matches_oracle(tmp_path, cf_type, oracle_lib):
    """40 steps driven from Fortran on the GPU against the oracle from the
    same initial state (dumped by the driver before the first step)."""
    print(f"fortran driver: cf_type {cf_type}, 1500 spheres, 40 steps", flush=True)
    prm = tmp_path / "case.prm"
    prm.write_text(prm_text(cf_type=cf_type))
    run = subprocess.run([driver(), str(prm), str(tmp_path)], capture_output=True, text=True, timeout=120,
                         cwd=str(tmp_path))
    assert run.returncode == 0, run.stdout + run.stderr
    assert "Good job" in run.stdout
    s0 = read_dump(str(tmp_path / "state_0000000000.bin"))
    s1 = read_dump(str(tmp_path / "state_0000000040.bin"))
    ref = native.build_engine(oracle_lib, scenario(s0, cf_type=cf_type), prefix=oracle.PREFIX)
    ref.iterate(40)
    r = ref.get_state()
    assert np.array_equal(s1["id"], r["id"])
    for k in ("posr", "lin_vel", "lin_acc", "theta", "rot_vel", "rot_acc"):
        scale = np.abs(r[k]).max() + 1e-300
        err = np.abs(s1[k] - r[k]) / (np.abs(r[k]) + scale)
        assert err.max() < 1e-10, (k, err.max())


def run_driver(d, name, text, *args):
    prm = d / name
    prm.write_text(text)
    run = subprocess.run([driver(), str(prm), *args], capture_output=True, text=True, timeout=120, cwd=str(d))
    assert run.returncode == 0, run.stdout + run.stderr
    assert "Good job" in run.stdout
    return run


@pytest.mark.gpu
def test_fortran_restart_and_visu(tmp_path):
    """The reference's restart and visualisation cycle through the Fortran
    host: 40 steps straight with backups every 20 and visualisation dumps
    every 20 against 20 steps, then RestartFlag = T from the file of step
    20 for steps 21-40.  The resumed run ends in the same state, bit for bit
    (the marks aside: the backup at step 20 turns the running context's
    marks into link counts, the file keeps the old ones, as in the
    reference); the dumps and the XDMF index read back as the reference
    lays them out."""
    print("fortran driver: restart at 20 and visu dumps", flush=True)
    visu = dict(saveXDMFOnce=False, save_ID=True, save_Diameter=False, save_Type=True, save_UsrMark=False,
                save_LinVel=True, save_LinAcc=False, save_Theta=False, save_RotVel=True, save_RotAcc=False,
                save_CntctForce=True, save_Torque=False)
    a, b = tmp_path / "straight", tmp_path / "resumed"
    a.mkdir()
    b.mkdir()
    run_driver(a, "case.prm", prm_text(ilast=40, save_visu=20, backup=20, visu=visu), str(a))
    run_driver(b, "first.prm", prm_text(ilast=20, backup=20))
    assert (b / "RestartForSettling0000000020").exists()
    run_driver(b, "second.prm", prm_text(restart=True, ifirst=21, ilast=40, backup=20), str(b))
    sa = read_dump(str(a / "state_0000000040.bin"))
    sb = read_dump(str(b / "state_0000000040.bin"))
    for k in sa:
        if k != "usr_mark":
            assert np.array_equal(sa[k], sb[k]), k
    # only the newest restart file stays (PIO_Delete_Prev_Restart)
    assert (a / "RestartForSettling0000000040").exists()
    assert not (a / "RestartForSettling0000000020").exists()
    # restart file of step 40 against the state dumped at 40
    st, links = R.read_restart(str(a / "RestartForSettling0000000040"), 2, 2, [1.0e-3, 1.5e-3])
    for k in ("id", "ptype", "lin_vel", "lin_acc", "theta", "rot_vel", "rot_acc"):
        assert np.array_equal(st[k], sa[k]), k
    assert np.array_equal(st["posr"], sa["posr"])
    # visualisation dumps of steps 0, 20, 40 and their index
    fields = {"ID", "Type", "LinVel", "RotVel", "CntctForce"}
    s0 = read_dump(str(a / "state_0000000000.bin"))
    for it, s in ((0, s0), (40, sa)):
        d = R.read_visu(str(a / f"PartVisuForSettling{it:010d}"), len(s["id"]), fields)
        assert np.array_equal(d["Pos"], s["posr"][:, :3]) and np.array_equal(d["ID"], s["id"])
        assert np.array_equal(d["LinVel"], s["lin_vel"][:, 0, :]) and np.array_equal(d["Type"], s["ptype"])
    assert (a / "PartVisuForSettling0000000020").exists()
    n = len(sa["id"])
    want = (R.xdmf_begin(1, 40, 20) + "".join(R.xdmf_grid("PartVisuForSettling", it, n, fields)
                                               for it in (0, 20, 40)) + R.xdmf_end())
    assert (a / "PartVisuForSettling.xmf").read_text() == want


@pytest.mark.gpu
def test_fortran_phase_by_phase_and_timers(tmp_path):
    """The driver's step as the reference's calls one at a time
    (SDEM_PHASES=1: sdem_pre_iteration ... sdem_remv_released) ends in the
    state of the sdem_step run bit for bit, and SDEM_TIMERS=1 prints the
    reference's step timer lines."""
    print("fortran driver: phase by phase against sdem_step, 40 steps", flush=True)
    a, b = tmp_path / "step", tmp_path / "phases"
    a.mkdir()
    b.mkdir()
    run_driver(a, "case.prm", prm_text(), str(a))
    prm = b / "case.prm"
    prm.write_text(prm_text())
    env = dict(os.environ, SDEM_PHASES="1", SDEM_TIMERS="1")
    run = subprocess.run([driver(), str(prm), str(b)], capture_output=True, text=True, timeout=120, cwd=str(b),
                         env=env)
    assert run.returncode == 0, run.stdout + run.stderr
    sa = read_dump(str(a / "state_0000000040.bin"))
    sb = read_dump(str(b / "state_0000000040.bin"))
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    for label in ("Execution time [tot, ave] [sec]:", "PreItertion time [tot, ave]", "CS and CF P-P time",
                  "CS and CF P-W time", "Acceleration time", "Integration time", "Comm_For_Exchange",
                  "Write to file time"):
        assert label in run.stdout, label


@pytest.mark.gpu
def test_fortran_hourglass_matches_oracle(tmp_path, oracle_lib):
    """The hourglass case (GeometrySource = 0, the sand box's materials, AB3)
    driven from Fortran on the GPU against the oracle from the same initial
    state: the neck's finite walls and the edges they share with the
    funnels are in contact."""
    print("fortran driver: hourglass (GeometrySource 0), 1200 spheres, 300 steps", flush=True)
    prm = tmp_path / "sandbox.prm"
    prm.write_text(sandbox_prm_text())
    run = subprocess.run([driver(), str(prm), str(tmp_path)], capture_output=True, text=True, timeout=120,
                         cwd=str(tmp_path))
    assert run.returncode == 0, run.stdout + run.stderr
    s0 = read_dump(str(tmp_path / "state_0000000000.bin"))
    s1 = read_dump(str(tmp_path / "state_0000000300.bin"))
    ref = native.build_engine(oracle_lib, sandbox_scenario(s0), prefix=oracle.PREFIX)
    pw = 0
    for _ in range(30):
        ref.iterate(10)
        pw += int(ref.get_counts()[3])
    print(f"  wall records summed over 30 samples: {pw}", flush=True)
    assert pw > 50
    r = ref.get_state()
    assert np.array_equal(s1["id"], r["id"])
    for k in ("posr", "lin_vel", "lin_acc", "theta", "rot_vel", "rot_acc"):
        scale = np.abs(r[k]).max() + 1e-300
        err = np.abs(s1[k] - r[k]) / (np.abs(r[k]) + scale)
        assert err.max() < 1e-10, (k, err.max())
