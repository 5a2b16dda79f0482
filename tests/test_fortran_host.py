"""The Fortran host (sdem_amd/fortran): the reference's .prm namelists and
call sequence driving libsdem_hip.so through ISO_C_BINDING."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from fortran_case import prm_text, read_dump, scenario
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


@pytest.mark.gpu
@pytest.mark.parametrize("cf_type", [native.CFT_NLIN, native.CFT_LINEAR])
def test_fortran_run_matches_oracle(tmp_path, cf_type, oracle_lib):
    """40 steps driven from Fortran on the GPU against the oracle from the
    same initial state (dumped by the driver before the first step)."""
    print(f"fortran driver: cf_type {cf_type}, 1500 spheres, 40 steps", flush=True)
    prm = tmp_path / "case.prm"
    prm.write_text(prm_text(cf_type=cf_type))
    run = subprocess.run([driver(), str(prm), str(tmp_path)], capture_output=True, text=True, timeout=120)
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
