"""The C-ABI library loads and exports every entry point include/sdem.h
declares; argument checks that need no device behave as documented."""
import ctypes
import os
import re

import pytest

from sdem_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdrs = [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))
            if f.endswith(".h")]
    names = set()
    for h in hdrs:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(sdem\w*)\s*\(", src))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    for n in ("sdem_create", "sdem_step", "sdem_set_state", "sdem_get_contacts", "sdem_destroy"):
        assert n in names


def test_library_exports_every_declared_symbol(hip_lib):
    missing = [n for n in declared_symbols() if not hasattr(hip_lib, n)]
    assert not missing, missing


def test_options_default_matches_reference_defaults(hip_lib):
    o = native.default_options(hip_lib)
    # DEMS_Options defaults, Prtcl_Parameters.f90:45-95
    assert o.dt == 1.0e-5 and tuple(o.gravity) == (0.0, -9.81, 0.0)
    assert o.cntct_list_size == 6 and o.base_wall_id == 10000000
    assert o.wall_max_update_iter == 100 and o.wall_neighbor_ratio == 2.0
    assert (o.cs_method, o.cf_type, o.ct_model, o.pi_method, o.pri_method) == (1, 1, 1, 1, 1)


def test_options_layout_matches_c_struct():
    # doubles first (12 x 8 bytes), then 14 int32: no padding anywhere
    assert ctypes.sizeof(native.Options) == 12 * 8 + 14 * 4


def test_create_rejects_bad_options(hip_lib):
    o = native.default_options(hip_lib)
    o.pi_method = 7
    h = ctypes.c_void_p()
    assert hip_lib.sdem_create(ctypes.byref(o), ctypes.byref(h)) == -1


def test_null_context_is_rejected(hip_lib):
    assert hip_lib.sdem_step(None, 1) == -1
    n = ctypes.c_int64(7)
    assert hip_lib.sdem_get_rebuilds(None, ctypes.byref(n)) == -1
    assert hip_lib.sdem_last_error(None) == b"null context"


def test_cell_factors_are_three_per_call():
    """sdem_set_tuning reads three cell factors: the binding broadcasts one
    number to every axis and refuses any other length rather than hand the
    library a shorter array."""
    assert native.cell_factors(0.5).tolist() == [0.5, 0.5, 0.5]
    assert native.cell_factors((0.5, 0.75, 1.0)).tolist() == [0.5, 0.75, 1.0]
    for bad in ((1.0, 2.0), (1.0, 2.0, 3.0, 4.0), [[1.0, 2.0, 3.0]]):
        with pytest.raises(ValueError):
            native.cell_factors(bad)
