"""ctypes binding of the sDEM step C ABI (include/sdem.h).

`Engine` drives one context of a library that exports the entry points of
include/sdem.h under some symbol prefix.  The product library is
``libsdem_hip.so`` with prefix ``sdem_``; `load_hip_library` finds it in the
package directory and raises if it has not been built -- there is no other
implementation to fall back on.

The method names mirror the reference's procedures where one exists
(Prtcl_DEMSystem.f90, Prtcl_Property.f90, Prtcl_Geometry.f90).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libsdem_hip.so"

SDEM_OK = 0
SDEM_E_ARG, SDEM_E_HIP, SDEM_E_CAPACITY, SDEM_E_STATE, SDEM_E_NODEV = -1, -2, -3, -4, -5
ERRORS = {-1: "SDEM_E_ARG", -2: "SDEM_E_HIP", -3: "SDEM_E_CAPACITY", -4: "SDEM_E_STATE",
          -5: "SDEM_E_NODEV"}

CSM_NBS_MUNJIZA, CSM_NBS_MUNJIZA_HRCHL = 1, 2
CFT_LSD, CFT_NLIN, CFT_LINEAR = 1, 2, 3
CTM_CONSTANT_TORQUE, CTM_VARIABLE_TORQUE = 1, 2
PIM_FE, PIM_AB2, PIM_AB3 = 1, 2, 3
KIND_PP, KIND_PPG, KIND_PPFIX, KIND_PW = 1, 2, 3, 4
PHASE_NAMES = ("control", "bin_sort", "search", "contact", "force")
# the calls of DEMS_iterate (Prtcl_DEMSystem.f90:213-248), one entry point each
STEP_PHASES = ("pre_iteration", "comm_for_cntct", "find_contacts", "find_contacts_pw", "clc_acceleration",
               "integrate", "comm_for_exchange", "remv_released")
# the reference's step timers (Prtcl_DEMSystem.f90:29-37), sdem_get_timers order
TIMER_NAMES = ("total", "pre_iteration", "comm_for_contact", "cs_and_cf_pp", "cs_and_cf_pw", "acceleration",
               "integration", "comm_for_exchange", "write_to_file")


class SdemError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where}: {ERRORS.get(code, code)}: {msg}")
        self.code = code


class Options(C.Structure):
    """sdem_options (include/sdem.h); fields of DEMS_Options,
    Prtcl_Parameters.f90:45-95."""

    _fields_ = [
        ("dt", C.c_double),
        ("gravity", C.c_double * 3),
        ("domain_min", C.c_double * 3),
        ("domain_max", C.c_double * 3),
        ("prtcl_cs_ratio", C.c_double),
        ("wall_neighbor_ratio", C.c_double),
        ("is_periodic", C.c_int32 * 3),
        ("cs_method", C.c_int32),
        ("cf_type", C.c_int32),
        ("ct_model", C.c_int32),
        ("pi_method", C.c_int32),
        ("pri_method", C.c_int32),
        ("cs_numlvls", C.c_int32),
        ("cntct_list_size", C.c_int32),
        ("base_wall_id", C.c_int32),
        ("wall_max_update_iter", C.c_int32),
        ("num_prtcl", C.c_int32),
        ("ifirst", C.c_int32),
    ]


def level_size(method: int) -> int:
    return {PIM_FE: 1, PIM_AB2: 2, PIM_AB3: 3}[method]


def load_hip_library(path: Optional[str] = None) -> C.CDLL:
    """Load the built HIP library; raise if it is missing."""
    path = path or os.path.join(HERE, LIB_NAME)
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} not found: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    return C.CDLL(path)


_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
_LP = C.POINTER(C.c_int64)


def _declare(lib: C.CDLL, prefix: str) -> None:
    f = lambda name: getattr(lib, prefix + name)  # noqa: E731
    V = C.c_void_p
    sig = {
        "options_default": ([C.POINTER(Options)], C.c_int),
        "create": ([C.POINTER(Options), C.POINTER(V)], C.c_int),
        "set_particle_property": ([V, C.c_int32, _DP, _DP, _DP, _DP] + [C.c_double] * 5, C.c_int),
        "set_wall_property": ([V, C.c_int32, _DP, _DP] + [C.c_double] * 5, C.c_int),
        "set_linear_stiffness": ([V, C.c_double, C.c_double], C.c_int),
        "add_plane_wall": ([V, _DP, _DP, _DP, _DP, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _DP, _IP],
                           C.c_int),
        "set_fixed_particles": ([V, C.c_int32, _IP, _IP, _DP], C.c_int),
        "set_state": ([V, C.c_int32, _IP, _IP, _IP, _DP, _DP, _DP, _DP, _DP, _DP], C.c_int),
        "set_contacts": ([V, C.c_int64, _IP, _IP, _IP, _DP, _DP], C.c_int),
        "step": ([V, C.c_int32], C.c_int),
        "get_num_particles": ([V, _IP], C.c_int),
        "get_state": ([V, C.c_int32, _IP, _IP, _IP, _IP, _DP, _DP, _DP, _DP, _DP, _DP], C.c_int),
        "get_forces": ([V, C.c_int32, _IP, _DP, _DP], C.c_int),
        "get_num_contacts": ([V, _LP], C.c_int),
        "get_contacts": ([V, C.c_int64, _LP, _IP, _IP, _IP, _DP, _DP], C.c_int),
        "get_counts": ([V, _IP], C.c_int),
        "get_iter": ([V, _IP, _IP], C.c_int),
        "last_error": ([V], C.c_char_p),
        "destroy": ([V], None),
    }
    optional = {
        "set_profiling": ([V, C.c_int32], C.c_int),
        "get_timers": ([V, _DP, _LP], C.c_int),
        "get_profile": ([V, _DP, _LP], C.c_int),
        "get_device_bytes": ([V, _LP], C.c_int),
        "get_rebuilds": ([V, _LP], C.c_int),
        "set_tuning": ([V, _DP, C.c_double, C.c_int32], C.c_int),
        "set_capacity": ([V, C.c_int32, C.c_int32], C.c_int),
        "create_ensemble": ([C.c_int32, C.POINTER(Options), C.POINTER(V)], C.c_int),
        "get_num_boxes": ([V, _IP], C.c_int),
        "select_box": ([V, C.c_int32], C.c_int),
        "set_block_sizes": ([V, C.c_int32, C.c_int32], C.c_int),
        "set_force_build": ([V, C.c_int32], C.c_int),
        "get_tuning": ([V, _DP, _DP, _IP, _IP], C.c_int),
        "autotune": ([V, C.c_int32, _IP, _DP], C.c_int),
        "write_restart": ([V, C.c_char_p, C.c_int32], C.c_int),
        "read_restart": ([V, C.c_char_p], C.c_int),
        "write_visu": ([V, C.c_char_p, C.c_uint32], C.c_int),
        "read_fixed_coord": ([V, C.c_char_p, C.c_int32], C.c_int),
        "xdmf_begin": ([C.c_char_p, C.c_int32, C.c_int32, C.c_int32], C.c_int),
        "xdmf_grid": ([C.c_char_p, C.c_char_p, C.c_int32, C.c_int32, C.c_uint32], C.c_int),
        "xdmf_end": ([C.c_char_p], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = f(name)
        fn.argtypes, fn.restype = args, res
    for name in STEP_PHASES:
        optional[name] = ([V], C.c_int)
    for name, (args, res) in optional.items():
        if hasattr(lib, prefix + name):
            fn = f(name)
            fn.argtypes, fn.restype = args, res


def _dptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_DP)


def _iptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_IP)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def cell_factors(cell_factor) -> np.ndarray:
    """The three per-axis cell factors sdem_set_tuning reads: one number for
    every axis, or one per axis (any other length is an error, not a read
    past the array)."""
    a = _f64(cell_factor)
    if a.ndim > 1 or a.size not in (1, 3):
        raise ValueError(f"cell_factor: one number or three, not {a.size}")
    return np.ascontiguousarray(np.broadcast_to(a.reshape(-1), (3,)), dtype=np.float64)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


class Engine:
    """One context of a library exporting the include/sdem.h entry points."""

    def __init__(self, lib: C.CDLL, options, prefix: str = "sdem_"):
        """options: one Options (a plain context) or a sequence of them (an
        ensemble of independent boxes, sdem_create_ensemble)."""
        self.lib, self.prefix = lib, prefix
        if not getattr(lib, "_sdem_declared_" + prefix, False):
            _declare(lib, prefix)
            setattr(lib, "_sdem_declared_" + prefix, True)
        many = not isinstance(options, Options)
        first = options[0] if many else options
        self.opts = first
        self.tsize = level_size(first.pi_method)
        self.rsize = level_size(first.pri_method)
        h = C.c_void_p()
        if many:
            arr = (Options * len(options))(*options)
            rc = self._fn("create_ensemble")(len(options), arr, C.byref(h))
        else:
            rc = self._fn("create")(C.byref(options), C.byref(h))
        if rc != SDEM_OK:
            raise SdemError(rc, prefix + "create", "context creation failed")
        self.h = h
        self.nbox = len(options) if many else 1

    # ------------------------------------------------------------------ util
    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def _check(self, rc: int, where: str) -> None:
        if rc != SDEM_OK:
            msg = self._fn("last_error")(self.h)
            raise SdemError(rc, self.prefix + where, msg.decode() if msg else "")

    def close(self) -> None:
        if getattr(self, "h", None):
            self._fn("destroy")(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------- set-up
    def set_particle_property(self, diameter, density, youngs, poisson, fric_s, fric_k, fric_roll,
                              rest_n, rest_t) -> None:
        d, rho, e, nu = map(_f64, (diameter, density, youngs, poisson))
        self._check(self._fn("set_particle_property")(
            self.h, len(d), _dptr(d), _dptr(rho), _dptr(e), _dptr(nu),
            fric_s, fric_k, fric_roll, rest_n, rest_t), "set_particle_property")

    def set_linear_stiffness(self, kn: float, kt_over_kn: float) -> None:
        self._check(self._fn("set_linear_stiffness")(self.h, kn, kt_over_kn), "set_linear_stiffness")

    def set_wall_property(self, youngs, poisson, fric_s, fric_k, fric_roll, rest_n, rest_t) -> None:
        e, nu = _f64(youngs), _f64(poisson)
        self._check(self._fn("set_wall_property")(
            self.h, len(e), _dptr(e), _dptr(nu), fric_s, fric_k, fric_roll, rest_n, rest_t),
            "set_wall_property")

    def add_plane_wall(self, p1, p2, p3, p4, user_id=1, wall_type=1, both_side=False,
                       is_infinite=True, trans_vel=(0.0, 0.0, 0.0)) -> int:
        pts = [_f64(p) for p in (p1, p2, p3, p4)]
        tv = _f64(trans_vel)
        wid = C.c_int32(0)
        self._check(self._fn("add_plane_wall")(
            self.h, *[_dptr(p) for p in pts], int(user_id), int(wall_type), int(bool(both_side)),
            int(bool(is_infinite)), _dptr(tv), C.byref(wid)), "add_plane_wall")
        return wid.value

    def set_fixed_particles(self, ids, ptype, posr) -> None:
        ids, ptype, posr = _i32(ids), _i32(ptype), _f64(posr).reshape(-1, 4)
        self._check(self._fn("set_fixed_particles")(
            self.h, len(ids), _iptr(ids), _iptr(ptype), _dptr(posr)), "set_fixed_particles")

    def set_state(self, state: Dict[str, np.ndarray]) -> None:
        n = len(state["id"])
        ids, pt = _i32(state["id"]), _i32(state["ptype"])
        mark = _i32(state["usr_mark"]) if state.get("usr_mark") is not None else None
        posr = _f64(state["posr"]).reshape(n, 4)
        lv = _f64(state["lin_vel"]).reshape(n, self.tsize, 3)
        la = _f64(state["lin_acc"]).reshape(n, self.tsize, 3)
        th = _f64(state["theta"]).reshape(n, 3)
        rv = _f64(state["rot_vel"]).reshape(n, self.rsize, 3)
        ra = _f64(state["rot_acc"]).reshape(n, self.rsize, 3)
        self._check(self._fn("set_state")(
            self.h, n, _iptr(ids), _iptr(pt), _iptr(mark), _dptr(posr), _dptr(lv), _dptr(la),
            _dptr(th), _dptr(rv), _dptr(ra)), "set_state")

    def set_contacts(self, kind, owner, partner, tan_delta, vrel_init) -> None:
        kind, owner, partner = _i32(kind), _i32(owner), _i32(partner)
        td, vr = _f64(tan_delta).reshape(-1, 3), _f64(vrel_init)
        self._check(self._fn("set_contacts")(
            self.h, len(kind), _iptr(kind), _iptr(owner), _iptr(partner), _dptr(td), _dptr(vr)),
            "set_contacts")

    # ---------------------------------------------------------------- step
    def iterate(self, nsteps: int = 1) -> None:
        """nsteps x DEMS_iterate (Prtcl_DEMSystem.f90:201-248)."""
        self._check(self._fn("step")(self.h, int(nsteps)), "step")

    def phase(self, name: str) -> None:
        """One call of DEMS_iterate (STEP_PHASES; sdem_<name>)."""
        self._check(self._fn(name)(self.h), name)

    def iterate_phases(self, nsteps: int = 1) -> None:
        """nsteps steps, each as its eight phase calls in the reference's
        order; a step whose search or forces ran out of capacity is run again
        from its start (the library has grown the capacity)."""
        for _ in range(int(nsteps)):
            for attempt in range(13):
                try:
                    for name in STEP_PHASES:
                        self.phase(name)
                    break
                except SdemError as err:
                    if err.code != SDEM_E_CAPACITY or attempt == 12:
                        raise

    def get_timers(self) -> Dict[str, tuple]:
        """The reference's step timers: name -> (total ms, count)."""
        ms = np.zeros(len(TIMER_NAMES))
        n = np.zeros(len(TIMER_NAMES), np.int64)
        self._check(self._fn("get_timers")(self.h, _dptr(ms), n.ctypes.data_as(_LP)), "get_timers")
        return {name: (float(ms[i]), int(n[i])) for i, name in enumerate(TIMER_NAMES)}

    # ------------------------------------------------------------- readers
    def num_particles(self) -> int:
        n = C.c_int32(0)
        self._check(self._fn("get_num_particles")(self.h, C.byref(n)), "get_num_particles")
        return n.value

    def get_state(self) -> Dict[str, np.ndarray]:
        cap = self.num_particles()
        ts, rs = self.tsize, self.rsize
        out = dict(id=np.zeros(cap, np.int32), ptype=np.zeros(cap, np.int32),
                   usr_mark=np.zeros(cap, np.int32), posr=np.zeros((cap, 4)),
                   lin_vel=np.zeros((cap, ts, 3)), lin_acc=np.zeros((cap, ts, 3)),
                   theta=np.zeros((cap, 3)), rot_vel=np.zeros((cap, rs, 3)),
                   rot_acc=np.zeros((cap, rs, 3)))
        n = C.c_int32(0)
        self._check(self._fn("get_state")(
            self.h, cap, C.byref(n), _iptr(out["id"]), _iptr(out["ptype"]), _iptr(out["usr_mark"]),
            _dptr(out["posr"]), _dptr(out["lin_vel"]), _dptr(out["lin_acc"]), _dptr(out["theta"]),
            _dptr(out["rot_vel"]), _dptr(out["rot_acc"])), "get_state")
        return out

    def get_forces(self):
        cap = self.num_particles()
        ids, f, t = np.zeros(cap, np.int32), np.zeros((cap, 3)), np.zeros((cap, 3))
        self._check(self._fn("get_forces")(self.h, cap, _iptr(ids), _dptr(f), _dptr(t)), "get_forces")
        return ids, f, t

    def get_contacts(self) -> Dict[str, np.ndarray]:
        m = C.c_int64(0)
        self._check(self._fn("get_num_contacts")(self.h, C.byref(m)), "get_num_contacts")
        cap = m.value
        out = dict(kind=np.zeros(cap, np.int32), owner=np.zeros(cap, np.int32),
                   partner=np.zeros(cap, np.int32), tan_delta=np.zeros((cap, 3)),
                   vrel_init=np.zeros(cap))
        self._check(self._fn("get_contacts")(
            self.h, cap, C.byref(m), _iptr(out["kind"]), _iptr(out["owner"]), _iptr(out["partner"]),
            _dptr(out["tan_delta"]), _dptr(out["vrel_init"])), "get_contacts")
        return out

    # --------------------------------------------- files (include/sdem_io.h)
    def write_restart(self, path: str, mark_links: bool = False) -> None:
        """PIO_Write_Restart: the reference's restart file of this box."""
        self._check(self._fn("write_restart")(self.h, os.fsencode(path), int(mark_links)), "write_restart")

    def read_restart(self, path: str) -> None:
        """PIO_Read_Restart + PIO_RestartCL: state and contact records."""
        self._check(self._fn("read_restart")(self.h, os.fsencode(path)), "read_restart")

    def read_fixed_coord(self, path: str, num_fix: int) -> None:
        """PIO_ReadFixedCoord: fixed spheres from FixedSpheresCoord.dat."""
        self._check(self._fn("read_fixed_coord")(self.h, os.fsencode(path), int(num_fix)), "read_fixed_coord")

    def write_visu(self, path: str, fields: int = 0) -> None:
        """PIO_Dump_visu: positions and the VISU_* fields selected."""
        self._check(self._fn("write_visu")(self.h, os.fsencode(path), int(fields)), "write_visu")

    def get_counts(self) -> np.ndarray:
        c = np.zeros(4, np.int32)
        self._check(self._fn("get_counts")(self.h, _iptr(c)), "get_counts")
        return c

    def get_iter(self):
        it, nx = C.c_int32(0), C.c_int32(0)
        self._check(self._fn("get_iter")(self.h, C.byref(it), C.byref(nx)), "get_iter")
        return it.value, nx.value

    def get_rebuilds(self) -> int:
        """Re-sorts and candidate-list builds so far, summed over the boxes."""
        n = C.c_int64(0)
        self._check(self._fn("get_rebuilds")(self.h, C.byref(n)), "get_rebuilds")
        return n.value

    # ---------------------------------------------------------- profiling
    def set_profiling(self, enable: bool) -> None:
        self._check(self._fn("set_profiling")(self.h, int(bool(enable))), "set_profiling")

    def set_tuning(self, cell_factor=None, skin=-1.0, keep_forces=-1) -> None:
        cf = None if cell_factor is None else cell_factors(cell_factor)
        self._check(self._fn("set_tuning")(self.h, _dptr(cf), float(skin), int(keep_forces)),
                    "set_tuning")

    def select(self, box: int) -> None:
        """Address box `box` of an ensemble with the setters and readers."""
        self._check(self._fn("select_box")(self.h, int(box)), "select_box")

    def set_force_build(self, waves_per_simd: int = 0) -> None:
        """The force kernel's build: 0 chosen by contact density, 2 or 3 waves per SIMD."""
        self._check(self._fn("set_force_build")(self.h, int(waves_per_simd)), "set_force_build")

    def set_block_sizes(self, contact_block=256, force_block=256) -> None:
        self._check(self._fn("set_block_sizes")(self.h, int(contact_block), int(force_block)),
                    "set_block_sizes")

    def get_tuning(self) -> dict:
        cf = np.zeros(3)
        sk = C.c_double(0)
        bc, bf = C.c_int32(0), C.c_int32(0)
        self._check(self._fn("get_tuning")(self.h, _dptr(cf), C.byref(sk), C.byref(bc), C.byref(bf)),
                    "get_tuning")
        return dict(cell_factor=tuple(cf), skin=sk.value, contact_block=bc.value, force_block=bf.value)

    def autotune(self, steps_per_trial: int = 10):
        """Try the candidate tunings on real steps, keep the fastest; returns
        (steps run, best ms per step)."""
        n, ms = C.c_int32(0), C.c_double(0)
        self._check(self._fn("autotune")(self.h, int(steps_per_trial), C.byref(n), C.byref(ms)), "autotune")
        return n.value, ms.value

    def set_capacity(self, hist_slots=0, list_slots=0) -> None:
        self._check(self._fn("set_capacity")(self.h, int(hist_slots), int(list_slots)), "set_capacity")

    def get_profile(self):
        ms = np.zeros(len(PHASE_NAMES))
        n = np.zeros(len(PHASE_NAMES), np.int64)
        self._check(self._fn("get_profile")(self.h, _dptr(ms), n.ctypes.data_as(_LP)), "get_profile")
        return {name: (float(ms[i]), int(n[i])) for i, name in enumerate(PHASE_NAMES)}

    def device_bytes(self) -> int:
        b = C.c_int64(0)
        self._check(self._fn("get_device_bytes")(self.h, C.byref(b)), "get_device_bytes")
        return b.value


def default_options(lib: C.CDLL, prefix: str = "sdem_") -> Options:
    if not getattr(lib, "_sdem_declared_" + prefix, False):
        _declare(lib, prefix)
        setattr(lib, "_sdem_declared_" + prefix, True)
    o = Options()
    getattr(lib, prefix + "options_default")(C.byref(o))
    return o


def make_options(lib: C.CDLL, prefix: str = "sdem_", **kw) -> Options:
    o = default_options(lib, prefix)
    for k, v in kw.items():
        if k in ("gravity", "domain_min", "domain_max"):
            arr = getattr(o, k)
            for i in range(3):
                arr[i] = float(v[i])
        elif k == "is_periodic":
            for i in range(3):
                o.is_periodic[i] = int(bool(v[i]))
        else:
            setattr(o, k, v)
    return o


def _push(eng: Engine, scenario: dict) -> None:
    """Properties, walls, fixed particles, state and contacts of one box, in
    the reference's initialisation order (DEMS_Initialize :58-195)."""
    if "linear" in scenario:
        eng.set_linear_stiffness(*scenario["linear"])
    pt = scenario["ptypes"]
    eng.set_particle_property(pt["diameter"], pt["density"], pt["youngs"], pt["poisson"], *scenario["pp"])
    wt = scenario["wtypes"]
    eng.set_wall_property(wt["youngs"], wt["poisson"], *scenario["pw"])
    for w in scenario.get("walls", []):
        eng.add_plane_wall(**w)
    fx = scenario.get("fixed")
    if fx is not None and len(fx["id"]):
        eng.set_fixed_particles(fx["id"], fx["ptype"], fx["posr"])
    if scenario.get("fixed_file"):  # (path of FixedSpheresCoord.dat, numPrtclFix)
        eng.read_fixed_coord(*scenario["fixed_file"])
    eng.set_state(scenario["state"])
    ct = scenario.get("contacts")
    if ct is not None and len(ct["kind"]):
        eng.set_contacts(ct["kind"], ct["owner"], ct["partner"], ct["tan_delta"], ct["vrel_init"])


# fields of the visualisation dump (include/sdem_io.h), in file order
VISU_FIELDS = ("ID", "Diameter", "Type", "UsrMark", "LinVel", "LinAcc", "Theta", "RotVel", "RotAcc",
               "CntctForce", "Torque")
VISU = {name: 1 << k for k, name in enumerate(VISU_FIELDS)}


def xdmf_begin(lib: C.CDLL, xmf: str, ifirst: int, ilast: int, save_visu: int) -> None:
    _declare(lib, "sdem_")
    rc = lib.sdem_xdmf_begin(os.fsencode(xmf), int(ifirst), int(ilast), int(save_visu))
    if rc != SDEM_OK:
        raise SdemError(rc, "sdem_xdmf_begin", xmf)


def xdmf_grid(lib: C.CDLL, xmf: str, data_base: str, itime: int, np_: int, fields: int) -> None:
    _declare(lib, "sdem_")
    rc = lib.sdem_xdmf_grid(os.fsencode(xmf), os.fsencode(data_base), int(itime), int(np_), int(fields))
    if rc != SDEM_OK:
        raise SdemError(rc, "sdem_xdmf_grid", xmf)


def xdmf_end(lib: C.CDLL, xmf: str) -> None:
    _declare(lib, "sdem_")
    rc = lib.sdem_xdmf_end(os.fsencode(xmf))
    if rc != SDEM_OK:
        raise SdemError(rc, "sdem_xdmf_end", xmf)


def build_engine(lib: C.CDLL, scenario: dict, prefix: str = "sdem_", tuning: Optional[dict] = None,
                 capacity: Optional[dict] = None) -> Engine:
    """Create an engine and push a scenario (see sdem_amd.scenarios) into it."""
    o = make_options(lib, prefix, **scenario["options"])
    eng = Engine(lib, o, prefix)
    if tuning and hasattr(lib, prefix + "set_tuning"):
        eng.set_tuning(**tuning)
    if capacity and hasattr(lib, prefix + "set_capacity"):
        eng.set_capacity(**capacity)
    _push(eng, scenario)
    return eng


def build_ensemble(lib: C.CDLL, scenarios_: list, tuning: Optional[dict] = None) -> Engine:
    """An ensemble context (sdem_create_ensemble) holding one box per scenario."""
    opts = [make_options(lib, "sdem_", **sc["options"]) for sc in scenarios_]
    eng = Engine(lib, opts)
    if tuning:
        eng.set_tuning(**tuning)
    for b, sc in enumerate(scenarios_):
        eng.select(b)
        _push(eng, sc)
    eng.select(0)
    return eng
