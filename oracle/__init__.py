"""CPU oracle of the sDEM step (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
this package; the product library never does.  `load()` returns the
ctypes handle of liboracle_sdem.so, whose entry points mirror
include/sdem.h with the prefix ``oracle_``.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_sdem.so")
PREFIX = "oracle_"


def build() -> str:
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def load() -> ctypes.CDLL:
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    # entry points beyond the shared ABI
    lib.oracle_get_pair_prop.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double)]
    lib.oracle_get_num_walls.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    lib.oracle_get_wall.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
    lib.oracle_wall_contact.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    lib.oracle_get_num_cnsv.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(ctypes.c_int64)]
    return lib
