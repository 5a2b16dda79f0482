"""MI355X (gfx950) implementation of sDEM's per-timestep DEM path.

The compute path is libsdem_hip.so (HIP kernels behind the C ABI of
include/sdem.h); `sdem_amd.native` binds it with ctypes, `sdem_amd.scenarios`
builds seeded initial conditions, and sdem_amd/fortran holds the Fortran
host that keeps the reference's .prm input and call sequence.
"""
from .native import Engine, Options, SdemError, build_engine, load_hip_library, make_options  # noqa: F401

__all__ = ["Engine", "Options", "SdemError", "build_engine", "load_hip_library", "make_options"]
