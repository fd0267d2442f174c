"""xpir_amd -- MI355X (gfx950) engine for XPIR's lattice (NFLLWE) PIR reply generation.

The compute path is libxpir_amd.so: hand-written HIP kernels for database import
(absorb + NTT), the reply MAC and the recursion step, behind the C ABI declared in
include/xpir_amd.h.  ``xpir_amd.capi`` is its ctypes binding; ``xpir_amd.cpp`` holds the
C++ facade that mirrors the reference's libpir.hpp classes.
"""
from .capi import Context, Database, DeviceBuffer, Server, XpgError, keygen, lib, params  # noqa: F401

__all__ = ["Context", "Database", "DeviceBuffer", "Server", "XpgError", "keygen", "lib", "params"]
