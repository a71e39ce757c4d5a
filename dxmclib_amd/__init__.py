"""dxmclib_amd - MI355X (gfx950) photon transport for diagnostic x-ray dose.

Native pieces (built in-tree by dxmclib_amd.build):
  libdxmc_hip.so   HIP kernels behind the C ABI of include/dxmc_hip.h
  libdxmc_host.so  C++ host API (World / Source / Exposure / Transport) and
                   its C facade include/dxmc_host.h
Python: capi (ctypes declarations), backend (numpy wrappers), workloads
(synthetic phantoms and sources of the benchmark configurations), serving
(request loop with continuous batching).
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
