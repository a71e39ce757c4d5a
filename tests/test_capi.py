"""The C ABI libraries load without a GPU and export every declared symbol;
argument validation fails loudly (no compute calls here)."""
import ctypes as C
import os
import re

import pytest

from dxmclib_amd import backend, capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(dxmc_\w+)\s*\(", text, re.M)))


@pytest.mark.parametrize("header,lib", [("dxmc_hip.h", capi.HIP_LIB), ("dxmc_host.h", capi.HOST_LIB)])
def test_exports_every_declared_symbol(header, lib):
    names = _declared(header)
    assert len(names) >= 8
    dll = C.CDLL(lib)
    missing = [n for n in names if not hasattr(dll, n)]
    assert not missing, missing
    table = capi.HIP_SYMBOLS if header == "dxmc_hip.h" else capi.HOST_SYMBOLS
    assert set(names) == set(table), "ctypes declarations out of sync with the header"


def test_abi_version():
    assert capi.hip().dxmc_hip_abi_version() == 1


def test_struct_layouts():
    # sizes of the C structures as compiled (x86-64 SysV)
    assert C.sizeof(capi.Exposure) == 18 * 8 + 8 + 16
    assert C.sizeof(capi.Shell) == 11 * 8
    assert C.sizeof(capi.Heel) == 4 * 8 + 16
    assert C.sizeof(capi.RunCfg) == 64


def test_null_arguments_rejected():
    lib = capi.hip()
    assert lib.dxmc_hip_ctx_create(0, None) == capi.DXMC_ERR_ARGUMENT
    assert lib.dxmc_hip_ctx_destroy(None) == capi.DXMC_ERR_ARGUMENT
    assert lib.dxmc_hip_run(None, None, None, None, 0, None, None, None, None, None) == capi.DXMC_ERR_ARGUMENT
    assert lib.dxmc_hip_tally_download(None, None, None, None, None) == capi.DXMC_ERR_ARGUMENT


def test_host_rejects_bad_input():
    with pytest.raises(backend.DxmcError):
        backend.HostTables(["Not A Material"], [1.0], 100.0)
    spec = backend.source_spec(99)
    t = backend.HostTables(["Water, Liquid"], [1.0], 100.0)
    with pytest.raises(backend.DxmcError):
        t.exposures(spec)
