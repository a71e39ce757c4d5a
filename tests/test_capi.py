"""The C ABI libraries load without a GPU and export every declared symbol;
argument validation fails loudly (no compute calls here)."""
import ctypes as C
import os
import re
import subprocess

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
    assert C.sizeof(capi.RunCfg) == 72


C_NAMES = {"WorldDesc": "dxmc_world_desc", "Shell": "dxmc_shell", "Heel": "dxmc_heel", "TablesDesc": "dxmc_tables_desc",
           "Exposure": "dxmc_exposure", "RunCfg": "dxmc_run_cfg", "TuneReport": "dxmc_tune_report", "RunStats": "dxmc_run_stats",
           "RunMonitor": "dxmc_run_monitor", "SourceSpec": "dxmc_source_spec", "HostWorld": "dxmc_host_world",
           "HostTransportCfg": "dxmc_host_transport_cfg", "ServerCfg": "dxmc_server_cfg", "RequestInfo": "dxmc_request_info"}


def test_ctypes_structures_match_the_headers(tmp_path):
    """Every ctypes structure has the size and field offsets that gcc gives
    the C structure of the same fields in include/*.h."""
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "dxmc_hip.h"', '#include "dxmc_host.h"', "int main(void) {"]
    for py, c in C_NAMES.items():
        lines.append(f'printf("{py} %zu\\n", sizeof({c}));')
        lines += [f'printf("{py}.{f[0]} %zu\\n", offsetof({c}, {f[0]}));' for f in getattr(capi, py)._fields_]
    lines.append("return 0; }")
    (tmp_path / "layout.c").write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(tmp_path / "layout.c"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    wrong = []
    for key, value in zip(out[0::2], out[1::2]):
        py, _, field = key.partition(".")
        got = getattr(getattr(capi, py), field).offset if field else C.sizeof(getattr(capi, py))
        if got != int(value):
            wrong.append((key, got, int(value)))
    assert len(out) // 2 == len(C_NAMES) + sum(len(getattr(capi, p)._fields_) for p in C_NAMES)
    assert not wrong, wrong


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
