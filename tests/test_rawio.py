"""Raw binary arrays (phantoms, dose maps) as the reference's validation
program reads and writes them (validation.cpp:128-135, 1280-1306): exactly
the array's bytes, no header; a read of the wrong size fails."""
import ctypes as C

import numpy as np
import pytest

from dxmclib_amd import backend, capi


def _read(path, like):
    out = np.zeros_like(like)
    st = capi.host().dxmc_host_read_raw(str(path).encode(), out.ctypes.data_as(C.c_void_p), out.nbytes)
    return st, out


def test_round_trip_and_numpy_layout(tmp_path):
    rng = np.random.default_rng(4)
    phantom = rng.integers(0, 20, (26, 32, 50), dtype=np.uint8)  # z, y, x as the case-5 phantom's 500 x 320 x 260
    dose = rng.random(phantom.shape)
    lib = capi.host()
    for name, a in (("world.bin", phantom), ("dose.bin", dose)):
        p = tmp_path / name
        assert lib.dxmc_host_write_raw(str(p).encode(), a.ctypes.data_as(C.c_void_p), a.nbytes) == 0
        assert p.stat().st_size == a.nbytes
        assert np.array_equal(np.fromfile(p, dtype=a.dtype).reshape(a.shape), a)  # x fastest, no header
        st, back = _read(p, a)
        assert st == 0 and np.array_equal(back, a)


def test_wrong_size_and_missing_file(tmp_path):
    a = np.arange(100, dtype=np.uint8)
    p = tmp_path / "short.bin"
    a[:99].tofile(p)
    st, _ = _read(p, a)
    assert st == capi.DXMC_ERR_ARGUMENT
    assert b"exactly 100 bytes" in capi.host().dxmc_host_last_error()
    st, _ = _read(tmp_path / "missing.bin", a)
    assert st == capi.DXMC_ERR_ARGUMENT
    with pytest.raises(backend.DxmcError):
        backend._check_host(st)

