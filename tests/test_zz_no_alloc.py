"""ipl_global_sums_noalloc, the allocation-free variant of ipl_global_sums, through the C ABI.

GPU: on two calls in a row, the second on a smaller grid, its sums equal those of
ipl_global_sums bit for bit (same kernels, same order) on the fused and the plain path, FP64 and
FP32.  CPU: in the emulation build, the call makes no device allocation while ipl_global_sums on
the plain path makes one."""
import ctypes
import os
import sys

import numpy as np
import pytest

from ip_cu_lbm_amd import F32, F64, PATH_FUSED, PATH_PLAIN, Simulation, cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _sums(sim, name):
    out = np.zeros(3)
    rc = getattr(sim._lib, name)(sim._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, sim._lib.ipl_last_error_message(sim._h).decode()
    return out


def _pair(lib=None):
    big = cases.lock_exchange(nx=48, ny=8, nz=24, storage=F64, path=PATH_FUSED)
    small = cases.lock_exchange(nx=20, ny=6, nz=12, storage=F64, path=PATH_FUSED)
    return big, small


@pytest.mark.gpu
@pytest.mark.parametrize("storage,path", [(F64, PATH_FUSED), (F32, PATH_FUSED), (F64, PATH_PLAIN), (F32, PATH_PLAIN)])
def test_noalloc_sums_equal_global_sums(storage, path):
    for case in _pair():
        case = case.replace(storage=storage, path=path)
        with Simulation(case) as sim:
            sim.step(10)
            want = _sums(sim, "ipl_global_sums")
            got = _sums(sim, "ipl_global_sums_noalloc")
            again = _sums(sim, "ipl_global_sums_noalloc")
        print(case.name, storage, path, want, got, flush=True)
        assert np.array_equal(got, want), (case.name, got, want)
        assert np.array_equal(again, want)
        assert np.all(np.isfinite(got)) and got[1] > 0  # salinity mass


def test_noalloc_sums_allocate_nothing_in_the_emulation():
    sys.path.insert(0, os.path.join(ROOT, "tests", "emu"))
    import build_emu
    lib_path = build_emu.build(asan=False)
    lib = ctypes.CDLL(lib_path)
    lib.ipl_emu_alloc_count.restype = ctypes.c_longlong
    for case in _pair():
        for path in (PATH_PLAIN, PATH_FUSED):
            with Simulation(case.replace(path=path, storage=F64), lib_path) as sim:
                sim.step(2)
                before = lib.ipl_emu_alloc_count()
                got = _sums(sim, "ipl_global_sums_noalloc")
                assert lib.ipl_emu_alloc_count() == before
                want = _sums(sim, "ipl_global_sums")
                assert lib.ipl_emu_alloc_count() == before + (1 if path == PATH_PLAIN else 0)
                assert np.array_equal(got, want)
