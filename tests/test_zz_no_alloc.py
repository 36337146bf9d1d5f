"""pic_diagnose_device, the allocation-free variant of pic_diagnose, against
pic_diagnose through the C-ABI on the GPU.

The record lands in device memory the caller owns (hipMalloc'd here); both entry points reduce the same workgroup partials in the same
order, so every sum agrees to the rounding of the final square root
(relative 1e-14 on gauss_l2, exact elsewhere) and the counts are exact.
Two calls in a row per form, the second on a smaller state (fewer
particles, after a few steps), check that the shared static scratch starts
each call clean."""
import ctypes

import numpy as np
import pytest

import plasmasim_amd as ps
from plasmasim_amd import _lib

pytestmark = pytest.mark.gpu


def hip():
    """The HIP runtime the library links (already loaded with it)."""
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return h


def device_record(sim):
    """The caller owns the record's device memory: allocated here, filled by
    pic_diagnose_device, copied back after the context stream synchronises."""
    h = hip()
    n = ctypes.sizeof(_lib.PicDiag)
    buf = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(buf), n) == 0
    try:
        assert h.hipMemset(buf, 0xFF, n) == 0
        sim.diagnose_device(buf.value)
        sim.sync()
        rec = _lib.PicDiag()
        assert h.hipMemcpy(ctypes.byref(rec), buf, n, 2) == 0  # device to host
    finally:
        h.hipFree(buf)
    return rec.to_dict()


def same(a, b):
    for k in a:
        x, y = np.asarray(a[k], dtype=np.float64), np.asarray(b[k], dtype=np.float64)
        if k == "gauss_l2":
            assert np.allclose(x, y, rtol=1e-14, atol=0), (k, x, y)
        elif k == "count":
            assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), (k, a[k], b[k])
        else:
            assert np.array_equal(x, y), (k, x, y)


@pytest.mark.parametrize("precision", [ps.PIC_FP64, ps.PIC_FP32])
def test_device_diagnostics_match_pic_diagnose_twice(precision):
    c = ps.preset(1)
    c.precision = precision
    c.tile[0], c.tile[1], c.tile[2] = (4, 4, 8) if precision == ps.PIC_FP32 else (0, 0, 0)
    big = ps.Simulation(c)
    big.load_particles()
    big.begin()
    big.step(2)
    same(device_record(big), big.diagnose())
    # second call: a smaller state on the same GPU (fewer particles, other
    # fields), so stale partials or counts of the first call would show
    c.ppc = 4
    small = ps.Simulation(c)
    small.load_particles()
    small.begin()
    small.step(3)
    d = device_record(small)
    ref = small.diagnose()
    same(d, ref)
    assert d["count"][0] < big.diagnose()["count"][0]
    big.close()
    small.close()
