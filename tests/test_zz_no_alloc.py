"""rsfg_matvec_dev_noalloc (the product with its staging in a static device
array) against rsfg_matvec_dev (the same product with allocated staging),
through the C ABI on device buffers: bitwise the same results on two calls
in a row, the second with a smaller input, in every storage format, with
and without normal rows; an operator too large for the static array is
refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _runtime_of_librsfg():
    """path of the HIP runtime librsfg.so uses.  When librsfg is loaded
    first it binds to the system runtime, and a later torch import maps
    torch's own copy beside it, whose device buffers would belong to another
    runtime instance: prefer the system copy, else the only one mapped."""
    import os
    from interact_amd import lib
    lib.load()
    paths = set()
    with open("/proc/self/maps") as f:
        for line in f:
            p = line.split()[-1]
            if "libamdhip64.so" in os.path.basename(p):
                paths.add(p)
    assert paths
    system = [p for p in paths if "-packages" not in p]
    return sorted(system or paths)[0]


class _Hip:
    """the few HIP runtime calls the test needs for device buffers"""

    def __init__(self):
        self.L = C.CDLL(_runtime_of_librsfg())

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        return p

    def put(self, dst, a):
        assert self.L.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def get(self, a, src):
        assert self.L.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, C.c_size_t(a.nbytes), 2) == 0

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free(self, p):
        self.L.hipFree(p)


def _system(length_km, depth_km, with_normal, store):
    from interact_amd import faults, lib
    from interact_amd.model import FaultSystem
    g = faults.planar_grid(2.0, length_km, depth_km)
    fs = FaultSystem(g)
    fs.assemble(lib.SLIP_VECTOR, shear_modulus=32e9, with_normal=with_normal, store=store)
    return fs


def _both(hip, fs, prec, x):
    """y of rsfg_matvec_dev and of rsfg_matvec_dev_noalloc for input x"""
    from interact_amd import lib
    L = lib.load()
    n = fs.ncomp * fs.n
    dx, dy1, dy2 = hip.alloc(x.nbytes), hip.alloc(8 * n), hip.alloc(8 * n)
    try:
        hip.put(dx, x)
        lib.check(L.rsfg_matvec_dev(fs.h, prec, dx, dy1), "rsfg_matvec_dev")
        lib.check(L.rsfg_matvec_dev_noalloc(fs.h, prec, dx, dy2), "rsfg_matvec_dev_noalloc")
        hip.sync()
        y1, y2 = np.zeros(n), np.zeros(n)
        hip.get(y1, dy1)
        hip.get(y2, dy2)
    finally:
        for p in (dx, dy1, dy2):
            hip.free(p)
    return y1, y2


@pytest.mark.parametrize("with_normal", [False, True])
def test_noalloc_product_equals_allocating_product(with_normal):
    from interact_amd import lib
    hip = _Hip()
    store = lib.PREC_FP32 | lib.PREC_BF16X2 | lib.PREC_BF16NF
    big = _system(100.0, 40.0, with_normal, store)      # 2000 DOF
    small = _system(40.0, 20.0, with_normal, store)     # 400 DOF
    rng = np.random.default_rng(7)
    for prec in (lib.PREC_FP64, lib.PREC_FP32, lib.PREC_BF16X2, lib.PREC_BF16NF):
        for fs in (big, small):     # the second call has the smaller input
            x = rng.standard_normal(fs.ncomp * fs.n)
            y1, y2 = _both(hip, fs, prec, x)
            assert np.array_equal(y1, y2), (prec, fs.n)
            assert np.abs(y1).max() > 0
            # and the host-side product agrees (FP64)
            if prec == lib.PREC_FP64:
                assert np.abs(y1 - fs.matvec(x)).max() == 0.0
    big.close()
    small.close()


def test_noalloc_refuses_an_operator_larger_than_its_buffer():
    """0.5 km patches with normal rows: 32000 columns + 48000 rows exceed the
    65536-double static buffer; the allocating entry point still runs."""
    from interact_amd import faults, lib
    from interact_amd.model import FaultSystem
    g = faults.planar_grid(0.5)
    fs = FaultSystem(g)
    fs.assemble(lib.SLIP_VECTOR, shear_modulus=32e9, with_normal=True)
    hip = _Hip()
    n = fs.ncomp * fs.n
    dx, dy = hip.alloc(8 * n), hip.alloc(8 * n)
    try:
        L = lib.load()
        rc = L.rsfg_matvec_dev_noalloc(fs.h, lib.PREC_FP64, dx, dy)
        assert rc == lib.RSFG_EINVAL
        assert b"too large for the static staging buffer" in L.rsfg_last_error()
        hip.put(dx, np.ones(n))
        lib.check(L.rsfg_matvec_dev(fs.h, lib.PREC_FP64, dx, dy), "rsfg_matvec_dev")
        hip.sync()
    finally:
        hip.free(dx)
        hip.free(dy)
    fs.close()
