"""tcym_ncsdot_noalloc (the NCsDot entry point that allocates nothing: its
block partials live in a fixed device array) against tcym_ncsdot through the
C ABI: bit-identical values on two calls in a row, the second on a smaller
lattice, in both precisions; slots that hold no field and lattices beyond the
fixed scratch are refused."""
import ctypes

import numpy as np
import pytest

import topologicalcymsolver_amd as T
import _oracle as O

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    """the HIP runtime, for the test's own device buffer of the output"""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return _hip


class DeviceDouble:
    def __init__(self):
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), 8) == 0
        assert hip().hipMemset(self.p, 0xff, 8) == 0          # NaN until written

    def value(self):
        h = ctypes.c_double()
        assert hip().hipMemcpy(ctypes.byref(h), self.p, 8, 2) == 0   # hipMemcpyDeviceToHost
        return h.value

    def free(self):
        hip().hipFree(self.p)


def bits(v):
    return np.float64(v).view(np.uint64)


@pytest.mark.parametrize("prec", [T.FP64, T.FP32])
@pytest.mark.parametrize("nc", [2, 3])
def test_noalloc_ncsdot_equals_ncsdot_twice(nc, prec):
    out = DeviceDouble()
    try:
        for N, seed in (((20, 18, 16), 7), ((6, 5, 7), 8)):     # the second call is smaller
            lat = O.Lattice(nc, N)
            lat.random(seed, 0.5, 0.4, restore=False)
            gpu = T.Solver(nc, N, prec)
            try:
                gpu.set_links(lat.U)
                gpu.set_efield(lat.E)
                want = gpu.ncs_dot()
                print("noalloc SU(%d) %s prec %d: ncsdot %.17g" % (nc, N, prec, want), flush=True)
                gpu.ncs_dot_noalloc(out.p.value)
                gpu.sync()
                got = out.value()
                assert bits(got) == bits(want), (got, want)
                if prec == T.FP64:
                    assert abs(want - lat.ncsdot()) < 1e-12 * max(1.0, abs(want))
            finally:
                gpu.close()
    finally:
        out.free()


def test_noalloc_refuses_empty_slots_and_large_lattices():
    out = DeviceDouble()
    try:
        gpu = T.Solver(2, (8, 8, 8))
        try:
            with pytest.raises(T.TcymError) as ex:        # nothing uploaded: the slots hold no field
                gpu.ncs_dot_noalloc(out.p.value, T.U_AUX, T.E_AUX)
            assert ex.value.code == T.E_ARG
        finally:
            gpu.close()
        # 240^3 needs 162000 partial sums, beyond the 131072 of the fixed scratch
        gpu = T.Solver(2, (240, 240, 240), T.FP32)
        try:
            with pytest.raises(T.TcymError) as ex:
                gpu.ncs_dot_noalloc(out.p.value)
            assert ex.value.code == T.E_ARG and "scratch" in str(ex.value)
        finally:
            gpu.close()
        assert np.isnan(out.value())                      # nothing was written
    finally:
        out.free()
