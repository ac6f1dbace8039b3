"""The no-allocation solve (mfd_solve_device_noalloc) against mfd_solve_device
on the same factor: two calls in a row, the second on a smaller system with
fewer right-hand sides, so that it runs on what the first call left in the
library's fixed scratch array.  The solve kernels are deterministic (no
atomics, fixed summation order) and both entry points launch the same
kernels, so the solutions must agree bit for bit.  Device buffers come from
the HIP runtime the library itself uses (through ctypes), so the test does
not depend on any other GPU framework in the process."""
import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "multifrontal-dag_amd"))
sys.path.insert(0, ROOT)

import mfdag  # noqa: E402
from mfdag import problems as P  # noqa: E402

H2D, D2H = 1, 2


class DeviceArray:
    """A column-major float64 matrix in device memory (hipMalloc of the
    runtime libmfdag.so is linked against)."""

    def __init__(self, host):
        mfdag.lib()  # the library's HIP runtime is loaded first
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.shape = host.shape
        self.nbytes = host.nbytes
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(self.nbytes)) == 0
        self.upload(host)

    def upload(self, host):
        h = np.asfortranarray(host, dtype=np.float64)
        assert self.hip.hipMemcpy(self.ptr, h.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(self.nbytes), H2D) == 0

    def download(self):
        out = np.empty(self.shape, dtype=np.float64, order="F")
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(self.nbytes), D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _solve_both(p, nrhs, precision, seed):
    s = mfdag.Solver(nb=16, graph=True)
    s.analyze(p.ccs_col, p.ccs_row, 64, p.perm)
    assert s.factorize(p.val, precision) == 0
    _, b = P.rhs(p, nrhs, seed=seed)
    ldb = p.n + 3  # a leading dimension larger than n
    host = np.zeros((ldb, nrhs), order="F")
    host[:p.n] = b
    d_ref, d_new = DeviceArray(host), DeviceArray(host)
    assert s.solve_device(d_ref.ptr.value, nrhs, ldb) == 0
    bytes_before = s.stats()["device_bytes"]
    print(f"noalloc solve: n={p.n} nrhs={nrhs} precision={precision}", flush=True)
    rc = s.solve_device_noalloc(d_new.ptr.value, nrhs, ldb)
    assert rc == 0, rc
    assert s.stats()["device_bytes"] == bytes_before  # the handle allocated nothing
    ref, new = d_ref.download(), d_new.download()
    d_ref.free()
    d_new.free()
    assert np.array_equal(ref[:p.n], new[:p.n])
    assert np.array_equal(new[p.n:], host[p.n:])  # rows past n untouched
    assert P.backward_error(p, new[:p.n], b) <= 1e-14
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [mfdag.FP64, mfdag.FP32_IR, mfdag.FP16_IR])
def test_noalloc_solve_matches_solve_device(precision):
    _solve_both(P.weighted_laplacian_3d(20, seed=5), 4, precision, seed=1)
    # second call: a smaller system, fewer right-hand sides, same scratch
    p = P.laplacian_2d(30)
    s = _solve_both(p, 2, precision, seed=2)
    # a solve that needs more than the fixed scratch is refused before any launch
    d_b = DeviceArray(np.zeros((p.n, 1)))
    assert s.solve_device_noalloc(d_b.ptr.value, 60000, p.n) == mfdag.MFD_ERR_SCRATCH
    d_b.free()
    s.close()
