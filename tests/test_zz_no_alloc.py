"""pm3d_deposit_noalloc (the allocation-free deposit: device pointers in and out, fixed-point mesh in
the library's scratch array) against pm3d_deposit on an FP64 handle, through the C-ABI: bit for bit,
on two calls in a row, the second with a smaller input (fewer particles, smaller mesh), so a stale
mesh left by the first call would show.  Device buffers come from hipMalloc of the HIP runtime the
library itself links, called through ctypes."""
import ctypes
import sys

import numpy as np
import pytest

from pbh_amd import _lib

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    _lib.load()  # the library's own runtime is loaded first; the same one is found again here
    rt = ctypes.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


class DevBuf:
    def __init__(self, rt, nbytes):
        self.rt = rt
        self.ptr = ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(self.ptr), max(nbytes, 8)) == 0

    def put(self, a: np.ndarray):
        assert self.rt.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, a: np.ndarray):
        assert self.rt.hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, D2H) == 0

    def free(self):
        self.rt.hipFree(self.ptr)


def positions(N, n, L, seed):
    """Half jittered lattice points (cell edges and centres included), half uniform."""
    rng = np.random.default_rng(seed)
    X = rng.uniform(0, L, (n, 3))
    m = n // 2
    q = (rng.integers(0, 2 * N, (m, 3)) * 0.5) * (L / N)
    X[:m] = np.remainder(q + rng.normal(0, 1e-3 * L / N, (m, 3)) * (rng.random((m, 1)) < 0.5), L)
    return X


def test_noalloc_deposit_matches_handle_deposit(hip):
    from pbh_amd.pm3d import PM3D
    lib = _lib.load()
    cases = [(128, 2_097_152, 1.0, 1), (32, 5_000, 8 * np.sqrt(6.0), 2)]  # second call smaller
    for N, n, L, seed in cases:
        say("noalloc deposit", N, n, L)
        X = np.ascontiguousarray(positions(N, n, L, seed))
        dX, dR = DevBuf(hip, X.nbytes), DevBuf(hip, N ** 3 * 8)
        dX.put(X)
        rc = lib.pm3d_deposit_noalloc(dX.ptr, n, N, L, dR.ptr, None)
        assert rc == 0, lib.pm3d_last_error(None)
        assert hip.hipDeviceSynchronize() == 0
        got = np.empty((N, N, N))
        dR.get(got)
        dX.free()
        dR.free()
        ref = PM3D(N, L, 8, frac_bits=52).deposit(X)  # the same 2^-52 grid on 128-bit cells
        assert np.array_equal(got, ref)
        assert abs(got.sum() - n) <= 1e-9 * n


def test_noalloc_rejects_meshes_beyond_the_scratch():
    lib = _lib.load()
    assert lib.pm3d_deposit_noalloc(None, 0, 256, 1.0, ctypes.c_void_p(8), None) == -1
    assert b"128" in lib.pm3d_last_error(None)
    assert lib.pm3d_deposit_noalloc(None, 0, 12, 1.0, ctypes.c_void_p(8), None) == -1
