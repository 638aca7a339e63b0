"""The allocation-free network gradient (mhg_mlp_gradient_noalloc) through
the C-ABI: the same loss and gradient, bit for bit, as mhg_mlp_gradient on
two calls in a row, the second with fewer rows; the row limit is refused."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import helpers as H  # noqa: E402
from mahyco_amd import abi, load_gpu

sys.path.insert(0, os.path.join(H.ROOT, "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


def _params(seed):
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(0, np.sqrt(2 / 125), (64, 128)), "b1": rng.normal(0, 0.05, 64),
         "W2": rng.normal(0, np.sqrt(2 / 64), (64, 64)), "b2": rng.normal(0, 0.05, 64),
         "W3": rng.normal(0, np.sqrt(1 / 64), (2, 64)), "b3": np.array([0.8, 0.8])}
    return M.pack(d).astype(np.float32)


def _hip():
    """The HIP runtime's allocation and copy calls, for the test's own device
    buffer (the entry point under test allocates nothing)."""
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def test_gradient_noalloc_equals_gradient():
    gpu = load_gpu()
    hip = _hip()
    fn = gpu.lib.mhg_mlp_gradient_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p]
    net = abi.Mlp(gpu, _params(3))
    rng = np.random.default_rng(4)
    m_all = 5000
    X = rng.uniform(-1, 1, (m_all, 128)).astype(np.float32)
    X[:, 125:] = 0
    T = rng.normal(0, 1, (m_all, 16)).astype(np.float32)
    net.set_data(X, T)
    nbytes = (net.nparam + 1) * 4
    out = C.c_void_p()
    assert hip.hipMalloc(C.byref(out), nbytes) == 0
    got = np.empty(net.nparam + 1, dtype=np.float32)
    try:
        for row0, m in [(100, 4096), (7, 300)]:
            print("noalloc gradient rows", row0, m, flush=True)
            loss, grad = net.gradient(row0, m)
            got[:] = np.nan
            assert fn(net.h, row0, m, out) == abi.OK
            assert hip.hipMemcpy(got.ctypes.data, out, nbytes, 2) == 0  # device to host
            assert np.array_equal(got[:-1], grad), np.abs(got[:-1] - grad).max()
            assert got[-1] == np.float32(loss)
        assert fn(net.h, 0, 4097, out) == abi.E_ARG
        assert fn(net.h, 4000, 1001, out) == abi.E_ARG
    finally:
        hip.hipFree(out)
        net.close()
