"""aqs_sample_noalloc: sampling on caller-owned device buffers, no allocation.

Checked against aqs_sample (the entry point it varies) and the FP64 oracle's
restatement of the sampling rule, bit for bit, on two calls in a row, the
second on a smaller state with fewer shots (the shared device scratch keeps
the first call's tables).  The test owns the uniforms and index buffers: it
allocates them with hipMalloc through the HIP runtime libaqs_hip.so itself
loaded (a second runtime in one process, such as the one torch bundles,
cannot open the device beside it).
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
from afquantumsim_amd import hip


def _hip_runtime():
    """The libamdhip64 that libaqs_hip.so is linked against (already loaded)."""
    hip.load()
    with open("/proc/self/maps") as f:
        for line in f:
            path = line.split()[-1]
            if "libamdhip64.so" in path and os.path.exists(path):
                rt = ctypes.CDLL(path)
                rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
                rt.hipFree.argtypes = [ctypes.c_void_p]
                rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
                rt.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
                rt.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
                rt.hipStreamDestroy.argtypes = [ctypes.c_void_p]
                return rt
    raise RuntimeError("libamdhip64 is not loaded with libaqs_hip.so")


H2D, D2H = 1, 2


def test_noalloc_rejects_sizes_beyond_its_scratch():
    # host-side checks, before any device work
    L = hip.load()
    rc = L.aqs_sample_noalloc(None, hip.NOALLOC_MAX_QUBITS + 1, hip.C64, None, 1, None, None)
    assert rc == hip.AQS_ERR_INVALID_ARGUMENT
    assert b"nqubits" in L.aqs_last_error()
    assert L.aqs_sample_noalloc(None, 0, hip.C64, None, 1, None, None) == hip.AQS_ERR_INVALID_ARGUMENT


def _uniforms(rng, shots, cdf_points):
    u = rng.uniform(size=shots)
    # exact edges: 0, just below 1, and values equal to cumulative sums
    u[:3] = [0.0, np.nextafter(1.0, 0.0), 0.5]
    k = min(len(cdf_points), shots - 3)
    u[3:3 + k] = cdf_points[:k]
    return u


@pytest.mark.gpu
def test_noalloc_matches_aqs_sample_twice():
    hip.init(0)
    rt = _hip_runtime()
    L = hip.load()
    stream = ctypes.c_void_p()
    assert rt.hipStreamCreate(ctypes.byref(stream)) == 0
    rng = np.random.default_rng(11)
    cases = ((17, hip.C64, 4000, stream), (11, hip.C64, 257, None), (9, hip.C128, 100, stream))
    for n, precision, shots, strm in cases:
        st = hip.State(n, precision)
        st.gaussian(1000 + n)
        host = st.download()
        cdf = np.cumsum(np.abs(host.astype(np.complex128)) ** 2)
        u = _uniforms(rng, shots, cdf[:: max(1, len(cdf) // 50)])
        du, dout = ctypes.c_void_p(), ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(du), 8 * shots) == 0
        assert rt.hipMalloc(ctypes.byref(dout), 8 * shots) == 0
        try:
            assert rt.hipMemcpy(du, u.ctypes.data, 8 * shots, H2D) == 0
            amps = ctypes.c_void_p()
            hip.check(L.aqs_state_device_ptr(st.h, ctypes.byref(amps)))
            st.sync()
            print(f"noalloc: n={n} precision={precision} shots={shots} own_stream={strm is not None}", flush=True)
            hip.check(L.aqs_sample_noalloc(amps, n, precision, du, shots, dout, strm))
            if strm is not None:
                assert rt.hipStreamSynchronize(strm) == 0
            else:
                st.sync()  # the library's stream
            got = np.zeros(shots, dtype=np.uint64)
            assert rt.hipMemcpy(got.ctypes.data, dout, 8 * shots, D2H) == 0
        finally:
            rt.hipFree(du)
            rt.hipFree(dout)
        np.testing.assert_array_equal(got, st.sample(u))
        if precision == hip.C64:
            np.testing.assert_array_equal(got.astype(np.uint32), oracle.sample_c64(host, n, u))
        else:
            np.testing.assert_array_equal(got.astype(np.uint32), oracle.sample_c128(host, n, u))
        st.close()
    rt.hipStreamDestroy(stream)
