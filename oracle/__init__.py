"""FP64 CPU oracle for the afQuantumSim state-vector path (ctypes binding).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker / the timed CPU baseline.  The product path (``afquantumsim_amd`` and
``libaqs_hip.so``) never imports it.

The C restatement lives in ``aqs_oracle.c``; each gate there cites the
reference operator it follows (src/quantum.cpp line ranges).  Gates are given
as tuples ``(kind, q0, q1, q2, angle, phi, lam)`` or ``("circuit", ...)``
dicts, see :func:`gate`.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")

# Kinds: include/quantum.h:799-823 order, then U3.
KINDS = [
    "Barrier", "X", "Y", "Z", "H", "Phase", "Swap", "RotX", "RotY", "RotZ",
    "CX", "CY", "CZ", "CH", "CPhase", "CSwap", "CRotX", "CRotY", "CRotZ",
    "CCX", "Or", "Circuit", "ControlCircuit", "U3",
]
KIND = {k: i for i, k in enumerate(KINDS)}


class OrcGate(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("q0", ctypes.c_int32),
        ("q1", ctypes.c_int32),
        ("q2", ctypes.c_int32),
        ("angle", ctypes.c_float),
        ("phi", ctypes.c_float),
        ("lam", ctypes.c_float),
        ("k", ctypes.c_int32),
        ("u", ctypes.c_void_p),
    ]


_lib = None


def build() -> str:
    """Compile the oracle with its Makefile; returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        up = ctypes.POINTER(ctypes.c_uint32)
        gp = ctypes.POINTER(OrcGate)
        L.orc_apply.argtypes = [dp, ctypes.c_int, gp]
        L.orc_run.argtypes = [dp, ctypes.c_int, gp, ctypes.c_int]
        L.orc_run_threads.argtypes = [dp, ctypes.c_int, gp, ctypes.c_int, ctypes.c_int]
        L.orc_refmethod_run.argtypes = [dp, ctypes.c_int, gp, ctypes.c_int, ctypes.c_int]
        L.orc_refmethod_run.restype = ctypes.c_int
        L.orc_matrix_1q.argtypes = [gp, dp]
        L.orc_dense_operator.argtypes = [ctypes.c_int, gp, dp]
        L.orc_gemv.argtypes = [ctypes.c_int, dp, dp, dp]
        L.orc_basis_state.argtypes = [dp, ctypes.c_int, ctypes.c_uint64]
        L.orc_product_state.argtypes = [dp, ctypes.c_int, dp]
        L.orc_gaussian_state.argtypes = [dp, ctypes.c_int, ctypes.c_uint64]
        L.orc_gaussian_component.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]
        L.orc_gaussian_component.restype = ctypes.c_double
        L.orc_probabilities.argtypes = [dp, ctypes.c_int, dp]
        L.orc_norm2.argtypes = [dp, ctypes.c_int]
        L.orc_norm2.restype = ctypes.c_double
        L.orc_prob1.argtypes = [dp, ctypes.c_int, ctypes.c_int]
        L.orc_prob1.restype = ctypes.c_double
        L.orc_expect_sum_z.argtypes = [dp, ctypes.c_int]
        L.orc_expect_sum_z.restype = ctypes.c_double
        L.orc_sample_c64.argtypes = [fp, ctypes.c_int, dp, ctypes.c_int, up]
        L.orc_sample_c128.argtypes = [dp, ctypes.c_int, dp, ctypes.c_int, up]
        L.orc_sample_fp64_cdf.argtypes = [dp, ctypes.c_int, dp, ctypes.c_int, up]
        L.orc_max_threads.restype = ctypes.c_int
        L.csv_apply.argtypes = [fp, ctypes.c_int, gp, ctypes.c_int]
        L.csv_apply.restype = ctypes.c_int
        L.csv_run.argtypes = [fp, ctypes.c_int, gp, ctypes.c_int, ctypes.c_int]
        L.csv_run.restype = ctypes.c_int
        L.csv_sample.argtypes = [fp, ctypes.c_int, dp, dp, ctypes.c_int, up, ctypes.c_int]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def gate(kind, q0=0, q1=0, q2=0, angle=0.0, phi=0.0, lam=0.0, k=0, u=None):
    """A gate description shared by the oracle and the GPU tests."""
    return {"kind": kind, "q": (q0, q1, q2), "angle": float(np.float32(angle)),
            "phi": float(np.float32(phi)), "lam": float(np.float32(lam)),
            "k": k, "u": None if u is None else np.ascontiguousarray(u, dtype=np.complex128)}


def _pack(gates):
    arr = (OrcGate * max(1, len(gates)))()
    keep = []
    for i, g in enumerate(gates):
        o = arr[i]
        o.kind = KIND[g["kind"]]
        o.q0, o.q1, o.q2 = g["q"]
        o.angle, o.phi, o.lam = g["angle"], g["phi"], g["lam"]
        o.k = g["k"]
        if g["u"] is not None:
            keep.append(g["u"])
            o.u = g["u"].ctypes.data
        else:
            o.u = None
    return arr, keep


def run(sv: np.ndarray, n: int, gates, threads: int = 1) -> np.ndarray:
    """Apply gates in order to a complex128 state (copy returned)."""
    out = np.ascontiguousarray(sv, dtype=np.complex128).copy()
    arr, _keep = _pack(gates)
    lib().orc_run_threads(_dp(out.view(np.float64)), n, arr, len(gates), threads)
    return out


def cpu_sv_run(sv: np.ndarray, n: int, gates, threads: int = 1) -> np.ndarray:
    """The plain complex64 CPU baseline (cpu_sv.c): gates applied in place to
    ``sv`` (complex64, modified and returned)."""
    assert sv.dtype == np.complex64 and sv.flags.c_contiguous and sv.size == 1 << n
    arr, _keep = _pack(gates)
    done = lib().csv_run(sv.view(np.float32).ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n, arr, len(gates),
                         threads)
    if done != len(gates):
        raise ValueError(f"cpu_sv does not implement gate {gates[done]['kind']}")
    return sv


def cpu_sv_sample(sv: np.ndarray, n: int, u: np.ndarray, threads: int = 1, cdf: np.ndarray = None) -> np.ndarray:
    """Sampling on the complex64 CPU baseline: FP64 inclusive CDF, binary search per shot."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    if cdf is None:
        cdf = np.empty(1 << n, dtype=np.float64)
    out = np.zeros(len(u), dtype=np.uint32)
    lib().csv_sample(sv.view(np.float32).ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n, _dp(cdf), _dp(u),
                     len(u), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), threads)
    return out


def refmethod_run(sv: np.ndarray, n: int, gates, threads: int = 1) -> np.ndarray:
    """The reference's method (dense N x N products for dense gates), for timing."""
    out = np.ascontiguousarray(sv, dtype=np.complex128).copy()
    arr, _keep = _pack(gates)
    if lib().orc_refmethod_run(_dp(out.view(np.float64)), n, arr, len(gates), threads) != 0:
        raise MemoryError("N x N operator allocation failed")
    return out


def dense_operator(n: int, g) -> np.ndarray:
    N = 1 << n
    op = np.zeros((N, N), dtype=np.complex128)
    arr, _keep = _pack([g])
    lib().orc_dense_operator(n, arr, _dp(op.view(np.float64)))
    return op


def matrix_1q(g) -> np.ndarray:
    m = np.zeros(4, dtype=np.complex128)
    arr, _keep = _pack([g])
    ok = lib().orc_matrix_1q(arr, _dp(m.view(np.float64)))
    return m.reshape(2, 2) if ok else None


def basis_state(n: int, index: int) -> np.ndarray:
    sv = np.zeros(1 << n, dtype=np.complex128)
    lib().orc_basis_state(_dp(sv.view(np.float64)), n, index)
    return sv


def product_state(n: int, qamps) -> np.ndarray:
    qa = np.ascontiguousarray(qamps, dtype=np.complex128).reshape(n, 2)
    sv = np.zeros(1 << n, dtype=np.complex128)
    lib().orc_product_state(_dp(sv.view(np.float64)), n, _dp(qa.view(np.float64)))
    return sv


def gaussian_state(n: int, seed: int) -> np.ndarray:
    sv = np.zeros(1 << n, dtype=np.complex128)
    lib().orc_gaussian_state(_dp(sv.view(np.float64)), n, seed)
    return sv


def prob1(sv: np.ndarray, n: int, q: int) -> float:
    sv = np.ascontiguousarray(sv, dtype=np.complex128)
    return lib().orc_prob1(_dp(sv.view(np.float64)), n, q)


def norm2(sv: np.ndarray, n: int) -> float:
    sv = np.ascontiguousarray(sv, dtype=np.complex128)
    return lib().orc_norm2(_dp(sv.view(np.float64)), n)


def expect_sum_z(sv: np.ndarray, n: int) -> float:
    sv = np.ascontiguousarray(sv, dtype=np.complex128)
    return lib().orc_expect_sum_z(_dp(sv.view(np.float64)), n)


def sample_c64(amps: np.ndarray, n: int, u: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(amps, dtype=np.complex64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(len(u), dtype=np.uint32)
    lib().orc_sample_c64(a.view(np.float32).ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n,
                         _dp(u), len(u), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out


def sample_c128(sv: np.ndarray, n: int, u: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(sv, dtype=np.complex128)
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(len(u), dtype=np.uint32)
    lib().orc_sample_c128(_dp(a.view(np.float64)), n, _dp(u), len(u),
                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out


def sample_fp64_cdf(sv: np.ndarray, n: int, u: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(sv, dtype=np.complex128)
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(len(u), dtype=np.uint32)
    lib().orc_sample_fp64_cdf(_dp(a.view(np.float64)), n, _dp(u), len(u),
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out


def unitary(n: int, gates) -> np.ndarray:
    """The circuit's 2^n x 2^n unitary (column j = circuit applied to e_j)."""
    N = 1 << n
    U = np.zeros((N, N), dtype=np.complex128)
    for j in range(N):
        U[:, j] = run(basis_state(n, j), n, gates)
    return U


def infidelity(psi: np.ndarray, ref: np.ndarray) -> float:
    """1 - |<ref|psi>|^2 with both renormalised in FP64."""
    a = np.asarray(psi, dtype=np.complex128)
    b = np.asarray(ref, dtype=np.complex128)
    ov = np.vdot(b, a) / (np.linalg.norm(a) * np.linalg.norm(b))
    return float(max(0.0, 1.0 - abs(ov) ** 2))
