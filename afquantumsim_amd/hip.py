"""ctypes binding of the C ABI in include/aqs_hip.h (libaqs_hip.so).

This is the product's Python host layer: it loads the in-tree HIP library
and fails loudly when it is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libaqs_hip.so")

AQS_OK = 0
AQS_ERR_INVALID_ARGUMENT = 1
AQS_ERR_OUT_OF_RANGE = 2
AQS_ERR_DOMAIN = 3
AQS_ERR_RUNTIME = 4
AQS_ERR_DEVICE = 5
AQS_ERR_NO_MEMORY = 6

C64 = 0
C128 = 1
NOALLOC_MAX_QUBITS = 30

OP_MAT1 = 1
OP_DIAG1 = 2
OP_X = 3
OP_SWAP = 4
OP_MATK = 5
OP_PHASE = 6
OP_DIAGT = 7
OP_PERMB = 8
OP_TILE = 9
CTRL_ANY = 1

PROFILE_SLOTS = ["other", "MAT1", "DIAG1", "X", "SWAP", "MATK", "PHASE", "MATK_MFMA", "DIAGT", "PERMB", "TILE", "SMALL"]


class AqsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class InvalidArgument(AqsError, ValueError):
    pass


class OutOfRange(AqsError, IndexError):
    pass


class DomainError(AqsError, ValueError):
    pass


_EXC = {AQS_ERR_INVALID_ARGUMENT: InvalidArgument, AQS_ERR_OUT_OF_RANGE: OutOfRange,
        AQS_ERR_DOMAIN: DomainError}


class Op(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("ntargets", ctypes.c_int32),
        ("targets", ctypes.c_int32 * 5),
        ("ncontrols", ctypes.c_int32),
        ("controls", ctypes.c_int32 * 3),
        ("flags", ctypes.c_int32),
        ("m", ctypes.c_double * 8),
        ("u", ctypes.c_void_p),
        ("more_controls", ctypes.c_uint64),
    ]


EXPORTS = [
    "aqs_last_error", "aqs_init", "aqs_device_info", "aqs_state_create", "aqs_state_destroy",
    "aqs_state_copy", "aqs_state_qubits", "aqs_state_precision", "aqs_state_device_ptr",
    "aqs_state_set_stream", "aqs_state_sync", "aqs_state_set_basis", "aqs_state_product",
    "aqs_state_gaussian", "aqs_state_upload", "aqs_state_download", "aqs_apply", "aqs_fuse",
    "aqs_fuse_tiled", "aqs_fuse_auto", "aqs_schedule_cost", "aqs_schedule_destroy", "aqs_schedule_info", "aqs_schedule_get", "aqs_apply_schedule",
    "aqs_norm2", "aqs_normalize", "aqs_prob1", "aqs_prob1_all", "aqs_expect_sum_z",
    "aqs_probabilities", "aqs_amplitude", "aqs_sample", "aqs_measure_qubit", "aqs_collapse_to",
    "aqs_inner", "aqs_profile_enable", "aqs_profile_read", "aqs_timer_start", "aqs_timer_stop",
    "aqs_batch_create", "aqs_batch_autotune", "aqs_batch_configure", "aqs_batch_run", "aqs_batch_state", "aqs_batch_info",
    "aqs_batch_destroy", "aqs_autotune", "aqs_tune_set", "aqs_server_create", "aqs_server_set_starvation_limit",
    "aqs_server_submit", "aqs_server_wait", "aqs_server_stats", "aqs_server_destroy", "aqs_server_cancel",
    "aqs_sample_noalloc", "aqs_server_set_grouping",
]


class RequestResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int32), ("nqubits", ctypes.c_int32),
                ("shots", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("passes", ctypes.c_uint64),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double),
                ("norm2", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("admitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
                ("evictions", ctypes.c_uint64), ("pending", ctypes.c_uint64), ("running", ctypes.c_int32),
                ("max_concurrent", ctypes.c_int32), ("resident_bytes", ctypes.c_uint64),
                ("peak_resident_bytes", ctypes.c_uint64), ("budget_bytes", ctypes.c_uint64),
                ("cancelled", ctypes.c_uint64), ("groups", ctypes.c_uint64), ("grouped", ctypes.c_uint64)]

_lib = None


def load(path: str = LIB_PATH):
    """Load libaqs_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m afquantumsim_amd.build`")
    L = ctypes.CDLL(path)
    vp, sz, i32, u64, dp = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)
    st = ctypes.c_int
    sigs = {
        "aqs_init": [i32],
        "aqs_device_info": [ctypes.c_char_p, sz, ctypes.POINTER(i32), ctypes.POINTER(sz)],
        "aqs_state_create": [i32, i32, ctypes.POINTER(vp)],
        "aqs_state_destroy": [vp],
        "aqs_state_copy": [vp, vp],
        "aqs_state_qubits": [vp, ctypes.POINTER(i32)],
        "aqs_state_precision": [vp, ctypes.POINTER(i32)],
        "aqs_state_device_ptr": [vp, ctypes.POINTER(vp)],
        "aqs_state_set_stream": [vp, vp],
        "aqs_state_sync": [vp],
        "aqs_state_set_basis": [vp, u64],
        "aqs_state_product": [vp, dp],
        "aqs_state_gaussian": [vp, u64],
        "aqs_state_upload": [vp, vp, i32],
        "aqs_state_download": [vp, vp],
        "aqs_apply": [vp, ctypes.POINTER(Op), sz],
        "aqs_fuse": [ctypes.POINTER(Op), sz, i32, i32, ctypes.POINTER(vp)],
        "aqs_fuse_tiled": [ctypes.POINTER(Op), sz, i32, i32, ctypes.POINTER(vp)],
        "aqs_schedule_cost": [vp, dp],
        "aqs_fuse_auto": [ctypes.POINTER(Op), sz, i32, ctypes.POINTER(vp)],
        "aqs_schedule_destroy": [vp],
        "aqs_schedule_info": [vp, ctypes.POINTER(sz), ctypes.POINTER(sz)],
        "aqs_schedule_get": [vp, sz, ctypes.POINTER(Op)],
        "aqs_apply_schedule": [vp, vp],
        "aqs_norm2": [vp, dp],
        "aqs_normalize": [vp],
        "aqs_prob1": [vp, i32, dp],
        "aqs_prob1_all": [vp, dp],
        "aqs_expect_sum_z": [vp, dp],
        "aqs_probabilities": [vp, ctypes.POINTER(ctypes.c_float)],
        "aqs_amplitude": [vp, u64, dp],
        "aqs_sample": [vp, dp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)],
        "aqs_measure_qubit": [vp, i32, ctypes.c_double, ctypes.POINTER(i32), dp],
        "aqs_collapse_to": [vp, u64],
        "aqs_inner": [vp, vp, dp],
        "aqs_profile_enable": [vp, i32],
        "aqs_profile_read": [vp, dp, ctypes.POINTER(ctypes.c_uint64)],
        "aqs_timer_start": [vp],
        "aqs_timer_stop": [vp, dp],
        "aqs_batch_create": [i32, ctypes.POINTER(i32), ctypes.POINTER(Op), ctypes.POINTER(sz), ctypes.POINTER(vp)],
        "aqs_batch_autotune": [vp, ctypes.POINTER(i32), ctypes.POINTER(i32), dp],
        "aqs_batch_configure": [vp, i32, i32],
        "aqs_batch_run": [vp, dp, dp],
        "aqs_batch_state": [vp, i32, vp],
        "aqs_batch_info": [vp, ctypes.POINTER(sz), ctypes.POINTER(sz), ctypes.POINTER(i32), ctypes.POINTER(i32),
                           ctypes.POINTER(i32)],
        "aqs_batch_destroy": [vp],
        "aqs_autotune": [vp, i32, ctypes.POINTER(i32)],
        "aqs_tune_set": [vp, i32, i32],
        "aqs_server_create": [sz, i32, i32, ctypes.POINTER(vp)],
        "aqs_server_set_starvation_limit": [vp, ctypes.c_double],
        "aqs_server_set_grouping": [vp, i32, i32],
        "aqs_server_submit": [vp, u64, i32, ctypes.POINTER(Op), sz, dp, ctypes.c_uint32, i32],
        "aqs_server_wait": [vp, u64, i32, ctypes.POINTER(RequestResult), ctypes.POINTER(ctypes.c_uint64), vp],
        "aqs_server_stats": [vp, ctypes.POINTER(ServerStats)],
        "aqs_server_cancel": [vp, u64, ctypes.POINTER(i32)],
        "aqs_server_destroy": [vp],
        "aqs_sample_noalloc": [vp, i32, i32, vp, ctypes.c_uint32, vp, vp],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = st
    L.aqs_last_error.restype = ctypes.c_char_p
    L.aqs_last_error.argtypes = []
    _lib = L
    return L


def check(code: int):
    if code != AQS_OK:
        msg = _lib.aqs_last_error().decode(errors="replace")
        raise _EXC.get(code, AqsError)(code, msg)


def init(device: int = 0):
    check(load().aqs_init(device))


def device_info():
    L = load()
    name = ctypes.create_string_buffer(256)
    cu = ctypes.c_int()
    mem = ctypes.c_size_t()
    check(L.aqs_device_info(name, 256, ctypes.byref(cu), ctypes.byref(mem)))
    return {"arch": name.value.decode(), "cu": cu.value, "hbm_bytes": mem.value}


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class State:
    """A device-resident 2^n state vector (complex64 or complex128)."""

    def __init__(self, n: int, precision: int = C64):
        L = load()
        h = ctypes.c_void_p()
        check(L.aqs_state_create(n, precision, ctypes.byref(h)))
        self.h = h
        self.n = n
        self.precision = precision
        self.dtype = np.complex64 if precision == C64 else np.complex128

    def close(self):
        if self.h:
            check(_lib.aqs_state_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # preparation
    def set_basis(self, index: int):
        check(_lib.aqs_state_set_basis(self.h, index))

    def product(self, qamps):
        qa = np.ascontiguousarray(qamps, dtype=np.complex128).reshape(self.n, 2)
        check(_lib.aqs_state_product(self.h, _dp(qa.view(np.float64))))

    def gaussian(self, seed: int):
        check(_lib.aqs_state_gaussian(self.h, seed))

    def upload(self, host, normalize: bool = False):
        a = np.ascontiguousarray(host, dtype=self.dtype)
        assert a.size == 1 << self.n
        check(_lib.aqs_state_upload(self.h, a.ctypes.data, int(normalize)))

    def download(self):
        out = np.empty(1 << self.n, dtype=self.dtype)
        check(_lib.aqs_state_download(self.h, out.ctypes.data))
        return out

    def copy_from(self, other: "State"):
        check(_lib.aqs_state_copy(self.h, other.h))

    def sync(self):
        check(_lib.aqs_state_sync(self.h))

    # gates
    def apply(self, ops):
        p = as_packed(ops)
        check(_lib.aqs_apply(self.h, p.arr, p.count))

    def autotune(self, reps: int = 3):
        """Pick the fastest one-qubit launch shape per target qubit (state unchanged)."""
        out = (ctypes.c_int * self.n)()
        check(_lib.aqs_autotune(self.h, reps, out))
        return list(out)

    def tune_set(self, qubit: int, variant: int):
        """Pin the one-qubit launch shape (0..8) for this size and target qubit."""
        check(_lib.aqs_tune_set(self.h, qubit, variant))

    def apply_schedule(self, sch: "Schedule"):
        check(_lib.aqs_apply_schedule(self.h, sch.h))

    # measurement
    def norm2(self) -> float:
        r = ctypes.c_double()
        check(_lib.aqs_norm2(self.h, ctypes.byref(r)))
        return r.value

    def normalize(self):
        check(_lib.aqs_normalize(self.h))

    def prob1(self, q: int) -> float:
        r = ctypes.c_double()
        check(_lib.aqs_prob1(self.h, q, ctypes.byref(r)))
        return r.value

    def prob1_all(self):
        out = np.zeros(self.n, dtype=np.float64)
        check(_lib.aqs_prob1_all(self.h, _dp(out)))
        return out

    def expect_sum_z(self) -> float:
        r = ctypes.c_double()
        check(_lib.aqs_expect_sum_z(self.h, ctypes.byref(r)))
        return r.value

    def probabilities(self):
        out = np.empty(1 << self.n, dtype=np.float32)
        check(_lib.aqs_probabilities(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def amplitude(self, i: int) -> complex:
        r = (ctypes.c_double * 2)()
        check(_lib.aqs_amplitude(self.h, i, r))
        return complex(r[0], r[1])

    def sample(self, u) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(len(u), dtype=np.uint64)
        check(_lib.aqs_sample(self.h, _dp(u), len(u), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return out

    def measure_qubit(self, q: int, u: float):
        o = ctypes.c_int()
        p = ctypes.c_double()
        check(_lib.aqs_measure_qubit(self.h, q, u, ctypes.byref(o), ctypes.byref(p)))
        return o.value, p.value

    def collapse_to(self, i: int):
        check(_lib.aqs_collapse_to(self.h, i))

    def inner(self, other: "State") -> complex:
        r = (ctypes.c_double * 2)()
        check(_lib.aqs_inner(self.h, other.h, r))
        return complex(r[0], r[1])

    # timing
    def profile(self, enable: bool):
        check(_lib.aqs_profile_enable(self.h, int(enable)))

    def profile_read(self):
        k = len(PROFILE_SLOTS)
        ms = (ctypes.c_double * k)()
        cnt = (ctypes.c_uint64 * k)()
        check(_lib.aqs_profile_read(self.h, ms, cnt))
        return {PROFILE_SLOTS[i]: (ms[i], cnt[i]) for i in range(k) if cnt[i]}

    def timer_start(self):
        check(_lib.aqs_timer_start(self.h))

    def timer_stop(self) -> float:
        r = ctypes.c_double()
        check(_lib.aqs_timer_stop(self.h, ctypes.byref(r)))
        return r.value


class Schedule:
    """A fused gate schedule (compile() output).

    mode "blocks": dense blocks of <= max_qubits qubits (aqs_fuse);
    mode "tiled":  cache-blocked tile passes of 2^tile_qubits amplitudes
                   (aqs_fuse_tiled, tile_qubits 0 = 13);
    mode "auto":   aqs_fuse_auto (blocks below 20 qubits, else the lower
                   modelled cost of the two).
    """

    def __init__(self, ops, n: int, max_qubits: int = 5, mode: str = "blocks", tile_qubits: int = 0):
        L = load()
        p = as_packed(ops)
        arr = p.arr
        self.n = n
        self.h = None
        if mode not in ("blocks", "tiled", "auto"):
            raise ValueError(f"unknown schedule mode {mode!r}")
        h = ctypes.c_void_p()
        if mode == "blocks":
            check(L.aqs_fuse(arr, p.count, n, max_qubits, ctypes.byref(h)))
        elif mode == "tiled":
            check(L.aqs_fuse_tiled(arr, p.count, n, tile_qubits, ctypes.byref(h)))
        else:
            check(L.aqs_fuse_auto(arr, p.count, n, ctypes.byref(h)))
        self.h = h
        c = ctypes.c_double()
        check(L.aqs_schedule_cost(h, ctypes.byref(c)))
        self._cost = c.value
        self.mode = mode
        if mode == "auto":
            self.mode = "tiled" if any(k == OP_TILE for k, _ in self.passes()) else "blocks"

    def cost(self) -> float:
        """Modelled cost in full-state sweeps."""
        return self._cost

    def info(self):
        p, s = ctypes.c_size_t(), ctypes.c_size_t()
        check(_lib.aqs_schedule_info(self.h, ctypes.byref(p), ctypes.byref(s)))
        return p.value, s.value

    def passes(self):
        p, _ = self.info()
        out = []
        for i in range(p):
            op = Op()
            check(_lib.aqs_schedule_get(self.h, i, ctypes.byref(op)))
            out.append((op.kind, list(op.targets[:op.ntargets])))
        return out

    def close(self):
        if self.h:
            check(_lib.aqs_schedule_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many small independent states run in one persistent launch."""

    def __init__(self, circuits_ops):
        """circuits_ops: list of (nqubits, ops) pairs."""
        L = load()
        self.n = [int(n) for n, _ in circuits_ops]
        allops = []
        offs = [0]
        for _, o in circuits_ops:
            allops.extend(o)
            offs.append(len(allops))
        arr, keep = pack_ops(allops)
        nq = (ctypes.c_int * len(self.n))(*self.n)
        off = (ctypes.c_size_t * len(offs))(*offs)
        h = ctypes.c_void_p()
        check(L.aqs_batch_create(len(self.n), nq, arr, off, ctypes.byref(h)))
        del keep
        self.h = h

    def autotune(self):
        b, lq, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(_lib.aqs_batch_autotune(self.h, ctypes.byref(b), ctypes.byref(lq), ctypes.byref(ms)))
        return {"block": b.value, "lds_qubits": lq.value, "ms": ms.value}

    def configure(self, block, lds_qubits):
        """Pins the workgroup size and the LDS tile (12..14 qubits)."""
        check(_lib.aqs_batch_configure(self.h, int(block), int(lds_qubits)))

    def run(self):
        out = np.zeros(len(self.n), dtype=np.float64)
        ms = ctypes.c_double()
        check(_lib.aqs_batch_run(self.h, _dp(out), ctypes.byref(ms)))
        return out, ms.value

    def state(self, i):
        out = np.empty(1 << self.n[i], dtype=np.complex64)
        check(_lib.aqs_batch_state(self.h, i, out.ctypes.data))
        return out

    def info(self):
        p, s = ctypes.c_size_t(), ctypes.c_size_t()
        b, lq, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(_lib.aqs_batch_info(self.h, ctypes.byref(p), ctypes.byref(s), ctypes.byref(b), ctypes.byref(lq),
                                  ctypes.byref(g)))
        return {"passes": p.value, "source_ops": s.value, "block": b.value, "lds_qubits": lq.value, "grid": g.value}

    def close(self):
        if self.h:
            check(_lib.aqs_batch_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request server on one device."""

    def __init__(self, budget_bytes: int, workers: int = 4, fuse_qubits: int = 0):
        L = load()
        h = ctypes.c_void_p()
        check(L.aqs_server_create(int(budget_bytes), workers, fuse_qubits, ctypes.byref(h)))
        self.h = h
        self._shots = {}
        self._n = {}
        self._keep = {}

    def set_starvation_limit(self, seconds: float):
        check(_lib.aqs_server_set_starvation_limit(self.h, seconds))

    def set_grouping(self, max_qubits: int, max_requests: int = 64):
        """Coalesce requests of at most max_qubits qubits (0: off), at most
        max_requests per batched launch."""
        check(_lib.aqs_server_set_grouping(self.h, max_qubits, max_requests))

    def submit(self, rid: int, n: int, ops, uniforms, keep_state: bool = False):
        p = as_packed(ops)
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        check(_lib.aqs_server_submit(self.h, rid, n, p.arr, p.count, _dp(u), len(u), int(keep_state)))
        self._shots[rid] = len(u)
        self._n[rid] = n
        self._keep[rid] = keep_state

    def wait(self, rid: int, timeout_ms: int = -1):
        res = RequestResult()
        shots = self._shots.get(rid, 0)
        idx = np.zeros(max(1, shots), dtype=np.uint64)
        state = np.zeros(1 << self._n[rid], dtype=np.complex64) if self._keep.get(rid) else None
        rc = _lib.aqs_server_wait(self.h, rid, timeout_ms, ctypes.byref(res),
                                  idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                  state.ctypes.data if state is not None else None)
        if not (rc == AQS_ERR_RUNTIME and b"within the timeout" in _lib.aqs_last_error()):
            # the result was handed over (or never will be): forget the request
            for d in (self._shots, self._n, self._keep):
                d.pop(rid, None)
        check(rc)
        out = {f: getattr(res, f) for f, _ in RequestResult._fields_}
        return out, idx[:shots], state

    def cancel(self, rid: int) -> bool:
        c = ctypes.c_int()
        check(_lib.aqs_server_cancel(self.h, rid, ctypes.byref(c)))
        return bool(c.value)

    def stats(self):
        s = ServerStats()
        check(_lib.aqs_server_stats(self.h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            check(_lib.aqs_server_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PackedOps:
    """An op list packed once (pack_ops), for repeated or timed submission;
    State.apply, Server.submit and Schedule take it in place of the list."""

    def __init__(self, ops):
        self.arr, self.keep = pack_ops(ops)
        self.count = len(ops)

    def __len__(self):
        return self.count


def as_packed(ops) -> PackedOps:
    return ops if isinstance(ops, PackedOps) else PackedOps(ops)


def pack_ops(ops):
    """List of op dicts -> ctypes array (keeps matrices alive in `keep`)."""
    arr = (Op * max(1, len(ops)))()
    keep = []
    for i, o in enumerate(ops):
        x = arr[i]
        x.kind = o["kind"]
        t = o["targets"]
        x.ntargets = len(t)
        for j, q in enumerate(t):
            x.targets[j] = q
        c = list(o.get("controls", ()))
        x.ncontrols = min(3, len(c))
        for j, q in enumerate(c[:3]):
            x.controls[j] = q
        more = int(o.get("more_controls", 0))
        for q in c[3:]:  # the fourth control and beyond: a qubit mask
            more |= 1 << q
        x.more_controls = more
        x.flags = o.get("flags", 0)
        m = o.get("m")
        if m is not None:
            mm = np.asarray(m, dtype=np.complex128).reshape(-1)
            for j in range(len(mm)):
                x.m[2 * j] = mm[j].real
                x.m[2 * j + 1] = mm[j].imag
        u = o.get("u")
        if u is not None:
            uu = np.ascontiguousarray(u, dtype=np.complex128)
            keep.append(uu)
            x.u = uu.ctypes.data
        else:
            x.u = None
    return arr, keep
