"""Gates with more than three controls (aqs_op.more_controls).

The reference builds any number of controls (NControl_Gate and nested
ControlGates, src/quantum_gates.cpp:80-114; grover_oracle's N-controlled Z,
src/quantum_algo.cpp:16-41).  Expected states here come from an independent
numpy restatement: the state as a (2,)*n tensor (axis q = qubit q), the
sub-tensor with every control axis at 1 taken, the gate applied on its
target axes.  CPU tests run the block and tile schedulers through the host
emulation (tests/cpp/tile_emu.cpp); GPU tests run the plain, block and tile
paths in both precisions.
"""
import ctypes
import os

import numpy as np
import pytest

from afquantumsim_amd import hip
from tests.helpers import random_state
from tests.test_tiling import EMU

X2 = np.array([[0, 1], [1, 0]], dtype=complex)


def np_apply(sv, n, op):
    psi = sv.reshape((2,) * n).copy()
    ctl = list(op.get("controls", ()))
    tg = list(op["targets"])
    sel = [slice(None)] * n
    for c in ctl:
        sel[c] = 1
    kind = op["kind"]
    if kind == hip.OP_PHASE:
        for t in tg:
            sel[t] = 1
        psi[tuple(sel)] *= complex(np.asarray(op["m"]).reshape(-1)[0])
        return psi.reshape(-1)
    sub = psi[tuple(sel)]
    rest = [q for q in range(n) if q not in ctl]
    ax = [rest.index(t) for t in tg]
    if kind == hip.OP_SWAP:
        sub = np.swapaxes(sub, ax[0], ax[1])
    else:
        if kind == hip.OP_MATK:
            U = np.asarray(op["u"], dtype=complex)
        else:
            M = np.asarray(op.get("m", X2.reshape(-1)), dtype=complex).reshape(2, 2)
            U = X2 if kind == hip.OP_X else (np.diag(np.diag(M)) if kind == hip.OP_DIAG1 else M)
        k = len(tg)
        moved = np.moveaxis(sub, ax, list(range(k)))
        shp = moved.shape
        moved = (U @ moved.reshape(1 << k, -1)).reshape(shp)
        sub = np.moveaxis(moved, list(range(k)), ax)
    psi[tuple(sel)] = sub
    return psi.reshape(-1)


def np_run(sv, n, ops):
    v = np.asarray(sv, dtype=complex).copy()
    for op in ops:
        v = np_apply(v, n, op)
    return v


def unitary(k, rng):
    z = rng.normal(size=(k, k)) + 1j * rng.normal(size=(k, k))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def random_ops(n, count, seed, min_ctl=4):
    """Many-controlled ops of every kind mixed with plain gates and
    uncontrolled swaps (which the schedulers relabel)."""
    rng = np.random.default_rng(seed)
    H = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    ops = [{"kind": hip.OP_MAT1, "targets": (q,), "m": H.reshape(-1)} for q in range(n)]
    for _ in range(count):
        kind = int(rng.choice([hip.OP_MAT1, hip.OP_X, hip.OP_PHASE, hip.OP_DIAG1, hip.OP_SWAP, hip.OP_MATK,
                               0, 0]))
        if kind == 0:  # plain gates between the controlled ones
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            ops.append({"kind": hip.OP_SWAP, "targets": (a, b)} if rng.random() < 0.4 else
                       {"kind": hip.OP_MAT1, "targets": (a,), "m": unitary(2, rng).reshape(-1)})
            continue
        nt = 2 if kind in (hip.OP_SWAP, hip.OP_MATK) else (int(rng.integers(1, 3)) if kind == hip.OP_PHASE else 1)
        nc = int(rng.integers(min_ctl, n - nt + 1))
        qs = [int(x) for x in rng.choice(n, nt + nc, replace=False)]
        op = {"kind": kind, "targets": tuple(qs[:nt]), "controls": tuple(qs[nt:])}
        if kind == hip.OP_MAT1:
            op["m"] = unitary(2, rng).reshape(-1)
        elif kind == hip.OP_DIAG1:
            d = np.exp(1j * rng.uniform(-3, 3, size=2))
            op["m"] = np.array([d[0], 0, 0, d[1]])
        elif kind == hip.OP_PHASE:
            op["m"] = np.array([np.exp(1j * rng.uniform(-3, 3)), 0, 0, 0])
        elif kind == hip.OP_MATK:
            op["u"] = unitary(4, rng)
        ops.append(op)
    return ops


@pytest.fixture(scope="module")
def emu():
    if not os.path.exists(EMU):
        pytest.fail("tests/cpp/libtile_emu.so is not built (run __graft_entry__.build())")
    hip.load()
    lib = ctypes.CDLL(EMU)
    for f in (lib.tile_emulate, lib.block_emulate):
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.POINTER(hip.Op), ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                      ctypes.POINTER(ctypes.c_size_t)]
    return lib


def test_numpy_restatement_fixes_known_answers():
    """The restatement itself: a 9-controlled X permutes one basis state; a
    controlled phase touches exactly the all-ones amplitudes."""
    n = 10
    op = {"kind": hip.OP_X, "targets": (9,), "controls": tuple(range(9))}
    for idx in range(1 << n):
        v = np.zeros(1 << n, complex)
        v[idx] = 1
        out = np_apply(v, n, op)
        want = idx ^ 1 if idx >> 1 == (1 << 9) - 1 else idx
        assert out[want] == 1 and np.count_nonzero(out) == 1
    ph = {"kind": hip.OP_PHASE, "targets": (0, 3), "controls": (1, 2, 4, 5), "m": [-1, 0, 0, 0]}
    out = np_apply(np.ones(1 << n, complex), n, ph)
    idx = np.arange(1 << n)
    mask = sum(1 << (n - 1 - q) for q in (0, 1, 2, 3, 4, 5))
    assert np.array_equal(out, np.where((idx & mask) == mask, -1, 1))


@pytest.mark.parametrize("mode", ["blocks", "tiles"])
@pytest.mark.parametrize("seed", range(4))
def test_schedules_with_many_controls(emu, mode, seed):
    n = 12 + seed % 2
    ops = random_ops(n, 40, seed)
    sv = random_state(n, seed)
    want = np_run(sv, n, ops)
    arr, keep = hip.pack_ops(ops)
    v = sv.copy()
    counts = (ctypes.c_size_t * 4)()
    fn = emu.block_emulate if mode == "blocks" else emu.tile_emulate
    rc = fn(arr, len(ops), n, 5 if mode == "blocks" else 10 + seed % 3, v.ctypes.data, counts)
    del keep
    assert rc == 0, hip.load().aqs_last_error()
    np.testing.assert_allclose(v, want, atol=1e-12, rtol=0)


def test_validation_of_more_controls():
    L = hip.load()
    st = None
    arr, keep = hip.pack_ops([{"kind": hip.OP_X, "targets": (0,), "controls": (1, 2, 3, 0)}])
    h = ctypes.c_void_p()
    # control equal to the target, qubit outside the state: rejected before any device work
    assert L.aqs_fuse(arr, 1, 6, 5, ctypes.byref(h)) != 0
    arr2, keep2 = hip.pack_ops([{"kind": hip.OP_X, "targets": (0,), "controls": (1, 2, 3, 7)}])
    assert L.aqs_fuse(arr2, 1, 6, 5, ctypes.byref(h)) != 0
    arr3, keep3 = hip.pack_ops([{"kind": hip.OP_X, "targets": (0,), "controls": (1, 2, 3, 4), "flags": 1}])
    assert L.aqs_fuse(arr3, 1, 6, 5, ctypes.byref(h)) != 0
    del keep, keep2, keep3, st


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [hip.C64, hip.C128], ids=["c64", "c128"])
@pytest.mark.parametrize("mode", ["plain", "blocks", "tiled"])
def test_gpu_many_controls(precision, mode):
    n = 12
    ops = random_ops(n, 40, 7 + precision)
    sv = random_state(n, 3)
    want = np_run(sv, n, ops)
    st = hip.State(n, precision)
    st.upload(sv.astype(st.dtype))
    if mode == "plain":
        st.apply(ops)
    else:
        sch = hip.Schedule(ops, n, 5) if mode == "blocks" else hip.Schedule(ops, n, mode="tiled")
        st.apply_schedule(sch)
        sch.close()
    got = st.download()
    st.close()
    np.testing.assert_allclose(got, want, atol=5e-6 if precision == hip.C64 else 1e-12, rtol=0)


@pytest.mark.gpu
def test_gpu_many_controlled_x_is_a_permutation():
    """A 9-controlled X on 10 qubits moves exactly the two all-ones
    amplitudes, bit for bit, in both precisions and every path."""
    n = 10
    op = {"kind": hip.OP_X, "targets": (4,), "controls": (0, 1, 2, 3, 5, 6, 7, 8, 9)}
    sv = random_state(n, 11).astype(np.complex64)
    want = np_apply(sv.astype(complex), n, op).astype(np.complex64)
    for prec in (hip.C64, hip.C128):
        for mode in ("plain", "blocks", "tiled"):
            st = hip.State(n, prec)
            st.upload(sv.astype(st.dtype))
            if mode == "plain":
                st.apply([op])
            else:
                sch = hip.Schedule([op], n, 5) if mode == "blocks" else hip.Schedule([op], n, mode="tiled")
                st.apply_schedule(sch)
                sch.close()
            got = st.download().astype(np.complex64)
            st.close()
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (prec, mode)
