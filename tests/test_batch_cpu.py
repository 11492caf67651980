"""The batched-state program on the CPU (no GPU): states larger than the LDS
tile run as phases of LDS tiles (afquantumsim_amd/csrc/batch.hip).

tests/cpp/tile_emu.cpp's batch_emulate builds a state's program with the
library's host code (batch_absorb + batch_phases) and reads every phase the
way the batch kernel does: tile by tile, local index -> state index, controls
outside the tile resolved per tile, register segments group by group (16
amplitudes over 4 slot bits, swaps as three X gates), plain segments op by
op.  Expected states come from the numpy
restatement of tests/test_many_controls.py (the state as a (2,)*n tensor) and
from the FP64 oracle; the program rounds its matrices to FP32 once, so
amplitudes agree to 1e-5, and permutation circuits bit for bit.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
from afquantumsim_amd import circuits, hip, ops
from tests.helpers import random_state
from tests.test_many_controls import np_run, unitary
from tests.test_tiling import EMU


@pytest.fixture(scope="module")
def emu():
    if not os.path.exists(EMU):
        pytest.fail("tests/cpp/libtile_emu.so is not built (run __graft_entry__.build())")
    hip.load()
    lib = ctypes.CDLL(EMU)
    lib.batch_emulate.restype = ctypes.c_int
    lib.batch_emulate.argtypes = [ctypes.POINTER(hip.Op), ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
    return lib


def emulate(emu, op_list, n, m, sv):
    arr, keep = hip.pack_ops(op_list)
    v = np.ascontiguousarray(sv, dtype=np.complex128).copy()
    counts = (ctypes.c_size_t * 4)()
    rc = emu.batch_emulate(arr, len(op_list), n, m, v.ctypes.data, counts)
    del keep
    assert rc == 0, (rc, hip.load().aqs_last_error())
    return v, {"ops": counts[0], "phases": counts[1], "tile_loads": counts[2], "register_segments": counts[3]}


def batch_ops(n, count, seed, perm_only=False):
    """Every op kind the batch takes: (controlled) 2x2, controlled diagonal,
    phases on 1-2 targets, X with all-of and any-of controls, (controlled)
    swaps, many-controlled gates; runs of uncontrolled one-qubit gates that
    the lowering absorbs."""
    rng = np.random.default_rng(seed)
    H = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    out = [] if perm_only else [{"kind": hip.OP_MAT1, "targets": (q,), "m": H.reshape(-1)} for q in range(n)]
    kinds = ["x", "swap", "ccx", "or", "cswap", "orswap"] + ([] if perm_only else
                                                             ["u", "u", "cu", "cdiag", "phase", "diag"])
    for _ in range(count):
        k = str(rng.choice(kinds))
        qs = [int(x) for x in rng.choice(n, min(n, 6), replace=False)]
        need = {"cu": 2, "cdiag": 2, "swap": 2, "ccx": 3, "or": 3, "cswap": 3, "orswap": 4}.get(k, 1)
        if n < need:
            k = "u"
        if k == "u":
            out.append({"kind": hip.OP_MAT1, "targets": (qs[0],), "m": unitary(2, rng).reshape(-1)})
        elif k == "diag":
            d = np.exp(1j * rng.uniform(-3, 3, size=2))
            out.append({"kind": hip.OP_DIAG1, "targets": (qs[0],), "m": np.array([d[0], 0, 0, d[1]])})
        elif k == "cu":
            nc = int(rng.integers(1, min(5, n)))
            out.append({"kind": hip.OP_MAT1, "targets": (qs[0],), "controls": tuple(qs[1:1 + nc]),
                        "m": unitary(2, rng).reshape(-1)})
        elif k == "cdiag":
            d = np.exp(1j * rng.uniform(-3, 3, size=2))
            out.append({"kind": hip.OP_DIAG1, "targets": (qs[0],), "controls": (qs[1],),
                        "m": np.array([d[0], 0, 0, d[1]])})
        elif k == "phase":
            nt = int(rng.integers(1, min(3, n + 1)))
            nc = int(rng.integers(0, min(4, n - nt + 1)))
            out.append({"kind": hip.OP_PHASE, "targets": tuple(qs[:nt]), "controls": tuple(qs[nt:nt + nc]),
                        "m": np.array([np.exp(1j * rng.uniform(-3, 3)), 0, 0, 0])})
        elif k == "x":
            nc = int(rng.integers(0, len(qs)))
            out.append({"kind": hip.OP_X, "targets": (qs[0],), "controls": tuple(qs[1:1 + nc])})
        elif k == "ccx":
            out.append({"kind": hip.OP_X, "targets": (qs[0],), "controls": (qs[1], qs[2])})
        elif k == "or":
            out.append({"kind": hip.OP_X, "targets": (qs[0],), "controls": (qs[1], qs[2]), "flags": hip.CTRL_ANY})
        elif k == "swap":
            out.append({"kind": hip.OP_SWAP, "targets": (qs[0], qs[1])})
        elif k == "orswap":
            out.append({"kind": hip.OP_SWAP, "targets": (qs[0], qs[1]), "controls": (qs[2], qs[3]),
                        "flags": hip.CTRL_ANY})
        elif k == "cswap":
            out.append({"kind": hip.OP_SWAP, "targets": (qs[0], qs[1]), "controls": (qs[2],)})
    return out


def np_reference(op_list, n, sv):
    """numpy restatement; an any-of controlled X or swap (the reference's Or)
    as its all-of cases in turn: if a, if b, if a and b (an involution applied
    three times is applied once)."""
    want = np.asarray(sv, dtype=complex).copy()
    for op in op_list:
        if op.get("flags", 0) & hip.CTRL_ANY:
            a, b = op["controls"]
            for c in ((a,), (b,), (a, b)):
                want = np_run(want, n, [{"kind": op["kind"], "targets": op["targets"], "controls": c}])
        else:
            want = np_run(want, n, [op])
    return want


@pytest.mark.parametrize("n,m,seed", [(12, 12, 0), (13, 12, 1), (14, 12, 2), (15, 13, 3), (16, 14, 4),
                                      (17, 12, 5), (18, 13, 6), (16, 12, 7)])
def test_phases_match_restatement(emu, n, m, seed):
    op_list = batch_ops(n, 80, seed)
    sv = random_state(n, seed)
    got, cnt = emulate(emu, op_list, n, m, sv)
    want = np_reference(op_list, n, sv)
    np.testing.assert_allclose(got, want, atol=1e-5, rtol=0)
    if n > m:
        assert cnt["phases"] >= 1 and cnt["tile_loads"] == cnt["phases"] << (n - m)
    else:
        assert cnt["phases"] == 1 and cnt["tile_loads"] == 1


@pytest.mark.parametrize("n,m", [(15, 12), (17, 14), (18, 12)])
def test_permutation_programs_bit_exact(emu, n, m):
    op_list = batch_ops(n, 120, 10 + n, perm_only=True)
    sv = random_state(n, n).astype(np.complex64).astype(np.complex128)
    got, _ = emulate(emu, op_list, n, m, sv)
    want = np_reference(op_list, n, sv)
    assert np.array_equal(got, want)


def test_ansatz_phases_per_layer(emu):
    """The variational ansatz (RY/RZ layers and a CX ladder): each phase takes
    about a layer, so a 16-qubit state at 14-qubit tiles crosses memory a few
    times per layer instead of once per gate; the result is the oracle's."""
    for n, L, m in ((16, 8, 14), (15, 12, 12), (18, 6, 13)):
        rng = np.random.default_rng(n * 31 + L)
        gates = circuits.ry_rz_ladder(n, L, rng)
        op_list = ops.lower_all(gates)
        got, cnt = emulate(emu, op_list, n, m, oracle.basis_state(n, 0))
        want = oracle.run(oracle.basis_state(n, 0), n, gates)
        np.testing.assert_allclose(got, want, atol=1e-5, rtol=0)
        assert cnt["ops"] <= L * (2 * n) + n
        assert cnt["phases"] <= 2 * L + 2, cnt
        # the CX ladder moves 3 new qubits into the registers per segment
        assert cnt["register_segments"] <= cnt["phases"] + L * ((n + 2) // 3 + 1), cnt


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_tiny_states(emu, n):
    """Fewer than 4 qubits: plain segments; 4 and more: register segments
    that pad their slots with free bits."""
    op_list = batch_ops(n, 30, 40 + n)
    sv = random_state(n, n)
    got, cnt = emulate(emu, op_list, n, 12, sv)
    np.testing.assert_allclose(got, np_reference(op_list, n, sv), atol=1e-5, rtol=0)
    assert (cnt["register_segments"] == 0) == (n < 4)


def test_rejects_matk(emu):
    op = {"kind": hip.OP_MATK, "targets": (0, 1), "u": np.eye(4)}
    arr, keep = hip.pack_ops([op])
    v = np.zeros(1 << 13, complex)
    counts = (ctypes.c_size_t * 4)()
    assert emu.batch_emulate(arr, 1, 13, 12, v.ctypes.data, counts) != 0
