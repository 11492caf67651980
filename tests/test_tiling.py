"""CPU checks of the cache-blocked scheduler (aqs_fuse_tiled).

tests/cpp/tile_emu.cpp builds the tiled schedule with the library's host
code and runs it on an FP64 host state, interpreting every tile pass the way
the HIP kernel does (tiles, segments, register slots, thread-fixed and
global bits).  Each result must equal the oracle's per-gate FP64 run to
1e-12: a gate moved past one it does not commute with, a wrong slot or bit
class, a control dropped, or a segment that skips a gate all fail.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
from afquantumsim_amd import circuits, hip, ops
from tests.helpers import ALL_KINDS_4Q, kind_gate, random_state

EMU = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "libtile_emu.so")


@pytest.fixture(scope="module")
def emu():
    if not os.path.exists(EMU):
        pytest.fail("tests/cpp/libtile_emu.so is not built (run __graft_entry__.build())")
    hip.load()
    lib = ctypes.CDLL(EMU)
    lib.tile_emulate.restype = ctypes.c_int
    lib.tile_emulate.argtypes = [ctypes.POINTER(hip.Op), ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
    return lib


def emulate(lib, n, sv, gates, tile_qubits=0):
    arr, keep = hip.pack_ops(ops.lower_all(gates, exact=True))
    v = np.ascontiguousarray(sv, dtype=np.complex128).copy()
    counts = (ctypes.c_size_t * 4)()
    rc = lib.tile_emulate(arr, len(gates) and len(arr), n, tile_qubits, v.ctypes.data, counts)
    del keep
    assert rc == 0, (rc, hip.load().aqs_last_error())
    return v, tuple(counts)


def emulate_ops(lib, n, sv, lowered, tile_qubits=0):
    arr, keep = hip.pack_ops(lowered)
    v = np.ascontiguousarray(sv, dtype=np.complex128).copy()
    counts = (ctypes.c_size_t * 4)()
    rc = lib.tile_emulate(arr, len(lowered) and len(arr), n, tile_qubits, v.ctypes.data, counts)
    del keep
    assert rc == 0, (rc, hip.load().aqs_last_error())
    return v, tuple(counts)


def check(lib, n, gates, seed=0, tile_qubits=0):
    sv = random_state(n, seed)
    want = oracle.run(sv, n, gates, threads=8)
    got, counts = emulate(lib, n, sv, gates, tile_qubits)
    np.testing.assert_allclose(got, want, atol=1e-12, rtol=0)
    return counts


def random_all_kinds(n, count, seed):
    """Random placements of every gate kind (Swaps included)."""
    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(count):
        kind, q, *ang = ALL_KINDS_4Q[int(rng.integers(len(ALL_KINDS_4Q)))]
        qs = rng.choice(n, size=len(q), replace=False)
        angle = float(rng.uniform(-3, 3)) if ang else 0.0
        gates.append(kind_gate((kind, tuple(int(x) for x in qs), angle) if ang else (kind, tuple(int(x) for x in qs))))
    return gates


@pytest.mark.parametrize("n,tile", [(10, 0), (12, 10), (13, 11), (14, 13), (15, 12)])
def test_tiled_random_layered(emu, n, tile):
    passes, tiles, segs, _ = check(emu, n, circuits.random_layered(n, 8, seed=n), seed=1, tile_qubits=tile)
    assert tiles >= 1 and segs >= tiles


@pytest.mark.parametrize("n", [11, 14, 16])
def test_tiled_brickwork(emu, n):
    gates = circuits.brickwork(n, 6, seed=3)
    passes, tiles, segs, _ = check(emu, n, gates, seed=2)
    # one sweep per tile pass, far fewer than gates
    assert passes < len(gates) / 4


@pytest.mark.parametrize("seed", range(6))
def test_tiled_all_kinds(emu, seed):
    n = 12 + seed % 3
    check(emu, n, random_all_kinds(n, 120, seed), seed=seed, tile_qubits=10 + seed % 4)


def test_tiled_qft_with_swaps(emu):
    n = 13
    check(emu, n, circuits.fourier_transform(n, final_swaps=True), seed=5)


def test_tiled_controls_outside_tile(emu):
    """Controls and diagonal bits on global and thread-held positions."""
    n = 16
    g = []
    for t in range(n):
        g.append(oracle.gate("H", t))
    g += [oracle.gate("CX", 0, 15), oracle.gate("CX", 15, 0), oracle.gate("CCX", 1, 14, 7),
          oracle.gate("CCX", 7, 2, 13), oracle.gate("CPhase", 3, 12, angle=0.7), oracle.gate("CZ", 0, 1),
          oracle.gate("CRotY", 2, 11, angle=1.3), oracle.gate("CH", 13, 3), oracle.gate("Or", 0, 9, 5),
          oracle.gate("Or", 5, 6, 15), oracle.gate("CSwap", 0, 4, 12), oracle.gate("CSwap", 14, 2, 3),
          oracle.gate("CRotZ", 15, 8, angle=-0.4), oracle.gate("CY", 6, 0), oracle.gate("Phase", 0, angle=0.2)]
    for t in range(n):
        g.append(oracle.gate("RotX", t, angle=0.1 * t + 0.05))
    check(emu, n, g, seed=9)


def test_tiled_dense_blocks_stay_plain(emu):
    n = 12
    U = np.linalg.qr(np.random.default_rng(1).normal(size=(8, 8)) + 1j)[0]
    g = [oracle.gate("H", q) for q in range(n)]
    g.append(oracle.gate("Circuit", 4, k=3, u=U))
    g += [oracle.gate("RotY", q, angle=0.3 * q) for q in range(n)]
    g.append(oracle.gate("ControlCircuit", 0, 6, k=3, u=U))
    g += [oracle.gate("CZ", q, q + 1) for q in range(n - 1)]
    check(emu, n, g, seed=4)


def test_tiled_permutations_bit_exact(emu):
    """Swaps relabel, X / CX / Toffoli are exchanges: the amplitudes only move."""
    n = 14
    rng = np.random.default_rng(7)
    g = []
    for _ in range(60):
        k = int(rng.integers(4))
        a, b, c = (int(x) for x in rng.choice(n, 3, replace=False))
        g.append(oracle.gate(("X", "CX", "CCX", "Swap")[k], a, b, c))
    sv = random_state(n, 3)
    want = oracle.run(sv, n, g)
    got, _ = emulate(emu, n, sv, g)
    assert np.array_equal(got, want)


def test_tiled_empty_and_small(emu):
    sv = random_state(11, 0)
    got, counts = emulate(emu, 11, sv, [])
    assert counts[0] == 0 and np.array_equal(got, sv)
    # below 10 qubits the dense-block schedule is used
    gates = [oracle.gate("H", 0), oracle.gate("CX", 0, 5), oracle.gate("RotY", 3, angle=0.4)]
    sv9 = random_state(9, 1)
    got, _ = emulate(emu, 9, sv9, gates)
    np.testing.assert_allclose(got, oracle.run(sv9, 9, gates), atol=1e-12)


def test_tiled_bad_width():
    L = hip.load()
    arr, keep = hip.pack_ops([])
    h = ctypes.c_void_p()
    assert L.aqs_fuse_tiled(arr, 0, 12, 9, ctypes.byref(h)) == 1
    assert L.aqs_fuse_tiled(arr, 0, 12, 14, ctypes.byref(h)) == 1


def test_tiled_cancelling_phases(emu):
    """Diagonal runs whose terms multiply to exactly 1 leave no pass behind."""
    n = 12
    g = [oracle.gate("CZ", 0, 1), oracle.gate("CZ", 1, 0), oracle.gate("Z", 3), oracle.gate("Z", 3)]
    sv = random_state(n, 2)
    got, counts = emulate(emu, n, sv, g)
    assert counts[0] == 0
    np.testing.assert_allclose(got, oracle.run(sv, n, g), atol=1e-15)
    g.append(oracle.gate("H", 5))
    got, counts = emulate(emu, n, sv, g)
    assert counts[1] == 1
    np.testing.assert_allclose(got, oracle.run(sv, n, g), atol=1e-12)


def test_auto_mode():
    """aqs_fuse_auto: blocks below 20 qubits, else the lower modelled cost."""
    small = ops.lower_all(circuits.brickwork(19, 6, seed=1))
    assert hip.Schedule(small, 19, mode="auto").mode == "blocks"
    for n, gates in ((24, circuits.brickwork(24, 8, seed=1)), (24, circuits.fourier_transform(24, True))):
        lowered = ops.lower_all(gates)
        a = hip.Schedule(lowered, n, mode="blocks")
        b = hip.Schedule(lowered, n, mode="tiled")
        c = hip.Schedule(lowered, n, mode="auto")
        assert c.cost() == min(a.cost(), b.cost())
        assert c.mode == ("tiled" if b.cost() < a.cost() else "blocks")


def test_auto_mode_errors():
    """aqs_fuse_auto fuses the tile schedule on a thread of its own: a bad
    argument or op still reports its status and message on the caller's
    thread, and leaves no schedule behind."""
    L = hip.load()
    arr, keep = hip.pack_ops(ops.lower_all(circuits.brickwork(22, 2, seed=1)))
    assert L.aqs_fuse_auto(arr, len(arr), 22, None) == 1
    assert b"null output" in L.aqs_last_error()
    h = ctypes.c_void_p(1)
    bad, keep2 = hip.pack_ops([{"kind": hip.OP_MAT1, "targets": (30,), "controls": (), "m": [1, 0, 0, 1]}])
    assert L.aqs_fuse_auto(bad, 1, 22, ctypes.byref(h)) != 0
    assert L.aqs_last_error() and h.value is None
    # and a good call after them succeeds
    assert L.aqs_fuse_auto(arr, len(arr), 22, ctypes.byref(h)) == 0 and h.value
    assert L.aqs_schedule_destroy(h) == 0


@pytest.mark.parametrize("tile", [10, 12])
def test_tiled_folded_phase_terms(emu, tile):
    """Phase ladders between tile qubits and outer qubits are folded into one
    per-tile value per local bit (QFT, random CPhase / controlled-RZ mixes)."""
    n = 16
    counts = check(emu, n, circuits.fourier_transform(n, final_swaps=True), seed=3, tile_qubits=tile)
    assert counts[3] > 0
    rng = np.random.default_rng(tile)
    g = [oracle.gate("H", q) for q in range(n)]
    for _ in range(150):
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        kind = ("CPhase", "CRotZ", "CZ")[int(rng.integers(3))]
        g.append(oracle.gate(kind, a, b, angle=float(rng.uniform(-3, 3))))
    g += [oracle.gate("RotX", q, angle=0.2 * q) for q in range(n)]
    counts = check(emu, n, g, seed=4, tile_qubits=tile)
    assert counts[3] > 0


def _haar2(rng):
    z = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / abs(np.diag(r)))


@pytest.mark.parametrize("seed", range(4))
def test_tiled_one_qubit_phases_carried(emu, seed):
    """Uncontrolled one-qubit gates run as a phase on the target's 1-half
    followed by a real rotation; the rest of each gate's diagonal (global
    phases of general unitaries included) is carried to the qubit's next
    gate, swapped by X, kept across controls and diagonal gates, and written
    out as phase terms before any other gate on the qubit.  Chains of general
    unitaries (Y: |u00| = 0, diagonal and real matrices, a non-unitary
    matrix) mixed with every other kind on a few qubits must equal the
    oracle."""
    n = 12
    rng = np.random.default_rng(seed)
    few = [0, 5, 10, 11]
    gates, lowered = [], []
    for i in range(160):
        r = int(rng.integers(12))
        a, b, c = (int(x) for x in rng.choice(few, 3, replace=False))
        if r < 4:
            u = _haar2(rng)
            if r == 1:
                u = np.diag(np.exp(1j * rng.uniform(-3, 3, size=2)))  # diagonal, as a matrix
            if r == 2 and i % 7 == 0:
                u = 0.5 * u  # not unitary: stays a complex 2x2
            gates.append(oracle.gate("Circuit", a, k=1, u=u))
            lowered.append({"kind": hip.OP_MAT1, "targets": (a,), "controls": (), "m": u.reshape(4)})
            continue
        kind = ("Y", "H", "X", "RotZ", "Phase", "CX", "CZ", "CRotY")[r - 4]
        g = oracle.gate(kind, a, b, c, angle=float(rng.uniform(-3, 3)))
        gates.append(g)
        lowered += ops.lower_all([g], exact=True)
    sv = random_state(n, seed)
    want = oracle.run(sv, n, gates)
    for tile in (10, 12):
        got, counts = emulate_ops(emu, n, sv, lowered, tile)
        np.testing.assert_allclose(got, want, atol=1e-12, rtol=0)


def test_tiled_u3_chain_with_diagonals(emu):
    """A U3 chain on one qubit with RotZ and CZ between: the carried phases
    pass the diagonal gates and end as phase terms after the last U3."""
    n = 12
    g = []
    for k in range(6):
        g.append(oracle.gate("U3", 3, angle=0.3 + k, phi=0.7 * k, lam=-0.2 * k))
        g.append(oracle.gate("RotZ", 3, angle=0.1 * k + 0.05))
        g.append(oracle.gate("CZ", 3, 4))
    check(emu, n, g, seed=6)
