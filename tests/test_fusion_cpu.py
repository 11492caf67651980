"""CPU checks of the dense-block fuser (aqs_fuse) behind compile().

tests/cpp/tile_emu.cpp's block_emulate() builds the block schedule with the
library's host code and applies it to an FP64 host state: dense blocks as
their FP64 matrices, diagonal tables read the way diag.hip reads them, bit
permutations as index maps.  Each result must equal the oracle's per-gate
FP64 run to 1e-12, so a gate reordered past one it does not commute with, a
phase term put in the wrong table slot, or a lost relabelling fails.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
from afquantumsim_amd import circuits, hip, ops
from tests.helpers import random_state
from tests.test_tiling import EMU, random_all_kinds


@pytest.fixture(scope="module")
def emu():
    if not os.path.exists(EMU):
        pytest.fail("tests/cpp/libtile_emu.so is not built (run __graft_entry__.build())")
    hip.load()
    lib = ctypes.CDLL(EMU)
    lib.block_emulate.restype = ctypes.c_int
    lib.block_emulate.argtypes = [ctypes.POINTER(hip.Op), ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
    return lib


def check(lib, n, gates, kmax=5, seed=0):
    sv = random_state(n, seed)
    want = oracle.run(sv, n, gates, threads=8)
    arr, keep = hip.pack_ops(ops.lower_all(gates, exact=True))
    v = np.ascontiguousarray(sv, dtype=np.complex128).copy()
    counts = (ctypes.c_size_t * 4)()
    rc = lib.block_emulate(arr, len(arr) if gates else 0, n, kmax, v.ctypes.data, counts)
    del keep
    assert rc == 0, (rc, hip.load().aqs_last_error())
    np.testing.assert_allclose(v, want, atol=1e-12, rtol=0)
    return tuple(counts)


@pytest.mark.parametrize("n", [6, 11, 16])
def test_blocks_qft(emu, n):
    passes, blocks, tables, perms = check(emu, n, circuits.fourier_transform(n, final_swaps=True), seed=n)
    assert perms <= 2


@pytest.mark.parametrize("n", [20, 30])
def test_qft_tables_anchor_on_preceding_block(n):
    """Every table of the QFT schedule takes as anchor the qubits of the
    block right before it, so it can ride that block's store.  In the last
    table every qubit is in equally many terms (those of the last two
    blocks); that tie goes to the preceding block too."""
    hip.load()
    ps = hip.Schedule(ops.lower_all(circuits.fourier_transform(n, final_swaps=True)), n, 5).passes()
    tables = [i for i, (k, _) in enumerate(ps) if k == hip.OP_DIAGT]
    assert len(tables) == (n + 4) // 5 - 1
    for i in tables:
        assert ps[i - 1][0] == hip.OP_MATK
        assert sorted(ps[i][1]) == sorted(ps[i - 1][1]), (i, ps[i][1], ps[i - 1][1])


@pytest.mark.parametrize("n", [6, 11, 16])
def test_blocks_inverse_qft(emu, n):
    """Phases-then-Hadamard ladders are hoisted into Hadamard-then-phases
    order: one block per five qubits and one table between blocks, as for
    the forward transform (a per-qubit flush would give n tables)."""
    passes, blocks, tables, perms = check(emu, n, circuits.inverse_fourier_transform(n), seed=n + 1)
    assert blocks <= (n + 4) // 5 + 1 and tables <= (n + 4) // 5


@pytest.mark.parametrize("kmax", [2, 3, 4, 5])
@pytest.mark.parametrize("seed", range(3))
def test_blocks_all_kinds(emu, kmax, seed):
    n = 9 + seed
    check(emu, n, random_all_kinds(n, 150, seed + 10 * kmax), kmax=kmax, seed=seed)


def test_blocks_phase_heavy(emu):
    """Random mixes of phase gates around dense gates: deferred terms, their
    tables and the hoisted program order."""
    n = 13
    rng = np.random.default_rng(4)
    g = circuits.fourier_transform(n, final_swaps=False)
    for _ in range(200):
        kind = str(rng.choice(["RotZ", "Phase", "CZ", "CPhase", "CRotZ", "Z", "H", "RotY", "Swap", "CX"],
                              p=[0.12, 0.12, 0.12, 0.2, 0.12, 0.04, 0.08, 0.08, 0.06, 0.06]))
        q = [int(x) for x in rng.choice(n, size=3, replace=False)]
        g.append(circuits.gate(kind, q[0], q[1], q[2], angle=float(rng.uniform(-3, 3))))
    g += circuits.inverse_fourier_transform(n)
    check(emu, n, g, seed=7)


def test_blocks_hoisting_keeps_cheaper():
    """aqs_fuse keeps the cheaper of program order and hoisted order, so the
    inverse transform costs no more than the forward one without its swaps."""
    hip.load()
    for n in (16, 24):
        fwd = hip.Schedule(ops.lower_all(circuits.fourier_transform(n, final_swaps=False)), n, 5)
        inv = hip.Schedule(ops.lower_all(circuits.inverse_fourier_transform(n)), n, 5)
        assert inv.cost() <= fwd.cost() + 0.5


@pytest.mark.parametrize("n,depth,most", [(12, 8, 14), (17, 12, 32), (20, 20, 50)])
def test_blocks_brickwork_keep_growing(emu, n, depth, most):
    """A CZ between two open blocks that do not fit together closes only the
    full one and joins the other, so blocks keep growing instead of leaving
    as 2-3 qubit fragments.  Closing every touched block gave 21, 42 and 73
    passes here (17, 36 and 66 blocks, of which 5, 12 and 26 of 5 qubits);
    the state still equals the oracle's."""
    gates = circuits.brickwork(n, depth, seed=n)
    passes, blocks, tables, perms = check(emu, n, gates, seed=n)
    sch = hip.Schedule(ops.lower_all(gates), n, 5)
    sizes = [len(t) for k, t in sch.passes() if k == hip.OP_MATK]
    sch.close()
    assert passes <= most, passes
    assert sum(s == 5 for s in sizes) >= 0.6 * len(sizes), sizes


def test_blocks_keep_growing_random_controls(emu):
    """Random circuits of every gate kind with many 2-3 qubit gates across
    blocks (the partial close path with controlled swaps, Toffolis and
    any-of controls) equal the oracle at every block width."""
    for kmax in (2, 3, 4, 5):
        for seed in range(3):
            check(emu, 10, random_all_kinds(10, 150, seed=500 + seed), kmax=kmax, seed=seed)
