"""Small states applied by the LDS-resident op runs of aqs_apply.

On a state of at most 12 qubits (complex64) or 11 (complex128), aqs_apply
sends each run of two or more non-MATK ops to one workgroup that holds the
state in LDS (csrc/gates.hip, k_small_run), 64 (40) ops per launch.  The
expected states come from the numpy restatement of tests/test_many_controls.py
(pinned there by known answers), extended here with AQS_CTRL_ANY: the gate
applies on the subspace where at least one control is 1.  The per-op kernels
(one aqs_apply call per op) are a second, independent check.
"""
import numpy as np
import pytest

from afquantumsim_amd import hip
from tests.helpers import random_state
from tests.test_many_controls import np_apply, unitary

pytestmark = pytest.mark.gpu


def np_apply_any(sv, n, op):
    """AQS_CTRL_ANY: the gate on every index with at least one control set."""
    full = np_apply(sv, n, dict(op, controls=()))
    psi = full.reshape((2,) * n)
    sel = [slice(None)] * n
    for c in op["controls"]:
        sel[c] = 0
    psi[tuple(sel)] = np.asarray(sv).reshape((2,) * n)[tuple(sel)]  # no control set: unchanged
    return psi.reshape(-1)


def np_run_any(sv, n, ops):
    v = np.asarray(sv, dtype=complex).copy()
    for op in ops:
        v = np_apply_any(v, n, op) if op.get("flags", 0) & hip.CTRL_ANY else np_apply(v, n, op)
    return v


def mixed_ops(n, count, seed, perm_only=False):
    """Every non-MATK kind at random positions, with 0..n-targets controls
    (beyond three: more_controls) and some any-of controlled ops."""
    rng = np.random.default_rng(seed)
    kinds = [hip.OP_X, hip.OP_SWAP] if perm_only else [hip.OP_MAT1, hip.OP_X, hip.OP_DIAG1, hip.OP_PHASE,
                                                        hip.OP_SWAP]
    ops = []
    while len(ops) < count:
        kind = int(rng.choice(kinds))
        if kind == hip.OP_SWAP and n < 2:
            continue
        nt = 2 if kind == hip.OP_SWAP else (int(rng.integers(1, min(3, n) + 1)) if kind == hip.OP_PHASE else 1)
        nc = int(rng.integers(0, n - nt + 1))
        qs = [int(x) for x in rng.choice(n, nt + nc, replace=False)]
        op = {"kind": kind, "targets": tuple(qs[:nt]), "controls": tuple(qs[nt:])}
        if 1 <= nc <= 3 and kind != hip.OP_PHASE and rng.random() < 0.25:
            op["flags"] = hip.CTRL_ANY
        if kind == hip.OP_MAT1:
            op["m"] = unitary(2, rng).reshape(-1)
        elif kind == hip.OP_DIAG1:
            d = np.exp(1j * rng.uniform(-3, 3, size=2))
            op["m"] = np.array([d[0], 0, 0, d[1]])
        elif kind == hip.OP_PHASE:
            op["m"] = np.array([np.exp(1j * rng.uniform(-3, 3)), 0, 0, 0])
        ops.append(op)
    return ops


def run(n, precision, sv, ops, one_call=True):
    st = hip.State(n, precision)
    st.upload(sv.astype(st.dtype))
    st.profile(True)
    if one_call:
        st.apply(ops)
    else:
        for op in ops:
            st.apply([op])
    prof = st.profile_read()
    got = st.download()
    st.close()
    return got, prof


CASES = [(hip.C64, n) for n in (1, 2, 3, 7, 10, 12)] + [(hip.C128, n) for n in (1, 4, 9, 11)]


@pytest.mark.parametrize("precision,n", CASES, ids=[f"{'c64' if p == hip.C64 else 'c128'}-{n}q" for p, n in CASES])
def test_small_runs_match_restatement(precision, n):
    ops = mixed_ops(n, 150, 100 + n)
    sv = random_state(n, n)
    want = np_run_any(sv, n, ops)
    got, prof = run(n, precision, sv, ops)
    chunk = 64 if precision == hip.C64 else 40
    assert prof.get("SMALL", (0.0, 0))[1] == -(-len(ops) // chunk), prof
    assert set(prof) == {"SMALL"}, prof
    np.testing.assert_allclose(got, want, atol=5e-6 if precision == hip.C64 else 1e-12, rtol=0)


@pytest.mark.parametrize("precision", [hip.C64, hip.C128], ids=["c64", "c128"])
def test_small_runs_match_per_op_kernels(precision):
    n = 10
    ops = mixed_ops(n, 200, 5)
    sv = random_state(n, 2)
    got, prof = run(n, precision, sv, ops)
    per_op, prof1 = run(n, precision, sv, ops, one_call=False)
    assert "SMALL" in prof and "SMALL" not in prof1
    np.testing.assert_allclose(got, per_op, atol=2e-6 if precision == hip.C64 else 1e-13, rtol=0)


@pytest.mark.parametrize("n", [2, 6, 12])
def test_small_permutation_runs_bit_exact(n):
    """Controlled X and SWAP runs move amplitudes bit for bit."""
    ops = mixed_ops(n, 130, 40 + n, perm_only=True)
    sv = random_state(n, 9).astype(np.complex64)
    want = np_run_any(sv.astype(complex), n, ops).astype(np.complex64)
    got, _ = run(n, hip.C64, sv, ops)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_small_run_between_matk_ops():
    """MATK ops split a run: the ops around them still go through LDS runs."""
    n = 8
    rng = np.random.default_rng(3)
    ops = mixed_ops(n, 30, 1) + [{"kind": hip.OP_MATK, "targets": (2, 5), "u": unitary(4, rng)}] + \
        mixed_ops(n, 30, 2) + [{"kind": hip.OP_MATK, "targets": (0, 7, 3), "u": unitary(8, rng)}] + mixed_ops(n, 5, 3)
    sv = random_state(n, 4)
    want = np_run_any(sv, n, ops)
    got, prof = run(n, hip.C64, sv, ops)
    assert prof["SMALL"][1] == 3 and prof["MATK"][1] == 2, prof
    np.testing.assert_allclose(got, want, atol=5e-6, rtol=0)


def test_small_run_error_applies_the_prefix():
    """An invalid op stops the run with the reference's error after the ops
    before it are applied, as on the per-op path."""
    n = 5
    ops = [{"kind": hip.OP_X, "targets": (0,)}, {"kind": hip.OP_X, "targets": (3,)},
           {"kind": hip.OP_X, "targets": (7,)}, {"kind": hip.OP_X, "targets": (1,)}]
    st = hip.State(n, hip.C64)
    st.set_basis(0)
    with pytest.raises(hip.OutOfRange):
        st.apply(ops)
    got = st.download()
    st.close()
    want = np.zeros(1 << n, dtype=np.complex64)
    want[(1 << (n - 1)) | (1 << (n - 1 - 3))] = 1
    assert np.array_equal(got, want)
