"""Batched persistent kernel (many small states in one launch) against the
FP64 oracle: final amplitudes to 2e-6 and <sum Z> to 1e-5 absolute."""
import numpy as np
import pytest

import oracle
from afquantumsim_amd import circuits, hip, ops

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    hip.init(0)
    yield


def _check(batch_spec, batch, expect):
    for i, (n, gates) in enumerate(batch_spec):
        want = oracle.run(oracle.basis_state(n, 0), n, gates)
        got = batch.state(i)
        np.testing.assert_allclose(got, want, atol=2e-6, err_msg=f"state {i} n={n}")
        assert abs(expect[i] - oracle.expect_sum_z(want, n)) < 1e-5


def test_batch_variational_small():
    spec = circuits.variational_batch(24, nmin=3, nmax=16, lmin=1, lmax=5, seed=8)
    b = hip.Batch([(n, ops.lower_all(g)) for n, g in spec])
    expect, ms = b.run()
    _check(spec, b, expect)
    # a second replay of the captured graph gives the same answer
    again, _ = b.run()
    assert np.array_equal(expect, again)


def test_batch_all_gate_kinds_and_autotune():
    rng = np.random.default_rng(3)
    spec = []
    for s in range(40):
        n = int(rng.integers(3, 15))
        gates = []
        for _ in range(30):
            kind = str(rng.choice(["H", "X", "Y", "Z", "RotX", "RotY", "RotZ", "Phase", "CX", "CZ", "CY",
                                   "CPhase", "Swap", "CSwap", "CCX", "Or", "CH", "CRotY"]))
            q = [int(x) for x in rng.choice(n, size=3, replace=False)]
            gates.append(circuits.gate(kind, q[0], q[1], q[2], angle=float(rng.uniform(-3, 3))))
        spec.append((n, gates))
    b = hip.Batch([(n, ops.lower_all(g)) for n, g in spec])
    tuned = b.autotune()
    assert tuned["block"] in (256, 512, 1024) and tuned["lds_qubits"] in (12, 13, 14)
    expect, _ = b.run()
    _check(spec, b, expect)


def test_batch_matches_single_state_path():
    """Same circuits through the per-state kernels give the same <sum Z>."""
    spec = circuits.variational_batch(6, nmin=10, nmax=16, lmin=4, lmax=12, seed=4)
    b = hip.Batch([(n, ops.lower_all(g)) for n, g in spec])
    expect, _ = b.run()
    for i, (n, g) in enumerate(spec):
        st = hip.State(n, hip.C64)
        st.apply(ops.lower_all(g))
        assert abs(st.expect_sum_z() - expect[i]) < 1e-5
        st.close()


@pytest.mark.parametrize("block,lds_qubits", [(256, 12), (512, 13), (1024, 14), (1024, 12)])
def test_batch_states_larger_than_the_tile(block, lds_qubits):
    """States of up to 20 qubits run as phases of LDS tiles: ansatz circuits
    and every op kind (controls inside and outside the tile, any-of controls,
    many-controlled gates) against the FP64 oracle and the numpy restatement."""
    from tests.test_batch_cpu import batch_ops, np_reference
    spec = [(n, circuits.ry_rz_ladder(n, L, np.random.default_rng(n + L))) for n, L in
            ((11, 5), (14, 6), (15, 8), (16, 12), (17, 3), (20, 2))]
    b = hip.Batch([(n, ops.lower_all(g)) for n, g in spec])
    b.configure(block, lds_qubits)
    assert b.info()["lds_qubits"] == lds_qubits
    expect, _ = b.run()
    _check(spec, b, expect)
    b.close()
    mixed = [(n, batch_ops(n, 60, 100 + n)) for n in (12, 14, 15, 16, 18)]
    b = hip.Batch(mixed)
    b.configure(block, lds_qubits)
    expect, _ = b.run()
    for i, (n, op_list) in enumerate(mixed):
        want = np_reference(op_list, n, oracle.basis_state(n, 0))
        np.testing.assert_allclose(b.state(i), want, atol=2e-6, err_msg=f"state {i} n={n}")
        assert abs(expect[i] - oracle.expect_sum_z(want, n)) < 1e-5
    again, _ = b.run()
    assert np.array_equal(expect, again)
    b.close()


def test_batch_configure_rejects_bad_shapes():
    spec = circuits.variational_batch(2, nmin=4, nmax=5, lmin=1, lmax=2, seed=1)
    b = hip.Batch([(n, ops.lower_all(g)) for n, g in spec])
    for blk, lq in ((128, 13), (512, 11), (512, 15)):
        with pytest.raises(RuntimeError):
            b.configure(blk, lq)
    b.close()
