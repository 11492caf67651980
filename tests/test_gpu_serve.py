"""Request server: results are the same alone and in the mix (bit-exact:
the gate kernels are deterministic), match the FP64 oracle, and admission
respects the HBM budget."""
import numpy as np
import pytest

import oracle
from afquantumsim_amd import circuits, hip, ops

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    hip.init(0)
    yield


def _requests(count, nmin, nmax, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        depth = int(rng.integers(2, 8))
        gates = circuits.brickwork(n, depth, seed=1000 + i)
        u = rng.uniform(size=200)
        out.append((i, n, gates, u))
    return out


def _run(server, reqs, together):
    results = {}
    if together:
        for rid, n, g, u in reqs:
            server.submit(rid, n, ops.lower_all(g), u, keep_state=True)
        for rid, *_ in reqs:
            results[rid] = server.wait(rid)
    else:
        for rid, n, g, u in reqs:
            server.submit(rid, n, ops.lower_all(g), u, keep_state=True)
            results[rid] = server.wait(rid)
    return results


def test_alone_equals_in_the_mix_and_oracle():
    reqs = _requests(14, 6, 18, seed=1)
    alone = hip.Server(budget_bytes=8 << 30, workers=1)
    r1 = _run(alone, reqs, together=False)
    alone.close()
    mix = hip.Server(budget_bytes=8 << 30, workers=4)
    r2 = _run(mix, reqs, together=True)
    st = mix.stats()
    mix.close()
    assert st["completed"] == len(reqs) and st["max_concurrent"] > 1
    for rid, n, g, u in reqs:
        a, ia, sa = r1[rid]
        b, ib, sb = r2[rid]
        assert np.array_equal(ia, ib)
        assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))
        want = oracle.run(oracle.basis_state(n, 0), n, g)
        assert oracle.infidelity(sa, want) < 1e-9
        assert np.array_equal(ia.astype(np.uint32), oracle.sample_c64(sa, n, u))
        assert abs(a["norm2"] - 1.0) < 1e-5


def test_budget_admission_and_reuse():
    # room for about two 18-qubit states (2 MiB each plus scratch)
    reqs = [(i, 18, circuits.brickwork(18, 3, seed=i), np.linspace(0.01, 0.99, 50)) for i in range(6)]
    srv = hip.Server(budget_bytes=5 << 20, workers=4)
    res = _run(srv, reqs, together=True)
    st = srv.stats()
    assert st["completed"] == 6
    assert st["max_concurrent"] <= 2
    assert st["peak_resident_bytes"] <= st["budget_bytes"]
    with pytest.raises(hip.InvalidArgument):
        srv.submit(99, 22, [], [])  # 32 MiB state never fits a 5 MiB budget
    srv.close()
    ref = hip.Server(budget_bytes=1 << 30, workers=1)
    alone = _run(ref, reqs, together=False)
    ref.close()
    for rid, *_ in reqs:
        assert np.array_equal(res[rid][1], alone[rid][1])


def test_duplicate_id_rejected():
    srv = hip.Server(budget_bytes=1 << 30, workers=2)
    g = ops.lower_all(circuits.brickwork(8, 2, seed=0))
    srv.submit(7, 8, g, [0.5])
    with pytest.raises(hip.InvalidArgument):
        srv.submit(7, 8, g, [0.5])
    srv.wait(7)
    srv.close()


def test_cancel_queued_and_running():
    # one worker: a long 24-qubit request runs, a small one waits behind it
    srv = hip.Server(budget_bytes=1 << 30, workers=1)
    big = ops.lower_all(circuits.brickwork(24, 40, seed=5))
    small = ops.lower_all(circuits.brickwork(10, 3, seed=6))
    srv.submit(1, 24, big, [0.5])
    srv.submit(2, 10, small, [0.5])
    assert srv.cancel(2)
    with pytest.raises(hip.AqsError, match="cancelled"):
        srv.wait(2)
    srv.cancel(1)  # stops after its current stage, or it had already finished
    try:
        srv.wait(1)
    except hip.AqsError as e:
        assert "cancelled" in str(e)
    # the server keeps serving, with correct results
    srv.submit(3, 10, small, np.linspace(0.05, 0.95, 20), keep_state=True)
    res, idx, state = srv.wait(3)
    want = oracle.run(oracle.basis_state(10, 0), 10, circuits.brickwork(10, 3, seed=6))
    assert oracle.infidelity(state, want) < 1e-9
    assert np.array_equal(idx.astype(np.uint32), oracle.sample_c64(state, 10, np.linspace(0.05, 0.95, 20)))
    with pytest.raises(hip.OutOfRange):
        srv.wait(3)  # a result is handed over once
    assert srv.stats()["cancelled"] >= 1
    srv.close()
