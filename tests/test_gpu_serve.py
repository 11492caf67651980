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


def _mixed_gates(n, seed):
    """Brickwork plus swaps, controlled swaps, Toffoli, any-of-controls and
    controlled phases: every op kind a coalesced group runs."""
    rng = np.random.default_rng(seed)
    g = circuits.brickwork(n, 3, seed=seed)
    for _ in range(3 * n):
        a, b, c = (int(x) for x in rng.choice(n, size=3, replace=False))
        kind = ["Swap", "CSwap", "CCX", "Or", "CRotZ", "CRotY", "CX"][int(rng.integers(7))]
        if kind == "Swap":
            g.append(circuits.gate("Swap", a, b))
        elif kind in ("CSwap", "CCX", "Or"):
            g.append(circuits.gate(kind, a, b, c))
        else:
            g.append(circuits.gate(kind, a, b, angle=float(rng.uniform(0, 6.28))))
    return g


def test_small_requests_coalesced():
    # one worker held by a 22-qubit request: the small requests queued behind
    # it run as one group (one batched launch, no host fusion)
    rng = np.random.default_rng(11)
    reqs = []
    for rid in range(1, 21):
        n = int(rng.integers(6, 17))
        g = _mixed_gates(n, seed=200 + rid) if rid % 3 == 0 else circuits.brickwork(n, int(rng.integers(2, 9)),
                                                                                   seed=300 + rid)
        reqs.append((rid, n, g, rng.uniform(size=300), hip.PackedOps(ops.lower_all(g))))
    srv = hip.Server(budget_bytes=8 << 30, workers=1)
    srv.submit(0, 24, hip.PackedOps(ops.lower_all(circuits.brickwork(24, 60, seed=3))), [0.5])
    for rid, n, g, u, packed in reqs:  # packed beforehand: these submits take microseconds
        srv.submit(rid, n, packed, u, keep_state=True)
    srv.wait(0)
    got = {rid: srv.wait(rid) for rid, *_ in reqs}
    st = srv.stats()
    assert st["grouped"] == 20 and st["groups"] <= 2, st
    reqs = [r[:4] for r in reqs]
    for rid, n, g, u in reqs:
        res, idx, state = got[rid]
        want = oracle.run(oracle.basis_state(n, 0), n, g)
        assert oracle.infidelity(state, want) < 1e-9, (rid, n)
        assert np.array_equal(idx.astype(np.uint32), oracle.sample_c64(state, n, u)), (rid, n)
        assert abs(res["norm2"] - oracle.norm2(state.astype(np.complex128), n)) < 1e-9
        assert res["passes"] >= 1 and res["t_admit"] >= res["t_submit"]
    srv.close()
    # a member's result does not depend on its group: alone (a group of one)
    # it is bit-identical
    alone = hip.Server(budget_bytes=8 << 30, workers=1)
    for rid, n, g, u in reqs[:6]:
        alone.submit(rid, n, ops.lower_all(g), u, keep_state=True)
        _, idx, state = alone.wait(rid)
        assert np.array_equal(state.view(np.uint32), got[rid][2].view(np.uint32))
        assert np.array_equal(idx, got[rid][1])
    assert alone.stats()["grouped"] == 6
    alone.close()


def test_grouping_limits_and_off():
    big = hip.PackedOps(ops.lower_all(circuits.brickwork(24, 60, seed=3)))
    small = [(rid, 10, circuits.brickwork(10, 4, seed=rid)) for rid in range(1, 8)]
    packed = {rid: hip.PackedOps(ops.lower_all(g)) for rid, _, g in small}
    # a MATK op keeps a request out of groups; the cap splits the rest
    rng = np.random.default_rng(5)
    q, _ = np.linalg.qr(rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8)))
    matk = circuits.brickwork(10, 2, seed=9) + [circuits.gate("Circuit", 2, k=3, u=q)]
    matk_packed = hip.PackedOps(ops.lower_all(matk))
    srv = hip.Server(budget_bytes=8 << 30, workers=1)
    srv.set_grouping(16, 3)
    srv.submit(0, 24, big, [0.5])
    for rid, n, g in small:
        srv.submit(rid, n, packed[rid], [0.25, 0.75], keep_state=True)
    srv.submit(50, 10, matk_packed, [0.5], keep_state=True)
    for rid in [0] + [r for r, *_ in small] + [50]:
        res, _, state = srv.wait(rid)
        if rid == 50:
            want = oracle.run(oracle.basis_state(10, 0), 10, matk)
            assert oracle.infidelity(state, want) < 1e-9
    st = srv.stats()
    assert st["grouped"] == 7 and st["groups"] == 3, st
    with pytest.raises(hip.InvalidArgument):
        srv.set_grouping(17, 4)
    with pytest.raises(hip.InvalidArgument):
        srv.set_grouping(8, 0)
    srv.set_grouping(0, 64)  # off: the small requests take the fused path
    srv.submit(0, 24, big, [0.5])
    for rid, n, g in small:
        srv.submit(rid, n, packed[rid], [0.25, 0.75], keep_state=True)
    for rid, n, g in small:
        res, idx, state = srv.wait(rid)
        want = oracle.run(oracle.basis_state(n, 0), n, g)
        assert oracle.infidelity(state, want) < 1e-9
    srv.wait(0)
    assert srv.stats()["groups"] == 3
    srv.close()
