"""Parity of the HIP path (called through the C ABI) with the FP64 oracle.

Tolerances: complex64 states agree with the FP64 oracle to 2e-6 absolute
per amplitude after a few gates (FP32 rounding of matrices and products),
complex128 states to 1e-12; permutation circuits are bit-exact; sampled
indices are bit-exact given the same amplitudes and uniforms.
"""
import math

import numpy as np
import pytest

import oracle
from afquantumsim_amd import circuits, hip, ops
from tests.helpers import (ALL_KINDS_4Q, cfequal, fixture_gates, init_qamps, kind_gate, load_kat,
                           random_state)

pytestmark = pytest.mark.gpu

KAT = load_kat()
TOL64 = 2e-6
TOL128 = 1e-12


@pytest.fixture(scope="module", autouse=True)
def _device():
    hip.init(0)
    yield


def gpu_run(n, sv, gates, precision=hip.C64, fused=None):
    st = hip.State(n, precision)
    st.upload(sv.astype(st.dtype))
    # complex128 runs use the exact FP64 operator; complex64 runs use the
    # reference's float32 matrix entries
    lowered = ops.lower_all(gates, exact=(precision == hip.C128))
    if fused:
        sch = hip.Schedule(lowered, n, fused)
        st.apply_schedule(sch)
    else:
        st.apply(lowered)
    out = st.download()
    st.close()
    return out


@pytest.mark.parametrize("precision", [hip.C64, hip.C128], ids=["c64", "c128"])
@pytest.mark.parametrize("case", KAT["gate_cases"], ids=lambda c: c["name"])
def test_kat_gates_gpu(case, precision):
    """test/tests.cpp:414-907 known answers, state built by aqs_state_product."""
    n = case["n"]
    st = hip.State(n, precision)
    st.product(init_qamps(case["init"]))
    st.apply(ops.lower_all(fixture_gates(case["gates"], n)))
    for idx, re, im in case["expect"]:
        got = st.amplitude(idx)
        if case["tol"] == 0:
            assert got == complex(re, im), (idx, got)
        else:
            assert cfequal(got, complex(re, im), case["tol"]), (idx, got, re, im)


@pytest.mark.parametrize("eq", KAT["equalities"], ids=lambda c: c["name"])
def test_kat_circuit_equalities_gpu(eq):
    """test/tests.cpp:909-961 on the device: ControlGate/Gate (MATK passes)
    equal their flattened circuits column by column."""
    n = eq["n"]
    lhs = fixture_gates(eq["lhs"], n)
    rhs = fixture_gates(eq["rhs"], n)
    for j in range(1 << n):
        e = oracle.basis_state(n, j)
        a = gpu_run(n, e, lhs)
        b = gpu_run(n, e, rhs)
        np.testing.assert_allclose(a, b, atol=TOL64)


@pytest.mark.parametrize("case", KAT["ch_cases"], ids=lambda c: c["name"])
def test_kat_ch_gpu(case):
    n = case["n"]
    st = hip.State(n, hip.C64)
    st.product(init_qamps(case["init"]))
    st.apply(ops.lower_all(fixture_gates(case["gates"], n)))
    ref = oracle.product_state(n, init_qamps(case["expect_product"]))
    np.testing.assert_allclose(st.download(), ref, atol=1e-6)


@pytest.mark.parametrize("case", KAT["prob_cases"], ids=lambda c: c["name"])
def test_kat_probabilities_gpu(case):
    n = case["n"]
    st = hip.State(n, hip.C64)
    st.product(init_qamps(case["init"]))
    for q, p in enumerate(case["p1"]):
        got = st.prob1(q)
        if case["tol"] == 0:
            assert got == p
        else:
            assert abs(got - p) <= max(abs(got), p) * case["tol"]


@pytest.mark.parametrize("case", KAT["measure_cases"], ids=lambda c: c["name"])
def test_kat_peek_measure_all_gpu(case):
    n = case["n"]
    st = hip.State(n, hip.C64)
    st.product(init_qamps(case["init"]))
    u = np.array([0.0, 0.25, 0.5, 0.999999])
    assert (st.sample(u) == case["peek_all"]).all()


def _all_position_gates(n, rng):
    """Every gate kind at random positions, including bit 0 targets/controls."""
    out = []
    for spec in ALL_KINDS_4Q:
        g = kind_gate(spec)
        k = spec[0]
        arity = len(spec[1])
        for _ in range(3):
            q = rng.choice(n, size=arity, replace=False)
            q = tuple(int(x) for x in q) + (0,) * (3 - arity)
            if rng.random() < 0.3:
                # put something on the least significant bit (qubit n-1)
                q = (n - 1,) + q[1:] if arity >= 1 else q
                if len(set(q[:arity])) < arity:
                    continue
            g2 = dict(g)
            g2["q"] = q
            out.append(g2)
    return out


@pytest.mark.parametrize("n", [3, 5, 8, 13])
@pytest.mark.parametrize("precision", [hip.C64, hip.C128], ids=["c64", "c128"])
def test_every_kind_every_position(n, precision):
    rng = np.random.default_rng(100 + n)
    gates = _all_position_gates(n, rng)
    if n < 3:
        gates = [g for g in gates if g["kind"] not in ("CSwap", "CCX", "Or")]
    sv = random_state(n, 7 + n)
    for g in gates:
        if max(g["q"][: {"Swap": 2, "CX": 2, "CY": 2, "CZ": 2, "CH": 2, "CPhase": 2, "CRotX": 2,
                          "CRotY": 2, "CRotZ": 2, "CSwap": 3, "CCX": 3, "Or": 3}.get(g["kind"], 1)]) >= n:
            continue
        want = oracle.run(sv, n, [g])
        got = gpu_run(n, sv, [g], precision)
        tol = TOL64 if precision == hip.C64 else TOL128
        np.testing.assert_allclose(got, want, atol=tol, err_msg=str(g))


@pytest.mark.parametrize("n", [4, 9, 14])
def test_permutation_circuits_bit_exact(n):
    """X / CX / CCX / Swap / CSwap / Or only: the complex64 result is the
    exact permutation of the complex64 input."""
    rng = np.random.default_rng(31 + n)
    gates = []
    for _ in range(60):
        kind = str(rng.choice(["X", "CX", "CCX", "Swap", "CSwap", "Or"]))
        q = rng.choice(n, size=3, replace=False)
        gates.append(oracle.gate(kind, int(q[0]), int(q[1]), int(q[2])))
    sv = random_state(n, 5).astype(np.complex64)
    want = oracle.run(sv.astype(np.complex128), n, gates).astype(np.complex64)
    got = gpu_run(n, sv.astype(np.complex128), gates)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("n", [10, 16])
def test_random_layered_circuit(n):
    """The 20-qubit workload's gate vocabulary at smaller n, FP64 oracle."""
    gates = circuits.random_layered(n, 12, seed=n)
    sv = oracle.basis_state(n, 0)
    want = oracle.run(sv, n, gates)
    got = gpu_run(n, sv, gates)
    assert oracle.infidelity(got, want) < 1e-9
    got128 = gpu_run(n, sv, gates, hip.C128)
    # 1 - |<a|b>|^2 itself carries ~1e-16 of rounding in FP64
    assert oracle.infidelity(got128, want) < 1e-14
    np.testing.assert_allclose(got128, want, atol=1e-12)


@pytest.mark.parametrize("kmax", [2, 3, 4, 5])
@pytest.mark.parametrize("n", [9, 14])
def test_fused_brickwork(n, kmax):
    """compile() fusion: fused schedule == per-gate oracle."""
    gates = circuits.brickwork(n, 6, seed=3)
    sv = oracle.basis_state(n, 0)
    want = oracle.run(sv, n, gates)
    got = gpu_run(n, sv, gates, fused=kmax)
    assert oracle.infidelity(got, want) < 1e-9
    np.testing.assert_allclose(got, want, atol=5e-6)
    got128 = gpu_run(n, sv, gates, hip.C128, fused=kmax)
    np.testing.assert_allclose(got128, want, atol=1e-12)


@pytest.mark.parametrize("kmax", [3, 4, 5])
@pytest.mark.parametrize("precision", [hip.C64, hip.C128], ids=["c64", "c128"])
def test_fused_diagonal_deferral(kmax, precision):
    """Circuits dominated by phase gates (QFT ladders, RZ, Phase, CZ, CPhase,
    CRotZ with its DIAG1 global phase) exercise deferred terms and the
    diagonal table passes of compile()."""
    n = 13
    rng = np.random.default_rng(kmax)
    gates = circuits.fourier_transform(n, final_swaps=False)
    for _ in range(120):
        kind = str(rng.choice(["RotZ", "Phase", "CZ", "CPhase", "CRotZ", "Z", "H", "RotY", "Swap"],
                              p=[0.15, 0.15, 0.15, 0.2, 0.15, 0.05, 0.05, 0.05, 0.05]))
        q = [int(x) for x in rng.choice(n, size=3, replace=False)]
        gates.append(circuits.gate(kind, q[0], q[1], q[2], angle=float(rng.uniform(-3, 3))))
    gates += circuits.inverse_fourier_transform(n)
    sv = random_state(n, 19)
    want = oracle.run(sv, n, gates)
    sch = hip.Schedule(ops.lower_all(gates, exact=(precision == hip.C128)), n, kmax)
    kinds = {k for k, _ in sch.passes()}
    assert hip.OP_DIAGT in kinds
    got = gpu_run(n, sv, gates, precision, fused=kmax)
    tol = 5e-6 if precision == hip.C64 else 1e-12
    np.testing.assert_allclose(got, want, atol=tol)


@pytest.mark.parametrize("name,n", [("qft", 14), ("qft", 20), ("iqft", 16), ("iqft", 21)])
def test_fused_tables_ride_mfma_blocks(name, n):
    """A diagonal table next to a 4/5-qubit MFMA block whose tile holds its
    anchor bits is applied in that block's store loop (table after the
    block) or at its load (table before it): fewer table launches than
    table passes, same state as the oracle."""
    gates = (circuits.fourier_transform(n, final_swaps=True) if name == "qft"
             else circuits.inverse_fourier_transform(n))
    sv = random_state(n, n)
    want = oracle.run(sv, n, gates, threads=8)
    sch = hip.Schedule(ops.lower_all(gates), n, 5)
    tables = sum(k == hip.OP_DIAGT for k, _ in sch.passes())
    st = hip.State(n, hip.C64)
    st.upload(sv.astype(st.dtype))
    st.profile(True)
    st.apply_schedule(sch)
    launched = st.profile_read().get("DIAGT", (0.0, 0))[1]
    got = st.download()
    st.close()
    assert tables >= 1 and launched < tables
    assert oracle.infidelity(got, want) < 1e-9
    np.testing.assert_allclose(got, want, atol=5e-6)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_tables_ride_wide_blocks(seed):
    """A 5-qubit block that carries a diagonal table keeps the one-wave
    tiles at 19 qubits, where plain blocks take the wide ones.  QFT ladders
    without the final swaps (the last block's table comes before it) and
    random phase gates between them (anchors and required bits in and
    outside the tile; tables that cannot ride get a pass of their own),
    against the oracle."""
    n = 19
    rng = np.random.default_rng(seed)
    gates = circuits.fourier_transform(n, final_swaps=False)
    for _ in range(60):
        kind = str(rng.choice(["RotZ", "Phase", "CZ", "CPhase", "CRotZ", "H", "RotY"],
                              p=[0.15, 0.15, 0.15, 0.25, 0.2, 0.05, 0.05]))
        q = [int(x) for x in rng.choice(n, size=3, replace=False)]
        gates.append(circuits.gate(kind, q[0], q[1], q[2], angle=float(rng.uniform(-3, 3))))
    gates += circuits.fourier_transform(n, final_swaps=False)
    sv = random_state(n, 40 + seed)
    want = oracle.run(sv, n, gates, threads=8)
    sch = hip.Schedule(ops.lower_all(gates), n, 5)
    tables = sum(k == hip.OP_DIAGT for k, _ in sch.passes())
    st = hip.State(n, hip.C64)
    st.upload(sv.astype(st.dtype))
    st.profile(True)
    st.apply_schedule(sch)
    launched = st.profile_read().get("DIAGT", (0.0, 0))[1]
    got = st.download()
    st.close()
    assert tables >= 1 and launched < tables
    assert oracle.infidelity(got, want) < 1e-9
    np.testing.assert_allclose(got, want, atol=5e-6)


@pytest.mark.parametrize("n", [3, 7, 12, 17])
def test_fused_swaps_are_relabelled_bit_exact(n):
    """Swap-only circuits compile to at most two in-place bit-permutation
    passes; the result is the exact permutation of the input."""
    rng = np.random.default_rng(n)
    gates = []
    for _ in range(40):
        a, b = (int(x) for x in rng.choice(n, size=2, replace=False))
        gates.append(oracle.gate("Swap", a, b))
    sch = hip.Schedule(ops.lower_all(gates), n, 5)
    kinds = [k for k, _ in sch.passes()]
    assert set(kinds) <= {hip.OP_PERMB} and len(kinds) <= 2
    sv = random_state(n, 3).astype(np.complex64)
    want = oracle.run(sv.astype(np.complex128), n, gates).astype(np.complex64)
    for prec in (hip.C64, hip.C128):
        st = hip.State(n, prec)
        st.upload(sv.astype(st.dtype))
        st.apply_schedule(sch)
        got = st.download().astype(np.complex64)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def restore_case(name, n):
    """Circuits whose schedule ends with an MFMA block then the restore
    permutation: the QFT with its swaps (block with a table prologue, tile
    pairs), a bit reversal before a 5-qubit block (tile pairs), a swap of
    bits 0 and 5 before a block on bits 1..5 (every tile its own partner)."""
    U = np.linalg.qr(np.random.default_rng(n).normal(size=(32, 32)) + 1j * np.random.default_rng(n + 1).normal(size=(32, 32)))[0]
    if name == "qft":
        return circuits.fourier_transform(n, final_swaps=True)
    if name == "reverse":
        g = [oracle.gate("H", q) for q in range(n)]
        g += [oracle.gate("Swap", i, n - 1 - i) for i in range(n // 2)]
        return g + [oracle.gate("Circuit", n - 5, k=5, u=U)]
    g = [oracle.gate("RotY", q, angle=0.1 + 0.2 * q) for q in range(n)]
    g.append(oracle.gate("Swap", n - 1, n - 6))
    return g + [oracle.gate("Circuit", n - 5, k=5, u=U)]


@pytest.mark.parametrize("name,n,folded", [("qft", 20, True), ("qft", 14, False), ("reverse", 16, True),
                                           ("swap05", 16, True)])
def test_mfma_store_folds_restore_permutation(name, n, folded):
    """A restore permutation that maps the last MFMA block's tile onto itself
    runs in that block's store (no permutation launch); the state equals the
    oracle's."""
    gates = restore_case(name, n)
    sv = random_state(n, 2 * n)
    want = oracle.run(sv, n, gates, threads=8)
    sch = hip.Schedule(ops.lower_all(gates), n, 5)
    kinds = [k for k, _ in sch.passes()]
    assert kinds[-1] == hip.OP_PERMB and kinds[-2] == hip.OP_MATK
    st = hip.State(n, hip.C64)
    st.upload(sv.astype(st.dtype))
    st.profile(True)
    st.apply_schedule(sch)
    prof = st.profile_read()
    got = st.download()
    st.close()
    assert (prof.get("PERMB", (0.0, 0))[1] == 0) == folded, prof
    assert oracle.infidelity(got, want) < 1e-9
    np.testing.assert_allclose(got, want, atol=5e-6)


def structured_swaps(n, name):
    """Swap lists whose restore permutations need the bit-permutation tile
    grown beyond bits 0..4 and their partners."""
    if name == "reverse":
        return [oracle.gate("Swap", i, n - 1 - i) for i in range(n // 2)]
    if name == "rotate":
        return [oracle.gate("Swap", i, i + 1) for i in range(n - 1)]
    if name == "middle":  # index bits 5..n-6 reversed, bits 0..4 fixed
        return [oracle.gate("Swap", 5 + i, n - 6 - i) for i in range((n - 10) // 2)]
    # low qubits exchanged among themselves only
    return [oracle.gate("Swap", n - 1, n - 2), oracle.gate("Swap", n - 3, n - 5)]


@pytest.mark.parametrize("n", [11, 16, 21])
@pytest.mark.parametrize("name", ["reverse", "rotate", "middle", "low"])
def test_fused_structured_permutations_bit_exact(n, name):
    gates = structured_swaps(n, name)
    sch = hip.Schedule(ops.lower_all(gates), n, 5)
    assert {k for k, _ in sch.passes()} <= {hip.OP_PERMB}
    sv = random_state(n, 5).astype(np.complex64)
    want = oracle.run(sv.astype(np.complex128), n, gates).astype(np.complex64)
    for prec in (hip.C64, hip.C128):
        st = hip.State(n, prec)
        st.upload(sv.astype(st.dtype))
        st.apply_schedule(sch)
        got = st.download().astype(np.complex64)
        st.close()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (n, name, prec)


@pytest.mark.parametrize("kmax", [2, 5])
def test_fused_swaps_mixed_with_gates(kmax):
    n = 12
    rng = np.random.default_rng(40 + kmax)
    gates = []
    for _ in range(150):
        kind = str(rng.choice(["Swap", "U3", "CX", "CZ", "CSwap", "CPhase", "H"]))
        q = [int(x) for x in rng.choice(n, size=3, replace=False)]
        gates.append(circuits.gate(kind, q[0], q[1], q[2], angle=float(rng.uniform(-3, 3)),
                                   phi=float(rng.uniform(-3, 3)), lam=float(rng.uniform(-3, 3))))
    sv = random_state(n, 44)
    want = oracle.run(sv, n, gates)
    np.testing.assert_allclose(gpu_run(n, sv, gates, fused=kmax), want, atol=5e-6)
    np.testing.assert_allclose(gpu_run(n, sv, gates, hip.C128, fused=kmax), want, atol=1e-12)


def _haar(k, rng):
    z = (rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))) / math.sqrt(2)
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


@pytest.mark.parametrize("k", [2, 3, 4, 5])
@pytest.mark.parametrize("begin", [0, 3, "low"])
def test_matk_blocks(k, begin):
    """Dense k-qubit blocks (MFMA for k = 4, 5) at high, middle and lowest bits."""
    n = 12
    rng = np.random.default_rng(k * 10 + (7 if begin == "low" else begin))
    U = _haar(k, rng)
    b = n - k if begin == "low" else begin
    g = oracle.gate("Circuit", b, k=k, u=U)
    sv = random_state(n, 3)
    want = oracle.run(sv, n, [g])
    got = gpu_run(n, sv, [g])
    np.testing.assert_allclose(got, want, atol=TOL64)
    gc = oracle.gate("ControlCircuit", (b - 1) % n if b > 0 else n - 1, b, k=k, u=U)
    if gc["q"][0] >= b + k or gc["q"][0] < b:
        want = oracle.run(sv, n, [gc])
        np.testing.assert_allclose(gpu_run(n, sv, [gc]), want, atol=TOL64)


@pytest.mark.parametrize("k", [4, 5])
def test_mfma_scattered_targets(k):
    """MFMA block on non-contiguous target bits, through the fuse path."""
    n = 13
    rng = np.random.default_rng(k)
    targets = sorted(int(x) for x in rng.choice(n, size=k, replace=False))
    U = _haar(k, rng)
    op = {"kind": hip.OP_MATK, "targets": tuple(targets), "u": U}
    sv = random_state(n, 11)
    # oracle: embed via a permutation to contiguous qubits
    st = hip.State(n, hip.C128)
    st.upload(sv)
    st.apply([op])
    want = st.download()
    st64 = hip.State(n, hip.C64)
    st64.upload(sv.astype(np.complex64))
    st64.apply([op])
    np.testing.assert_allclose(st64.download(), want, atol=TOL64)
    # independent check of the complex128 path against explicit algebra
    idx = np.arange(1 << n)
    ref = np.zeros_like(sv)
    pos = [n - 1 - t for t in targets]
    sub = np.zeros(1 << n, dtype=np.int64)
    for j, p in enumerate(pos):
        sub |= ((idx >> p) & 1) << (k - 1 - j)
    rest = idx.copy()
    for p in pos:
        rest &= ~(1 << p)
    for r in range(1 << k):
        off = 0
        for j, p in enumerate(pos):
            if (r >> (k - 1 - j)) & 1:
                off |= 1 << p
        mask = sub == r
        src = rest[mask]
        acc = np.zeros(src.size, dtype=np.complex128)
        for c in range(1 << k):
            offc = 0
            for j, p in enumerate(pos):
                if (c >> (k - 1 - j)) & 1:
                    offc |= 1 << p
            acc += U[r, c] * sv[src | offc]
        ref[src | off] = acc
    np.testing.assert_allclose(want, ref, atol=1e-12)


@pytest.mark.parametrize("n", [1, 3, 10, 17])
def test_sampling_bit_exact(n):
    rng = np.random.default_rng(n)
    sv = random_state(n, 40 + n).astype(np.complex64)
    st = hip.State(n, hip.C64)
    st.upload(sv)
    u = np.concatenate([rng.uniform(size=4000), [0.0, 1.0 - 2 ** -53, 1.0, 2.0]])
    got = st.sample(u)
    want = oracle.sample_c64(st.download(), n, u)
    assert np.array_equal(got.astype(np.uint32), want)
    # complex128: one amplitude per 16-byte load in the block sums
    st128 = hip.State(n, hip.C128)
    st128.upload(random_state(n, 50 + n))
    got = st128.sample(u)
    assert np.array_equal(got.astype(np.uint32), oracle.sample_c128(st128.download(), n, u))
    st128.close()


def test_sampling_sparse_and_clamp():
    n = 12
    sv = np.zeros(1 << n, dtype=np.complex64)
    sv[5] = 0.6
    sv[3000] = 0.8j
    st = hip.State(n, hip.C64)
    st.upload(sv)
    u = np.array([0.0, 0.3599999, 0.3600001, 0.999, 1.0, 5.0])
    assert list(st.sample(u)) == [5, 5, 3000, 3000, 3000, 3000]


@pytest.mark.parametrize("precision", [hip.C64, hip.C128], ids=["c64", "c128"])
def test_measurement_reductions(precision):
    n = 11
    st = hip.State(n, precision)
    sv = random_state(n, 77).astype(st.dtype).astype(np.complex128)  # the values the state holds
    st.upload(sv.astype(st.dtype))
    assert abs(st.norm2() - oracle.norm2(sv, n)) < 1e-13
    for q in range(n):
        assert abs(st.prob1(q) - oracle.prob1(sv, n, q)) < 1e-13
    np.testing.assert_allclose(st.prob1_all(), [oracle.prob1(sv, n, q) for q in range(n)], atol=1e-13)
    assert abs(st.expect_sum_z() - oracle.expect_sum_z(sv, n)) < 1e-12
    p = st.probabilities()
    np.testing.assert_allclose(p, np.abs(sv) ** 2, rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("precision", [hip.C64, hip.C128], ids=["c64", "c128"])
def test_measure_qubit_collapse(precision):
    """Qubit n-1 (bit 0) splits each 16-byte pair of complex64 amplitudes
    between kept and zeroed; the zeroed ones are exactly +0."""
    n = 6
    tol = 1e-13 if precision == hip.C128 else 2e-7
    for q in (0, 3, 5):
        for u in (0.01, 0.99):
            st = hip.State(n, precision)
            sv = random_state(n, 12).astype(st.dtype)
            st.upload(sv)
            sv = sv.astype(np.complex128)
            outcome, p1 = st.measure_qubit(q, u)
            assert abs(p1 - oracle.prob1(sv, n, q)) < 1e-13
            assert outcome == int(u < p1)
            m = 1 << (n - 1 - q)
            idx = np.arange(1 << n)
            keep = ((idx & m) != 0) == bool(outcome)
            ref = np.where(keep, sv, 0) / math.sqrt(p1 if outcome else 1 - p1)
            got = st.download()
            np.testing.assert_allclose(got, ref, atol=tol)
            assert not got[~keep].view(np.uint32 if precision == hip.C64 else np.uint64).any()


def test_state_preparation():
    n = 9
    qa = np.stack([np.array([math.cos(0.1 * q), np.exp(1j * q) * math.sin(0.1 * q)]) for q in range(n)])
    st = hip.State(n, hip.C128)
    st.product(qa)
    np.testing.assert_allclose(st.download(), oracle.product_state(n, qa), atol=1e-14)
    st.gaussian(2024)
    np.testing.assert_allclose(st.download(), oracle.gaussian_state(n, 2024), atol=1e-13)
    st.set_basis(17)
    e = st.download()
    assert e[17] == 1 and np.count_nonzero(e) == 1
    with pytest.raises(hip.InvalidArgument):
        st.upload(np.zeros(1 << n), normalize=True)


def test_errors_match_reference_types():
    st = hip.State(4, hip.C64)
    with pytest.raises(hip.OutOfRange):
        st.apply([{"kind": hip.OP_X, "targets": (4,)}])
    with pytest.raises(hip.InvalidArgument):
        st.apply([{"kind": hip.OP_X, "targets": (1,), "controls": (1,)}])
    with pytest.raises(hip.InvalidArgument):
        st.apply([{"kind": hip.OP_SWAP, "targets": (2, 2)}])
    with pytest.raises(hip.OutOfRange):
        st.prob1(9)


@pytest.mark.parametrize("fused", [None, 4, 5])
def test_qft_closed_form_gpu(fused):
    """fourier_transform(n) + swaps on a Gaussian state = bitrev(ifft(bitrev(v))) sqrt(N)."""
    n = 16
    N = 1 << n
    st = hip.State(n, hip.C64)
    st.gaussian(77)
    v = oracle.gaussian_state(n, 77)
    gates = circuits.fourier_transform(n, final_swaps=True)
    lowered = ops.lower_all(gates)
    if fused:
        st.apply_schedule(hip.Schedule(lowered, n, fused))
    else:
        st.apply(lowered)
    rev = np.array([int(format(i, f"0{n}b")[::-1], 2) for i in range(N)])
    ref = np.fft.ifft(v[rev])[rev] * math.sqrt(N)
    assert oracle.infidelity(st.download(), ref) < 1e-9


def test_autotune_keeps_state_and_results():
    """The launch-shape tuner runs the identity and leaves the state as is;
    every tuned variant then gives the oracle's answer."""
    n = 14
    sv = random_state(n, 90).astype(np.complex64)
    st = hip.State(n, hip.C64)
    st.upload(sv)
    variants = st.autotune(reps=2)
    assert len(variants) == n and all(0 <= v < 9 for v in variants)
    np.testing.assert_array_equal(st.download(), sv)
    gates = [oracle.gate("U3", q, angle=0.3 + q, phi=0.1 * q, lam=-0.2) for q in range(n)]
    want = oracle.run(sv.astype(np.complex128), n, gates)
    st.apply(ops.lower_all(gates))
    np.testing.assert_allclose(st.download(), want, atol=TOL64)


@pytest.mark.parametrize("n", [3, 13, 20])
def test_every_launch_shape(n):
    """Each one-qubit launch shape, pinned for every target, gives the
    oracle's answer (controls included); bad pins are rejected."""
    sv = random_state(n, 91).astype(np.complex64)
    gates = [oracle.gate("U3", q, angle=0.3 + q, phi=0.1 * q, lam=-0.2) for q in range(n)]
    gates += [oracle.gate("CRotY", q, (q + 1) % n, angle=0.7 - 0.1 * q) for q in range(n)]
    want = oracle.run(sv.astype(np.complex128), n, gates)
    st = hip.State(n, hip.C64)
    for v in range(9):
        for q in range(n):
            st.tune_set(q, v)
        st.upload(sv)
        st.apply(ops.lower_all(gates))
        np.testing.assert_allclose(st.download(), want, atol=TOL64, err_msg=f"variant {v}")
    with pytest.raises(hip.AqsError):
        st.tune_set(0, 9)
    with pytest.raises(hip.AqsError):
        st.tune_set(n, 0)
    for q in range(n):
        st.tune_set(q, 0)
    st.close()


def test_33_qubit_launch_sizes():
    """A 33-qubit complex64 state (64 GiB): 2^31..2^32 work units per pass,
    where the one-wave launch shapes give way to 64-bit multi-unit shapes.
    H on bits 0 and 1, CZ on them (a low-bit phase pass over 2^32 float4s),
    RotZ on bit 32 and X on bit 27 from |0..0> leave four nonzero
    amplitudes, known in closed form."""
    n = 33
    st = hip.State(n, hip.C64)
    st.set_basis(0)
    theta = 0.6
    gates = [oracle.gate("H", 32), oracle.gate("H", 31), oracle.gate("CZ", 31, 32),
             oracle.gate("RotZ", 0, angle=theta), oracle.gate("X", 5)]
    st.apply(ops.lower_all(gates))
    c, s = math.cos(np.float32(theta) / 2), math.sin(np.float32(theta) / 2)
    rz0 = complex(c, -s)
    base = 1 << 27
    for low, sign in ((0, 1), (1, 1), (2, 1), (3, -1)):
        got = st.amplitude(base | low)
        assert abs(got - 0.5 * sign * rz0) < 1e-6, (low, got)
    for idx in (0, 3, base | 4, (1 << 32) | base | 3, (1 << 33) - 1):
        assert abs(st.amplitude(idx)) == 0.0, idx
    assert abs(st.norm2() - 1.0) < 1e-6
    st.close()


@pytest.mark.parametrize("targets", [(13, 14, 15, 16, 17), (0, 1, 2, 3, 4), (2, 5, 9, 11, 16), (0, 6, 7, 8, 17),
                                     (17, 3, 12, 8, 1)])
def test_mfma_wide_tiles_18q(targets):
    """A 5-qubit block on 18 qubits runs on the wide tiles (its targets and
    the 7 lowest other bits): qubits at the top, the bottom, scattered, and
    listed out of order (the sub-index follows the op's target order)
    against the complex128 path, which multiplies group by group."""
    n = 18
    rng = np.random.default_rng(sum(targets))
    U = _haar(5, rng)
    op = {"kind": hip.OP_MATK, "targets": tuple(targets), "u": U}
    sv = random_state(n, 17)
    st = hip.State(n, hip.C128)
    st.upload(sv)
    st.apply([op])
    want = st.download()
    st.close()
    st64 = hip.State(n, hip.C64)
    st64.upload(sv.astype(np.complex64))
    st64.profile(True)
    st64.apply([op])
    prof = st64.profile_read()
    got = st64.download()
    st64.close()
    assert prof.get("MATK_MFMA", (0.0, 0))[1] == 1, prof
    np.testing.assert_allclose(got, want, atol=TOL64)
    assert oracle.infidelity(got, want) < 1e-10


@pytest.mark.parametrize("schedule", ["plain", "blocks5", "tiles"])
def test_brickwork_20q_depth20_against_oracle(schedule):
    """The measured workload's circuit (U3 on every qubit, CZ on alternating
    pairs, depth 20) at 20 qubits, on every execution path, against the FP64
    oracle gate for gate."""
    n = 20
    gates = circuits.brickwork(n, 20, seed=2)
    want = oracle.run(oracle.basis_state(n, 0), n, gates, threads=8)
    st = hip.State(n, hip.C64)
    st.set_basis(0)
    lowered = ops.lower_all(gates)
    if schedule == "plain":
        st.apply(lowered)
    elif schedule == "blocks5":
        st.apply_schedule(hip.Schedule(lowered, n, 5))
    else:
        st.apply_schedule(hip.Schedule(lowered, n, mode="tiled"))
    got = st.download()
    assert oracle.infidelity(got, want) < 1e-10
    assert np.abs(got - want).max() < 2e-6
    st.close()


def test_random_layers_20q_depth40_against_oracle():
    """The 20-qubit random-circuit workload (1200 gates) plain, against the oracle."""
    n = 20
    gates = circuits.random_layered(n, 40, seed=1)
    want = oracle.run(oracle.basis_state(n, 0), n, gates, threads=8)
    st = hip.State(n, hip.C64)
    st.set_basis(0)
    st.apply(ops.lower_all(gates))
    got = st.download()
    assert oracle.infidelity(got, want) < 1e-10
    assert np.abs(got - want).max() < 2e-6
    st.close()


def test_brickwork_28q_against_complex128():
    """At 28 qubits (the measured size), four brickwork layers in complex64
    (the tuned one-qubit and phase passes) against the same layers in
    complex128, which runs on the generic pair kernel: infidelity <= 1e-9.
    The layers must also move the state far from its input, so a pass that
    did nothing would fail."""
    n = 28
    gates = circuits.brickwork(n, 4, seed=9)
    ref = hip.State(n, hip.C128)
    ref.gaussian(5)
    st = hip.State(n, hip.C64)
    st.copy_from(ref)
    before = hip.State(n, hip.C128)
    before.copy_from(ref)
    lowered = ops.lower_all(gates)
    st.apply(lowered)
    ref.apply(lowered)
    ov = st.inner(ref)
    infid = 1.0 - abs(ov) ** 2 / (st.norm2() * ref.norm2())
    assert infid < 1e-9, infid
    moved = abs(before.inner(ref)) ** 2
    assert moved < 0.01, moved
    perm = [circuits.gate("CX", 0, n - 1), circuits.gate("Swap", 3, n - 2), circuits.gate("X", n - 1),
            circuits.gate("CCX", 1, 2, 20)]
    snap = hip.State(n, hip.C64)
    snap.copy_from(st)
    st.apply(ops.lower_all(perm))
    assert abs(st.inner(snap)) ** 2 < 0.5  # the permutation moved amplitudes
    st.apply(ops.lower_all(list(reversed(perm))))
    assert np.array_equal(st.download().view(np.uint32), snap.download().view(np.uint32))
    for s_ in (st, ref, before, snap):
        s_.close()
