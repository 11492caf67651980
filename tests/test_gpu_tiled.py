"""Tile passes on the GPU (aqs_fuse_tiled + the tile kernel) against the FP64 oracle.

Tolerances as in the other parity tests: complex64 to 5e-6 absolute per
amplitude (and infidelity < 1e-9), complex128 to 1e-12; permutation
circuits bit-exact.
"""
import math

import numpy as np
import pytest

import oracle
from afquantumsim_amd import circuits, hip, ops
from tests.helpers import random_state
from tests.test_tiling import random_all_kinds

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    hip.init(0)
    yield


def tiled_run(n, sv, gates, precision=hip.C64, tile_qubits=0):
    st = hip.State(n, precision)
    st.upload(sv.astype(st.dtype))
    sch = hip.Schedule(ops.lower_all(gates, exact=(precision == hip.C128)), n, mode="tiled",
                       tile_qubits=tile_qubits)
    kinds = [k for k, _ in sch.passes()]
    st.apply_schedule(sch)
    out = st.download()
    st.close()
    return out, kinds


@pytest.mark.parametrize("tile", [10, 11, 12, 13])
@pytest.mark.parametrize("n", [13, 17])
def test_tiled_random_layered_gpu(n, tile):
    gates = circuits.random_layered(n, 10, seed=n + tile)
    sv = random_state(n, tile)
    want = oracle.run(sv, n, gates, threads=8)
    got, kinds = tiled_run(n, sv, gates, tile_qubits=tile)
    assert hip.OP_TILE in kinds and len(kinds) < len(gates) / 4
    assert oracle.infidelity(got, want) < 1e-9
    np.testing.assert_allclose(got, want, atol=5e-6)
    got128, _ = tiled_run(n, sv, gates, hip.C128, tile_qubits=tile)
    np.testing.assert_allclose(got128, want, atol=1e-12)


@pytest.mark.parametrize("n", [10, 14, 20])
def test_tiled_brickwork_gpu(n):
    gates = circuits.brickwork(n, 8, seed=5)
    sv = oracle.basis_state(n, 0)
    want = oracle.run(sv, n, gates, threads=8)
    got, kinds = tiled_run(n, sv, gates)
    assert oracle.infidelity(got, want) < 1e-9
    np.testing.assert_allclose(got, want, atol=5e-6)
    got128, _ = tiled_run(n, sv, gates, hip.C128)
    np.testing.assert_allclose(got128, want, atol=1e-12)


@pytest.mark.parametrize("seed", range(6))
def test_tiled_all_kinds_gpu(seed):
    n = 12 + seed % 4
    gates = random_all_kinds(n, 150, seed)
    sv = random_state(n, seed + 100)
    want = oracle.run(sv, n, gates)
    got, _ = tiled_run(n, sv, gates, tile_qubits=10 + seed % 4)
    np.testing.assert_allclose(got, want, atol=5e-6)
    got128, _ = tiled_run(n, sv, gates, hip.C128, tile_qubits=10 + seed % 4)
    np.testing.assert_allclose(got128, want, atol=1e-12)


def test_tiled_permutations_bit_exact_gpu():
    n = 18
    rng = np.random.default_rng(8)
    gates = []
    for _ in range(80):
        k = int(rng.integers(5))
        a, b, c = (int(x) for x in rng.choice(n, 3, replace=False))
        gates.append(oracle.gate(("X", "CX", "CCX", "Swap", "CSwap")[k], a, b, c))
    sv = random_state(n, 5).astype(np.complex64)
    want = oracle.run(sv.astype(np.complex128), n, gates).astype(np.complex64)
    for prec in (hip.C64, hip.C128):
        got, _ = tiled_run(n, sv.astype(np.complex128), gates, prec)
        assert np.array_equal(got.astype(np.complex64).view(np.uint32), want.view(np.uint32))


def test_tiled_qft_closed_form_gpu():
    """fourier_transform(n) + swaps on a Gaussian state = bitrev(ifft(bitrev(v))) sqrt(N)."""
    n = 18
    N = 1 << n
    st = hip.State(n, hip.C64)
    st.gaussian(71)
    v = oracle.gaussian_state(n, 71)
    sch = hip.Schedule(ops.lower_all(circuits.fourier_transform(n, final_swaps=True)), n, mode="tiled")
    st.apply_schedule(sch)
    rev = np.array([int(format(i, f"0{n}b")[::-1], 2) for i in range(N)])
    ref = np.fft.ifft(v[rev])[rev] * math.sqrt(N)
    assert oracle.infidelity(st.download(), ref) < 1e-9


def test_tiled_matches_plain_at_24_qubits():
    """Full-size tiles (13 local bits, 2^11 tiles) against the per-gate FP64 path."""
    n = 24
    gates = circuits.brickwork(n, 4, seed=6)
    ref = hip.State(n, hip.C128)
    ref.set_basis(0)
    ref.apply(ops.lower_all(gates, exact=True))
    want = ref.download()
    ref.close()
    st = hip.State(n, hip.C64)
    st.set_basis(0)
    sch = hip.Schedule(ops.lower_all(gates), n, mode="tiled")
    st.apply_schedule(sch)
    got = st.download()
    assert oracle.infidelity(got, want) < 1e-9
    g = got.astype(np.complex128)
    assert abs(np.vdot(g, g).real - 1.0) < 1e-5

