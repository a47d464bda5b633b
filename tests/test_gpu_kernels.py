"""Kernel-level parity of the gfx950 library against the CPU oracle and numpy.

GEMM: vmul! (real.jl:7-102); UDT: udt_AVX_pivot! (UDT.jl:216-334); rdivp! (real.jl:198-226);
calculate_greens_AVX! (stack.jl:442-496).
"""
import numpy as np
import pytest

from conftest import rel_maxdiff

pytestmark = pytest.mark.gpu


def graded(n, rng, decades=12):
    """Q1 diag(10^k) Q2: the widely graded matrices a DQMC slice chain produces."""
    q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = 10.0 ** np.linspace(0, -decades, n)
    rng.shuffle(s)
    return np.asfortranarray(q1 @ np.diag(s) @ q2)


@pytest.mark.parametrize("n", [4, 16, 36, 64, 100, 256, 300, 576])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_fp64(hip, n, ta, tb):
    rng = np.random.default_rng(n + 10 * ta + 20 * tb)
    A = np.asfortranarray(rng.standard_normal((n, n)))
    B = np.asfortranarray(rng.standard_normal((n, n)))
    C = hip.la_gemm(A, B, ta, tb)
    ref = (A.T if ta else A) @ (B.T if tb else B)
    # FP64 accumulation: error bound ~ n eps |A||B|
    assert np.abs(C - ref).max() <= 4 * n * np.finfo(float).eps * (np.abs(A).max() * np.abs(B).max() * n)


@pytest.mark.parametrize("n", [16, 100, 256, 576])
def test_gemm_fp32(hip, n):
    rng = np.random.default_rng(n)
    A = np.asfortranarray(rng.standard_normal((n, n)))
    B = np.asfortranarray(rng.standard_normal((n, n)))
    C = hip.la_gemm(A, B, 0, 1, precision="mixed")
    ref = A.astype(np.float32).astype(np.float64) @ B.astype(np.float32).astype(np.float64).T
    # FP32 accumulation tolerance
    assert rel_maxdiff(C, ref) < 5e-6


@pytest.mark.parametrize("n", [1, 2, 16, 36, 64, 144, 256, 300, 576, 1100])
def test_udt_reconstruction(hip, n):
    rng = np.random.default_rng(7 * n)
    X = graded(n, rng, decades=8) if n > 2 else np.asfortranarray(rng.standard_normal((n, n)))
    U, D, T, piv = hip.la_udt(X, True)
    assert np.abs(U.T @ U - np.eye(n)).max() < 1e-13 * max(1, n / 16)
    recon = U @ np.diag(D) @ T
    assert np.abs(recon - X).max() < 1e-13 * np.abs(X).max() * max(1, n / 16)
    # T = D^-1 R P^T: upper triangular after undoing the permutation, unit diagonal modulus
    R = T[:, piv]
    assert np.abs(np.tril(R, -1)).max() == 0.0
    assert np.allclose(np.abs(np.diag(R)), 1.0, atol=1e-14)
    assert sorted(piv.tolist()) == list(range(n))


@pytest.mark.parametrize("n", [16, 64, 256, 576])
def test_udt_matches_oracle_on_graded_input(hip, oracle, n):
    """Widely graded columns fix the pivot order, so GPU and oracle factors coincide."""
    rng = np.random.default_rng(n)
    q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = 10.0 ** (-0.2 * rng.permutation(n))  # column scales 1.58x apart, shuffled
    X = np.asfortranarray(q1 @ (np.eye(n) + 0.01 * rng.standard_normal((n, n))) @ np.diag(s))
    U, D, T, piv = hip.la_udt(X, True)
    Uo, Do, To, pivo = oracle.udt(X, True)
    assert piv.tolist() == pivo.tolist()
    assert rel_maxdiff(D, Do) < 1e-12
    assert np.abs(U - Uo).max() < 1e-11
    assert np.abs(T - To).max() < 1e-11


@pytest.mark.parametrize("n", [16, 100, 256, 576])
def test_udt_dirty_upper(hip, n):
    rng = np.random.default_rng(3 * n)
    X = graded(n, rng, decades=6)
    U, D, T, piv = hip.la_udt(X, False)
    P = np.zeros((n, n))
    P[piv, np.arange(n)] = 1.0
    recon = U @ np.diag(D) @ np.triu(T) @ P.T
    assert np.abs(recon - X).max() < 1e-13 * np.abs(X).max() * max(1, n / 16)


@pytest.mark.parametrize("n,unit", [(1, True), (16, True), (36, True), (256, True), (300, True),
                                    (576, True), (16, False), (256, False)])
def test_rdivp(hip, oracle, n, unit):
    rng = np.random.default_rng(n)
    A = np.asfortranarray(rng.standard_normal((n, n)))
    # a UDT's T has a unit-modulus diagonal; other diagonals exercise the plain division
    diag = rng.choice([-1.0, 1.0], n) * (1.0 if unit else rng.choice([0.3, 0.7, 1.5, 3.0], n))
    T = np.asfortranarray(np.triu(rng.uniform(-0.5, 0.5, (n, n)), 1) + np.diag(diag)
                          + np.tril(rng.standard_normal((n, n)), -1))  # dirty lower part
    piv = rng.permutation(n).astype(np.int32)
    out = hip.la_rdivp(A, T, piv)
    ref = oracle.rdivp(A, T, piv)
    dense = A[:, piv] @ np.linalg.inv(np.triu(T))
    assert rel_maxdiff(out, ref) < 1e-12
    assert rel_maxdiff(out, dense) < 1e-10


@pytest.mark.parametrize("n", [16, 64, 256])
def test_calc_greens_udt(hip, oracle, n):
    rng = np.random.default_rng(11 * n)
    Ul, Dl, Tl, _ = oracle.udt(graded(n, rng, 10), True)
    Ur, Dr, Tr, _ = oracle.udt(graded(n, rng, 10), True)
    G = hip.la_calc_greens_udt(Ul, Dl, Tl, Ur, Dr, Tr)
    Go = oracle.calc_greens_udt(Ul, Dl, Tl, Ur, Dr, Tr)
    assert rel_maxdiff(G, Go) < 1e-10
