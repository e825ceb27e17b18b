"""Host dense symmetric eigensolver used by the Krylov solvers
(ls_amd_dense_eigh: Householder tridiagonalisation + implicit QR), against
numpy.linalg.eigh and the defining identities."""
import ctypes as C

import numpy as np
import pytest

from lsamd import _abi


def eigh(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    w = np.zeros(max(n, 1))
    Z = np.zeros((max(n, 1), max(n, 1)))
    p = C.POINTER(C.c_double)
    rc = _abi.load().ls_amd_dense_eigh(n, A.ctypes.data_as(p), w.ctypes.data_as(p), Z.ctypes.data_as(p))
    assert rc == 0, _abi.last_error()
    return w[:n], Z[:n, :n]


def check(A, tol=1e-12):
    w, Z = eigh(A)
    n = A.shape[0]
    scale = max(1.0, np.abs(A).max() * n)
    assert np.all(np.diff(w) >= 0)
    assert np.abs(w - np.linalg.eigvalsh(A)).max() <= tol * scale
    assert np.abs(A @ Z - Z * w).max() <= 1e-11 * scale
    assert np.abs(Z.T @ Z - np.eye(n)).max() <= 1e-12 * n


@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 64, 160])
def test_random_symmetric(n):
    rng = np.random.default_rng(n)
    M = rng.standard_normal((n, n))
    check(M + M.T)


def test_structured_matrices():
    rng = np.random.default_rng(5)
    # real embedding of a complex Hermitian matrix: every eigenvalue twice
    A = rng.standard_normal((30, 30))
    B = rng.standard_normal((30, 30))
    Hr, Hi = A + A.T, B - B.T
    R = np.block([[Hr, -Hi], [Hi, Hr]])
    w, _ = eigh(R)
    assert np.allclose(w[0::2], w[1::2], atol=1e-10)
    check(R)
    # already diagonal / tridiagonal / block diagonal with exact zeros
    check(np.diag(rng.standard_normal(40)))
    T = np.diag(rng.standard_normal(40)) + np.diag(rng.standard_normal(39), 1)
    check(T + np.triu(T, 1).T)
    C2 = np.zeros((50, 50))
    C2[:20, :20] = (lambda m: m + m.T)(rng.standard_normal((20, 20)))
    C2[20:, 20:] = (lambda m: m + m.T)(rng.standard_normal((30, 30)))
    check(C2)
    # highly degenerate: rank-one update of the identity, and the zero matrix
    v = rng.standard_normal(25)
    check(np.eye(25) * 3.0 + np.outer(v, v))
    check(np.zeros((12, 12)))
    # block-tridiagonal projected matrix of a block Lanczos run (widely spread spectrum)
    blocks = 12
    b = 8
    T = np.zeros((blocks * b, blocks * b))
    for j in range(blocks):
        Aj = rng.standard_normal((b, b))
        T[j * b:(j + 1) * b, j * b:(j + 1) * b] = (Aj + Aj.T) * 10 - 40 * np.eye(b)
        if j + 1 < blocks:
            Bj = np.triu(rng.standard_normal((b, b)))
            T[(j + 1) * b:(j + 2) * b, j * b:(j + 1) * b] = Bj
            T[j * b:(j + 1) * b, (j + 1) * b:(j + 2) * b] = Bj.T
    check(T)


def eigh_hermitian(H):
    H = np.ascontiguousarray(H, dtype=np.complex128)
    n = H.shape[0]
    w = np.zeros(max(n, 1))
    Z = np.zeros((max(n, 1), max(n, 1)), dtype=np.complex128)
    p = C.POINTER(C.c_double)
    rc = _abi.load().ls_amd_dense_eigh_hermitian(n, H.ctypes.data_as(p), w.ctypes.data_as(p), Z.ctypes.data_as(p))
    assert rc == 0, _abi.last_error()
    return w[:n], Z[:n, :n]


def check_hermitian(H, tol=1e-12):
    w, Z = eigh_hermitian(H)
    n = H.shape[0]
    scale = max(1.0, np.abs(H).max() * n)
    assert np.all(np.diff(w) >= 0)
    assert np.abs(w - np.linalg.eigvalsh(H)).max() <= tol * scale
    assert np.abs(H @ Z - Z * w).max() <= 1e-11 * scale
    assert np.abs(Z.conj().T @ Z - np.eye(n)).max() <= 1e-12 * n


@pytest.mark.parametrize("n", [1, 2, 3, 8, 33, 96])
def test_random_hermitian(n):
    """ls_amd_dense_eigh_hermitian (complex Householder + implicit QR) against numpy."""
    rng = np.random.default_rng(100 + n)
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    check_hermitian(M + M.conj().T)


def test_structured_hermitian():
    rng = np.random.default_rng(9)
    # degenerate: a rank-two update of a multiple of the identity
    u = rng.standard_normal((20, 2)) + 1j * rng.standard_normal((20, 2))
    check_hermitian(3.0 * np.eye(20) + u @ u.conj().T)
    check_hermitian(np.zeros((6, 6)))
    # a real symmetric input gives the real solver's eigenvalues
    A = rng.standard_normal((25, 25))
    A = A + A.T
    w, _ = eigh_hermitian(A.astype(complex))
    assert np.abs(w - eigh(A)[0]).max() <= 1e-12 * 25 * np.abs(A).max()
    # block diagonal with an exactly decoupled complex block
    H = np.zeros((10, 10), dtype=complex)
    H[:4, :4] = (lambda m: m + m.conj().T)(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))
    H[4:, 4:] = np.diag(np.arange(6.0))
    check_hermitian(H)
