"""Host eigensolver of the Lanczos driver against numpy.linalg.eigh.

Matrices have the thick-restart shape (diagonal Ritz block, arrow column,
tridiagonal tail), plus dense, tridiagonal, degenerate and 1x1/2x2 cases.
Eigenvectors are checked through A z = w z and orthonormality (sign-free), and
the last-row mode against the last components of the full vectors."""
import numpy as np
import pytest

from emulator import sym_eig


def arrow_tridiag(rng, k, n):
    T = np.zeros((n, n))
    T[np.arange(k), np.arange(k)] = np.sort(rng.normal(size=k))
    T[:k, k] = T[k, :k] = rng.normal(size=k)
    for j in range(k, n):
        T[j, j] = rng.normal()
        if j + 1 < n:
            T[j, j + 1] = T[j + 1, j] = abs(rng.normal())
    return T


def cases():
    rng = np.random.default_rng(7)
    out = [np.array([[3.0]]), np.array([[1.0, 2.0], [2.0, -1.0]]), np.zeros((3, 3)), np.eye(4) * 2.5]
    for k, n in [(0, 10), (3, 12), (10, 50), (20, 50), (49, 50)]:
        out.append(arrow_tridiag(rng, k, n))
    G = rng.normal(size=(30, 30))
    out.append(G + G.T)
    Q, _ = np.linalg.qr(rng.normal(size=(20, 20)))
    out.append(Q @ np.diag(np.repeat([1.0, 2.0, 3.0, 4.0], 5)) @ Q.T)  # 5-fold degenerate
    # Wilkinson's W21+: tridiagonal, diagonal |10 - i|, unit couplings; its top
    # eigenvalues come in pairs that agree to ~1e-14
    W = np.diag(np.abs(np.arange(21) - 10.0)) + np.diag(np.ones(20), 1) + np.diag(np.ones(20), -1)
    out.append(W)
    # graded tridiagonal (diagonal spans 12 decades) and one with exact zero couplings (split)
    g = 10.0 ** -np.arange(13)
    out.append(np.diag(g) + np.diag(g[1:] * 0.3, 1) + np.diag(g[1:] * 0.3, -1))
    S = arrow_tridiag(rng, 0, 40)
    S[13, 14] = S[14, 13] = S[27, 28] = S[28, 27] = 0.0
    out.append(S)
    return out


@pytest.mark.parametrize("A", cases(), ids=lambda A: f"n{A.shape[0]}")
def test_sym_eig_matches_eigh(A):
    n = A.shape[0]
    w, Z = sym_eig(A)
    wr = np.linalg.eigvalsh(A)
    scale = max(1.0, np.abs(wr).max())
    assert np.abs(w - wr).max() < 1e-12 * scale * n
    assert np.abs(A @ Z - Z * w).max() < 1e-12 * scale * n
    assert np.abs(Z.T @ Z - np.eye(n)).max() < 1e-12 * n
    w2, last = sym_eig(A, full=False)
    assert np.abs(w2 - w).max() < 1e-12 * scale * n
    # last components agree with the full vectors up to sign (exact when eigenvalues are simple)
    simple = np.ones(n, bool)
    if n > 1:
        gap = np.diff(w)
        simple[:-1] &= gap > 1e-8
        simple[1:] &= gap > 1e-8
    assert np.abs(np.abs(last[simple]) - np.abs(Z[-1, simple])).max(initial=0) < 1e-10
