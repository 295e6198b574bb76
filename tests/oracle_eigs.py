"""Lowest eigenvalues of the oracle's H (test helper): dense eigvalsh of the
matrixs-style dense H for small bases, ARPACK (scipy eigsh) driven by the
oracle's H*v for large ones, both restricted to the states the solver spans
(the parity mask, or the N-hbar-omega window)."""
import numpy as np

from oracle.refcpu import RefCPU


def oracle_lowest(args, k=5, dense_max=2600, seed=1, dopfile=None):
    from scipy.sparse.linalg import LinearOperator, eigsh
    ref = RefCPU(*args, fp64=True, dopfile=dopfile)
    _, keep = ref.trunc()
    idx = np.flatnonzero(keep)
    if idx.size <= dense_max:
        H = ref.dense()[np.ix_(idx, idx)]
        ev = np.linalg.eigvalsh(0.5 * (H + H.T))[:k]
    else:
        # shifted by -sigma on the kept states, so that the masked directions
        # (eigenvalue 0 of the operator) lie above every wanted eigenvalue
        mask = keep.astype(float)
        sigma = 1000.0
        op = LinearOperator((ref.N, ref.N),
                            matvec=lambda x: mask * (ref.apply(mask * np.ravel(x)) - sigma * np.ravel(x)),
                            dtype=np.float64)
        v0 = mask * np.random.default_rng(seed).standard_normal(ref.N)
        ev = np.sort(eigsh(op, k=k, which="SA", tol=1e-13, v0=v0)[0]) + sigma
    ref.close()
    return ev
