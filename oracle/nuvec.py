"""ORACLE (test infrastructure only): NuVec's reading of the Lanczos files.

Restates src/NuVec/NuVec.f90:113-224 for the .lnz variant: read <prefix>.abz
(isave, alpha, beta / imax, no_level, tt, no_state, basd / bas), diagonalise
the tridiagonal matrix with diagonal alpha and subdiagonal beta(2:isave)
(Numerical Recipes tqli convention, src/NuCommon/ICommon.FI:188-251, where
e(i) couples i-1 and i and e(1) is ignored), sort the eigenvalues (indexx)
and form each state as sum_ix eigv(ix, i) * record ix of <prefix>.lnz
(direct access, no_state REAL*4 per record).  numpy's eigh stands in for tqli.
"""
from __future__ import annotations

import struct

import numpy as np


def read_records(path: str):
    """The payloads of a Fortran unformatted sequential file: each record is a
    4-byte little-endian byte count, the bytes, and the count again.  Written
    here, apart from the product's readers, so that a reader bug cannot hide
    in both."""
    with open(path, "rb") as f:
        data = f.read()
    out, pos = [], 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        if n < 0 or pos + 8 + n > len(data):
            raise ValueError(f"{path}: bad record length {n} at byte {pos}")
        (n2,) = struct.unpack_from("<i", data, pos + 4 + n)
        if n2 != n:
            raise ValueError(f"{path}: record markers differ at byte {pos}")
        out.append(data[pos + 4:pos + 4 + n])
        pos += 8 + n
    return out


def read_abz_lnz(prefix: str):
    recs = list(read_records(prefix + ".abz"))
    r1 = recs[0]
    isave = int(np.frombuffer(r1[:4], dtype="<i4")[0])
    alpha = np.frombuffer(r1[4:4 + 4 * isave], dtype="<f4").astype(np.float64)
    beta = np.frombuffer(r1[4 + 4 * isave:4 + 8 * isave], dtype="<f4").astype(np.float64)
    r2 = recs[1]
    imax, no_level = np.frombuffer(r2[:8], dtype="<i4")
    tt = float(np.frombuffer(r2[8:12], dtype="<f4")[0])
    no_state, basd = np.frombuffer(r2[12:20], dtype="<i4")
    bas = [bytes(recs[2][8 * b:8 * b + 8]) for b in range(int(basd))]
    return {"isave": isave, "alpha": alpha, "beta": beta, "imax": int(imax), "no_level": int(no_level),
            "tt": tt, "no_state": int(no_state), "basd": int(basd), "bas": bas}


def nuvec_states(prefix: str):
    """(energies, states) as NuVec forms them: min(no_level, isave) states."""
    abz = read_abz_lnz(prefix)
    n, N = abz["isave"], abz["no_state"]
    T = np.diag(abz["alpha"])
    for i in range(1, n):
        T[i - 1, i] = T[i, i - 1] = abz["beta"][i]
    w, Z = np.linalg.eigh(T)
    recs = np.fromfile(prefix + ".lnz", dtype="<f4")
    assert recs.size >= n * N, "vector file shorter than isave records"
    recs = recs[:n * N].reshape(n, N).astype(np.float64)
    nlev = min(abz["no_level"], n)
    return w[:nlev], Z[:, :nlev].T @ recs
