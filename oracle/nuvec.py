"""ORACLE (test infrastructure only): NuVec's reading of the Lanczos files and
its analysis of the states.

Restates src/NuVec/NuVec.f90:113-224.  The .lnz variant: read <prefix>.abz
(isave, alpha, beta / imax, no_level, tt, no_state, basd / bas), diagonalise
the tridiagonal matrix with diagonal alpha and subdiagonal beta(2:isave)
(Numerical Recipes tqli convention, src/NuCommon/ICommon.FI:188-251, where
e(i) couples i-1 and i and e(1) is ignored), sort the eigenvalues (indexx)
and form each state as sum_ix eigv(ix, i) * record ix of <prefix>.lnz
(direct access, no_state REAL*4 per record).  The .trl variant (:124-130,
:174-190): the .abz first record also holds matrix(isave, isave) and the second
one count; count = 1 shifts betat up by one (betat(i) couples i and i+1), and
count > 1 diagonalises the matrix (tred2 + tqli).  numpy's eigh stands in for
tqli/tred2.  The analysis (:136-152, :224-276): each state's squared amplitudes
summed per (a partition, b partition, 2Ja, 2Jb) in percent, the average shell
occupations of both sides, and the share of positive and negative parity
partition pairs, from the .nus shells and the .tri partitions read here.
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


def read_abz(prefix: str):
    """Both .abz layouts: the .lnz one (three records) or the .trl one (four)."""
    recs = read_records(prefix + ".abz")
    import os
    trl = not os.path.exists(prefix + ".lnz") and os.path.exists(prefix + ".trl")
    out = read_abz_lnz(prefix)
    out["trl"] = trl
    if trl:
        n = out["isave"]
        r1 = recs[0]
        out["matrix"] = np.frombuffer(r1[4 + 8 * n:4 + 8 * n + 4 * n * n], dtype="<f4").astype(np.float64) \
            .reshape(n, n, order="F")
        out["count"] = int(np.frombuffer(recs[1][20:24], dtype="<i4")[0])
        out["rec4"] = recs[3]
    return out


def nuvec_states_any(prefix: str):
    """(energies, states) from either layout."""
    abz = read_abz(prefix)
    if not abz["trl"]:
        return nuvec_states(prefix)
    n, N = abz["isave"], abz["no_state"]
    if abz["count"] == 1:
        T = np.diag(abz["alpha"])
        for i in range(n - 1):
            T[i, i + 1] = T[i + 1, i] = abz["beta"][i]
    else:
        T = abz["matrix"]
    w, Z = np.linalg.eigh(0.5 * (T + T.T))
    recs = np.fromfile(prefix + ".trl", dtype="<f4")
    assert recs.size >= n * N, "vector file shorter than isave records"
    recs = recs[:n * N].reshape(n, N).astype(np.float64)
    nlev = min(abz["no_level"], n, N)
    return w[:nlev], Z[:, :nlev].T @ recs


def read_nus_shells(nus_prefix: str):
    """Shell parities of the .nus (lines: name, three descriptions, particles,
    parity, shell count, shell types, then one line per shell ending in its parity)."""
    lines = [ln for ln in open(nus_prefix + ".nus") if not ln.startswith("!")]
    ns = int(lines[6].split()[0])
    return [int(lines[8 + k].split()[5]) for k in range(ns)]


def read_tri(opfile: str):
    """.tri header: partition count, 2J range, proton flag; nJT(p, J); partition occupations."""
    recs = read_records(opfile + ".tri")
    h = np.frombuffer(recs[0][:20], dtype="<i4")
    nsp, minJ2, maxJ2, proton = int(h[0]), int(h[1]), int(h[2]), bool(h[3])
    nJ = maxJ2 // 2 - minJ2 // 2 + 1
    nJT = np.frombuffer(recs[1][:4 * nsp * nJ], dtype="<i4").reshape(nJ, nsp).T  # [p, J - minJ2/2]
    ptdim = len(recs[2]) // (2 * nsp)
    occ = np.frombuffer(recs[2], dtype="<i2").reshape(nsp, ptdim)
    return {"nsp": nsp, "minJ2": minJ2, "maxJ2": maxJ2, "proton": proton, "nJT": nJT, "occ": occ}


def analyse(states, nus_prefix, opfile_a, opfile_b, bas):
    """NuVec's per-state analysis (:224-262): percentages per (a partition,
    b partition, Ja, Jb) (spartpc, Fortran order), average occupations of every
    shell for the a and b partitions (unnormalised, as .xva holds them), the
    total squared amplitude, and the positive / negative parity amplitudes."""
    par = read_nus_shells(nus_prefix)
    ns = len(par)
    A, B = read_tri(opfile_a), read_tri(opfile_b)

    def cparity(occ):
        p = 1
        for k in range(ns):
            if occ[k] > 0:
                p *= par[k] ** int(occ[k])
        return p
    pa = [cparity(A["occ"][p]) for p in range(A["nsp"])]
    pb = [cparity(B["occ"][p]) for p in range(B["nsp"])]
    # per basis state: (anpl, bnpl, ja2, jb2), the order of :136-152
    keys = []
    for key in bas:
        ja2, jb2 = key[0], key[1]
        for bp in range(B["nsp"]):
            for _ in range(B["nJT"][bp, jb2 // 2 - B["minJ2"] // 2]):
                for ap in range(A["nsp"]):
                    keys += [(ap, bp, ja2, jb2)] * int(A["nJT"][ap, ja2 // 2 - A["minJ2"] // 2])
    keys = np.array(keys)
    nJa = A["maxJ2"] // 2 - A["minJ2"] // 2 + 1
    nJb = B["maxJ2"] // 2 - B["minJ2"] // 2 + 1
    out = []
    for v in np.atleast_2d(states):
        a2 = v.astype(np.float64) ** 2
        t = a2.sum()
        pc = np.zeros((A["nsp"], B["nsp"], nJa, nJb), order="F")
        np.add.at(pc, (keys[:, 0], keys[:, 1], keys[:, 2] // 2 - A["minJ2"] // 2, keys[:, 3] // 2 - B["minJ2"] // 2), a2)
        ava = (A["occ"][keys[:, 0], :ns].astype(np.float64) * a2[:, None]).sum(axis=0)
        avb = (B["occ"][keys[:, 1], :ns].astype(np.float64) * a2[:, None]).sum(axis=0)
        sign = np.array([pa[a] * pb[b] for a, b in zip(keys[:, 0], keys[:, 1])])
        out.append({"tampl": t, "spartpc": pc * 100.0 / t, "ava": ava, "avb": avb,
                    "paramp": a2[sign >= 0].sum(), "nparamp": a2[sign < 0].sum()})
    return out


def read_xva(prefix: str, ns: int, shape):
    """.xva records: energy, ava(ns), avb(ns), spartpc(shape, Fortran order), REAL*4."""
    out = []
    for r in read_records(prefix + ".xva"):
        x = np.frombuffer(r, dtype="<f4").astype(np.float64)
        out.append({"energy": x[0], "ava": x[1:1 + ns], "avb": x[1 + ns:1 + 2 * ns],
                    "spartpc": x[1 + 2 * ns:].reshape(shape, order="F")})
    return out


def read_lev(prefix: str):
    """The last line of .lev: '(5(f10.3,a1))' energies and parity marks."""
    lines = [ln.rstrip("\n") for ln in open(prefix + ".lev")]
    tail = []
    for ln in reversed(lines):
        if not ln.strip():
            break
        tail.insert(0, ln)
    vals = []
    for ln in tail:
        for k in range(0, len(ln), 11):
            f = ln[k:k + 11]
            if len(f) == 11:
                vals.append((float(f[:10]), f[10]))
    return vals
