"""Fortran unformatted sequential-record I/O (4-byte little-endian length markers).

NuShellX writes every operator file (``.tri .trd .oph .mat .sgt``) as unformatted
sequential records: a 4-byte byte count, the payload, the same count again
(Intel Fortran and flang agree on this layout).  The solver frequently reads
only a prefix of a record (e.g. ``src/NuTRL/NuCoreX.f90:1146`` reads 3 of the 7
header integers of a ``.trd``), so readers here return the whole payload and
let the caller decode what it needs.
"""
from __future__ import annotations

import struct
from typing import BinaryIO, Iterator, List

import numpy as np


class RecordWriter:
    def __init__(self, path: str):
        self._f: BinaryIO = open(path, "wb")

    def write(self, *parts) -> None:
        """Write one record made of the concatenation of ``parts``.

        Each part is bytes, a numpy array (written in its own dtype, Fortran
        column-major order for 2-D arrays) or a python int (written as int32).
        """
        chunks: List[bytes] = []
        for p in parts:
            if isinstance(p, (bytes, bytearray)):
                chunks.append(bytes(p))
            elif isinstance(p, np.ndarray):
                chunks.append(np.asfortranarray(p).tobytes(order="F"))
            elif isinstance(p, (bool, np.bool_)):
                chunks.append(struct.pack("<i", -1 if p else 0))
            elif isinstance(p, (int, np.integer)):
                chunks.append(struct.pack("<i", int(p)))
            elif isinstance(p, np.float32):
                chunks.append(struct.pack("<f", float(p)))
            elif isinstance(p, (float, np.float64)):
                chunks.append(struct.pack("<d", float(p)))
            else:
                raise TypeError(f"cannot write {type(p)}")
        payload = b"".join(chunks)
        n = len(payload)
        self._f.write(struct.pack("<i", n))
        self._f.write(payload)
        self._f.write(struct.pack("<i", n))

    def close(self) -> None:
        self._f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_records(path: str) -> Iterator[bytes]:
    """Yield the payload of each record of a sequential unformatted file."""
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        pos += 4
        payload = data[pos:pos + n]
        pos += n
        (n2,) = struct.unpack_from("<i", data, pos)
        if n2 != n:
            raise ValueError(f"{path}: record marker mismatch {n} != {n2}")
        pos += 4
        yield payload
