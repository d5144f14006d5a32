"""Reader/writer of the vortex solver's snapshot files (s_NNNNN_step.dat).

The solver writes them as Fortran sequential unformatted records (4-byte
length marker, payload, the same marker again) in f_data_out.f90:19-124 and
reads them back in b_continued_file.f90:15-108.  The blob block is a record
holding nvor_b (int32) followed by seven records per blob
(f_data_out.f90:97-107):

    (x, y)  (u, v)_1  (u, v)_2  (u, v)_3   cir   cor   blob_id

Header layouts differ between solver versions (the shipped s_00198 has one
header record more than f_data_out.f90:30-43 writes), so the blob block is
located by that record pattern rather than by position.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

_BLOB_PATTERN = (16, 16, 16, 16, 8, 8, 4)


@dataclass
class Snapshot:
    xy: np.ndarray        # (N, 2) blob positions
    vel: np.ndarray       # (N, 3, 2) velocity history vor_vb(:, 1..3, :)
    cir: np.ndarray       # (N,) circulation
    cor: np.ndarray       # (N,) core radius
    blob_id: np.ndarray   # (N,) int32
    records: list         # every record payload (bytes), for rewriting
    blob_record: int      # index of the nvor_b record


def read_records(path: str) -> list:
    data = open(path, "rb").read()
    out, off = [], 0
    while off < len(data):
        (n,) = struct.unpack_from("<i", data, off)
        payload = data[off + 4:off + 4 + n]
        (n2,) = struct.unpack_from("<i", data, off + 4 + n)
        if n2 != n or n < 0:
            raise ValueError(f"{path}: broken record marker at byte {off}")
        out.append(payload)
        off += 8 + n
    return out


def _find_blob_block(recs: list) -> int:
    for i, r in enumerate(recs):
        if len(r) != 4:
            continue
        (n,) = struct.unpack("<i", r)
        if n < 0 or i + 1 + 7 * n > len(recs):
            continue
        if n == 0:
            continue
        first = all(len(recs[i + 1 + k]) == _BLOB_PATTERN[k] for k in range(7))
        last = all(len(recs[i + 1 + 7 * (n - 1) + k]) == _BLOB_PATTERN[k] for k in range(7))
        if first and last:
            return i
    raise ValueError("no blob block found")


def read_snapshot(path: str) -> Snapshot:
    recs = read_records(path)
    i = _find_blob_block(recs)
    (n,) = struct.unpack("<i", recs[i])
    blk = recs[i + 1:i + 1 + 7 * n]
    xy = np.frombuffer(b"".join(blk[0::7]), dtype="<f8").reshape(n, 2).copy()
    v1 = np.frombuffer(b"".join(blk[1::7]), dtype="<f8").reshape(n, 2)
    v2 = np.frombuffer(b"".join(blk[2::7]), dtype="<f8").reshape(n, 2)
    v3 = np.frombuffer(b"".join(blk[3::7]), dtype="<f8").reshape(n, 2)
    cir = np.frombuffer(b"".join(blk[4::7]), dtype="<f8").copy()
    cor = np.frombuffer(b"".join(blk[5::7]), dtype="<f8").copy()
    bid = np.frombuffer(b"".join(blk[6::7]), dtype="<i4").copy()
    vel = np.stack([v1, v2, v3], axis=1).copy()
    return Snapshot(xy, vel, cir, cor, bid, recs, i)


def write_snapshot(path: str, snap: Snapshot) -> None:
    """Write the snapshot back, with the blob block rebuilt from the arrays
    (so new velocities in snap.vel[:, 0] land in vor_vb(:, 1, :))."""
    n = snap.xy.shape[0]
    i = snap.blob_record
    (n_old,) = struct.unpack("<i", snap.records[i])
    blob = [struct.pack("<i", n)]
    for j in range(n):
        blob.append(struct.pack("<2d", *snap.xy[j]))
        for k in range(3):
            blob.append(struct.pack("<2d", *snap.vel[j, k]))
        blob.append(struct.pack("<d", snap.cir[j]))
        blob.append(struct.pack("<d", snap.cor[j]))
        blob.append(struct.pack("<i", int(snap.blob_id[j])))
    recs = snap.records[:i] + blob + snap.records[i + 1 + 7 * n_old:]
    with open(path, "wb") as f:
        for r in recs:
            m = struct.pack("<i", len(r))
            f.write(m + r + m)
