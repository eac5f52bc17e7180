"""Writes tests/golden/ttdb_reference.json from the reference's own test data
file tests/_data/particles.ttdb (a 1,660-particle, 4-frame run): its table
definitions, frame times, per-array names, type ids and element counts, the
first and last bytes of the blobs (frame headers), and for the first frame's
r and rho the decoded values' sums and first entries.  The blobs are
decompressed with the system's libzstd (libzstd.so.1).

Run from the repository root:  python tests/golden/make_ttdb_fixture.py
"""
import ctypes as C
import hashlib
import json
import sqlite3
from pathlib import Path

import numpy as np

SRC = "/root/reference/tests/_data/particles.ttdb"
OUT = Path(__file__).resolve().parent / "ttdb_reference.json"


def libzstd_decode(blob: bytes) -> bytes:
    z = C.CDLL("libzstd.so.1")
    z.ZSTD_getFrameContentSize.restype = C.c_ulonglong
    z.ZSTD_getFrameContentSize.argtypes = [C.c_void_p, C.c_size_t]
    z.ZSTD_decompress.restype = C.c_size_t
    z.ZSTD_decompress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]
    n = z.ZSTD_getFrameContentSize(blob, len(blob))
    buf = C.create_string_buffer(max(int(n), 1))
    got = z.ZSTD_decompress(buf, n, blob, len(blob))
    assert got == n, (got, n)
    return buf.raw[:n]


def main():
    con = sqlite3.connect(f"file:{SRC}?immutable=1", uri=True)
    tables = {}
    for name, in con.execute("SELECT name FROM sqlite_master WHERE type = 'table' "
                             "AND name != 'sqlite_sequence'"):
        cols = con.execute(f"PRAGMA table_info({name})").fetchall()
        tables[name] = [[c[1], c[2], c[3], c[5]] for c in cols]  # name, type, notnull, pk
    frames = [list(r) for r in con.execute("SELECT id, series_id, time FROM DataFrames")]
    arrays = []
    for aid, fid, name, tid, size, blob in con.execute(
            "SELECT id, frame_id, name, type, size, data FROM DataArrays ORDER BY id"):
        arrays.append({"frame": fid, "name": name, "type": tid, "size": size,
                       "head_hex": blob[:7].hex(), "tail_hex": blob[-9:].hex()})
    decoded = {}
    for name in ("r", "rho"):
        tid, size, blob = con.execute(
            "SELECT type, size, data FROM DataArrays WHERE frame_id = 1 AND name = ?",
            (name,)).fetchone()
        raw = libzstd_decode(blob)
        x = np.frombuffer(raw, dtype=np.float64)
        decoded[name] = {"bytes": len(raw), "sha256": hashlib.sha256(raw).hexdigest(),
                         "sum": float(x.sum()), "first": x[:6].tolist()}
    OUT.write_text(json.dumps({"source": "tests/_data/particles.ttdb", "tables": tables,
                               "settings": con.execute("SELECT * FROM Settings").fetchall(),
                               "frames": frames, "arrays": arrays, "decoded": decoded},
                              indent=1) + "\n")


if __name__ == "__main__":
    main()
