"""TitSolver's particle output: the `.ttdb` storage (data::Storage).

A `.ttdb` file is an SQLite database with the tables of
source/tit/data/storage.cpp:37-60: `Settings` (the series limit),
`DataSeries`, `DataFrames` (one per output time) and `DataArrays` (one per
particle field and frame: name, type id, element count and a zstd-framed
blob of the field's elements).  `ParticleArray::write`
(sph/particle_array.hpp:165-172) creates one frame per call and one array per
varying field; titwcsph writes a frame every 100 steps and keeps one series
(wcsph.cpp:158-164, 183-189).

Type ids follow data/type.hpp:185-189: (kind + 1) | rank << 8 | dim << 16,
kinds int8, uint8, int16, uint16, int32, uint32, int64, uint64, float32,
float64 in that order, ranks scalar, vector, matrix.  A vector field stores
dim values per particle, a matrix field dim x dim, row by row
(core/serialization.hpp:101-129).

Blobs are zstd frames (RFC 8878).  The reference streams each field through
libzstd (core/zstd.cpp) and leaves a second, empty frame behind the data
frame when the stream is flushed at its end.  This writer frames the bytes
itself, with no compression library: one single-segment frame whose blocks
are raw (stored) blocks of at most 128 KiB, or RLE blocks where a block's
bytes are all equal (fields that are still zero), followed by the same empty
frame.  Any zstd decoder reads them.

`FrameWriter` moves the writing off the stepping thread: `submit` hands a
downloaded frame to a worker thread, and the next steps run on the device
while it is written.
"""
from __future__ import annotations

import queue
import sqlite3
import struct
import threading
from pathlib import Path

import numpy as np

# storage.cpp:37-60
SCHEMA = """
PRAGMA journal_mode = WAL;
PRAGMA foreign_keys = ON;

CREATE TABLE IF NOT EXISTS Settings (
  id INTEGER PRIMARY KEY CHECK (id = 0),
  max_series INTEGER
) STRICT;
INSERT OR IGNORE INTO Settings (id, max_series) VALUES (0, 5);

CREATE TABLE IF NOT EXISTS DataSeries (
  id   INTEGER PRIMARY KEY AUTOINCREMENT,
  name TEXT NOT NULL
) STRICT;

CREATE TABLE IF NOT EXISTS DataFrames (
  id        INTEGER PRIMARY KEY AUTOINCREMENT,
  series_id INTEGER NOT NULL,
  time      REAL NOT NULL,
  FOREIGN KEY (series_id) REFERENCES DataSeries(id) ON DELETE CASCADE
) STRICT;

CREATE TABLE IF NOT EXISTS DataArrays (
  id       INTEGER PRIMARY KEY AUTOINCREMENT,
  frame_id INTEGER NOT NULL,
  name     TEXT NOT NULL,
  type     INTEGER,
  size     INTEGER,
  data     BLOB,
  FOREIGN KEY (frame_id) REFERENCES DataFrames(id) ON DELETE CASCADE
) STRICT;
"""

# ---------------------------------------------------------------------------
# Data types (data/type.hpp)

KINDS = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64,
         np.uint64, np.float32, np.float64]
RANK_SCALAR, RANK_VECTOR, RANK_MATRIX = 0, 1, 2


def type_id(dtype, rank: int, dim: int) -> int:
    """Integer id of a data type: (kind + 1) | rank << 8 | dim << 16."""
    kind = [np.dtype(k) for k in KINDS].index(np.dtype(dtype))
    if rank not in (RANK_SCALAR, RANK_VECTOR, RANK_MATRIX) or dim < 1:
        raise ValueError("invalid data type rank or dimensionality")
    if rank == RANK_SCALAR and dim != 1:
        raise ValueError("a scalar has dimensionality 1")
    return (kind + 1) | rank << 8 | dim << 16


def parse_type(tid: int):
    """(numpy dtype, rank, dim) of a type id."""
    kind = (tid - 1) & 0xFF
    rank = (tid >> 8) & 0xFF
    dim = (tid >> 16) & 0xFF
    if kind >= len(KINDS) or rank > RANK_MATRIX or dim < 1:
        raise ValueError(f"invalid data type id {tid}")
    return np.dtype(KINDS[kind]), rank, dim


def type_of(arr: np.ndarray) -> int:
    """Type id of a particle field held as (n,), (n, d) or (n, d, d)."""
    if arr.ndim == 1:
        return type_id(arr.dtype, RANK_SCALAR, 1)
    if arr.ndim == 2:
        return type_id(arr.dtype, RANK_VECTOR, arr.shape[1])
    if arr.ndim == 3 and arr.shape[1] == arr.shape[2]:
        return type_id(arr.dtype, RANK_MATRIX, arr.shape[1])
    raise ValueError(f"no data type for an array of shape {arr.shape}")


# ---------------------------------------------------------------------------
# zstd framing (RFC 8878), stored and RLE blocks only

ZSTD_MAGIC = 0xFD2FB528
BLOCK_MAX = 128 * 1024
BLOCK_RAW, BLOCK_RLE, BLOCK_COMPRESSED = 0, 1, 2


def _frame_header(size: int) -> bytes:
    """Magic and a single-segment frame header whose content-size field is
    the smallest that holds `size` (RFC 8878 3.1.1.1.1-3.1.1.1.4)."""
    if size < 256:
        return struct.pack("<IBB", ZSTD_MAGIC, 0x20, size)
    if size < 65536 + 256:
        return struct.pack("<IBH", ZSTD_MAGIC, 0x60, size - 256)
    if size < 1 << 32:
        return struct.pack("<IBI", ZSTD_MAGIC, 0xA0, size)
    return struct.pack("<IBQ", ZSTD_MAGIC, 0xE0, size)


def _block_header(last: bool, btype: int, size: int) -> bytes:
    v = (1 if last else 0) | btype << 1 | size << 3
    return bytes((v & 0xFF, (v >> 8) & 0xFF, (v >> 16) & 0xFF))


def zstd_frame(data) -> bytes:
    """One zstd frame holding `data` in stored blocks (RLE where a block's
    bytes are all equal), then an empty frame, as the reference's stream
    compressor leaves its blobs."""
    mv = memoryview(np.ascontiguousarray(data)).cast("B")
    n = len(mv)
    out = [_frame_header(n)]
    if n == 0:
        out.append(_block_header(True, BLOCK_RAW, 0))
    for start in range(0, n, BLOCK_MAX):
        chunk = mv[start:start + BLOCK_MAX]
        last = start + BLOCK_MAX >= n
        arr = np.frombuffer(chunk, dtype=np.uint8)
        if len(arr) > 1 and arr[0] == arr[-1] and not np.any(arr != arr[0]):
            out += [_block_header(last, BLOCK_RLE, len(arr)), bytes((int(arr[0]),))]
        else:
            out += [_block_header(last, BLOCK_RAW, len(arr)), chunk.tobytes()]
    out.append(_frame_header(0) + _block_header(True, BLOCK_RAW, 0))
    return b"".join(out)


def zstd_decode(blob: bytes) -> bytes:
    """Contents of every frame in `blob` (zstd frames with stored and RLE
    blocks, and skippable frames).  A compressed block raises ValueError:
    that needs a full zstd decoder."""
    out = bytearray()
    pos = 0
    blob = bytes(blob)
    while pos < len(blob):
        magic, = struct.unpack_from("<I", blob, pos)
        pos += 4
        if 0x184D2A50 <= magic <= 0x184D2A5F:  # skippable frame
            size, = struct.unpack_from("<I", blob, pos)
            pos += 4 + size
            continue
        if magic != ZSTD_MAGIC:
            raise ValueError("not a zstd frame")
        fhd = blob[pos]
        pos += 1
        fcs_flag, single, checksum, did_flag = fhd >> 6, (fhd >> 5) & 1, (fhd >> 2) & 1, fhd & 3
        if not single:
            pos += 1  # window descriptor
        pos += (0, 1, 2, 4)[did_flag]
        fcs_len = (1 if single else 0, 2, 4, 8)[fcs_flag]
        fcs = None
        if fcs_len:
            fcs = int.from_bytes(blob[pos:pos + fcs_len], "little")
            if fcs_len == 2:
                fcs += 256
        pos += fcs_len
        start = len(out)
        while True:
            v = int.from_bytes(blob[pos:pos + 3], "little")
            pos += 3
            last, btype, size = v & 1, (v >> 1) & 3, v >> 3
            if btype == BLOCK_RAW:
                out += blob[pos:pos + size]
                pos += size
            elif btype == BLOCK_RLE:
                out += bytes((blob[pos],)) * size
                pos += 1
            else:
                raise ValueError("compressed zstd block: needs a zstd library")
            if last:
                break
        if checksum:
            pos += 4
        if fcs is not None and len(out) - start != fcs:
            raise ValueError("zstd frame content size mismatch")
    return bytes(out)


# ---------------------------------------------------------------------------
# Storage (data/storage.hpp, storage.cpp)


class Storage:
    """A `.ttdb` file: series of frames of named arrays (data::Storage)."""

    def __init__(self, path, read_only: bool = False):
        self.path = Path(path)
        if read_only:
            self.db = sqlite3.connect(f"file:{self.path}?mode=ro", uri=True,
                                      check_same_thread=False)
        else:
            self.db = sqlite3.connect(str(self.path), check_same_thread=False)
            self.db.executescript(SCHEMA)
            self.db.commit()

    def close(self):
        if self.db is not None:
            self.db.commit()
            self.db.close()
            self.db = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # settings
    @property
    def max_series(self) -> int:
        return self.db.execute("SELECT max_series FROM Settings").fetchone()[0]

    def set_max_series(self, value: int):
        """Storage::set_max_series (storage.cpp:81-95): drops the oldest
        series beyond the limit."""
        if value < 1:
            raise ValueError("maximum number of series must be positive")
        self.db.execute("UPDATE Settings SET max_series = ?", (value,))
        extra = self.num_series() - value
        if extra > 0:
            self.db.execute("DELETE FROM DataSeries WHERE id IN (SELECT id FROM DataSeries "
                            "ORDER BY id ASC LIMIT ?)", (extra,))
        self.db.commit()

    # series
    def num_series(self) -> int:
        return self.db.execute("SELECT COUNT(*) FROM DataSeries").fetchone()[0]

    def series_ids(self) -> list[int]:
        return [r[0] for r in self.db.execute("SELECT id FROM DataSeries ORDER BY id ASC")]

    def create_series(self, name: str = "") -> "Series":
        """Storage::create_series_id (storage.cpp:129-145): the oldest series
        goes when the limit is reached."""
        if self.num_series() >= self.max_series:
            self.db.execute("DELETE FROM DataSeries WHERE id IN (SELECT id FROM DataSeries "
                            "ORDER BY id ASC LIMIT 1)")
        cur = self.db.execute("INSERT INTO DataSeries (name) VALUES (?)", (name,))
        self.db.commit()
        return Series(self, cur.lastrowid)

    def last_series(self) -> "Series":
        row = self.db.execute("SELECT id FROM DataSeries ORDER BY id DESC LIMIT 1").fetchone()
        if row is None:
            raise LookupError("no series in the storage")
        return Series(self, row[0])


class Series:
    def __init__(self, storage: Storage, sid: int):
        self.storage, self.id = storage, sid

    @property
    def name(self) -> str:
        return self.storage.db.execute("SELECT name FROM DataSeries WHERE id = ?",
                                       (self.id,)).fetchone()[0]

    def frames(self) -> list["Frame"]:
        return [Frame(self.storage, r[0]) for r in self.storage.db.execute(
            "SELECT id FROM DataFrames WHERE series_id = ? ORDER BY id ASC", (self.id,))]

    def create_frame(self, time: float) -> "Frame":
        """Storage::series_create_frame_id (storage.cpp:232-245): frame
        times increase."""
        db = self.storage.db
        last = db.execute("SELECT time FROM DataFrames WHERE series_id = ? ORDER BY id DESC "
                          "LIMIT 1", (self.id,)).fetchone()
        if last is not None and not float(time) > last[0]:
            raise ValueError("frame time must be greater than the last frame time")
        cur = db.execute("INSERT INTO DataFrames (series_id, time) VALUES (?, ?)",
                         (self.id, float(time)))
        return Frame(self.storage, cur.lastrowid)


class Frame:
    def __init__(self, storage: Storage, fid: int):
        self.storage, self.id = storage, fid

    @property
    def time(self) -> float:
        return self.storage.db.execute("SELECT time FROM DataFrames WHERE id = ?",
                                       (self.id,)).fetchone()[0]

    def array_names(self) -> list[str]:
        return [r[0] for r in self.storage.db.execute(
            "SELECT name FROM DataArrays WHERE frame_id = ? ORDER BY id ASC", (self.id,))]

    def write_array(self, name: str, data: np.ndarray):
        """Frame::create_array + Array::write (storage.cpp:324-366): the
        array's type id, element count and framed bytes."""
        if not name:
            raise ValueError("array name must not be empty")
        db = self.storage.db
        if db.execute("SELECT id FROM DataArrays WHERE frame_id = ? AND name = ?",
                      (self.id, name)).fetchone():
            raise ValueError(f"array {name!r} already exists")
        data = np.ascontiguousarray(data)
        db.execute("INSERT INTO DataArrays (frame_id, name, type, size, data) "
                   "VALUES (?, ?, ?, ?, ?)",
                   (self.id, name, type_of(data), int(data.shape[0]), zstd_frame(data)))

    def read_array(self, name: str, decode=zstd_decode) -> np.ndarray:
        row = self.storage.db.execute(
            "SELECT type, size, data FROM DataArrays WHERE frame_id = ? AND name = ?",
            (self.id, name)).fetchone()
        if row is None:
            raise KeyError(name)
        dtype, rank, dim = parse_type(row[0])
        shape = (row[1],) + (() if rank == 0 else (dim,) * rank)
        flat = np.frombuffer(decode(row[2]), dtype=dtype)
        return flat.reshape(shape).copy()


# ---------------------------------------------------------------------------
# Particle output


def particle_fields(state: dict, dim: int, masses, equations=None) -> dict:
    """The varying fields written per frame, in the reference layout (fluid
    first, then fixed): r, v, rho, p, m and, with the equations' constants,
    the sound speed cs of the Tait equation (equation_of_state.hpp:49-52)."""
    n = state["rho"].shape[0]
    out = {"r": state["r"].reshape(n, dim), "v": state["v"].reshape(n, dim),
           "rho": state["rho"], "p": state["p"],
           "m": np.broadcast_to(np.asarray(masses, dtype=np.float64), (n,))}
    if equations is not None:
        eq = equations
        out["cs"] = eq.cs0 * (state["rho"] / eq.rho0) ** ((eq.xi - 1.0) / 2.0)
    return out


def write_particles(series: Series, time: float, fields: dict):
    """ParticleArray::write (particle_array.hpp:165-172): one frame at
    `time`, one array per field."""
    frame = series.create_frame(time)
    for name, arr in fields.items():
        frame.write_array(name, np.ascontiguousarray(arr, dtype=np.float64))
    series.storage.db.commit()
    return frame


class FrameWriter:
    """Writes frames on a worker thread so the device keeps stepping: the
    caller downloads the state and `submit`s it; `close` waits for every
    frame to be written."""

    def __init__(self, series: Series):
        self.series = series
        self.q: queue.Queue = queue.Queue(maxsize=4)
        self.error: BaseException | None = None
        self.t = threading.Thread(target=self._run, daemon=True)
        self.t.start()

    def _run(self):
        while True:
            item = self.q.get()
            if item is None:
                return
            if self.error is None:
                try:
                    write_particles(self.series, *item)
                except BaseException as e:  # reported by close()
                    self.error = e

    def submit(self, time: float, fields: dict):
        if self.error is not None:
            raise self.error
        self.q.put((time, fields))

    def close(self):
        self.q.put(None)
        self.t.join()
        if self.error is not None:
            raise self.error
