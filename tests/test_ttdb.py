"""The .ttdb writer (titsolver_amd/ttdb.py) against the reference's format.

Pinned by the reference's own data file tests/_data/particles.ttdb, through
the fixture tests/golden/ttdb_reference.json (table definitions, type ids,
element counts and blob framing of a 1,660-particle run), and by an
independent zstd decoder, the system's libzstd, where it can be loaded.
"""
from __future__ import annotations

import ctypes as C
import json
import sqlite3
from pathlib import Path

import numpy as np
import pytest

from titsolver_amd import ttdb

FIXTURE = json.loads((Path(__file__).parent / "golden" / "ttdb_reference.json").read_text())


def _libzstd():
    try:
        z = C.CDLL("libzstd.so.1")
    except OSError:
        return None
    z.ZSTD_getFrameContentSize.restype = C.c_ulonglong
    z.ZSTD_getFrameContentSize.argtypes = [C.c_void_p, C.c_size_t]
    z.ZSTD_decompress.restype = C.c_size_t
    z.ZSTD_decompress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]
    z.ZSTD_isError.restype = C.c_uint
    z.ZSTD_isError.argtypes = [C.c_size_t]
    return z


def _libzstd_decode(z, blob: bytes) -> bytes:
    n = z.ZSTD_getFrameContentSize(blob, len(blob))
    buf = C.create_string_buffer(max(int(n), 1))
    got = z.ZSTD_decompress(buf, n, blob, len(blob))
    assert not z.ZSTD_isError(got) and got == n
    return buf.raw[:n]


def test_schema_matches_reference(tmp_path):
    st = ttdb.Storage(tmp_path / "a.ttdb")
    st.close()
    con = sqlite3.connect(tmp_path / "a.ttdb")
    for name, cols in FIXTURE["tables"].items():
        got = [[c[1], c[2], c[3], c[5]] for c in con.execute(f"PRAGMA table_info({name})")]
        want = [list(c) for c in cols]
        if name == "DataSeries":
            # the fixture predates storage.cpp:43, which made the name NOT NULL
            want[1][2] = 1
        assert got == want, name
    assert con.execute("SELECT max_series FROM Settings").fetchone()[0] == 5


def test_type_ids_match_reference():
    ids = {a["name"]: a["type"] for a in FIXTURE["arrays"] if a["frame"] == 1}
    assert ttdb.type_of(np.zeros((3, 2))) == ids["r"] == 131338
    assert ttdb.type_of(np.zeros(3)) == ids["rho"] == 65546
    assert ttdb.type_of(np.zeros((3, 2, 2))) == ids["L"] == 131594
    assert ttdb.type_of(np.zeros(3, dtype=np.uint64)) == ids["parinfo"] == 65544
    assert ttdb.type_of(np.zeros((3, 3))) == 196874
    for tid in set(ids.values()):
        dtype, rank, dim = ttdb.parse_type(tid)
        assert ttdb.type_id(dtype, rank, dim) == tid


def test_frame_header_and_trailer_match_reference():
    """A 1,660-particle vector field (26,560 bytes) gets the reference's
    single-segment header with a two-byte content size, a scalar field
    (13,280 bytes) likewise, and every blob ends with the reference's empty
    frame."""
    by_name = {a["name"]: a for a in FIXTURE["arrays"] if a["frame"] == 1}
    vec = ttdb.zstd_frame(np.zeros((1660, 2)) + 0.5)
    sca = ttdb.zstd_frame(np.arange(1660, dtype=np.float64))
    assert vec[:7].hex() == by_name["r"]["head_hex"]
    assert sca[:7].hex() == by_name["rho"]["head_hex"]
    for blob in (vec, sca):
        assert blob[-9:].hex() == by_name["r"]["tail_hex"]


@pytest.mark.parametrize("nbytes", [0, 1, 255, 256, 65791, 65792, 131072, 131073, 400000])
def test_zstd_frames_round_trip(nbytes):
    rng = np.random.default_rng(nbytes)
    data = rng.integers(0, 256, size=nbytes, dtype=np.uint8).tobytes()
    blob = ttdb.zstd_frame(np.frombuffer(data, dtype=np.uint8))
    assert ttdb.zstd_decode(blob) == data
    z = _libzstd()
    if z is not None and nbytes > 0:
        assert _libzstd_decode(z, blob) == data


def test_rle_blocks_for_constant_fields():
    """An all-zero field (v at rest) shrinks to RLE blocks; a field that is
    constant only in its first block keeps stored blocks after it."""
    zeros = ttdb.zstd_frame(np.zeros((200000, 2)))
    assert len(zeros) < 4 * (3_200_000 // 131072 + 1) + 32  # 4 bytes per 128 KiB block
    mixed = np.zeros(40000)
    mixed[20000:] = np.arange(20000)
    blob = ttdb.zstd_frame(mixed)
    assert ttdb.zstd_decode(blob) == mixed.tobytes()
    z = _libzstd()
    if z is not None:
        assert _libzstd_decode(z, zeros) == np.zeros((200000, 2)).tobytes()
        assert _libzstd_decode(z, blob) == mixed.tobytes()


def test_compressed_block_is_refused():
    blob = bytes.fromhex("28b52ffd20050500000100" + "00" * 8)
    with pytest.raises(ValueError):
        ttdb.zstd_decode(blob)


def test_particles_round_trip_and_series_limit(tmp_path):
    """ParticleArray::write's frames (one per call, one array per field) read
    back bit for bit; frame times must increase; the oldest series goes when
    the limit (wcsph.cpp:162 keeps one) is reached."""
    path = tmp_path / "particles.ttdb"
    rng = np.random.default_rng(3)
    n = 1000
    st = ttdb.Storage(path)
    st.set_max_series(1)
    st.create_series("old")
    series = st.create_series()
    assert st.num_series() == 1
    states = []
    for k in range(3):
        s = {"r": rng.normal(size=(n, 2)), "v": rng.normal(size=(n, 2)),
             "rho": 1000 + rng.normal(size=n), "p": rng.normal(size=n)}
        states.append(s)
        ttdb.write_particles(series, 0.1 * k, ttdb.particle_fields(s, 2, 0.4))
    with pytest.raises(ValueError):
        series.create_frame(0.1)
    st.close()
    ro = ttdb.Storage(path, read_only=True)
    ser = ro.last_series()
    frames = ser.frames()
    assert [f.time for f in frames] == [0.0, 0.1, 0.2]
    for f, s in zip(frames, states):
        assert f.array_names() == ["r", "v", "rho", "p", "m"]
        for k in ("r", "v", "rho", "p"):
            assert np.array_equal(f.read_array(k), s[k]), k
        assert np.all(f.read_array("m") == 0.4)
    z = _libzstd()
    if z is not None:
        got = frames[1].read_array("r", decode=lambda b: _libzstd_decode(z, b))
        assert np.array_equal(got, states[1]["r"])
    ro.close()


def test_frame_writer_thread(tmp_path):
    st = ttdb.Storage(tmp_path / "w.ttdb")
    w = ttdb.FrameWriter(st.create_series("run"))
    xs = [np.full((500, 3), float(k)) for k in range(6)]
    for k, x in enumerate(xs):
        w.submit(float(k), {"r": x, "rho": x[:, 0].copy()})
    w.close()
    frames = st.last_series().frames()
    assert len(frames) == 6
    assert all(np.array_equal(f.read_array("r"), xs[k]) for k, f in enumerate(frames))
    st.close()


def test_reference_file_decodes(tmp_path):
    """The reference's own particles.ttdb, where it is present, read through
    this module with libzstd for its compressed blocks: frame 1's r and rho
    match the fixture's sums."""
    src = Path("/root/reference/tests/_data/particles.ttdb")
    z = _libzstd()
    if not src.exists() or z is None:
        pytest.skip("reference data file or libzstd not present")
    st = ttdb.Storage(src, read_only=True)
    st.db.close()
    st.db = sqlite3.connect(f"file:{src}?immutable=1", uri=True)
    fr = st.last_series().frames()[0]
    r = fr.read_array("r", decode=lambda b: _libzstd_decode(z, b))
    rho = fr.read_array("rho", decode=lambda b: _libzstd_decode(z, b))
    assert r.shape == (1660, 2) and rho.shape == (1660,)
    assert r.sum() == pytest.approx(FIXTURE["decoded"]["r"]["sum"], rel=1e-15)
    assert rho.sum() == pytest.approx(FIXTURE["decoded"]["rho"]["sum"], rel=1e-15)
    st.db.close()
