"""STORMBIN reading on the host (storm_hip_stormbin_read): the reference
loader's row selection (bench/bench_allpairs.cpp:44-72) on hand-built files
whose expected rows are written out literally."""
import os
import struct

import numpy as np
import pytest

import corpora


def _write_raw(path, rows, n_bits, n_rows_hdr=None, magic=b"STORMBIN", tail=b""):
    with open(path, "wb") as f:
        f.write(magic)
        f.write(struct.pack("<IIII", 1, len(rows) if n_rows_hdr is None else n_rows_hdr,
                            n_bits, 0))
        for r in rows:
            f.write(struct.pack("<I", len(r)))
            f.write(np.asarray(r, dtype="<u4").tobytes())
        f.write(tail)


def _rows(offs, pos):
    return [pos[offs[i]:offs[i + 1]].tolist() for i in range(len(offs) - 1)]


ROWS = [[1, 2], [], [3], [4, 5, 6], [7], [], [8, 9]]


def test_reads_every_nonempty_row(sh, tmp_path):
    p = tmp_path / "a.bin"
    _write_raw(p, ROWS, 10)
    offs, pos, u = sh.stormbin_read(str(p))
    assert u == 10
    assert _rows(offs, pos) == [[1, 2], [3], [4, 5, 6], [7], [8, 9]]


def test_stride_counts_empty_rows_and_max_rows_caps(sh, tmp_path):
    p = tmp_path / "a.bin"
    _write_raw(p, ROWS, 10)
    # seen indices 0, 2, 4, 6 are candidates; none of them is empty
    offs, pos, _ = sh.stormbin_read(str(p), stride=2)
    assert _rows(offs, pos) == [[1, 2], [3], [7], [8, 9]]
    # stride 3: indices 0, 3, 6
    offs, pos, _ = sh.stormbin_read(str(p), stride=3)
    assert _rows(offs, pos) == [[1, 2], [4, 5, 6], [8, 9]]
    offs, pos, _ = sh.stormbin_read(str(p), max_rows=3)
    assert _rows(offs, pos) == [[1, 2], [3], [4, 5, 6]]
    # stride 0 means 1
    offs, pos, _ = sh.stormbin_read(str(p), stride=0)
    assert len(offs) == 6


def test_truncated_record_ends_the_file(sh, tmp_path):
    p = tmp_path / "t.bin"
    # a record announcing 5 positions with only 2 present
    _write_raw(p, [[1], [2, 3]], 10, tail=struct.pack("<III", 5, 7, 8))
    offs, pos, _ = sh.stormbin_read(str(p))
    assert _rows(offs, pos) == [[1], [2, 3]]
    # header row count is not used (as the loader)
    _write_raw(p, [[1], [2, 3], [4]], 10, n_rows_hdr=99)
    offs, pos, _ = sh.stormbin_read(str(p))
    assert _rows(offs, pos) == [[1], [2, 3], [4]]


def test_rejects_bad_files(sh, tmp_path):
    p = tmp_path / "bad.bin"
    _write_raw(p, ROWS, 10, magic=b"STORMBOX")
    with pytest.raises(sh.StormHipError):
        sh.stormbin_read(str(p))
    _write_raw(p, [[1], []], 10)          # one non-empty row
    with pytest.raises(sh.StormHipError):
        sh.stormbin_read(str(p))
    with pytest.raises(sh.StormHipError):
        sh.stormbin_read(str(tmp_path / "missing.bin"))
    p.write_bytes(b"STORMBIN")            # header cut short
    with pytest.raises(sh.StormHipError):
        sh.stormbin_read(str(p))


def test_writer_round_trips_generated_rows(sh, tmp_path):
    rows = corpora.mixed_corpus(3, 300, 5000, 0.02, structure="runs")
    rows = [r for r in rows if len(r)]
    offs, pos = corpora.csr(rows)
    p = tmp_path / "g.bin"
    sh.write_stormbin(str(p), offs, pos, 5000)
    assert os.path.getsize(p) == 24 + 4 * (len(rows) + len(pos))
    o2, p2, u = sh.stormbin_read(str(p))
    assert u == 5000
    assert np.array_equal(o2, offs) and np.array_equal(p2, pos)
