"""Input path on the CPU: the parallel LIBSVM parser reads every record as the reference's
sequential fgets + sscanf loop does (parser.h:18-62, restated by the oracle's line parser),
the native writer round-trips every float exactly, and the binary query formats are read and
size-checked."""
import os

import numpy as np
import pytest

import pyoracle
import song_amd


def _oracle_rows(text, dim):
    rows = []
    for line in text.splitlines(keepends=True):
        r = np.zeros(dim, dtype=np.float32)
        for i, v in pyoracle.parse_libsvm_line(line):
            r[i] = v
        rows.append(r)
    return np.array(rows, dtype=np.float32).reshape(-1, dim)


def test_parallel_parser_equals_the_record_by_record_reading(tmp_path):
    rng = np.random.default_rng(3)
    dim = 12
    lines = []
    for i in range(20000):
        kind = i % 7
        nz = rng.choice(dim, size=rng.integers(0, dim + 1), replace=False)
        vals = rng.standard_normal(len(nz)) * 10.0 ** rng.integers(-3, 4)
        body = " ".join(f"{j + 1}:{float(v)!r}" for j, v in zip(nz, vals))
        if kind == 1 and len(nz):
            body += " "          # trailing space: an odd token, ignored
        elif kind == 2:
            body = ""            # empty record
        elif kind == 3:
            body = body.replace(":", ":+", 1)  # a signed value
        elif kind == 4:
            body += " 3:1e-45 4:-0"
        lines.append(f"{i % 5} {body}\n" if kind != 2 else "\n")
    text = "".join(lines).rstrip("\n")  # the last record has no newline
    p = tmp_path / "q.svm"
    p.write_text(text)
    want = _oracle_rows(text, dim)
    for threads in (1, 3, 8):
        got, n = song_amd.parse_libsvm(str(p), len(want) + 5, dim, threads=threads)
        assert n == len(want)
        assert np.array_equal(got[:n].view(np.uint32), want.view(np.uint32)), threads
        assert not got[n:].any()


def test_parallel_parser_reports_the_first_error_in_file_order(tmp_path):
    dim = 4
    rows = ["0 1:1 2:2\n"] * 30000
    rows[21000] = "0 9:1\n"
    rows[25000] = "0 0:1\n"
    p = tmp_path / "bad.svm"
    p.write_text("".join(rows))
    for threads in (1, 8):
        with pytest.raises(song_amd.SongError, match="line 21001: feature index 9"):
            song_amd.parse_libsvm(str(p), 30000, dim, threads=threads)
    p.write_text("0 1:1\n" * 11)
    with pytest.raises(song_amd.SongError, match="more than 10 lines"):
        song_amd.parse_libsvm(str(p), 10, dim, threads=4)


def test_writer_round_trips_every_float(tmp_path):
    rng = np.random.default_rng(4)
    x = (rng.standard_normal((5000, 33)) * 10.0 ** rng.integers(-30, 30, size=(5000, 33))).astype(np.float32)
    x[::7] = np.rint(x[::7] * 0 + rng.integers(-300, 300, size=(1, 33)))
    x[3, :5] = [np.float32(1e-45), -np.float32(3.4e38), 0.0, np.float32(0.1), np.float32(16777217.0)]
    p = str(tmp_path / "w.svm")
    song_amd.write_libsvm(p, x, threads=4)
    got, n = song_amd.parse_libsvm(p, len(x), x.shape[1])
    assert n == len(x)
    assert np.array_equal(got.view(np.uint32), (x + np.float32(0.0)).view(np.uint32))


def test_binary_query_formats(tmp_path):
    rng = np.random.default_rng(5)
    q = rng.standard_normal((37, 24)).astype(np.float32)
    q.tofile(tmp_path / "q.f32")
    with open(tmp_path / "q.fvecs", "wb") as f:
        for r in q:
            np.int32(24).tofile(f)
            r.tofile(f)
    with open(tmp_path / "q.fbin", "wb") as f:
        np.array([37, 24], dtype=np.uint32).tofile(f)
        q.tofile(f)
    for name in ("q.f32", "q.fvecs", "q.fbin"):
        assert np.array_equal(song_amd.read_queries(str(tmp_path / name), 24), q), name
    # sizes that do not match the format are errors
    with pytest.raises(song_amd.SongError, match="whole number"):
        song_amd.read_queries(str(tmp_path / "q.f32"), 25)
    with pytest.raises(song_amd.SongError, match="fbin dim"):
        song_amd.read_queries(str(tmp_path / "q.fbin"), 25)
    with open(tmp_path / "bad.fvecs", "wb") as f:
        for i, r in enumerate(q[:3]):
            np.int32(24 if i != 1 else 23).tofile(f)
            r.tofile(f)
    with pytest.raises(song_amd.SongError, match="row 1: dimension 23"):
        song_amd.read_queries(str(tmp_path / "bad.fvecs"), 24)
    with open(tmp_path / "short.fbin", "wb") as f:
        np.array([38, 24], dtype=np.uint32).tofile(f)
        q.tofile(f)
    with pytest.raises(song_amd.SongError, match="header says 38 rows"):
        song_amd.read_queries(str(tmp_path / "short.fbin"), 24)
    # LIBSVM through the same entry point
    song_amd.write_libsvm(str(tmp_path / "q.svm"), q)
    assert np.array_equal(song_amd.read_queries(str(tmp_path / "q.svm"), 24), q)
