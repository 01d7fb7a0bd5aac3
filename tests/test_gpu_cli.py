"""`song test` on the GPU: the printed ids ("%zu " each, one line per query, main.cu:28-33)
equal the reference song.cpu's ids for the same files."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import golden_cases
import pyoracle
from song_amd import datasets

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SONG = os.path.join(ROOT, "song_amd", "bin", "song")
CLI = os.path.join(ROOT, "song_amd", "cli")


@pytest.mark.parametrize("name", ["tiny_cos_500", "c1_cos_15k", "tie_ip_2k"])
def test_song_test_prints_reference_ids(tmp_path, name):
    meta, base, query, ids, _ = golden_cases.load(name)
    n, dim = base.shape
    datasets.write_libsvm(str(tmp_path / "base.svm"), base)
    datasets.write_libsvm(str(tmp_path / "query.svm"), query)
    m = meta["metric"]
    r = subprocess.run([SONG, "build", "base.svm", "0", "0", str(n), str(dim), "0", m], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    pq = meta["pq"]
    r = subprocess.run([SONG, "test", "0", "query.svm", str(pq), str(n), str(dim), str(pq), m], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    assert lines[-1] == "" and len(lines) == len(query) + 1
    for i, line in enumerate(lines[:-1]):
        want = [int(v) for v in ids[i] if v >= 0]
        assert line == "".join(f"{v} " for v in want), (name, i)


def test_test_query_script_display(tmp_path):
    meta, base, query, ids, _ = golden_cases.load("tiny_l2_500")
    n, dim = base.shape
    datasets.write_libsvm(str(tmp_path / "b.svm"), base)
    datasets.write_libsvm(str(tmp_path / "q.svm"), query)
    for f in ("build_graph.sh", "test_query.sh"):
        shutil.copy(os.path.join(CLI, f), tmp_path / f)
    shutil.copy(SONG, tmp_path / "song")
    r = subprocess.run(["./build_graph.sh", "b.svm", str(n), str(dim), "l2"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run(["./test_query.sh", "q.svm", str(n), str(dim), "l2", "5"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = [[int(v) for v in line.split()] for line in r.stdout.strip("\n").split("\n")]
    assert len(got) == len(query)
    assert all(len(g) == 5 for g in got)
    # the script searches at pq 100; the oracle gives the reference's ids at that budget
    graph = pyoracle.build(base, "l2")
    o_ids = pyoracle.search_batch(base, graph, "l2", query, 100, threads=4)[0]
    assert np.array_equal(np.array(got), np.asarray(o_ids)[:, :5].astype(np.int64))


def test_query_file_without_trailing_newline(tmp_path):
    """The last query line has no '\\n': it is still searched (parser.h:57 reads it with fgets)."""
    meta, base, query, ids, _ = golden_cases.load("tiny_cos_500")
    n, dim = base.shape
    datasets.write_libsvm(str(tmp_path / "base.svm"), base)
    datasets.write_libsvm(str(tmp_path / "query.svm"), query)
    (tmp_path / "query.svm").write_text((tmp_path / "query.svm").read_text().rstrip("\n"))
    m, pq = meta["metric"], meta["pq"]
    r = subprocess.run([SONG, "build", "base.svm", "0", "0", str(n), str(dim), "0", m], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([SONG, "test", "0", "query.svm", str(pq), str(n), str(dim), str(pq), m], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    assert len(lines) == len(query) + 1
    for i, line in enumerate(lines[:-1]):
        assert line == "".join(f"{v} " for v in ids[i] if v >= 0), i
