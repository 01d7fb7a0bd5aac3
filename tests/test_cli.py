"""The drop-in `song` CLI (reference main.cu / main.cc argv and files) on the CPU side:
`song build` writes bfsg.data / bfsg.graph byte-identical to the reference song.cpu."""
import hashlib
import os
import subprocess

import pytest

import golden_cases
from song_amd import datasets

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SONG = os.path.join(ROOT, "song_amd", "bin", "song")
CLI = os.path.join(ROOT, "song_amd", "cli")


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


@pytest.mark.parametrize("name", ["tiny_cos_500", "tie_l2_3k", "tie_ip_2k", "c1_cos_15k"])
def test_song_build_matches_reference_files(tmp_path, name):
    meta, base, _, _, _ = golden_cases.load(name)
    n, dim = base.shape
    datasets.write_libsvm(str(tmp_path / "base.svm"), base)
    r = subprocess.run([SONG, "build", "base.svm", "0", "0", str(n), str(dim), "0", meta["metric"]],
                       cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert sha(tmp_path / "bfsg.graph") == meta["graph_sha256"]
    assert sha(tmp_path / "bfsg.data") == meta["data_sha256"]


def test_build_graph_script(tmp_path):
    """The scripts run beside a copy of the binary, as fill_parameters.sh lays them out."""
    import shutil

    meta, base, _, _, _ = golden_cases.load("tiny_l2_500")
    datasets.write_libsvm(str(tmp_path / "b.svm"), base)
    for f in ("build_graph.sh", "test_query.sh"):
        shutil.copy(os.path.join(CLI, f), tmp_path / f)
    shutil.copy(SONG, tmp_path / "song")
    r = subprocess.run(["./build_graph.sh", "b.svm", "500", str(base.shape[1]), "l2"], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert sha(tmp_path / "bfsg.graph") == meta["graph_sha256"]
    r = subprocess.run(["./build_graph.sh", "b.svm"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr


def test_usage_and_errors(tmp_path):
    r = subprocess.run([SONG, "build"], capture_output=True, text=True)
    assert r.returncode == 1 and r.stdout.startswith("Usage:")
    r = subprocess.run([SONG, "build", "missing.svm", "0", "0", "10", "4", "0", "l2"], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode == 1 and "not found" in r.stderr
    r = subprocess.run([SONG, "build", "x", "0", "0", "10", "4", "0", "hamming"], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode == 1 and r.stdout.startswith("Usage:")
    (tmp_path / "bad.svm").write_text("0 1:1 9:2\n")
    r = subprocess.run([SONG, "build", "bad.svm", "0", "0", "10", "4", "0", "l2"], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode == 1 and "feature index" in r.stderr


def test_query_file_without_trailing_newline_is_read_whole(tmp_path):
    """`song test` counts query records as fgets reads them (parser.h:57): a last line without
    '\\n' is a record.  Without a GPU the run gets past the parse and stops at the index upload."""
    meta, base, query, _, _ = golden_cases.load("tiny_l2_500")
    n, dim = base.shape
    datasets.write_libsvm(str(tmp_path / "base.svm"), base)
    r = subprocess.run([SONG, "build", "base.svm", "0", "0", str(n), str(dim), "0", "l2"], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    datasets.write_libsvm(str(tmp_path / "q.svm"), query[:3])
    text = (tmp_path / "q.svm").read_text()
    (tmp_path / "q.svm").write_text(text.rstrip("\n"))
    r = subprocess.run([SONG, "test", "0", "q.svm", "20", str(n), str(dim), "5", "l2"], cwd=tmp_path,
                       capture_output=True, text=True)
    assert "more than" not in r.stderr and "parse" not in r.stderr, r.stderr
    if r.returncode != 0:
        assert "GPU" in r.stderr or "index" in r.stderr, r.stderr
    else:
        assert len(r.stdout.strip("\n").split("\n")) == 3
