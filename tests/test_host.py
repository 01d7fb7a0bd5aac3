"""Host-side (CPU) parts of the drop-in: builder, LIBSVM parser, bfsg files, ABI exports."""
import ctypes
import hashlib
import os
import re

import numpy as np
import pytest

import golden_cases
import pyoracle
import song_amd
from song_amd import datasets

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("name", ["tiny_cos_500", "tiny_l2_500", "tie_l2_3k", "tie_ip_2k", "ip_10k",
                                  "c1_cos_15k", "mix_l2_20k"])
@pytest.mark.parametrize("threads", [1, 4])
def test_parallel_builder_matches_reference_graph(name, threads):
    meta, base, _, _, graph = golden_cases.load(name)
    g = song_amd.build_graph(base, meta["metric"], threads=threads)
    assert golden_cases.sha256(g) == meta["graph_sha256"]


def test_builder_1m_matches_reference_graph():
    """Full-scale pin: the 1M x 128 mixture graph equals the reference song.cpu build."""
    meta, _ = golden_cases.load_c2_1m()
    base, _ = datasets.mixture_l2()
    assert golden_cases.sha256(base) == meta["data_sha256"]
    g, st = song_amd.build_graph(base, "l2", threads=0, return_stats=True)
    assert golden_cases.sha256(g) == meta["graph_sha256"]
    assert st["redone"] < st["searches"] // 4


def test_builder_matches_oracle_on_random_cases():
    rng = np.random.default_rng(123)
    for metric in ("l2", "ip", "cos"):
        base = rng.standard_normal((700, 9)).astype(np.float32)
        assert np.array_equal(song_amd.build_graph(base, metric, threads=3), pyoracle.build(base, metric))


def test_libsvm_parser_matches_oracle(tmp_path):
    base, _ = datasets.uniform_cos(n_base=200, n_query=1, dim=16, seed=4)
    path = str(tmp_path / "x.svm")
    datasets.write_libsvm(path, base)
    got, rows = song_amd.parse_libsvm(path, 200, 16)
    assert rows == 200
    assert np.array_equal(got, base)
    with open(path) as f:
        for i, line in enumerate(f):
            ref = pyoracle.parse_libsvm_line(line)
            dense = np.zeros(16, np.float32)
            for j, v in ref:
                dense[j] = v
            assert np.array_equal(dense, got[i])


def test_libsvm_parser_errors(tmp_path):
    p = tmp_path / "bad.svm"
    p.write_text("0 1:1 40:2\n")
    with pytest.raises(song_amd.SongError, match="feature index"):
        song_amd.parse_libsvm(str(p), 1, 16)
    p.write_text("0 1:1\n0 2:2\n")
    with pytest.raises(song_amd.SongError, match="more than"):
        song_amd.parse_libsvm(str(p), 1, 16)
    with pytest.raises(song_amd.SongError, match="not found"):
        song_amd.parse_libsvm(str(tmp_path / "missing.svm"), 1, 16)


def test_file_roundtrip_and_size_check(tmp_path):
    L = song_amd.host_lib()
    a = np.arange(100, dtype=np.uint64)
    p = str(tmp_path / "g").encode()
    assert L.song_file_write(p, a.ctypes.data, a.nbytes) == 0
    b = np.empty_like(a)
    assert L.song_file_read(p, b.ctypes.data, b.nbytes) == 0
    assert np.array_equal(a, b)
    c = np.empty(101, dtype=np.uint64)
    assert L.song_file_read(p, c.ctypes.data, c.nbytes) != 0
    assert b"expected" in L.song_host_last_error()


def _header_functions(path):
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(song_\w+)\s*\(", txt, flags=re.M)))


def test_libraries_export_every_declared_symbol():
    hip = ctypes.CDLL(song_amd.LIB_PATH)
    host = ctypes.CDLL(song_amd.HOST_LIB_PATH)
    for header, L in (("song_amd.h", hip), ("song_learned.h", hip), ("song_host.h", host)):
        names = _header_functions(os.path.join(ROOT, "include", header))
        assert names, header
        for n in names:
            assert hasattr(L, n), f"{n} declared in {header} but not exported"


def test_search_without_gpu_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    base = np.zeros((4, 4), np.float32)
    g = np.zeros((4, 32), np.uint64)
    with pytest.raises(song_amd.SongError, match="no GPU"):
        song_amd.Index(base, g, "l2")


def test_index_create_validates_graph_without_gpu_side_effects():
    """Validation errors are reported before any device work (checked with no GPU too)."""
    import torch

    base = np.zeros((4, 4), np.float32)
    g = np.zeros((4, 32), np.uint64)
    g[1, 0] = 40  # degree beyond 31
    with pytest.raises(song_amd.SongError, match="degree"):
        song_amd.Index(base, g, "l2")
    g[1, 0] = 1
    g[1, 1] = 9  # id beyond n
    with pytest.raises(song_amd.SongError, match="neighbour id"):
        song_amd.Index(base, g, "l2")


def test_ctypes_structs_match_the_c_header(tmp_path):
    """The Python mirrors of song_search_opts, song_tier and song_batch_config have the sizes and
    field offsets the C compiler gives include/song_amd.h (a mismatch corrupts every call)."""
    import ctypes
    import shutil
    import subprocess

    if not shutil.which("gcc"):
        pytest.skip("gcc not installed")
    fields = {"song_search_opts": (song_amd.SearchOpts, ["mode", "hash_log2", "rerank", "visited_bits"]),
              "song_tier": (song_amd.Tier, ["pq_max", "groups_per_cu", "visited_bits"]),
              "song_batch_config": (song_amd.BatchConfig, ["tier", "retry", "parallel_tiers", "priority_first"])}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "song_amd.h"', "int main(void) {"]
    for cname, (_, names) in fields.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in names:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "abi.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "abi"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n"):
        if line:
            cname, what, val = line.split()
            got[(cname, what)] = int(val)
    for cname, (cls, names) in fields.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for f in names:
            assert got[(cname, f)] == getattr(cls, f).offset, (cname, f)
