"""Structure of the compiled persistent search kernels (gfx950 assembly of this build).

The persistent kernel's loop over work-list positions must stay wave-uniform: one loop
at depth 1 whose head takes the next position with one atomic, broadcasts it
(readfirstlane) and leaves on a scalar compare.  A lane-0-only branch at the head let
the compiler thread the back edge from the query's lane-0 epilogue, so that lanes 1..63
started the next iteration without lane 0 (and re-read position 0 for ever); this test
fails on that shape.
"""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "song_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"


def _asm(metric, out_dir):
    out = os.path.join(out_dir, f"{metric}.s")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include"), "--cuda-device-only", "-S",
           os.path.join(CSRC, f"search_{metric}.hip"), "-o", out]
    return subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE), out


def _functions(text, prefix):
    out, cur, name = {}, None, None
    for line in text.split("\n"):
        m = re.match(r"^(" + prefix + r"\w+):", line)
        if m:
            name, cur = m.group(1), []
            continue
        if cur is not None:
            if line.startswith(".Lfunc_end"):
                out[name], cur = cur, None
            else:
                cur.append(line)
    return out


def loop_head_is_uniform(body):
    """True if the work loop is one depth-1 loop: atomic, readfirstlane, scalar compare,
    and no exec-narrowing latch in front of its header."""
    d1 = [k for k, x in enumerate(body) if "This Loop Header: Depth=1" in x]
    at = [k for k, x in enumerate(body) if "global_atomic_add" in x]
    if len(d1) != 1 or not at or at[0] < d1[0]:
        return False
    seg = "\n".join(body[at[0]:at[0] + 25])
    if "v_readfirstlane_b32" not in seg or "s_cmp_ge_u32" not in seg:
        return False
    return "s_andn2_b64 exec, exec" not in "\n".join(body[max(0, d1[0] - 8):d1[0]])


@pytest.fixture(scope="module")
def search_asm():
    """Assembly text of search_{l2,ip,cos}.hip for gfx950, compiled once for this module."""
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not installed")
    tmp = tempfile.mkdtemp(prefix="song_asm_")
    try:
        jobs = [_asm(m, tmp) for m in ("l2", "ip", "cos")]
        texts = []
        for p, out in jobs:
            _, err = p.communicate(timeout=600)
            assert p.returncode == 0, err.decode(errors="replace")[-2000:]
            texts.append(open(out).read())
        yield texts
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def work_loop_is_uniform(body):
    """As loop_head_is_uniform for a kernel with other loops beside its work loop (the helper
    wave's in search_duo): the depth-1 loop whose head holds the first atomic add."""
    at = [k for k, x in enumerate(body) if "global_atomic_add" in x]
    if not at:
        return False
    d1 = [k for k in range(at[0]) if "This Loop Header: Depth=1" in body[k]]
    if not d1:
        return False
    seg = "\n".join(body[at[0]:at[0] + 25])
    # the exit compare: scalar, or a vector compare of the broadcast position with the count
    # (loaded into a vector register, equal in every lane) and a branch on vcc
    scalar_exit = "s_cmp_ge_u32" in seg
    vector_exit = re.search(r"v_cmp_ge_u32_e32 vcc, s\d+, v\d+", seg) is not None and "s_cbranch_vccnz" in seg
    if "v_readfirstlane_b32" not in seg or not (scalar_exit or vector_exit):
        return False
    return "s_andn2_b64 exec, exec" not in "\n".join(body[max(0, d1[-1] - 8):d1[-1]])


def test_persistent_loop_is_wave_uniform(search_asm):
    kernels = {}
    for text in search_asm:
        kernels.update(_functions(text, r"_ZN4song17search_persistent"))
    # 3 metrics x (exact, fp32) x 3 row storages x 5 row shapes, and the LDS-bag and global
    # best-list forms (float32 rows) of each metric, mode and shape
    assert len(kernels) == 120
    bad = [k for k, body in kernels.items() if not loop_head_is_uniform(body)]
    assert not bad, f"{len(bad)} persistent kernels with a split work loop, e.g. {bad[:3]}"


def test_two_wave_kernel_work_loop_is_wave_uniform(search_asm):
    kernels = {}
    for text in search_asm:
        kernels.update(_functions(text, r"_ZN4song10search_duo"))
    assert len(kernels) == 30  # 3 metrics x 5 row shapes x best list front in LDS / global (FP32 mode)
    bad = [k for k, body in kernels.items() if not work_loop_is_uniform(body)]
    assert not bad, f"{len(bad)} two-wave kernels with a split work loop, e.g. {bad[:3]}"


def test_checker_flags_a_split_loop():
    """The shape check itself: a latch that narrows exec before the header is rejected."""
    good = ["s_branch .LBB0_3", ".LBB0_3:  ; =>This Loop Header: Depth=1", "global_atomic_add v1, v3, v1, s[4:5] sc0",
            "v_readfirstlane_b32 s0, v1", "s_cmp_ge_u32 s58, s0"]
    assert loop_head_is_uniform(good)
    split = ["s_andn2_b64 exec, exec, s[0:1]", "s_cbranch_execz .LBB0_298"] + good[1:3] + ["s_nop 0"] * 30 + good[3:]
    assert not loop_head_is_uniform(split)
    assert work_loop_is_uniform(["s_cbranch_scc1 .LBB0_9", ".LBB0_2:  ; Inner Loop Header: Depth=1", "s_sleep 2",
                                 "s_branch .LBB0_2"] + good)
    assert not work_loop_is_uniform(["s_cbranch_scc1 .LBB0_9"] + split)


def _kernel_metadata(text):
    """{kernel name: {field: int}} from the code object metadata of an assembly listing."""
    meta = text[text.index("amdhsa.kernels"):]
    out = {}
    for block in re.split(r"\n  - \.", meta):
        m = re.search(r"\.name:\s+(\S+)", block)
        if not m:
            continue
        fields = {k: int(v) for k, v in re.findall(r"\.(vgpr_spill_count|private_segment_fixed_size|vgpr_count):\s+(\d+)",
                                                   block)}
        out[m.group(1)] = fields
    return out


def test_search_kernels_keep_every_vector_register_in_registers(search_asm):
    """No search kernel spills vector registers to scratch memory.  The kernels keep per-lane state
    (the queue head, the best list's tail) live across divergent regions; a spill under a partial
    exec mask would save only the active lanes' values, and scratch round trips would also lengthen
    every expansion."""
    kernels = {}
    for text in search_asm:
        kernels.update({k: v for k, v in _kernel_metadata(text).items() if "search" in k})
    assert len(kernels) == 3 * 80
    bad = {k: v for k, v in kernels.items() if v.get("vgpr_spill_count", 0) or v.get("private_segment_fixed_size", 0)}
    assert not bad, f"{len(bad)} search kernels use scratch, e.g. {list(bad.items())[:2]}"
