"""Batching form on the GPU: persistent per-tier kernels with uneven k and pq per query.

Bar: every query's ids, distances and flags equal the plain search's at the same
(k, pq) -- one plain call per distinct pq over that pq's queries -- under every tier
configuration (waves per group, filter size with global spills, lanes per row, heap
depth, grid size, overflow re-runs), through hipGraph replays, and after autotuning;
the queries searched at the golden pq equal the reference CPU search's ids.
"""
import numpy as np
import pytest

import golden_cases
import pyoracle
import song_amd
from song_amd import autotune, datasets, tier

pytestmark = pytest.mark.gpu

PQS = (32, 64, 100, 128, 256, 512)


def _dev():
    import torch

    return torch.device("cuda:0")


@pytest.fixture(scope="module")
def c1():
    """15k x 16 cosine (exact mode), 2000 queries with uneven (k, pq)."""
    import torch

    meta, base, _, ref_ids, graph = golden_cases.load("c1_cos_15k")
    graph = pyoracle.build(base, "cos") if graph is None else graph
    assert golden_cases.sha256(graph) == meta["graph_sha256"]
    _, query = datasets.uniform_cos(**meta["kwargs"])
    query = query[:2000]
    k, pq = datasets.hetero_budgets(len(query), seed=11, ks=(1, 10, 100), pqs=PQS)
    # the golden queries keep the golden pq (100) and k = 100 to compare with the reference
    ng = meta["n_query"]
    pq[:ng], k[:ng] = meta["pq"], meta["pq"]
    ix = song_amd.Index(base, graph, "cos")
    dev = _dev()
    q = torch.from_numpy(np.ascontiguousarray(query)).to(dev)
    return dict(meta=meta, base=base, graph=graph, ix=ix, q=q, k=k, pq=pq, ref_ids=ref_ids, ng=ng,
                k_d=torch.from_numpy(k).to(dev), pq_d=torch.from_numpy(pq).to(dev))


def _plain(ix, q, k, pq, mode, storage=song_amd.STORE_F32, entry=None, steps=None, k_ld=100):
    """Plain search at each query's (k, pq): one call per distinct pq; -1 pads beyond k."""
    import torch

    nq = q.shape[0]
    ids = torch.full((nq, k_ld), -1, dtype=torch.int32, device=q.device)
    ds = torch.full((nq, k_ld), float("inf"), dtype=torch.float32, device=q.device)
    fl = torch.zeros(nq, dtype=torch.int32, device=q.device)
    for p in sorted(set(pq.tolist())):
        sel = np.nonzero(pq == p)[0]
        idx = torch.from_numpy(sel).to(q.device)
        kk = min(p, k_ld)
        i2 = torch.empty((len(sel), kk), dtype=torch.int32, device=q.device)
        d2 = torch.empty((len(sel), kk), dtype=torch.float32, device=q.device)
        f2 = torch.empty(len(sel), dtype=torch.int32, device=q.device)
        ix.search_device(q[idx].contiguous(), p, kk, i2, d2, f2, mode=mode, storage=storage,
                         entry=entry[idx].contiguous() if entry is not None else None,
                         steps=steps[idx].contiguous() if steps is not None else None)
        ids[idx, :kk], ds[idx, :kk], fl[idx] = i2, d2, f2
    # keep only each query's first k
    kt = torch.from_numpy(k).to(q.device)[:, None]
    col = torch.arange(k_ld, device=q.device)[None, :]
    ids = torch.where(col < kt, ids, torch.full_like(ids, -7))
    ds = torch.where(col < kt, ds, torch.zeros_like(ds))
    return ids, ds, fl


def _batch_run(b, q, k_d, pq_d, k_ld=100, graph=False, entry=None, steps=None):
    import torch

    nq = q.shape[0]
    ids = torch.full((nq, k_ld), -7, dtype=torch.int32, device=q.device)
    ds = torch.zeros((nq, k_ld), dtype=torch.float32, device=q.device)
    fl = torch.full((nq,), -1, dtype=torch.int32, device=q.device)
    if graph:  # captured on a side stream (graph runs need a stream other than the default)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        for _ in range(2):  # capture, then one replay
            b.run(q, ids, fl, pq_arr=pq_d, k_arr=k_d, dists=ds, graph=True, entry=entry, steps=steps,
                  stream=side.cuda_stream)
    else:
        b.run(q, ids, fl, pq_arr=pq_d, k_arr=k_d, dists=ds, graph=graph, entry=entry, steps=steps)
    torch.cuda.synchronize()
    return ids, ds, fl


def _same(got, want, what):
    import torch

    for name, a, w in zip(("ids", "dists", "flags"), got, want):
        if not torch.equal(a, w):
            bad = torch.nonzero((a != w).reshape(a.shape[0], -1).any(1)).flatten()[:5].tolist()
            raise AssertionError(f"{what}: {name} differ at queries {bad}")


CONFIGS = {
    "tier_per_pq": [tier(p) for p in PQS],
    "one_tier_4waves_spill": [tier(512, hash_log2=9, slack=0, waves_per_group=4)],
    "two_tiers_shape3_grid1": [tier(128, shape=3, waves_per_group=2, groups_per_cu=1),
                               tier(512, hash_log2=10, shape=0, groups_per_cu=2)],
    # tables up to 2^15: the re-run's default (the largest tier's table x4) would not fit one
    # wave's LDS, so it takes the largest table that does
    "large_tables": [tier(128, hash_log2=14), tier(512, hash_log2=15)],
    # the queue's overflow bag in LDS, with little tie room so that refills and truncations run
    "lds_bag": [tier(100, bag_in_lds=1, slack=0), tier(512, bag_in_lds=1, waves_per_group=2, slack=8, shape=2)],
}


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_batch_equals_plain_per_query(c1, name):
    import torch

    print(f"[batch] config {name}", flush=True)
    want = _plain(c1["ix"], c1["q"], c1["k"], c1["pq"], song_amd.MODE_EXACT)
    b = song_amd.Batch(c1["ix"], CONFIGS[name], c1["q"].shape[0], mode=song_amd.MODE_EXACT)
    got = _batch_run(b, c1["q"], c1["k_d"], c1["pq_d"])
    counts = b.counts()
    print(f"[batch] {name}: per tier {counts[:-1].tolist()} re-runs {int(counts[-1])} grids "
          f"{[b.grid(t) for t in range(len(CONFIGS[name]) + 1)]}", flush=True)
    assert counts[:-1].sum() == c1["q"].shape[0]
    _same(got, want, name)
    assert int(got[2].max()) == 0 or (got[2] & song_amd.FLAG_STEP_CAP).equal(got[2])
    # the golden queries (pq 100, k 100) equal the reference CPU search
    ng = c1["ng"]
    ids = got[0][:ng].cpu().numpy().astype(np.int64)
    assert np.array_equal(ids, c1["ref_ids"][:, :100])
    del torch


def test_serial_and_parallel_tiers_agree(c1):
    want = _plain(c1["ix"], c1["q"], c1["k"], c1["pq"], song_amd.MODE_EXACT)
    for par in (False, True):
        b = song_amd.Batch(c1["ix"], [tier(64), tier(128, waves_per_group=2), tier(512)], c1["q"].shape[0],
                           mode=song_amd.MODE_EXACT, parallel_tiers=par)
        _same(_batch_run(b, c1["q"], c1["k_d"], c1["pq_d"]), want, f"parallel={par}")
        _same(_batch_run(b, c1["q"], c1["k_d"], c1["pq_d"], graph=True), want, f"graph parallel={par}")


def test_batch_overflow_reruns_on_ties():
    """Tie-heavy integer data with tiny visited tables: queries that overflow them are re-run
    by the retry kernel and the results equal the plain search's (and, at the golden pq,
    the reference's)."""
    import torch

    meta, base, query, ref_ids, graph = golden_cases.load("tie_l2_3k")
    ix = song_amd.Index(base, graph, "l2")
    dev = _dev()
    q = torch.from_numpy(query).to(dev)
    nq = q.shape[0]
    for pq in (meta["pq"], 400):
        b = song_amd.Batch(ix, [tier(pq, hash_log2=8, slack=0)], nq, mode=song_amd.MODE_EXACT, spill_log2=10,
                           retry=tier(pq, hash_log2=12, slack=2048))
        ids = torch.empty((nq, pq), dtype=torch.int32, device=dev)
        fl = torch.empty(nq, dtype=torch.int32, device=dev)
        b.run(q, ids, fl, pq=pq, k=pq)
        counts = b.counts()
        print(f"[batch] ties pq={pq}: re-runs {int(counts[-1])} of {nq}", flush=True)
        assert int(fl.abs().max()) == 0
        want = torch.empty_like(ids)
        ix.search_device(q, pq, pq, want, mode=song_amd.MODE_EXACT)
        assert torch.equal(ids, want)
        if pq == meta["pq"]:
            assert np.array_equal(ids.cpu().numpy().astype(np.int64), ref_ids)
        else:
            assert counts[-1] > 0  # the visited tables overflowed: the re-run path ran


def test_batch_graph_replays_follow_inputs(c1):
    """Captured once, replayed with new query contents and eager work in between."""
    import torch

    nq = 1000
    q = c1["q"][:nq].clone()
    k_d, pq_d = c1["k_d"][:nq].clone(), c1["pq_d"][:nq].clone()
    b = song_amd.Batch(c1["ix"], [tier(128), tier(512, waves_per_group=2)], nq, mode=song_amd.MODE_EXACT)
    ids = torch.empty((nq, 100), dtype=torch.int32, device=q.device)
    ds = torch.empty((nq, 100), dtype=torch.float32, device=q.device)
    fl = torch.empty(nq, dtype=torch.int32, device=q.device)
    side = torch.cuda.Stream()
    sources = [c1["q"][:nq], c1["q"][nq:2 * nq], c1["q"][:nq].flip(0)]
    for rep, src in enumerate(sources):
        q.copy_(src)
        ids.fill_(-7)
        ds.zero_()
        x = torch.randn(256, 256, device=q.device)
        (x @ x).sum().item()  # eager work between replays
        main = torch.cuda.current_stream()
        side.wait_stream(main)
        b.run(q, ids, fl, pq_arr=pq_d, k_arr=k_d, dists=ds, graph=True, stream=side.cuda_stream)
        main.wait_stream(side)
        torch.cuda.synchronize()
        got = (ids.clone(), ds.clone(), fl.clone())
        want = _batch_run(b, q, k_d, pq_d)
        _same(got, want, f"replay {rep}")


def test_batch_bad_budgets_flagged(c1):
    import torch

    nq = 64
    q = c1["q"][:nq]
    pq = torch.full((nq,), 64, dtype=torch.int32, device=q.device)
    pq[3], pq[10] = 0, 600
    k = torch.full((nq,), 10, dtype=torch.int32, device=q.device)
    b = song_amd.Batch(c1["ix"], [tier(512)], nq, mode=song_amd.MODE_EXACT)
    ids, _, fl = _batch_run(b, q, k, pq, k_ld=10)
    fl = fl.cpu().numpy()
    assert fl[3] == song_amd.FLAG_BAD_PARAMS and fl[10] == song_amd.FLAG_BAD_PARAMS
    assert (ids[3] == -1).all() and (ids[10] == -1).all()
    ok = np.setdiff1d(np.arange(nq), [3, 10])
    assert (fl[ok] == 0).all()
    plain_ids, _, _ = _plain(c1["ix"], q[ok].contiguous(), np.full(len(ok), 10), np.full(len(ok), 64),
                             song_amd.MODE_EXACT, k_ld=10)
    assert torch.equal(ids[torch.from_numpy(ok).to(q.device)], plain_ids)


def test_batch_entries_steps_and_bf16(c1):
    """Per-query start lists (padded with -1), expansion caps and BF16 rows + re-rank."""
    import torch

    ix = c1["ix"]
    ix.add_storage(song_amd.STORE_BF16)
    nq = 1500
    q = c1["q"][:nq].contiguous()
    k, pq = c1["k"][:nq], c1["pq"][:nq]
    g = torch.Generator().manual_seed(5)
    entry = torch.randint(0, c1["base"].shape[0], (nq, 4), generator=g, dtype=torch.int32)
    entry[::3, 2:] = -1  # shorter lists
    entry[::7, :] = -1   # no entry at all: vertex 0
    entry = entry.to(q.device)
    steps = torch.randint(5, 200, (nq,), generator=g, dtype=torch.int32).to(q.device)
    for storage, mode in ((song_amd.STORE_F32, song_amd.MODE_EXACT), (song_amd.STORE_BF16, song_amd.MODE_FP32)):
        want = _plain(ix, q, k, pq, mode, storage=storage, entry=entry, steps=steps)
        b = song_amd.Batch(ix, [tier(64), tier(256, waves_per_group=2), tier(512)], nq, mode=mode, storage=storage)
        got = _batch_run(b, q, c1["k_d"][:nq], c1["pq_d"][:nq], entry=entry, steps=steps)
        _same(got, want, f"storage {storage}")
        assert (got[2] & song_amd.FLAG_STEP_CAP).sum() > 0


def test_batch_per_tier_storage(c1):
    """A tier's own row storage: with BF16 batch storage, the middle tier reads FP32 rows; every
    query equals the plain search at its tier's storage (non-integer cosine data, so the two
    storages give different results and a mix-up shows)."""
    import torch

    ix = c1["ix"]
    ix.add_storage(song_amd.STORE_BF16)
    nq = 1500
    q = c1["q"][:nq].contiguous()
    k, pq = c1["k"][:nq], c1["pq"][:nq]
    w_bf = _plain(ix, q, k, pq, song_amd.MODE_FP32, storage=song_amd.STORE_BF16)
    w_32 = _plain(ix, q, k, pq, song_amd.MODE_FP32, storage=song_amd.STORE_F32)
    mid = torch.from_numpy((pq > 64) & (pq <= 256)).to(q.device)
    assert bool(mid.any()) and bool((~mid).any())
    # the storages differ on these queries
    assert not (torch.equal(w_bf[0][mid], w_32[0][mid]) and torch.equal(w_bf[1][mid], w_32[1][mid]))
    want = tuple(torch.where(mid[:, None] if a.dim() == 2 else mid, a, b_) for a, b_ in zip(w_32, w_bf))
    b = song_amd.Batch(ix, [tier(64), tier(256, waves_per_group=2, storage=song_amd.STORE_F32), tier(512)], nq,
                       mode=song_amd.MODE_FP32, storage=song_amd.STORE_BF16)
    _same(_batch_run(b, q, c1["k_d"][:nq], c1["pq_d"][:nq]), want, "per-tier storage")
    _same(_batch_run(b, q, c1["k_d"][:nq], c1["pq_d"][:nq], graph=True), want, "per-tier storage, graph")


def test_autotuned_tiers_keep_results(c1):
    import torch

    nq = 2000
    table = autotune.tune(c1["ix"], c1["q"][:nq], c1["pq_d"][:nq], c1["k_d"][:nq], mode=song_amd.MODE_EXACT, reps=1,
                          max_queries_per_tier=400, k_ld=100, verbose=True)
    assert [t["pq_max"] for t in table["tiers"]] == sorted(set(c1["pq"][:nq].tolist()))
    for entry in table["log"]:
        assert entry["ms"] > 0 and len(entry["trials"]) > 10
    b = song_amd.Batch(c1["ix"], table["tiers"], nq, mode=song_amd.MODE_EXACT)
    got = _batch_run(b, c1["q"][:nq], c1["k_d"][:nq], c1["pq_d"][:nq])
    want = _plain(c1["ix"], c1["q"][:nq], c1["k"][:nq], c1["pq"][:nq], song_amd.MODE_EXACT)
    _same(got, want, "autotuned")
    del torch


def test_batch_fp32_integer_mixture():
    """Integer-valued L2 data in FP32 mode (exact there) through a multi-tier batch."""
    import torch

    base, query = datasets.mixture_l2(n_base=20000, n_query=1200, dim=64, n_centers=50, seed=21)
    graph = song_amd.build_graph(base, "l2", threads=8)
    ix = song_amd.Index(base, graph, "l2")
    dev = _dev()
    q = torch.from_numpy(query).to(dev)
    assert song_amd.resolve_mode(ix, q) == song_amd.MODE_FP32
    k, pq = datasets.hetero_budgets(len(query), seed=3)
    want = _plain(ix, q, k, pq, song_amd.MODE_FP32)
    b = song_amd.Batch(ix, [tier(64), tier(512, waves_per_group=4)], len(query), mode=song_amd.MODE_FP32)
    got = _batch_run(b, q, torch.from_numpy(k).to(dev), torch.from_numpy(pq).to(dev))
    _same(got, want, "mixture fp32")


# ---------------------------------------------------------------- inner product (the 2M x 96 generator)

@pytest.fixture(scope="module")
def ip96():
    """Unit-norm Gaussian rows, 96-d, inner product (the batching workload's generator) on the
    golden ip_10k case: its graph, its 300 reference-searched queries at pq 64 (k = 64), and
    1,700 more queries of the same generator with uneven (k, pq)."""
    import torch

    meta, base, _, ref_ids, graph = golden_cases.load("ip_10k")
    graph = song_amd.build_graph(base, "ip", threads=8) if graph is None else graph
    assert golden_cases.sha256(graph) == meta["graph_sha256"]
    _, query = datasets.normalized_ip(n_base=meta["kwargs"]["n_base"], n_query=2000, dim=96, seed=4)
    k, pq = datasets.hetero_budgets(len(query), seed=4, ks=(1, 10, 100), pqs=(32, 64, 128, 256, 512))
    ng = meta["n_query"]
    pq[:ng], k[:ng] = meta["pq"], meta["pq"]
    ix = song_amd.Index(base, graph, "ip")
    dev = _dev()
    q = torch.from_numpy(np.ascontiguousarray(query)).to(dev)
    return dict(meta=meta, base=base, ix=ix, q=q, k=k, pq=pq, ref_ids=ref_ids, ng=ng,
                k_d=torch.from_numpy(k).to(dev), pq_d=torch.from_numpy(pq).to(dev))


IP_CONFIGS = {
    "tier_per_pq": [tier(p) for p in (32, 64, 128, 256, 512)],
    "two_tiers_2waves_spill": [tier(64, waves_per_group=2, hash_log2=9),
                               tier(512, waves_per_group=4, slack=8, groups_per_cu=1)],
}


@pytest.mark.parametrize("name", sorted(IP_CONFIGS))
def test_ip_batch_equals_plain_and_reference(ip96, name):
    """search_persistent on inner product (exact mode): every query equals the plain search at its
    (k, pq), and the golden queries equal the reference song.cpu's ids."""
    print(f"[batch-ip] config {name}", flush=True)
    c = ip96
    want = _plain(c["ix"], c["q"], c["k"], c["pq"], song_amd.MODE_EXACT)
    b = song_amd.Batch(c["ix"], IP_CONFIGS[name], c["q"].shape[0], mode=song_amd.MODE_EXACT)
    for graph in (False, True):
        got = _batch_run(b, c["q"], c["k_d"], c["pq_d"], graph=graph)
        _same(got, want, f"{name} graph={graph}")
    ng, pq = c["ng"], c["meta"]["pq"]
    ids = got[0][:ng, :pq].cpu().numpy().astype(np.int64)
    assert np.array_equal(ids, c["ref_ids"][:, :pq])
    b.close()


def test_ip_batch_fp32_mode_equals_plain_fp32(ip96):
    """The FP32-accumulation persistent kernel on inner product equals the FP32 plain search."""
    c = ip96
    want = _plain(c["ix"], c["q"], c["k"], c["pq"], song_amd.MODE_FP32)
    b = song_amd.Batch(c["ix"], IP_CONFIGS["tier_per_pq"], c["q"].shape[0], mode=song_amd.MODE_FP32)
    _same(_batch_run(b, c["q"], c["k_d"], c["pq_d"]), want, "ip fp32")
    b.close()


# ---------------------------------------------------------------- several queries per wave

GROUP_CONFIGS = {
    "qpw4": [tier(p, queries_per_wave=4) for p in PQS],
    "qpw2_two_tiers": [tier(100, queries_per_wave=2, shape=1, waves_per_group=2),
                       tier(512, queries_per_wave=4, slack=8, hash_log2=9)],
    "qpw4_mixed_with_one_query_tier": [tier(128, queries_per_wave=4, shape=3, groups_per_cu=1),
                                       tier(512, waves_per_group=2)],
}


@pytest.mark.parametrize("name", sorted(GROUP_CONFIGS))
def test_several_queries_per_wave_equal_plain_and_reference(c1, name):
    """Sub-groups of 16 / 32 lanes each search their own query (exact mode, cosine 16-d): every
    query equals the plain search at its (k, pq) -- ids, distances, flags -- through graph
    replays, and the golden queries equal the reference's ids."""
    print(f"[batch-group] config {name}", flush=True)
    want = _plain(c1["ix"], c1["q"], c1["k"], c1["pq"], song_amd.MODE_EXACT)
    b = song_amd.Batch(c1["ix"], GROUP_CONFIGS[name], c1["q"].shape[0], mode=song_amd.MODE_EXACT)
    for graph in (False, True):
        got = _batch_run(b, c1["q"], c1["k_d"], c1["pq_d"], graph=graph)
        _same(got, want, f"{name} graph={graph}")
    ng, pq = c1["ng"], c1["meta"]["pq"]
    assert np.array_equal(got[0][:ng, :pq].cpu().numpy().astype(np.int64), c1["ref_ids"][:, :pq])
    b.close()


@pytest.mark.parametrize("qpw", [2, 4])
def test_several_queries_per_wave_on_ip_and_fp32(ip96, qpw):
    c = ip96
    for mode in (song_amd.MODE_EXACT, song_amd.MODE_FP32):
        want = _plain(c["ix"], c["q"], c["k"], c["pq"], mode)
        b = song_amd.Batch(c["ix"], [tier(p, queries_per_wave=qpw) for p in (32, 64, 128, 256, 512)],
                           c["q"].shape[0], mode=mode)
        _same(_batch_run(b, c["q"], c["k_d"], c["pq_d"]), want, f"ip qpw={qpw} mode={mode}")
        b.close()


def test_several_queries_per_wave_with_entries_caps_and_spills():
    """Learned-form inputs (several start vertices per query, expansion caps) and a visited
    table small enough to spill, on L2 integer data in FP32 mode and BF16 rows."""
    import torch

    meta, base, query, _, graph = golden_cases.load("tie_l2_3k")
    ix = song_amd.Index(base, graph, "l2")
    ix.add_storage(song_amd.STORE_BF16)
    dev = _dev()
    q = torch.from_numpy(np.ascontiguousarray(query)).to(dev)
    nq = q.shape[0]
    rng = np.random.default_rng(9)
    entry = torch.from_numpy(rng.integers(0, base.shape[0], size=(nq, 8)).astype(np.int32)).to(dev)
    entry[::5, 3:] = -1
    steps = torch.from_numpy(rng.integers(5, 200, size=nq).astype(np.int32)).to(dev)
    k_h = np.full(nq, 10, dtype=np.int32)
    pq_h = np.where(np.arange(nq) % 3 == 0, 40, 100).astype(np.int32)
    k_d, pq_d = torch.from_numpy(k_h).to(dev), torch.from_numpy(pq_h).to(dev)
    for storage in (song_amd.STORE_F32, song_amd.STORE_BF16):
        want = _plain(ix, q, k_h, pq_h, song_amd.MODE_FP32, storage=storage, entry=entry, steps=steps)
        for tiers in ([tier(40, queries_per_wave=4), tier(100, queries_per_wave=4, hash_log2=8)],
                      [tier(100, queries_per_wave=2, hash_log2=8, slack=0)]):
            b = song_amd.Batch(ix, tiers, nq, mode=song_amd.MODE_FP32, storage=storage)
            got = _batch_run(b, q, k_d, pq_d, entry=entry, steps=steps)
            _same(got, want, f"storage={storage} tiers={tiers}")
            b.close()


# ---------------------------------------------------------------- two waves per query

def _mix20k():
    import torch

    meta, base, query, ref_ids, graph = golden_cases.load("mix_l2_20k")
    graph = song_amd.build_graph(base, "l2", threads=8) if graph is None else graph
    ix = song_amd.Index(base, graph, "l2")
    return meta, base, ix, torch.from_numpy(np.ascontiguousarray(query)).to(_dev()), ref_ids


@pytest.mark.parametrize("pq", [100, 200])
def test_two_waves_per_query_equal_plain_and_reference(pq):
    """A helper wave per query computes the predicted next expansion's distances ahead of the
    search wave (FP32 mode, integer L2 data where FP32 is exact): ids, distances and flags equal
    the plain search's, the golden queries equal the reference's ids at the golden pq, and the
    search waves do take their distances from the helper's cache (stats column 7)."""
    import torch

    meta, base, ix, q, ref_ids = _mix20k()
    nq = q.shape[0]
    k = pq if pq == meta["pq"] else 20  # the full list at the golden pq
    k_h = np.full(nq, k, dtype=np.int32)
    pq_h = np.full(nq, pq, dtype=np.int32)
    want = _plain(ix, q, k_h, pq_h, song_amd.MODE_FP32)
    for tiers in ([tier(pq, waves_per_query=2)], [tier(pq, waves_per_query=2, hash_log2=9, slack=0, groups_per_cu=1)]):
        b = song_amd.Batch(ix, tiers, nq, mode=song_amd.MODE_FP32)
        st = torch.zeros((nq, 8), dtype=torch.int32, device=q.device)
        for graph in (False, True):
            ids = torch.full((nq, 100), -7, dtype=torch.int32, device=q.device)
            ds = torch.zeros((nq, 100), dtype=torch.float32, device=q.device)
            fl = torch.full((nq,), -1, dtype=torch.int32, device=q.device)
            kd = torch.from_numpy(k_h).to(q.device)
            pd = torch.from_numpy(pq_h).to(q.device)
            if graph:
                side = torch.cuda.Stream()
                side.wait_stream(torch.cuda.current_stream())
                for _ in range(2):
                    b.run(q, ids, fl, pq_arr=pd, k_arr=kd, dists=ds, graph=True, stream=side.cuda_stream)
            else:
                b.run(q, ids, fl, pq_arr=pd, k_arr=kd, dists=ds, stats=st)
            torch.cuda.synchronize()
            _same((ids, ds, fl), want, f"duo pq={pq} {tiers} graph={graph}")
            if pq == meta["pq"]:
                assert np.array_equal(ids[:, :pq].cpu().numpy().astype(np.int64), ref_ids[:, :pq])
        hits = int(st[:, 7].sum())
        exps = int(st[:, 0].sum())
        print(f"[duo] pq {pq}: {hits} of {exps} expansions took their distances from the helper", flush=True)
        assert hits > 0
        b.close()


@pytest.mark.parametrize("mode", ["exact", "fp32"])
def test_best_list_in_global_memory_equals_plain(mode):
    """Tiers with the best list's front in each wave slot's global scratch (best_in_global 1) or in
    LDS (0): ids, distances and flags equal the plain search's at a budget where the front holds
    most of the list; with small tables and no slack (spills, overflow re-runs) and, in FP32
    mode, with two waves per query; eager runs and graph replays."""
    import torch

    meta, base, ix, q, ref_ids = _mix20k()
    m = song_amd.MODE_EXACT if mode == "exact" else song_amd.MODE_FP32
    nq = q.shape[0]
    pq, k = 500, 100
    k_h = np.full(nq, k, dtype=np.int32)
    pq_h = np.full(nq, pq, dtype=np.int32)
    want = _plain(ix, q, k_h, pq_h, m)
    variants = [[tier(pq, best_in_global=1)], [tier(pq, best_in_global=1, hash_log2=9, slack=0, groups_per_cu=1)],
                [tier(100, best_in_global=1), tier(pq, best_in_global=1, waves_per_group=2)],
                [tier(pq, best_in_global=0)], [tier(pq, best_in_global=0, hash_log2=9, slack=0, groups_per_cu=1)]]
    if m == song_amd.MODE_FP32:
        variants += [[tier(pq, best_in_global=1, waves_per_query=2)], [tier(pq, best_in_global=0, waves_per_query=2)]]
    kd = torch.from_numpy(k_h).to(q.device)
    pd = torch.from_numpy(pq_h).to(q.device)
    for tiers in variants:
        print(f"[best in global] {mode} {tiers}", flush=True)
        b = song_amd.Batch(ix, tiers, nq, mode=m)
        for graph in (False, True):
            ids = torch.full((nq, k), -7, dtype=torch.int32, device=q.device)
            ds = torch.zeros((nq, k), dtype=torch.float32, device=q.device)
            fl = torch.full((nq,), -1, dtype=torch.int32, device=q.device)
            if graph:
                side = torch.cuda.Stream()
                side.wait_stream(torch.cuda.current_stream())
                for _ in range(2):
                    b.run(q, ids, fl, pq_arr=pd, k_arr=kd, dists=ds, graph=True, stream=side.cuda_stream)
            else:
                b.run(q, ids, fl, pq_arr=pd, k_arr=kd, dists=ds)
            torch.cuda.synchronize()
            _same((ids, ds, fl), want, f"best in global {mode} {tiers} graph={graph}")
        b.close()


def test_two_waves_per_query_on_inner_product(ip96):
    c = ip96
    want = _plain(c["ix"], c["q"], c["k"], c["pq"], song_amd.MODE_FP32)
    b = song_amd.Batch(c["ix"], [tier(p, waves_per_query=2) for p in (32, 64, 128, 256, 512)], c["q"].shape[0],
                       mode=song_amd.MODE_FP32)
    _same(_batch_run(b, c["q"], c["k_d"], c["pq_d"]), want, "ip duo")
    b.close()
