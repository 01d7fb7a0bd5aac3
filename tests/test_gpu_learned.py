"""Learned form on the GPU: MFMA GEMM, MLP gradients and fused inference against the
float64 numpy oracle, training, and the ablation learned-vs-plain."""
import numpy as np
import pytest

import mlp_ref
import song_amd
from song_amd import datasets, groundtruth, learned

pytestmark = pytest.mark.gpu


def _bf16_tensor(a):
    import torch

    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    code = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)
    return torch.from_numpy(code.view(np.int16)).cuda(), mlp_ref.bf16(a32)


@pytest.mark.parametrize("M,N,K", [(70, 45, 37), (64, 64, 16), (1, 130, 300), (257, 3, 5)])
def test_gemm_matches_numpy(M, N, K):
    rng = np.random.default_rng(M * 1000 + N)
    a, ar = _bf16_tensor(rng.standard_normal((M, K)))
    b, br = _bf16_tensor(rng.standard_normal((K, N)))
    c = learned.gemm_bf16(a, b, K, 1, N, 1, M, N, K).cpu().numpy()
    ref = ar @ br
    assert np.allclose(c, ref, rtol=1e-4, atol=1e-4 * np.sqrt(K))
    # transposed operands through strides: A given as [K, M], B as [N, K]
    at, _ = _bf16_tensor(ar.T.copy())
    bt, _ = _bf16_tensor(br.T.copy())
    c2 = learned.gemm_bf16(at, bt, 1, M, 1, K, M, N, K).cpu().numpy()
    assert np.allclose(c2, ref, rtol=1e-4, atol=1e-4 * np.sqrt(K))


def _small_problem(seed=0, width_in=20, hidden=32, classes=50, B=64):
    import torch

    rng = np.random.default_rng(seed)
    m = learned.MLP(width_in, hidden, classes, max_batch=B, seed=seed)
    p = m.get_params() * 1.5  # move away from the initial scale
    m.set_params(p)
    x, xr = _bf16_tensor(np.pad(rng.standard_normal((B, width_in)), ((0, 0), (0, m.width - width_in))))
    labels = rng.integers(0, classes, B).astype(np.int32)
    stop = rng.standard_normal(B).astype(np.float32)
    params = mlp_ref.split(p, m.width, hidden, classes)
    return m, x, xr, torch.from_numpy(labels).cuda(), torch.from_numpy(stop).cuda(), params, labels, stop


def test_mlp_gradients_match_oracle():
    m, x, xr, lab_d, stop_d, params, labels, stop = _small_problem()
    w = 0.3
    xe, ms, g = m.loss_grad(x, lab_d, stop_d, w, want_grads=True)
    rxe, rms, rg = mlp_ref.loss_grad(params, xr, labels, stop, w, emulate_bf16=True)
    assert abs(xe - rxe) < 1e-3 * max(1, rxe) and abs(ms - rms) < 1e-3 * max(1, rms)
    got = mlp_ref.split(g, m.width, m.hidden, m.n_classes)
    for t, (a, b) in enumerate(zip(got, rg)):
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert err < 2e-2, (t, err)


# (batch, classes, hidden): 70 rows leave the last 32-row group partial; 300 classes make two
# 256-column chunks and a padded interleaved weight copy; 64 hidden units are four k-steps
SHAPES = [(64, 50, 32), (70, 300, 64)]


@pytest.mark.parametrize("B,classes,hidden", SHAPES)
def test_fused_inference_matches_oracle(B, classes, hidden):
    m, x, xr, _, _, params, _, _ = _small_problem(seed=1, hidden=hidden, classes=classes, B=B)
    top2, conf, stop, logits = m.infer(x, logits=True)
    _, _, Z = mlp_ref.forward(params, xr, emulate_bf16=True)
    order, rconf, rstop = mlp_ref.top2_conf(Z)
    assert np.allclose(logits.cpu().numpy(), Z, rtol=1e-3, atol=1e-3)
    srt = np.sort(Z[:, :-1], axis=1)
    clear = (srt[:, -1] - srt[:, -2] > 1e-2) & (srt[:, -2] - srt[:, -3] > 1e-2)
    assert np.array_equal(top2.cpu().numpy()[clear], order[clear])
    assert np.allclose(conf.cpu().numpy(), rconf, atol=2e-3)
    assert np.allclose(stop.cpu().numpy(), rstop, rtol=1e-3, atol=1e-3)


@pytest.mark.parametrize("B,classes,hidden", SHAPES)
def test_fused_inference_neighbourhood_confidence(B, classes, hidden):
    import torch

    m, x, xr, _, _, params, _, _ = _small_problem(seed=2, hidden=hidden, classes=classes, B=B)
    C, E = m.n_classes, 3
    nbr = np.stack([(np.arange(C) + j * 7) % C for j in range(E)], 1).astype(np.int32)
    top2, conf, _ = m.infer(x, nbr=torch.from_numpy(nbr).cuda())
    _, _, Z = mlp_ref.forward(params, xr, emulate_bf16=True)
    logits = Z[:, :-1]
    p = np.exp(logits - logits.max(1, keepdims=True))
    p /= p.sum(1, keepdims=True)
    best = top2[:, 0].cpu().numpy()
    ref = np.array([p[i, nbr[best[i]]].sum() for i in range(len(best))])
    assert np.allclose(conf.cpu().numpy(), np.minimum(ref, 1.0), atol=2e-3)


def test_training_learns_nearest_centre_classes():
    import torch

    rng = np.random.default_rng(7)
    classes, dim, n = 64, 16, 8192
    centres = rng.standard_normal((classes, dim)) * 3
    lab = rng.integers(0, classes, n)
    xs = centres[lab] + rng.standard_normal((n, dim)) * 0.3
    m = learned.MLP(dim, 64, classes, max_batch=512, seed=3)
    xpad = np.pad(xs, ((0, 0), (0, m.width - dim)))
    x, _ = _bf16_tensor(xpad)
    ld = torch.from_numpy(lab.astype(np.int32)).cuda()
    st = torch.zeros(n, dtype=torch.float32, device="cuda")
    first, _, _ = m.loss_grad(x[:512], ld[:512], st[:512])
    for ep in range(6):
        for b in range(0, n, 512):
            m.train_step(x[b:b + 512], ld[b:b + 512], st[b:b + 512], 3e-3)
    last, _, _ = m.loss_grad(x[:512], ld[:512], st[:512])
    top2, conf, _ = m.infer(x)
    acc = (top2[:, 0].cpu().numpy() == lab).mean()
    print(f"[learned] xent {first:.3f} -> {last:.3f}, top1 accuracy {acc:.3f}", flush=True)
    assert last < 0.5 * first and acc > 0.9


# 32-d rows, and 128-d rows of the 1M x 128 workload's generator (integer-valued Gaussian mixture)
@pytest.fixture(scope="module", params=[32, 128], ids=["dim32", "dim128"])
def mix50k(request):
    dim = request.param
    centres = 100 if dim == 32 else 200
    print(f"[learned] mixture 50k x {dim}, {centres} centres", flush=True)
    base, _ = datasets.mixture_l2(n_base=50000, n_query=1, dim=dim, n_centers=centres, seed=12)
    train = datasets.mixture_l2_queries(20000, dim=dim, n_centers=centres, seed=12, stream=1)
    test = datasets.mixture_l2_queries(1000, dim=dim, n_centers=centres, seed=12, stream=2)
    graph = song_amd.build_graph(base, "l2", threads=8)
    ix = song_amd.Index(base, graph, "l2")
    return base, graph, ix, train, test


def test_learned_entries_beat_vertex_zero(mix50k):
    """Ablation: same index, queries and pq; learned start vertices vs vertex 0."""
    import torch

    base, graph, ix, train, test = mix50k
    ls = learned.LearnedSearch(ix, base, n_pivots=256, hidden=64, seed=0)
    log = ls.fit(train, pq_label=64, epochs=6, batch=512, verbose=False)
    print(f"[learned] fit {log}", flush=True)
    gt_i, gt_d = groundtruth.knn_fp64(base, test, 10, "l2")
    q = torch.from_numpy(test).cuda()
    k, pq = 10, 24
    plain = torch.empty((q.shape[0], k), dtype=torch.int32, device="cuda")
    st_p = torch.empty((q.shape[0], 8), dtype=torch.int32, device="cuda")
    ix.search_device(q, pq, k, plain, stats=st_p)
    r_plain = groundtruth.recall_at_k(plain.cpu().numpy().astype(np.int64), gt_d, 10, base, test)
    lrn = torch.empty_like(plain)
    st_l = torch.empty_like(st_p)
    plan = ls.search(q, pq, k, lrn, conf_threshold=0.0, step_scale=1e6, stats=st_l)  # entries only, no cap
    r_learn = groundtruth.recall_at_k(lrn.cpu().numpy().astype(np.int64), gt_d, 10, base, test)
    capped = torch.empty_like(plain)
    st_c = torch.empty_like(st_p)
    ls.search(q, pq, k, capped, conf_threshold=0.0, step_scale=1.5, stats=st_c)
    r_cap = groundtruth.recall_at_k(capped.cpu().numpy().astype(np.int64), gt_d, 10, base, test)
    e_p, e_l, e_c = (s[:, 0].float().mean().item() for s in (st_p, st_l, st_c))
    print(f"[learned] pq={pq} recall plain={r_plain:.4f} learned={r_learn:.4f} learned+stop={r_cap:.4f} "
          f"expansions {e_p:.1f} / {e_l:.1f} / {e_c:.1f}", flush=True)
    assert r_learn > r_plain
    assert e_c <= e_l
    # the confidence split: every query is served, fallback ones by the plain search
    entry, steps, conf_ids, fb_ids = ls.plan(q, conf_threshold=0.9)
    assert conf_ids.numel() + fb_ids.numel() == q.shape[0]


def test_learned_batch_equals_split_search_and_graph_replays(mix50k):
    """LearnedBatch (one batched search, per-query pq / entries / caps made on the GPU)
    equals the confident / fallback split run as two plain calls with the same plan; the
    captured pipeline replays to the same ids with eager work between replays."""
    import torch

    base, graph, ix, train, test = mix50k
    ls = learned.LearnedSearch(ix, base, n_pivots=256, hidden=64, seed=1)
    ls.fit(train, pq_label=64, epochs=6, batch=512, verbose=False)
    q = torch.from_numpy(test).cuda()
    nq, k = q.shape[0], 10
    _, conf, _ = ls.predict(q)
    thr = float(conf.median().item())  # both groups non-empty
    lb = learned.LearnedBatch(ls, nq, pq_learned=24, pq_fallback=64, k=k, conf_threshold=thr, step_scale=1.5,
                              mode=song_amd.MODE_FP32)
    pq_arr, entry, caps, ok = lb.run(q)
    torch.cuda.synchronize()
    got = lb.ids.clone()
    conf_idx = torch.nonzero(ok).flatten().to(torch.int32)
    fb_idx = torch.nonzero(~ok).flatten().to(torch.int32)
    print(f"[learned] batch: {conf_idx.numel()} confident, {fb_idx.numel()} fallback", flush=True)
    assert conf_idx.numel() > 0 and fb_idx.numel() > 0
    assert bool((pq_arr[ok] == 24).all()) and bool((pq_arr[~ok] == 64).all())
    assert bool((entry[~ok] == -1).all()) and bool((entry[ok] >= 0).all())
    ref = torch.full_like(got, -3)
    ix.search_device(q, 64, k, ref, subset=fb_idx, mode=song_amd.MODE_FP32, retry=0)
    ix.search_device(q, 24, k, ref, entry=entry, steps=caps, subset=conf_idx, mode=song_amd.MODE_FP32, retry=0)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
    assert int(lb.flags.abs().max()) in (0, song_amd.FLAG_STEP_CAP)

    lb.capture(q)
    for rep in range(3):
        lb.ids.fill_(-5)
        x = torch.randn(512, 512, device=q.device)
        (x @ x).sum().item()
        lb.replay()
        torch.cuda.synchronize()
        assert torch.equal(lb.ids, got), f"replay {rep}"


def test_fallback_cap_keeps_settled_queries_equal_to_plain(mix50k):
    """The fallback expansion cap calibrated on training queries: the cap is the quantile of the
    plain search's settle steps over the low-confidence training queries; a fallback query is
    flagged as capped when the uncapped plain search expands past the cap, and
    every fallback query whose top-k settled within the cap returns the plain search's top-k.
    The batched pipeline equals the split search under the same cap."""
    import torch

    base, graph, ix, train, test = mix50k
    ls = learned.LearnedSearch(ix, base, n_pivots=256, hidden=64, seed=2)
    ls.fit(train, pq_label=64, epochs=6, batch=512, verbose=False)
    q = torch.from_numpy(test).cuda()
    nq, k, pqf = q.shape[0], 10, 64
    _, conf, _ = ls.predict(q)
    thr = float(conf.quantile(0.7).item())  # a large fallback group
    cal = ls.calibrate_fallback_cap(train, pqf, k=k, conf_threshold=thr, quantile=0.9, max_queries=5000)
    print(f"[learned] fallback cap {cal}", flush=True)
    # the calibration, restated: plain search of the same training queries, settle quantile
    tq = torch.from_numpy(np.ascontiguousarray(train[:5000])).cuda()
    _, tconf, _ = ls.predict(tq)
    low = torch.nonzero(tconf < thr).flatten()
    assert low.numel() >= 500 and cal["low_confidence_only"]
    tq = tq[low].contiguous()
    tst = torch.empty((tq.shape[0], 8), dtype=torch.int32, device="cuda")
    tids = torch.empty((tq.shape[0], k), dtype=torch.int32, device="cuda")
    ix.search_device(tq, pqf, k, tids, stats=tst, mode=song_amd.MODE_FP32)
    torch.cuda.synchronize()
    settle = tst[:, 4].cpu().numpy().astype(np.int64)
    assert cal["cap"] == int(np.ceil(np.quantile(settle, 0.9)))
    cap = cal["cap"]

    lb = learned.LearnedBatch(ls, nq, pq_learned=24, pq_fallback=pqf, k=k, conf_threshold=thr, step_scale=1.5,
                              mode=song_amd.MODE_FP32, fallback_cap=cap)
    pq_arr, entry, caps, ok = lb.run(q)
    torch.cuda.synchronize()
    assert bool((caps[~ok] == cap).all())
    fb_idx = torch.nonzero(~ok).flatten().to(torch.int32)
    ref = torch.empty((nq, k), dtype=torch.int32, device="cuda")
    st = torch.zeros((nq, 8), dtype=torch.int32, device="cuda")
    ix.search_device(q, pqf, k, ref, stats=st, subset=fb_idx, mode=song_amd.MODE_FP32, retry=0)
    torch.cuda.synchronize()
    fb = fb_idx.long()
    exp_plain, settle_plain = st[fb, 0], st[fb, 4]
    capped = (lb.flags[fb] & song_amd.FLAG_STEP_CAP) != 0
    print(f"[learned] {fb.numel()} fallback queries, {int(capped.sum())} capped at {cap}", flush=True)
    assert int(capped.sum()) > 0
    # flagged: the plain search expanded past the cap, or stopped right at it (the cap is tested
    # before the stop rule)
    assert bool((capped >= (exp_plain > cap)).all()) and bool((capped <= (exp_plain >= cap)).all())
    settled = settle_plain <= cap
    assert torch.equal(lb.ids[fb][settled], ref[fb][settled])
    # the eager split search under the same cap gives the same ids
    split = torch.full_like(lb.ids, -3)
    ls.search(q, 24, k, split, conf_threshold=thr, step_scale=1.5, pq_fallback=pqf, fallback_cap=cap)
    torch.cuda.synchronize()
    assert torch.equal(split, lb.ids)


def test_plan_kernel_matches_its_restatement():
    """song_mlp_plan against the tensor formula it replaces: confident = conf >= threshold with a
    valid best pivot; entries from the best pivot's neighbourhood, caps from the stop output
    (integer stop values, so 2^stop is exact in both), fallback rows -1 / the fallback cap."""
    import ctypes
    import torch

    dev = torch.device("cuda")
    g = torch.Generator(device="cuda").manual_seed(3)
    B, C, E = 3000, 500, 8
    top2 = torch.randint(0, C, (B, 2), device=dev, dtype=torch.int32, generator=g)
    top2[::97, 0] = 0x7FFFFFFF  # no valid best pivot (non-finite logits)
    top2[::89, 0] = -1
    conf = torch.rand(B, device=dev, generator=g)
    conf[::13] = 0.9  # exactly at the threshold: confident
    stop = torch.randint(0, 12, (B,), device=dev, generator=g).float()
    nbr = torch.randint(0, C, (C, E), device=dev, dtype=torch.int32, generator=g)
    pv = torch.randint(0, 10 ** 6, (C,), device=dev, dtype=torch.int32, generator=g)
    L = learned._mlp_lib()
    for cap, scale in ((0, 1.5), (700, 2.5)):
        pq = torch.empty(B, dtype=torch.int32, device=dev)
        entry = torch.empty((B, E), dtype=torch.int32, device=dev)
        caps = torch.empty(B, dtype=torch.int32, device=dev)
        ok = torch.empty(B, dtype=torch.bool, device=dev)
        dp = learned._dp
        learned._mcheck(L.song_mlp_plan(dp(top2), dp(conf), dp(stop), B, dp(nbr), E, dp(pv), C, ctypes.c_float(0.9),
                                        ctypes.c_float(scale), ctypes.c_float(8.0), cap, 192, 768, dp(pq), dp(entry),
                                        dp(caps), dp(ok), None))
        torch.cuda.synchronize()
        best = top2[:, 0].long()
        valid = (best >= 0) & (best < C)
        want_ok = (conf >= 0.9) & valid
        safe = torch.where(valid, best, torch.zeros_like(best))
        want_entry = torch.where(want_ok[:, None], pv[nbr[safe].long()], torch.full_like(entry, -1))
        want_caps = torch.where(want_ok, (scale * (torch.exp2(stop) - 1.0) + 8.0).clamp(1, 1e6).to(torch.int32),
                                torch.full_like(caps, cap if cap else 1000000))
        want_pq = torch.where(want_ok, torch.full_like(pq, 192), torch.full_like(pq, 768))
        assert torch.equal(ok, want_ok) and bool(want_ok.any()) and bool((~want_ok).any())
        assert torch.equal(entry, want_entry)
        assert torch.equal(caps, want_caps)
        assert torch.equal(pq, want_pq)
