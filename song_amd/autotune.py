"""Autotuner of the batching form: per-tier launch knobs picked by measurement.

For one index, storage, mode and workload (device queries with per-query k and pq), the
tuner splits the queries into tiers by pq and tunes each tier on its own queries by
coordinate descent over the knobs of a tier (song_tier in include/song_amd.h):

  waves_per_group  queries in flight per workgroup (1, 2, 4)
  hash_log2        visited filter size in LDS (smaller: more waves fit, more spills)
  shape            lanes per candidate row (the sub-wave width of one distance)
  slack            heap depth beyond pq (tie room; overflows cost a re-run)
  groups_per_cu    persistent grid size (0 = all that fit; a smaller grid must be 3% faster)
  visited_bits     visited-table slot width (0 = compact 16-bit slots where the ids fit, 32 = 32-bit:
                   fewer instructions per probe, twice the LDS per slot)
  storage          with low-precision batch storage: the tier's rows (-1 = the batch's, or FP32 rows,
                   whose decode-free distances shorten a latency-bound tier); taken only where the
                   results stay identical, e.g. BF16 rows of integer data
  queries_per_wave 1, or 2 / 4 queries per 64-lane wave on sub-groups of 32 / 16 lanes (same results)
  bloom_k          (recall-judged tiers only, see approx_ok) a blocked Bloom filter in place of the
                   exact visited table: less LDS per wave, approximate visited set
  multiprobe       (recall-judged tiers with several queries per wave) two expansions per step
  bag_in_lds       one query per wave on float32 rows: the queue's overflow bag in LDS (same results)
  waves_per_query  1, or 2: a helper wave per query computes the predicted next expansion's distances
                   ahead of the search wave (FP32 mode, float32 rows; same results)
  best_in_global   the best list's front in the slot's global scratch (1), in LDS (0), or by size (-1):
                   global frees pq keys of LDS per wave for throughput-bound tiers; a tier of few
                   queries is faster with it in LDS (same results)

Each candidate runs the tier's queries through song_batch_run and is timed with events
(median of `reps`); a candidate is accepted only if every result equals the default
configuration's (ids, distances and flags), so tuning never changes results.  The
table is a JSON-able dict that `Batch(index, table["tiers"], ...)` takes directly.
"""
from __future__ import annotations

import json
import time

import numpy as np

from . import MODE_EXACT, STORE_F32, Batch, SongError, tier

SHAPES = (0, 1, 2, 3, 4)


def default_tiers(pqs):
    """One tier per distinct pq value (ascending)."""
    return [tier(int(p)) for p in sorted(set(int(x) for x in pqs))]


def _time_run(batch, q, ids, flags, pq_arr, k_arr, dists, reps, entry=None, steps=None, min_ms=30.0, max_reps=15):
    """Median of at least `reps` timed runs, more (up to max_reps) for short runs, so that a tier of
    a few milliseconds is timed over about min_ms."""
    import torch

    batch.run(q, ids, flags, pq_arr=pq_arr, k_arr=k_arr, dists=dists, entry=entry, steps=steps)  # warm-up
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    batch.run(q, ids, flags, pq_arr=pq_arr, k_arr=k_arr, dists=dists, entry=entry, steps=steps)
    e1.record()
    e1.synchronize()
    first = e0.elapsed_time(e1)
    reps = int(min(max_reps, max(reps, np.ceil(min_ms / max(first, 1e-3)))))
    ts = [first]
    for _ in range(reps - 1):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        batch.run(q, ids, flags, pq_arr=pq_arr, k_arr=k_arr, dists=dists, entry=entry, steps=steps)
        e1.record()
        e1.synchronize()
        ts.append(e0.elapsed_time(e1))
    return float(np.median(ts))


def _neighbours(value, choices):
    return [c for c in choices if c != value]


def tune(index, queries, pq_arr, k_arr, tiers=None, mode=MODE_EXACT, storage=STORE_F32, rerank=1, reps=3,
         max_queries_per_tier=20000, k_ld=None, entry=None, steps=None, verbose=False, freeze=(), approx_ok=None):
    """Tune every tier on its own queries; returns the table (dict).  freeze: knob names kept at
    the given tiers' values (e.g. "storage", where the rows are part of what is measured).

    approx_ok: {tier index: judge} for tiers whose results are judged by recall rather than
    equality.  Those tiers also try the Bloom filter in place of the exact visited table
    (bloom_k -1: the reference's auto_tune_bloom sizing, and its size knob hash_log2); a
    candidate that changes the results is accepted only if judge(ids, query_rows) is true, where
    ids are the tier's results (int32 cuda [n, k_ld]) and query_rows the tier's rows in
    `queries`; knobs that keep the results are accepted, as elsewhere, only on equal results."""
    import torch

    t_start = time.time()
    pq_h = pq_arr.cpu().numpy()
    tiers = [dict(t) for t in (tiers or default_tiers(pq_h))]
    bounds = [t["pq_max"] for t in tiers]
    k_ld = k_ld or int(k_arr.max().item())
    log = []
    lo = 0
    for ti, t in enumerate(tiers):
        sel = np.nonzero((pq_h > lo) & (pq_h <= t["pq_max"]))[0][:max_queries_per_tier]
        lo = t["pq_max"]
        if sel.size == 0:
            continue
        idx = torch.from_numpy(sel.astype(np.int64)).to(queries.device)
        q = queries[idx].contiguous()
        pa, ka = pq_arr[idx].contiguous(), k_arr[idx].contiguous()
        en = entry[idx].contiguous() if entry is not None else None
        st = steps[idx].contiguous() if steps is not None else None
        nq = q.shape[0]
        ids = torch.empty((nq, k_ld), dtype=torch.int32, device=q.device)
        dists = torch.empty((nq, k_ld), dtype=torch.float32, device=q.device)
        flags = torch.empty(nq, dtype=torch.int32, device=q.device)

        judge = (approx_ok or {}).get(ti)

        def measure(cfg, want=None, approx=False):
            try:
                b = Batch(index, [cfg], nq, mode=mode, storage=storage, rerank=rerank)
            except SongError as e:  # e.g. LDS beyond the limit
                return None, str(e)
            ids.fill_(-7)
            ms = _time_run(b, q, ids, flags, pa, ka, dists, reps, en, st)
            out = (ids.clone(), dists.clone(), flags.clone())
            b.close()
            if want is not None:
                same = all(torch.equal(a, w) for a, w in zip(out, want))
                if not same and not (approx and judge(out[0], sel)):
                    return None, "results differ"
            return ms, out

        best = dict(t)
        best_ms, ref = measure(best)
        if best_ms is None:
            raise SongError(f"tier {ti}: default configuration failed: {ref}")
        trials = [(dict(best), best_ms)]
        # two rounds of coordinate descent: a knob's best value can depend on another's (a larger
        # visited table only pays off with fewer waves per group, which the first round may have
        # rejected at the smaller table)
        knobs = [("queries_per_wave", (1, 2, 4)),
                 ("waves_per_query", (1, 2)),
                 ("waves_per_group", (1, 2, 4)),
                 ("hash_log2", tuple(range(8, 15))),
                 ("shape", SHAPES),
                 ("slack", (0, 8, 32, 128)),
                 ("groups_per_cu", (0, 1, 2, 4, 8)),
                 ("visited_bits", (0, 32)),
                 ("bag_in_lds", (0, 1)),
                 ("best_in_global", (-1, 0, 1))]
        if storage != STORE_F32:
            knobs.append(("storage", (-1, STORE_F32)))
        if judge is not None:
            knobs.append(("bloom_k", (0, -1)))
            knobs.append(("multiprobe", (1, 2)))
        knobs = [kc for kc in knobs if kc[0] not in freeze]
        best.setdefault("storage", -1)
        best.setdefault("bloom_k", 0)
        best.setdefault("queries_per_wave", 1)
        best.setdefault("multiprobe", 1)
        best.setdefault("bag_in_lds", 0)
        best.setdefault("waves_per_query", 1)
        best.setdefault("best_in_global", -1)
        rows_f32 = storage == STORE_F32 or best.get("storage", -1) == STORE_F32
        for _round in range(2):
            for key, choices in knobs:
                cur = best[key]
                bloom = best.get("bloom_k", 0) != 0
                if key == "hash_log2" and bloom:
                    # the filter's words: around the reference sizing for the tier's pq
                    from . import bloom_size

                    h0 = bloom_size(t["pq_max"])[0]
                    choices = (0,) + tuple(h for h in (h0 - 1, h0 + 1, h0 + 2) if 1 <= h <= 15)
                elif key == "hash_log2" and cur == 0:
                    cur = -1
                if key == "visited_bits" and bloom:
                    continue
                # sub-groups use the exact table; multi-probe needs sub-groups
                if (key == "queries_per_wave" and bloom) or (key == "bloom_k" and best["queries_per_wave"] > 1):
                    continue
                if key == "multiprobe" and best["queries_per_wave"] == 1:
                    continue
                if key == "bag_in_lds" and (best["queries_per_wave"] > 1 or not rows_f32 or best["waves_per_query"] > 1
                                            or best["best_in_global"] == 1):
                    continue
                # the best list's front in global memory: one query per wave, float32 rows, global bag
                if key == "best_in_global" and (best["queries_per_wave"] > 1 or not rows_f32 or best["bag_in_lds"]):
                    continue
                # the helper wave: FP32 mode, float32 rows, exact table, one query per wave, global bag
                if key == "waves_per_query" and (mode == MODE_EXACT or not rows_f32 or bloom or best["queries_per_wave"] > 1
                                                 or best["bag_in_lds"]):
                    continue
                if key in ("queries_per_wave", "bag_in_lds", "bloom_k") and best["waves_per_query"] > 1:
                    continue
                for v in _neighbours(cur, choices):
                    cand = dict(best)
                    cand[key] = v
                    if key == "bloom_k":
                        cand["hash_log2"] = 0  # the exact table's or the filter's default size
                    if key == "queries_per_wave" and v > 1:
                        cand["shape"] = min(cand["shape"], 3) if cand["shape"] >= 0 else -1
                    if key == "queries_per_wave" and v == 1:
                        cand["multiprobe"] = 1
                    if key == "queries_per_wave" and v > 1:
                        cand["bag_in_lds"] = 0
                        cand["best_in_global"] = -1
                    if key == "shape" and cand["queries_per_wave"] > 1 and v > 3:
                        continue
                    approx = (key in ("bloom_k", "multiprobe") or (key == "hash_log2" and bloom)
                              or (key == "queries_per_wave" and best["multiprobe"] > 1))
                    ms, out = measure(cand, want=ref, approx=approx)
                    if ms is None and key == "hash_log2" and cand["waves_per_group"] > 1:
                        # a table too large for the group's LDS: try it with one wave per group
                        cand["waves_per_group"] = 1
                        ms, out = measure(cand, want=ref, approx=approx)
                    trials.append((dict(cand), ms))
                    # a candidate must be 1% faster (timing noise); a grid smaller than all that
                    # fit must win clearly: it caps the queries in flight, which costs a batch
                    # larger than the tuning sample
                    margin = 0.97 if key == "groups_per_cu" and v != 0 else 0.99
                    if ms is not None and ms < best_ms * margin:
                        best, best_ms = cand, ms
                        if approx:
                            ref = out  # later knobs must keep the accepted filter's results
        tiers[ti] = best
        entry_log = {"tier": ti, "pq_max": t["pq_max"], "queries": int(nq), "ms": best_ms,
                     "qps": nq / (best_ms * 1e-3), "config": best,
                     "trials": [{"config": c, "ms": m} for c, m in trials]}
        log.append(entry_log)
        if verbose:
            print(f"[autotune] tier {ti} pq<={t['pq_max']}: {nq} queries, {best_ms:.2f} ms, {best}", flush=True)
    return {"metric": index.metric, "n": index.n, "dim": index.dim, "mode": mode, "storage": storage,
            "bounds": bounds, "tiers": tiers, "log": log, "seconds": time.time() - t_start}


def save(table, path):
    with open(path, "w") as f:
        json.dump(table, f, indent=1)


def load(path):
    with open(path) as f:
        return json.load(f)
