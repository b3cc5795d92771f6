"""Benchmark of the MI355X matching path; prints one JSON line.

Headline (value): embeddings/s of the batched path on the large cyclic
workload: a Chung-Lu power-law graph of 16,000,000 vertices from 256M edge
samples (degree exponent 2.1, maximum degree 50,000; loops and duplicates
removed), 64 labels drawn Zipf(0.8), resident in HBM; 128 random-walk
queries of 12..24 vertices with 2..6 extra (non-tree) edges each,
count-only with a cap of 1e9 embeddings per query.  One step = one
csm_batch_match (filter, candidate indexes and one persistent join kernel)
over a group of 16 of the 128 queries, the groups taken in turn; value =
embeddings counted in the timed steps / their time.  The same 128 queries
are also run once per query on the plain path (csm_match) and once in a
single batch call, for queries/s and every count compared.  The roofline
entry is the persistent join kernel, timed with HIP events on its stream:
algorithmic bytes = the kernel's own logical reads and writes, counted by
the kernel (rows read, list locators, list entries examined and searched,
hub-bitmap words, rows written).  The plain path's join classes also carry
the reference's extension bytes for the same rows (join.cu:99-177: row,
offsets, and id + candidate word per adjacency entry of the row's parent
vertex, rows written) as `ref_bytes_per_launch`; the batch kernel does not
read its rows' parent degrees, so its entry has null there.

rmat16_full: the plain path on R-MAT scale 16 (65,536 vertices, 2^20 edge
samples, a=0.57 b=0.19 c=0.19), 16 labels Zipf(1.0), two 8-vertex
random-walk queries (tree-like; with 3 extra cycle edges), full enumeration
(no cap) under a time limit per query: an unfinished query reports the
count reached as a lower bound.

batched: R-MAT scale 20 (2^24 edge samples), 32 labels Zipf(1.0); 256
random-walk queries of 4..32 vertices with average degree 2..4, per-query
cap 1e8, in the reference's matching order.  The plain path runs them query
by query and one csm_batch_match runs all 256, both with the same 2 s limit
per query (a query over it reports a lower bound on both paths); queries/s
compared over all 256 and every finished count must agree.  `reordered`
repeats the comparison with the candidate-aware matching order.

served: R-MAT scale 22 (2^26 edge samples), 32 labels Zipf(1.0), resident;
128 random-walk queries (4..32 vertices, geometric with mean 10) arriving
as Poisson streams at 32, 64 and 96 queries/s; p50/p99 latency from arrival
to result of the servers beside a FIFO plain per-query server.

cpu_baseline times the CPU oracle (same graph, queries, matching order and
cap) on a bounded sample of the headline workload (first queries of the
first group, a time budget each), with up to 16 threads, and names the host
CPU.
"""
import argparse
import json
import os
import random
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import clustersm_amd as csm  # noqa: E402

METRIC = "embeddings/s + batched queries/s vs per-query launch; p50/p99 latency"
HBM_PEAK_GBS = 8000.0


HEADLINE_QUERIES = 128
HEADLINE_CAP = 1_000_000_000


def chung_lu_16m_graph():
    """Chung-Lu 16M vertices / 256M edge samples, 64 Zipf(0.8) labels, and its
    128 random-walk queries (12..24 vertices, 2..6 extra edges)."""
    hg = csm.gen_chung_lu(16_000_000, 256_000_000, 2.1, 50_000, 64, 0.8, 41)
    return hg


def headline_queries(host, n):
    r = random.Random(4)
    out = []
    for i in range(n):
        nv, extra = r.randint(12, 24), r.randint(2, 6)
        out.append(csm.gen_query(host, nv, csm.QUERY_EXTRA, 300 + i, extra_edges=extra))
    return out


def kernel_totals(result_lists):
    tot = {}
    for results in result_lists:
        for r in results:
            for c in range(csm.K_NUM):
                k = r.kernels[c]
                t = tot.setdefault(c, [0, 0, 0, 0])
                t[0] += k.launches
                t[1] += k.ns
                t[2] += k.bytes
                t[3] += k.bytes_ref
    return tot


def measured_traffic(profile):
    """HBM bytes per launch by kernel class, from a committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE, calibrated with a copy kernel)."""
    out = {}
    path = os.path.join(ROOT, "profiles", profile)
    if os.path.exists(path):
        for ln in open(path):
            if ln.startswith("traffic_bytes_per_launch:"):
                cls, val = ln.split(":", 1)[1].split()
                out[cls] = int(val)
    return out


def roofline_from(tot, profile, cls=None):
    """Dominant kernel class (or `cls`): achieved = algorithmic bytes (the
    kernel's logical reads and writes) / device time (bytes/ns = GB/s);
    ref_equivalent = the reference algorithm's bytes for the same rows / time,
    which exceeds any bandwidth where the candidate index skips adjacency
    scans (not a roofline figure)."""
    dom = cls if cls is not None else max(tot, key=lambda c: tot[c][1])
    launches, ns, byts, ref = tot[dom]
    achieved = byts / ns if ns else 0.0
    traffic = measured_traffic(profile).get(csm.KERNEL_NAMES[dom])
    return {
        "kernel": csm.KERNEL_NAMES[dom],
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "launches": launches,
        "avg_launch_us": round(ns / launches / 1e3, 2) if launches else None,
        "bytes_per_launch": byts // launches if launches else None,
        "ref_bytes_per_launch": ref // launches if launches and ref else None,
        "ref_equivalent_GBps": round(ref / ns, 1) if ns and ref else None,
    }


def kernels_json(tot):
    return {csm.KERNEL_NAMES[c]: {"launches": v[0], "ms": round(v[1] / 1e6, 3), "GB": round(v[2] / 1e9, 3),
                                  "ref_GB": round(v[3] / 1e9, 3) if v[3] else None}
            for c, v in tot.items() if v[0]}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(hg, host_queries, cap, seconds):
    """The CPU backtracking matcher (same matching order, filter and cap) on
    the same graph and queries, each query stopped by a time budget, with up
    to 16 threads over root candidates."""
    from oracle import oracle as O

    t0 = time.time()
    labels, edges = hg.arrays()
    data = O.Data.from_arrays(labels, edges)
    del labels, edges
    build_s = time.time() - t0
    threads = min(os.cpu_count() or 1, 16)
    per = seconds / len(host_queries)
    total, secs, timed_out, capped = 0, 0.0, 0, 0
    for hq in host_queries:
        ql, qe = hq.arrays()
        oq = O.Query.from_arrays(ql, qe)
        c, cp, _, to, sec = O.match(data, oq, oq.plan(), cap=cap, threads=threads, budget_s=per, refine=False)
        total += c
        secs += sec
        timed_out += to
        capped += cp
    return {
        "value": round(total / secs, 1) if secs else 0.0,
        "unit": "embeddings/s",
        "cores": threads,
        "kind": "port",
        "sample": f"same graph and cap, first {len(host_queries)} queries of the first group, backtracking oracle "
                  f"with {threads} threads, {per:.1f} s budget per query ({timed_out} stopped by the budget, "
                  f"{capped} reached the cap): {total} embeddings in {secs:.1f} s (oracle CSR build "
                  f"{build_s:.1f} s, not timed)",
        "host": {"cpu_model": cpu_model(), "nproc": os.cpu_count()},
    }


def headline_section(ctx, args):
    """The headline: timed steps of batched groups, then one pass of the plain
    path and one whole-set batch call over the same 128 queries."""
    t0 = time.time()
    hg = chung_lu_16m_graph()
    g = ctx.upload(hg)
    host = g.to_host()
    hqs = headline_queries(host, HEADLINE_QUERIES)
    del host
    qs = [g.prepare(h.plan()) for h in hqs]
    nv, ne, maxdeg = g.info()
    print(f"# headline setup {time.time() - t0:.1f}s: |V|={nv} |E|={ne} max_degree={maxdeg}", file=sys.stderr, flush=True)
    group = args.group
    groups = [qs[i:i + group] for i in range(0, len(qs), group)]

    def step(i):
        return csm.batch_match(g, groups[i % len(groups)], cap=HEADLINE_CAP)

    for i in range(args.warmup):
        step(i)
    ctx.set_profiling(True)
    t1 = time.perf_counter()
    steps = [step(args.warmup + i) for i in range(args.steps)]
    elapsed = time.perf_counter() - t1
    ctx.set_profiling(False)
    embeddings = sum(x.count for st in steps for x in st)
    anomalies = sum(x.anomalies for st in steps for x in st)
    recounted = sum(1 for st in steps for x in st if x.path in (csm.PATH_BATCH_BUDGET, csm.PATH_BATCH_ABORT))
    print(f"# headline timed {args.steps} steps in {elapsed:.2f}s: {embeddings} embeddings", file=sys.stderr, flush=True)
    tot = kernel_totals([[st[0]] for st in steps])  # a batch call's kernel totals are on its first result
    # one pass of the plain path, and one batch call over all 128
    plain, t_plain = [], 0.0
    for q in qs:
        t2 = time.perf_counter()
        plain.append(q.match(cap=HEADLINE_CAP)[0])
        t_plain += time.perf_counter() - t2
    print(f"# headline plain path: {len(qs)} queries in {t_plain:.2f}s", file=sys.stderr, flush=True)
    t3 = time.perf_counter()
    whole = csm.batch_match(g, qs, cap=HEADLINE_CAP)
    t_whole = time.perf_counter() - t3
    print(f"# headline one batch call: {len(qs)} queries in {t_whole:.2f}s", file=sys.stderr, flush=True)
    mism = [i for i, (x, p) in enumerate(zip(whole, plain)) if (x.count, x.capped, x.status) != (p.count, p.capped, 0)]
    by_step = {}
    for i, st in enumerate(steps):
        base = ((args.warmup + i) % len(groups)) * group
        for j, x in enumerate(st):
            p = plain[base + j]
            if (x.count, x.capped, x.status) != (p.count, p.capped, 0):
                by_step.setdefault(base + j, [x.count, int(x.capped), x.status, p.count, int(p.capped)])
    out = {
        "elapsed": elapsed,
        "embeddings": embeddings,
        "tot": tot,
        "graph": (nv, ne, maxdeg),
        "detail": {
            "queries_per_s": round(args.steps * group / elapsed, 2),
            "plain_queries_per_s": round(len(qs) / t_plain, 2),
            "plain_embeddings_per_s": round(sum(p.count for p in plain) / t_plain, 1),
            "one_call_queries_per_s": round(len(qs) / t_whole, 2),
            "one_call_embeddings_per_s": round(sum(x.count for x in whole) / t_whole, 1),
            "speedup_vs_plain": round((args.steps * group / elapsed) / (len(qs) / t_plain), 2),
            "one_call_speedup_vs_plain": round(t_plain / t_whole, 2),
            "capped_queries": sum(1 for p in plain if p.capped),
            "counts_equal_plain": not mism and not by_step,
            "mismatches": mism[:8] + [[k] + v for k, v in list(by_step.items())[:8]],
            "batch_anomalies": anomalies,
            "recounted_on_plain_path": recounted,
        },
        "hg": hg,
        "hqs": hqs,
    }
    for q in qs:
        q.free()
    g.free()
    return out


def rmat16_section(ctx, limits_ms):
    """The plain path on R-MAT 16 with the two 8-vertex queries, no cap."""
    hg = csm.gen_rmat(16, 1 << 20, 16, 1.0, 1)
    g = ctx.upload(hg)
    out = []
    named = [("tree", csm.gen_query(hg, 8, csm.QUERY_TREE, 11)),
             ("cycles+3", csm.gen_query(hg, 8, csm.QUERY_EXTRA, 12, extra_edges=3))]
    for (name, hq), lim in zip(named, limits_ms):
        q = g.prepare(hq.plan())
        ctx.set_profiling(True)
        t = time.perf_counter()
        r, _ = q.match(cap=0, time_limit_ms=lim)
        dt = time.perf_counter() - t
        ctx.set_profiling(False)
        done = r.status == 0
        print(f"# rmat16 {name}: {'exact' if done else 'lower bound'} {r.count} in {dt:.2f}s", file=sys.stderr,
              flush=True)
        tot = kernel_totals([[r]])
        entry = {"name": name, "count": r.count, "exact": done, "seconds": round(dt, 3),
                 "time_limit_ms": lim, "embeddings_per_s": round(r.count / dt, 1), "items": r.items,
                 "roofline": roofline_from(tot, "count_kernel_rmat16_plain.md")}
        if name == "cycles+3":
            # the same first 5 s of the search with the count level's lists
            # copied into LDS (lds_lists option), beside the default
            rate = {}
            for lds in (False, True):
                t = time.perf_counter()
                rs, _ = q.match(cap=0, time_limit_ms=5000, lds_lists=lds)
                rate[lds] = rs.items / (time.perf_counter() - t)
            entry["lds_lists_sample"] = {"seconds": 5, "default_items_per_s": round(rate[False], 1),
                                         "lds_lists_items_per_s": round(rate[True], 1),
                                         "ratio": round(rate[True] / rate[False], 3)}
            print(f"# rmat16 {name}: lds_lists {rate[True]:.3g} items/s, default {rate[False]:.3g}", file=sys.stderr,
                  flush=True)
        out.append(entry)
        q.free()
    g.free()
    return {"workload": "plain path, R-MAT scale 16 (2^20 samples, 16 Zipf(1.0) labels), two 8-vertex random-walk "
                        "queries (tree-like; +3 cycle edges), full enumeration (no cap) under a time limit; an "
                        "unfinished query's count is a lower bound",
            "queries": out}


def batched_section(ctx, nqueries, cap, item_budget, sample=0):
    """The R-MAT scale 20 query set over all its queries, in the reference's
    matching order: the plain path query by query, then one batch call, both
    with the same per-query item budget (a query over it stops with a lower
    bound on both paths); queries/s over every query, and every count that
    both paths finish must agree.

    sample > 0 is for counter profiles: only the batch, with default options,
    over the first `sample` queries (warm-up call and timed call), with no
    plain leg and no autotuner."""
    t0 = time.time()
    hg = csm.gen_rmat(20, 1 << 24, 32, 1.0, 3)
    g = ctx.upload(hg)
    host = g.to_host()
    r = random.Random(5)
    plans = []
    for i in range(nqueries):
        nv = r.randint(4, 32)
        avg = r.uniform(2, 4)
        plans.append(csm.gen_query(host, nv, csm.QUERY_AVG_DEGREE, 1000 + i, avg_degree=avg).plan())
    qs = [g.prepare(p) for p in plans]
    print(f"# batched setup {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    if sample:
        kq = qs[:sample]
        csm.batch_match(g, kq, cap=cap)
        ctx.set_profiling(True)
        t2 = time.perf_counter()
        res = csm.batch_match(g, kq, cap=cap)
        t_batch = time.perf_counter() - t2
        ctx.set_profiling(False)
        print(f"# batched sample: {len(kq)} queries in {t_batch:.2f}s", file=sys.stderr, flush=True)
        out = {"profile_sample": f"first {len(kq)} of the {nqueries} queries, default options, batch only",
               "queries_per_s": round(len(kq) / t_batch, 2), "kernels": kernels_json(kernel_totals([[res[0]]]))}
        for q in qs:
            q.free()
        g.free()
        return out
    plain, plain_s = [], []
    for q in qs:
        t1 = time.perf_counter()
        plain.append(q.match(cap=cap, item_budget=item_budget)[0])
        plain_s.append(time.perf_counter() - t1)
    t_plain = sum(plain_s)
    p_budget = [i for i in range(nqueries) if plain[i].status == csm.E_BUDGET]
    print(f"# plain path: {nqueries} queries in {t_plain:.2f}s, {len(p_budget)} stopped by the item budget",
          file=sys.stderr, flush=True)
    t3 = time.perf_counter()
    tuned = csm.batch_autotune(g, qs, cap=cap, sample=64, item_budget=item_budget)  # untimed: once per graph and mix
    tune_s = time.perf_counter() - t3
    print(f"# autotune {tune_s:.1f}s: wgs_per_cu={tuned.wgs_per_cu} max_out={tuned.max_out} "
          f"block_kb={tuned.block_kb} ilp={tuned.ilp}", file=sys.stderr, flush=True)
    csm.batch_match(g, qs, opts=tuned)  # warm-up (code objects, arena layout)
    ctx.set_profiling(True)
    t2 = time.perf_counter()
    res = csm.batch_match(g, qs, opts=tuned)
    t_batch = time.perf_counter() - t2
    ctx.set_profiling(False)
    b_budget = [i for i in range(nqueries) if res[i].status == csm.E_BUDGET]
    print(f"# batched: {nqueries} queries in {t_batch:.2f}s, {len(b_budget)} stopped by the item budget",
          file=sys.stderr, flush=True)
    both = [i for i in range(nqueries) if plain[i].status == 0 and res[i].status == 0]
    # finished on both: identical; stopped on one path: its count is a lower
    # bound of the other's (checked exactly below with the reordered plain path)
    mism = [[i, res[i].count, int(res[i].capped), res[i].status, res[i].path, plain[i].count, int(plain[i].capped),
             plain[i].status]
            for i in range(nqueries)
            if (i in both and (res[i].count, res[i].capped) != (plain[i].count, plain[i].capped))
            or (plain[i].status == csm.E_BUDGET and res[i].status == 0 and res[i].count < plain[i].count)
            or (res[i].status == csm.E_BUDGET and plain[i].status == 0 and plain[i].count < res[i].count)]
    paths = {}
    for x in res:
        paths[x.path] = paths.get(x.path, 0) + 1
    tot = kernel_totals([[res[0]]])
    bk = tot.get(csm.K_BATCH, [0, 0, 0, 0])
    out = {
        "workload": f"R-MAT scale 20 (2^24 samples, 32 Zipf(1.0) labels), {nqueries} random-walk queries of "
                    f"4..32 vertices, average degree 2..4, cap {cap:.0e} per query, count-only, reference "
                    f"matching order; both paths bound each query by {item_budget:.0e} items",
        "queries": nqueries,
        "queries_per_s": round(nqueries / t_batch, 2),
        "plain_queries_per_s": round(nqueries / t_plain, 2),
        "speedup_vs_plain": round(t_plain / t_batch, 2),
        "stopped_by_item_budget": {"plain": p_budget, "batch": b_budget,
                                   "rule": "a query over the budget reports a lower bound; both paths stop the "
                                           "same queries (none of them has an embedding in the plain path's "
                                           "first items)"},
        "counts_equal_plain": not mism,
        "compared": len(both),
        "mismatches": mism,
        "autotuned": {"wgs_per_cu": tuned.wgs_per_cu, "max_out": tuned.max_out, "block_kb": tuned.block_kb,
                      "ilp": tuned.ilp, "sample_queries": min(64, nqueries), "tuning_s": round(tune_s, 1)},
        "capped_queries": sum(1 for x in res if x.capped),
        "embeddings_per_s": round(sum(x.count for x in res) / t_batch, 1),
        "finished_by": {{csm.PATH_PLAIN: "plain", csm.PATH_HOST: "host", csm.PATH_BATCH: "batch kernel",
                         csm.PATH_BATCH_BUDGET: "plain after the item budget",
                         csm.PATH_BATCH_ABORT: "plain after an early stop"}[k]: v for k, v in paths.items()},
        "batch_anomalies": sum(x.anomalies for x in res),
        "kernels": kernels_json(tot),
        "join_kernel": {"ms": round(bk[1] / 1e6, 1), "algorithmic_GB": round(bk[2] / 1e9, 2),
                        "GB_per_s": round(bk[2] / bk[1], 1) if bk[1] else None,
                        "ref_GB": round(bk[3] / 1e9, 2) if bk[3] else None},
        # candidate tests (row x driver entry) done by each path; capped
        # queries overshoot their cap by different amounts on the two paths
        "items": {"batch": sum(x.items for x in res), "plain": sum(p.items for p in plain),
                  "top_batch": [[i, x.items, plain[i].items, int(x.capped)]
                                for i, x in sorted(enumerate(res), key=lambda t: -t[1].items)[:8]]},
    }
    # the same comparison with the candidate-aware matching order on both
    # paths (the reorder option: each path orders from its own filter counts,
    # inside the timed calls); counts and cap flags do not depend on the
    # order, so they must equal the default order's
    plain_r, t_plain_r = [], 0.0
    for q in qs:
        t1 = time.perf_counter()
        plain_r.append(q.match(cap=cap, item_budget=item_budget, reorder=True)[0])
        t_plain_r += time.perf_counter() - t1
    csm.batch_match(g, qs, opts=tuned, reorder=True)
    t2 = time.perf_counter()
    res_r = csm.batch_match(g, qs, opts=tuned, reorder=True)
    t_batch_r = time.perf_counter() - t2
    print(f"# reordered: plain {nqueries} queries in {t_plain_r:.2f}s, batched in {t_batch_r:.2f}s",
          file=sys.stderr, flush=True)
    same_r = all((x.count, x.capped, x.status) == (p.count, p.capped, 0) for x, p in zip(res_r, plain_r))
    same_default = all((plain_r[i].count, plain_r[i].capped) == (plain[i].count, plain[i].capped) for i in both)
    # the queries the default order stopped: the reference-order batch's
    # result against the reordered plain path's exact one (counts and cap
    # flags do not depend on the order)
    stopped = [i for i in range(nqueries) if plain[i].status != 0 or res[i].status != 0]
    out["stopped_checked_by_reordered_plain"] = {
        "queries": stopped,
        "batch_equals_reordered_plain": all(
            (res[i].count, res[i].capped, res[i].status) == (plain_r[i].count, plain_r[i].capped, 0)
            for i in stopped if res[i].status == 0 and plain_r[i].status == 0),
        "results": [[i, res[i].count, int(res[i].capped), res[i].status, plain_r[i].count, int(plain_r[i].capped),
                     plain_r[i].status] for i in stopped],
    }
    out["reordered"] = {
        "order": "candidate-aware (reorder option) on both paths; the default is the reference's BFS edge order",
        "queries": nqueries,
        "queries_per_s": round(nqueries / t_batch_r, 2),
        "plain_queries_per_s": round(nqueries / t_plain_r, 2),
        "speedup_vs_plain": round(t_plain_r / t_batch_r, 2),
        "speedup_vs_default_plain": round(t_plain / t_batch_r, 2),
        "counts_equal_plain": same_r,
        "counts_equal_default_order": same_default,
        "items_batch": sum(x.items for x in res_r),
    }
    for q in qs:
        q.free()
    g.free()
    return out


def footprint(plan, label_count, n, avg_degree, words):
    """Admission estimate of one query's device footprint in bytes: its
    candidate bitmaps plus, per query edge in both directions, the expected
    candidate-neighbour list bytes before filtering (label shares of the data
    graph)."""
    est = 8.0 * words * plan.num_v
    for e in range(plan.num_e):
        cs = float(label_count[plan.labels[plan.edge_src[e]]]) if plan.labels[plan.edge_src[e]] < len(label_count) else 0.0
        ct = float(label_count[plan.labels[plan.edge_dst[e]]]) if plan.labels[plan.edge_dst[e]] < len(label_count) else 0.0
        est += cs * (8.0 + 4.0 * avg_degree * ct / n) + ct * (8.0 + 4.0 * avg_degree * cs / n)
    return int(est) + 4096


def served_section(ctx, nqueries, qps, cap, limit_ms, max_batch, budget_gb, seed=7, extra_qps=None):
    """Served workload: a Poisson stream of random-walk queries (4..32
    vertices, geometric with mean 10) against a resident R-MAT scale 22 graph.

    Batched server (live, on the GPU): whenever the previous batch is done,
    every request that has arrived is admitted in arrival order, up to
    max_batch and while the footprint estimates fit the budget; the batch runs
    as one csm_batch_match (filter, indexes, persistent join; exact, no time
    limit), a batch of one as one csm_match, and each request's latency runs
    from its arrival to the batch's end.  Plain server: the same arrivals served one query at a time, FIFO,
    with each query's measured plain-path service time (prepare + match).
    Queries the plain path cannot finish within limit_ms are screened out
    first, and every served count must equal its plain count.  extra_qps
    adds arrival streams at other rates over the same queries (other_loads)."""
    t0 = time.time()
    hg = csm.gen_rmat(22, 1 << 26, 32, 1.0, 11)
    g = ctx.upload(hg)
    host = g.to_host()
    labels, _ = host.arrays()
    label_count = np.bincount(labels)
    n, m, _ = g.info()
    words = (n + 63) // 64
    avg_degree = 2.0 * m / max(n, 1)
    r = random.Random(seed)
    plans = []
    for i in range(nqueries):
        nv = 4
        while nv < 32 and r.random() > 1.0 / 7.0:  # 4 + geometric: mean 10
            nv += 1
        plans.append(csm.gen_query(host, nv, csm.QUERY_AVG_DEGREE, 7000 + i, avg_degree=r.uniform(2, 4)).plan())
    print(f"# served setup {time.time() - t0:.1f}s: |V|={n} |E|={m}", file=sys.stderr, flush=True)
    # plain service times (and screening)
    plain, service = [], []
    for pl in plans:
        t1 = time.perf_counter()
        q = g.prepare(pl)
        plain.append(q.match(cap=cap, time_limit_ms=limit_ms)[0])
        q.free()
        service.append(time.perf_counter() - t1)
    keep = [i for i in range(nqueries) if plain[i].status == 0]
    print(f"# served: plain path finishes {len(keep)} of {nqueries} within {limit_ms} ms, "
          f"{sum(service[i] for i in keep):.2f}s of service", file=sys.stderr, flush=True)
    wq = g.prepare(plans[keep[0]])
    csm.batch_match(g, [wq], cap=cap)  # warm-up (code objects)
    wq.free()
    budget = budget_gb * (1 << 30)
    pct = lambda v, f: round(1e3 * float(np.percentile(v, f)), 1)
    mism = []

    def run_load(rate, rs):
        # one Poisson arrival stream over the kept queries
        arr, t = [], 0.0
        for _ in keep:
            t += rs.expovariate(rate)
            arr.append(t)
        # plain server: FIFO over measured service times
        lat_plain, free_at = [], 0.0
        for a, i in zip(arr, keep):
            start = max(a, free_at)
            free_at = start + service[i]
            lat_plain.append(free_at - a)
        # batched server, live
        lat = [0.0] * len(keep)
        pending, nxt, done, sizes = [], 0, 0, []
        singles = 0
        ts = time.perf_counter()
        while done < len(keep):
            now = time.perf_counter() - ts
            while nxt < len(keep) and arr[nxt] <= now:
                pending.append(nxt)
                nxt += 1
            if not pending:
                time.sleep(max(0.0, arr[nxt] - now))
                continue
            batch, used = [], 0
            for j in pending:
                est = footprint(plans[keep[j]], label_count, n, avg_degree, words)
                if batch and (len(batch) >= max_batch or used + est > budget):
                    break
                batch.append(j)
                used += est
            pending = pending[len(batch):]
            qs = [g.prepare(plans[keep[j]]) for j in batch]
            if len(qs) == 1:  # a batch of one runs on the plain path: no batch set-up
                res = [qs[0].match(cap=cap)[0]]
                singles += 1
            else:
                res = csm.batch_match(g, qs, cap=cap)
            tend = time.perf_counter() - ts
            for j, x, q in zip(batch, res, qs):
                lat[j] = tend - arr[j]
                p = plain[keep[j]]
                if (x.status, x.count, x.capped) != (0, p.count, p.capped):
                    mism.append([keep[j], x.count, int(x.capped), x.status, x.path, p.count, int(p.capped),
                                 len(batch), rate])
                q.free()
            sizes.append(len(batch))
            done += len(batch)
        span = time.perf_counter() - ts
        print(f"# served at {rate} queries/s: {len(sizes)} batches in {span:.2f}s", file=sys.stderr, flush=True)
        return {"offered_qps": rate, "arrival_span_s": round(arr[-1], 3),
                "batched": {"p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "queries_per_s": round(len(keep) / span, 2),
                            "drain_ms": round(1e3 * (span - arr[-1]), 1),
                            "batches": len(sizes), "mean_batch": round(sum(sizes) / max(1, len(sizes)), 1),
                            "batches_of_one_on_plain_path": singles},
                "plain_fifo": {"p50_ms": pct(lat_plain, 50), "p99_ms": pct(lat_plain, 99),
                               "queries_per_s": round(len(keep) / max(free_at - arr[0], 1e-9), 2),
                               "mean_service_ms": round(1e3 * sum(service[i] for i in keep) / max(1, len(keep)), 1)}}

    main = run_load(qps, r)
    more = [run_load(x, random.Random(seed + 1 + k)) for k, x in enumerate(extra_qps or [])]

    def run_cpp(rate, rs):
        """The overlapped C++ server (csm_serve_*): requests submitted at their
        arrival times; latency is the server's own arrival-to-result time."""
        arr, t = [], 0.0
        for _ in keep:
            t += rs.expovariate(rate)
            arr.append(t)
        ctx.serve_start(g, budget_bytes=budget, max_inflight=max_batch, cap=cap)
        tickets = []
        ts = time.perf_counter()
        try:
            for a, i in zip(arr, keep):
                wait = a - (time.perf_counter() - ts)
                if wait > 0:
                    time.sleep(wait)
                tickets.append(ctx.serve_submit(plans[i]))
            res = [ctx.serve_poll(tk) for tk in tickets]
        finally:
            ctx.serve_stop()
        span = time.perf_counter() - ts
        lat = [x.ms_latency / 1e3 for x in res]
        for x, i in zip(res, keep):
            p = plain[i]
            if (x.status, x.count, x.capped) != (0, p.count, p.capped):
                mism.append([i, x.count, int(x.capped), x.status, "cpp_server", p.count, int(p.capped), rate])
        print(f"# C++ server at {rate} queries/s: {span:.2f}s", file=sys.stderr, flush=True)
        # queries/s over first arrival to last result is bounded by the
        # stream's own span (len / arr[-1]); drain_ms: last result after the
        # last arrival
        return {"offered_qps": rate, "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
                "queries_per_s": round(len(keep) / span, 2), "arrival_span_s": round(arr[-1], 3),
                "arrival_bound_qps": round(len(keep) / arr[-1], 2), "drain_ms": round(1e3 * (span - arr[-1]), 1)}

    rates = [qps] + list(extra_qps or [])
    cpp = [run_cpp(x, random.Random(seed + 201 + k)) for k, x in enumerate(rates)]
    g.free()
    return {
        "workload": f"R-MAT scale 22 (2^26 samples, 32 Zipf(1.0) labels) resident; Poisson arrivals at {qps} "
                    f"queries/s of random-walk queries, 4..32 vertices (geometric, mean 10), average degree 2..4, "
                    f"cap {cap:.0e}; admission by footprint estimate within {budget_gb} GB, at most {max_batch} "
                    f"per batch",
        "server": "batched: continuous batching at batch granularity written over csm_batch_match, each batch "
                  "takes every request that arrived while the previous one ran (a batch of one runs as csm_match); "
                  "cpp_server: the library's server (csm_serve_*) on the same kind of arrivals, each request admitted "
                  "into the running resident join kernel",
        "queries": len(keep),
        "screened_out": [i for i in range(nqueries) if i not in keep],
        "offered_qps": qps,
        "batched": main["batched"],
        "plain_fifo": main["plain_fifo"],
        "other_loads": more,
        "cpp_server": {"server": "csm_serve_*: a resident persistent join kernel; arriving requests are admitted "
                                 "by footprint and slots, prepared (filter, candidate index) on a second stream beside "
                                 "the running kernel and handed to it through a doorbell in pinned memory; each "
                                 "result is published as its query's last task ends", "loads": cpp},
        "counts_equal_plain": not mism,
        "mismatches": mism,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--group", type=int, default=16, help="headline queries per step (one batch call)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rmat16-limits-ms", type=str, default="5000,45000",
                    help="time limits of the two uncapped R-MAT 16 queries (tree, cycles+3)")
    ap.add_argument("--no-rmat16", action="store_true")
    ap.add_argument("--batched-queries", type=int, default=256)
    ap.add_argument("--batched-cap", type=int, default=100_000_000)
    ap.add_argument("--batched-item-budget", type=int, default=20_000_000_000)
    ap.add_argument("--batched-limit-ms", type=int, default=2000, help="served workload: plain-path screening limit")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-sample", type=int, default=0,
                    help="profiling: batch only over the first N queries (no plain leg, no autotuner)")
    ap.add_argument("--arena-gb", type=float, default=96.0)
    ap.add_argument("--served-queries", type=int, default=128)
    ap.add_argument("--served-qps", type=float, default=32.0)
    ap.add_argument("--served-more-qps", type=str, default="64,96",
                    help="comma-separated further arrival rates for the served workload ('' for none)")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--headline-only", action="store_true", help="profiling: the headline steps only")
    ap.add_argument("--sections", type=str, default="headline,cpu,rmat16,batched,served",
                    help="comma-separated sections to run (the headline always runs unless omitted here)")
    args = ap.parse_args()
    sections = set(x.strip() for x in args.sections.split(","))
    if args.headline_only:
        sections = {"headline"}
    if args.gpus != 1:
        print(json.dumps({"error": "the library runs on one GPU; multi-GPU sharding is a design only "
                                   "(queries or root candidates split across GPUs, one u64 count reduction)"}))
        sys.exit(2)

    ctx = csm.Context(0, int(args.arena_gb * (1 << 30)))
    ms_copy = ctx.copy_probe(1 << 30, 10)
    copy_gbs = 2.0 * (1 << 30) / (ms_copy * 1e6)
    if "headline" not in sections:  # partial runs for profiling and development
        line = {"metric": METRIC, "partial": sorted(sections)}
        if "rmat16" in sections:
            line["rmat16_full"] = rmat16_section(ctx, [int(x) for x in args.rmat16_limits_ms.split(",")])
        if "batched" in sections:
            line["batched"] = batched_section(ctx, args.batched_queries, args.batched_cap, args.batched_item_budget,
                                              args.batched_sample)
        if "served" in sections:
            more = [float(x) for x in args.served_more_qps.split(",") if x.strip()]
            line["served"] = served_section(ctx, args.served_queries, args.served_qps, args.batched_cap,
                                            args.batched_limit_ms, 64, 64, extra_qps=more)
        ctx.close()
        print(json.dumps(line), flush=True)
        return
    hl = headline_section(ctx, args)
    elapsed, tot = hl["elapsed"], hl["tot"]
    nv, ne, maxdeg = hl["graph"]
    line = {
        "metric": METRIC,
        "value": round(hl["embeddings"] / elapsed, 1),
        "unit": "embeddings/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"batched path, Chung-Lu power-law graph (16M vertices, 256M edge samples, exponent 2.1, "
                        f"max degree 50,000, 64 Zipf(0.8) labels), {HEADLINE_QUERIES} random-walk queries of 12..24 "
                        f"vertices with 2..6 non-tree edges, count-only, cap 1e9 per query; one step = one "
                        f"csm_batch_match over {args.group} of the queries (groups in turn)",
            "graph_vertices": nv,
            "graph_edges": ne,
            "max_degree": maxdeg,
            "queries": HEADLINE_QUERIES,
            "queries_per_step": args.group,
            "cap_per_query": HEADLINE_CAP,
            "parallelism": "single GPU",
            **hl["detail"],
        },
        "roofline": roofline_from(tot, "batch_join_chung_lu_16m.md", cls=csm.K_BATCH),
        "kernels": kernels_json(tot),
        "copy_probe_GBps": round(copy_gbs, 1),
    }
    hg, hqs = hl["hg"], hl["hqs"]
    del hl
    if "cpu" in sections and not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(hg, hqs[:4], HEADLINE_CAP, args.cpu_seconds)
    del hg, hqs
    if "rmat16" in sections and not args.no_rmat16:
        lim = [int(x) for x in args.rmat16_limits_ms.split(",")]
        line["rmat16_full"] = rmat16_section(ctx, lim)
    if "batched" in sections and not args.no_batched:
        line["batched"] = batched_section(ctx, args.batched_queries, args.batched_cap, args.batched_item_budget,
                                          args.batched_sample)
    if "served" in sections and not args.no_served:
        more = [float(x) for x in args.served_more_qps.split(",") if x.strip()]
        line["served"] = served_section(ctx, args.served_queries, args.served_qps, args.batched_cap,
                                        args.batched_limit_ms, 64, 64, extra_qps=more)
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
