#!/usr/bin/env python3
"""Collision-query throughput of the MI355X engine on one GPU.

Default workload (`--workload c2`): a single large layout — a 4000x1500 rectangle with 5
rectangular holes, 2000 placed star/concave polygons of 8-256 vertices (min separation 0.5),
quadtree depth 8 — queried by 64 items of 8-512 vertices at 16384 continuous-rotation
transforms each: 1,048,576 collision queries per step (verdict = surrogate fail-fast ||
polygon query, the composition of lbf/benches/ci_bench.rs:107-120).  Inputs are resident in
HBM before timing.  One step = one pass of the batched persistent kernels over all queries,
with the schedule the autotuner picked (tuning runs before the timed region).

The JSON line also carries the plain form (one launch per query set), the hipGraph-replayed
form, the overlap gradient (4096 items x 16-64 poles, forward + backward per evaluation), a
roofline entry for the dominant kernel and the CPU oracle's throughput on the host cores (a
reported baseline).

`--workload c3`: 256 independent layouts (star containers of 64-2048 vertices, 50-400 items),
16 query items per layout x 4096 transforms (16.8M queries) with a per-query random 10%
HazardFilter, all in one launch.  `--workload c5`: one 6000x2000 layout with 20 holes and
about 20000 concave items of 64-1024 vertices (depth 10), 256 query items x 65536
transforms (16.8M queries) detected in one launch, the hazard sets collected, then one
overlap-gradient pass over all placed items (32 poles each).
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "jagua-rs_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
METRIC = "bit-exact collision queries/sec on 1 MI355X, batched vs plain; overlap-grad evals/s"


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(counters, n_queries, n_poly, v_q, n_ff):
    """Bytes one batch of queries must touch: transform (24 B), ff poles (12 B each), query
    vertices and POI when the polygon stage runs (8 B each), quadtree nodes (32 B), hazard
    entries (32 B), hazard edges (16 B), containment polygon vertices (8 B), verdict (1 B);
    the traversal counts come from the oracle's run of the same queries."""
    nodes, entries, edges, pip = (int(x) for x in counters)
    return (24 * n_queries + 12 * n_ff * n_queries + (8 * v_q + 8) * n_poly + 32 * nodes + 32 * entries
            + 16 * edges + 8 * pip + n_queries)


def timed(fn, reps, stream):
    """Average ms of fn() over reps, HIP events recorded on the stream the kernels run on."""
    import torch
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(stream)
    for _ in range(reps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def pmc_profile(name):
    """The counter summary committed under profiles/ for this workload: HBM bytes per step
    (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by the same counters' reading for a copy
    of known size), the bound the counters show, the per-kernel counters and the schedule they
    were taken under; {} when absent."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return {}
    text = open(path).read()
    out = {}
    m = re.search(r"traffic_bytes_per_step:\s*([0-9.eE+]+)", text)
    if m:
        out["traffic"] = float(m.group(1))
    m = re.search(r"bound:\s*([a-z-]+)", text)
    if m:
        out["bound"] = m.group(1)
    m = re.search(r"schedule:\s*(\{.*?\})", text)
    if m:
        out["schedule"] = json.loads(m.group(1))
    m = re.search(r"counters_json:\s*(\{.*\})", text)
    if m:
        out["counters"] = json.loads(m.group(1))
    return out


def base_result(args, workload, total, ms, dtype="f32"):
    return {
        "metric": METRIC, "value": total / (ms / 1e3), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": workload,
    }


def tune_and_time(plan, mats, status, args, stream):
    """Autotune the plan's schedule (outside the timed region), then time `steps` launches."""
    import torch
    log("autotuning the batched schedule")
    plan.detect(mats, status)
    default_ms = timed(lambda: plan.detect(mats, status), 3, stream)
    cfg, tuned_ms = plan.autotune(mats, status, reps=5)
    log(f"schedule {cfg}: {tuned_ms:.3f} ms (default schedule {default_ms:.3f} ms)")
    for _ in range(args.warmup):
        plan.detect(mats, status)
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    ms = timed(lambda: plan.detect(mats, status), args.steps, stream)
    return ms, {"schedule": cfg, "default_schedule_ms": default_ms}


# ---------------------------------------------------------------------------- config 2
def host_info():
    """Host cores this process may use and the CPU model (what lscpu reports as Model name)."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return cores, model


def run_c2(args):
    import numpy as np
    import torch

    import jagua_amd as J
    from jagua_amd import synth as S

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    log("generating layout")
    scene = S.config2_scene()
    log(f"layout: {len(scene.items)} items, {len(scene.holes)} holes, {scene.n_edges} edges; building engine")
    cde, _ = S.build_cde(scene)
    cde.sync()
    st = cde.stats()
    log(f"engine: {st['n_nodes']} nodes, {st['n_entries']} entries, {st['n_edge_refs']} edge refs, "
        f"{st['device_bytes'] / 1e6:.1f} MB in HBM")
    polys, dts = S.config2_queries(n_items=args.items, n_per_item=args.per_item)
    items = S.make_items(polys)
    mats_np = [J.compose(d) for d in dts]
    total = sum(len(m) for m in mats_np)
    offs = np.cumsum([0] + [len(m) for m in mats_np])
    mats_all = torch.from_numpy(np.concatenate(mats_np)).to(dev)
    status = torch.empty(total, dtype=torch.uint8, device=dev)
    sets = [(cde, it, int(offs[i]), int(offs[i + 1] - offs[i]), None) for i, it in enumerate(items)]
    if args.schedule:
        sched = json.loads(args.schedule)
        plan = J.QueryPlan(sets, **{k: sched[k] for k in ("chunk", "workgroups", "lanes_per_query", "stages")
                                    if k in sched})
        plan.detect(mats_all, status)
        for _ in range(args.warmup):
            plan.detect(mats_all, status)
        ms = timed(lambda: plan.detect(mats_all, status), args.steps, stream)
        tune = {"schedule": plan.config(), "autotuned": False}
    else:
        plan = J.QueryPlan(sets)
        ms, tune = tune_and_time(plan, mats_all, status, args, stream)
        tune["autotuned"] = True
    qps = total / (ms / 1e3)
    status_np = status.cpu().numpy()
    log(f"batched: {ms:.3f} ms/step, {qps / 1e6:.1f} M queries/s, collision fraction {status_np.mean():.4f}")
    launch_ms = timed(lambda: plan.detect(mats_all, status), 5, stream)
    schedule = {k: v for k, v in tune["schedule"].items() if k != "flags"}
    if args.profile == "collect":
        c_, h_, k_ = plan.collect(mats_all, capacity=8)
        ms_c = timed(lambda: plan.collect(mats_all, capacity=8, count=c_, hash=h_, keys=k_), args.steps, stream)
        log(f"profile: {args.steps} batched collections, {ms_c:.3f} ms each")
        return {"profile": "collect", "ms": ms_c, "steps": args.steps}
    if args.profile == "overlap":
        out = overlap_c4(stream)
        return {"profile": "overlap", **out}

    result = base_result(args, {
        "workload": "single large layout: 4000x1500 rectangle + 5 holes, 2000 polygons (8-256 vertices), "
                    f"quadtree depth 8; {args.items} query items (8-512 vertices) x {args.per_item} transforms",
        "form": "batched persistent kernels (surrogate stage, polygon stage)", "queries_per_step": total,
        "layout_items": len(scene.items), **tune}, total, ms)

    cap = 8
    cnt = torch.empty(total, dtype=torch.int32, device=dev)
    hsh = torch.empty(total, dtype=torch.int64, device=dev)
    keys = torch.empty((total, cap), dtype=torch.int32, device=dev)
    if not args.no_extra:
        # plain forms: the same wave kernels, one plan per (layout, item) set launched one after
        # another (like for like), and the one-thread-per-query kernel per set
        status_plain = torch.empty_like(status)
        set_plans = [J.QueryPlan([(cde, it, 0, int(offs[i + 1] - offs[i]), None)], **schedule)
                     for i, it in enumerate(items)]

        def plain_wave_step():
            for i, p in enumerate(set_plans):
                p.detect(mats_all[offs[i]:offs[i + 1]], status_plain[offs[i]:offs[i + 1]])

        plain_wave_step()
        ms_plain_wave = timed(plain_wave_step, args.steps, stream)
        plain_wave_ok = bool(torch.equal(status_plain, status))
        status_scalar = torch.empty_like(status)

        def plain_scalar_step():
            for i, it in enumerate(items):
                cde.detect_batch(it, mats_all[offs[i]:offs[i + 1]], out=status_scalar[offs[i]:offs[i + 1]])

        plain_scalar_step()
        ms_scalar = timed(plain_scalar_step, 3, stream)
        scalar_ok = bool(torch.equal(status_scalar, status))
        status_graph = torch.empty_like(status)
        plan.detect_graph(mats_all, status_graph, replays=1)
        ms_graph = timed(lambda: plan.detect_graph(mats_all, status_graph, replays=1), args.steps, stream)
        graph_ok = bool(torch.equal(status_graph, status))
        # hazard sets of the same step (collect_surrogate_collisions U collect_poly_collisions,
        # ci_bench.rs:56-69): count, hash and the first 8 keys per query, all sets in one
        # persistent launch; the per-query kernel (one launch per item) beside it
        plan.collect(mats_all, capacity=cap, count=cnt, hash=hsh, keys=keys)
        ms_collect = timed(lambda: plan.collect(mats_all, capacity=cap, count=cnt, hash=hsh, keys=keys),
                           args.steps, stream)
        cnt_p = torch.empty_like(cnt)
        hsh_p = torch.empty_like(hsh)
        keys_p = torch.empty_like(keys)

        def collect_plain_step():
            for i, it in enumerate(items):
                a, b = int(offs[i]), int(offs[i + 1])
                cde.collect_batch(it, mats_all[a:b], capacity=cap, out=(cnt_p[a:b], hsh_p[a:b], keys_p[a:b]))

        collect_plain_step()
        ms_collect_plain = timed(collect_plain_step, 2, stream)
        collect_forms_agree = bool(torch.equal(cnt, cnt_p) and torch.equal(hsh, hsh_p) and torch.equal(keys, keys_p))
        log(f"collect: batched {ms_collect:.3f} ms/step ({total / (ms_collect / 1e3) / 1e6:.1f} M queries/s), "
            f"per-query kernel {ms_collect_plain:.3f} ms; same sets: {collect_forms_agree}")
        result["forms"] = {
            "plain_queries_per_s": total / (ms_plain_wave / 1e3), "plain_ms_per_step": ms_plain_wave,
            "plain_form": f"the batched kernels, one plan per (layout, item) set, {len(items)} plans launched "
                          "one after another",
            "batched_over_plain": ms_plain_wave / ms, "plain_matches_batched": plain_wave_ok,
            "plain_scalar_queries_per_s": total / (ms_scalar / 1e3), "plain_scalar_ms_per_step": ms_scalar,
            "batched_over_plain_scalar": ms_scalar / ms, "plain_scalar_matches_batched": scalar_ok,
            "batched_queries_per_s": qps, "graph_queries_per_s": total / (ms_graph / 1e3),
            "graph_ms_per_step": ms_graph, "graph_matches_batched": graph_ok,
            "collect_queries_per_s": total / (ms_collect / 1e3), "collect_ms_per_step": ms_collect,
            "collect_plain_queries_per_s": total / (ms_collect_plain / 1e3),
            "collect_plain_ms_per_step": ms_collect_plain, "collect_forms_agree": collect_forms_agree,
            "collect_over_collect_plain": ms_collect_plain / ms_collect}
        cprof = pmc_profile("pmc_collect_config2.md")
        if cprof:
            ck = cprof["counters"]["k_collect_units"]
            result["forms"]["collect_profile"] = {
                "bound": cprof.get("bound"), "traffic": cprof.get("traffic"), "lane_util": ck["lane_util"],
                "valu_issue": ck["valu_issue"], "wait_share": ck["wait_share"], "file": "profiles/pmc_collect_config2.md"}
        log(f"plain (per-set plans) {ms_plain_wave:.3f} ms/step, scalar {ms_scalar:.3f} ms, graph {ms_graph:.3f} ms; "
            f"same verdicts: {plain_wave_ok and scalar_ok and graph_ok}")
        result["overlap"] = overlap_c4(stream)
        result["search"] = search_c2(cde, items, stream)
        result["update"] = update_c2(cde, scene, st)
        result["surrogates"] = surrogates_piers()

    roof, cpu = None, None
    if not args.no_cpu:
        from oracle import oracle as O
        cores, model = host_info()
        threads = args.cpu_threads or cores
        log(f"oracle: engine, traversal counters, CPU baseline ({threads} threads, {model})")
        oc = O.OracleCDE(scene)
        oitems = [O.OracleItem(p) for p in polys]
        tot_bytes = 0
        cpu_sec = 0.0
        n_match = 0
        for i, (oi, m) in enumerate(zip(oitems, mats_np)):
            r = O.query(oc, oi, mats=m, mode=1, threads=threads, counters=True)
            n_poly = int((r["surr"] == 0).sum())
            tot_bytes += algorithmic_bytes(r["counters"], len(m), n_poly, len(polys[i]), 2)
            n_match += int((r["status"] == status_np[offs[i]:offs[i + 1]]).sum())
            cpu_sec += min(O.query(oc, oi, mats=m, mode=1, threads=threads)["seconds"] for _ in range(2))
        # one thread on the same mix: the first 1024 transforms of every item
        one_n, one_sec = 0, 0.0
        for oi, m in zip(oitems, mats_np):
            one_sec += O.query(oc, oi, mats=m[:1024], mode=1, threads=1)["seconds"]
            one_n += len(m[:1024])
        log(f"oracle agreement {n_match}/{total}; CPU {total / cpu_sec / 1e6:.2f} M queries/s on {threads} threads, "
            f"{one_n / one_sec / 1e6:.3f} M on one")
        result["bit_exact"] = n_match == total
        if not args.no_extra:
            # the hazard sets of the whole step against the oracle's (count and hash of every query)
            c_np, h_np = cnt.cpu().numpy().astype(np.uint32), hsh.cpu().numpy().view(np.uint64)
            sets_match, coll_sec = 0, 0.0
            for i, (oi, m) in enumerate(zip(oitems, mats_np)):
                r = O.query(oc, oi, mats=m, mode=2, threads=threads)
                coll_sec += r["seconds"]
                a, b = int(offs[i]), int(offs[i + 1])
                sets_match += int(((r["count"] == c_np[a:b]) & (r["hash"] == h_np[a:b])).sum())
            result["bit_exact_sets"] = sets_match == total
            result["forms"]["collect_cpu_queries_per_s"] = total / coll_sec
            log(f"hazard sets: {sets_match}/{total} equal to the oracle's; CPU collect {total / coll_sec / 1e6:.2f} "
                f"M queries/s on {threads} threads")
        achieved = tot_bytes / (launch_ms / 1e3) / 1e9
        prof = pmc_profile("pmc_batched_config2.md")
        roof = {"bound": prof.get("bound", "latency"), "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": prof.get("traffic"),
                "kernel": ("k_fused<KIND_CIRCLE> (one batched step)" if tune["schedule"]["stages"] == 1 else
                           "k_surrogate<false, KIND_CIRCLE> + k_poly (one batched step, two launches)"),
                "launch_ms": launch_ms, "algorithmic_bytes_per_launch": tot_bytes,
                "counters": prof.get("counters"), "counter_schedule": prof.get("schedule")}
        cpu = {"value": total / cpu_sec, "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"the full step ({total} detect queries) on the C++ oracle, {threads} threads, best of 2",
               "one_thread_queries_per_s": one_n / one_sec,
               "one_thread_sample": f"the first 1024 transforms of each of the {len(oitems)} items ({one_n} queries)",
               "host_cores": cores, "cpu_model": model}
        result["vs_cpu_all_cores"] = qps / cpu["value"]
    result["roofline"] = roof
    result["cpu_baseline"] = cpu
    return result


def search_c2(cde, items, stream):
    """lbf's placement search for each of the query items on the layout, all at once on the
    GPU: 5000 samples per item (LBFConfig::default) as 8 uniform and 2 local rounds of 500,
    continuous rotation, one captured graph per run; then 32 independent searches per item
    in one run, whose rounds fill the GPU like a detection step."""
    import jagua_amd as J
    search = J.PlacementSearch([(cde, it, "continuous", None) for it in items], samples_per_round=500,
                               uniform_rounds=8, local_rounds=2, seed=1)
    search.run(graph=True)
    ms = timed(lambda: search.run(graph=True), 10, stream)
    res = search.results()
    n_samples = sum(r["samples"] for r in res)
    log(f"placement search: {len(items)} items x 5000 samples in {ms:.3f} ms, {sum(r['found'] for r in res)} placed")
    # the same searches 32 times over (2048 searches, each with its own sample stream): a round
    # is then 1,024,000 queries, the size of one batched detection step
    reps = 32
    wide = J.PlacementSearch([(cde, it, "continuous", None) for _ in range(reps) for it in items],
                             samples_per_round=500, uniform_rounds=8, local_rounds=2, seed=2)
    wide.run(graph=True)
    ms_w = timed(lambda: wide.run(graph=True), 5, stream)
    res_w = wide.results()
    n_w = sum(r["samples"] for r in res_w)
    log(f"placement search, {len(res_w)} searches: {ms_w:.3f} ms, {n_w / ms_w / 1e3:.1f} M samples/s")
    return {"searches": len(res_w), "samples_per_search": 5000, "ms_per_run": ms_w,
            "searches_per_s": len(res_w) / (ms_w / 1e3), "samples_per_s": n_w / (ms_w / 1e3),
            "placed": int(sum(r["found"] for r in res_w)), "form": "hipGraph",
            "rounds": "8 uniform + 2 local, 500 samples per search each (1,024,000 queries per round)",
            "searches_64": {"items": len(items), "samples_per_item": 5000, "ms_per_run": ms,
                            "searches_per_s": len(items) / (ms / 1e3), "samples_per_s": n_samples / (ms / 1e3),
                            "placed": int(sum(r["found"] for r in res))}}


def update_c2(cde, scene, st0, n=100):
    """The engine updates of lbf placement steps on the 2000-polygon layout: n random items
    are each removed and placed again at a random pose (deregister + register), with the
    device image brought up to date (a patch of the touched nodes and slots) after every
    change; st0 holds the full upload of the same engine for comparison."""
    import numpy as np

    import jagua_amd as J
    rng = np.random.default_rng(5)
    idx = rng.choice(len(scene.items), n, replace=False)
    (x0, y0), (x1, y1) = scene.container.min(0), scene.container.max(0)
    poses = np.stack([rng.uniform(0, 2 * np.pi, n), rng.uniform(x0, x1, n), rng.uniform(y0, y1, n)], 1)
    poses = poses.astype(np.float32)
    shapes = [J.SPolygon.new(scene.items[i][0]).transform_clone(tuple(map(float, poses[j]))) for j, i in enumerate(idx)]
    full0 = cde.stats()["n_full_uploads"]
    t_host, t_sync, patch_ms, patch_bytes = 0.0, 0.0, [], []
    for j, i in enumerate(idx):
        old = J.HazardEntity.placed_item(int(i), scene.items[i][1], int(i))
        new = J.HazardEntity.placed_item(int(i), tuple(map(float, poses[j])), 100000 + int(i))
        for change in (lambda: cde.deregister_hazard_by_entity(old),
                       lambda: cde.register_hazard(J.Hazard(new, shapes[j], True))):
            t0 = time.perf_counter()
            change()
            t1 = time.perf_counter()
            cde.sync()
            t2 = time.perf_counter()
            t_host += t1 - t0
            t_sync += t2 - t1
            s = cde.stats()
            patch_ms.append(s["patch_ms"])
            patch_bytes.append(s["patch_bytes"])
    full = cde.stats()["n_full_uploads"] - full0
    k = 2 * n
    out = {"changes": k, "quadtree_ms_per_change": t_host * 1e3 / k, "sync_ms_per_change": t_sync * 1e3 / k,
           "patch_ms_median": float(np.median(patch_ms)), "patch_bytes_median": float(np.median(patch_bytes)),
           "full_uploads": int(full), "full_upload_ms": st0["flatten_ms"] + st0["upload_ms"],
           "image_bytes": st0["device_bytes"]}
    log(f"update: {k} changes, quadtree {out['quadtree_ms_per_change']:.3f} ms + sync "
        f"{out['sync_ms_per_change']:.3f} ms per change (median patch {out['patch_bytes_median'] / 1e3:.1f} kB); "
        f"full upload {out['full_upload_ms']:.1f} ms; {full} full uploads")
    return out


def surrogates_piers(n=512, n_sample=4):
    """SPolygon::new + generate_surrogate with 2 piers (sp_surrogate.rs:33-53, piers.rs:25-98)
    for n star polygons of 8-256 vertices: one batch with pole and pier searches on the GPU,
    against the host routine (one thread) timed on the first n_sample polygons."""
    import jagua_amd as J
    from jagua_amd import synth as S
    ps = [S.polygon(50_000 + i, 8, 256, 5.0, 30.0, i % 3) for i in range(n)]
    cfg = J.SPSurrogateConfig(n_ff_poles=2, n_ff_piers=2)
    cfg0 = J.SPSurrogateConfig(n_ff_poles=2, n_ff_piers=0)
    J.SPolygon.new_batch(ps[:4], cfg)
    t0 = time.perf_counter()
    got, n_host = J.SPolygon.new_batch(ps, cfg)
    t1 = time.perf_counter()
    J.SPolygon.new_batch(ps, cfg0)
    t2 = time.perf_counter()
    ref = [J.SPolygon.new(p).generate_surrogate(cfg) for p in ps[:n_sample]]
    t3 = time.perf_counter()
    same = all((g.piers.view("u4") == r.piers.view("u4")).all() for g, r in zip(got, ref))
    out = {"polygons": n, "n_ff_piers": 2, "batch_s": t1 - t0, "batch_without_piers_s": t2 - t1,
           "host_s_per_polygon": (t3 - t2) / n_sample, "host_sample": n_sample, "host_completed": int(n_host),
           "sample_bit_exact": bool(same)}
    log(f"surrogates with piers: {n} polygons in {out['batch_s']:.3f} s on the GPU "
        f"({out['batch_without_piers_s']:.3f} s without piers); host routine {out['host_s_per_polygon']:.3f} s "
        f"per polygon; sample bit-exact: {same}")
    return out


def overlap_work(poles, off, pairs, n_items, n_cedges):
    """Algorithmic flops and bytes of one forward and one backward evaluation of the overlap
    quantity over candidate pairs and container edges (what k_ov_* compute, counted per pole
    pair and per pole-edge test, independent of which pole pairs overlap): forward 9 flops per
    pole pair (offset 4, squared distance 3, radius sum and square 2) and 6 per pole rotation;
    backward 6 more per pole pair for the gradient terms; per pole and container edge 14
    (clamped projection and distance).  Bytes: pair indices, both poses, both items' poles
    (16 B each), the partial or the 6 gradient floats."""
    import numpy as np
    n = np.diff(off.astype(np.int64))
    na, nb = n[pairs[:, 0]], n[pairs[:, 1]]
    pp = int((na * nb).sum())
    rot = int((na + nb).sum())
    total_poles = int(n.sum())
    flops_fwd = 9 * pp + 6 * rot + (6 + 14 * n_cedges) * total_poles
    flops_bwd = 15 * pp + 6 * rot + (6 + 20 * n_cedges) * total_poles
    bytes_fwd = int(len(pairs)) * (8 + 24 + 4) + 16 * rot + n_items * (12 + 4) + 16 * total_poles
    bytes_bwd = int(len(pairs)) * (8 + 24 + 24) + 16 * rot + n_items * (12 + 12 + 12) + 16 * total_poles
    return {"pole_pairs": pp, "flops_fwd": flops_fwd, "flops_bwd": flops_bwd, "bytes_fwd": bytes_fwd,
            "bytes_bwd": bytes_bwd}


FP32_PEAK_TFLOPS = 157.3


def overlap_c4(stream):
    import torch

    import jagua_amd as J
    from jagua_amd import synth as S
    log("overlap workload: 4096 items x 16-64 poles, pair candidates from the engine's quadtree")
    w = S.config4_workload()
    cde, keys = S.overlap_engine(w)
    ov = J.PoleOverlap(w["poles"], w["off"], w["pairs"][:0], w["cedges"])
    pose = torch.from_numpy(w["pose"]).cuda()
    n_pairs = ov.pairs_from_cde(cde, keys, pose, w["rad"])
    pairs = ov.pairs()
    qbuf = torch.empty(1, dtype=torch.float32, device="cuda")
    gbuf = torch.empty((len(w["pose"]), 3), dtype=torch.float32, device="cuda")
    q1, g1 = torch.empty_like(qbuf), torch.empty_like(gbuf)

    def ov_step():   # one evaluation: Q and its gradient in one pass over the pairs
        ov.forward_backward(pose, qbuf, gbuf)

    ov_step()
    ov.forward(pose, q1)
    ov.backward(pose, g1)
    torch.cuda.synchronize()
    one_pass_ok = bool(torch.equal(q1, qbuf) and torch.equal(g1, gbuf))
    ms_fwd = timed(lambda: ov.forward(pose, q1), 50, stream)
    ms_bwd = timed(lambda: ov.backward(pose, g1), 50, stream)
    ms_sep = timed(lambda: (ov.forward(pose, q1), ov.backward(pose, g1)), 50, stream)
    ms_ov = timed(ov_step, 50, stream)
    # the evaluation as one graph, replayed
    side = torch.cuda.Stream()
    qg, gg = torch.empty_like(qbuf), torch.empty_like(gbuf)
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        ov.forward_backward(pose, qg, gg, stream=side)
        side.synchronize()
        with J.Graph.capture(side) as graph:
            ov.forward_backward(pose, qg, gg, stream=side)
    graph.launch(side)
    side.synchronize()
    graph_ok = bool(torch.equal(qg, qbuf) and torch.equal(gg, gbuf))
    ms_graph = timed(lambda: graph.launch(side), 50, side)
    work = overlap_work(w["poles"], w["off"], pairs, len(w["pose"]), len(w["cedges"]))
    tf_fwd = work["flops_fwd"] / (ms_fwd / 1e3) / 1e12
    tf_bwd = work["flops_bwd"] / (ms_bwd / 1e3) / 1e12
    tf_eval = work["flops_bwd"] / (ms_ov / 1e3) / 1e12
    log(f"overlap: {n_pairs} quadtree pairs (grid {len(w['pairs'])}); one pass {ms_ov * 1e3:.1f} us "
        f"({1e3 / ms_ov:.0f} evals/s, {tf_eval:.2f} TFLOP/s, equal to the separate passes: {one_pass_ok}); "
        f"fwd {ms_fwd:.4f} ms ({tf_fwd:.2f} TFLOP/s), bwd {ms_bwd:.4f} ms ({tf_bwd:.2f} TFLOP/s); "
        f"as one graph {1e3 / ms_graph:.0f} evals/s (equal: {graph_ok})")
    return {"evals_per_s": 1e3 / ms_ov, "ms_per_eval": ms_ov, "form": "forward and backward in one pass "
            "(jg_overlap_fwd_bwd: two launches)", "one_pass_matches_separate": one_pass_ok,
            "fwd_ms": ms_fwd, "bwd_ms": ms_bwd, "separate_ms_per_eval": ms_sep,
            "graph_evals_per_s": 1e3 / ms_graph, "graph_ms_per_eval": ms_graph, "graph_matches_eager": graph_ok,
            "n_items": int(len(w["pose"])), "n_poles": int(len(w["poles"])), "n_pairs": int(n_pairs),
            "pairs": "engine quadtree, on the GPU", "dtype": "f32",
            "roofline": {"bound": "valu", "unit": "TFLOP/s", "peak": FP32_PEAK_TFLOPS,
                         "achieved_eval": tf_eval, "frac_eval": tf_eval / FP32_PEAK_TFLOPS,
                         "flops_eval": work["flops_bwd"],
                         "achieved_fwd": tf_fwd, "achieved_bwd": tf_bwd, "frac_fwd": tf_fwd / FP32_PEAK_TFLOPS,
                         "frac_bwd": tf_bwd / FP32_PEAK_TFLOPS,
                         "gbs_fwd": work["bytes_fwd"] / (ms_fwd / 1e3) / 1e9,
            "gbs_bwd": work["bytes_bwd"] / (ms_bwd / 1e3) / 1e9, **work}}


# ---------------------------------------------------------------------------- config 3
def run_c3(args, summary=False):
    import numpy as np
    import torch

    import jagua_amd as J
    from jagua_amd import synth as S

    stream = torch.cuda.current_stream()
    n_layouts = args.layouts
    log(f"generating {n_layouts} layouts")
    scenes, queries, filters = S.config3_workload(n_layouts=n_layouts, n_per_item=args.c3_per_item)
    log(f"layouts: {sum(len(s.items) for s in scenes)} items, {sum(s.n_edges for s in scenes)} edges; building engines")
    cdes = [S.build_cde(sc)[0] for sc in scenes]
    for c in cdes:
        c.sync()
    polys = [p for ps, _ in queries for p in ps]
    log(f"{len(polys)} query items (pole searches on the GPU)")
    items, n_host = J.SPolygon.new_batch(polys)
    mats_np = [J.compose(d) for _, ds in queries for d in ds]
    total = sum(len(m) for m in mats_np)
    offs = np.cumsum([0] + [len(m) for m in mats_np])
    mats = torch.from_numpy(np.concatenate(mats_np)).cuda()
    status = torch.empty(total, dtype=torch.uint8, device="cuda")
    per = len(queries[0][0])
    sets = [(cdes[k // per], items[k], int(offs[k]), int(offs[k + 1] - offs[k]), filters[k // per])
            for k in range(len(items))]
    plan = J.QueryPlan(sets)
    if summary:
        plan.detect(mats, status)
        ms = timed(lambda: plan.detect(mats, status), 3, stream)
        tune = {"schedule": plan.config(), "autotuned": False}
    else:
        ms, tune = tune_and_time(plan, mats, status, args, stream)
        tune["autotuned"] = True
    st_np = status.cpu().numpy()
    log(f"batched: {ms:.3f} ms/step, {total / ms / 1e3:.1f} M queries/s, collision fraction {st_np.mean():.4f}")
    result = base_result(args, {
        "workload": f"{n_layouts} layouts (star containers 64-2048 vertices, 50-400 items of 8-256 vertices, depth 7),"
                    f" {per} query items per layout x {args.c3_per_item} transforms, per-query random 10% HazardFilter",
        "form": "batched persistent kernels, one launch for all layouts", "queries_per_step": total, **tune},
        total, ms)
    status_plain = torch.empty_like(status)

    def plain_step():
        for k, (c, it, a, n, f) in enumerate(sets):
            c.detect_batch(it, mats[a:a + n], f, out=status_plain[a:a + n])

    plain_step()
    ms_plain = timed(plain_step, 1 if summary else 2, stream)
    sg = torch.empty_like(status)
    plan.detect_graph(mats, sg, replays=1)
    ms_graph = timed(lambda: plan.detect_graph(mats, sg, replays=1), args.steps, stream)
    result["forms"] = {"plain_queries_per_s": total / (ms_plain / 1e3), "plain_ms_per_step": ms_plain,
                       "graph_queries_per_s": total / (ms_graph / 1e3), "graph_ms_per_step": ms_graph,
                       "batched_over_plain": ms_plain / ms, "plain_matches_batched": bool(torch.equal(status_plain, status)),
                       "graph_matches_batched": bool(torch.equal(sg, status)), "items_searched_on_host": n_host}
    log(f"plain {ms_plain:.1f} ms ({len(sets)} launches), graph {ms_graph:.3f} ms")
    if not args.no_cpu:
        from oracle import oracle as O
        threads = args.cpu_threads or host_info()[0]
        k_lay = max(1, min(n_layouts, 4 if summary else 16))
        log(f"oracle sample: {k_lay} layouts x {per} items x {args.c3_per_item} transforms")
        sec, n, n_match = 0.0, 0, 0
        for l in range(k_lay):
            oc = O.OracleCDE(scenes[l])
            of = O.make_filter(2, filters[l].seed, filters[l].threshold)
            for j in range(per):
                k = l * per + j
                r = O.query(oc, O.OracleItem(polys[k]), mats=mats_np[k], mode=1, filt=of, threads=threads)
                sec += r["seconds"]
                n += len(mats_np[k])
                n_match += int((r["status"] == st_np[offs[k]:offs[k + 1]]).sum())
        result["cpu_baseline"] = {"value": n / sec, "unit": "queries/s", "cores": threads, "kind": "port",
                                  "sample": f"{n} queries ({k_lay} layouts) on the C++ oracle, {threads} threads"}
        result["bit_exact_sample"] = {"queries": n, "matching": n_match}
        log(f"oracle agreement {n_match}/{n}; CPU {n / sec / 1e6:.2f} M queries/s")
    result["roofline"] = None
    return result


# ---------------------------------------------------------------------------- config 5
def run_c5(args, summary=False):
    """The full-scale workload: batched detection of 16.8M queries, their hazard sets
    (batched collection) and one overlap-gradient pass over all placed items (32 poles each,
    pair candidates from the engine's quadtree on the GPU), the three captured in one HIP
    graph and replayed.  With --c5-oracle, a sample of the queries (at least 1M) and the
    gradient are checked against the oracle (its engine takes minutes to build)."""
    import numpy as np
    import torch

    import jagua_amd as J
    from jagua_amd import synth as S

    stream = torch.cuda.current_stream()
    log("c5: generating the 6000x2000 layout (about 20000 concave items)")
    t0 = time.time()
    scene = S.config5_scene()
    log(f"c5 layout: {len(scene.items)} items, {len(scene.holes)} holes, {scene.n_edges} edges "
        f"({time.time() - t0:.0f} s); building engine")
    t0 = time.time()
    cde, lay_keys = S.build_cde(scene)
    cde.sync()
    st = cde.stats()
    log(f"c5 engine ({time.time() - t0:.0f} s): {st['n_nodes']} nodes, {st['n_edge_refs']} edge refs, "
        f"{st['device_bytes'] / 1e6:.0f} MB in HBM")
    polys, dts = S.config5_queries(n_items=args.c5_items, n_per_item=args.c5_per_item)
    t0 = time.time()
    items, n_host = J.SPolygon.new_batch(polys)
    log(f"c5: {len(items)} query items in {time.time() - t0:.1f} s ({n_host} searched on the host)")
    mats_np = [J.compose(d) for d in dts]
    total = sum(len(m) for m in mats_np)
    offs = np.cumsum([0] + [len(m) for m in mats_np])
    mats = torch.from_numpy(np.concatenate(mats_np)).cuda()
    status = torch.empty(total, dtype=torch.uint8, device="cuda")
    sets = [(cde, it, int(offs[i]), int(offs[i + 1] - offs[i]), None) for i, it in enumerate(items)]
    plan = J.QueryPlan(sets)
    steps = 3 if summary else args.steps
    if summary:
        plan.detect(mats, status)
        ms = timed(lambda: plan.detect(mats, status), steps, stream)
        tune = {"schedule": plan.config(), "autotuned": False}
    else:
        ms, tune = tune_and_time(plan, mats, status, args, stream)
        tune["autotuned"] = True
    log(f"c5 batched detect: {ms:.3f} ms/step, {total / ms / 1e3:.1f} M queries/s, "
        f"collision fraction {status.float().mean().item():.4f}")
    cap = 8
    cnt = torch.empty(total, dtype=torch.int32, device="cuda")
    hsh = torch.empty(total, dtype=torch.int64, device="cuda")
    keys = torch.empty((total, cap), dtype=torch.int32, device="cuda")
    plan.collect(mats, capacity=cap, count=cnt, hash=hsh, keys=keys)
    ms_collect = timed(lambda: plan.collect(mats, capacity=cap, count=cnt, hash=hsh, keys=keys), steps, stream)
    log(f"c5 batched collect: {ms_collect:.1f} ms, {total / ms_collect / 1e3:.1f} M queries/s")
    # overlap gradient over all placed items, 32 poles each (pole searches on the GPU), pair
    # candidates from the engine's quadtree on the GPU
    cfg32 = J.SPSurrogateConfig(((32, 0.0), (32, 0.0), (32, 0.0)), n_ff_poles=2)
    t0 = time.time()
    lay, n_host2 = J.SPolygon.new_batch([v for v, _ in scene.items], cfg32)
    t_poles = time.time() - t0
    log(f"c5: {len(lay)} layout items x 32 poles in {t_poles:.1f} s ({n_host2} on the host)")
    poles = [p.poles for p in lay]
    off = np.zeros(len(poles) + 1, dtype=np.uint32)
    off[1:] = np.cumsum([len(p) for p in poles])
    # overlap pose (x, y, theta) from each placed item's DTransformation (rotation, tx, ty),
    # jittered as one step of an optimiser would move the items (the placed layout itself has
    # no overlap, so its Q and gradient are zero)
    pose = np.stack([(dt[1], dt[2], dt[0]) for _, dt in scene.items]).astype(np.float64)
    jit = np.random.default_rng(55)
    pose[:, :2] += jit.normal(0.0, 1.0, (len(pose), 2))
    pose[:, 2] += jit.normal(0.0, 0.02, len(pose))
    pose = pose.astype(np.float32)
    rad = np.array([np.max(np.hypot(p[:, 0], p[:, 1]) + p[:, 2]) for p in poles], dtype=np.float32)
    cedges = np.array([[0, 0, 6000, 0], [6000, 0, 6000, 2000], [6000, 2000, 0, 2000], [0, 2000, 0, 0]], np.float32)
    ov = J.PoleOverlap(np.concatenate(poles), off, np.zeros((0, 2), np.uint32), cedges)
    pose_t = torch.from_numpy(pose).cuda()
    t0 = time.time()
    n_pairs = ov.pairs_from_cde(cde, lay_keys, pose_t, rad)
    t_pairs = time.time() - t0
    qb = torch.empty(1, dtype=torch.float32, device="cuda")
    gb = torch.empty((len(pose), 3), dtype=torch.float32, device="cuda")
    ov.forward_backward(pose_t, qb, gb)
    ms_grad = timed(lambda: ov.forward_backward(pose_t, qb, gb), 20, stream)
    log(f"c5 gradient pass: {ms_grad:.3f} ms ({n_pairs} candidate pairs from the quadtree in {t_pairs:.2f} s), "
        f"Q = {qb.item():.4g}")
    # the three in one graph: captured on a side stream after one eager call each
    side = torch.cuda.Stream()
    status_g = torch.empty_like(status)
    cnt_g, hsh_g, keys_g = torch.empty_like(cnt), torch.empty_like(hsh), torch.empty_like(keys)
    qg, gg = torch.empty_like(qb), torch.empty_like(gb)
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        plan.detect(mats, status_g, stream=side)
        plan.collect(mats, capacity=cap, count=cnt_g, hash=hsh_g, keys=keys_g, stream=side)
        ov.forward_backward(pose_t, qg, gg, stream=side)
        side.synchronize()
        with J.Graph.capture(side) as graph:
            plan.detect(mats, status_g, stream=side)
            plan.collect(mats, capacity=cap, count=cnt_g, hash=hsh_g, keys=keys_g, stream=side)
            ov.forward_backward(pose_t, qg, gg, stream=side)
    graph_ok = True
    for _ in range(2):
        status_g.fill_(255)
        cnt_g.fill_(-1)
        torch.cuda.synchronize()
        graph.launch(side)
        side.synchronize()
        graph_ok = graph_ok and bool(torch.equal(status_g, status) and torch.equal(cnt_g, cnt) and
                                     torch.equal(hsh_g, hsh) and torch.equal(keys_g, keys) and
                                     torch.equal(qg, qb) and torch.equal(gg, gb))
    ms_graph = timed(lambda: graph.launch(side), steps, side)
    log(f"c5 one graph (detect + collect + gradient): {ms_graph:.1f} ms; replays equal the eager calls: {graph_ok}")
    result = base_result(args, {
        "workload": f"6000x2000 rectangle + {len(scene.holes)} holes, {len(scene.items)} concave items (64-1024 "
                    f"vertices), depth 10; {len(items)} query items x {args.c5_per_item} transforms; collection; "
                    "overlap gradient over all items (32 poles each)",
        "form": "batched detect + batched collect + overlap forward/backward, one HIP graph",
        "queries_per_step": total, **tune}, total, ms)
    result["collect"] = {"queries_per_s": total / (ms_collect / 1e3), "ms": ms_collect, "keys_capacity": cap}
    result["overlap"] = {"ms_per_eval": ms_grad, "evals_per_s": 1e3 / ms_grad, "n_items": len(pose),
                         "n_poles": int(off[-1]), "n_pairs": int(n_pairs), "pairs": "engine quadtree, on the GPU",
                         "pair_search_s": t_pairs, "pole_search_s": t_poles}
    result["graph"] = {"ms": ms_graph, "replays_match_eager": graph_ok,
                       "queries_per_s": total / (ms_graph / 1e3)}
    result["full_pass_ms"] = ms_graph
    result["layout_items"] = len(scene.items)
    result["roofline"] = None
    result["cpu_baseline"] = None
    if args.c5_oracle and not summary:
        from oracle import oracle as O
        cores, model = host_info()
        threads = args.cpu_threads or cores
        t0 = time.time()
        oc = O.OracleCDE(scene)
        log(f"c5 oracle engine in {time.time() - t0:.0f} s")
        k_items = max(1, min(len(items), (1 << 20) // args.c5_per_item + (1 if (1 << 20) % args.c5_per_item else 0)))
        st_np = status.cpu().numpy()
        c_np, h_np = cnt.cpu().numpy().astype(np.uint32), hsh.cpu().numpy().view(np.uint64)
        n, n_det, n_sets, sec = 0, 0, 0, 0.0
        for i in range(k_items):
            oi = O.OracleItem(polys[i])
            r = O.query(oc, oi, mats=mats_np[i], mode=3, threads=threads)
            a, b = int(offs[i]), int(offs[i + 1])
            n += b - a
            n_det += int((r["status"] == st_np[a:b]).sum())
            n_sets += int(((r["count"] == c_np[a:b]) & (r["hash"] == h_np[a:b])).sum())
            sec += r["seconds"]
            log(f"c5 oracle item {i + 1}/{k_items}: {n_det}/{n} verdicts, {n_sets}/{n} sets equal")
        q64, g64 = O.overlap64(pose.astype(np.float64), np.concatenate(poles), off, ov.pairs(), cedges)
        g = gb.cpu().numpy().astype(np.float64)
        result["bit_exact_sample"] = {"queries": n, "verdicts_matching": n_det, "sets_matching": n_sets,
                                      "items": k_items}
        result["overlap"]["q_fp64"] = q64
        result["overlap"]["q_rel_err_fp64"] = abs(qb.item() - q64) / max(q64, 1e-300)
        result["overlap"]["grad_rel_err_fp64"] = float(np.linalg.norm(g - g64) / np.linalg.norm(g64))
        result["cpu_baseline"] = {"value": n / sec, "unit": "queries/s", "cores": threads, "kind": "port",
                                  "sample": f"{n} queries (detect and collect) on the C++ oracle, {threads} threads",
                                  "cpu_model": model}
        log(f"c5 oracle: verdicts {n_det}/{n}, sets {n_sets}/{n}; Q rel err {result['overlap']['q_rel_err_fp64']:.2e}, "
            f"grad {result['overlap']['grad_rel_err_fp64']:.2e}")
    return result


def summary_of(r):
    """The keys of a workload's result carried in the default line."""
    keys = ("value", "ms_per_step", "config", "forms", "collect", "overlap", "graph", "full_pass_ms",
            "bit_exact_sample", "cpu_baseline")
    return {k: r[k] for k in keys if k in r and r[k] is not None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["c2", "c3", "c5"], default="c2")
    ap.add_argument("--items", type=int, default=64)
    ap.add_argument("--per-item", type=int, default=16384)
    ap.add_argument("--layouts", type=int, default=256)
    ap.add_argument("--c3-per-item", type=int, default=4096)
    ap.add_argument("--c5-items", type=int, default=256)
    ap.add_argument("--c5-per-item", type=int, default=65536)
    ap.add_argument("--c5-oracle", action="store_true",
                    help="c5: check 1M queries and the gradient against the oracle (minutes of CPU)")
    ap.add_argument("--quick", action="store_true", help="c2 only (no many-layout and full-scale summaries)")
    ap.add_argument("--profile", choices=["", "collect", "overlap"], default="",
                    help="c2: after the timed detection steps, run only `steps` batched collections "
                         "(collect) or overlap evaluations (overlap), for counter runs of those kernels")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores")
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle (CPU baseline, bit-exact check)")
    ap.add_argument("--no-extra", action="store_true", help="only the batched form (for profiling)")
    ap.add_argument("--schedule", default="",
                    help='a fixed batched schedule as JSON (keys chunk, workgroups, lanes_per_query, stages), '
                         'used instead of the autotuner (profiled runs keep the timed schedule)')
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device (the known-size kernel that scales "
                         "FETCH_SIZE / WRITE_SIZE counter readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the query path has no "
              "multi-GPU form in this build", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.calib_copy_mib:
        n = args.calib_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calib_copy_mib} MiB read + {args.calib_copy_mib} MiB written")
        del src, dst
    result = {"c2": run_c2, "c3": run_c3, "c5": run_c5}[args.workload](args)
    if args.workload == "c2" and not args.quick and not args.no_extra:
        # the many-layout and full-scale workloads, timed with their default schedules
        result["c3"] = summary_of(run_c3(args, summary=True))
        result["c5"] = summary_of(run_c5(args, summary=True))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
