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


def pmc_traffic(name):
    """HBM bytes per launch measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this
    workload, scaled by the same counters' reading for a copy of known size; read from the
    summary committed under profiles/ (None when absent)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_step:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


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
    plan = J.QueryPlan(sets)
    ms, tune = tune_and_time(plan, mats_all, status, args, stream)
    qps = total / (ms / 1e3)
    status_np = status.cpu().numpy()
    log(f"batched: {ms:.3f} ms/step, {qps / 1e6:.1f} M queries/s, collision fraction {status_np.mean():.4f}")
    launch_ms = timed(lambda: plan.detect(mats_all, status), 5, stream)

    result = base_result(args, {
        "workload": "single large layout: 4000x1500 rectangle + 5 holes, 2000 polygons (8-256 vertices), "
                    f"quadtree depth 8; {args.items} query items (8-512 vertices) x {args.per_item} transforms",
        "form": "batched persistent kernels (surrogate stage, polygon stage)", "queries_per_step": total,
        "layout_items": len(scene.items), **tune}, total, ms)

    if not args.no_extra:
        status_plain = torch.empty_like(status)

        def plain_step():
            for i, it in enumerate(items):
                cde.detect_batch(it, mats_all[offs[i]:offs[i + 1]], out=status_plain[offs[i]:offs[i + 1]])

        plain_step()
        ms_plain = timed(plain_step, 3, stream)
        plain_ok = bool(torch.equal(status_plain, status))
        status_graph = torch.empty_like(status)
        plan.detect_graph(mats_all, status_graph, replays=1)
        ms_graph = timed(lambda: plan.detect_graph(mats_all, status_graph, replays=1), args.steps, stream)
        graph_ok = bool(torch.equal(status_graph, status))
        # hazard sets of the same step (collect_surrogate_collisions U collect_poly_collisions,
        # ci_bench.rs:56-69): count, hash and the first 8 keys per query, one launch per item
        cnt = torch.empty(total, dtype=torch.int32, device=dev)
        hsh = torch.empty(total, dtype=torch.int64, device=dev)
        keys = torch.empty((total, 8), dtype=torch.int32, device=dev)

        def collect_step():
            for i, it in enumerate(items):
                a, b = int(offs[i]), int(offs[i + 1])
                cde.collect_batch(it, mats_all[a:b], capacity=8, out=(cnt[a:b], hsh[a:b], keys[a:b]))

        collect_step()
        ms_collect = timed(collect_step, 3, stream)
        sets_agree = bool(torch.equal(cnt > 0, status.to(torch.int32) > 0))
        log(f"collect: {ms_collect:.3f} ms/step ({total / (ms_collect / 1e3) / 1e6:.1f} M queries/s); "
            f"non-empty set <=> collision on every query: {sets_agree}")
        result["forms"] = {
            "plain_queries_per_s": total / (ms_plain / 1e3), "plain_ms_per_step": ms_plain,
            "batched_queries_per_s": qps, "graph_queries_per_s": total / (ms_graph / 1e3),
            "graph_ms_per_step": ms_graph, "batched_over_plain": ms_plain / ms,
            "plain_matches_batched": plain_ok, "graph_matches_batched": graph_ok,
            "collect_queries_per_s": total / (ms_collect / 1e3), "collect_ms_per_step": ms_collect,
            "collect_sets_match_verdicts": sets_agree}
        log(f"plain: {ms_plain:.3f} ms/step; graph: {ms_graph:.3f} ms/step; same verdicts: {plain_ok and graph_ok}")
        result["overlap"] = overlap_c4(stream)
        result["search"] = search_c2(cde, items, stream)
        result["update"] = update_c2(cde, scene, st)
        result["surrogates"] = surrogates_piers()

    roof, cpu = None, None
    if not args.no_cpu:
        from oracle import oracle as O
        threads = args.cpu_threads or (os.cpu_count() or 1)
        log(f"oracle: engine, traversal counters, CPU baseline ({threads} threads)")
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
        one = O.query(oc, oitems[0], mats=mats_np[0][:4096], mode=1, threads=1)["seconds"]
        log(f"oracle agreement {n_match}/{total}; CPU {total / cpu_sec / 1e6:.2f} M queries/s on {threads} threads")
        achieved = tot_bytes / (launch_ms / 1e3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": pmc_traffic("pmc_batched_config2.md"),
                "kernel": ("k_fused<KIND_CIRCLE> (one batched step)" if tune["schedule"]["stages"] == 1 else
                           "k_surrogate<false, KIND_CIRCLE> + k_poly (one batched step, two launches)"),
                "launch_ms": launch_ms,
                "algorithmic_bytes_per_launch": tot_bytes}
        cpu = {"value": total / cpu_sec, "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"the full step ({total} detect queries) on the C++ oracle, {threads} threads, best of 2",
               "one_thread_queries_per_s": 4096 / one}
        result["bit_exact"] = n_match == total
        result["vs_cpu_all_cores"] = qps / cpu["value"]
    result["roofline"] = roof
    result["cpu_baseline"] = cpu
    return result


def search_c2(cde, items, stream):
    """lbf's placement search for each of the query items on the layout, all at once on the
    GPU: 5000 samples per item (LBFConfig::default) as 8 uniform and 2 local rounds of 500,
    continuous rotation, one captured graph per run."""
    import jagua_amd as J
    search = J.PlacementSearch([(cde, it, "continuous", None) for it in items], samples_per_round=500,
                               uniform_rounds=8, local_rounds=2, seed=1)
    search.run(graph=True)
    ms = timed(lambda: search.run(graph=True), 10, stream)
    res = search.results()
    n_samples = sum(r["samples"] for r in res)
    log(f"placement search: {len(items)} items x 5000 samples in {ms:.3f} ms, {sum(r['found'] for r in res)} placed")
    return {"items": len(items), "samples_per_item": 5000, "ms_per_run": ms, "searches_per_s": len(items) / (ms / 1e3),
            "samples_per_s": n_samples / (ms / 1e3), "placed": int(sum(r["found"] for r in res)), "form": "hipGraph"}


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


def overlap_c4(stream):
    import torch

    import jagua_amd as J
    from jagua_amd import synth as S
    log("overlap workload: 4096 items x 16-64 poles")
    w = S.config4_workload()
    ov = J.PoleOverlap(w["poles"], w["off"], w["pairs"], w["cedges"])
    pose = torch.from_numpy(w["pose"]).cuda()
    qbuf = torch.empty(1, dtype=torch.float32, device="cuda")
    gbuf = torch.empty((len(w["pose"]), 3), dtype=torch.float32, device="cuda")

    def ov_step():
        ov.forward(pose, qbuf)
        ov.backward(pose, gbuf)

    ov_step()
    ms_fwd = timed(lambda: ov.forward(pose, qbuf), 50, stream)
    ms_bwd = timed(lambda: ov.backward(pose, gbuf), 50, stream)
    ms_ov = timed(ov_step, 50, stream)
    log(f"overlap: fwd {ms_fwd:.3f} ms, bwd {ms_bwd:.3f} ms, {1e3 / ms_ov:.0f} evals/s")
    return {"evals_per_s": 1e3 / ms_ov, "ms_per_eval": ms_ov, "fwd_ms": ms_fwd, "bwd_ms": ms_bwd,
            "n_items": int(len(w["pose"])), "n_poles": int(len(w["poles"])), "n_pairs": int(len(w["pairs"])),
            "dtype": "f32"}


# ---------------------------------------------------------------------------- config 3
def run_c3(args):
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
    ms, tune = tune_and_time(plan, mats, status, args, stream)
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
    ms_plain = timed(plain_step, 2, stream)
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
        threads = args.cpu_threads or (os.cpu_count() or 1)
        k_lay = max(1, min(n_layouts, 16))
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
def run_c5(args):
    import numpy as np
    import torch

    import jagua_amd as J
    from jagua_amd import synth as S

    stream = torch.cuda.current_stream()
    log("generating the 6000x2000 layout (about 20000 concave items)")
    t0 = time.time()
    scene = S.config5_scene()
    log(f"layout: {len(scene.items)} items, {len(scene.holes)} holes, {scene.n_edges} edges "
        f"({time.time() - t0:.0f} s); building engine")
    t0 = time.time()
    cde, _ = S.build_cde(scene)
    cde.sync()
    st = cde.stats()
    log(f"engine ({time.time() - t0:.0f} s): {st['n_nodes']} nodes, {st['n_edge_refs']} edge refs, "
        f"{st['device_bytes'] / 1e6:.0f} MB in HBM")
    polys, dts = S.config5_queries(n_items=args.c5_items, n_per_item=args.c5_per_item)
    t0 = time.time()
    items, n_host = J.SPolygon.new_batch(polys)
    log(f"{len(items)} query items in {time.time() - t0:.1f} s ({n_host} searched on the host)")
    mats_np = [J.compose(d) for d in dts]
    total = sum(len(m) for m in mats_np)
    offs = np.cumsum([0] + [len(m) for m in mats_np])
    mats = torch.from_numpy(np.concatenate(mats_np)).cuda()
    del mats_np
    status = torch.empty(total, dtype=torch.uint8, device="cuda")
    sets = [(cde, it, int(offs[i]), int(offs[i + 1] - offs[i]), None) for i, it in enumerate(items)]
    plan = J.QueryPlan(sets)
    ms, tune = tune_and_time(plan, mats, status, args, stream)
    log(f"batched detect: {ms:.3f} ms/step, {total / ms / 1e3:.1f} M queries/s, "
        f"collision fraction {status.float().mean().item():.4f}")
    cap = 8
    cnt = torch.empty(total, dtype=torch.int32, device="cuda")
    hsh = torch.empty(total, dtype=torch.int64, device="cuda")
    keys = torch.empty((total, cap), dtype=torch.int32, device="cuda")

    def collect_all():
        for i, it in enumerate(items):
            a, b = int(offs[i]), int(offs[i + 1])
            c, h, k = cde.collect_batch(it, mats[a:b], capacity=cap)
            cnt[a:b], hsh[a:b], keys[a:b] = c, h, k

    collect_all()
    ms_collect = timed(collect_all, 1, stream)
    log(f"collect: {ms_collect:.1f} ms, {total / ms_collect / 1e3:.1f} M queries/s")
    # overlap gradient over all placed items, 32 poles each (pole searches on the GPU)
    cfg32 = J.SPSurrogateConfig(((32, 0.0), (32, 0.0), (32, 0.0)), n_ff_poles=2)
    t0 = time.time()
    lay, n_host2 = J.SPolygon.new_batch([v for v, _ in scene.items], cfg32)
    t_poles = time.time() - t0
    log(f"{len(lay)} layout items x 32 poles in {t_poles:.1f} s ({n_host2} on the host)")
    poles = [p.poles for p in lay]
    off = np.zeros(len(poles) + 1, dtype=np.uint32)
    off[1:] = np.cumsum([len(p) for p in poles])
    # overlap pose (x, y, theta) from each placed item's DTransformation (rotation, tx, ty)
    pose = np.stack([(dt[1], dt[2], dt[0]) for _, dt in scene.items]).astype(np.float32)
    rad = np.array([np.max(np.hypot(p[:, 0], p[:, 1]) + p[:, 2]) for p in poles], dtype=np.float32)
    pairs = J.candidate_pairs(pose, rad, [0.0, 0.0, 6000.0, 2000.0], 10)
    cedges = np.array([[0, 0, 6000, 0], [6000, 0, 6000, 2000], [6000, 2000, 0, 2000], [0, 2000, 0, 0]], np.float32)
    ov = J.PoleOverlap(np.concatenate(poles), off, pairs, cedges)
    pose_t = torch.from_numpy(pose).cuda()
    qb = torch.empty(1, dtype=torch.float32, device="cuda")
    gb = torch.empty((len(pose), 3), dtype=torch.float32, device="cuda")
    ov.forward(pose_t, qb)
    ov.backward(pose_t, gb)
    ms_grad = timed(lambda: (ov.forward(pose_t, qb), ov.backward(pose_t, gb)), 20, stream)
    log(f"gradient pass: {ms_grad:.3f} ms ({len(pairs)} candidate pairs)")
    result = base_result(args, {
        "workload": f"6000x2000 rectangle + {len(scene.holes)} holes, {len(scene.items)} concave items (64-1024 "
                    f"vertices), depth 10; {len(items)} query items x {args.c5_per_item} transforms; collection; "
                    "overlap gradient over all items (32 poles each)",
        "form": "batched detect + plain collect + overlap forward/backward", "queries_per_step": total, **tune},
        total, ms)
    result["collect"] = {"queries_per_s": total / (ms_collect / 1e3), "ms": ms_collect, "keys_capacity": cap}
    result["overlap"] = {"ms_per_eval": ms_grad, "evals_per_s": 1e3 / ms_grad, "n_items": len(pose),
                         "n_poles": int(off[-1]), "n_pairs": int(len(pairs)), "pole_search_s": t_poles}
    result["full_pass_ms"] = ms + ms_collect + ms_grad
    result["roofline"] = None
    result["cpu_baseline"] = None
    return result


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
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores")
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle (CPU baseline, bit-exact check)")
    ap.add_argument("--no-extra", action="store_true", help="only the batched form (for profiling)")
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
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
