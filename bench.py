#!/usr/bin/env python3
"""Collision-query throughput of the MI355X engine on one GPU.

Workload (default): a single large layout — a 4000x1500 rectangle with 5 rectangular holes,
2000 placed star/concave polygons of 8-256 vertices (min separation 0.5), quadtree depth 8 —
queried by 64 items of 8-512 vertices at 16384 continuous-rotation transforms each:
1,048,576 collision queries per step (verdict = surrogate fail-fast || polygon query, the
composition of lbf/benches/ci_bench.rs:107-120).  Inputs are resident in HBM before timing.

One step = one pass of the batched persistent kernel over all queries.  The JSON line also
carries the plain form (one launch per query set), the hipGraph-replayed form, the overlap
gradient (4096 items x 16-64 poles, forward + backward per evaluation), a roofline entry for
the dominant kernel and the CPU oracle's throughput on the host cores (a reported baseline).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "jagua-rs_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--items", type=int, default=64)
    ap.add_argument("--per-item", type=int, default=16384)
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

    import numpy as np
    import torch

    import jagua_amd as J
    from jagua_amd import synth as S

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    if args.calib_copy_mib:
        n = args.calib_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calib_copy_mib} MiB read + {args.calib_copy_mib} MiB written")
        del src, dst
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

    def step():
        plan.detect(mats_all, status)

    log(f"warmup {args.warmup} steps (batched form, {total} queries per step)")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    ms = timed(step, args.steps, stream)
    qps = total / (ms / 1e3)
    status_np = status.cpu().numpy()
    log(f"batched: {ms:.3f} ms/step, {qps / 1e6:.1f} M queries/s, collision fraction {status_np.mean():.4f}")
    # average duration of one launch of the dominant kernel, HIP events on its stream
    launch_ms = timed(step, 5, stream)

    result = {
        "metric": "bit-exact collision queries/sec on 1 MI355X, batched vs plain; overlap-grad evals/s",
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "single large layout: 4000x1500 rectangle + 5 holes, 2000 polygons (8-256 "
                               f"vertices), quadtree depth 8; {args.items} query items (8-512 vertices) x "
                               f"{args.per_item} transforms",
                   "form": "batched persistent kernel", "queries_per_step": total,
                   "layout_items": len(scene.items)},
    }

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
        result["forms"] = {
            "plain_queries_per_s": total / (ms_plain / 1e3), "plain_ms_per_step": ms_plain,
            "batched_queries_per_s": qps, "graph_queries_per_s": total / (ms_graph / 1e3),
            "graph_ms_per_step": ms_graph, "batched_over_plain": ms_plain / ms,
            "plain_matches_batched": plain_ok, "graph_matches_batched": graph_ok}
        log(f"plain: {ms_plain:.3f} ms/step; graph: {ms_graph:.3f} ms/step; same verdicts: {plain_ok and graph_ok}")

        log("overlap workload: 4096 items x 16-64 poles")
        w = S.config4_workload()
        ov = J.PoleOverlap(w["poles"], w["off"], w["pairs"], w["cedges"])
        pose = torch.from_numpy(w["pose"]).to(dev)
        qbuf = torch.empty(1, dtype=torch.float32, device=dev)
        gbuf = torch.empty((len(w["pose"]), 3), dtype=torch.float32, device=dev)

        def ov_step():
            ov.forward(pose, qbuf)
            ov.backward(pose, gbuf)

        ov_step()
        ms_fwd = timed(lambda: ov.forward(pose, qbuf), 50, stream)
        ms_bwd = timed(lambda: ov.backward(pose, gbuf), 50, stream)
        ms_ov = timed(ov_step, 50, stream)
        result["overlap"] = {"evals_per_s": 1e3 / ms_ov, "ms_per_eval": ms_ov, "fwd_ms": ms_fwd, "bwd_ms": ms_bwd,
                             "n_items": int(len(w["pose"])), "n_poles": int(len(w["poles"])),
                             "n_pairs": int(len(w["pairs"])), "dtype": "f32"}
        log(f"overlap: fwd {ms_fwd:.3f} ms, bwd {ms_bwd:.3f} ms, {1e3 / ms_ov:.0f} evals/s")

    roof = None
    cpu = None
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
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_detect_batched",
                "launch_ms": launch_ms, "algorithmic_bytes_per_launch": tot_bytes}
        cpu = {"value": total / cpu_sec, "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"the full step ({total} detect queries) on the C++ oracle, {threads} threads, best of 2",
               "one_thread_queries_per_s": 4096 / one}
        result["bit_exact"] = n_match == total
        result["vs_cpu_all_cores"] = qps / cpu["value"]
    result["roofline"] = roof
    result["cpu_baseline"] = cpu
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
