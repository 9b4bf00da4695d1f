#!/usr/bin/env python3
"""Collision-query throughput of the MI355X engine on one GPU.

Workload (default): a single large layout — a 4000x1500 rectangle with 5 rectangular holes,
2000 placed star/concave polygons of 8-256 vertices (min separation 0.5), quadtree depth 8 —
queried by 64 items of 8-512 vertices at 16384 continuous-rotation transforms each:
1,048,576 collision queries per step (verdict = surrogate fail-fast || polygon query, the
composition of lbf/benches/ci_bench.rs:107-120).  Inputs are resident in HBM before timing.

Prints one JSON line with the throughput, a roofline entry for the dominant kernel and the
CPU oracle's throughput on the host cores as a reported baseline.
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
    """Bytes one query must touch (transform, query geometry, nodes, hazard entries and edges,
    containment vertices, verdict), from the oracle's traversal counters."""
    nodes, entries, edges, pip = (int(x) for x in counters)
    return (24 * n_queries + 12 * n_ff * n_queries + (8 * v_q + 8) * n_poly + 32 * nodes + 32 * entries
            + 16 * edges + 8 * pip + n_queries)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--form", choices=["batched", "plain"], default="batched")
    ap.add_argument("--items", type=int, default=64)
    ap.add_argument("--per-item", type=int, default=16384)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores")
    ap.add_argument("--no-cpu", action="store_true")
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
    mats_all = torch.from_numpy(np.concatenate(mats_np)).to(dev)
    status = torch.empty(total, dtype=torch.uint8, device=dev)
    offs = np.cumsum([0] + [len(m) for m in mats_np])
    stream = torch.cuda.current_stream()

    plan = None
    if args.form == "batched":
        sets = [(cde, it, int(offs[i]), int(offs[i + 1] - offs[i]), None) for i, it in enumerate(items)]
        plan = J.QueryPlan(sets)

    def step():
        if plan is not None:
            plan.detect(mats_all, status)
        else:
            for i, it in enumerate(items):
                cde.detect_batch(it, mats_all[offs[i]:offs[i + 1]], out=status[offs[i]:offs[i + 1]])

    log(f"warmup {args.warmup} steps ({args.form} form, {total} queries per step)")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record(stream)
    for _ in range(args.steps):
        step()
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / args.steps
    qps = total / (ms / 1e3)
    status_np = status.cpu().numpy()
    log(f"{ms:.3f} ms/step, {qps / 1e6:.1f} M queries/s, collision fraction {status_np.mean():.4f}")

    # per-launch durations of the dominant kernel, on the stream it runs on
    ev = []
    if plan is not None:
        for _ in range(5):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            plan.detect(mats_all, status)
            b.record(stream)
            ev.append((a, b))
        torch.cuda.synchronize()
        launch_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
        launches_per_step = 1
    else:
        per = []
        for i, it in enumerate(items):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            cde.detect_batch(it, mats_all[offs[i]:offs[i + 1]], out=status[offs[i]:offs[i + 1]])
            b.record(stream)
            per.append((a, b))
        torch.cuda.synchronize()
        launch_ms = float(np.mean([a.elapsed_time(b) for a, b in per]))
        launches_per_step = len(items)

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
        "config": {"workload": "single large layout: 4000x1500 + 5 holes, 2000 polygons (8-256 v), depth 8; "
                               f"{args.items} query items (8-512 v) x {args.per_item} transforms",
                   "form": args.form, "queries_per_step": total, "layout_items": len(scene.items)},
    }

    # algorithmic bytes from the oracle's traversal of the same queries, and the CPU baseline
    roof = None
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        threads = args.cpu_threads or (os.cpu_count() or 1)
        log(f"oracle: building engine and measuring traversal counters ({threads} threads)")
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
            t = O.query(oc, oi, mats=m, mode=1, threads=threads)
            cpu_sec += t["seconds"]
        log(f"oracle agreement {n_match}/{total}; CPU {total / cpu_sec / 1e6:.2f} M queries/s on {threads} threads")
        bytes_per_launch = tot_bytes / launches_per_step
        achieved = bytes_per_launch / (launch_ms / 1e3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "kernel": "k_detect_batched" if plan is not None else "k_detect_plain",
                "launch_ms": launch_ms, "algorithmic_bytes_per_launch": bytes_per_launch}
        cpu = {"value": total / cpu_sec, "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"the full step ({total} queries, detect) on the C++ oracle, {threads} threads"}
        result["bit_exact"] = n_match == total
    result["roofline"] = roof
    result["cpu_baseline"] = cpu
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
