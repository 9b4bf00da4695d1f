"""Writes tests/golden/config1_seed0.npz: the 1000x400 layout of 40 twelve-vertex stars
(quadtree depth 5) and one 16-vertex query star at 4096 transforms (rotations k*45 deg),
with the oracle's outputs for every transform: detect status, surrogate verdict, collected
hazard-set size, order-independent set hash and the first 32 slot keys ascending, plus the
oracle's quadtree dump.

    python tests/golden/make_config1.py

The inputs are stored explicitly, so tests that read the fixture depend neither on the
workload generator nor on the oracle: the oracle is checked against it on the CPU and the
GPU engine is checked against it through the C ABI.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "jagua-rs_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

from jagua_amd import synth as S  # noqa: E402
from oracle import oracle as O  # noqa: E402

KEYS_CAP = 32
OUT = os.path.join(HERE, "config1_seed0.npz")


def main():
    sc = S.config1_scene(0)
    polys, dts = S.config1_queries(0, n=4096)
    oc = O.OracleCDE(sc)
    oitem = O.OracleItem(polys[0])
    mats = O.compose(dts)
    r = O.query(oc, oitem, mats=mats, mode=3, keys_cap=KEYS_CAP, threads=8)
    nodes, haz = oc.dump()
    np.savez_compressed(
        OUT,
        container=sc.container.astype(np.float32),
        item_n=np.array([len(v) for v, _ in sc.items], dtype=np.uint32),
        item_verts=np.concatenate([v for v, _ in sc.items]).astype(np.float32),
        item_dt=np.stack([dt for _, dt in sc.items]).astype(np.float32),
        depth=np.uint32(sc.depth), cd_threshold=np.uint32(sc.cd_threshold),
        query=polys[0].astype(np.float32), dts=dts.astype(np.float32), mats=mats,
        status=r["status"], surr=r["surr"], count=r["count"], hash=r["hash"], keys=r["keys"],
        tree_nodes=nodes, tree_haz=haz)
    print(f"wrote {OUT}: {len(mats)} queries, {int(r['status'].sum())} colliding, "
          f"{int((r['count'] > 0).sum())} with a non-empty set, max set {int(r['count'].max())}")


if __name__ == "__main__":
    main()
