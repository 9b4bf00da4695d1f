"""Reduced forms of the many-layout and full-scale workloads against the oracle, bit for bit:
star-shaped containers of 64-2048 vertices with a per-query random 10% HazardFilter, all
layouts in one batched launch (and its hipGraph replay); depth-10 layouts of concave
64-1024-vertex polygons around star-shaped holes, with hazard collection."""
import numpy as np
import pytest
import torch

import jagua_amd as J
from jagua_amd import synth as S
from oracle import oracle as O

from helpers import engines, items

pytestmark = pytest.mark.gpu


def test_many_layouts_filtered(cuda):
    layouts = [0, 4, 9]                      # 60, 323 and 382 placed items
    sets, refs, mats_all, keep = [], [], [], []
    first = 0
    for l in layouts:
        sc = S.config3_layout(3, l)
        cde, keys, oc = engines(sc)
        polys, dts = S.config3_queries(sc, 3, l, n_items=4, n_per_item=1024)
        its, oits = items(polys)
        filt = J.HashedFilter((3 << 16) + l, 0.1)
        ofilt = O.make_filter(2, seed=filt.seed, threshold=filt.threshold)
        keep += [cde, oc, its, oits]
        for it, oit, d in zip(its, oits, dts):
            m = J.compose(d)
            ref = O.query(oc, oit, mats=m, mode=1, filt=ofilt, threads=8)["status"]
            # plain form, one launch per set
            st = cde.detect_batch(it, torch.from_numpy(m).to(cuda), filt).cpu().numpy()
            np.testing.assert_array_equal(st, ref)
            sets.append((cde, it, first, len(m), filt))
            refs.append(ref)
            mats_all.append(m)
            first += len(m)
    ref = np.concatenate(refs)
    assert 0 < ref.sum() < len(ref)
    M = torch.from_numpy(np.concatenate(mats_all)).to(cuda)
    out = torch.full((len(M),), 9, dtype=torch.uint8, device=cuda)
    plan = J.QueryPlan(sets)
    plan.detect(M, out)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
    for r in range(3):                     # graph replays with eager work in between
        out.fill_(9)
        plan.detect_graph(M, out, replays=1)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), ref, err_msg=f"replay {r}")


def test_full_scale_reduced_with_collection(cuda):
    sc = S.config5_scene(seed=6, n_items=350, width=1200.0, height=400.0, n_holes=4)
    assert sc.depth == 10 and len(sc.holes) >= 1
    cde, keys, oc = engines(sc)
    polys, dts = S.config5_queries(seed=6, n_items=3, n_per_item=2048, r_lo=3.0, r_hi=6.0, width=1200.0,
                                   height=400.0, nv_hi=256)
    its, oits = items(polys)
    cap = 24
    for it, oit, d in zip(its, oits, dts):
        m = J.compose(d)
        mt = torch.from_numpy(m).to(cuda)
        ref = O.query(oc, oit, mats=m, mode=3, keys_cap=cap, threads=8)
        np.testing.assert_array_equal(cde.detect_batch(it, mt).cpu().numpy(), ref["status"])
        cnt, hsh, ks = cde.collect_batch(it, mt, capacity=cap)
        np.testing.assert_array_equal(cnt.cpu().numpy().astype(np.uint32), ref["count"])
        np.testing.assert_array_equal(hsh.cpu().numpy().view(np.uint64), ref["hash"])
        np.testing.assert_array_equal(ks.cpu().numpy().view(np.uint32), ref["keys"])
        assert 0 < ref["status"].sum() < len(m)
        out = torch.empty(len(m), dtype=torch.uint8, device=cuda)
        J.QueryPlan([(cde, it, 0, len(m), None)]).detect(mt, out)
        np.testing.assert_array_equal(out.cpu().numpy(), ref["status"])
