"""The batched hazard collection (jg_plan_collect: one persistent launch over many layouts and
items, with the exact subtree skip and the per-query overflow collector) against the oracle's
collect_surrogate_collisions ∪ collect_poly_collisions (cd_engine.rs:261-310), bit for bit:
set sizes, set hashes and ascending keys."""
import numpy as np
import pytest
import torch

import jagua_amd as J
from jagua_amd import synth as S
from oracle import oracle as O

from helpers import engines, items

pytestmark = pytest.mark.gpu


def check_plan(plan, mats_all, expect, cap, cuda):
    m = torch.from_numpy(np.concatenate(mats_all)).to(cuda)
    cnt, hsh, keys = plan.collect(m, capacity=cap)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(cnt.cpu().numpy().astype(np.uint32), np.concatenate([e["count"] for e in expect]))
    np.testing.assert_array_equal(hsh.cpu().numpy().view(np.uint64), np.concatenate([e["hash"] for e in expect]))
    if cap:
        np.testing.assert_array_equal(keys.cpu().numpy().view(np.uint32), np.concatenate([e["keys"] for e in expect]))
    return cnt.cpu().numpy()


def test_collect_config1(cuda):
    sc = S.config1_scene(0)
    cde, keys, oc = engines(sc)
    polys, dts = S.config1_queries(0)
    (it,), (oit,) = items(polys)
    mats = J.compose(dts)
    ref = O.query(oc, oit, mats=mats, mode=2, keys_cap=8, threads=8)
    plan = J.QueryPlan([(cde, it, 0, len(mats), None)])
    cnt = check_plan(plan, [mats], [ref], 8, cuda)
    assert cnt.max() >= 2 and (cnt == 0).sum() > 0


@pytest.mark.parametrize("depth", [0, 3, 10])
def test_collect_depths(cuda, depth):
    sc = S.config1_scene(1)
    sc.depth = depth
    cde, keys, oc = engines(sc)
    polys, dts = S.config1_queries(1, n=2048)
    (it,), (oit,) = items(polys)
    mats = J.compose(dts)
    ref = O.query(oc, oit, mats=mats, mode=2, keys_cap=8, threads=8)
    check_plan(J.QueryPlan([(cde, it, 0, len(mats), None)]), [mats], [ref], 8, cuda)


def test_collect_many_layouts_items_and_seeds(cuda):
    """Three lbf layouts and a scaled configuration-2 layout, six items (one with piers, one
    large enough for sets beyond the LDS collector), pre-seeded collectors on some sets."""
    scenes = [S.config1_scene(10 + i) for i in range(3)] + [S.config2_scene(seed=11, n_items=250)]
    built = [engines(s) for s in scenes]
    polys = [S.polygon(100 + i, 8, 200, 8.0, 30.0, 2) for i in range(4)] + [S.config1_queries(3)[0][0]]
    prod, orc = items(polys)
    pier_poly = S.polygon(150, 14, 14, 12.0, 30.0, 0)
    (pit,), (opit,) = items([pier_poly], n_ff_piers=2)
    big = S.polygon(151, 64, 64, 300.0, 450.0, 0)
    (bit,), (obit,) = items([big])
    prod += [pit, bit]
    orc += [opit, obit]
    sets, expect, mats_all, first = [], [], [], 0
    for li, (cde, keys, oc) in enumerate(built):
        box = scenes[li].cde_bbox()
        for ii in range(len(prod)):
            n = 200 + 61 * ii + 7 * li if ii < len(prod) - 1 else 60
            mats = J.compose(S.transforms(2000 * li + ii, n, box, 0))
            seeded = (li + ii) % 3 == 1
            if seeded:
                sub = keys[1::7][:5]
                pf, of, mode = J.HazKeyFilter(sub), O.make_filter(1, keys=[k.idx for k in sub]), 2 | 4
            else:
                pf, of, mode = None, None, 2
            sets.append((cde, prod[ii], first, n, pf))
            expect.append(O.query(oc, orc[ii], mats=mats, mode=mode, filt=of, keys_cap=12, threads=8))
            mats_all.append(mats)
            first += n
    plan = J.QueryPlan(sets)
    cnt = check_plan(plan, mats_all, expect, 12, cuda)
    assert cnt.max() > 32, "no set went through the overflow collector"
    # again, without keys, and after an engine update
    cde, keys, oc = built[0]
    cde.deregister_hazard_by_key(keys[0])
    oc.deregister_key(oc.keys[0])
    e0 = O.query(oc, orc[0], mats=mats_all[0], mode=2, keys_cap=0, threads=8)
    m = torch.from_numpy(np.concatenate(mats_all)).to(cuda)
    c, h, _ = plan.collect(m, capacity=0)
    n0 = len(mats_all[0])
    np.testing.assert_array_equal(c[:n0].cpu().numpy().astype(np.uint32), e0["count"])
    np.testing.assert_array_equal(h[:n0].cpu().numpy().view(np.uint64), e0["hash"])


def test_collect_star_container_and_holes(cuda):
    """A many-layout workload's layout (star container of hundreds of vertices) and a layout
    with star holes at depth 10: containment scans over long candidate lists."""
    sc3 = S.config3_layout(3, 5)
    sc5 = S.config5_scene(seed=5, n_items=300, width=1200.0, height=400.0, n_holes=4)
    sets, expect, mats_all, first = [], [], [], 0
    for sc, seed in ((sc3, 31), (sc5, 32)):
        cde, keys, oc = engines(sc)
        polys = [S.polygon(seed * 10 + i, 8, 120, 4.0, 30.0, 2) for i in range(3)]
        prod, orc = items(polys)
        c = sc.container
        box = (float(c[:, 0].min()), float(c[:, 1].min()), float(c[:, 0].max()), float(c[:, 1].max()))
        for ii in range(3):
            mats = J.compose(S.transforms(seed * 100 + ii, 1500, box, 0))
            sets.append((cde, prod[ii], first, len(mats), None))
            expect.append(O.query(oc, orc[ii], mats=mats, mode=2, keys_cap=6, threads=8))
            mats_all.append(mats)
            first += len(mats)
    check_plan(J.QueryPlan(sets), mats_all, expect, 6, cuda)
