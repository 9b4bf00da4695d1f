"""The batched persistent form (one launch over many layouts, items and filters, fed by a
device work queue) against the oracle, bit for bit; hipGraph replays checked one by one."""
import numpy as np
import pytest
import torch

import jagua_amd as J
from jagua_amd import synth as S
from oracle import oracle as O

from helpers import engines, items

pytestmark = pytest.mark.gpu


def test_batched_many_layouts(cuda):
    scenes = [S.config1_scene(10 + i) for i in range(3)] + [S.config2_scene(seed=11, n_items=250)]
    built = [engines(s) for s in scenes]
    polys = [S.polygon(100 + i, 8, 200, 8.0, 30.0, 2) for i in range(5)] + [S.config1_queries(3)[0][0]]
    prod, orc = items(polys)
    sets, expect, first = [], [], 0
    mats_all = []
    filters = [None, "keys", "hashed"]
    for li, (cde, keys, oc) in enumerate(built):
        box = scenes[li].cde_bbox()
        for ii in range(len(polys)):
            n = 300 + 97 * ii + 13 * li
            dts = S.transforms(1000 * li + ii, n, box, 0)
            mats = J.compose(dts)
            kind = filters[(li + ii) % 3]
            if kind == "keys":
                sub = keys[::4]
                pf, of = J.HazKeyFilter(sub), O.make_filter(1, keys=[k.idx for k in sub])
            elif kind == "hashed":
                pf = J.HashedFilter(77 + li, 0.1)
                of = O.make_filter(2, pf.seed, pf.threshold)
            else:
                pf, of = None, None
            sets.append((cde, prod[ii], first, n, pf))
            expect.append(O.query(oc, orc[ii], mats=mats, mode=1, filt=of, threads=8)["status"])
            mats_all.append(mats)
            first += n
    plan = J.QueryPlan(sets)
    m = torch.from_numpy(np.concatenate(mats_all)).to(cuda)
    out = torch.full((first,), 255, dtype=torch.uint8, device=cuda)
    plan.detect(m, out)
    np.testing.assert_array_equal(out.cpu().numpy(), np.concatenate(expect))
    # graph replays, with eager work and synchronisation between them
    for r in range(3):
        out.fill_(255)
        torch.cuda.synchronize()
        plan.detect_graph(m, out, replays=1)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), np.concatenate(expect))
    # the plan follows an engine update (re-flattened image)
    cde, keys, oc = built[0]
    cde.deregister_hazard_by_key(keys[0])
    oc.deregister_key(oc.keys[0])
    exp0 = O.query(oc, orc[0], mats=mats_all[0], mode=1, threads=8)["status"]
    plan.detect(m, out)
    np.testing.assert_array_equal(out[: len(exp0)].cpu().numpy(), exp0)


@pytest.mark.parametrize("stages,chunk,n", [(2, 1, 3000), (2, 32, 3000), (1, 7, 3000), (1, 64, 3000),
                                             (2, 64, 100), (1, 16, 100), (0, 0, 1)])
def test_batched_chunks(cuda, stages, chunk, n):
    """Both stage layouts, every unit size, 7 workgroups (not a multiple of the 8 claim
    streams) and plans with fewer units than streams: every query is claimed exactly once."""
    sc = S.config1_scene(20)
    cde, keys, oc = engines(sc)
    polys, dts = S.config1_queries(20, n=n)
    prod, orc = items(polys)
    mats = J.compose(dts)
    plan = J.QueryPlan([(cde, prod[0], 0, len(mats), None)], chunk=chunk, workgroups=7, stages=stages)
    out = torch.empty(len(mats), dtype=torch.uint8, device=cuda)
    plan.detect(torch.from_numpy(mats).to(cuda), out)
    ref = O.query(oc, orc[0], mats=mats, mode=1)["status"]
    np.testing.assert_array_equal(out.cpu().numpy(), ref)


def test_batched_poly_only_and_surrogate_only(cuda):
    sc = S.config1_scene(21)
    cde, keys, oc = engines(sc)
    polys, dts = S.config1_queries(21, n=2048)
    prod, orc = items(polys)
    mats = J.compose(dts)
    m = torch.from_numpy(mats).to(cuda)
    ref = O.query(oc, orc[0], mats=mats, mode=1)
    for flags in (J.QUERY_POLY, J.QUERY_SURROGATE):
        plan = J.QueryPlan([(cde, prod[0], 0, len(mats), None)], flags=flags)
        out = torch.empty(len(mats), dtype=torch.uint8, device=cuda)
        plan.detect(m, out)
        plain = cde.detect_batch(prod[0], m, flags=flags).cpu().numpy()
        np.testing.assert_array_equal(out.cpu().numpy(), plain)
        if flags == J.QUERY_SURROGATE:
            np.testing.assert_array_equal(plain, ref["surr"])


def test_autotune_keeps_verdicts(cuda):
    sc = S.config2_scene(seed=12, n_items=300)
    cde, keys, oc = engines(sc)
    polys = [S.polygon(300 + i, 8, 300, 10.0, 30.0, 2) for i in range(4)]
    prod, orc = items(polys)
    box = sc.cde_bbox()
    sets, expect, mats_all, first = [], [], [], 0
    for i in range(len(polys)):
        mats = J.compose(S.transforms(500 + i, 3000 + 111 * i, box, 0))
        filt = J.HashedFilter(9, 0.1) if i % 2 else None
        of = O.make_filter(2, filt.seed, filt.threshold) if filt else None
        sets.append((cde, prod[i], first, len(mats), filt))
        expect.append(O.query(oc, orc[i], mats=mats, mode=1, filt=of, threads=8)["status"])
        mats_all.append(mats)
        first += len(mats)
    m = torch.from_numpy(np.concatenate(mats_all)).to(cuda)
    out = torch.full((first,), 255, dtype=torch.uint8, device=cuda)
    plan = J.QueryPlan(sets, flags=0)
    cfg, ms = plan.autotune(m, out, reps=2)
    torch.cuda.synchronize()
    assert ms > 0 and cfg == plan.config()
    assert cfg["lanes_per_query"] in (0, 8, 16, 32, 64) and cfg["chunk"] in (16, 32, 64) and cfg["stages"] in (1, 2)
    np.testing.assert_array_equal(out.cpu().numpy(), np.concatenate(expect))
    out.fill_(255)
    plan.detect(m, out)
    np.testing.assert_array_equal(out.cpu().numpy(), np.concatenate(expect))
    out.fill_(255)
    plan.detect_graph(m, out, replays=2)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), np.concatenate(expect))


def test_batched_fractional_small_layout(cuda):
    """A layout tens of units wide (leaf nodes about 1.25 units) with 48-64-vertex query
    items, so each query fills a wavefront in the polygon stage: the query bbox and virtual
    root must be computed from exact f32 values (a bbox rounded to integers picks a wrong
    virtual root here), plain and batched forms against the oracle."""
    sc = S.generate_scene(depth=5, seed=41, container_kind=0, cw=40.0, ch=16.0, n_items=40, item_nv_lo=12,
                          item_nv_hi=40, item_r_lo=0.8, item_r_hi=1.6, item_shape=0, min_sep=0.0)
    cde, keys, oc = engines(sc)
    polys = [S.polygon(42 + i, 48, 64, 0.6, 1.4, i % 3) for i in range(4)]
    prod, orc = items(polys)
    sets, refs, mats_all, first = [], [], [], 0
    for i, (it, oit) in enumerate(zip(prod, orc)):
        m = J.compose(S.transforms(43 + i, 4096, (0, 0, 40, 16), 0))
        refs.append(O.query(oc, oit, mats=m, mode=1, threads=8)["status"])
        sets.append((cde, it, first, len(m), None))
        mats_all.append(m)
        first += len(m)
    ref = np.concatenate(refs)
    assert 0 < ref.sum() < len(ref)
    M = torch.from_numpy(np.concatenate(mats_all)).to(cuda)
    for stages in (1, 2):
        out = torch.full((len(M),), 9, dtype=torch.uint8, device=cuda)
        J.QueryPlan(sets, stages=stages).detect(M, out)
        np.testing.assert_array_equal(out.cpu().numpy(), ref, err_msg=f"stages {stages}")


@pytest.mark.parametrize("filtered", [False, True])
def test_batched_shallow_virtual_roots(cuda, filtered):
    """Small query items (8-24 vertices, so G <= 32 lanes per query) placed across the root's
    centre lines: their virtual root is the root or a node just below it, whose Partial-hazard
    lists hold hundreds of containment candidates, scanned by the whole wave one query at a
    time. Verdicts equal the oracle's with and without a hashed filter, under both stage
    layouts."""
    sc = S.config2_scene(seed=13, n_items=400)
    cde, keys, oc = engines(sc)
    box = sc.cde_bbox()
    cx, cy = 0.5 * (box[0] + box[2]), 0.5 * (box[1] + box[3])
    polys = [S.polygon(700 + i, 8, 24, 10.0, 30.0, 2) for i in range(3)]
    prod, orc = items(polys)
    rng = np.random.default_rng(5)
    sets, expect, mats_all, first = [], [], [], 0
    for i in range(len(polys)):
        n = 2048
        dt = np.empty((n, 3), dtype=np.float32)
        dt[:, 0] = rng.uniform(0, 6.28, n)
        across_x = rng.random(n) < 0.5                     # straddle the vertical or horizontal centre line
        dt[:, 1] = np.where(across_x, cx + rng.uniform(-15, 15, n), rng.uniform(box[0], box[2], n))
        dt[:, 2] = np.where(across_x, rng.uniform(box[1], box[3], n), cy + rng.uniform(-15, 15, n))
        mats = J.compose(dt)
        filt = J.HashedFilter(21 + i, 0.1) if filtered else None
        of = O.make_filter(2, filt.seed, filt.threshold) if filt else None
        sets.append((cde, prod[i], first, len(mats), filt))
        expect.append(O.query(oc, orc[i], mats=mats, mode=1, filt=of, threads=8)["status"])
        mats_all.append(mats)
        first += len(mats)
    ref = np.concatenate(expect)
    assert 0 < (ref == 0).sum() and (ref == 1).sum() > 0
    m = torch.from_numpy(np.concatenate(mats_all)).to(cuda)
    for stages in (2, 1):
        out = torch.full((first,), 9, dtype=torch.uint8, device=cuda)
        J.QueryPlan(sets, stages=stages).detect(m, out)
        np.testing.assert_array_equal(out.cpu().numpy(), ref)


def test_empty_and_ragged_sets(cuda):
    """Query sets of zero queries among non-empty ones (ragged), a plan of only empty sets and
    empty plain batches: the other sets' verdicts and hazard sets are the oracle's, and the empty
    ones launch nothing that writes."""
    sc = S.config1_scene(23)
    cde, keys, oc = engines(sc)
    polys, dts = S.config1_queries(23, n=600)
    prod, orc = items(polys)
    mats = J.compose(dts)
    m = torch.from_numpy(mats).to(cuda)
    ref = O.query(oc, orc[0], mats=mats, mode=3)
    sets = [(cde, prod[0], 0, 0, None), (cde, prod[0], 0, 250, None), (cde, prod[0], 250, 0, None),
            (cde, prod[0], 250, 350, None), (cde, prod[0], 600, 0, None)]
    plan = J.QueryPlan(sets)
    out = torch.full((len(mats),), 255, dtype=torch.uint8, device=cuda)
    plan.detect(m, out)
    np.testing.assert_array_equal(out.cpu().numpy(), ref["status"])
    cnt, hsh, _ = plan.collect(m, capacity=0)
    np.testing.assert_array_equal(cnt.cpu().numpy().astype(np.uint32), ref["count"])
    np.testing.assert_array_equal(hsh.cpu().numpy().view(np.uint64), ref["hash"])
    # only empty sets, and empty plain batches
    none = J.QueryPlan([(cde, prod[0], 0, 0, None)])
    none.detect(m[:0], out[:0])
    assert cde.detect_batch(prod[0], m[:0]).numel() == 0
    c0, h0, _ = cde.collect_batch(prod[0], m[:0], capacity=0)
    assert c0.numel() == 0 and h0.numel() == 0
    torch.cuda.synchronize()
