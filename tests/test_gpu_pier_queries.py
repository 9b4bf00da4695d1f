"""Query items whose surrogate carries fail-fast piers (SPSurrogateConfig::n_ff_piers > 0,
sp_surrogate.rs:55-61): the pier segments descend from the quadtree root in
detect_surrogate_collision and collect_surrogate_collisions (cd_engine.rs:204-223,
296-310).  Plain detect / collect and both batched stage layouts against the oracle, bit for
bit, with pier-only surrogates (the edge-entity kernels) and pole + pier surrogates (the
mixed-entity kernels) in one plan, with and without filters.  Also: a planned item whose
surrogate is regenerated, and key-filtered queries issued back to back on one stream."""
import numpy as np
import pytest
import torch

import jagua_amd as J
from jagua_amd import synth as S
from oracle import oracle as O

from helpers import assert_same_item, engines

pytestmark = pytest.mark.gpu


def _items(polys, poles, piers):
    cfg = J.SPSurrogateConfig(n_ff_poles=poles, n_ff_piers=piers)
    prod = S.make_items(polys, cfg)
    orc = [O.OracleItem(p, n_ff_poles=poles, n_ff_piers=piers) for p in polys]
    for it, p in zip(prod, polys):
        assert_same_item(it, p, n_ff_poles=poles, n_ff_piers=piers)
    return prod, orc


def _polys():
    return [S.polygon(600 + i, nv, nv, 12.0, 30.0, 2) for i, nv in enumerate((9, 16, 40))]


def test_pier_items_plain_detect_and_collect(cuda):
    sc = S.config1_scene(60)
    cde, keys, oc = engines(sc)
    polys = _polys()
    for poles, piers in ((0, 2), (2, 1)):
        prod, orc = _items(polys, poles, piers)
        for i, (it, oit) in enumerate(zip(prod, orc)):
            mats = J.compose(S.transforms(610 + i, 3000, (0, 0, 1000, 400), 0))
            m = torch.from_numpy(mats).to(cuda)
            ref = O.query(oc, oit, mats=mats, mode=3, keys_cap=8, threads=8)
            np.testing.assert_array_equal(cde.detect_batch(it, m).cpu().numpy(), ref["status"])
            # the surrogate stage alone: pier (and pole) descents from the root
            sur = cde.detect_batch(it, m, flags=J.QUERY_SURROGATE).cpu().numpy()
            np.testing.assert_array_equal(sur, ref["surr"])
            assert ref["surr"].sum() > 0
            cnt, hsh, k = cde.collect_batch(it, m, capacity=8)
            np.testing.assert_array_equal(cnt.cpu().numpy().astype(np.uint32), ref["count"])
            np.testing.assert_array_equal(hsh.cpu().numpy().view(np.uint64), ref["hash"])
            np.testing.assert_array_equal(k.cpu().numpy().view(np.uint32), ref["keys"])


@pytest.mark.parametrize("stages", [2, 1])
@pytest.mark.parametrize("filtered", [False, True])
def test_pier_items_batched(cuda, stages, filtered):
    """Pier-only and pole + pier surrogates in one plan (mixed-entity kernels), and a plan of
    pier-only surrogates alone (edge-entity kernels)."""
    sc = S.config2_scene(seed=61, n_items=250)
    cde, keys, oc = engines(sc)
    polys = _polys()
    pier_only, o_pier_only = _items(polys, 0, 2)
    mixed, o_mixed = _items(polys[:2], 2, 1)
    box = sc.cde_bbox()
    for group in ((pier_only, o_pier_only), (pier_only + mixed, o_pier_only + o_mixed)):
        sets, expect, mats_all, first = [], [], [], 0
        for i, (it, oit) in enumerate(zip(*group)):
            n = 1500 + 101 * i
            mats = J.compose(S.transforms(620 + i, n, box, 0))
            if filtered and i % 2 == 0:
                pf = J.HashedFilter(31 + i, 0.1)
                of = O.make_filter(2, pf.seed, pf.threshold)
            elif filtered:
                sub = keys[1::4]
                pf, of = J.HazKeyFilter(sub), O.make_filter(1, keys=[k.idx for k in sub])
            else:
                pf, of = None, None
            sets.append((cde, it, first, n, pf))
            expect.append(O.query(oc, oit, mats=mats, mode=1, filt=of, threads=8)["status"])
            mats_all.append(mats)
            first += n
        plan = J.QueryPlan(sets, stages=stages)
        m = torch.from_numpy(np.concatenate(mats_all)).to(cuda)
        out = torch.full((first,), 255, dtype=torch.uint8, device=cuda)
        plan.detect(m, out)
        np.testing.assert_array_equal(out.cpu().numpy(), np.concatenate(expect))


def test_plan_follows_a_regenerated_surrogate(cuda):
    """A planned item whose surrogate is regenerated with other pole and pier counts: the
    plan's table (and its captured graph) must be rebuilt from the item's new device record."""
    sc = S.config1_scene(62)
    cde, keys, oc = engines(sc)
    poly = S.polygon(630, 14, 14, 12.0, 30.0, 0)
    (it,), (oit,) = _items([poly], 2, 0)
    mats = J.compose(S.transforms(631, 4000, (0, 0, 1000, 400), 0))
    m = torch.from_numpy(mats).to(cuda)
    plan = J.QueryPlan([(cde, it, 0, len(mats), None)])
    out = torch.full((len(mats),), 255, dtype=torch.uint8, device=cuda)
    plan.detect(m, out)
    np.testing.assert_array_equal(out.cpu().numpy(), O.query(oc, oit, mats=mats, mode=1)["status"])
    plan.detect_graph(m, out)
    for poles, piers in ((1, 2), (3, 0)):
        it.generate_surrogate(J.SPSurrogateConfig(n_ff_poles=poles, n_ff_piers=piers))
        oit2 = O.OracleItem(poly, n_ff_poles=poles, n_ff_piers=piers)
        ref = O.query(oc, oit2, mats=mats, mode=1)
        out.fill_(255)
        plan.detect(m, out)
        np.testing.assert_array_equal(out.cpu().numpy(), ref["status"])
        out.fill_(255)
        plan.detect_graph(m, out)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), ref["status"])
        # the surrogate stage sees the new poles and piers
        sur = torch.full((len(mats),), 255, dtype=torch.uint8, device=cuda)
        J.QueryPlan([(cde, it, 0, len(mats), None)], flags=J.QUERY_SURROGATE).detect(m, sur)
        np.testing.assert_array_equal(sur.cpu().numpy(), ref["surr"])


def test_key_filtered_calls_back_to_back_on_one_stream(cuda):
    """Two key-filtered plain detections issued on one non-blocking stream without a
    synchronisation between them: the second call's filter mask must not replace the first's
    while the first kernel may still read it."""
    sc = S.config2_scene(seed=63, n_items=400)
    cde, keys, oc = engines(sc)
    poly = S.polygon(640, 12, 12, 10.0, 25.0, 2)
    (it,), (oit,) = _items([poly], 2, 0)
    mats = J.compose(S.transforms(641, 200000, sc.cde_bbox(), 0))
    m = torch.from_numpy(mats).to(cuda)
    subs = [keys[0::2], keys[1::2]]
    refs = [O.query(oc, oit, mats=mats, mode=1, filt=O.make_filter(1, keys=[k.idx for k in s]), threads=8)["status"]
            for s in subs]
    assert not np.array_equal(refs[0], refs[1])
    stream = torch.cuda.Stream()
    outs = [torch.full((len(mats),), 255, dtype=torch.uint8, device=cuda) for _ in subs]
    with torch.cuda.stream(stream):
        for r in range(3):
            for s, o in zip(subs, outs):
                cde.detect_batch(it, m, J.HazKeyFilter(s), out=o, stream=stream)
    stream.synchronize()
    for o, ref in zip(outs, refs):
        np.testing.assert_array_equal(o.cpu().numpy(), ref)
