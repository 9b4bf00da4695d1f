"""GPU collision queries against the CPU oracle, bit for bit (verdicts, statuses, hazard
sets), through the C ABI.  Cases follow the reference's test strategy: the lbf-style
layout of the benchmark, quadtree depths 0/3/10 (lbf/tests/tests.rs:18), hazard filters
(filter.rs), place/remove/restore sequences (tests.rs:48-66), and edge cases (empty layout,
queries outside the container, huge hazard sets, the filtered-Entire status)."""
import numpy as np
import pytest
import torch

import jagua_amd as J
from jagua_amd import synth as S
from oracle import oracle as O

from helpers import assert_same_item, engines, items

pytestmark = pytest.mark.gpu


def run_both(cde, oc, item, oitem, mats, filt=None, ofilt=None, cap=16, cuda=None, collect=True):
    m = torch.from_numpy(mats).to(cuda)
    st = cde.detect_batch(item, m, filt).cpu().numpy()
    ref = O.query(oc, oitem, mats=mats, mode=3 if collect else 1, filt=ofilt, keys_cap=cap, threads=8)
    np.testing.assert_array_equal(st, ref["status"])
    if collect:
        cnt, hsh, keys = cde.collect_batch(item, m, capacity=cap)
        np.testing.assert_array_equal(cnt.cpu().numpy().astype(np.uint32), ref["count"])
        np.testing.assert_array_equal(hsh.cpu().numpy().view(np.uint64), ref["hash"])
        np.testing.assert_array_equal(keys.cpu().numpy().view(np.uint32), ref["keys"])
    return st, ref


def test_compose_matches_oracle():
    dt = S.transforms(5, 20000, (0, 0, 1000, 400), 0)
    assert np.array_equal(J.compose(dt).view(np.uint32), O.compose(dt).view(np.uint32))


def test_config1_full(cuda):
    sc = S.config1_scene(0)
    cde, keys, oc = engines(sc)
    polys, dts = S.config1_queries(0)
    (item,), (oitem,) = items(polys)
    assert_same_item(item, polys[0])
    st, ref = run_both(cde, oc, item, oitem, J.compose(dts), cuda=cuda)
    assert 0 < st.sum() < len(st)          # a mix of clear and colliding samples
    assert (ref["count"] > 0).sum() > 0


@pytest.mark.parametrize("depth", [0, 3, 10])
def test_depths(cuda, depth):
    sc = S.config1_scene(1)
    sc.depth = depth
    cde, keys, oc = engines(sc)
    polys, dts = S.config1_queries(1, n=2048)
    (item,), (oitem,) = items(polys)
    run_both(cde, oc, item, oitem, J.compose(dts), cuda=cuda)


def test_config2_small(cuda):
    sc = S.config2_scene(seed=3, n_items=300)
    cde, keys, oc = engines(sc)
    polys, dts = S.config2_queries(seed=3, n_items=4, n_per_item=1024)
    its, oits = items(polys)
    for it, oit, p, d in zip(its, oits, polys, dts):
        assert_same_item(it, p)
        run_both(cde, oc, it, oit, J.compose(d), cuda=cuda)


def test_filters(cuda):
    sc = S.config1_scene(2)
    cde, keys, oc = engines(sc)
    polys, dts = S.config1_queries(2, n=4096)
    (item,), (oitem,) = items(polys)
    mats = J.compose(dts)
    # HazKey self filter (Layout::is_feasible, layout.rs:128-136) and a HazKeyFilter subset
    run_both(cde, oc, item, oitem, mats, J.HazKeyFilter([keys[3]]), O.make_filter(1, keys=[keys[3].idx]),
             cuda=cuda, collect=False)
    sub = keys[::3]
    run_both(cde, oc, item, oitem, mats, J.HazKeyFilter(sub), O.make_filter(1, keys=[k.idx for k in sub]),
             cuda=cuda, collect=False)
    # a random 10% of hazards per query
    hf = J.HashedFilter(1234, 0.10)
    st, ref = run_both(cde, oc, item, oitem, mats, hf, O.make_filter(2, hf.seed, hf.threshold), cuda=cuda,
                       collect=False)


def test_filtered_entire_status(cuda):
    """The reference aborts (cd_engine.rs:179) when a filtered Entire hazard sits in the
    virtual root; both sides report JG_Q_PANIC_FILTERED_ENTIRE for exactly the same queries."""
    sc = S.generate_scene(depth=6, seed=9, cw=1000.0, ch=400.0, n_holes=3, hole_kind=0, hole_lo=150.0,
                          hole_hi=250.0, n_items=10, item_nv_lo=8, item_nv_hi=8, item_r_lo=15.0, item_r_hi=25.0)
    cde, keys, oc = engines(sc)
    poly = S.polygon(77, 8, 8, 3.0, 6.0, 0)
    (item,), (oitem,) = items([poly])
    holes = [2, 3, 4]
    dts = S.transforms(78, 8192, (0, 0, 1000, 400), 0)
    f = J.HazKeyFilter([J.HazKey(h, 1) for h in holes])
    st, ref = run_both(cde, oc, item, oitem, J.compose(dts), f, O.make_filter(1, keys=holes), cuda=cuda,
                       collect=False)
    assert (st == J.Q_PANIC_FILTERED_ENTIRE).sum() > 0


def test_empty_layout_and_outside(cuda):
    sc = S.Scene(np.array([[0, 0], [500, 0], [500, 200], [0, 200]], np.float32), [], [], 4, 16)
    cde, keys, oc = engines(sc)
    poly = S.polygon(5, 10, 10, 5.0, 20.0, 0)
    (item,), (oitem,) = items([poly])
    dts = S.transforms(6, 4096, (-100, -300, 600, 500), 0)    # many samples outside the CDE bbox
    st, ref = run_both(cde, oc, item, oitem, J.compose(dts), cuda=cuda)
    assert st.sum() > 0


def test_large_sets(cuda):
    """A query polygon covering many hazards: sets larger than the key capacity are still
    counted and hashed exactly."""
    sc = S.config1_scene(3)
    cde, keys, oc = engines(sc)
    poly = S.polygon(8, 64, 64, 150.0, 300.0, 0)
    (item,), (oitem,) = items([poly])
    dts = S.transforms(9, 512, (0, 0, 1000, 400), 0)
    run_both(cde, oc, item, oitem, J.compose(dts), cap=4, cuda=cuda)


def test_update_sequence(cuda):
    """place / remove / save / restore (lbf/tests/tests.rs:48-66): queries stay bit-exact."""
    sc = S.config1_scene(4)
    cde, keys, oc = engines(sc)
    polys, dts = S.config1_queries(4, n=2048)
    (item,), (oitem,) = items(polys)
    mats = J.compose(dts)
    snap, osnap = cde.save(), oc.save()
    rng = np.random.default_rng(0)
    for i in rng.choice(len(keys), 5, replace=False):
        cde.deregister_hazard_by_key(keys[i])
        oc.deregister_key(oc.keys[i])
    run_both(cde, oc, item, oitem, mats, cuda=cuda)
    cde.restore(snap)
    O.lib().or_cde_restore(oc.h, osnap)
    run_both(cde, oc, item, oitem, mats, cuda=cuda)


def test_collect_sets_beyond_the_slot_list(cuda):
    """Hazard sets of more than 256 slots (the collector's slot list in LDS) are counted, hashed
    and listed exactly through the bitmask scan."""
    sc = S.config2_scene(seed=14, n_items=1200)
    cde, keys, oc = engines(sc)
    poly = S.polygon(21, 64, 64, 600.0, 800.0, 0)
    (item,), (oitem,) = items([poly])
    dts = S.transforms(22, 48, (500, 400, 3500, 1100), 0)
    st, ref = run_both(cde, oc, item, oitem, J.compose(dts), cap=300, cuda=cuda)
    assert ref["count"].max() > 256 and ref["count"].min() < 256


@pytest.mark.parametrize("depth", [3, 8])
def test_random_update_sequences(cuda, depth):
    """Random batches of deregister / register / save-restore (lbf/tests/tests.rs:48-66 in a
    loop).  Each batch reaches the device image as an in-place patch of the touched nodes and
    slots; after every batch the plain detect / collect and a batched plan's detect and collect
    agree with the oracle bit for bit, and the keys handed out agree with the oracle's slot map."""
    sc = S.config1_scene(40)
    sc.depth = depth
    cde, keys, oc = engines(sc)
    polys, dts = S.config1_queries(40, n=2048)
    (item,), (oitem,) = items(polys)
    mats = J.compose(dts)
    m = torch.from_numpy(mats).to(cuda)
    plan = J.QueryPlan([(cde, item, 0, len(mats), None)])
    out = torch.empty(len(mats), dtype=torch.uint8, device=cuda)
    run_both(cde, oc, item, oitem, mats, cuda=cuda)
    full0 = cde.stats()["n_full_uploads"]
    extra = [it for s in (41, 42, 43) for it in S.config1_scene(s).items]
    live = [(k, ok, i, dt) for i, (k, ok, (_, dt)) in enumerate(zip(keys, oc.keys, sc.items))]
    rng = np.random.default_rng(depth)
    snap = None
    for rnd in range(10):
        for _ in range(int(rng.integers(1, 6))):
            if live and rng.random() < 0.5:
                k, ok, _, _ = live.pop(int(rng.integers(len(live))))
                cde.deregister_hazard_by_key(k)
                oc.deregister_key(ok)
            elif extra:
                loc, dt = extra.pop()
                pk = 1000 + len(extra)
                k = cde.register_hazard(J.Hazard(J.HazardEntity.placed_item(pk, dt, pk),
                                                 J.SPolygon.new(loc).transform_clone(dt), True))
                ok = oc.register_item(pk, loc, dt, pk)
                assert (k.idx, k.version) == tuple(ok)
                live.append((k, ok, pk, dt))
        if rnd == 3:
            snap, osnap = cde.save(), oc.save()
        if rnd == 7:
            cde.restore(snap)
            oc.restore(osnap)
            live = []          # keys after a restore are not tracked; later rounds only register
        st, ref = run_both(cde, oc, item, oitem, mats, cuda=cuda)
        plan.detect(m, out)
        np.testing.assert_array_equal(out.cpu().numpy(), ref["status"])
        # the batched collection starts query edges through the cell tables' first collecting
        # depths, which the patches keep current
        cnt, hsh, _ = plan.collect(m, capacity=0)
        np.testing.assert_array_equal(cnt.cpu().numpy().astype(np.uint32), ref["count"])
        np.testing.assert_array_equal(hsh.cpu().numpy().view(np.uint64), ref["hash"])
    s = cde.stats()
    assert s["n_patches"] >= 8 and s["n_full_uploads"] - full0 <= 2, s
