"""Pins the oracle's whole collision query, not only its predicates, against results that do
not come from the oracle:

* an exact brute force in integer arithmetic (every f32 is an integer multiple of 2^-149):
  a query collides with an interior-scope hazard iff their boundaries cross or one polygon
  lies inside the other, and with the exterior hazard iff it crosses the container's
  boundary or lies outside it.  That is what detect_poly_collision || detect_surrogate_collision
  decides (cd_engine.rs:160-255, lbf/benches/ci_bench.rs:111-119) and what the collected set
  holds (cd_engine.rs:261-310), for every sample whose polygons stay more than TOL apart
  wherever they do not cross (vertex-to-edge distances, both ways).  Samples closer than that
  are decided by f32 rounding and are skipped; the test asserts that few are.
* the oracle at quadtree depth 0 (one node, no virtual root, every hazard Partial:
  qt_node.rs:61 never splits level 0, so detection is a plain loop over all hazard edges and
  containments) against depths 3, 5 and 8 on the same samples: the quadrant descent
  (qt_traits.rs), the presence classification of constrict (qt_hazard.rs), the virtual root and
  strongest-hazard logic must not change a verdict or a set.

This covers the composition of the query, which the predicate tests cannot: virtual root,
strongest, the containment loop and its scope rule, collection of Entire hazards, the
exterior insert when the bbox leaves the engine's, and filters."""
import numpy as np
import pytest

import jagua_amd as J
from jagua_amd import synth as S
from oracle import oracle as O

F32 = np.float32
SCALE = 150                      # 2^SCALE * any finite f32 is an integer
TOL = 1e-2                       # world units (layouts span 1000)


def to_int(v):
    n, d = float(v).as_integer_ratio()
    return n * ((1 << SCALE) // d)


def orient(a, b, c):
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])


def proper_cross(p1, p2, p3, p4):
    """Exact: the closed segments p1p2 and p3p4 cross at a point interior to both (None when
    an endpoint lies on the other segment's line: a degenerate case)."""
    d1, d2 = orient(p3, p4, p1), orient(p3, p4, p2)
    d3, d4 = orient(p1, p2, p3), orient(p1, p2, p4)
    if d1 == 0 or d2 == 0 or d3 == 0 or d4 == 0:
        return None
    return (d1 > 0) != (d2 > 0) and (d3 > 0) != (d4 > 0)


def inside(poly, p):
    """Exact crossing-number test of point p (integers) in polygon poly (list of integer
    points); None on the boundary."""
    c = False
    n = len(poly)
    for i in range(n):
        a, b = poly[i], poly[(i + 1) % n]
        if (a[1] > p[1]) != (b[1] > p[1]):
            o = orient(a, b, p)
            if o == 0:
                return None
            if (o > 0) == (b[1] > a[1]):
                c = not c
    return c


def seg_dist(P, A, B):
    """min distance (f64) from points P [n,2] to segments A->B [m,2]."""
    P = P.astype(np.float64)[:, None, :]
    A = A.astype(np.float64)[None]
    B = B.astype(np.float64)[None]
    AB = B - A
    L2 = np.maximum((AB * AB).sum(-1), 1e-300)
    t = np.clip(((P - A) * AB).sum(-1) / L2, 0.0, 1.0)
    D = P - (A + t[..., None] * AB)
    return float(np.sqrt((D * D).sum(-1)).min())


def transform32(m, v):
    """point.rs:26-31 in f32, written order, no contraction."""
    m = m.astype(F32)
    x, y = v[:, 0].astype(F32), v[:, 1].astype(F32)
    tx = (m[0] * x + m[1] * y) + m[2] * F32(1.0)
    ty = (m[3] * x + m[4] * y) + m[5] * F32(1.0)
    return np.stack([tx, ty], 1).astype(F32)


class Haz:
    def __init__(self, slot, interior, v32):
        self.slot, self.interior = slot, interior
        self.v = v32
        self.A, self.B = v32, np.roll(v32, -1, axis=0)
        self.vi = [(to_int(x), to_int(y)) for x, y in v32]
        self.lo, self.hi = v32.min(0).astype(np.float64), v32.max(0).astype(np.float64)


def hazards_of(scene, oc):
    """The engine's hazards as the reference registers them (Container::new, then
    Layout::place_item): exterior, holes, placed items transformed by compose(dt)."""
    hz = [Haz(1, False, O.spolygon_new(scene.container)[0])]
    for i, h in enumerate(scene.holes):
        hz.append(Haz(2 + i, True, O.spolygon_new(h)[0]))
    for (loc, dt), key in zip(scene.items, oc.keys):
        m = O.compose(np.asarray(dt, F32)[None])[0]
        hz.append(Haz(key[0], True, transform32(m, O.spolygon_new(loc)[0])))
    return hz


def exact_query(hz, qv):
    """(robust, set of colliding slots, any crossing) of the query polygon qv (f32 [n,2])."""
    qA, qB = qv, np.roll(qv, -1, axis=0)
    qi = [(to_int(x), to_int(y)) for x, y in qv]
    qlo, qhi = qv.min(0).astype(np.float64), qv.max(0).astype(np.float64)
    hits, crossed = set(), False
    for h in hz:
        if h.interior and ((h.lo > qhi + TOL).any() or (h.hi < qlo - TOL).any()):
            continue
        if min(seg_dist(qv, h.A, h.B), seg_dist(h.v, qA, qB)) <= TOL:
            return False, None, None
        cross = False
        nq, nh = len(qi), len(h.vi)
        for i in range(nq):
            a, b = qv[i], qv[(i + 1) % nq]
            elo, ehi = np.minimum(a, b), np.maximum(a, b)
            cand = np.nonzero(((np.minimum(h.A, h.B) <= ehi).all(1)) & ((np.maximum(h.A, h.B) >= elo).all(1)))[0]
            for j in cand:
                c = proper_cross(qi[i], qi[(i + 1) % nq], h.vi[j], h.vi[(j + 1) % nh])
                if c is None:
                    return False, None, None
                if c:
                    cross = True
                    break
            if cross:
                break
        if cross:
            hits.add(h.slot)
            crossed = True
            continue
        q_in_h = inside(h.vi, qi[0])
        if q_in_h is None:
            return False, None, None
        if not h.interior:
            if not q_in_h:
                hits.add(h.slot)
            continue
        h_in_q = inside(qi, h.vi[0])
        if h_in_q is None:
            return False, None, None
        if q_in_h or h_in_q:
            hits.add(h.slot)
    return True, hits, crossed


def scenes():
    holes = S.generate_scene(depth=6, seed=9, cw=1000.0, ch=400.0, n_holes=3, hole_kind=0, hole_lo=150.0,
                             hole_hi=250.0, n_items=10, item_nv_lo=8, item_nv_hi=8, item_r_lo=15.0, item_r_hi=25.0)
    return [("lbf layout", S.config1_scene(0)), ("holes", holes)]


def queries(seed):
    """Query items of three sizes: small ones fit inside holes and gaps, large ones enclose
    placed items; transforms reach beyond the container bbox on every side."""
    rng = np.random.default_rng(seed)
    polys = [S.polygon(seed * 10 + 1, 8, 16, 2.0, 5.0, 0), S.polygon(seed * 10 + 2, 12, 16, 15.0, 35.0, 0),
             S.polygon(seed * 10 + 3, 10, 14, 60.0, 90.0, 0)]
    out = []
    for k, p in enumerate(polys):
        n = 900
        dt = np.zeros((n, 3), F32)
        dt[:, 0] = rng.uniform(0, 2 * np.pi, n)
        dt[: n // 4, 0] = (rng.integers(0, 8, n // 4) * (np.pi / 4)).astype(F32)   # lbf's discrete rotations
        dt[:, 1] = rng.uniform(-120, 1120, n)
        dt[:, 2] = rng.uniform(-120, 520, n)
        out.append((p, J.compose(dt)))
    return out


@pytest.mark.parametrize("which", [0, 1])
def test_oracle_query_equals_exact_brute_force(which):
    name, sc = scenes()[which]
    oc = O.OracleCDE(sc)
    hz = hazards_of(sc, oc)
    n_slots = max(h.slot for h in hz) + 1
    key_filter = [h.slot for h in hz if h.interior][::3]
    stats = dict(samples=0, skipped=0, collide=0, clear=0, contain_only=0, exterior=0)
    for p, mats in queries(which + 1):
        oit = O.OracleItem(p)
        qloc = oit.info()[0]
        ref = O.query(oc, oit, mats=mats, mode=3, keys_cap=n_slots, threads=8)
        reff = O.query(oc, oit, mats=mats, mode=1, filt=O.make_filter(1, keys=key_filter), threads=8)
        for q in range(len(mats)):
            stats["samples"] += 1
            robust, hits, crossed = exact_query(hz, transform32(mats[q], qloc))
            if not robust:
                stats["skipped"] += 1
                continue
            want = 1 if hits else 0
            assert ref["status"][q] == want, (name, q, sorted(hits))
            assert ref["count"][q] == len(hits), (name, q, sorted(hits), ref["keys"][q])
            keys = ref["keys"][q][: ref["count"][q]]
            assert sorted(keys.tolist()) == sorted(hits), (name, q)
            # HazKeyFilter: the verdict over the relevant hazards only (cd_engine.rs:160-223);
            # a filtered Entire hazard in the virtual root is the reference's unreachable!()
            if reff["status"][q] != J.Q_PANIC_FILTERED_ENTIRE:
                assert reff["status"][q] == (1 if hits - set(key_filter) else 0), (name, q)
            stats["collide" if hits else "clear"] += 1
            stats["contain_only"] += bool(hits) and not crossed
            stats["exterior"] += 1 in hits
    print(name, stats)
    assert stats["skipped"] <= 0.05 * stats["samples"], stats
    for k in ("collide", "clear", "contain_only", "exterior"):
        assert stats[k] >= 20, stats


@pytest.mark.parametrize("which", [0, 1])
def test_depth0_equals_deeper_trees(which):
    name, sc = scenes()[which]
    base = O.OracleCDE(S.Scene(sc.container, sc.holes, sc.items, 0, sc.cd_threshold))
    hz = hazards_of(sc, base)
    deep = [O.OracleCDE(S.Scene(sc.container, sc.holes, sc.items, d, sc.cd_threshold)) for d in (3, 5, 8)]
    mismatched = robust_mismatch = total = 0
    for p, mats in queries(which + 11):
        oit = O.OracleItem(p)
        qloc = oit.info()[0]
        r0 = O.query(base, oit, mats=mats, mode=3, threads=8)
        for oc in deep:
            r = O.query(oc, oit, mats=mats, mode=3, threads=8)
            bad = np.nonzero((r["status"] != r0["status"]) | (r["hash"] != r0["hash"]))[0]
            total += len(mats)
            mismatched += len(bad)
            for q in bad:
                robust, _, _ = exact_query(hz, transform32(mats[q], qloc))
                robust_mismatch += robust
                print(name, "depth mismatch at sample", q, "robust" if robust else "within TOL of a boundary")
    assert robust_mismatch == 0
    assert mismatched <= 1e-3 * total, (mismatched, total)


def test_preseeded_collector_is_the_union():
    """A collector that arrives holding some keys (the collector as filter, filter.rs:69-76)
    ends with those keys together with the unseeded set: collection only adds
    (collector.rs:44-67), so seeding can change nothing else."""
    name, sc = scenes()[1]
    oc = O.OracleCDE(sc)
    hz = hazards_of(sc, oc)
    n_slots = max(h.slot for h in hz) + 1
    seed = [h.slot for h in hz][1::4]
    for p, mats in queries(21):
        oit = O.OracleItem(p)
        plain = O.query(oc, oit, mats=mats, mode=2, keys_cap=n_slots, threads=8)
        seeded = O.query(oc, oit, mats=mats, mode=2 | 4, filt=O.make_filter(1, keys=seed), keys_cap=n_slots,
                         threads=8)
        for q in range(len(mats)):
            a = set(plain["keys"][q][: plain["count"][q]].tolist())
            b = set(seeded["keys"][q][: seeded["count"][q]].tolist())
            assert b == a | set(seed), (q, sorted(a), sorted(b))
