"""Shared builders for the parity tests: one scene registered into both the product engine
and the oracle, query items built on both sides, and the comparison of their outputs."""
import numpy as np

import jagua_amd as J
from jagua_amd import synth as S
from oracle import oracle as O


def engines(scene, surrogate_config=None):
    cde, keys = S.build_cde(scene, surrogate_config)
    oc = O.OracleCDE(scene)
    return cde, keys, oc


def items(polys, n_ff_poles=2, n_ff_piers=0):
    cfg = J.SPSurrogateConfig(n_ff_poles=n_ff_poles, n_ff_piers=n_ff_piers)
    prod = S.make_items(polys, cfg)
    orc = [O.OracleItem(p, n_ff_poles=n_ff_poles, n_ff_piers=n_ff_piers) for p in polys]
    return prod, orc


def assert_same_item(prod_item, poly, n_ff_poles=2, n_ff_piers=0):
    """Product SPolygon::new + surrogate == oracle, bit for bit."""
    verts, meta = O.spolygon_new(poly)
    assert np.array_equal(prod_item.vertices, verts)
    i = prod_item.info()
    mine = np.array([i.area, i.diameter, *i.bbox, *i.poi], dtype=np.float32)
    assert np.array_equal(mine.view(np.uint32), meta.view(np.uint32)), (mine, meta)
    poles, piers = O.surrogate(poly, n_ff_poles=n_ff_poles, n_ff_piers=n_ff_piers)
    assert np.array_equal(prod_item.poles.view(np.uint32), poles.view(np.uint32))
    assert np.array_equal(prod_item.piers.view(np.uint32), piers.view(np.uint32))


def tree_groups(nodes, haz):
    out, o = [], 0
    for r in nodes:
        k = int(r[2])
        out.append(sorted(map(tuple, haz[o:o + k].tolist())))
        o += k
    return out


def assert_same_tree(cde, oc):
    pn, ph = cde.dump()
    on, oh = oc.dump()
    assert np.array_equal(pn, on), "node records differ"
    assert tree_groups(pn, ph) == tree_groups(on, oh), "hazard entries differ"
