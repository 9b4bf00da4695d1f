"""Collision queries on the reference's own instances (tests/golden/assets, the instances
lbf/tests/tests.rs runs), bit for bit with the oracle.  Items are imported with io.Importer
and placed one after another into a strip (or the first bin, with its holes and quality zones)
by the GPU placement search; each placement is checked clear by the oracle and registered in
both engines (the oracle builds every hazard as the reference does: SPolygon::new of the
original shape, centred, then placed).  Then every item is queried at random poses: detect,
and collect with the hazard sets."""
import json
import os

import numpy as np
import pytest
import torch

import jagua_amd as J
from jagua_amd import io as I
from jagua_amd import synth as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ASSETS = os.path.join(os.path.dirname(__file__), "golden", "assets")


def _engines(cont, depth):
    cde = cont.base_cde(J.CDEConfig(depth, 16))
    oc = O.OracleCDE.from_hazards(cont.cde_bbox(), depth, 16, [(h.entity.kind, h.entity.a, h.entity.b, h.shape.vertices)
                                                                 for h in cont.hazards()])
    return cde, oc


@pytest.mark.parametrize("name,depth", [("shirts", 5), ("swim", 5), ("trousers", 3), ("baldacci1", 5)])
def test_instance_layout_queries(cuda, name, depth):
    ext = json.load(open(os.path.join(ASSETS, name + ".json")))
    imp = I.Importer()
    if "bins" in ext:
        inst = I.import_bpp_instance(imp, ext)
        cont = inst.bins[0][0]
    else:
        inst = I.import_spp_instance(imp, ext)
        cont = inst.strip(np.float32(inst.base_width * np.float32(1.3)))
    cde, oc = _engines(cont, depth)
    items = [it for it, _ in inst.items]
    oitems = {it.id: O.OracleItem(it.shape_orig.vertices, pre=it.pre_transform) for it in items}
    order = [it for it, d in inst.items for _ in range(d)][:20]
    placed = []
    for k, it in enumerate(order):
        rot = it.allowed_orientations
        rots = "continuous" if rot == "continuous" else np.array([0.0] if rot is None else rot, np.float32)
        s = J.PlacementSearch([(cde, it.shape, rots, None)], samples_per_round=256, uniform_rounds=6,
                              local_rounds=2, seed=1000 + k)
        s.run(graph=False)
        r = s.results()[0]
        if not r["found"]:
            continue
        dt = np.array([r["rotation"], r["tx"], r["ty"]], dtype=np.float32)
        assert O.query(oc, oitems[it.id], mats=J.compose(dt[None]), mode=1)["status"][0] == 0
        I.place_items(cde, inst, [(it.id, dt)], first_pk=len(placed))
        oc.register_item(it.id, it.shape_orig.vertices, dt, len(placed), pre=it.pre_transform)
        placed.append((it.id, dt))
    assert len(placed) >= 3
    box = cont.cde_bbox()
    hits = total = 0
    for i, it in enumerate(items):
        mats = J.compose(S.transforms(50 + i, 2048, box, 0))
        m = torch.from_numpy(mats).to(cuda)
        st = cde.detect_batch(it.shape, m).cpu().numpy()
        ref = O.query(oc, oitems[it.id], mats=mats, mode=3, keys_cap=16, threads=8)
        np.testing.assert_array_equal(st, ref["status"])
        cnt, hsh, keys = cde.collect_batch(it.shape, m, capacity=16)
        np.testing.assert_array_equal(cnt.cpu().numpy().astype(np.uint32), ref["count"])
        np.testing.assert_array_equal(hsh.cpu().numpy().view(np.uint64), ref["hash"])
        np.testing.assert_array_equal(keys.cpu().numpy().view(np.uint32), ref["keys"])
        hits, total = hits + int(st.sum()), total + len(st)
    assert 0 < hits < total
    lay = I.export_layout(cont, inst, placed)
    assert len(lay["placed_items"]) == len(placed) and 0 < lay["density"] < 1
