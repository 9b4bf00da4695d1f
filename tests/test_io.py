"""Instance import and layout export (io/import.rs, io/export.rs, probs/spp/io, probs/bpp/io)
on the reference's own instance files (tests/golden/assets: the instances lbf/tests/tests.rs
runs), against the oracle's restatement of the import, and pinned by hand-made cases and
closed forms: the import rules on small polygons, centroids of shapes whose centroid is exact,
translation-only transformation chains, and export -> import round trips."""
import glob
import json
import os

import numpy as np
import pytest

import jagua_amd as J
from jagua_amd import io as I
from oracle import oracle as O

from helpers import assert_same_tree

ASSETS = os.path.join(os.path.dirname(__file__), "golden", "assets")
SPP = sorted(os.path.basename(f)[:-5] for f in glob.glob(os.path.join(ASSETS, "*.json"))
             if "bins" not in json.load(open(f)))
F32 = np.float32


def _load(name):
    return json.load(open(os.path.join(ASSETS, name + ".json")))


def _points(shape):
    """The outer points of an ExtShape as given in the file."""
    d = shape["data"]
    if shape["type"] == "simple_polygon":
        return np.array(d, dtype=F32)
    if shape["type"] == "polygon":
        return np.array(d["outer"], dtype=F32)
    x0, y0 = F32(d["x_min"]), F32(d["y_min"])
    x1, y1 = F32(x0 + F32(d["width"])), F32(y0 + F32(d["height"]))
    return np.array([[x0, y0], [x1, y0], [x1, y1], [x0, y1]], dtype=F32)


def _u(a):
    return np.asarray(a, dtype=F32).view(np.uint32)


@pytest.mark.parametrize("name", SPP)
def test_spp_items_match_oracle(name):
    """Every item of every strip packing instance: the imported original shape and its centroid,
    the centering pre-transformation and the internal shape (vertices, area, bbox, POI, poles)
    equal the oracle's import bit for bit; demands and the 100%-density strip width follow the
    file."""
    ext = _load(name)
    inst = I.import_spp_instance(I.Importer(), ext)
    by_id = {int(e["id"]): e for e in ext["items"]}
    total = F32(0.0)
    for it, demand in inst.items:
        e = by_id[it.id]
        assert demand == int(e["demand"])
        ov, oc = O.import_polygon(_points(e["shape"]))
        assert np.array_equal(_u(it.shape_orig.vertices), _u(ov))
        assert np.array_equal(_u(I.centroid(it.shape_orig)), _u(oc))
        assert np.array_equal(_u(it.pre_transform), _u([0.0, -oc[0], -oc[1]]))
        oi = O.OracleItem(ov, pre=it.pre_transform)
        xy, meta, poles = oi.info()
        inf = it.shape.info()
        assert np.array_equal(_u(it.shape.vertices), _u(xy))
        mine = np.array([inf.area, inf.diameter, *inf.bbox, *inf.poi], dtype=F32)
        assert np.array_equal(_u(mine), _u(meta))
        assert np.array_equal(_u(it.shape.poles[:, :3]), _u(poles))
        total = F32(total + F32(it.area * F32(demand)))
        ao = e.get("allowed_orientations")
        if ao is None:
            assert it.allowed_orientations == "continuous"
        elif len(ao) == 0 or ao == [0.0]:
            assert it.allowed_orientations is None
        else:
            assert [float(a) for a in it.allowed_orientations] == [float(I.to_radians(a)) for a in ao]
    assert inst.base_width == F32(total / F32(ext["strip_height"]))
    assert inst.strip_height == F32(ext["strip_height"])


def test_bpp_bins_and_engine_match_oracle():
    """baldacci1: bins with quality zones become engines whose hazards (Exterior, holes as zone 0,
    then zones by quality) and quadtrees equal the oracle's engine on the same hazards."""
    ext = _load("baldacci1")
    inst = I.import_bpp_instance(I.Importer(), ext)
    assert len(inst.bins) == len(ext["bins"])
    for (cont, stock, cost), eb in zip(inst.bins, sorted(ext["bins"], key=lambda b: b["id"])):
        assert (stock, cost) == (int(eb["stock"]), int(eb["cost"]))
        hz = cont.hazards()
        n_holes = len(eb["shape"]["data"].get("inner", [])) if eb["shape"]["type"] == "polygon" else 0
        assert len(hz) == 1 + len(eb.get("zones", [])) + n_holes
        assert sum(h.entity.kind == J.ENTITY_HOLE for h in hz) == n_holes + sum(z["quality"] == 0 for z in eb.get("zones", []))
        qualities = [h.entity.a if h.entity.kind == J.ENTITY_INFERIOR_QUALITY_ZONE else 0 for h in hz[1:]]
        assert qualities == sorted(qualities)
        for depth in (0, 5):
            cfg = J.CDEConfig(depth, 16)
            cde = cont.base_cde(cfg)
            oc = O.OracleCDE.from_hazards(cont.cde_bbox(), depth, 16,
                                          [(h.entity.kind, h.entity.a, h.entity.b, h.shape.vertices) for h in hz])
            assert_same_tree(cde, oc)
    mq = {int(e["id"]): e.get("min_quality") for e in ext["items"]}
    assert all(it.min_quality == mq[it.id] for it, _ in inst.items)


def test_import_rules():
    """import_simple_polygon (io/import.rs:198-260): the closing vertex is stripped, a vertex
    approximately equal to its successor is removed, other duplicates (by bit pattern, as
    itertools' unique() hashes Point) are rejected; product and oracle agree on each case."""
    sq = [(0, 0), (4, 0), (4, 3), (0, 3)]
    cases = {
        "closed": (sq + [(0, 0)], 4),
        "near_duplicate": ([(0, 0), (4, 0), (4.0000005, 0), (4, 3), (0, 3)], 4),
        "exact_duplicate": ([(0, 0), (4, 0), (4, 0), (4, 3), (0, 3)], 4),
        "signed_zero": ([(0.0, 0.0), (4, 0), (4, 4), (-0.0, 0.0), (-4, 4)], 5),
        "repeat": ([(0.0, 0.0), (4, 0), (4, 4), (0.0, 0.0), (-4, 4)], None),
    }
    for name, (pts, n) in cases.items():
        o = O.import_polygon(np.array(pts, dtype=F32))
        if n is None:
            assert o is None, name
            with pytest.raises(J.JaguaError, match="duplicate"):
                I.import_simple_polygon(pts)
            continue
        p = I.import_simple_polygon(pts)
        assert p.n_vertices == n == len(o[0]), name
        assert np.array_equal(_u(p.vertices), _u(o[0])), name
    p = I.import_simple_polygon(cases["near_duplicate"][0])
    assert F32(4.0000005) in p.vertices[:, 0] and not (p.vertices[:, 0] == 4.0).all()
    r = I.rect_polygon(1, 2, 3, 4)
    assert np.array_equal(r.vertices, np.array([[1, 2], [4, 2], [4, 6], [1, 6]], dtype=F32))


def test_centroid_closed_forms():
    """SPolygon::centroid (simple_polygon.rs:173-192) where every step is exact: rectangles and
    a right triangle; product and oracle agree on every asset shape."""
    for pts, c in [([(0, 0), (4, 0), (4, 2), (0, 2)], (2, 1)), ([(0, 0), (3, 0), (0, 3)], (1, 1)),
                   ([(-6, -2), (2, -2), (2, 6), (-6, 6)], (-2, 2))]:
        p = J.SPolygon.new(np.array(pts, dtype=F32))
        assert tuple(float(v) for v in I.centroid(p)) == c
        assert tuple(float(v) for v in O.import_polygon(np.array(pts, dtype=F32))[1]) == c


def test_transformation_chain():
    """ext_to_int / int_to_ext (io/import.rs:222-233, io/export.rs:34-45): equal to the oracle bit
    for bit; translation-only chains are exact sums; a round trip returns the transformation
    within f32 rounding."""
    rng = np.random.default_rng(0)
    for _ in range(500):
        a = np.array([rng.uniform(-7, 7), rng.uniform(-500, 500), rng.uniform(-500, 500)], dtype=F32)
        b = np.array([rng.uniform(-7, 7), rng.uniform(-500, 500), rng.uniform(-500, 500)], dtype=F32)
        assert np.array_equal(_u(I.ext_to_int_transformation(b, a)), _u(O.ext_to_int(b, a)))
        assert np.array_equal(_u(I.int_to_ext_transformation(b, a)), _u(O.int_to_ext(b, a)))
        back = I.ext_to_int_transformation(I.int_to_ext_transformation(b, a), a)
        assert abs(np.angle(np.exp(1j * (float(back[0]) - float(b[0]))))) < 1e-4
        assert np.allclose(back[1:], b[1:], atol=2e-3)
    for pre, ext in [((0, -1.5, 2.25), (0, 10, 20)), ((0, 3, -4), (0, -0.5, 0.125))]:
        pre, ext = np.array(pre, dtype=F32), np.array(ext, dtype=F32)
        assert np.array_equal(I.ext_to_int_transformation(ext, pre), [0, ext[1] - pre[1], ext[2] - pre[2]])
        assert np.array_equal(I.int_to_ext_transformation(ext, pre), [0, ext[1] + pre[1], ext[2] + pre[2]])
    assert abs(float(I.to_radians(180)) - np.pi) < 1e-6 and abs(float(I.to_degrees(np.pi)) - 180) < 1e-4
    assert I.to_radians(90) == F32(F32(90) * F32(F32(np.pi) / F32(180)))


def test_export_import_round_trip():
    """export_layout (io/export.rs) then import_spp_solution (probs/spp/io/import.rs:63-80):
    the placements come back within f32 rounding and the density is the placed area over the
    container's."""
    inst = I.import_spp_instance(I.Importer(), _load("shirts"))
    cont = inst.strip(F32(inst.base_width * F32(1.5)))
    rng = np.random.default_rng(1)
    placed = [(it.id, np.array([rng.uniform(0, 6.28), rng.uniform(0, 50), rng.uniform(0, 40)], dtype=F32))
              for it, _ in inst.items]
    lay = I.export_layout(cont, inst, placed)
    sol = json.loads(json.dumps({"strip_width": float(F32(inst.base_width * F32(1.5))), "layout": lay,
                                 "density": lay["density"], "run_time_sec": 0}))
    width, back = I.import_spp_solution(inst, sol)
    assert width == F32(inst.base_width * F32(1.5))
    for (i0, d0), (i1, d1) in zip(placed, back):
        assert i0 == i1
        assert abs(np.angle(np.exp(1j * (float(d1[0]) - float(d0[0]))))) < 1e-4
        assert np.allclose(d1[1:], d0[1:], atol=1e-3)
    area = F32(0.0)
    for it, _ in inst.items:
        area = F32(area + it.area)
    assert lay["density"] == float(F32(area / cont.area))


def test_importer_refuses_shape_modification():
    with pytest.raises(NotImplementedError):
        I.Importer(simplify_tolerance=0.001)
    with pytest.raises(NotImplementedError):
        I.Importer(min_item_separation=0.5)
    I.Importer(min_item_separation=0.0)
    with pytest.raises(J.JaguaError):
        I.import_spp_instance(I.Importer(), {"name": "x", "strip_height": 10.0, "items": [
            {"id": 1, "demand": 1, "shape": {"type": "rectangle", "data": {"x_min": 0, "y_min": 0, "width": 1,
                                                                           "height": 1}}}]})
