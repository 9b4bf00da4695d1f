"""Instances and layouts in the reference's JSON format.

* ``ExtItem`` / ``ExtContainer`` / ``ExtShape`` / ``ExtLayout`` (io/ext_repr.rs:6-121), strip
  packing (probs/spp/io/ext_repr.rs) and bin packing (probs/bpp/io/ext_repr.rs) instances, read
  with Python's json module (the reference uses serde_json).
* ``Importer`` (io/import.rs:19-260): shapes become SPolygons by the reference's rules (closing
  vertex stripped, consecutive near-duplicates eliminated, other duplicates rejected); items are
  centred on their centroid by a pre-transformation; containers become the engine the reference
  builds for an empty layout (Container::new, entities/container.rs:27-80): the Exterior hazard,
  then the quality zones in ascending quality, holes being the zones of quality 0.
* Strip and bin packing instance import (probs/spp/io/import.rs:12-60, probs/bpp/io/import.rs),
  solution import (placements through ext_to_int_transformation) and layout export
  (io/export.rs:6-45).

Shapes are imported unmodified.  The importer's shape modifications (simplification, offsetting,
closing narrow concavities: geometry/shape_modification.rs) are not part of this library; an
Importer asked for one raises NotImplementedError.  An offset of 0, which the reference skips
(geometry/original_shape.rs:27-32), is accepted.
"""
import json
import math

import numpy as np

from . import CDEConfig, CDEngine, Hazard, HazardEntity, SPolygon, _f32, _ptr
from . import _native as N
from ._native import JaguaError, check

F32 = np.float32
N_QUALITIES = 10                                     # entities/container.rs:89

# f32::to_radians / f32::to_degrees as Rust's core library defines them (the multiplier of
# to_radians is computed in f32, the one of to_degrees is the f32 literal of 180/pi)
_RAD = F32(F32(math.pi) / F32(180.0))
_DEG = F32(57.2957795130823208767981548141051703)


def to_radians(deg):
    return F32(F32(deg) * _RAD)


def to_degrees(rad):
    return F32(F32(rad) * _DEG)


def _bits(x):
    return int(np.array(x, dtype=F32).view(np.uint32))


def approx_eq(a, b):
    """float-cmp approx_eq!(f32, a, b) with the default margin (util/fpa.rs:18-22)."""
    a, b = F32(a), F32(b)
    if a == b:
        return True
    with np.errstate(over="ignore", invalid="ignore"):
        if abs(F32(a - b)) <= F32(1.1920929e-07):
            return True
    d = (_bits(a) - _bits(b)) & 0xFFFFFFFF
    d = d - (1 << 32) if d >= (1 << 31) else d
    ad = (1 << 31) - 1 if d == -(1 << 31) else abs(d)
    return ad <= 4


# ----------------------------------------------------------------------------- shapes
def eliminate_degenerate_vertices(points):
    """io/import.rs:235-260: removes every vertex approximately equal to its successor."""
    n = len(points)
    remove = [i for i in range(n)
              if approx_eq(points[i][0], points[(i + 1) % n][0]) and approx_eq(points[i][1], points[(i + 1) % n][1])]
    for i in sorted(remove, reverse=True):
        if i < len(points):
            del points[i]
    return points


def import_simple_polygon(data):
    """io/import.rs:198-212: points of an ExtSPolygon to an SPolygon."""
    pts = [(F32(x), F32(y)) for x, y in data]
    if len(pts) > 1 and pts[0][0] == pts[-1][0] and pts[0][1] == pts[-1][1]:
        pts.pop()
    eliminate_degenerate_vertices(pts)
    if len({(_bits(x), _bits(y)) for x, y in pts}) != len(pts):     # unique() hashes the bits
        raise JaguaError(-2, "Simple polygon has non-consecutive duplicate vertices")
    return SPolygon.new(np.array(pts, dtype=F32).reshape(-1, 2))


def rect_polygon(x_min, y_min, width, height):
    """Rect::try_new(x_min, y_min, x_min + width, y_min + height) then SPolygon::from(Rect)
    (geometry/primitives/rect.rs, simple_polygon.rs:330-344)."""
    x0, y0 = F32(x_min), F32(y_min)
    x1, y1 = F32(x0 + F32(width)), F32(y0 + F32(height))
    if not (x0 < x1 and y0 < y1):
        raise JaguaError(-2, "invalid rectangle")
    return SPolygon.new(np.array([[x0, y0], [x1, y0], [x1, y1], [x0, y1]], dtype=F32))


def _shape(ext, what, allow_polygon_holes):
    kind, data = ext["type"], ext.get("data")
    if kind == "rectangle":
        return rect_polygon(data["x_min"], data["y_min"], data["width"], data["height"]), []
    if kind == "simple_polygon":
        return import_simple_polygon(data), []
    if kind == "polygon":
        holes = [import_simple_polygon(h) for h in data.get("inner", [])] if allow_polygon_holes else []
        return import_simple_polygon(data["outer"]), holes
    if kind == "multi_polygon":
        raise JaguaError(-2, f"No support for multipolygon {what} shapes")
    raise JaguaError(-1, f"unknown shape type {kind!r}")


def centroid(poly):
    out = np.zeros(2, dtype=F32)
    check(N.lib().jg_spolygon_centroid(poly._h, _ptr(out)))
    return out


def dt_chain(a, invert_a, b):
    a, b = _f32(a), _f32(b)
    out = np.zeros(3, dtype=F32)
    check(N.lib().jg_dt_chain(_ptr(a), 1 if invert_a else 0, _ptr(b), _ptr(out)))
    return out


def ext_to_int_transformation(ext_dt, pre_dt):
    """io/import.rs:222-233: (rotation, tx, ty) of the original shape -> of the internal shape."""
    return dt_chain(pre_dt, True, ext_dt)


def int_to_ext_transformation(int_dt, pre_dt):
    """io/export.rs:34-45: (rotation, tx, ty) of the internal shape -> of the original shape."""
    return dt_chain(pre_dt, False, int_dt)


# ----------------------------------------------------------------------------- entities
class Item:
    """An imported item (entities/item.rs): the original shape, the pre-transformation that
    centres it, the internal shape used for collision detection (with its surrogate), the
    allowed rotations (None = only 0, "continuous", or a list of radians) and the minimum
    quality."""

    def __init__(self, id, shape_orig, pre_transform, shape, allowed_orientations, min_quality):
        self.id, self.shape_orig, self.pre_transform, self.shape = id, shape_orig, pre_transform, shape
        self.allowed_orientations, self.min_quality = allowed_orientations, min_quality

    @property
    def area(self):
        return F32(self.shape_orig.area)


class Container:
    """An imported container (entities/container.rs:13-80): the outer shape and the quality
    zones (quality -> list of SPolygons, holes in zone 0)."""

    def __init__(self, id, outer, zones, cde_config):
        self.id, self.outer, self.zones, self.cde_config = id, outer, zones, cde_config

    def hazards(self):
        """The base engine's hazards in the reference's order (container.rs:58-68,
        InferiorQualityZone::to_hazards :121-132)."""
        hz = [Hazard(HazardEntity.exterior(), self.outer, False)]
        for q in range(N_QUALITIES):
            for idx, s in enumerate(self.zones.get(q, [])):
                ent = HazardEntity.hole(idx) if q == 0 else HazardEntity.inferior_quality_zone(q, idx)
                hz.append(Hazard(ent, s, False))
        return hz

    def cde_bbox(self):
        """outer.bbox.inflate_to_square() (container.rs:67, rect.rs:93-109)."""
        x0, y0, x1, y1 = (F32(v) for v in self.outer.bbox)
        w, h = F32(x1 - x0), F32(y1 - y0)
        dx = dy = F32(0.0)
        if h < w:
            dy = F32(F32(w - h) / F32(2.0))
        elif w < h:
            dx = F32(F32(h - w) / F32(2.0))
        return (F32(x0 - dx), F32(y0 - dy), F32(x1 + dx), F32(y1 + dy))

    def base_cde(self, cde_config=None):
        """The engine of an empty layout in this container."""
        return CDEngine(self.cde_bbox(), self.hazards(), cde_config or self.cde_config)

    @property
    def area(self):
        """Container::area (container.rs:83-86): outer minus the zones of quality 0."""
        a = F32(self.outer.area)
        holes = F32(0.0)
        for s in self.zones.get(0, []):
            holes = F32(holes + F32(s.area))
        return F32(a - holes) if self.zones.get(0) else a


class Importer:
    """io/import.rs:19-44."""

    def __init__(self, cde_config=None, simplify_tolerance=None, min_item_separation=None,
                 narrow_concavity_cutoff_ratio=None):
        if simplify_tolerance is not None or narrow_concavity_cutoff_ratio is not None or \
                (min_item_separation is not None and F32(min_item_separation) / F32(2.0) != 0.0):
            raise NotImplementedError("shape modification (simplify / offset / close concavities) is not supported; "
                                      "import the unmodified shapes")
        self.cde_config = cde_config or CDEConfig()

    def import_item(self, ext):
        """Importer::import_item (io/import.rs:46-96)."""
        shape, _ = _shape(ext["shape"], "item", False)
        c = centroid(shape)
        pre = np.array([0.0, -c[0], -c[1]], dtype=F32)        # centering_transformation (:214-218)
        internal = shape.transform_clone(pre)
        internal.generate_surrogate(self.cde_config.item_surrogate_config)
        ao = ext.get("allowed_orientations")
        if ao is None:
            rot = "continuous"
        elif len(ao) == 0 or (len(ao) == 1 and F32(ao[0]) == 0.0):
            rot = None
        else:
            rot = [to_radians(a) for a in ao]
        return Item(int(ext["id"]), shape, pre, internal, rot, ext.get("min_quality"))

    def import_container(self, ext):
        """Importer::import_container (io/import.rs:98-196)."""
        zones_ext = ext.get("zones", [])
        if not all(int(z["quality"]) < N_QUALITIES for z in zones_ext):
            raise JaguaError(-1, "All quality zones must have lower quality than N_QUALITIES")
        outer, holes = _shape(ext["shape"], "container", True)
        zones = {}
        for q in range(N_QUALITIES):
            shapes = []
            for z in zones_ext:
                if int(z["quality"]) != q:
                    continue
                k = z["shape"]["type"]
                if k not in ("rectangle", "simple_polygon"):
                    raise JaguaError(-2, f"No support for {k} quality zone shapes")
                shapes.append(_shape(z["shape"], "zone", False)[0])
            if q == 0:
                shapes += holes
            if shapes:
                zones[q] = shapes
        return Container(int(ext["id"]), outer, zones, self.cde_config)


# ----------------------------------------------------------------------------- instances
def read_json(src):
    """A path, a JSON string or an already parsed object."""
    if isinstance(src, (dict, list)):
        return src
    if isinstance(src, str) and src.lstrip().startswith("{"):
        return json.loads(src)
    with open(src) as f:
        return json.load(f)


class SPInstance:
    """Strip packing instance (probs/spp/entities): items with demands (ids 0..n-1, demand > 0)
    and a strip of fixed height whose starting width gives 100% density."""

    def __init__(self, name, items, strip_height, base_width, cde_config):
        self.name, self.items, self.strip_height = name, items, strip_height
        self.base_width, self.cde_config = base_width, cde_config

    def item(self, item_id):
        return self.items[item_id][0]

    def strip(self, width=None):
        """The strip as a Container (probs/spp/entities/strip.rs:40-56): the rectangle
        (0, 0) - (width, strip_height); id = width's bit pattern."""
        w = F32(self.base_width if width is None else width)
        if not w > 0:
            raise JaguaError(-1, "strip width must be positive")
        outer = SPolygon.new(np.array([[0, 0], [w, 0], [w, self.strip_height], [0, self.strip_height]], dtype=F32))
        return Container(_bits(w), outer, {}, self.cde_config)


def import_spp_instance(importer, src):
    """probs/spp/io/import.rs:12-60."""
    ext = read_json(src)
    items = sorted(((importer.import_item(e), int(e["demand"])) for e in ext["items"]), key=lambda t: t[0].id)
    items = [t for t in items if t[1] > 0]
    if [it.id for it, _ in items] != list(range(len(items))):
        raise JaguaError(-1, "All items should have consecutive IDs starting from 0")
    if not items:
        raise JaguaError(-1, "ExtSPInstance must have at least one item with positive demand")
    total = F32(0.0)
    for it, d in items:
        total = F32(total + F32(it.area * F32(d)))
    h = F32(ext["strip_height"])
    if not h > 0:
        raise JaguaError(-1, "strip height must be positive")
    return SPInstance(ext.get("name", ""), items, h, F32(total / h), importer.cde_config)


class BPInstance:
    """Bin packing instance (probs/bpp/entities): items with demands and bins (container,
    stock, cost)."""

    def __init__(self, name, items, bins):
        self.name, self.items, self.bins = name, items, bins

    def item(self, item_id):
        return self.items[item_id][0]


def import_bpp_instance(importer, src):
    """probs/bpp/io/import.rs: items as for strip packing, bins as containers."""
    ext = read_json(src)
    items = sorted(((importer.import_item(e), int(e["demand"])) for e in ext["items"]), key=lambda t: t[0].id)
    items = [t for t in items if t[1] > 0]
    if [it.id for it, _ in items] != list(range(len(items))):
        raise JaguaError(-1, "All items should have consecutive IDs starting from 0")
    bins = sorted(((importer.import_container(b), int(b["stock"]), int(b["cost"])) for b in ext["bins"]),
                  key=lambda t: t[0].id)
    if [b.id for b, _, _ in bins] != list(range(len(bins))):
        raise JaguaError(-1, "All bins should have consecutive IDs starting from 0")
    return BPInstance(ext.get("name", ""), items, bins)


# ----------------------------------------------------------------------------- layouts
def place_items(cde, instance, placements, first_pk=0):
    """Registers placed items (item_id, internal (rotation, tx, ty)) into an engine as
    Layout::place_item does (entities/layout.rs): the item's internal shape transformed by the
    placement, a dynamic PlacedItem hazard.  Returns the hazard keys."""
    keys = []
    for k, (item_id, dt) in enumerate(placements):
        it = instance.item(item_id)
        d = _f32(dt)
        shape = it.shape.transform_clone(d)
        pk = first_pk + k
        keys.append(cde.register_hazard(Hazard(HazardEntity.placed_item(item_id, d, pk), shape, True)))
    return keys


def import_spp_solution(instance, src):
    """The placements of an ExtSPSolution (probs/spp/io/import.rs:63-80): strip width and
    (item_id, internal transformation) per placed item, via ext_to_int_transformation."""
    ext = read_json(src)
    out = []
    for p in ext["layout"]["placed_items"]:
        t = p["transformation"]
        ext_dt = np.array([to_radians(t["rotation"]), t["translation"][0], t["translation"][1]], dtype=F32)
        it = instance.item(int(p["item_id"]))
        out.append((it.id, ext_to_int_transformation(ext_dt, it.pre_transform)))
    return F32(ext["strip_width"]), out


def export_layout(container, instance, placements):
    """export_layout_snapshot (io/export.rs:6-28) as a JSON-ready dict: container id, placed
    items with their transformation of the original shape (rotation in degrees), density."""
    placed, area = [], F32(0.0)
    for item_id, dt in placements:
        it = instance.item(item_id)
        e = int_to_ext_transformation(dt, it.pre_transform)
        placed.append({"item_id": int(item_id),
                       "transformation": {"rotation": float(to_degrees(e[0])),
                                          "translation": [float(e[1]), float(e[2])]}})
        area = F32(area + it.area)
    return {"container_id": int(container.id), "placed_items": placed, "density": float(F32(area / container.area))}
