"""Synthetic nesting workloads (libjagua_synth.so) and their assembly into CDEngines.

A Scene is a container (plus holes) and placed items given as local polygons with their
DTransformation; `build_cde` registers it exactly as jagua-rs builds a layout:
Container::new (entities/container.rs:57-70: Exterior then holes, on the square-inflated
container bbox), then Layout::place_item per item (entities/layout.rs:73-85).
"""
import concurrent.futures as cf
import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from . import (CDEConfig, CDEngine, Hazard, HazardEntity, SPSurrogateConfig, SPolygon)

F32 = np.float32


@dataclass
class Scene:
    container: np.ndarray
    holes: list = field(default_factory=list)
    items: list = field(default_factory=list)       # (local vertices [n,2] f32, dt [3] f32)
    depth: int = 5
    cd_threshold: int = 16

    def cde_bbox(self):
        """Container::new: outer.bbox.inflate_to_square() (container.rs:68, rect.rs:93-109), in f32."""
        c = self.container.astype(F32)
        x0, y0 = F32(c[:, 0].min()), F32(c[:, 1].min())
        x1, y1 = F32(c[:, 0].max()), F32(c[:, 1].max())
        w, h = F32(x1 - x0), F32(y1 - y0)
        dx = dy = F32(0.0)
        if h < w:
            dy = F32(F32(w - h) / F32(2.0))
        elif w < h:
            dx = F32(F32(h - w) / F32(2.0))
        return np.array([F32(x0 - dx), F32(y0 - dy), F32(x1 + dx), F32(y1 + dy)], dtype=F32)

    @property
    def n_edges(self):
        return sum(len(v) for v, _ in self.items) + len(self.container) + sum(len(h) for h in self.holes)


def layout_params(**kw):
    p = N.LayoutParams()
    defaults = dict(seed=0, container_kind=0, cw=1000.0, ch=400.0, container_nv_lo=4, container_nv_hi=4, n_holes=0,
                    hole_kind=0, hole_lo=50.0, hole_hi=300.0, hole_nv_lo=8, hole_nv_hi=64, n_items=40,
                    item_nv_lo=12, item_nv_hi=12, item_r_lo=20.0, item_r_hi=40.0, item_shape=0, min_sep=0.0,
                    rot_kind=0, placement=0, max_attempts=20000, lbf_samples=2000)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, v)
    return p


def generate_scene(depth=5, cd_threshold=16, **kw):
    L = N.synth()
    p = layout_params(**kw)
    h = L.js_layout_generate(C.byref(p))
    try:
        sz = (C.c_uint32 * 3)()
        L.js_layout_sizes(h, sz)
        cont = np.empty((sz[0], 2), dtype=F32)
        L.js_layout_container(h, cont.ctypes.data_as(N.c_f32p))
        holes = []
        for i in range(sz[1]):
            a = np.empty((L.js_layout_hole_n(h, i), 2), dtype=F32)
            L.js_layout_hole(h, i, a.ctypes.data_as(N.c_f32p))
            holes.append(a)
        items = []
        for i in range(sz[2]):
            a = np.empty((L.js_layout_item_n(h, i), 2), dtype=F32)
            dt = np.empty(3, dtype=F32)
            L.js_layout_item(h, i, a.ctypes.data_as(N.c_f32p), dt.ctypes.data_as(N.c_f32p))
            items.append((a, dt))
    finally:
        L.js_layout_free(h)
    return Scene(cont, holes, items, depth, cd_threshold)


def polygon(seed, nv_lo, nv_hi, r_lo, r_hi, shape=0):
    xy = np.empty((nv_hi, 2), dtype=F32)
    n = N.synth().js_polygon(seed, nv_lo, nv_hi, r_lo, r_hi, shape, xy.ctypes.data_as(N.c_f32p))
    return xy[:n].copy()


def transforms(seed, n, box, rot_kind=0):
    """[n, 3] DTransformations: (rotation, tx, ty) uniform in box; rot_kind 1 = multiples of 45 deg."""
    dt = np.empty((n, 3), dtype=F32)
    N.synth().js_transforms(seed, n, float(box[0]), float(box[1]), float(box[2]), float(box[3]), rot_kind,
                            dt.ctypes.data_as(N.c_f32p))
    return dt


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def build_cde(scene, surrogate_config=None):
    """The product CDEngine of a scene (Exterior, holes, then placed items in order)."""
    cfg = CDEConfig(scene.depth, scene.cd_threshold, surrogate_config or SPSurrogateConfig())
    cont = SPolygon.new(scene.container)
    haz = [Hazard(HazardEntity.exterior(), cont, False)]
    haz += [Hazard(HazardEntity.hole(i), SPolygon.new(h), False) for i, h in enumerate(scene.holes)]
    cde = CDEngine(scene.cde_bbox(), haz, cfg)
    with cf.ThreadPoolExecutor(_threads()) as ex:
        shapes = list(ex.map(lambda it: SPolygon.new(it[0]).transform_clone(it[1]), scene.items))
    keys = []
    for i, ((_, dt), s) in enumerate(zip(scene.items, shapes)):
        keys.append(cde.register_hazard(Hazard(HazardEntity.placed_item(i, dt, i), s, True)))
    return cde, keys


def make_items(polys, surrogate_config=None):
    """Query items: SPolygon::new + generate_surrogate for each local polygon (threaded)."""
    cfg = surrogate_config or SPSurrogateConfig()
    with cf.ThreadPoolExecutor(_threads()) as ex:
        return list(ex.map(lambda p: SPolygon.new(p).generate_surrogate(cfg), polys))


# ---------------------------------------------------------------------------- workloads
# Configurations of the collision-query benchmark (sizes in the docstrings).

def config1_scene(seed=0):
    """1000x400 rectangle, 40 twelve-vertex stars (radii U[20,40]) placed left-bottom-fill
    with rotations in multiples of 45 degrees; quadtree depth 5."""
    return generate_scene(depth=5, seed=seed, container_kind=0, cw=1000.0, ch=400.0, n_items=40, item_nv_lo=12,
                          item_nv_hi=12, item_r_lo=20.0, item_r_hi=40.0, item_shape=0, rot_kind=1, placement=1,
                          lbf_samples=4000, min_sep=0.0)


def config1_queries(seed=0, n=4096):
    """One 16-vertex star (radii U[15,35]) at n transforms over the container, rotations k*45 deg."""
    q = polygon(seed * 7919 + 11, 16, 16, 15.0, 35.0, 0)
    return [q], transforms(seed * 7919 + 12, n, (0.0, 0.0, 1000.0, 400.0), rot_kind=1)


def config2_scene(seed=2, n_items=2000):
    """4000x1500 rectangle with 5 rectangular holes (sides U[50,300]); n_items star or concave
    polygons (radius 10-30), log-uniform 8-256 vertices, minimum separation 0.5; quadtree depth 8."""
    return generate_scene(depth=8, seed=seed, container_kind=0, cw=4000.0, ch=1500.0, n_holes=5, hole_kind=0,
                          hole_lo=50.0, hole_hi=300.0, n_items=n_items, item_nv_lo=8, item_nv_hi=256,
                          item_r_lo=10.0, item_r_hi=30.0, item_shape=2, min_sep=0.5, rot_kind=0, placement=0,
                          max_attempts=20000)


def config2_queries(seed=2, n_items=64, n_per_item=16384):
    """n_items query polygons (log-uniform 8-512 vertices) with n_per_item continuous-rotation
    transforms each over the container bbox."""
    polys = [polygon(seed * 100003 + i, 8, 512, 10.0, 30.0, 2) for i in range(n_items)]
    dts = [transforms(seed * 100003 + 7 * i + 1, n_per_item, (0.0, 0.0, 4000.0, 1500.0), 0) for i in range(n_items)]
    return polys, dts


def config4_workload(seed=4, n_items=4096, width=2000.0, height=800.0, r_lo=6.0, r_hi=11.5, jitter=3.0,
                     poles_lo=16, poles_hi=64, depth=8):
    """Differentiable overlap: n_items star polygons (12-24 vertices, radius U[r_lo, r_hi]) each
    carrying 16-64 surrogate poles (SPSurrogateConfig with n_pole_limits (n, 0.0): exactly n
    inscribed circles), on a jittered grid over a width x height rectangle.  Returns a dict
    with local poles [total, 3], pole_off, pose [n, 3], bounding radius, candidate pairs (grid
    broad phase at quadtree depth `depth`) and the container's 4 edges."""
    from . import candidate_pairs
    rng = np.random.default_rng(seed)
    nx = int(np.ceil(np.sqrt(n_items * width / height)))
    ny = int(np.ceil(n_items / nx))
    gx, gy = width / nx, height / ny
    polys, n_poles = [], []
    for i in range(n_items):
        polys.append(polygon(seed * 1_000_003 + i, 12, 24, r_lo, r_hi, 0))
        n_poles.append(int(rng.integers(poles_lo, poles_hi + 1)))

    def mk(a):
        p, n = a
        sp = SPolygon.new(p).generate_surrogate(SPSurrogateConfig(((n, 0.0), (n, 0.0), (n, 0.0)), n_ff_poles=2))
        return sp.poles

    with cf.ThreadPoolExecutor(_threads()) as ex:
        poles = list(ex.map(mk, zip(polys, n_poles)))
    off = np.zeros(n_items + 1, dtype=np.uint32)
    off[1:] = np.cumsum([len(p) for p in poles])
    allp = np.concatenate(poles).astype(F32)
    pose = np.zeros((n_items, 3), dtype=F32)
    for i in range(n_items):
        cx, cy = (i % nx + 0.5) * gx, (i // nx + 0.5) * gy
        pose[i] = (cx + rng.uniform(-jitter, jitter), cy + rng.uniform(-jitter, jitter), rng.uniform(0, 2 * np.pi))
    rad = np.array([np.max(np.hypot(p[:, 0], p[:, 1]) + p[:, 2]) for p in poles], dtype=F32)
    pairs = candidate_pairs(pose, rad, [0.0, 0.0, width, height], depth)
    cedges = np.array([[0, 0, width, 0], [width, 0, width, height], [width, height, 0, height],
                       [0, height, 0, 0]], dtype=F32)
    return dict(poles=allp, off=off, pose=pose, rad=rad, pairs=pairs, cedges=cedges, polys=polys)


def overlap_engine(w, depth=8, cd_threshold=16):
    """A CDEngine holding the overlap workload's items as placed hazards at their poses
    (DTransformation (theta, x, y) of pose (x, y, theta)) in its width x height rectangle, for
    quadtree pair candidates.  Returns (engine, item keys)."""
    ce = w["cedges"]
    width, height = float(ce[:, [0, 2]].max()), float(ce[:, [1, 3]].max())
    cont = np.array([[0, 0], [width, 0], [width, height], [0, height]], dtype=F32)
    items = [(w["polys"][i], np.array([w["pose"][i, 2], w["pose"][i, 0], w["pose"][i, 1]], dtype=F32))
             for i in range(len(w["polys"]))]
    return build_cde(Scene(cont, [], items, depth, cd_threshold))


def _log_uniform_int(rng, lo, hi):
    return int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))


def config3_layout(seed=3, layout=0):
    """One of the many independent bin-packing layouts: a star-shaped container of 64-2048
    vertices (radii U[0.7R, R], R = 40 sqrt(n_items)) holding n_items ~ log-uniform [50, 400]
    star or concave polygons of 8-256 vertices (radius 10-30); quadtree depth 7."""
    rng = np.random.default_rng([seed, layout])
    n_items = _log_uniform_int(rng, 50, 400)
    big_r = 40.0 * float(np.sqrt(n_items))
    return generate_scene(depth=7, seed=(seed << 20) + layout, container_kind=1, cw=0.7 * big_r, ch=big_r,
                          container_nv_lo=64, container_nv_hi=2048, n_items=n_items, item_nv_lo=8,
                          item_nv_hi=256, item_r_lo=10.0, item_r_hi=30.0, item_shape=2, min_sep=0.0, rot_kind=0,
                          placement=0, max_attempts=4000)


def config3_queries(scene, seed=3, layout=0, n_items=16, n_per_item=4096):
    """n_items query polygons (8-256 vertices, radius 10-30) per layout, each at n_per_item
    continuous-rotation transforms over the container's bbox."""
    c = scene.container
    box = (float(c[:, 0].min()), float(c[:, 1].min()), float(c[:, 0].max()), float(c[:, 1].max()))
    base = ((seed << 20) + layout) * 1009
    polys = [polygon(base + 2 * i, 8, 256, 10.0, 30.0, 2) for i in range(n_items)]
    dts = [transforms(base + 2 * i + 1, n_per_item, box, 0) for i in range(n_items)]
    return polys, dts


def config3_workload(seed=3, n_layouts=256, n_items=16, n_per_item=4096):
    """Many layouts in one launch: n_layouts config3 layouts (generated in parallel), their
    query polygons and transforms, and a HashedFilter per layout that deems a random 10% of
    the hazards irrelevant per query.  Returns (scenes, [(polys, dts)], filters)."""
    from . import HashedFilter
    with cf.ThreadPoolExecutor(_threads()) as ex:
        scenes = list(ex.map(lambda l: config3_layout(seed, l), range(n_layouts)))
    queries = [config3_queries(sc, seed, l, n_items, n_per_item) for l, sc in enumerate(scenes)]
    filters = [HashedFilter((seed << 16) + l, 0.1) for l in range(n_layouts)]
    return scenes, queries, filters


def config5_scene(seed=5, n_items=20000, item_r_lo=6.0, item_r_hi=12.0, width=6000.0, height=2000.0, n_holes=20):
    """Full scale: a 6000x2000 rectangle with 20 star-shaped holes of 8-64 vertices (radius
    60-200) and n_items concave polygons (3-octave value-noise radii, amplitude 0.3) of
    log-uniform 64-1024 vertices; quadtree depth 10."""
    return generate_scene(depth=10, seed=seed, container_kind=0, cw=width, ch=height, n_holes=n_holes, hole_kind=1,
                          hole_lo=60.0, hole_hi=200.0, hole_nv_lo=8, hole_nv_hi=64, n_items=n_items, item_nv_lo=64,
                          item_nv_hi=1024, item_r_lo=item_r_lo, item_r_hi=item_r_hi, item_shape=1, min_sep=0.0,
                          rot_kind=0, placement=0, max_attempts=2000)


def config5_queries(seed=5, n_items=256, n_per_item=65536, r_lo=6.0, r_hi=12.0, width=6000.0, height=2000.0,
                    nv_hi=1024):
    """n_items query polygons of mixed shape and size (8-1024 vertices) at n_per_item
    continuous-rotation transforms over the 6000x2000 container."""
    polys = [polygon(seed * 100003 + i, 8, nv_hi, r_lo, r_hi, 2) for i in range(n_items)]
    dts = [transforms(seed * 100003 + 7 * i + 1, n_per_item, (0.0, 0.0, width, height), 0) for i in range(n_items)]
    return polys, dts
