"""jagua_amd — the collision-query hot path of jagua-rs on one MI355X (gfx950).

Python mirror of the reference API over the C ABI in include/jagua_amd.h:

    SPolygon, SPSurrogateConfig         geometry/primitives/simple_polygon.rs, fail_fast/sp_surrogate.rs
    DTransformation                     geometry/d_transformation.rs
    HazardEntity, Hazard, HazKey        collision_detection/hazards/hazard.rs
    NoFilter, HazKeyFilter, HashedFilter collision_detection/hazards/filter.rs
    CDEConfig, CDEngine, CDESnapshot    collision_detection/cd_engine.rs

Queries are batched: one call evaluates a whole array of transforms of one item on the GPU
(lbf samples them one by one in lbf/src/opt/search.rs:41-101).  Transforms and outputs are
torch tensors on the GPU (device plumbing only) or numpy arrays (copied in and out).
"""
import ctypes as C

import numpy as np

from . import _native as N
from ._native import JaguaError, check

Q_CLEAR, Q_COLLISION, Q_PANIC_FILTERED_ENTIRE = 0, 1, 2
QUERY_SURROGATE, QUERY_POLY = 1, 2

ENTITY_PLACED_ITEM, ENTITY_EXTERIOR, ENTITY_HOLE, ENTITY_INFERIOR_QUALITY_ZONE = 0, 1, 2, 3


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, t=C.c_float):
    return a.ctypes.data_as(C.POINTER(t))


class SPSurrogateConfig:
    """SPSurrogateConfig (sp_surrogate.rs:120-131); defaults are LBFConfig::default (lbf/src/config.rs:36-40)."""

    def __init__(self, n_pole_limits=((100, 0.0), (20, 0.75), (10, 0.90)), n_ff_poles=2, n_ff_piers=0):
        self.n_pole_limits = tuple((int(n), float(c)) for n, c in n_pole_limits)
        self.n_ff_poles = int(n_ff_poles)
        self.n_ff_piers = int(n_ff_piers)

    def _c(self):
        c = N.SurrogateConfig()
        for i, (n, cov) in enumerate(self.n_pole_limits):
            c.n_pole_limits_n[i] = n
            c.n_pole_limits_cov[i] = cov
        c.n_ff_poles = self.n_ff_poles
        c.n_ff_piers = self.n_ff_piers
        return c


class CDEConfig:
    """CDEConfig (cd_engine.rs:345-353); defaults from LBFConfig::default (lbf/src/config.rs:33-41)."""

    def __init__(self, quadtree_depth=5, cd_threshold=16, item_surrogate_config=None):
        self.quadtree_depth = int(quadtree_depth)
        self.cd_threshold = int(cd_threshold)
        self.item_surrogate_config = item_surrogate_config or SPSurrogateConfig()

    def _c(self):
        c = N.CDEConfigC()
        c.quadtree_depth = self.quadtree_depth
        c.cd_threshold = self.cd_threshold
        c.item_surrogate_config = self.item_surrogate_config._c()
        return c


class DTransformation:
    """DTransformation (d_transformation.rs:11-47): a rotation followed by a translation."""

    def __init__(self, rotation, translation):
        self.rotation = float(np.float32(rotation))
        self.translation = (float(np.float32(translation[0])), float(np.float32(translation[1])))

    def as_array(self):
        return np.array([self.rotation, *self.translation], dtype=np.float32)

    def compose(self):
        """Transformation matrix rows (m00 m01 m02 m10 m11 m12), transformation.rs:131-139."""
        return compose(self.as_array()[None])[0]


def compose(dt):
    """DTransformation::compose for an array [n, 3] of (rotation, tx, ty) -> [n, 6] matrices."""
    dt = _f32(dt).reshape(-1, 3)
    out = np.empty((dt.shape[0], 6), dtype=np.float32)
    check(N.lib().jg_transforms_from_dt(_ptr(dt), dt.shape[0], _ptr(out)))
    return out


class SPolygon:
    """SPolygon (simple_polygon.rs:24-37), optionally carrying its SPSurrogate."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    @classmethod
    def new(cls, points):
        """SPolygon::new (simple_polygon.rs:41-71)."""
        pts = _f32(points).reshape(-1, 2)
        h = C.c_void_p()
        check(N.lib().jg_spolygon_new(_ptr(pts), pts.shape[0], C.byref(h)))
        return cls(h.value)

    @classmethod
    def new_batch(cls, polygons, surrogate_config=None, with_surrogate=True):
        """SPolygon::new (+ generate_surrogate) for many polygons, pole searches on the GPU
        (jg_spolygon_new_batch).  Returns (list of SPolygon, number searched on the host)."""
        arrs = [_f32(p).reshape(-1, 2) for p in polygons]
        off = np.zeros(len(arrs) + 1, dtype=np.uint32)
        off[1:] = np.cumsum([len(a) for a in arrs])
        xy = np.ascontiguousarray(np.concatenate(arrs) if arrs else np.zeros((0, 2), np.float32))
        hs = (C.c_void_p * max(1, len(arrs)))()
        n_host = C.c_uint32()
        cfg = (surrogate_config or SPSurrogateConfig())._c() if with_surrogate else None
        check(N.lib().jg_spolygon_new_batch(_ptr(xy), _ptr(off, C.c_uint32), len(arrs),
                                            C.byref(cfg) if cfg is not None else None, hs, C.byref(n_host)))
        return [cls(hs[i]) for i in range(len(arrs))], n_host.value

    def __del__(self, _N=N):
        if getattr(self, "_h", None) and self._h.value and _N._lib is not None:
            _N._lib.jg_spolygon_free(self._h)
            self._h = C.c_void_p()

    def generate_surrogate(self, config=None):
        """SPolygon::generate_surrogate (simple_polygon.rs:73-80)."""
        cfg = (config or SPSurrogateConfig())._c()
        check(N.lib().jg_spolygon_generate_surrogate(self._h, C.byref(cfg)))
        return self

    def transform_clone(self, dt):
        """transform_clone by dt.compose() (placed_item.rs:27-36)."""
        d = dt.as_array() if isinstance(dt, DTransformation) else _f32(dt)
        h = C.c_void_p()
        check(N.lib().jg_spolygon_transform_clone(self._h, _ptr(d), C.byref(h)))
        return SPolygon(h.value)

    def info(self):
        i = N.SPolygonInfoC()
        check(N.lib().jg_spolygon_get_info(self._h, C.byref(i)))
        return i

    @property
    def n_vertices(self):
        return self.info().n_vertices

    @property
    def vertices(self):
        out = np.empty((self.n_vertices, 2), dtype=np.float32)
        check(N.lib().jg_spolygon_get_vertices(self._h, _ptr(out)))
        return out

    @property
    def bbox(self):
        return np.array(self.info().bbox[:], dtype=np.float32)

    @property
    def area(self):
        return np.float32(self.info().area)

    @property
    def diameter(self):
        return np.float32(self.info().diameter)

    @property
    def poi(self):
        return np.array(self.info().poi[:], dtype=np.float32)

    @property
    def poles(self):
        out = np.empty((self.info().n_poles, 3), dtype=np.float32)
        if len(out):
            check(N.lib().jg_spolygon_get_poles(self._h, _ptr(out)))
        return out

    @property
    def piers(self):
        out = np.empty((self.info().n_piers, 4), dtype=np.float32)
        if len(out):
            check(N.lib().jg_spolygon_get_piers(self._h, _ptr(out)))
        return out


class HazardEntity:
    """HazardEntity (hazard.rs:38-51)."""

    def __init__(self, kind, a=0, b=0, dt=(0.0, 0.0, 0.0)):
        self.kind, self.a, self.b = int(kind), int(a), int(b)
        self.dt = tuple(float(np.float32(x)) for x in dt)

    @classmethod
    def placed_item(cls, item_id, dt, pk):
        d = dt.as_array() if isinstance(dt, DTransformation) else _f32(dt)
        return cls(ENTITY_PLACED_ITEM, item_id, pk, tuple(d))

    @classmethod
    def exterior(cls):
        return cls(ENTITY_EXTERIOR)

    @classmethod
    def hole(cls, idx):
        return cls(ENTITY_HOLE, idx)

    @classmethod
    def inferior_quality_zone(cls, quality, idx):
        return cls(ENTITY_INFERIOR_QUALITY_ZONE, quality, idx)

    def scope_interior(self):
        """HazardEntity::scope (hazard.rs:55-62)."""
        return self.kind != ENTITY_EXTERIOR

    def _c(self):
        e = N.HazardEntityC()
        e.kind, e.a, e.b = self.kind, self.a, self.b
        for i in range(3):
            e.dt[i] = self.dt[i]
        return e

    def __eq__(self, o):
        return (self.kind, self.a, self.b, self.dt) == (o.kind, o.a, o.b, o.dt)

    def __hash__(self):
        return hash((self.kind, self.a, self.b, self.dt))


class Hazard:
    """Hazard (hazard.rs:16-23)."""

    def __init__(self, entity, shape, dynamic):
        self.entity, self.shape, self.dynamic = entity, shape, bool(dynamic)


class HazKey:
    """HazKey (hazard.rs:8-11); as a filter it deems only itself irrelevant (filter.rs:51-55)."""

    def __init__(self, idx, version):
        self.idx, self.version = int(idx), int(version)

    def __eq__(self, o):
        return (self.idx, self.version) == (o.idx, o.version)

    def __hash__(self):
        return hash((self.idx, self.version))

    def __repr__(self):
        return f"HazKey({self.idx}v{self.version})"

    def _filter(self):
        return HazKeyFilter([self])._filter()


class NoFilter:
    """NoFilter (filter.rs:57-65)."""

    def _filter(self):
        f = N.FilterC()
        f.kind = 0
        return f


class HazKeyFilter:
    """HazKeyFilter (filter.rs:13-48): the given keys are irrelevant."""

    def __init__(self, keys):
        self.keys = list(keys)
        self._arr = (N.HazKeyC * max(1, len(self.keys)))()
        for i, k in enumerate(self.keys):
            self._arr[i].idx, self._arr[i].version = k.idx, k.version

    def _filter(self):
        f = N.FilterC()
        f.kind = 1
        f.n_keys = len(self.keys)
        f.keys = C.cast(self._arr, C.POINTER(N.HazKeyC))
        return f


class HashedFilter:
    """A per-query pseudo-random HazardFilter: each query deems a `fraction` of the hazard slots
    irrelevant, chosen by the hash specified in include/jagua_amd.h."""

    def __init__(self, seed, fraction):
        self.seed = int(seed) & 0xFFFFFFFF
        self.threshold = min(0xFFFFFFFF, int(round(float(fraction) * 4294967296.0)))

    def _filter(self):
        f = N.FilterC()
        f.kind = 2
        f.seed, f.threshold = self.seed, self.threshold
        return f


class CDESnapshot:
    """CDESnapshot (cd_engine.rs:356-359)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    def __del__(self, _N=N):
        if getattr(self, "_h", None) and self._h.value and _N._lib is not None:
            _N._lib.jg_snapshot_free(self._h)
            self._h = C.c_void_p()


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _stream_ptr(stream):
    """A hipStream_t: a torch stream's handle, a raw handle, or torch's current stream."""
    if stream is not None:
        return C.c_void_p(getattr(stream, "cuda_stream", stream))
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class CDEngine:
    """CDEngine (cd_engine.rs:21-342) with its quadtree flattened into HBM."""

    def __init__(self, bbox, static_hazards, config=None):
        self.config = config or CDEConfig()
        bb = _f32(bbox).reshape(4)
        n = len(static_hazards)
        ents = (N.HazardEntityC * max(1, n))()
        shapes = (C.c_void_p * max(1, n))()
        dyn = (C.c_uint8 * max(1, n))()
        for i, h in enumerate(static_hazards):
            ents[i] = h.entity._c()
            shapes[i] = h.shape._h.value
            dyn[i] = 1 if h.dynamic else 0
        self._h = C.c_void_p()
        cfg = self.config._c()
        check(N.lib().jg_cde_new(_ptr(bb), C.byref(cfg), ents, shapes, dyn, n, C.byref(self._h)))

    def __del__(self, _N=N):
        if getattr(self, "_h", None) and self._h.value and _N._lib is not None:
            _N._lib.jg_cde_free(self._h)
            self._h = C.c_void_p()

    def register_hazard(self, hazard):
        """CDEngine::register_hazard (cd_engine.rs:59-69); returns the new HazKey."""
        k = N.HazKeyC()
        e = hazard.entity._c()
        check(N.lib().jg_cde_register_hazard(self._h, C.byref(e), hazard.shape._h, 1 if hazard.dynamic else 0, C.byref(k)))
        return HazKey(k.idx, k.version)

    def deregister_hazard_by_entity(self, entity):
        e = entity._c()
        check(N.lib().jg_cde_deregister_hazard_by_entity(self._h, C.byref(e)))

    def deregister_hazard_by_key(self, key):
        check(N.lib().jg_cde_deregister_hazard_by_key(self._h, N.HazKeyC(key.idx, key.version)))

    def save(self):
        h = C.c_void_p()
        check(N.lib().jg_cde_save(self._h, C.byref(h)))
        return CDESnapshot(h.value)

    def restore(self, snapshot):
        check(N.lib().jg_cde_restore(self._h, snapshot._h))

    def bbox(self):
        out = np.empty(4, dtype=np.float32)
        check(N.lib().jg_cde_bbox(self._h, _ptr(out)))
        return out

    @property
    def hkey_exterior(self):
        k = N.HazKeyC()
        check(N.lib().jg_cde_hkey_exterior(self._h, C.byref(k)))
        return HazKey(k.idx, k.version)

    def sync(self):
        check(N.lib().jg_cde_sync(self._h))

    def stats(self):
        s = N.CDEStatsC()
        check(N.lib().jg_cde_get_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in N.CDEStatsC._fields_}

    def dump(self):
        """Flattened tree records (node [n,8] u32, hazard [m,8] u32), BFS order."""
        nn, nh = C.c_uint64(), C.c_uint64()
        check(N.lib().jg_cde_dump(self._h, None, 0, None, 0, C.byref(nn), C.byref(nh)))
        nodes = np.zeros((nn.value, 8), dtype=np.uint32)
        haz = np.zeros((max(1, nh.value), 8), dtype=np.uint32)
        check(N.lib().jg_cde_dump(self._h, _ptr(nodes, C.c_uint32), nn.value, _ptr(haz, C.c_uint32), nh.value,
                                  C.byref(nn), C.byref(nh)))
        return nodes, haz[: nh.value]

    # -- queries ------------------------------------------------------------------------
    def detect_batch(self, item, mats, filter=None, flags=0, out=None, stream=None):
        """status[i] = detect_surrogate_collision || detect_poly_collision for transform i
        (ci_bench.rs:107-120).  mats: [n, 6] f32, a torch GPU tensor or a numpy array."""
        f = (filter or NoFilter())._filter()
        if _is_torch(mats):
            import torch
            n = mats.shape[0]
            if out is None:
                out = torch.empty(n, dtype=torch.uint8, device=mats.device)
            check(N.lib().jg_detect_batch(self._h, item._h, C.c_void_p(mats.data_ptr()), n, C.byref(f), flags,
                                          C.c_void_p(out.data_ptr()), _stream_ptr(stream)))
            return out
        m = _f32(mats).reshape(-1, 6)
        st = np.empty(m.shape[0], dtype=np.uint8)
        check(N.lib().jg_detect_batch_host(self._h, item._h, _ptr(m), m.shape[0], C.byref(f), flags,
                                           _ptr(st, C.c_uint8)))
        return st

    def collect_batch(self, item, mats, capacity=16, stream=None, out=None):
        """Per transform i: (set size, set hash, first `capacity` slot keys ascending) of
        collect_surrogate_collisions ∪ collect_poly_collisions (ci_bench.rs:56-69).  With
        device transforms, `out` may give the (count int32 [n], hash int64 [n], keys int32
        [n, capacity]) tensors to write into (contiguous rows)."""
        if _is_torch(mats):
            import torch
            n = mats.shape[0]
            if out is not None:
                cnt, hsh, keys = out
                if cnt.shape != (n,) or hsh.shape != (n,) or keys.shape != (n, max(1, capacity)) or \
                        (cnt.dtype, hsh.dtype, keys.dtype) != (torch.int32, torch.int64, torch.int32) or \
                        any(t.device != mats.device or not t.is_contiguous() for t in (cnt, hsh, keys)):
                    raise ValueError("out must be contiguous int32 [n], int64 [n], int32 [n, capacity] on the "
                                     "transforms' device")
            else:
                cnt = torch.empty(n, dtype=torch.int32, device=mats.device)
                hsh = torch.empty(n, dtype=torch.int64, device=mats.device)
                keys = torch.empty((n, max(1, capacity)), dtype=torch.int32, device=mats.device)
            check(N.lib().jg_collect_batch(self._h, item._h, C.c_void_p(mats.data_ptr()), n, capacity,
                                           C.c_void_p(cnt.data_ptr()), C.c_void_p(hsh.data_ptr()),
                                           C.c_void_p(keys.data_ptr()) if capacity else None, _stream_ptr(stream)))
            return cnt, hsh, keys
        m = _f32(mats).reshape(-1, 6)
        n = m.shape[0]
        cnt = np.empty(n, dtype=np.uint32)
        hsh = np.empty(n, dtype=np.uint64)
        keys = np.empty((n, max(1, capacity)), dtype=np.uint32)
        check(N.lib().jg_collect_batch_host(self._h, item._h, _ptr(m), n, capacity, _ptr(cnt, C.c_uint32),
                                            _ptr(hsh, C.c_uint64), _ptr(keys, C.c_uint32) if capacity else None))
        return cnt, hsh, keys


def _query_sets(sets):
    """jg_query_set array of [(engine, item, first, count, filter)] and the objects it refers to."""
    arr = (N.QuerySetC * len(sets))()
    keep = []
    for i, (cde, item, first, count, filt) in enumerate(sets):
        arr[i].cde = cde._h.value
        arr[i].item = item._h.value
        arr[i].first = first
        arr[i].count = count
        f = (filt or NoFilter())
        keep.append(f)
        arr[i].filter = f._filter()
    return arr, keep


# JG_NOALLOC_SCRATCH_BYTES: the fixed per-GPU scratch of detect_sets_noalloc
NOALLOC_SCRATCH_BYTES = 32 << 20


def detect_sets_noalloc(sets, mats, out, chunk=0, workgroups=0, lanes_per_query=0, flags=0, stages=0,
                        stream=None):
    """The batched detection of QueryPlan(sets).detect(mats, out) in one call that allocates
    no device memory (jg_detect_sets_noalloc): its tables live in a fixed per-GPU scratch
    array.  Every engine must be synced (CDEngine.sync()) after its last update."""
    arr, keep = _query_sets(sets)
    cfg = N.PlanConfigC(chunk, workgroups, lanes_per_query, flags, stages)
    check(N.lib().jg_detect_sets_noalloc(arr, len(sets), C.byref(cfg), C.c_void_p(mats.data_ptr()),
                                         C.c_void_p(out.data_ptr()), _stream_ptr(stream)))
    return out


class Graph:
    """The launches of the calls made inside `with Graph.capture(stream) as g:` on that stream
    (a torch.cuda.Stream), captured once into a HIP graph (jg_graph_*) and replayed by
    g.launch().  Call every captured operation once before capturing."""

    def __init__(self):
        self._h = C.c_void_p()
        self._stream = None

    @classmethod
    def capture(cls, stream):
        import contextlib

        @contextlib.contextmanager
        def cm():
            g = cls()
            g._stream = stream
            check(N.lib().jg_graph_begin(_stream_ptr(stream)))
            try:
                yield g
            except BaseException:
                h = C.c_void_p()
                N.lib().jg_graph_end(_stream_ptr(stream), C.byref(h))
                if h.value:
                    N.lib().jg_graph_free(h)
                raise
            check(N.lib().jg_graph_end(_stream_ptr(stream), C.byref(g._h)))
        return cm()

    def launch(self, stream=None):
        check(N.lib().jg_graph_launch(self._h, _stream_ptr(stream if stream is not None else self._stream)))

    def __del__(self, _N=N):
        if getattr(self, "_h", None) and self._h.value and _N._lib is not None:
            _N._lib.jg_graph_free(self._h)
            self._h = C.c_void_p()


class QueryPlan:
    """Batched persistent form: many (engine, item, transform range) sets in one launch.

    Schedule (jg_plan_config): chunk = queries per work unit (16, 32, 64; 0 = 64); workgroups
    = persistent one-wave workgroups (0 = 16 per compute unit); lanes_per_query = lanes per
    query in the polygon stage (0 = per item, 8..32); stages = 0 or 2 for two kernels (the
    surrogate stage, then the polygon stage over one-pass survivor groups), 1 for one kernel
    running both stages of a unit."""

    def __init__(self, sets, chunk=0, workgroups=0, lanes_per_query=0, flags=0, stages=0):
        n = len(sets)
        arr, self._keep = _query_sets(sets)
        cfg = N.PlanConfigC(chunk, workgroups, lanes_per_query, flags, stages)
        self._h = C.c_void_p()
        self._refs = [s[0] for s in sets] + [s[1] for s in sets]
        check(N.lib().jg_plan_create(arr, n, C.byref(cfg), C.byref(self._h)))

    def __del__(self, _N=N):
        if getattr(self, "_h", None) and self._h.value and _N._lib is not None:
            _N._lib.jg_plan_free(self._h)
            self._h = C.c_void_p()

    def detect(self, mats, out, stream=None):
        check(N.lib().jg_plan_detect(self._h, C.c_void_p(mats.data_ptr()), C.c_void_p(out.data_ptr()),
                                     _stream_ptr(stream)))
        return out

    def detect_graph(self, mats, out, replays=1, stream=None):
        check(N.lib().jg_plan_detect_graph(self._h, C.c_void_p(mats.data_ptr()), C.c_void_p(out.data_ptr()),
                                           replays, _stream_ptr(stream)))
        return out

    def collect(self, mats, capacity=16, count=None, hash=None, keys=None, stream=None):
        """Hazard sets of every query (jg_plan_collect): (count u32 [n], hash as int64 [n],
        keys int32 [n, capacity] ascending, padded with -1) as torch tensors on the device.
        A set's HazKeyFilter is a pre-seeded collector: its keys are in every set."""
        import torch
        n = mats.shape[0]
        dev = mats.device
        count = count if count is not None else torch.empty(n, dtype=torch.int32, device=dev)
        hash = hash if hash is not None else torch.empty(n, dtype=torch.int64, device=dev)
        if capacity and keys is None:
            keys = torch.empty((n, capacity), dtype=torch.int32, device=dev)
        check(N.lib().jg_plan_collect(self._h, C.c_void_p(mats.data_ptr()), capacity, C.c_void_p(count.data_ptr()),
                                      C.c_void_p(hash.data_ptr()),
                                      C.c_void_p(keys.data_ptr()) if capacity else None, _stream_ptr(stream)))
        return count, hash, keys

    def autotune(self, mats, out, reps=3, stream=None):
        """Times candidate schedules on these inputs and keeps the fastest (jg_plan_autotune).
        Returns (schedule dict, its median ms); `out` holds the verdicts afterwards."""
        best = N.PlanConfigC()
        ms = C.c_float()
        check(N.lib().jg_plan_autotune(self._h, C.c_void_p(mats.data_ptr()), C.c_void_p(out.data_ptr()), reps,
                                       _stream_ptr(stream), C.byref(best), C.byref(ms)))
        return self._cfg_dict(best), ms.value

    def config(self):
        c = N.PlanConfigC()
        check(N.lib().jg_plan_get_config(self._h, C.byref(c)))
        return self._cfg_dict(c)

    @staticmethod
    def _cfg_dict(c):
        return {"chunk": c.chunk, "workgroups": c.workgroups, "lanes_per_query": c.lanes_per_query,
                "flags": c.flags, "stages": c.stages}


class PlacementSearch:
    """lbf's sampled placement search (lbf/src/opt/search.rs:17-106) for many independent
    (engine, item) pairs at once, resident on the GPU (jg_search_*, include/jagua_amd.h).

    searches: list of (cde, item, rotations, filter); rotations is None (no rotation),
    "continuous", or a list of allowed rotations in radians; filter may be None.  Defaults
    draw lbf's 5000 samples per item (LBFConfig::default, lbf/src/config.rs:33-41) as 8 uniform
    and 2 local rounds of 500."""

    def __init__(self, searches, samples_per_round=500, uniform_rounds=8, local_rounds=2, seed=0,
                 ls_sd_transl=(0.01, 0.0005), ls_sd_rot=(2.0 * np.pi / 180.0, 0.5 * np.pi / 180.0)):
        n = len(searches)
        arr = (N.SearchItemC * n)()
        self._keep = []
        self.n = n
        for i, (cde, item, rots, filt) in enumerate(searches):
            arr[i].cde = cde._h.value
            arr[i].item = item._h.value
            if isinstance(rots, str):
                if rots != "continuous":
                    raise ValueError("rotations: None, 'continuous' or a list of radians")
                arr[i].continuous = 1
            elif rots is not None and len(rots):
                r = _f32(rots).reshape(-1)
                self._keep.append(r)
                arr[i].rotations = _ptr(r)
                arr[i].n_rotations = len(r)
            f = filt or NoFilter()
            self._keep.append(f)
            arr[i].filter = f._filter()
        cfg = N.SearchConfigC()
        cfg.samples_per_round, cfg.uniform_rounds, cfg.local_rounds = samples_per_round, uniform_rounds, local_rounds
        cfg.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        cfg.ls_sd_transl[0], cfg.ls_sd_transl[1] = ls_sd_transl
        cfg.ls_sd_rot[0], cfg.ls_sd_rot[1] = ls_sd_rot
        self.samples_per_round = samples_per_round
        self.rounds = uniform_rounds + local_rounds
        self.uniform_rounds = uniform_rounds
        self._refs = [s[0] for s in searches] + [s[1] for s in searches]
        self._h = C.c_void_p()
        check(N.lib().jg_search_create(arr, n, C.byref(cfg), C.byref(self._h)))

    def __del__(self, _N=N):
        if getattr(self, "_h", None) and self._h.value and _N._lib is not None:
            _N._lib.jg_search_free(self._h)
            self._h = C.c_void_p()

    def run(self, graph=False, stream=None):
        check(N.lib().jg_search_run(self._h, 1 if graph else 0, _stream_ptr(stream)))
        return self

    def results(self):
        out = (N.SearchResultC * self.n)()
        check(N.lib().jg_search_results(self._h, out))
        return [{"found": bool(r.found), "rotation": r.rotation, "tx": r.tx, "ty": r.ty, "cost": r.cost,
                 "x_max": r.x_max, "y_max": r.y_max, "samples": r.samples} for r in out]

    def samples(self):
        """(dt [rounds, n, K, 3], status [rounds, n, K]) of the last run."""
        shape = (self.rounds, self.n, self.samples_per_round)
        dt = np.empty(shape + (3,), dtype=np.float32)
        st = np.empty(shape, dtype=np.uint8)
        check(N.lib().jg_search_samples(self._h, _ptr(dt), _ptr(st, C.c_uint8)))
        return dt, st


class PoleOverlap:
    """The pole-pair overlap quantity Q (include/jagua_amd.h, jg_overlap_*) as a forward /
    backward op pair over per-item poses (x, y, theta), on the GPU.

    poles: [total, 3] local (px, py, r); pole_off: [n_items + 1]; pairs: [n_pairs, 2] item
    indices; cedges: [n_edges, 4] container edges (x1, y1, x2, y2)."""

    def __init__(self, poles, pole_off, pairs, cedges):
        poles = _f32(poles).reshape(-1, 3)
        off = np.ascontiguousarray(pole_off, dtype=np.uint32)
        pr = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        ce = _f32(cedges).reshape(-1, 4)
        self.n_items = len(off) - 1
        self._h = C.c_void_p()
        check(N.lib().jg_overlap_create(_ptr(poles), _ptr(off, C.c_uint32), self.n_items, _ptr(pr, C.c_uint32), len(pr),
                                        _ptr(ce), len(ce), C.byref(self._h)))

    def __del__(self, _N=N):
        if getattr(self, "_h", None) and self._h.value and _N._lib is not None:
            _N._lib.jg_overlap_free(self._h)
            self._h = C.c_void_p()

    def set_pairs(self, pairs):
        pr = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        check(N.lib().jg_overlap_set_pairs(self._h, _ptr(pr, C.c_uint32), len(pr)))

    def pairs_from_cde(self, cde, item_keys, pose, radius, stream=None):
        """Candidate pairs from the engine's quadtree on the GPU (jg_overlap_pairs_from_cde):
        item k is the hazard item_keys[k] (HazKey) of cde, its bounding circle centred at
        pose[k] (x, y) (torch GPU f32 [n, 3]) with radius radius[k].  Returns the pair count."""
        slots = np.ascontiguousarray([k.idx for k in item_keys], dtype=np.uint32)
        rad = _f32(radius).reshape(-1)
        n = C.c_uint64()
        check(N.lib().jg_overlap_pairs_from_cde(self._h, cde._h, _ptr(slots, C.c_uint32), C.c_void_p(pose.data_ptr()),
                                                _ptr(rad), C.byref(n), _stream_ptr(stream)))
        return n.value

    def pairs(self):
        """The current candidate pairs [n, 2] (host copy)."""
        n = C.c_uint64()
        check(N.lib().jg_overlap_get_pairs(self._h, None, 0, C.byref(n)))
        out = np.empty((max(1, n.value), 2), dtype=np.uint32)
        check(N.lib().jg_overlap_get_pairs(self._h, _ptr(out, C.c_uint32), n.value, C.byref(n)))
        return out[: n.value]

    def forward(self, pose, out=None, stream=None):
        """Q for pose (torch GPU f32 [n_items, 3]); returns a GPU f32 tensor of shape [1]."""
        import torch
        if out is None:
            out = torch.empty(1, dtype=torch.float32, device=pose.device)
        check(N.lib().jg_overlap_fwd(self._h, C.c_void_p(pose.data_ptr()), C.c_void_p(out.data_ptr()),
                                     _stream_ptr(stream)))
        return out

    def backward(self, pose, out=None, stream=None):
        """dQ/d(x, y, theta) per item (GPU f32 [n_items, 3])."""
        import torch
        if out is None:
            out = torch.empty((self.n_items, 3), dtype=torch.float32, device=pose.device)
        check(N.lib().jg_overlap_bwd(self._h, C.c_void_p(pose.data_ptr()), C.c_void_p(out.data_ptr()),
                                     _stream_ptr(stream)))
        return out

    def forward_backward(self, pose, q=None, grad=None, stream=None):
        """(Q, dQ/d(x, y, theta)) in one pass, equal to forward() and backward()."""
        import torch
        if q is None:
            q = torch.empty(1, dtype=torch.float32, device=pose.device)
        if grad is None:
            grad = torch.empty((self.n_items, 3), dtype=torch.float32, device=pose.device)
        check(N.lib().jg_overlap_fwd_bwd(self._h, C.c_void_p(pose.data_ptr()), C.c_void_p(q.data_ptr()),
                                         C.c_void_p(grad.data_ptr()), _stream_ptr(stream)))
        return q, grad


def candidate_pairs(pose, radius, bbox, depth=8):
    """Items whose bounding circles overlap (uniform grid of 2^depth cells per side)."""
    pose = _f32(pose).reshape(-1, 3)
    rad = _f32(radius).reshape(-1)
    bb = _f32(bbox).reshape(4)
    n = C.c_uint64()
    check(N.lib().jg_overlap_candidate_pairs(_ptr(pose), _ptr(rad), len(pose), _ptr(bb), depth, None, 0, C.byref(n)))
    out = np.empty((max(1, n.value), 2), dtype=np.uint32)
    check(N.lib().jg_overlap_candidate_pairs(_ptr(pose), _ptr(rad), len(pose), _ptr(bb), depth,
                                             _ptr(out, C.c_uint32), n.value, C.byref(n)))
    return out[: n.value]


def version():
    return N.lib().jg_version().decode()
