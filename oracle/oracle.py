"""TEST INFRASTRUCTURE ONLY — ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker
and the CPU baseline; never by the product (jagua-rs_amd/).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

f32p = C.POINTER(C.c_float)
u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class HazardDesc(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("n_verts", C.c_uint32), ("a", C.c_uint64), ("b", C.c_uint64),
                ("dt", C.c_float * 3), ("dynamic", C.c_uint32), ("has_dt", C.c_uint32), ("has_pre", C.c_uint32),
                ("pre", C.c_float * 3), ("_pad", C.c_uint32)]


class Filter(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("seed", C.c_uint32), ("threshold", C.c_uint32), ("n_keys", C.c_uint32),
                ("keys", u32p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        L = C.CDLL(LIB)
        L.or_spolygon_new.argtypes = [f32p, C.c_uint32, f32p, f32p]
        L.or_surrogate.argtypes = [f32p, C.c_uint32, u32p, f32p, C.c_uint32, C.c_uint32, f32p, C.c_uint32, u32p,
                                   f32p, C.c_uint32, u32p]
        L.or_pip.argtypes = [f32p, C.c_uint32, C.c_float, C.c_float]
        for n in ("or_edge_edge", "or_rect_edge", "or_circle_edge", "or_circle_rect", "or_edge_quadrants"):
            getattr(L, n).argtypes = [f32p, f32p]
        L.or_fpa_eq.argtypes = [C.c_float, C.c_float]
        L.or_transform_from_dt.argtypes = [f32p, C.c_uint64, f32p]
        L.or_cde_create.restype = C.c_void_p
        L.or_cde_create.argtypes = [f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(HazardDesc), f32p]
        L.or_cde_destroy.argtypes = [C.c_void_p]
        L.or_cde_register.argtypes = [C.c_void_p, C.POINTER(HazardDesc), f32p, u32p]
        L.or_cde_deregister_entity.argtypes = [C.c_void_p, C.POINTER(HazardDesc)]
        L.or_cde_deregister_key.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
        L.or_cde_save.restype = C.c_void_p
        L.or_cde_save.argtypes = [C.c_void_p]
        L.or_cde_restore.argtypes = [C.c_void_p, C.c_void_p]
        L.or_snapshot_free.argtypes = [C.c_void_p]
        L.or_cde_invariants.argtypes = [C.c_void_p]
        L.or_cde_trees_match.argtypes = [C.c_void_p, C.c_void_p]
        L.or_cde_dump.restype = C.c_int64
        L.or_cde_dump.argtypes = [C.c_void_p, u32p, C.c_uint64, u32p, C.c_uint64, u64p]
        L.or_item_create.restype = C.c_void_p
        L.or_item_create.argtypes = [f32p, C.c_uint32, u32p, f32p, C.c_uint32, C.c_uint32]
        L.or_item_create_pre.restype = C.c_void_p
        L.or_item_create_pre.argtypes = [f32p, C.c_uint32, f32p, u32p, f32p, C.c_uint32, C.c_uint32]
        L.or_import_polygon.argtypes = [f32p, C.c_uint32, f32p, f32p]
        L.or_item_info.argtypes = [C.c_void_p, f32p, C.c_uint32, f32p, f32p, C.c_uint32, u32p]
        L.or_ext_to_int.argtypes = [f32p, f32p, f32p]
        L.or_int_to_ext.argtypes = [f32p, f32p, f32p]
        L.or_item_from_parts.restype = C.c_void_p
        L.or_item_from_parts.argtypes = [f32p, C.c_uint32, f32p, f32p, C.c_uint32, f32p, C.c_uint32, C.c_uint32]
        L.or_item_destroy.argtypes = [C.c_void_p]
        L.or_overlap64.argtypes = [C.POINTER(C.c_double), f32p, u32p, C.c_uint32, u32p, C.c_uint64, f32p,
                                   C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_query_batch.argtypes = [C.c_void_p, C.c_void_p, f32p, f32p, C.c_uint64, C.c_uint32, C.POINTER(Filter),
                                     u8p, u8p, u32p, u64p, u32p, C.c_uint32, C.c_uint32, u64p,
                                     C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def fp(a):
    return a.ctypes.data_as(f32p)


DEFAULT_LIMITS = ((100, 0.0), (20, 0.75), (10, 0.90))


def spolygon_new(xy):
    """SPolygon::new -> (vertices, meta[area, diameter, bbox4, poi3]) or None when it bails."""
    xy = _f(xy).reshape(-1, 2)
    out = np.empty_like(xy)
    meta = np.empty(9, dtype=np.float32)
    r = lib().or_spolygon_new(fp(xy), len(xy), fp(out), fp(meta))
    return None if r != 0 else (out, meta)


def surrogate(xy, limits=DEFAULT_LIMITS, n_ff_poles=2, n_ff_piers=0):
    xy = _f(xy).reshape(-1, 2)
    ln = (C.c_uint32 * 3)(*[n for n, _ in limits])
    lc = (C.c_float * 3)(*[c for _, c in limits])
    poles = np.empty((1000, 3), dtype=np.float32)
    piers = np.empty((max(1, n_ff_piers), 4), dtype=np.float32)
    npo, npi = C.c_uint32(), C.c_uint32()
    r = lib().or_surrogate(fp(xy), len(xy), ln, lc, n_ff_poles, n_ff_piers, fp(poles), 1000, C.byref(npo), fp(piers),
                           len(piers), C.byref(npi))
    if r != 0:
        raise RuntimeError("oracle surrogate generation failed")
    return poles[: npo.value].copy(), piers[: npi.value].copy()


def compose(dt):
    dt = _f(dt).reshape(-1, 3)
    out = np.empty((len(dt), 6), dtype=np.float32)
    lib().or_transform_from_dt(fp(dt), len(dt), fp(out))
    return out


KIND_PLACED, KIND_EXTERIOR, KIND_HOLE, KIND_IQZ = 0, 1, 2, 3


class OracleCDE:
    """The oracle's CDEngine built from a Scene (same registration order as build_cde)."""

    def __init__(self, scene):
        descs = [self._desc(KIND_EXTERIOR, scene.container)]
        verts = [_f(scene.container).ravel()]
        for i, h in enumerate(scene.holes):
            descs.append(self._desc(KIND_HOLE, h, a=i))
            verts.append(_f(h).ravel())
        arr = (HazardDesc * len(descs))(*descs)
        v = np.concatenate(verts)
        bb = _f(scene.cde_bbox())
        self.h = lib().or_cde_create(fp(bb), scene.depth, scene.cd_threshold, len(descs), arr, fp(v))
        if not self.h:
            raise RuntimeError("oracle CDE construction failed")
        self.keys = []
        for i, (loc, dt) in enumerate(scene.items):
            self.keys.append(self.register_item(i, loc, dt, pk=i))

    @classmethod
    def from_hazards(cls, bbox, depth, cd_threshold, hazards):
        """An engine from static hazards [(kind, a, b, vertices)], in order."""
        self = cls.__new__(cls)
        descs = [cls._desc(k, xy, a=a, b=b) for k, a, b, xy in hazards]
        arr = (HazardDesc * len(descs))(*descs)
        v = np.concatenate([_f(xy).ravel() for _, _, _, xy in hazards])
        self.h = lib().or_cde_create(fp(_f(bbox)), depth, cd_threshold, len(descs), arr, fp(v))
        if not self.h:
            raise RuntimeError("oracle CDE construction failed")
        self.keys = []
        return self

    @staticmethod
    def _desc(kind, xy, a=0, b=0, dt=None, dynamic=False, pre=None):
        d = HazardDesc()
        d.kind, d.n_verts, d.a, d.b = kind, len(xy), a, b
        if dt is not None:
            for i in range(3):
                d.dt[i] = float(dt[i])
            d.has_dt = 1
        if pre is not None:
            for i in range(3):
                d.pre[i] = float(pre[i])
            d.has_pre = 1
        d.dynamic = 1 if dynamic else 0
        return d

    def register_item(self, item_id, local, dt, pk, pre=None):
        """A placed item: SPolygon::new(local), transformed by pre (an imported item's
        centering) and then by dt."""
        d = self._desc(KIND_PLACED, local, a=item_id, b=pk, dt=dt, dynamic=True, pre=pre)
        key = (C.c_uint32 * 2)()
        loc = _f(local)
        if lib().or_cde_register(self.h, C.byref(d), fp(loc), key) != 0:
            raise RuntimeError("oracle register failed")
        return (key[0], key[1])

    def deregister_item(self, item_id, dt, pk, n_verts=3):
        d = self._desc(KIND_PLACED, np.zeros((n_verts, 2)), a=item_id, b=pk, dt=dt, dynamic=True)
        if lib().or_cde_deregister_entity(self.h, C.byref(d)) != 0:
            raise RuntimeError("oracle deregister failed")

    def deregister_key(self, key):
        if lib().or_cde_deregister_key(self.h, key[0], key[1]) != 0:
            raise RuntimeError("oracle deregister by key failed")

    def save(self):
        return lib().or_cde_save(self.h)

    def restore(self, snap):
        if lib().or_cde_restore(self.h, snap) != 0:
            raise RuntimeError("oracle restore failed")

    def invariants(self):
        return lib().or_cde_invariants(self.h)

    def dump(self):
        nh = C.c_uint64()
        nn = lib().or_cde_dump(self.h, None, 0, None, 0, C.byref(nh))
        nodes = np.zeros((nn, 8), dtype=np.uint32)
        haz = np.zeros((max(1, nh.value), 8), dtype=np.uint32)
        lib().or_cde_dump(self.h, nodes.ctypes.data_as(u32p), nn, haz.ctypes.data_as(u32p), nh.value, C.byref(nh))
        return nodes, haz[: nh.value]

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_cde_destroy(self.h)
            self.h = None


class OracleItem:
    def __init__(self, xy, limits=DEFAULT_LIMITS, n_ff_poles=2, n_ff_piers=0, pre=None):
        """SPolygon::new(xy) (transformed by pre first when given: an imported item) with
        its surrogate."""
        xy = _f(xy).reshape(-1, 2)
        ln = (C.c_uint32 * 3)(*[n for n, _ in limits])
        lc = (C.c_float * 3)(*[c for _, c in limits])
        if pre is None:
            self.h = lib().or_item_create(fp(xy), len(xy), ln, lc, n_ff_poles, n_ff_piers)
        else:
            self.h = lib().or_item_create_pre(fp(xy), len(xy), fp(_f(pre)), ln, lc, n_ff_poles, n_ff_piers)
        self._n = len(xy)
        if not self.h:
            raise RuntimeError("oracle item construction failed")

    def info(self):
        """(vertices, meta [area, diameter, bbox (4), poi (3)], poles [n][3]) of the item's shape."""
        xy = np.zeros((self._n, 2), dtype=np.float32)
        meta = np.zeros(9, dtype=np.float32)
        poles = np.zeros((1000, 3), dtype=np.float32)
        npl = C.c_uint32()
        lib().or_item_info(self.h, fp(xy), self._n, fp(meta), fp(poles), 1000, C.byref(npl))
        return xy, meta, poles[: npl.value]

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_item_destroy(self.h)
            self.h = None


def make_filter(kind=0, seed=0, threshold=0, keys=()):
    f = Filter()
    f.kind, f.seed, f.threshold = kind, seed, threshold
    arr = (C.c_uint32 * max(1, len(keys)))(*keys)
    f.n_keys = len(keys)
    f.keys = C.cast(arr, u32p)
    f._keep = arr
    return f


def query(cde, item, mats=None, dts=None, mode=3, filt=None, keys_cap=0, threads=1, counters=False):
    """Runs n collision queries on the oracle.  mode: 1 detect, 2 collect, 3 both; | 4 seeds
    the collector with a key filter's slots (a pre-seeded collector).  Returns a dict with
    status, surr, count, hash, keys (when requested), counters (nodes, entries, edges, pip
    vertices), seconds."""
    src = _f(mats).reshape(-1, 6) if mats is not None else _f(dts).reshape(-1, 3)
    n = len(src)
    status = np.zeros(n, dtype=np.uint8)
    surr = np.zeros(n, dtype=np.uint8)
    cnt = np.zeros(n, dtype=np.uint32)
    hsh = np.zeros(n, dtype=np.uint64)
    keys = np.zeros((n, keys_cap), dtype=np.uint32) if keys_cap else None
    ct = np.zeros(4, dtype=np.uint64)
    sec = C.c_double()
    r = lib().or_query_batch(cde.h, item.h, None if mats is not None else fp(src), fp(src) if mats is not None else None,
                             n, mode, C.byref(filt) if filt is not None else None, status.ctypes.data_as(u8p),
                             surr.ctypes.data_as(u8p), cnt.ctypes.data_as(u32p), hsh.ctypes.data_as(u64p),
                             keys.ctypes.data_as(u32p) if keys is not None else None, keys_cap, threads,
                             ct.ctypes.data_as(u64p) if counters else None, C.byref(sec))
    if r != 0:
        raise RuntimeError("oracle query failed (the reference would have panicked)")
    return dict(status=status, surr=surr, count=cnt, hash=hsh, keys=keys, counters=ct, seconds=sec.value)


def overlap64(pose, poles, off, pairs, cedges, grad=True):
    """FP64 Q and dQ/d(x, y, theta) of the pole-pair overlap quantity."""
    pose = np.ascontiguousarray(pose, dtype=np.float64).reshape(-1, 3)
    poles = _f(poles).reshape(-1, 3)
    off = np.ascontiguousarray(off, dtype=np.uint32)
    pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    cedges = _f(cedges).reshape(-1, 4)
    q = C.c_double()
    g = np.zeros((len(pose), 3), dtype=np.float64) if grad else None
    lib().or_overlap64(pose.ctypes.data_as(C.POINTER(C.c_double)), fp(poles), off.ctypes.data_as(u32p), len(pose),
                       pairs.ctypes.data_as(u32p), len(pairs), fp(cedges), len(cedges), C.byref(q),
                       g.ctypes.data_as(C.POINTER(C.c_double)) if grad else None)
    return q.value, g


def import_polygon(xy):
    """import_simple_polygon (io/import.rs:198-212) + SPolygon::new: (vertices, centroid), or
    None when the reference bails."""
    xy = _f(xy).reshape(-1, 2)
    out = np.zeros_like(xy)
    c = np.zeros(2, dtype=np.float32)
    n = lib().or_import_polygon(fp(xy), len(xy), fp(out), fp(c))
    return None if n < 0 else (out[:n], c)


def ext_to_int(ext, pre):
    out = np.zeros(3, dtype=np.float32)
    if lib().or_ext_to_int(fp(_f(ext)), fp(_f(pre)), fp(out)) != 0:
        raise RuntimeError("NaN transformation")
    return out


def int_to_ext(int_dt, pre):
    out = np.zeros(3, dtype=np.float32)
    if lib().or_int_to_ext(fp(_f(int_dt)), fp(_f(pre)), fp(out)) != 0:
        raise RuntimeError("NaN transformation")
    return out
