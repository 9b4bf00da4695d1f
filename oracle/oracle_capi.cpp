// TEST INFRASTRUCTURE ONLY — C entry points of the CPU oracle, loaded with ctypes by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg (oracle/liboracle.so).
//
// A collision query is composed exactly as lbf/benches/ci_bench.rs composes it:
//   verdict = detect_surrogate_collision || detect_poly_collision       (ci_bench.rs:107-120)
//   set     = collect_surrogate_collisions ∪ collect_poly_collisions    (ci_bench.rs:56-69)
// The per-query hashed filter and the set hash are restated from include/jagua_amd.h, where
// the product declares them (jg_filter_hash, jg_set_hash_mix); the oracle keeps its own copy.
#include <algorithm>
#include <chrono>
#include <thread>

#include "jg_oracle.hpp"

using namespace oracle;

namespace {

// murmur3 fmix32 / splitmix64 finalisers — the two mixing functions the ABI specifies.
inline uint32_t fmix32(uint32_t h) {
    h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
    return h;
}
inline uint64_t mix64(uint64_t z) {
    z ^= z >> 30; z *= 0xbf58476d1ce4e5b9ull; z ^= z >> 27; z *= 0x94d049bb133111ebull; z ^= z >> 31;
    return z;
}
inline bool hashed_irrelevant(uint32_t seed, uint32_t threshold, uint64_t q, uint32_t slot) {
    uint32_t h = fmix32(seed ^ fmix32((uint32_t)q ^ fmix32((uint32_t)(q >> 32) + 0x9e3779b9u)) ^ (slot * 0x85ebca77u));
    return h < threshold;
}

struct OrItem {
    SPolygon shape;
};

struct OrFilter {           // mirrors jg_hazard_filter
    uint32_t kind;          // 0 none, 1 key mask, 2 hashed per query
    uint32_t seed;
    uint32_t threshold;
    uint32_t n_keys;
    const uint32_t* keys;   // slot indices (kind 1)
};

SPSurrogateConfig make_cfg(const uint32_t* lim_n, const float* lim_cov, uint32_t n_ff_poles, uint32_t n_ff_piers) {
    SPSurrogateConfig c{};
    for (int i = 0; i < 3; i++) { c.n_pole_limits_n[i] = lim_n[i]; c.n_pole_limits_cov[i] = lim_cov[i]; }
    c.n_ff_poles = n_ff_poles;
    c.n_ff_piers = n_ff_piers;
    return c;
}

std::vector<Point> to_points(const float* xy, uint32_t n) {
    std::vector<Point> p(n);
    for (uint32_t i = 0; i < n; i++) p[i] = {xy[2 * i], xy[2 * i + 1]};
    return p;
}

}  // namespace

extern "C" {

struct or_hazard_desc {
    uint32_t kind;
    uint32_t n_verts;
    uint64_t a;
    uint64_t b;
    float dt[3];
    uint32_t dynamic;
    uint32_t has_dt;     // placed items: shape = SPolygon::new(local) transformed by dt
    uint32_t has_pre;    // imported items: transformed by pre (the centering) before dt
    float pre[3];
    uint32_t _pad;
};

// SPolygon::new — writes the (CCW-normalised) vertices and
// meta = [area, diameter, bbox x_min, y_min, x_max, y_max, poi cx, cy, r].
int or_spolygon_new(const float* xy, uint32_t n, float* out_xy, float* meta) {
    try {
        SPolygon s = spolygon_new(to_points(xy, n));
        for (uint32_t i = 0; i < n; i++) { out_xy[2 * i] = s.vertices[i].x; out_xy[2 * i + 1] = s.vertices[i].y; }
        float m[9] = {s.area, s.diameter, s.bbox.x_min, s.bbox.y_min, s.bbox.x_max, s.bbox.y_max,
                      s.poi.center.x, s.poi.center.y, s.poi.radius};
        std::memcpy(meta, m, sizeof m);
        return 0;
    } catch (...) {
        return -1;
    }
}

// SPSurrogate::new — poles [n][3] (cx, cy, r), piers [n][4].
int or_surrogate(const float* xy, uint32_t n, const uint32_t* lim_n, const float* lim_cov, uint32_t n_ff_poles,
                 uint32_t n_ff_piers, float* poles, uint32_t poles_cap, uint32_t* n_poles, float* piers,
                 uint32_t piers_cap, uint32_t* n_piers) {
    try {
        SPolygon s = spolygon_new(to_points(xy, n));
        spolygon_generate_surrogate(s, make_cfg(lim_n, lim_cov, n_ff_poles, n_ff_piers));
        *n_poles = (uint32_t)s.surrogate.poles.size();
        *n_piers = (uint32_t)s.surrogate.piers.size();
        for (uint32_t i = 0; i < *n_poles && i < poles_cap; i++) {
            poles[3 * i] = s.surrogate.poles[i].center.x;
            poles[3 * i + 1] = s.surrogate.poles[i].center.y;
            poles[3 * i + 2] = s.surrogate.poles[i].radius;
        }
        for (uint32_t i = 0; i < *n_piers && i < piers_cap; i++) {
            const Edge& e = s.surrogate.piers[i];
            piers[4 * i] = e.start.x; piers[4 * i + 1] = e.start.y; piers[4 * i + 2] = e.end.x; piers[4 * i + 3] = e.end.y;
        }
        return 0;
    } catch (...) {
        return -1;
    }
}

// Point-in-polygon and primitive predicates, exposed for the predicate tests.
int or_pip(const float* xy, uint32_t n, float px, float py) {
    try {
        SPolygon s = spolygon_new(to_points(xy, n));
        return spolygon_collides_point(s, {px, py}) ? 1 : 0;
    } catch (...) {
        return -1;
    }
}
int or_edge_edge(const float* e1, const float* e2) {
    return edge_intersection(Edge{{e1[0], e1[1]}, {e1[2], e1[3]}}, Edge{{e2[0], e2[1]}, {e2[2], e2[3]}}) ? 1 : 0;
}
int or_rect_edge(const float* r, const float* e) {
    return rect_collides_edge(Rect{r[0], r[1], r[2], r[3]}, Edge{{e[0], e[1]}, {e[2], e[3]}}) ? 1 : 0;
}
int or_circle_edge(const float* c, const float* e) {
    return circle_collides_edge(Circle{{c[0], c[1]}, c[2]}, Edge{{e[0], e[1]}, {e[2], e[3]}}) ? 1 : 0;
}
int or_circle_rect(const float* c, const float* r) {
    return circle_collides_rect(Circle{{c[0], c[1]}, c[2]}, Rect{r[0], r[1], r[2], r[3]}) ? 1 : 0;
}
int or_fpa_eq(float a, float b) { return fpa_eq(a, b) ? 1 : 0; }
int or_edge_quadrants(const float* e, const float* r) {
    Rect rr{r[0], r[1], r[2], r[3]};
    auto qs = rect_quadrants(rr);
    auto cq = edge_collides_with_quadrants(Edge{{e[0], e[1]}, {e[2], e[3]}}, rr, qs);
    return (cq[0] ? 1 : 0) | (cq[1] ? 2 : 0) | (cq[2] ? 4 : 0) | (cq[3] ? 8 : 0);
}
void or_transform_from_dt(const float* dt, uint64_t n, float* out6) {
    for (uint64_t i = 0; i < n; i++) {
        Transformation t = transformation_from_dt(dt[3 * i], dt[3 * i + 1], dt[3 * i + 2]);
        float* o = out6 + 6 * i;
        o[0] = t.m[0][0]; o[1] = t.m[0][1]; o[2] = t.m[0][2]; o[3] = t.m[1][0]; o[4] = t.m[1][1]; o[5] = t.m[1][2];
    }
}

static Hazard make_hazard(const or_hazard_desc& d, const float* verts) {
    Hazard h;
    h.entity.kind = (EntityKind)d.kind;
    h.entity.a = d.a;
    h.entity.b = d.b;
    std::memcpy(h.entity.dt, d.dt, sizeof d.dt);
    h.dynamic = d.dynamic != 0;
    SPolygon s = spolygon_new(to_points(verts, d.n_verts));
    if (d.has_pre) s = spolygon_transform_clone(s, transformation_from_dt(d.pre[0], d.pre[1], d.pre[2]));
    if (d.has_dt) s = spolygon_transform_clone(s, transformation_from_dt(d.dt[0], d.dt[1], d.dt[2]));
    h.shape = std::move(s);
    return h;
}

void* or_cde_create(const float* bbox4, uint32_t depth, uint32_t cdt, uint32_t n, const or_hazard_desc* descs,
                    const float* verts) {
    try {
        std::vector<Hazard> hs;
        uint64_t off = 0;
        for (uint32_t i = 0; i < n; i++) {
            hs.push_back(make_hazard(descs[i], verts + off));
            off += 2ull * descs[i].n_verts;
        }
        CDEConfig cfg{};
        cfg.quadtree_depth = (uint8_t)depth;
        cfg.cd_threshold = (uint8_t)cdt;
        return new CDEngine(Rect{bbox4[0], bbox4[1], bbox4[2], bbox4[3]}, std::move(hs), cfg);
    } catch (...) {
        return nullptr;
    }
}
void or_cde_destroy(void* c) { delete (CDEngine*)c; }

int or_cde_register(void* c, const or_hazard_desc* d, const float* verts, uint32_t* key_out) {
    try {
        HazKey k = ((CDEngine*)c)->register_hazard(make_hazard(*d, verts));
        key_out[0] = k.idx;
        key_out[1] = k.version;
        return 0;
    } catch (...) {
        return -1;
    }
}
int or_cde_deregister_entity(void* c, const or_hazard_desc* d) {
    try {
        HazardEntity e;
        e.kind = (EntityKind)d->kind;
        e.a = d->a;
        e.b = d->b;
        std::memcpy(e.dt, d->dt, sizeof d->dt);
        ((CDEngine*)c)->deregister_hazard_by_entity(e);
        return 0;
    } catch (...) {
        return -1;
    }
}
int or_cde_deregister_key(void* c, uint32_t idx, uint32_t version) {
    try {
        ((CDEngine*)c)->deregister_hazard_by_key(HazKey{idx, version});
        return 0;
    } catch (...) {
        return -1;
    }
}
void* or_cde_save(void* c) { return new std::vector<Hazard>(((CDEngine*)c)->save()); }
int or_cde_restore(void* c, void* snap) {
    try {
        ((CDEngine*)c)->restore(*(std::vector<Hazard>*)snap);
        return 0;
    } catch (...) {
        return -1;
    }
}
void or_snapshot_free(void* s) { delete (std::vector<Hazard>*)s; }

// Invariant bits: 1 = no dangling hazards, 2 = hazard vectors sorted with correct edge caches.
int or_cde_invariants(void* c) {
    const CDEngine& e = *(CDEngine*)c;
    return (qt_contains_no_dangling_hazards(e) ? 1 : 0) | (qt_hazard_vecs_sorted_and_cached(e) ? 2 : 0);
}
// util/assertions.rs:91-104 — two engines hold the same tree and the same hazard entities.
int or_cde_trees_match(void* a, void* b) {
    const CDEngine& x = *(CDEngine*)a;
    const CDEngine& y = *(CDEngine*)b;
    if (!qt_nodes_match(&x.quadtree, &y.quadtree)) return 0;
    std::vector<HazardEntity> ex, ey;
    x.hazards_map.for_each([&](HazKey, const Hazard& h) { ex.push_back(h.entity); });
    y.hazards_map.for_each([&](HazKey, const Hazard& h) { ey.push_back(h.entity); });
    if (ex.size() != ey.size()) return 0;
    for (auto& e : ex) {
        bool f = false;
        for (auto& g : ey) f |= e == g;
        if (!f) return 0;
    }
    return 1;
}

// Tree dump in BFS order, for structural comparison with the product's flattened tree.
// Per node (8 words): level, n_children(0/4), n_hazards, n_active_edges, bbox as 4 float bits.
// Per hazard (8 words, nodes in BFS order, hazards in vector order): slot idx, presence,
// n_edges, 0, ff_bbox as 4 float bits (zero unless Partial).
// Returns the node count; hazard count in *n_haz.  Call with null buffers to size.
int64_t or_cde_dump(void* c, uint32_t* nodes, uint64_t nodes_cap, uint32_t* haz, uint64_t haz_cap, uint64_t* n_haz) {
    const CDEngine& e = *(CDEngine*)c;
    std::vector<const QTNode*> q{&e.quadtree};
    uint64_t hn = 0;
    for (size_t i = 0; i < q.size(); i++) {
        const QTNode* nd = q[i];
        if (nodes && i < nodes_cap) {
            uint32_t* o = nodes + 8 * i;
            o[0] = nd->level;
            o[1] = nd->children ? 4 : 0;
            o[2] = (uint32_t)nd->hazards.hazards.size();
            o[3] = (uint32_t)nd->hazards.n_active_edges;
            std::memcpy(o + 4, &nd->bbox, 16);
        }
        for (const auto& h : nd->hazards.hazards) {
            if (haz && hn < haz_cap) {
                uint32_t* o = haz + 8 * hn;
                o[0] = h.hkey.idx;
                o[1] = (uint32_t)h.presence;
                o[2] = (uint32_t)h.n_edges();
                o[3] = 0;
                Rect z{0, 0, 0, 0};
                std::memcpy(o + 4, h.presence == Presence::Partial ? &h.partial.ff_bbox : &z, 16);
            }
            hn++;
        }
        if (nd->children)
            for (const auto& ch : *nd->children) q.push_back(&ch);
    }
    *n_haz = hn;
    return (int64_t)q.size();
}

void* or_item_create(const float* xy, uint32_t n, const uint32_t* lim_n, const float* lim_cov, uint32_t n_ff_poles,
                     uint32_t n_ff_piers) {
    try {
        auto* it = new OrItem;
        it->shape = spolygon_new(to_points(xy, n));
        spolygon_generate_surrogate(it->shape, make_cfg(lim_n, lim_cov, n_ff_poles, n_ff_piers));
        return it;
    } catch (...) {
        return nullptr;
    }
}
// An imported item's internal shape (Item::new, entities/item.rs:22-42): SPolygon::new of the
// original, transformed by pre (the centering), then the surrogate.
void* or_item_create_pre(const float* xy, uint32_t n, const float* pre, const uint32_t* lim_n, const float* lim_cov,
                         uint32_t n_ff_poles, uint32_t n_ff_piers) {
    try {
        auto* it = new OrItem;
        it->shape = spolygon_transform_clone(spolygon_new(to_points(xy, n)), transformation_from_dt(pre[0], pre[1], pre[2]));
        spolygon_generate_surrogate(it->shape, make_cfg(lim_n, lim_cov, n_ff_poles, n_ff_piers));
        return it;
    } catch (...) {
        return nullptr;
    }
}

// import_simple_polygon's vertices (io/import.rs:198-212) passed through SPolygon::new; the
// centroid of the polygon (simple_polygon.rs:173-192).  Returns the vertex count, -1 on error.
int or_import_polygon(const float* xy, uint32_t n, float* out_xy, float* centroid2) {
    try {
        SPolygon s = spolygon_new(import_simple_polygon_points(to_points(xy, n)));
        for (size_t i = 0; i < s.vertices.size(); i++) { out_xy[2 * i] = s.vertices[i].x; out_xy[2 * i + 1] = s.vertices[i].y; }
        const Point c = spolygon_centroid(s);
        centroid2[0] = c.x;
        centroid2[1] = c.y;
        return (int)s.vertices.size();
    } catch (...) {
        return -1;
    }
}

int or_ext_to_int(const float* ext, const float* pre, float* out) { return ext_to_int_transformation(ext, pre, out) ? 0 : -1; }
int or_int_to_ext(const float* in, const float* pre, float* out) { return int_to_ext_transformation(in, pre, out) ? 0 : -1; }

// Item built from already-known vertices, POI and surrogate (to check query parity on
// shared inputs independently of the construction path).
void* or_item_from_parts(const float* xy, uint32_t n, const float* meta9, const float* poles, uint32_t n_poles,
                         const float* piers, uint32_t n_piers, uint32_t n_ff_poles) {
    auto* it = new OrItem;
    it->shape.vertices = to_points(xy, n);
    it->shape.area = meta9[0];
    it->shape.diameter = meta9[1];
    it->shape.bbox = Rect{meta9[2], meta9[3], meta9[4], meta9[5]};
    it->shape.poi = Circle{{meta9[6], meta9[7]}, meta9[8]};
    for (uint32_t i = 0; i < n_poles; i++)
        it->shape.surrogate.poles.push_back(Circle{{poles[3 * i], poles[3 * i + 1]}, poles[3 * i + 2]});
    for (uint32_t i = 0; i < n_piers; i++)
        it->shape.surrogate.piers.push_back(Edge{{piers[4 * i], piers[4 * i + 1]}, {piers[4 * i + 2], piers[4 * i + 3]}});
    it->shape.surrogate.config.n_ff_poles = n_ff_poles;
    it->shape.surrogate.present = true;
    return it;
}
void or_item_destroy(void* it) { delete (OrItem*)it; }
// An item's shape: vertices (returns their count), meta = [area, diameter, bbox (4), poi (3)],
// poles [n][3].
int or_item_info(const void* it, float* out_xy, uint32_t xy_cap, float* meta9, float* poles, uint32_t poles_cap,
                 uint32_t* n_poles) {
    const SPolygon& s = ((const OrItem*)it)->shape;
    for (size_t i = 0; i < s.vertices.size() && i < xy_cap; i++) { out_xy[2 * i] = s.vertices[i].x; out_xy[2 * i + 1] = s.vertices[i].y; }
    const float m[9] = {s.area, s.diameter, s.bbox.x_min, s.bbox.y_min, s.bbox.x_max, s.bbox.y_max,
                        s.poi.center.x, s.poi.center.y, s.poi.radius};
    std::memcpy(meta9, m, sizeof m);
    *n_poles = (uint32_t)s.surrogate.poles.size();
    for (uint32_t i = 0; i < *n_poles && i < poles_cap; i++) {
        poles[3 * i] = s.surrogate.poles[i].center.x;
        poles[3 * i + 1] = s.surrogate.poles[i].center.y;
        poles[3 * i + 2] = s.surrogate.poles[i].radius;
    }
    return (int)s.vertices.size();
}

// One batch of collision queries over `n` DTransformations (dt[i] = rotation, tx, ty) or, when
// `mat` is non-null, over the 2x3 matrices mat[i] (m00 m01 m02 m10 m11 m12).
// mode bit 0 = detect (status), bit 1 = collect (set count/hash/keys).
// Work is split into `n_threads` static contiguous blocks (one collector per thread).
// counters (optional, 4 words) receive the summed traversal work; *seconds the wall time.
int or_query_batch(void* cde_p, void* item_p, const float* dt, const float* mat, uint64_t n, uint32_t mode,
                   const OrFilter* filt, uint8_t* status, uint8_t* surr_hit, uint32_t* set_count, uint64_t* set_hash,
                   uint32_t* set_keys, uint32_t keys_cap, uint32_t n_threads, uint64_t* counters, double* seconds) {
    const CDEngine& cde = *(CDEngine*)cde_p;
    const OrItem& item = *(OrItem*)item_p;
    if (n_threads == 0) n_threads = 1;
    size_t n_slots = cde.hazards_map.slots.size();
    HazardFilter base_filter;
    if (filt && filt->kind == 1) {
        base_filter.irrelevant.assign(n_slots, 0);
        for (uint32_t i = 0; i < filt->n_keys; i++)
            if (filt->keys[i] < n_slots) base_filter.irrelevant[filt->keys[i]] = 1;
    }
    std::vector<Counters> cts(n_threads);
    int err = 0;
    auto worker = [&](uint32_t tid) {
        uint64_t lo = n * tid / n_threads, hi = n * (tid + 1) / n_threads;
        HazardFilter f = base_filter;
        HazardCollector col;
        col.reset(n_slots);
        SPolygon buffer = item.shape;
        Counters* ct = counters ? &cts[tid] : nullptr;
        try {
            for (uint64_t q = lo; q < hi; q++) {
                Transformation t;
                if (mat) {
                    const float* m = mat + 6 * q;
                    t.m[0][0] = m[0]; t.m[0][1] = m[1]; t.m[0][2] = m[2];
                    t.m[1][0] = m[3]; t.m[1][1] = m[4]; t.m[1][2] = m[5];
                } else {
                    t = transformation_from_dt(dt[3 * q], dt[3 * q + 1], dt[3 * q + 2]);
                }
                if (filt && filt->kind == 2) {
                    f.irrelevant.assign(n_slots, 0);
                    for (uint32_t s = 1; s < n_slots; s++)
                        f.irrelevant[s] = hashed_irrelevant(filt->seed, filt->threshold, q, s) ? 1 : 0;
                }
                bool transformed = false;
                auto do_transform = [&]() {
                    if (transformed) return;
                    for (size_t i = 0; i < buffer.vertices.size(); i++)
                        buffer.vertices[i] = transform_point(item.shape.vertices[i], t);
                    buffer.poi.center = transform_point(item.shape.poi.center, t);
                    buffer.bbox = spolygon_bbox(buffer.vertices);
                    transformed = true;
                };
                if (mode & 1) {
                    bool sh = cde.detect_surrogate_collision(item.shape.surrogate, t, f, ct);
                    if (surr_hit) surr_hit[q] = sh ? 1 : 0;
                    uint8_t st;
                    if (sh) {
                        st = Q_COLLISION;
                    } else {
                        do_transform();
                        st = cde.detect_poly_collision(buffer, f, ct);
                    }
                    status[q] = st;
                }
                if (mode & 2) {
                    col.reset(n_slots);
                    // mode bit 4: the collector arrives pre-seeded with the key filter's slots
                    // (a collector used as the filter, filter.rs:69-76)
                    if ((mode & 4) && filt && filt->kind == 1)
                        for (uint32_t i = 0; i < filt->n_keys; i++)
                            if (filt->keys[i] < n_slots) col.insert(HazKey{filt->keys[i], 0});
                    cde.collect_surrogate_collisions(item.shape.surrogate, t, col, ct);
                    do_transform();
                    cde.collect_poly_collisions(buffer, col, ct);
                    uint64_t h = 0;
                    for (uint32_t k : col.keys) h += mix64((uint64_t)k);
                    set_count[q] = (uint32_t)col.len();
                    set_hash[q] = h;
                    if (set_keys) {
                        std::vector<uint32_t> ks = col.keys;
                        std::sort(ks.begin(), ks.end());
                        for (uint32_t i = 0; i < keys_cap; i++)
                            set_keys[q * keys_cap + i] = i < ks.size() ? ks[i] : 0xffffffffu;
                    }
                }
            }
        } catch (...) {
            err = -1;
        }
    };
    auto t0 = std::chrono::steady_clock::now();
    if (n_threads == 1) {
        worker(0);
    } else {
        std::vector<std::thread> th;
        for (uint32_t i = 0; i < n_threads; i++) th.emplace_back(worker, i);
        for (auto& x : th) x.join();
    }
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (counters) {
        for (int i = 0; i < 4; i++) counters[i] = 0;
        for (auto& c : cts) {
            counters[0] += c.nodes; counters[1] += c.entries; counters[2] += c.edges; counters[3] += c.pip_vertices;
        }
    }
    return err;
}

}  // extern "C"
