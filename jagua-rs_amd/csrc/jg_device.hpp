// Device-side views of the flattened engine and the query item, and the traversal
// primitives shared by every query kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jg_geom.hpp"

namespace jg {

constexpr uint32_t ENTIRE_BIT = 0x80000000u;
constexpr uint32_t SLOT_MASK = 0x7fffffffu;
constexpr uint32_t COUNT_MASK = 0x3fffffffu;
constexpr uint32_t HIT = 0xffffffffu;

struct DevNode {                      // mirrors FlatNode (32 B)
    float4 bbox;
    int32_t first_child;
    uint32_t haz_begin, haz_info, n_active_edges;
};

struct DevEntry {                     // mirrors FlatEntry (32 B)
    float4 ff;
    uint32_t slot_flags, edge_begin, edge_count, n_edges_total;
};

struct CdeView {
    const DevNode* nodes;
    const uint32_t* node_up;          // parent << 2 | sibling index
    const DevEntry* ent;
    const float4* edges;
    const float4* haz_bbox;
    const float4* haz_poi;            // cx, cy, r, interior
    const uint2* haz_vr;
    const float2* haz_verts;
    uint32_t n_slots, exterior_slot, cd_threshold, n_nodes;
};

struct ItemView {
    const float2* v;                  // local vertices (CCW)
    const float4* ff;                 // ff poles (cx, cy, r, 0) then piers (x1 y1 x2 y2)
    float4 poi;
    uint32_t n_v, n_poles, n_piers, pad;
};

struct FilterView {
    const uint32_t* mask;             // JG_FILTER_KEYS: bit per slot
    uint32_t kind, seed, threshold, n_words;
};

__device__ __forceinline__ bool irrelevant(const FilterView& f, uint64_t q, uint32_t slot) {
    if (f.kind == 0) return false;
    if (f.kind == 1) return (slot >> 5) < f.n_words && ((f.mask[slot >> 5] >> (slot & 31u)) & 1u);
    return filter_hash(f.seed, q, slot) < f.threshold;
}

__device__ __forceinline__ Mat load_mat(const float* m6, uint64_t i) {
    const float* p = m6 + 6 * i;
    return Mat{p[0], p[1], p[2], p[3], p[4], p[5]};
}

// QTHazardVec::strongest (qt_hazard_vec.rs:60-63): index of the first relevant entry of
// node `nd`, or -1.  Without a filter it is the first entry, whose presence the node record
// carries, so no entry is read.
__device__ __forceinline__ int strongest(const CdeView& c, const DevNode& nd, const FilterView& f, uint64_t q,
                                         bool& entire) {
    const uint32_t cnt = nd.haz_info & COUNT_MASK;
    if (f.kind == 0) {
        entire = (nd.haz_info & ENTIRE_BIT) != 0u;
        return cnt ? (int)nd.haz_begin : -1;
    }
    for (uint32_t i = 0; i < cnt; i++) {
        const uint32_t sf = c.ent[nd.haz_begin + i].slot_flags;
        if (!irrelevant(f, q, sf & SLOT_MASK)) {
            entire = (sf & ENTIRE_BIT) != 0u;
            return (int)(nd.haz_begin + i);
        }
    }
    return -1;
}

// Depth-first descent in the reference's quadrant order without a memory stack: the
// pending-sibling masks of the current path live 4 bits per level in one 64-bit register;
// moving to a sibling or the parent uses the node_up links of the BFS image.
// visit(node, &first_child) returns HIT to stop, or the mask of children to descend into.
template <class Visit>
__device__ __forceinline__ bool descend(const CdeView& c, int start, Visit visit) {
    int n = start;
    uint32_t d = 0;
    uint64_t pend = 0;
    for (;;) {
        int fc = -1;
        const uint32_t m = visit(n, fc);
        if (m == HIT) return true;
        if (m != 0u) {
            const uint32_t k = (uint32_t)__builtin_ctz(m);
            pend |= (uint64_t)(m & ~(1u << k)) << (4u * d);
            d++;
            n = fc + (int)k;
            continue;
        }
        for (;;) {
            if (d == 0) return false;
            const uint32_t sh = 4u * (d - 1u);
            const uint32_t nib = (uint32_t)(pend >> sh) & 0xfu;
            const uint32_t up = c.node_up[n];
            if (nib != 0u) {
                const uint32_t k = (uint32_t)__builtin_ctz(nib);
                pend &= ~((uint64_t)1 << (sh + k));
                n = n - (int)(up & 3u) + (int)k;
                break;
            }
            n = (int)(up >> 2);
            d--;
        }
    }
}

// CDEngine::get_virtual_root (cd_engine.rs:314-327); children's boxes are the quadrants of
// the node's box (rect.rs:141-151), exactly as the tree built them.
__device__ __forceinline__ int virtual_root(const CdeView& c, float4 bb) {
    int n = 0;
    for (;;) {
        const DevNode nd = c.nodes[n];
        if (nd.first_child < 0) return n;
        int next = -1;
        for (int k = 0; k < 4; k++)
            if (box_relation(box_quadrant(nd.bbox, k), bb) == REL_SURROUNDING) { next = nd.first_child + k; break; }
        if (next < 0) return n;
        n = next;
    }
}

// QTNode::collides for a circle (qt_node.rs:99-147; Circle quadrant test qt_traits.rs:10-13)
__device__ __forceinline__ bool collides_circle(const CdeView& c, int start, float cx, float cy, float r,
                                                const FilterView& f, uint64_t q) {
    return descend(c, start, [&](int n, int& fc) -> uint32_t {
        const DevNode nd = c.nodes[n];
        bool entire = false;
        const int s = strongest(c, nd, f, q, entire);
        if (s < 0) return 0u;
        if (entire) return HIT;
        if (nd.first_child >= 0) {
            fc = nd.first_child;
            uint32_t m = 0;
            for (int k = 0; k < 4; k++) m |= circle_box(cx, cy, r, box_quadrant(nd.bbox, k)) ? (1u << k) : 0u;
            return m;
        }
        const int he = (int)(nd.haz_begin + (nd.haz_info & COUNT_MASK));
        for (int i = s; i < he; i++) {
            const DevEntry e = c.ent[i];
            if (irrelevant(f, q, e.slot_flags & SLOT_MASK)) continue;
            if (!circle_box(cx, cy, r, e.ff)) continue;
            for (uint32_t j = 0; j < e.edge_count; j++)
                if (circle_edge(cx, cy, r, c.edges[e.edge_begin + j])) return HIT;
        }
        return 0u;
    });
}

// QTNode::collides for an edge (qt_node.rs:99-147; Edge quadrant test qt_traits.rs:19-106)
__device__ __forceinline__ bool collides_edge(const CdeView& c, int start, float4 ed, const FilterView& f, uint64_t q) {
    return descend(c, start, [&](int n, int& fc) -> uint32_t {
        const DevNode nd = c.nodes[n];
        bool entire = false;
        const int s = strongest(c, nd, f, q, entire);
        if (s < 0) return 0u;
        if (entire) return HIT;
        if (nd.first_child >= 0) {
            fc = nd.first_child;
            float4 qs[4];
            for (int k = 0; k < 4; k++) qs[k] = box_quadrant(nd.bbox, k);
            return edge_quadrants(ed, nd.bbox, qs);
        }
        const int he = (int)(nd.haz_begin + (nd.haz_info & COUNT_MASK));
        for (int i = s; i < he; i++) {
            const DevEntry e = c.ent[i];
            if (irrelevant(f, q, e.slot_flags & SLOT_MASK)) continue;
            if (!box_edge(e.ff, ed)) continue;
            for (uint32_t j = 0; j < e.edge_count; j++)
                if (edge_edge(ed, c.edges[e.edge_begin + j])) return HIT;
        }
        return 0u;
    });
}

// ---------------------------------------------------------------------------------------
// Deferred leaf tests.  A wave's lanes descend independently; a lane that reaches a leaf
// whose strongest relevant hazard is Partial records (lane, leaf) in a per-wave LDS list
// instead of testing it inline, and the wave then tests all recorded leaves with all of its
// lanes.  The verdict of a lane is the OR of its Entire hits and its leaf hits, exactly the
// result of QTNode::collides (qt_node.rs:99-147), whose recursion only decides which hazard
// is reported first.

struct Deferred {                     // per-wave LDS
    uint32_t* items;                  // lane << 26 | leaf node
    uint32_t* count;
    uint32_t cap;
    float4* geom;                     // per lane: circle (cx, cy, r, 1) or edge (x1 y1 x2 y2)
    uint32_t* kind;                   // per lane: 1 circle, 0 edge
    uint64_t* qidx;                   // per lane: query index (hashed filters)
    uint32_t* hit;                    // per lane: leaf hit flag
};

// Entity kinds of a descent: KIND_EDGE and KIND_CIRCLE are fixed at compile time;
// KIND_MIXED reads the per-lane kind (surrogates with both poles and piers).
enum : int { KIND_EDGE = 0, KIND_CIRCLE = 1, KIND_MIXED = 2 };

template <bool FILT>
__device__ __forceinline__ bool irrel(const FilterView& f, uint64_t q, uint32_t slot) {
    return FILT && irrelevant(f, q, slot);
}

template <bool FILT>
__device__ __forceinline__ int strongest_t(const CdeView& c, const DevNode& nd, const FilterView& f, uint64_t q,
                                           bool& entire) {
    if (!FILT) {
        entire = (nd.haz_info & ENTIRE_BIT) != 0u;
        return (nd.haz_info & COUNT_MASK) ? (int)nd.haz_begin : -1;
    }
    return strongest(c, nd, f, q, entire);
}

// QTHazPartial::collides_with over the relevant Partial entries of a leaf, from entry s on
// (qt_node.rs:126-141, qt_partial_hazard.rs:67-72).
template <bool FILT, int KIND>
__device__ __forceinline__ bool leaf_hits(const CdeView& c, const DevNode& nd, int s, bool circle_rt, float4 g,
                                          const FilterView& f, uint64_t q) {
    const bool circle = KIND == KIND_CIRCLE || (KIND == KIND_MIXED && circle_rt);
    const int he = (int)(nd.haz_begin + (nd.haz_info & COUNT_MASK));
    for (int i = s; i < he; i++) {
        const DevEntry e = c.ent[i];
        if (irrel<FILT>(f, q, e.slot_flags & SLOT_MASK)) continue;
        if (circle) {
            if (!circle_box(g.x, g.y, g.z, e.ff)) continue;
            for (uint32_t j = 0; j < e.edge_count; j++)
                if (circle_edge(g.x, g.y, g.z, c.edges[e.edge_begin + j])) return true;
        } else {
            if (!box_edge(e.ff, g)) continue;
            for (uint32_t j = 0; j < e.edge_count; j++)
                if (edge_edge(g, c.edges[e.edge_begin + j])) return true;
        }
    }
    return false;
}

// Descent of one entity (circle: g = (cx, cy, r, -); edge: g = segment) from `start`,
// returning true on an Entire hit; leaves are deferred (tested inline when the list is full).
template <bool FILT, int KIND>
__device__ __forceinline__ bool descend_deferring(const CdeView& c, int start, bool circle_rt, float4 g,
                                                  const FilterView& f, uint64_t q, uint32_t lane, const Deferred& D) {
    const bool circle = KIND == KIND_CIRCLE || (KIND == KIND_MIXED && circle_rt);
    return descend(c, start, [&](int n, int& fc) -> uint32_t {
        const DevNode nd = c.nodes[n];
        bool entire = false;
        const int s = strongest_t<FILT>(c, nd, f, q, entire);
        if (s < 0) return 0u;
        if (entire) return HIT;
        if (nd.first_child >= 0) {
            fc = nd.first_child;
            if (circle) {
                uint32_t m = 0;
                for (int k = 0; k < 4; k++) m |= circle_box(g.x, g.y, g.z, box_quadrant(nd.bbox, k)) ? (1u << k) : 0u;
                return m;
            }
            float4 qs[4];
            for (int k = 0; k < 4; k++) qs[k] = box_quadrant(nd.bbox, k);
            return edge_quadrants(g, nd.bbox, qs);
        }
        const uint32_t at = atomicAdd(D.count, 1u);
        if (at < D.cap) {
            D.items[at] = (lane << 26) | (uint32_t)n;
            return 0u;
        }
        return leaf_hits<FILT, KIND>(c, nd, s, circle_rt, g, f, q) ? HIT : 0u;
    });
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Tests every deferred leaf with the wave's lanes (round-robin) and returns this lane's flag.
template <bool FILT, int KIND>
__device__ __forceinline__ bool test_deferred(const CdeView& c, const FilterView& f, const Deferred& D, uint32_t lane) {
    wave_sync_lds();
    const uint32_t n = min(*D.count, D.cap);
    for (uint32_t i = lane; i < n; i += 64) {
        const uint32_t it = D.items[i];
        const uint32_t owner = it >> 26;
        if (D.hit[owner]) continue;
        const DevNode nd = c.nodes[it & 0x3ffffffu];
        const uint64_t q = FILT ? D.qidx[owner] : 0ull;
        bool entire = false;
        const int s = strongest_t<FILT>(c, nd, f, q, entire);
        if (s >= 0 && leaf_hits<FILT, KIND>(c, nd, s, KIND == KIND_MIXED && D.kind[owner] != 0u, D.geom[owner], f, q))
            D.hit[owner] = 1u;
    }
    wave_sync_lds();
    return D.hit[lane] != 0u;
}

// Hazard-polygon point-in-polygon (simple_polygon.rs:255-292) over the hazard's full shape.
__device__ __forceinline__ bool pip_hazard(const CdeView& c, uint32_t slot, float px, float py) {
    const float4 bb = c.haz_bbox[slot];
    if (!box_has_point(bb, px, py)) return false;
    const float ray_x = bb.z + (bb.z - bb.x);
    const uint2 vr = c.haz_vr[slot];
    const float2* v = c.haz_verts + vr.x;
    int cnt = 0;
    float2 a = v[0];
    for (uint32_t i = 0; i < vr.y; i++) {
        const float2 b = v[i + 1 == vr.y ? 0 : i + 1];
        cnt += pip_edge(px, py, ray_x, a.x, a.y, b.x, b.y);
        a = b;
    }
    return (cnt & 1) == 1;
}

// Query-polygon point-in-polygon on the transformed query (vertices transformed on the fly).
__device__ __forceinline__ bool pip_query(const ItemView& it, const Mat& m, float4 qbb, float px, float py) {
    if (!box_has_point(qbb, px, py)) return false;
    const float ray_x = qbb.z + (qbb.z - qbb.x);
    int cnt = 0;
    float2 a = xform(m, it.v[0].x, it.v[0].y);
    const float2 first = a;
    for (uint32_t i = 0; i < it.n_v; i++) {
        const float2 b = i + 1 == it.n_v ? first : xform(m, it.v[i + 1].x, it.v[i + 1].y);
        cnt += pip_edge(px, py, ray_x, a.x, a.y, b.x, b.y);
        a = b;
    }
    return (cnt & 1) == 1;
}

// CDEngine::detect_containment_collision (cd_engine.rs:231-255)
__device__ __forceinline__ bool containment(const CdeView& c, uint32_t slot, const ItemView& it, const Mat& m,
                                            float4 qbb, float2 qpoi) {
    const int rel = box_almost_relation(c.haz_bbox[slot], qbb);
    bool contained = false;
    if (rel == REL_SURROUNDING) {
        contained = pip_hazard(c, slot, qpoi.x, qpoi.y);
    } else if (rel == REL_ENCLOSED) {
        const float4 hp = c.haz_poi[slot];
        contained = pip_query(it, m, qbb, hp.x, hp.y);
    }
    const bool interior = c.haz_poi[slot].w != 0.0f;
    return interior ? contained : !contained;
}

// Transformed bounding box of the query (simple_polygon.rs:122-133 after transform_from).
__device__ __forceinline__ float4 query_bbox(const ItemView& it, const Mat& m) {
    float4 b = make_float4(3.40282347e+38f, 3.40282347e+38f, -3.40282347e+38f, -3.40282347e+38f);
    for (uint32_t i = 0; i < it.n_v; i++) {
        const float2 p = xform(m, it.v[i].x, it.v[i].y);
        b.x = mn(b.x, p.x);
        b.y = mn(b.y, p.y);
        b.z = mx(b.z, p.x);
        b.w = mx(b.w, p.y);
    }
    return b;
}

}  // namespace jg
