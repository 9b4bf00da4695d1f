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

// Reads of the engine image and the item records.  Their pointers come out of descriptor
// tables, so the compiler sees generic addresses and would issue flat loads, which count
// against the LDS wait counter as well: every LDS read of the traversal would then wait for
// the node and edge loads in flight.  These read through global-address-space pointers
// (global_load) instead.
#if defined(__HIP_DEVICE_COMPILE__)
typedef int jg_i4 __attribute__((ext_vector_type(4)));
typedef int jg_i2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const jg_i4 jg_gi4;
typedef __attribute__((address_space(1))) const jg_i2 jg_gi2;
typedef __attribute__((address_space(1))) const int jg_gi1;
__device__ __forceinline__ float4 ldg4(const float4* p) {
    const jg_i4 v = *(jg_gi4*)p;
    return make_float4(__int_as_float(v.x), __int_as_float(v.y), __int_as_float(v.z), __int_as_float(v.w));
}
__device__ __forceinline__ float2 ldg2(const float2* p) {
    const jg_i2 v = *(jg_gi2*)p;
    return make_float2(__int_as_float(v.x), __int_as_float(v.y));
}
__device__ __forceinline__ uint2 ldg2u(const uint2* p) {
    const jg_i2 v = *(jg_gi2*)p;
    return make_uint2((uint32_t)v.x, (uint32_t)v.y);
}
__device__ __forceinline__ uint32_t ldg1(const uint32_t* p) { return (uint32_t)*(jg_gi1*)p; }
__device__ __forceinline__ DevNode ldg_node(const DevNode* p) {
    const jg_i4 a = ((jg_gi4*)p)[0], b = ((jg_gi4*)p)[1];
    DevNode n;
    n.bbox = make_float4(__int_as_float(a.x), __int_as_float(a.y), __int_as_float(a.z), __int_as_float(a.w));
    n.first_child = b.x;
    n.haz_begin = (uint32_t)b.y;
    n.haz_info = (uint32_t)b.z;
    n.n_active_edges = (uint32_t)b.w;
    return n;
}
__device__ __forceinline__ DevEntry ldg_ent(const DevEntry* p) {
    const jg_i4 a = ((jg_gi4*)p)[0], b = ((jg_gi4*)p)[1];
    DevEntry e;
    e.ff = make_float4(__int_as_float(a.x), __int_as_float(a.y), __int_as_float(a.z), __int_as_float(a.w));
    e.slot_flags = (uint32_t)b.x;
    e.edge_begin = (uint32_t)b.y;
    e.edge_count = (uint32_t)b.z;
    e.n_edges_total = (uint32_t)b.w;
    return e;
}
#else
__device__ inline float4 ldg4(const float4* p) { return *p; }
__device__ inline float2 ldg2(const float2* p) { return *p; }
__device__ inline uint2 ldg2u(const uint2* p) { return *p; }
__device__ inline uint32_t ldg1(const uint32_t* p) { return *p; }
__device__ inline DevNode ldg_node(const DevNode* p) { return *p; }
__device__ inline DevEntry ldg_ent(const DevEntry* p) { return *p; }
#endif
// The slot | Entire word of an entry (its fifth word).
__device__ __forceinline__ uint32_t ldg_slot_flags(const DevEntry* p) { return ldg1((const uint32_t*)p + 4); }

struct CdeView {
    const DevNode* nodes;
    const uint32_t* node_up;          // parent << 2 | sibling index
    const DevEntry* ent;
    const float4* edges;
    const float4* haz_bbox;
    const float4* haz_poi;            // cx, cy, r, interior
    const uint2* haz_vr;
    const float2* haz_verts;
    const uint32_t* cells;            // the deepest node on the path to each cell of a
                                      // 2^cell_level grid over the root box (row iy, column ix)
    uint32_t n_slots, exterior_slot, cd_threshold, n_nodes;
    uint32_t cell_level, pad_[3];
    float4 ext_inner;                 // a box inside the container (the exterior hazard's shape):
                                      // its bbox when that shape is an axis-aligned rectangle,
                                      // else empty (x_min > x_max)
};

// A cell table entry: cdepth << 28 | depth << 24 | node, for the deepest node on the path to the
// cell and the depth of the first node on that path where QTNode::collect_collisions tests
// entries (15: none on the path).
constexpr uint32_t CELL_NODE = 0xffffffu;
__device__ __forceinline__ uint32_t cell_depth(uint32_t e) { return (e >> 24) & 15u; }
__device__ __forceinline__ uint32_t cell_cdepth(uint32_t e) { return e >> 28; }

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
// moving to a sibling or the parent uses the node_up links of the device image.
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
// the node's box (rect.rs:141-151), exactly as the tree built them.  depth: the node's depth.
__device__ __forceinline__ int virtual_root(const CdeView& c, float4 bb, uint32_t& depth) {
    int n = 0;
    depth = 0;
    for (;;) {
        const DevNode nd = ldg_node(c.nodes + n);
        if (nd.first_child < 0) return n;
        const int k = quadrant_surrounding(nd.bbox, bb);
        if (k < 0) return n;
        n = nd.first_child + k;
        depth++;
    }
}
__device__ __forceinline__ int virtual_root(const CdeView& c, float4 bb) {
    uint32_t depth;
    return virtual_root(c, bb, depth);
}

// virtual_root through the cell tables: the quadrants of the root box (root: node 0's box) are
// followed arithmetically (box_quadrant, the first quadrant that surrounds bb, as the search
// takes it) down to the tables' deepest level or the first level where none surrounds bb; the
// table entry of that cell is the deepest node on the path, where the search ends unless it
// lies at the tables' deepest level, from which it goes on through node records.
__device__ __forceinline__ int virtual_root_cells(const CdeView& c, float4 root, float4 bb, uint32_t& depth) {
    int n = 0;
    depth = 0;
    if (c.cell_level > 0u) {
        float4 r = root;
        uint32_t ix = 0, iy = 0, l = 0;
        for (; l < c.cell_level; l++) {
            const int k = quadrant_surrounding(r, bb);
            if (k < 0) break;
            r = box_quadrant(r, k);
            ix = 2u * ix + ((k == 0 || k == 3) ? 1u : 0u);
            iy = 2u * iy + ((k == 0 || k == 1) ? 1u : 0u);
        }
        const uint32_t e = ldg1(c.cells + ((1u << (2u * l)) - 1u) / 3u + ((iy << l) | ix));
        n = (int)(e & CELL_NODE);
        depth = cell_depth(e);
        if (depth < l || l < c.cell_level) return n;
    }
    for (;;) {
        const DevNode nd = ldg_node(c.nodes + n);
        if (nd.first_child < 0) return n;
        const int k = quadrant_surrounding(nd.bbox, bb);
        if (k < 0) return n;
        n = nd.first_child + k;
        depth++;
    }
}

// The cell (ix, iy) at level `depth` of a node whose box `box` is a quadrant `depth` levels below
// the root box: the quadrants holding the box's centre (strictly on one side of every midline
// above it, as the box is one of the subdivision's).
__device__ __forceinline__ void node_cell(float4 root, float4 box, uint32_t depth, uint32_t& ix, uint32_t& iy) {
    const float cx = (box.x + box.z) / 2.0f, cy = (box.y + box.w) / 2.0f;
    float4 r = root;
    ix = 0;
    iy = 0;
    for (uint32_t l = 0; l < depth; l++) {
        const float mxm = (r.x + r.z) / 2.0f, mym = (r.y + r.w) / 2.0f;
        const bool xh = cx > mxm, yh = cy > mym;
        r = make_float4(xh ? mxm : r.x, yh ? mym : r.y, xh ? r.z : mxm, yh ? r.w : mym);
        ix = 2u * ix + (xh ? 1u : 0u);
        iy = 2u * iy + (yh ? 1u : 0u);
    }
}

// A node from which the descent of an edge with bounding box eb, starting at a node n0 that
// holds it (depth d0 below the cell tables' deepest level, box box0, cell (x0, y0)), may start
// instead: following the quadrant subdivision below n0 (box_quadrant's midpoints, the boxes the
// tree's nodes have) while eb lies strictly inside one quadrant, down to the cell tables'
// deepest level, the cell table gives the deepest node on that path (n0 when eb meets one of
// n0's midlines).  No memory is read until that one entry.
//
// QTNode::collides from n0 visits this node with the same outcome.  Every hazard present below
// a node is an entry of it (qt_node.rs:39-81: hazards reach children by constrict, and an
// Entire hazard is Entire in every quadrant, qt_hazard.rs:58), and Entire entries come first
// (qt_hazard_vec.rs), so along the path: an ancestor without relevant hazards leaves this node
// none either, and an ancestor whose strongest relevant hazard is Entire leaves it Entire here
// too.  An edge strictly inside each quadrant of the path collides with no quadrant off the
// path at any ancestor (the box test of collides_with_quadrants finds those quadrants
// disjoint, qt_traits.rs:19-106), so the ancestors' visits reduce to this one.
__device__ __forceinline__ uint32_t start_node_below(const CdeView& c, uint32_t n0, uint32_t d0, float4 box0,
                                                     uint32_t x0, uint32_t y0, float4 eb) {
    float4 r = box0;
    uint32_t ix = x0, iy = y0, l = d0;
    for (; l < c.cell_level; l++) {
        const float mxm = (r.x + r.z) / 2.0f, mym = (r.y + r.w) / 2.0f;
        const bool xh = eb.x > mxm, xl = eb.z < mxm, yh = eb.y > mym, yl = eb.w < mym;
        if (!(xh || xl) || !(yh || yl)) break;
        r = make_float4(xh ? mxm : r.x, yh ? mym : r.y, xh ? r.z : mxm, yh ? r.w : mym);
        ix = 2u * ix + (xh ? 1u : 0u);
        iy = 2u * iy + (yh ? 1u : 0u);
    }
    if (l == d0) return n0;
    return ldg1(c.cells + ((1u << (2u * l)) - 1u) / 3u + ((iy << l) | ix)) & CELL_NODE;
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
            return circle_quadrants(cx, cy, r, nd.bbox);
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
            return edge_quadrants_node(ed, nd.bbox);
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
// Wave-cooperative traversal.  QTNode::collides (qt_node.rs:99-147) answers whether any
// relevant hazard of the subtree collides with the entity; its recursion order only decides
// which hazard is found first.  A wave therefore turns all descents of its entities (up to
// WL_ENT circles or edges, each belonging to one query slot) into two LDS queues: node visits
// (entity, node) and edge tests (entity, hazard edge).  Each step takes 64 items of one kind:
//   node step: the visit's strongest relevant hazard (qt_hazard_vec.rs:60-63) is Entire -> hit;
//              an internal node pushes the colliding quadrants (qt_traits.rs); a leaf pushes one
//              edge test per edge of every relevant Partial hazard whose ff_bbox the entity
//              collides with (qt_partial_hazard.rs:67-72);
//   edge step: one Circle-Edge or Edge-Edge test per lane (circle.rs:109-113, edge.rs:162-228).
// Edge steps run whenever 64 tests are pending, so both kinds of step keep every lane busy.
// A query's verdict is the OR of its entities' hits, exactly the reference's verdict.
//
// Kernels using it run one wavefront per workgroup: lanes exchange queue entries and hit flags
// through LDS only across wave_sync(), a workgroup barrier, which the compiler must reach with
// all lanes together; queue depths are wave-uniform values read from lane 63 of scans.

// Entity kinds of a traversal: KIND_EDGE and KIND_CIRCLE are fixed at compile time;
// KIND_MIXED reads the per-entity kind (surrogates with both poles and piers).
enum : int { KIND_EDGE = 0, KIND_CIRCLE = 1, KIND_MIXED = 2 };

constexpr uint32_t WL_ENT = 128;          // entities per traversal
constexpr uint32_t WL_NCAP = 512;         // node-visit queue entries per wave
constexpr uint32_t WL_ECAP = 640;         // edge-test queue entries per wave (LDS within 8 KB: 20 waves per CU)
constexpr int WL_EDGE_BITS = 10;          // edge tests one leaf visit pushes: fewer than 2^WL_EDGE_BITS
static_assert(64u * ((1u << WL_EDGE_BITS) - 1u) < 65536u, "a wave's edge pushes must fit 16 bits");
constexpr uint32_t WL_RESERVE = 3u * (JG_MAX_QT_DEPTH + 1u);   // node slots a depth-first walk may need
static_assert(WL_NCAP >= WL_ENT + WL_RESERVE + 3u, "node queue must hold the roots and the reserve");
constexpr uint32_t WL_NODE_MASK = 0xffffffu;
constexpr uint32_t WL_MAX_NODES = 1u << 24;
constexpr uint32_t WL_CIRCLE = 0x80u;     // WorkList::ent_slot flag: the entity is a circle
// An edge-test entry names up to 8 consecutive hazard edges: the first edge's index in its low
// 29 bits, the count less one in the top 3.
constexpr uint32_t WL_EDGE_INDEX = 0x1fffffffu;
constexpr uint32_t WL_MAX_EDGES = WL_EDGE_INDEX + 1u;

struct WorkList {                         // per-wave LDS
    uint32_t* nodes;                      // [WL_NCAP] entity << 24 | node
    uint32_t* edges;                      // [WL_ECAP] hazard edges (WL_EDGE_INDEX)
    uint8_t* edge_ent;                    // [WL_ECAP] entity of the edge test
    float4* geom;                         // [WL_ENT] circle (cx, cy, r, -) or edge (x1 y1 x2 y2)
    uint8_t* ent_slot;                    // [WL_ENT] query slot (0..63) | WL_CIRCLE
    uint32_t* hit;                        // [64] per query slot
    uint32_t* qidx;                       // [64] per query slot: query index (hashed filters)
};

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

// QTHazPartial::collides_with's ff_bbox pre-test (qt_partial_hazard.rs:67-72)
__device__ __forceinline__ bool ff_pass(const DevEntry& e, bool circle, float4 g) {
    return circle ? circle_box(g.x, g.y, g.z, e.ff) : box_edge(e.ff, g);
}

__device__ __forceinline__ bool edge_hit(bool circle, float4 g, float4 he) {
    return circle ? circle_edge(g.x, g.y, g.z, he) : edge_edge(g, he);
}

// QTHazPartial::collides_with over the relevant Partial entries of a leaf, from entry s on
// (qt_node.rs:126-141, qt_partial_hazard.rs:67-72).
template <bool FILT>
__device__ __forceinline__ bool leaf_hits(const CdeView& c, const DevNode& nd, int s, bool circle, float4 g,
                                          const FilterView& f, uint64_t q) {
    const int he = (int)(nd.haz_begin + (nd.haz_info & COUNT_MASK));
    for (int i = s; i < he; i++) {
        const DevEntry e = c.ent[i];
        if (irrel<FILT>(f, q, e.slot_flags & SLOT_MASK) || !ff_pass(e, circle, g)) continue;
        for (uint32_t j = 0; j < e.edge_count; j++)
            if (edge_hit(circle, g, c.edges[e.edge_begin + j])) return true;
    }
    return false;
}

// One visit of QTNode::collides: HIT, or the mask of children to visit (first child in fc).
template <bool FILT>
__device__ __forceinline__ uint32_t visit_node(const CdeView& c, int n, bool circle, float4 g, const FilterView& f,
                                               uint64_t q, int& fc) {
    const DevNode nd = c.nodes[n];
    bool entire = false;
    const int s = strongest_t<FILT>(c, nd, f, q, entire);
    if (s < 0) return 0u;
    if (entire) return HIT;
    if (nd.first_child >= 0) {
        fc = nd.first_child;
        if (circle) return circle_quadrants(g.x, g.y, g.z, nd.bbox);
        return edge_quadrants_node(g, nd.bbox);
    }
    return leaf_hits<FILT>(c, nd, s, circle, g, f, q) ? HIT : 0u;
}

// The workgroup is a single wavefront: a convergent barrier ordering its lanes' LDS accesses.
__device__ __forceinline__ void wave_sync() { __syncthreads(); }

// Number of set bits of m below this lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m, uint32_t lane) {
    return (uint32_t)__builtin_popcountll(m & ((1ull << lane) - 1ull));
}

// Exclusive prefix sum over the wave of per-lane counts x < 2^BITS, one ballot per bit;
// total (the sum) is wave-uniform.
template <int BITS>
__device__ __forceinline__ uint32_t wave_prefix(uint32_t x, uint32_t lane, uint32_t& total) {
    uint32_t off = 0, tot = 0;
    for (int b = 0; b < BITS; b++) {
        const uint64_t bb = __ballot((x >> b) & 1u);
        off += lanes_below(bb, lane) << b;
        tot += (uint32_t)__builtin_popcountll(bb) << b;
    }
    total = tot;
    return off;
}

// Inclusive prefix sum of x over the 64 lanes of a wave with every lane active: DPP row shifts
// within rows of 16 lanes, then the row broadcasts of lanes 15 and 31 (GFX9 DPP).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, uint32_t lane) {
    const uint32_t rl = lane & 15u;
    uint32_t v = x;
    uint32_t t = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x111, 0xf, 0xf, false);   // row_shr:1
    if (rl >= 1u) v += t;
    t = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x112, 0xf, 0xf, false);            // row_shr:2
    if (rl >= 2u) v += t;
    t = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x114, 0xf, 0xf, false);            // row_shr:4
    if (rl >= 4u) v += t;
    t = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x118, 0xf, 0xf, false);            // row_shr:8
    if (rl >= 8u) v += t;
    t = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x142, 0xf, 0xf, false);            // row_bcast:15
    if ((lane & 31u) >= 16u) v += t;
    t = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x143, 0xf, 0xf, false);            // row_bcast:31
    if (lane >= 32u) v += t;
    return v;
}

// Pushes one node visit per lane with want set onto a node queue of depth n (callers set
// geom/ent_slot first); returns the new depth, the same in every lane.
__device__ __forceinline__ uint32_t wave_push_roots(const WorkList& W, uint32_t n, bool want, uint32_t ent,
                                                    uint32_t node, uint32_t lane) {
    uint32_t total;
    const uint32_t off = wave_prefix<1>(want ? 1u : 0u, lane, total);
    if (want) W.nodes[n + off] = (ent << 24) | node;
    return n + total;
}

// Runs the wave's nn pending node visits (and the edge tests they create) to completion;
// W.hit[slot] is set for every query slot one of whose entities collides.  Callers push at
// most WL_ENT roots (WL_ENT <= WL_NCAP).  Each lane handles IPL items per step (node visits or
// edge tests): their memory reads are issued together, so a step waits for one round trip
// instead of IPL.  An edge-test item holds up to EG consecutive edges of one hazard entry: a
// leaf visit pushes one item per EG edges, so the lanes of a node step loop over a few pushes
// each instead of one per edge (a leaf of the benchmark layout holds about 7 edges a circle
// may meet), and an edge step tests up to EG edges per item.
template <bool FILT, int KIND, int IPL = 1, int EG = 1>
__device__ __forceinline__ void wave_traverse(const CdeView& c, const FilterView& f, const WorkList& W, uint32_t nn,
                                              uint32_t lane) {
    static_assert(IPL == 1 || IPL == 2, "one or two items per lane");
    static_assert(EG >= 1 && EG <= 8, "one to eight edges per edge-test item");
    // edge tests one item may push; the per-lane sum over IPL items stays below 2^16 / 64
    constexpr uint32_t EDGE_LIMIT = IPL == 1 ? (1u << WL_EDGE_BITS) : (1u << (WL_EDGE_BITS - 1));
    static_assert(64u * IPL * (EDGE_LIMIT - 1u) < 65536u, "a wave's edge pushes must fit 16 bits");
    constexpr uint32_t STEP = 64u * IPL;
    uint32_t ne = 0;
    while (nn + ne != 0u) {
        wave_sync();                              // pushes and hit flags of the last step are visible
        if (ne >= STEP || nn == 0u) {
            // edge step: both tests' edges are read before either is evaluated
            const uint32_t take = ne < STEP ? ne : STEP;
            uint32_t eo[IPL], cnt[IPL];
            float4 hedge[IPL][EG];
#pragma unroll
            for (int r = 0; r < IPL; r++) {
                const uint32_t idx = lane + 64u * (uint32_t)r;
                const uint32_t at = idx < take ? ne - 1u - idx : 0u;
                eo[r] = W.edge_ent[at];
                const uint32_t v = idx < take ? W.edges[at] : 0u;
                const uint32_t e0 = v & WL_EDGE_INDEX;
                cnt[r] = (v >> 29) + 1u;
#pragma unroll
                for (int k = 0; k < EG; k++) hedge[r][k] = ldg4(c.edges + e0 + ((uint32_t)k < cnt[r] ? (uint32_t)k : 0u));
            }
#pragma unroll
            for (int r = 0; r < IPL; r++) {
                if (lane + 64u * (uint32_t)r >= take) continue;
                const uint32_t es = W.ent_slot[eo[r]];
                const uint32_t slot = es & 63u;
                if (W.hit[slot]) continue;
                const bool circle = KIND == KIND_CIRCLE || (KIND == KIND_MIXED && (es & WL_CIRCLE) != 0u);
                const float4 gm = W.geom[eo[r]];
                bool h = false;
#pragma unroll
                for (int k = 0; k < EG; k++) h = h || ((uint32_t)k < cnt[r] && edge_hit(circle, gm, hedge[r][k]));
                if (h) W.hit[slot] = 1u;
            }
            ne -= take;
            continue;
        }
        // node step.  An item pushes at most 4 children, 3 more slots than it frees, so taking t
        // items grows the queue by at most 3t: steps take no more than keeps WL_RESERVE slots
        // free.  Within the reserve, single-item steps walk depth-first, which grows the queue by
        // at most 3 per level below where the walk began; the reserve covers the deepest tree.
        // The queue therefore never overflows and every step makes progress.
        uint32_t take = 1u;
        if (nn + 3u + WL_RESERVE <= WL_NCAP) {
            const uint32_t room = (WL_NCAP - WL_RESERVE - nn) / 3u;
            take = nn < STEP ? nn : STEP;
            take = take < room ? take : room;
        }
        const uint32_t base = nn - take;
        uint32_t item[IPL], m[IPL], n_edges[IPL], slot[IPL];
        uint32_t p_beg[IPL], p_cnt[IPL], n_pass[IPL];   // a leaf's first passing entry's edges; passing entries
                                                        // (kept with one item per lane)
        int fc[IPL], s[IPL], he[IPL];
        bool circle[IPL];
        float4 g[IPL];
        uint64_t q[IPL];
        DevNode nd[IPL];
#pragma unroll
        for (int r = 0; r < IPL; r++) {                // the node records, all in flight together
            const uint32_t idx = lane + 64u * (uint32_t)r;
            item[r] = idx < take ? W.nodes[nn - 1u - idx] : 0u;
            nd[r] = ldg_node(c.nodes + (item[r] & WL_NODE_MASK));
        }
        uint32_t packed = 0;
#pragma unroll
        for (int r = 0; r < IPL; r++) {
            const uint32_t o = item[r] >> 24;
            m[r] = 0;
            n_edges[r] = 0;
            if constexpr (IPL == 1) {
                p_beg[r] = 0;
                p_cnt[r] = 0;
                n_pass[r] = 0;
            }
            fc[r] = 0;
            s[r] = -1;
            he[r] = 0;
            slot[r] = 0;
            circle[r] = KIND == KIND_CIRCLE;
            g[r] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            q[r] = 0;
            if (lane + 64u * (uint32_t)r < take) {
                const uint32_t es = W.ent_slot[o];
                slot[r] = es & 63u;
                if (!W.hit[slot[r]]) {
                    g[r] = W.geom[o];
                    if (KIND == KIND_MIXED) circle[r] = (es & WL_CIRCLE) != 0u;
                    if (FILT) q[r] = W.qidx[slot[r]];
                    bool entire = false;
                    s[r] = strongest_t<FILT>(c, nd[r], f, q[r], entire);
                    if (s[r] >= 0 && entire) {
                        W.hit[slot[r]] = 1u;
                        s[r] = -1;
                    } else if (s[r] >= 0 && nd[r].first_child >= 0) {
                        fc[r] = nd[r].first_child;
                        m[r] = circle[r] ? circle_quadrants(g[r].x, g[r].y, g[r].z, nd[r].bbox)
                                         : edge_quadrants_node(g[r], nd[r].bbox);
                        s[r] = -1;
                    } else if (s[r] >= 0) {
                        // leaf: count the edge tests of its relevant Partial hazards
                        he[r] = (int)(nd[r].haz_begin + (nd[r].haz_info & COUNT_MASK));
                        if constexpr (IPL == 1) {
                            for (int i = s[r]; i < he[r]; i++) {
                                const DevEntry e = ldg_ent(c.ent + i);
                                if (!irrel<FILT>(f, q[r], e.slot_flags & SLOT_MASK) && ff_pass(e, circle[r], g[r])) {
                                    if (n_pass[r] == 0u) {
                                        p_beg[r] = e.edge_begin;
                                        p_cnt[r] = e.edge_count;
                                    }
                                    n_pass[r]++;
                                    n_edges[r] += (e.edge_count + (uint32_t)EG - 1u) / (uint32_t)EG;
                                }
                            }
                        } else {
                            for (int i = s[r]; i < he[r]; i++) {
                                const DevEntry e = ldg_ent(c.ent + i);
                                if (!irrel<FILT>(f, q[r], e.slot_flags & SLOT_MASK) && ff_pass(e, circle[r], g[r]))
                                    n_edges[r] += (e.edge_count + (uint32_t)EG - 1u) / (uint32_t)EG;
                            }
                        }
                    }
                }
            }
            if (n_edges[r] >= EDGE_LIMIT) {
                // more edge tests than one push carries: this lane tests the leaf's edges itself
                bool h = false;
                for (int i = s[r]; i < he[r] && !h; i++) {
                    const DevEntry e = ldg_ent(c.ent + i);
                    if (irrel<FILT>(f, q[r], e.slot_flags & SLOT_MASK) || !ff_pass(e, circle[r], g[r])) continue;
                    for (uint32_t j = 0; j < e.edge_count && !h; j++) h = edge_hit(circle[r], g[r], ldg4(c.edges + e.edge_begin + j));
                }
                if (h) W.hit[slot[r]] = 1u;
                n_edges[r] = 0;
            }
            // one scan for both pushes: child counts in the low 16 bits, edge counts in the high
            // 16; neither sum over the wave carries out of its half
            packed += (uint32_t)__builtin_popcount(m[r]) | (n_edges[r] << 16);
        }
        const uint32_t incl = wave_incl_scan(packed, lane);
        const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        const uint32_t excl = incl - packed;
        const uint32_t n_tot = tot & 0xffffu, e_tot = tot >> 16;
        wave_sync();                              // every lane has read its items before pushes overwrite
        {
            uint32_t at = base + (excl & 0xffffu);     // base + n_tot <= WL_NCAP by the choice of take
#pragma unroll
            for (int r = 0; r < IPL; r++)
                for (uint32_t b = m[r]; b; b &= b - 1u)
                    W.nodes[at++] = (item[r] & ~WL_NODE_MASK) | (uint32_t)(fc[r] + __builtin_ctz(b));
            nn = base + n_tot;
        }
        const bool fits = ne + e_tot <= WL_ECAP;
        uint32_t at = ne + (excl >> 16);
#pragma unroll
        for (int r = 0; r < IPL; r++) {
            if (!n_edges[r]) continue;
            if (IPL == 1 && fits && n_pass[r] == 1u) {
                // one passing entry (the common leaf): its edges from the count above (measured
                // faster with one item per lane; with two, the extra registers cost more)
                const uint32_t o = item[r] >> 24;
                for (uint32_t j = 0; j < p_cnt[r]; j += EG, at++) {
                    const uint32_t k = p_cnt[r] - j < (uint32_t)EG ? p_cnt[r] - j : (uint32_t)EG;
                    W.edges[at] = (p_beg[r] + j) | ((k - 1u) << 29);
                    W.edge_ent[at] = (uint8_t)o;
                }
            } else if (fits) {
                const uint32_t o = item[r] >> 24;
                for (int i = s[r]; i < he[r]; i++) {
                    const DevEntry e = ldg_ent(c.ent + i);
                    if (irrel<FILT>(f, q[r], e.slot_flags & SLOT_MASK) || !ff_pass(e, circle[r], g[r])) continue;
                    for (uint32_t j = 0; j < e.edge_count; j += EG, at++) {
                        const uint32_t k = e.edge_count - j < (uint32_t)EG ? e.edge_count - j : (uint32_t)EG;
                        W.edges[at] = (e.edge_begin + j) | ((k - 1u) << 29);
                        W.edge_ent[at] = (uint8_t)o;
                    }
                }
            } else {
                // edge queue full: this lane tests the leaf's edges itself
                bool h = false;
                for (int i = s[r]; i < he[r] && !h; i++) {
                    const DevEntry e = ldg_ent(c.ent + i);
                    if (irrel<FILT>(f, q[r], e.slot_flags & SLOT_MASK) || !ff_pass(e, circle[r], g[r])) continue;
                    for (uint32_t j = 0; j < e.edge_count && !h; j++) h = edge_hit(circle[r], g[r], ldg4(c.edges + e.edge_begin + j));
                }
                if (h) W.hit[slot[r]] = 1u;
            }
        }
        if (fits) ne += e_tot;
    }
    wave_sync();
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

// Sum of v over the 64 lanes (all active).
__device__ __forceinline__ int wave_sum_i32(int v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// pip_hazard / pip_query with the wave's 64 lanes splitting the polygon's edges (lane l takes
// edges l, l + 64, ...).  Each edge's crossing count is the sequential loop's, and their sum is
// an integer, so the parity is the same.  All 64 lanes call them with the same arguments.
__device__ __forceinline__ bool pip_hazard_wave(const CdeView& c, uint32_t slot, float px, float py, uint32_t lane) {
    const float4 hb = ldg4(c.haz_bbox + slot);
    if (!box_has_point(hb, px, py)) return false;
    const float ray_x = hb.z + (hb.z - hb.x);
    const uint2 vr = ldg2u(c.haz_vr + slot);
    const float2* v = c.haz_verts + vr.x;
    int cnt = 0;
    for (uint32_t i = lane; i < vr.y; i += 64u) {
        const float2 a = ldg2(v + i), b = ldg2(v + (i + 1 == vr.y ? 0 : i + 1));
        cnt += pip_edge(px, py, ray_x, a.x, a.y, b.x, b.y);
    }
    return (wave_sum_i32(cnt) & 1) == 1;
}
__device__ __forceinline__ bool pip_query_wave(const ItemView& it, const Mat& m, float4 qbb, float px, float py,
                                               uint32_t lane) {
    if (!box_has_point(qbb, px, py)) return false;
    const float ray_x = qbb.z + (qbb.z - qbb.x);
    int cnt = 0;
    for (uint32_t i = lane; i < it.n_v; i += 64u) {
        const uint32_t k = i + 1 == it.n_v ? 0 : i + 1;
        const float2 vi = ldg2(it.v + i), vk = ldg2(it.v + k);
        const float2 a = xform(m, vi.x, vi.y), b = xform(m, vk.x, vk.y);
        cnt += pip_edge(px, py, ray_x, a.x, a.y, b.x, b.y);
    }
    return (wave_sum_i32(cnt) & 1) == 1;
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
