// Batched hazard collection (jg_plan_collect): the hazard set of every query of a plan, as
// ci_bench composes it (lbf/benches/ci_bench.rs:56-69): collect_surrogate_collisions
// (cd_engine.rs:296-310) then collect_poly_collisions (cd_engine.rs:261-289), in persistent
// one-wave workgroups that claim the plan's work units.
//
// A wave collects CP queries of a unit at a time.  All their entities (ff poles and piers from
// the root, query edges from the first collecting node of their root path below the query's
// virtual root, collect_start_cells) share one wave traversal in the manner of the detection's
// (node visits and edge tests queued in LDS, 64 per step), with
// QTNode::collect_collisions (qt_node.rs:151-191) as the visit: a node whose active edges number
// at most cd_threshold, or a leaf, inserts its Entire hazards and queues edge tests for its
// Partial hazards; other nodes queue their colliding quadrants.  Each query's collector is a
// short list of slots in LDS with union semantics (collector.rs:44-67).
//
// Subtree skip.  A set is a union, so a visit that cannot add a slot can be left out.  Below a
// node, an entity collects a hazard only where it meets it: through an edge test, which fails
// unless the two edges' bounding boxes overlap (edge.rs:187-196), or at an Entire node of the
// hazard that the entity reaches (inside the hazard).  Every hazard present below a node is an
// entry of the node (constrict only narrows).  Hazard boxes are widened by 2^-14 of their
// coordinates' magnitude, far above the few ulps by which an Entire node or a quadrant flag may
// stray past a box in f32.  The exterior hazard's Entire nodes lie outside the container, so it
// can be ruled out only for an entity inside a box that is inside the container (the
// container's bbox when it is an axis-aligned rectangle), shrunk by the same margin.
//  - query edges: everything they can collect is an entry of the query's virtual root whose
//    widened box meets the query's bbox.  Those candidates (up to NC) are listed per query when
//    the pass starts, with each edge's mask of candidates its box meets; a visit is skipped
//    once every candidate of its edge's mask is in the set.  The skip reads only LDS.
//  - ff poles and piers descend from the root: a visit to a node of at most PRUNE_MAX entries
//    is skipped when no entry outside the set may be collected by the entity's box.
// The candidates also hold every hazard the containment loop can add (it needs a bbox
// relation, or the exterior), so that loop runs over them.  On the configuration-2 layout
// the skip takes the oracle's 525 node visits per query to about 196, and the start nodes to 78.
//
// A query whose set outgrows the LDS list (SET_CAP slots) is handed to the per-query collector
// (k_collect_overflow), which keeps a bit per slot.  A query with more than NC candidates skips
// no visit of its edges and scans all its virtual root's entries for containment.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/jagua_amd.h"
#include "jg_internal.hpp"
#include "jg_plan.hpp"

using namespace jg;

namespace {

constexpr uint32_t CL_WAVES = 6;         // waves per SIMD: VGPRs (80) and LDS per wave are sized for it
constexpr uint32_t CL_ENT = 64;          // entities per traversal (below 256: 8-bit indices)
constexpr uint32_t CL_NCAP = 320;        // node-visit queue entries
static_assert(CL_NCAP >= CL_ENT + WL_RESERVE + 3u, "node queue must hold the roots and the reserve");
constexpr uint32_t CP = 8;              // queries per collection pass
constexpr uint32_t CG = WAVE / CP;       // lanes per query in the set-up, containment and result phases
constexpr uint32_t SET_CAP = 16;         // slots per query collector in LDS
constexpr uint32_t NC = 8;              // candidates per query (bits of an edge's mask)
constexpr uint32_t CE_CAP = 128;         // edge-test queue entries
constexpr uint32_t CEG = 4;              // hazard edges per edge-test entry (first edge | (count - 1) << 29)
constexpr uint32_t CE_LIMIT = 128;       // edge tests one visit pushes: fewer (also below the scan's 16 bits)
static_assert(64u * (CE_LIMIT - 1u) < 65536u, "a wave's edge pushes must fit 16 bits");
constexpr bool INLINE_EDGES = false;   // test a visit's hazard edges in the visiting lane (no edge queue)
constexpr uint32_t PRUNE_MAX = 4;        // surrogate visits: nodes with at most this many entries are checked
constexpr uint32_t SURR_MASK = 0u;       // emask of a pole or pier: the node-entry check applies
constexpr uint32_t C_EXT = 0x80000000u;  // candidate record flags: exterior hazard, Entire in the virtual root
constexpr uint32_t C_ENTIRE = 0x40000000u;
constexpr uint32_t C_SLOT = 0x3fffffffu;
constexpr uint32_t VR_COLLECTS = 0x80000000u;   // vcell flag: the virtual root collects

struct CollectLds {
    uint32_t nodes[CL_NCAP];             // entity << 24 | node
    uint32_t edges[CE_CAP];              // hazard edges: first edge | (count - 1) << 29 (CEG)
    uint32_t eslot[CE_CAP];              // its hazard's slot
    float4 geom[CL_ENT];                 // circle (cx, cy, r, -) or edge (x1 y1 x2 y2)
    uint32_t emask[CL_ENT];              // query edge: its candidates; pole or pier: SURR_MASK
    uint32_t set[CP * SET_CAP];          // per query: slot + 1, filled from the front, 0 = free
    uint32_t cand[CP * NC];              // per query: candidate slot | C_EXT | C_ENTIRE
    float4 cbox[CP * NC];                // its widened box (the exterior: the inner box, shrunk)
    uint32_t ncand[CP];                  // candidates found (more than NC: none are used)
    uint32_t cdone[CP];                  // bit i: candidate i is in the set
    uint32_t ovf[CP];                    // per query: its set outgrew SET_CAP
    int32_t vroot[CP];
    uint32_t vdepth[CP];                 // its depth
    uint32_t vcell[CP];                  // its cell at its level, iy << 16 | ix (levels up to the tables'),
                                         // | VR_COLLECTS when it collects (a leaf, or at most cd_threshold active edges)
    float4 vbox[CP];                     // its box
    float mat[CP][6];                    // per query: its transformation
    float4 qbox[CP];                     // per query: its transformed bbox
    uint32_t vbeg[CP], vcnt[CP];         // its virtual root's entry range
    uint32_t vpre[CP + 1];               // prefix sums of vcnt over the pass
    uint8_t edge_ent[CE_CAP];
    uint8_t ent_slot[CL_ENT];            // query (0..CP-1) | WL_CIRCLE
};
static_assert(sizeof(CollectLds) * 4u * CL_WAVES <= 160u * 1024u, "a CU's waves share its 160 KB of LDS");

__device__ __forceinline__ bool set_has(const CollectLds& L, uint32_t q, uint32_t slot) {
    const uint32_t key = slot + 1u;
    const uint32_t* t = L.set + q * SET_CAP;
    for (uint32_t i = 0; i < SET_CAP; i++) {
        const uint32_t v = __atomic_load_n(&t[i], __ATOMIC_RELAXED);
        if (v == key) return true;
        if (v == 0u) return false;
    }
    return false;
}

// Inserts at the first free place, so a set stays a prefix; concurrent inserts of one slot keep
// one copy (the compare-and-swap sees the other's key).  A new slot that is one of the query's
// candidates marks it done.
__device__ __forceinline__ void set_put(CollectLds& L, uint32_t q, uint32_t slot) {
    const uint32_t key = slot + 1u;
    uint32_t* t = L.set + q * SET_CAP;
    for (uint32_t i = 0; i < SET_CAP; i++) {
        const uint32_t old = atomicCAS(&t[i], 0u, key);
        if (old == key) return;
        if (old == 0u) {
            const uint32_t nc = L.ncand[q];
            if (nc <= NC)
                for (uint32_t k = 0; k < nc; k++)
                    if ((L.cand[q * NC + k] & C_SLOT) == slot) atomicOr(&L.cdone[q], 1u << k);
            return;
        }
    }
    atomicOr(&L.ovf[q], 1u);
}

// The widening of a hazard box: 2^-14 of its coordinates' magnitude plus its size.
__device__ __forceinline__ float widen(float4 b) {
    const float mag = fmaxf(fmaxf(fabsf(b.x), fabsf(b.z)), fmaxf(fabsf(b.y), fabsf(b.w)));
    return (mag + (b.z - b.x) + (b.w - b.y)) * 0x1p-14f;
}
__device__ __forceinline__ float4 widened(float4 b) {
    const float w = widen(b);
    return make_float4(b.x - w, b.y - w, b.z + w, b.w + w);
}
// The container's inner box shrunk by the margin (empty when there is none).
__device__ __forceinline__ float4 inner_shrunk(const CdeView& c) {
    const float4 b = c.ext_inner;
    if (!(b.x < b.z)) return b;
    const float w = widen(b);
    return make_float4(b.x + w, b.y + w, b.z - w, b.w - w);
}
__device__ __forceinline__ bool boxes_meet(float4 a, float4 b) {
    return a.x <= b.z && a.z >= b.x && a.y <= b.w && a.w >= b.y;
}
__device__ __forceinline__ bool box_inside(float4 in, float4 out) {
    return in.x > out.x && in.z < out.z && in.y > out.y && in.w < out.w;
}
// Whether an entity with box eb may collect hazard `slot` anywhere below a node that holds it.
__device__ __forceinline__ bool may_collect(const CdeView& c, uint32_t slot, float4 eb) {
    if (slot == c.exterior_slot) return !box_inside(eb, inner_shrunk(c));
    return boxes_meet(eb, widened(ldg4(c.haz_bbox + slot)));
}

// The node from which a query edge with box eb, starting at node n, reaches the same visits
// of QTNode::collect_collisions (qt_node.rs:151-191): while n is an internal node of the
// collection (it has children and more than cd_threshold active edges) and eb lies strictly
// inside one of its quadrants, that quadrant's child.  An edge strictly inside a quadrant
// collides with no other quadrant (the box test of collides_with_quadrants,
// qt_traits.rs:19-106), so the visit of n would push just that child.
__device__ __forceinline__ uint32_t collect_start(const CdeView& c, uint32_t n, float4 eb) {
    for (;;) {
        const DevNode nd = ldg_node(c.nodes + n);
        if (nd.first_child < 0 || nd.n_active_edges <= c.cd_threshold) return n;
        const float4 r = nd.bbox;
        const float mxm = (r.x + r.z) / 2.0f, mym = (r.y + r.w) / 2.0f;
        const bool xh = eb.x > mxm, xl = eb.z < mxm, yh = eb.y > mym, yl = eb.w < mym;
        if (!(xh || xl) || !(yh || yl)) return n;
        // children in quadrant order: top right, top left, bottom left, bottom right
        n = (uint32_t)nd.first_child + (xh ? (yh ? 0u : 3u) : (yh ? 1u : 2u));
    }
}

// collect_start for a query edge with box eb, of a query whose virtual root vr (depth vd, box
// vbox, cell (vx, vy) of its level) does not collect, through the cell tables: below vr the
// edge's root path is followed arithmetically while eb lies strictly inside one quadrant (no
// memory read), down to the tables' deepest level; that cell's entry gives the deepest node on
// the path and the depth of the first collecting node on it (below vr, as vr does not
// collect), and at most one more entry gives that node.  Without one, the start is the
// deepest node, from which collect_start goes on below the tables' levels.  The path's boxes
// below vr hold the edge: vr's box does, and eb is strictly on one side of each midline.
__device__ __forceinline__ uint32_t collect_start_cells(const CdeView& c, float4 eb, uint32_t vr, uint32_t vd,
                                                        float4 vbox, uint32_t vx, uint32_t vy) {
    if (vd >= c.cell_level) return collect_start(c, vr, eb);
    float4 r = vbox;
    uint32_t ix = vx, iy = vy, l = vd;
    for (; l < c.cell_level; l++) {
        const float mxm = (r.x + r.z) / 2.0f, mym = (r.y + r.w) / 2.0f;
        const bool xh = eb.x > mxm, xl = eb.z < mxm, yh = eb.y > mym, yl = eb.w < mym;
        if (!(xh || xl) || !(yh || yl)) break;
        r = make_float4(xh ? mxm : r.x, yh ? mym : r.y, xh ? r.z : mxm, yh ? r.w : mym);
        ix = 2u * ix + (xh ? 1u : 0u);
        iy = 2u * iy + (yh ? 1u : 0u);
    }
    if (l == vd) return vr;                       // eb meets a midline of vr: vr's visit starts it
    const uint32_t e = ldg1(c.cells + ((1u << (2u * l)) - 1u) / 3u + ((iy << l) | ix));
    const uint32_t cd = cell_cdepth(e), n = e & CELL_NODE;
    if (cd == 15u) return l == c.cell_level ? collect_start(c, n, eb) : n;
    if (cd == cell_depth(e)) return n;
    const uint32_t s = l - cd;                    // the first collecting node, above the cell's node
    return ldg1(c.cells + ((1u << (2u * cd)) - 1u) / 3u + (((iy >> s) << cd) | (ix >> s))) & CELL_NODE;
}

__device__ __forceinline__ float4 entity_box(bool circle, float4 g) {
    return circle ? make_float4(g.x - g.z, g.y - g.z, g.x + g.z, g.y + g.z)
                  : make_float4(mn(g.x, g.z), mn(g.y, g.w), mx(g.x, g.z), mx(g.y, g.w));
}

// The Partial entries of a detection-now node whose edges an entity tests (qt_node.rs:174-189:
// the ff_bbox pre-test of qt_partial_hazard.rs:67-72).  The count and the pushes use this same
// predicate, so they agree whatever the sets do meanwhile.
__device__ __forceinline__ bool tests_edges(const DevEntry& e, bool circle, float4 g) {
    return !(e.slot_flags & ENTIRE_BIT) && ff_pass(e, circle, g);
}

// This lane tests an entry's edges itself and inserts its slot on a hit (queue overflow).
__device__ __forceinline__ void test_inline(const CdeView& c, CollectLds& L, const DevNode& nd, uint32_t qs,
                                            bool circle, float4 g) {
    const uint32_t he = nd.haz_begin + (nd.haz_info & COUNT_MASK);
    for (uint32_t i = nd.haz_begin; i < he; i++) {
        const DevEntry e = ldg_ent(c.ent + i);
        if (!tests_edges(e, circle, g)) continue;
        const uint32_t sl = e.slot_flags & SLOT_MASK;
        if (set_has(L, qs, sl)) continue;
        for (uint32_t j = 0; j < e.edge_count; j++)
            if (edge_hit(circle, g, ldg4(c.edges + e.edge_begin + j))) {
                set_put(L, qs, sl);
                break;
            }
    }
}

// Whether a visit of a pole or pier to node nd may still add to its query's set: some entry
// outside the set may be collected by the entity's box (entries read four at a time).
__device__ __forceinline__ bool surrogate_live(const CdeView& c, const CollectLds& L, const DevNode& nd, uint32_t cnt,
                                               uint32_t qs, float4 eb) {
    for (uint32_t i0 = 0; i0 < cnt; i0 += 4) {
        uint32_t sl[4];
#pragma unroll
        for (uint32_t r = 0; r < 4; r++) sl[r] = ldg_slot_flags(c.ent + nd.haz_begin + (i0 + r < cnt ? i0 + r : 0u)) & SLOT_MASK;
        float4 hb[4];
#pragma unroll
        for (uint32_t r = 0; r < 4; r++) hb[r] = ldg4(c.haz_bbox + sl[r]);
#pragma unroll
        for (uint32_t r = 0; r < 4; r++) {
            if (i0 + r >= cnt || set_has(L, qs, sl[r])) continue;
            if (sl[r] == c.exterior_slot ? !box_inside(eb, inner_shrunk(c)) : boxes_meet(eb, widened(hb[r]))) return true;
        }
    }
    return false;
}

// Runs the wave's nn pending node visits of QTNode::collect_collisions and the edge tests they
// create to completion, inserting into the queries' sets.  Same queue discipline as the
// detection's traversal (the node queue cannot overflow; an edge push that would overflow is
// tested inline).
__device__ __forceinline__ void wave_collect(const CdeView& c, CollectLds& L, uint32_t nn, uint32_t lane) {
    uint32_t ne = 0;
    while (nn + ne != 0u) {
        wave_sync();
        if (ne >= 64u || nn == 0u) {
            const uint32_t take = ne < 64u ? ne : 64u;
            if (lane < take) {
                const uint32_t at = ne - 1u - lane;
                const uint32_t o = L.edge_ent[at], sl = L.eslot[at];
                const uint32_t es = L.ent_slot[o], qs = es & 63u;
                const uint32_t v = L.edges[at], e0 = v & WL_EDGE_INDEX, k_n = (v >> 29) + 1u;
                float4 he[CEG];
#pragma unroll
                for (uint32_t k = 0; k < CEG; k++) he[k] = ldg4(c.edges + e0 + (k < k_n ? k : 0u));
                if (!L.ovf[qs] && !set_has(L, qs, sl)) {
                    const bool circle = (es & WL_CIRCLE) != 0u;
                    const float4 g = L.geom[o];
                    bool h = false;
#pragma unroll
                    for (uint32_t k = 0; k < CEG; k++) h = h || (k < k_n && edge_hit(circle, g, he[k]));
                    if (h) set_put(L, qs, sl);
                }
            }
            ne -= take;
            continue;
        }
        uint32_t take = 1u;
        if (nn + 3u + WL_RESERVE <= CL_NCAP) {
            const uint32_t room = (CL_NCAP - WL_RESERVE - nn) / 3u;
            take = nn < 64u ? nn : 64u;
            take = take < room ? take : room;
        }
        const uint32_t base = nn - take;
        const uint32_t item = lane < take ? L.nodes[nn - 1u - lane] : 0u;
        const uint32_t o = item >> 24;
        uint32_t qs = 0, em = 0;
        bool live = false;
        if (lane < take) {
            qs = L.ent_slot[o] & 63u;
            em = L.emask[o];
            // a query edge whose candidates are all collected adds nothing below any node
            live = !L.ovf[qs] && (em == SURR_MASK || (em & ~L.cdone[qs]) != 0u);
        }
        const DevNode nd = ldg_node(c.nodes + (live ? (item & WL_NODE_MASK) : 0u));
        uint32_t m = 0, n_edges = 0;
        uint32_t n_pass = 0, p_beg = 0, p_cnt = 0, p_slot = 0;   // passing Partial entries; the first one's edges
        int fc = 0;
        bool circle = false;
        float4 g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const uint32_t cnt = nd.haz_info & COUNT_MASK;
        if (live) {
            const uint32_t es = L.ent_slot[o];
            circle = (es & WL_CIRCLE) != 0u;
            g = L.geom[o];
            live = cnt > 0u;
            if (live && em == SURR_MASK && cnt <= PRUNE_MAX) live = surrogate_live(c, L, nd, cnt, qs, entity_box(circle, g));
            if (live) {
                if (nd.first_child >= 0 && nd.n_active_edges > c.cd_threshold) {
                    fc = nd.first_child;
                    m = circle ? circle_quadrants(g.x, g.y, g.z, nd.bbox) : edge_quadrants_node(g, nd.bbox);
                } else if (INLINE_EDGES) {
                    // the visit inserts its Entire hazards and tests its Partial hazards' edges itself
                    for (uint32_t i = nd.haz_begin; i < nd.haz_begin + cnt; i++) {
                        const uint32_t sf = ldg_slot_flags(c.ent + i);
                        const uint32_t sl = sf & SLOT_MASK;
                        if ((sf & ENTIRE_BIT) && !set_has(L, qs, sl)) set_put(L, qs, sl);
                    }
                    test_inline(c, L, nd, qs, circle, g);
                } else {
                    for (uint32_t i = nd.haz_begin; i < nd.haz_begin + cnt; i++) {
                        const DevEntry e = ldg_ent(c.ent + i);
                        const uint32_t sl = e.slot_flags & SLOT_MASK;
                        if (e.slot_flags & ENTIRE_BIT) {
                            if (!set_has(L, qs, sl)) set_put(L, qs, sl);
                        } else if (ff_pass(e, circle, g)) {
                            if (n_pass == 0u) {
                                p_beg = e.edge_begin;
                                p_cnt = e.edge_count;
                                p_slot = sl;
                            }
                            n_pass++;
                            n_edges += (e.edge_count + CEG - 1u) / CEG;
                        }
                    }
                    if (n_edges >= CE_LIMIT) {
                        test_inline(c, L, nd, qs, circle, g);
                        n_edges = 0;
                    }
                }
            }
        }
        const uint32_t packed = (uint32_t)__builtin_popcount(m) | (n_edges << 16);
        const uint32_t incl = wave_incl_scan(packed, lane);
        const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        const uint32_t excl = incl - packed;
        wave_sync();                              // every lane has read its item before pushes overwrite
        {
            uint32_t at = base + (excl & 0xffffu);
            for (uint32_t b = m; b; b &= b - 1u) L.nodes[at++] = (item & ~WL_NODE_MASK) | (uint32_t)(fc + __builtin_ctz(b));
            nn = base + (tot & 0xffffu);
        }
        const uint32_t e_tot = tot >> 16;
        const bool fits = ne + e_tot <= CE_CAP;
        if (n_edges) {
            if (fits && n_pass == 1u) {
                // one passing entry (the common visit): its edges from the count above
                uint32_t at = ne + (excl >> 16);
                for (uint32_t j = 0; j < p_cnt; j += CEG, at++) {
                    const uint32_t k = p_cnt - j < CEG ? p_cnt - j : CEG;
                    L.edges[at] = (p_beg + j) | ((k - 1u) << 29);
                    L.eslot[at] = p_slot;
                    L.edge_ent[at] = (uint8_t)o;
                }
            } else if (fits) {
                uint32_t at = ne + (excl >> 16);
                for (uint32_t i = nd.haz_begin; i < nd.haz_begin + cnt; i++) {
                    const DevEntry e = ldg_ent(c.ent + i);
                    if (!tests_edges(e, circle, g)) continue;
                    const uint32_t sl = e.slot_flags & SLOT_MASK;
                    for (uint32_t j = 0; j < e.edge_count; j += CEG, at++) {
                        const uint32_t k = e.edge_count - j < CEG ? e.edge_count - j : CEG;
                        L.edges[at] = (e.edge_begin + j) | ((k - 1u) << 29);
                        L.eslot[at] = sl;
                        L.edge_ent[at] = (uint8_t)o;
                    }
                }
            } else {
                test_inline(c, L, nd, qs, circle, g);
            }
        }
        if (fits) ne += e_tot;
    }
    wave_sync();
}

__device__ __forceinline__ uint64_t group_sum_u64(uint64_t v) {
    for (uint32_t o = CG >> 1; o > 0; o >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, (int)o, 64);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), (int)o, 64);
        v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}

// detect_containment_collision (cd_engine.rs:231-255) for one candidate of the query, by the
// query's CG lanes (the one holding it in `pend`): inserts it when it collides.
__device__ __forceinline__ void containment_group(const CdeView& c, CollectLds& L, const ItemView& it, const Mat& m,
                                                  float4 bb, float2 qpoi, uint32_t qs, uint32_t sub, uint32_t lane,
                                                  bool pend, uint32_t sl, int rel) {
    const uint64_t gmask = (1ull << CG) - 1ull;
    const uint32_t gbase = qs * CG;
    for (uint64_t pm = (ballot(pend) >> gbase) & gmask; pm; pm &= pm - 1ull) {
        const int l = (int)gbase + __builtin_ctzll(pm);
        const uint32_t s2 = (uint32_t)__shfl((int)sl, l, 64);
        const int rl = __shfl(rel, l, 64);
        bool contained;
        if (rl == REL_SURROUNDING) {
            contained = pip_hazard_group(c, s2, qpoi.x, qpoi.y, sub, CG);
        } else {
            const float4 hp = ldg4(c.haz_poi + s2);
            contained = pip_query_group(it, m, bb, hp.x, hp.y, sub, CG);
        }
        if (lane == (uint32_t)l && (ldg4(c.haz_poi + s2).w != 0.0f ? contained : !contained)) set_put(L, qs, s2);
    }
}

// The hazard sets of queries q0 .. q0 + P - 1 of set sd (P <= CP).
__device__ __forceinline__ void collect_pass(const SetDesc& sd, uint32_t set_index, uint32_t q0, uint32_t P,
                                             CollectLds& L, const float* __restrict__ m6, uint32_t cap,
                                             uint32_t* __restrict__ count, uint64_t* __restrict__ hash,
                                             uint32_t* __restrict__ keys, uint32_t* __restrict__ counters,
                                             uint64_t* __restrict__ ovf, uint32_t lane) {
    const CdeView c = sd.cde;
    const ItemView it = sd.item;
    const uint32_t qs = lane / CG, sub = lane % CG;
    const bool active = qs < P;
    const uint64_t out = sd.first + q0 + qs;
    Mat m{};
    if (active) m = load_mat(m6, out);
    // the transformed bbox (simple_polygon.rs:122-133), reduced over the query's CG lanes
    float4 bb = make_float4(3.40282347e+38f, 3.40282347e+38f, -3.40282347e+38f, -3.40282347e+38f);
    if (active)
        for (uint32_t i = sub; i < it.n_v; i += CG) {
            const float2 vi = ldg2(it.v + i);
            const float2 p = xform(m, vi.x, vi.y);
            bb.x = mn(bb.x, p.x); bb.y = mn(bb.y, p.y); bb.z = mx(bb.z, p.x); bb.w = mx(bb.w, p.y);
        }
    for (uint32_t o = CG >> 1; o > 0; o >>= 1) {
        bb.x = fminf(bb.x, __shfl_xor(bb.x, (int)o, 64));
        bb.y = fminf(bb.y, __shfl_xor(bb.y, (int)o, 64));
        bb.z = fmaxf(bb.z, __shfl_xor(bb.z, (int)o, 64));
        bb.w = fmaxf(bb.w, __shfl_xor(bb.w, (int)o, 64));
    }
    for (uint32_t i = sub; i < SET_CAP; i += CG) L.set[qs * SET_CAP + i] = 0u;
    if (sub == 0) {
        L.ovf[qs] = 0u;
        L.ncand[qs] = 0u;
        L.cdone[qs] = 0u;
    }
    int vr = 0;
    uint32_t vcnt = 0, ve0 = 0, vdepth = 0, vcell = 0, vcoll = 1;
    float4 vbox = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 root_bb = ldg4((const float4*)c.nodes);
    if (active) {
        vr = virtual_root_cells(c, root_bb, bb, vdepth);
        const DevNode vn = ldg_node(c.nodes + vr);
        ve0 = vn.haz_begin;
        vcnt = vn.haz_info & COUNT_MASK;
        vbox = vn.bbox;
        vcoll = (vn.first_child < 0 || vn.n_active_edges <= c.cd_threshold) ? 1u : 0u;
        uint32_t ix = 0, iy = 0;
        if (vdepth < c.cell_level) node_cell(root_bb, vbox, vdepth, ix, iy);   // (else unused)
        vcell = (iy << 16) | ix;
    }
    if (sub == 0) {
        L.vroot[qs] = vr;
        L.vdepth[qs] = vdepth;
        L.vcell[qs] = vcell | (vcoll ? VR_COLLECTS : 0u);
        L.vbox[qs] = vbox;
        L.mat[qs][0] = m.a; L.mat[qs][1] = m.b; L.mat[qs][2] = m.c;
        L.mat[qs][3] = m.d; L.mat[qs][4] = m.e; L.mat[qs][5] = m.f;
        L.qbox[qs] = bb;
        L.vbeg[qs] = ve0;
        L.vcnt[qs] = vcnt;
    }
    wave_sync();
    // the candidates: the virtual root's entries whose widened box meets the query's bbox, and
    // the exterior unless the bbox is inside the inner box.  The pass's (query, entry) pairs are
    // scanned by the whole wave, four per lane in flight, so that one query of a shallow virtual
    // root (hundreds of entries) does not hold the other queries' lanes
    if (lane <= CP) {
        uint32_t pre = 0;
        for (uint32_t j = 0; j < lane; j++) pre += L.vcnt[j];
        L.vpre[lane] = pre;
    }
    wave_sync();
    {
        const uint32_t T = L.vpre[CP];
        const float4 inner = inner_shrunk(c);
        for (uint32_t p0 = 0; p0 < T; p0 += 4u * WAVE) {
            uint32_t sf[4], qq[4];
#pragma unroll
            for (uint32_t r = 0; r < 4; r++) {
                const uint32_t p = p0 + r * WAVE + lane;
                uint32_t lo = 0;                  // the query of pair p: last lo with vpre[lo] <= p
                for (uint32_t step = CP / 2; step > 0; step >>= 1)
                    if (L.vpre[lo + step] <= p) lo += step;
                qq[r] = lo;
                sf[r] = p < T ? ldg_slot_flags(c.ent + L.vbeg[lo] + (p - L.vpre[lo])) : 0u;
            }
            float4 hb[4];
#pragma unroll
            for (uint32_t r = 0; r < 4; r++) hb[r] = ldg4(c.haz_bbox + (sf[r] & SLOT_MASK));
#pragma unroll
            for (uint32_t r = 0; r < 4; r++) {
                if (p0 + r * WAVE + lane >= T) continue;
                const uint32_t q = qq[r];
                const float4 qb = L.qbox[q];
                const uint32_t sl = sf[r] & SLOT_MASK;
                const bool ext = sl == c.exterior_slot;
                const float4 wb = ext ? inner : widened(hb[r]);
                if (ext ? box_inside(qb, inner) : !boxes_meet(qb, wb)) continue;
                const uint32_t at = atomicAdd(&L.ncand[q], 1u);
                if (at < NC) {
                    L.cand[q * NC + at] = sl | (ext ? C_EXT : 0u) | ((sf[r] & ENTIRE_BIT) ? C_ENTIRE : 0u);
                    L.cbox[q * NC + at] = wb;
                }
            }
        }
    }
    wave_sync();
    if (active) {
        // a pre-seeded collector, then the exterior when the shape leaves the engine's bbox
        // (cd_engine.rs:262-264); the traversal goes on regardless
        for (uint32_t i = sub; i < sd.n_seed; i += CG) set_put(L, qs, ldg1(sd.seed + i));
        if (sub == 0 && box_relation(ldg4((const float4*)c.nodes), bb) != REL_SURROUNDING) set_put(L, qs, c.exterior_slot);
    }
    // every query's ff poles / piers from the root and edges from its virtual root.  A cursor
    // walks the entities 64 at a time; those that may add to their query's set take the next
    // free places of the traversal (CL_ENT), which runs when they are full or the entities end
    const uint32_t n_ff = it.n_poles + it.n_piers;
    const uint32_t n_ent = n_ff + it.n_v;
    const uint32_t total = P * n_ent;
    const float inv_ent = 1.0f / (float)n_ent;
    uint32_t cursor = 0;
    while (cursor < total) {
        wave_sync();
        uint32_t n = 0;
        while (cursor < total && n < CL_ENT) {
            const uint32_t e = cursor + lane;
            bool want = false;
            uint32_t root = 0, em = SURR_MASK, qi = 0;
            bool circle = false;
            float4 gm = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (e < total) {
                // e / n_ent through the float reciprocal: the quotient is below CP, so the
                // estimate is off by at most one, corrected here
                qi = (uint32_t)((float)e * inv_ent);
                qi = qi * n_ent > e ? qi - 1u : ((qi + 1u) * n_ent <= e ? qi + 1u : qi);
                const uint32_t k = e - qi * n_ent;
                if (!L.ovf[qi]) {
                    const Mat mq{L.mat[qi][0], L.mat[qi][1], L.mat[qi][2], L.mat[qi][3], L.mat[qi][4], L.mat[qi][5]};
                    if (k < n_ff) {
                        const float4 p = ldg4(it.ff + k);
                        circle = k < it.n_poles;
                        if (circle) {
                            const float2 cc = xform(mq, p.x, p.y);
                            gm = make_float4(cc.x, cc.y, p.z, 0.0f);
                        } else {
                            const float2 a = xform(mq, p.x, p.y), b = xform(mq, p.z, p.w);
                            gm = make_float4(a.x, a.y, b.x, b.y);
                        }
                        want = true;
                    } else {
                        const uint32_t i = k - n_ff, i1 = i + 1 == it.n_v ? 0 : i + 1;
                        const float2 va = ldg2(it.v + i), vb = ldg2(it.v + i1);
                        const float2 a = xform(mq, va.x, va.y), b = xform(mq, vb.x, vb.y);
                        gm = make_float4(a.x, a.y, b.x, b.y);
                        root = (uint32_t)L.vroot[qi];
                        // the candidates this edge's box may collect (all of them when the
                        // query has more than NC); an edge whose candidates are all in the
                        // set already adds nothing
                        const uint32_t nc = L.ncand[qi];
                        if (nc > NC) {
                            em = 0xffffffffu;
                        } else {
                            const float4 eb = entity_box(false, gm);
                            em = 0u;
                            for (uint32_t t = 0; t < nc; t++) {
                                const uint32_t cf = L.cand[qi * NC + t];
                                const float4 cb = L.cbox[qi * NC + t];
                                const bool may = (cf & C_EXT) ? !box_inside(eb, cb) : boxes_meet(eb, cb);
                                em |= may ? (1u << t) : 0u;
                            }
                        }
                        want = (em & ~L.cdone[qi]) != 0u;
                        const uint32_t vc = L.vcell[qi];
                        if (want && !(vc & VR_COLLECTS))
                            root = collect_start_cells(c, entity_box(false, gm), root, L.vdepth[qi], L.vbox[qi],
                                                       vc & 0xffffu, (vc >> 16) & 0x7fffu);
                    }
                }
            }
            // the first CL_ENT - n wanted entities of these 64 take places n, n + 1, ...
            const uint64_t wb = __ballot(want);
            const uint32_t rank = lanes_below(wb, lane);
            const uint32_t room = CL_ENT - n;
            const uint32_t n_want = (uint32_t)__builtin_popcountll(wb);
            if (want && rank < room) {
                const uint32_t o = n + rank;
                L.geom[o] = gm;
                L.emask[o] = em;
                L.ent_slot[o] = (uint8_t)(qi | (circle ? WL_CIRCLE : 0u));
                L.nodes[o] = (o << 24) | root;
            }
            if (n_want <= room) {
                n += n_want;
                cursor += WAVE;
            } else {
                // resume at the first wanted entity that found no place
                uint64_t r = wb;
                for (uint32_t t = 0; t < room; t++) r &= r - 1ull;
                cursor += (uint32_t)__builtin_ctzll(r);
                n = CL_ENT;
            }
        }
        wave_collect(c, L, n, lane);
    }
    // containment over the virtual root's Partial hazards not yet collected
    // (cd_engine.rs:276-287), by the query's CG lanes: a candidate without bbox relation
    // collides iff its scope is exterior; the point-in-polygon tests left are split over the
    // CG lanes (containment()'s result).  Only candidates can collide; with more than NC the
    // virtual root's entries are scanned.
    if (active && !L.ovf[qs]) {
        // the query's transformation, bbox and virtual root entries from LDS, so that no
        // registers hold them over the traversal
        const Mat m{L.mat[qs][0], L.mat[qs][1], L.mat[qs][2], L.mat[qs][3], L.mat[qs][4], L.mat[qs][5]};
        const float4 bb = L.qbox[qs];
        const uint32_t vcnt = L.vcnt[qs], ve0 = L.vbeg[qs];
        const float2 qpoi = xform(m, it.poi.x, it.poi.y);
        const uint32_t nc = L.ncand[qs];
        const bool listed = nc <= NC;
        const uint32_t n_scan = listed ? nc : vcnt;
        for (uint32_t k0 = 0; k0 < n_scan; k0 += CG) {
            const uint32_t k = k0 + sub;
            uint32_t sl = 0;
            int rel = REL_DISJOINT;
            bool pend = false;
            if (k < n_scan) {
                uint32_t sf;
                if (listed) {
                    const uint32_t cf = L.cand[qs * NC + k];
                    sf = (cf & C_SLOT) | ((cf & C_ENTIRE) ? ENTIRE_BIT : 0u);
                } else {
                    sf = ldg_slot_flags(c.ent + ve0 + k);
                }
                sl = sf & SLOT_MASK;
                if (!(sf & ENTIRE_BIT) && !set_has(L, qs, sl)) {
                    rel = box_almost_relation(ldg4(c.haz_bbox + sl), bb);
                    if (rel == REL_SURROUNDING || rel == REL_ENCLOSED) pend = true;
                    else if (ldg4(c.haz_poi + sl).w == 0.0f) set_put(L, qs, sl);
                }
            }
            containment_group(c, L, it, m, bb, qpoi, qs, sub, lane, pend, sl, rel);
        }
    }
    wave_sync();
    // results: count, order-independent hash, the first `cap` slots ascending
    if (active) {
        const uint32_t* t = L.set + qs * SET_CAP;
        if (L.ovf[qs]) {
            if (sub == 0) {
                const uint32_t at = atomicAdd(&counters[C_OVERFLOW], 1u);
                ovf[at] = ((uint64_t)set_index << 32) | (uint32_t)(q0 + qs);
            }
        } else {
            uint32_t n = 0;
            while (n < SET_CAP && t[n] != 0u) n++;
            uint64_t h = 0;
            for (uint32_t i = sub; i < n; i += CG) {
                const uint32_t si = t[i] - 1u;
                h += mix64((uint64_t)si);
                if (keys && cap) {
                    uint32_t rank = 0;
                    for (uint32_t j = 0; j < n; j++) rank += t[j] - 1u < si ? 1u : 0u;
                    if (rank < cap) keys[out * cap + rank] = si;
                }
            }
            if (keys)
                for (uint32_t k = n + sub; k < cap; k += CG) keys[out * cap + k] = 0xffffffffu;
            h = group_sum_u64(h);
            if (sub == 0) {
                count[out] = n;
                hash[out] = h;
            }
        }
    }
    wave_sync();
}

__global__ void k_collect_reset(uint32_t* counters) {
    for (uint32_t i = threadIdx.x; i < N_COUNTER_WORDS; i += blockDim.x) counters[i] = 0u;
}

// CL_WAVES waves per SIMD (the VGPR limit that sets), 4 CL_WAVES per CU.
__global__ __launch_bounds__(WAVE, CL_WAVES) void k_collect_units(const SetDesc* __restrict__ sets, const Unit* __restrict__ units,
                                                       uint32_t n_units, uint32_t* __restrict__ counters,
                                                       const float* __restrict__ m6, uint32_t cap,
                                                       uint32_t* __restrict__ count, uint64_t* __restrict__ hash,
                                                       uint32_t* __restrict__ keys, uint64_t* __restrict__ ovf) {
    __shared__ CollectLds lds;
    const uint32_t lane = threadIdx.x;
    uint32_t ux = blockIdx.x & (N_STREAMS - 1u), utried = 0;
    for (;;) {
        const uint32_t u = claim_next(counters, C_USTREAM0, ux, utried, n_units, lane);
        if (u >= n_units) return;
        const Unit un = units[u];
        const uint32_t si = wave_bcast(un.set);
        for (uint32_t p0 = 0; p0 < un.count; p0 += CP) {
            const uint32_t P = un.count - p0 < CP ? un.count - p0 : CP;
            collect_pass(sets[si], si, un.q0 + p0, P, lds, m6, cap, count, hash, keys, counters, ovf, lane);
        }
    }
}

}  // namespace

namespace jg {

jg_status collect_launch(const TableLayout& L, uint8_t* base, const float* d_m6, uint32_t cap, uint32_t* d_count,
                         uint64_t* d_hash, uint32_t* d_keys, uint64_t* ovf, uint32_t n_cus, hipStream_t stream) {
    uint32_t* counters = (uint32_t*)(base + L.o_ctr);
    const SetDesc* sets = (const SetDesc*)base;
    const Unit* units = (const Unit*)(base + L.o_units);
    hipLaunchKernelGGL(k_collect_reset, dim3(1), dim3(64), 0, stream, counters);
    const uint32_t grid = std::max<uint32_t>(1u, std::min<uint32_t>(n_cus * 4u * CL_WAVES, L.n_units));
    hipLaunchKernelGGL(k_collect_units, dim3(grid), dim3(WAVE), 0, stream, sets, units, L.n_units, counters, d_m6,
                       d_keys ? cap : 0u, d_count, d_hash, d_keys, ovf);
    jg_status st = hip_check(hipGetLastError(), "batched collection launch");
    if (st != JG_OK) return st;
    collect_overflow_launch(sets, counters, ovf, d_m6, L.max_slot_words, d_keys ? cap : 0u, d_count, d_hash, d_keys,
                            n_cus * 4u, stream);
    return hip_check(hipGetLastError(), "collection overflow launch");
}

}  // namespace jg
