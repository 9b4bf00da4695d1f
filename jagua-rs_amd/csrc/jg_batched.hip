// Batched persistent query form (jg_plan_*): many (engine, item, transform range) sets run
// in one launch sequence fed by device work queues.
//   k_surrogate: work units of up to 64 queries; lanes = (query, ff pole/pier) pairs, each
//                lane descends from the root (cd_engine.rs:204-223).  Collisions are written
//                out; the survivors of a unit are appended to the survivor queue in groups of
//                one polygon-stage pass each (64/G queries, (set, first query, mask)).
//   k_poly:      survivor groups; each query is handled by G lanes (G = the item's vertex
//                count rounded up to a power of two, 8..32 by default), 64/G queries per
//                wavefront: transformed bbox, virtual root, the item's edges descending from it
//                with a group-wide early exit, then containment over the virtual root's hazards
//                (cd_engine.rs:160-197).
// The default schedule is these two kernels.  k_fused runs both stages of a work unit in one
// wave instead: no survivor queue, but a unit with many survivors keeps its wave busy long
// after the others finish; the autotuner weighs it.  Units are ordered by item size (largest
// first).  Work is claimed from 8 interleaved streams of counters, never from one address.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/jagua_amd.h"
#include "jg_internal.hpp"
#include "jg_plan.hpp"

using namespace jg;

namespace {

// LDS of a workgroup (one wavefront): its work list (within 8 KB, so that 20 waves share a CU)
struct WaveLds {
    uint32_t nodes[WL_NCAP];
    uint32_t edges[WL_ECAP];
    float4 geom[WL_ENT];
    uint32_t hit[64];
    uint32_t qidx[64];
    uint8_t edge_ent[WL_ECAP];
    uint8_t ent_slot[WL_ENT];
};
static_assert(sizeof(WaveLds) <= 8192, "20 waves of a CU share its 160 KB of LDS");

__device__ __forceinline__ WorkList wave_list(WaveLds& L) {
    return WorkList{L.nodes, L.edges, L.edge_ent, L.geom, L.ent_slot, L.hit, L.qidx};
}

}  // namespace

// Hazard edges per edge-test item of the surrogate stage's traversals (pole circles and piers)
// and of the polygon stage's (query edges).
constexpr int SURR_EG = 4;
constexpr int POLY_EG = 1;

// The surrogate stage of one work unit (cd_engine.rs:204-223 for each of its queries): writes
// the verdicts it settles and returns the mask of queries (bit i = query un.q0 + i) that go on
// to the polygon stage.
template <bool FILT, int KIND, int IPL>
__device__ __forceinline__ uint64_t surrogate_unit(const SetDesc& sd, const Unit& un, const WorkList& W,
                                                   const float* __restrict__ m6, uint8_t* __restrict__ status,
                                                   uint32_t lane) {
    const CdeView c = sd.cde;
    const ItemView it = sd.item;
    const FilterView f = sd.filter;
    const uint32_t flags = sd.flags;
    const uint64_t base = sd.first + un.q0;
    const uint32_t cnt = un.count;
    const uint64_t valid = cnt == 64 ? ~0ull : ((1ull << cnt) - 1ull);

    uint64_t hit_mask = 0;
    const uint32_t n_ent = (flags & JG_QUERY_SURROGATE) ? it.n_poles + it.n_piers : 0u;
    if (n_ent > 0) {
        // every (query, ff pole / pier) of the unit descends from the root
        // (cd_engine.rs:204-223), WL_ENT entities per traversal
        W.hit[lane] = 0u;
        if (FILT) W.qidx[lane] = un.q0 + lane;
        const uint32_t total = cnt * n_ent;
        for (uint32_t e0 = 0; e0 < total; e0 += WL_ENT) {
            wave_sync();
            uint32_t n = 0;
            for (uint32_t j = 0; j < WL_ENT; j += WAVE) {
                const uint32_t o = j + lane;
                const uint32_t e = e0 + o;
                bool want = false;
                if (e < total) {
                    const uint32_t qi = e / n_ent, k = e - qi * n_ent;
                    if (!W.hit[qi]) {
                        const Mat m = load_mat(m6, base + qi);
                        const float4 p = ldg4(it.ff + k);
                        const bool circle = k < it.n_poles;
                        float4 gm;
                        if (circle) {
                            const float2 cc = xform(m, p.x, p.y);
                            gm = make_float4(cc.x, cc.y, p.z, 0.0f);
                        } else {
                            const float2 a = xform(m, p.x, p.y), bb = xform(m, p.z, p.w);
                            gm = make_float4(a.x, a.y, bb.x, bb.y);
                        }
                        W.geom[o] = gm;
                        W.ent_slot[o] = (uint8_t)(qi | (circle ? WL_CIRCLE : 0u));
                        want = true;
                    }
                }
                n = wave_push_roots(W, n, want, o, 0u, lane);
            }
            wave_traverse<FILT, KIND, IPL, SURR_EG>(c, f, W, n, lane);
        }
        hit_mask = __ballot(W.hit[lane] != 0u) & valid;
    }
    const uint64_t todo = (flags & JG_QUERY_POLY) ? (valid & ~hit_mask) : 0ull;
    if (lane < cnt) {
        const bool h = (hit_mask >> lane) & 1ull;
        if (h) status[base + lane] = JG_Q_COLLISION;
        else if (!(flags & JG_QUERY_POLY)) status[base + lane] = JG_Q_CLEAR;
    }
    return todo;
}

template <bool FILT, int KIND>
__global__ __launch_bounds__(WAVE, 5) void k_surrogate(const SetDesc* __restrict__ sets, const Unit* __restrict__ units,
                                                   uint32_t n_units, uint32_t* __restrict__ counters,
                                                   Group* __restrict__ groups, const uint32_t* __restrict__ gbase,
                                                   const float* __restrict__ m6, uint8_t* __restrict__ status) {
    const uint32_t lane = threadIdx.x;
    __shared__ WaveLds lds;
    const WorkList W = wave_list(lds);
    // the polygon stage of this launch claims groups after this kernel: clear its claim counters
    if (blockIdx.x == 0 && lane < N_STREAMS) counters[C_STREAM0 + STREAM_STRIDE * lane] = 0u;
    uint32_t ux = blockIdx.x & (N_STREAMS - 1u), utried = 0;
    for (;;) {
        const uint32_t u = claim_next(counters, C_USTREAM0, ux, utried, n_units, lane);
        if (u >= n_units) return;
        const Unit un = units[u];
        const uint64_t todo = surrogate_unit<FILT, KIND, 2>(sets[un.set], un, W, m6, status, lane);
        if (todo != 0ull && lane == 0) {
            // one survivor group per polygon-stage pass (64 / G queries)
            const uint32_t per = 64u / sets[un.set].lanes;
            const uint32_t n_g = ((uint32_t)__builtin_popcountll(todo) + per - 1u) / per;
            // the unit's stream (u mod 8) appends to its own region of the group array
            const uint32_t xs = u & (N_STREAMS - 1u);
            const uint32_t g0 = gbase[xs] + atomicAdd(&counters[C_GCOUNT0 + STREAM_STRIDE * xs], n_g);
            uint64_t rest = todo;
            for (uint32_t j = 0; j < n_g; j++) {
                uint64_t r = rest;
                for (uint32_t i = 0; i < per && r; i++) r &= r - 1ull;
                groups[g0 + j] = Group{un.set, un.q0, rest & ~r};
                rest = r;
            }
        }
        wave_sync();
    }
}

// detect_containment_collision (cd_engine.rs:231-255) over the candidates of one query, the
// entries e0 .. e0 + cnt, by the query's G lanes (lane sub of the group starting at lane
// base): true at the first that collides.  Each lane filters four candidates in flight (their
// entries, then their bboxes and interior flags, read before any test): a shallow virtual root
// holds up to every hazard of the layout.  A candidate without bbox relation collides iff its
// scope is exterior; the point-in-polygon tests the filter leaves (hazard bbox surrounding, or
// enclosed by, the query's) run on all G lanes, each polygon's edges split over them, since one
// lane walking a container of thousands of vertices would hold the group.  The result is
// containment()'s: interior ? contained : !contained, OR over the candidates.
template <bool FILT>
__device__ __forceinline__ bool scan_containment(const CdeView& c, const FilterView& f, uint64_t q, uint32_t e0,
                                                 uint32_t cnt, uint32_t sub, uint32_t G, uint32_t base,
                                                 const ItemView& it, const Mat& m, float4 bb, float2 qpoi) {
    constexpr uint32_t CU = 4;
    const uint64_t gmask = G >= 64u ? ~0ull : ((1ull << G) - 1ull);
    for (uint32_t k0 = 0; k0 < cnt; k0 += CU * G) {
        uint32_t sl[CU];
        float4 hb[CU];
        float iw[CU];
#pragma unroll
        for (uint32_t r = 0; r < CU; r++) {
            const uint32_t i = k0 + r * G + sub;
            sl[r] = ldg_slot_flags(c.ent + e0 + (i < cnt ? i : 0u)) & SLOT_MASK;
        }
#pragma unroll
        for (uint32_t r = 0; r < CU; r++) {
            hb[r] = ldg4(c.haz_bbox + sl[r]);
            iw[r] = ldg4(c.haz_poi + sl[r]).w;
        }
        bool direct = false;               // exterior scope, no bbox relation: !contained
        uint32_t pend = 0;                 // bit r: candidate r needs its point-in-polygon test
        int prel[CU];
#pragma unroll
        for (uint32_t r = 0; r < CU; r++) {
            prel[r] = REL_DISJOINT;
            if (k0 + r * G + sub >= cnt || irrel<FILT>(f, q, sl[r])) continue;
            prel[r] = box_almost_relation(hb[r], bb);
            if (prel[r] == REL_SURROUNDING || prel[r] == REL_ENCLOSED) pend |= 1u << r;
            else if (iw[r] == 0.0f) direct = true;
        }
        if ((ballot(direct) >> base) & gmask) return true;
#pragma unroll
        for (uint32_t r = 0; r < CU; r++) {
            for (uint64_t pm = (ballot((pend >> r) & 1u) >> base) & gmask; pm; pm &= pm - 1ull) {
                const int l = (int)base + __builtin_ctzll(pm);
                const uint32_t s2 = (uint32_t)__shfl((int)sl[r], l, 64);
                const int rl = __shfl(prel[r], l, 64);
                bool contained;
                if (rl == REL_SURROUNDING) {
                    contained = pip_hazard_group(c, s2, qpoi.x, qpoi.y, sub, G);
                } else {
                    const float4 hp = ldg4(c.haz_poi + s2);
                    contained = pip_query_group(it, m, bb, hp.x, hp.y, sub, G);
                }
                if (ldg4(c.haz_poi + s2).w != 0.0f ? contained : !contained) return true;
            }
        }
    }
    return false;
}

// A query whose virtual root holds more containment candidates than this has them scanned by
// the whole wave after the pass's other work, not by its own G lanes.
constexpr uint32_t WAVE_SCAN_MIN = 64;

// scan_containment for one query by the whole wave (all 64 lanes, the same arguments but
// lane): candidates are filtered four per lane in flight; the point-in-polygon tests the
// filter leaves (hazard bbox surrounding, or enclosed by, the query's) run on all 64 lanes,
// since one lane walking a container of thousands of vertices would hold the wave.  The
// result is the same: a candidate collides iff interior ? contained : !contained.
template <bool FILT>
__device__ __forceinline__ bool scan_containment_wave(const CdeView& c, const FilterView& f, uint64_t q, uint32_t e0,
                                                      uint32_t cnt, uint32_t lane, const ItemView& it, const Mat& m,
                                                      float4 bb, float2 qpoi) {
    constexpr uint32_t CU = 4;
    for (uint32_t k0 = 0; k0 < cnt; k0 += CU * 64u) {
        uint32_t sl[CU];
        float4 hb[CU];
        float iw[CU];
#pragma unroll
        for (uint32_t r = 0; r < CU; r++) {
            const uint32_t i = k0 + r * 64u + lane;
            sl[r] = ldg_slot_flags(c.ent + e0 + (i < cnt ? i : 0u)) & SLOT_MASK;
        }
#pragma unroll
        for (uint32_t r = 0; r < CU; r++) {
            hb[r] = ldg4(c.haz_bbox + sl[r]);
            iw[r] = ldg4(c.haz_poi + sl[r]).w;
        }
        bool direct = false;               // exterior scope, no bbox relation: !contained
        uint32_t pend = 0;                 // bit r: candidate r needs its point-in-polygon test
        int prel[CU];
#pragma unroll
        for (uint32_t r = 0; r < CU; r++) {
            prel[r] = REL_DISJOINT;
            if (k0 + r * 64u + lane >= cnt || irrel<FILT>(f, q, sl[r])) continue;
            prel[r] = box_almost_relation(hb[r], bb);
            if (prel[r] == REL_SURROUNDING || prel[r] == REL_ENCLOSED) pend |= 1u << r;
            else if (iw[r] == 0.0f) direct = true;
        }
        if (ballot(direct) != 0ull) return true;
#pragma unroll
        for (uint32_t r = 0; r < CU; r++) {
            for (uint64_t pm = ballot((pend >> r) & 1u); pm; pm &= pm - 1ull) {
                const int l = __builtin_ctzll(pm);
                const uint32_t s = (uint32_t)__shfl((int)sl[r], l, 64);
                const int rl = __shfl(prel[r], l, 64);
                bool contained;
                if (rl == REL_SURROUNDING) {
                    contained = pip_hazard_wave(c, s, qpoi.x, qpoi.y, lane);
                } else {
                    const float4 hp = ldg4(c.haz_poi + s);
                    contained = pip_query_wave(it, m, bb, hp.x, hp.y, lane);
                }
                if (ldg4(c.haz_poi + s).w != 0.0f ? contained : !contained) return true;
            }
        }
    }
    return false;
}

// The polygon stage (cd_engine.rs:160-197) of the queries q0 + i for the set bits i of mask,
// all of one set: 64/G queries per pass, G lanes each.
template <bool FILT>
__device__ __forceinline__ void poly_queries(const SetDesc& sd, uint32_t q0, uint64_t mask, const WorkList& W,
                                             const float* __restrict__ m6, uint8_t* __restrict__ status,
                                             uint32_t lane) {
    const CdeView c = sd.cde;
    const ItemView it = sd.item;
    const FilterView f = sd.filter;
    const uint32_t G = sd.lanes;
    const uint32_t per = 64u / G;                     // queries per pass
    const uint32_t slot = lane / G, sub = lane % G;
    const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    const float4 root_bb = ldg_node(c.nodes).bbox;
    const uint32_t nv = it.n_v;
    uint64_t todo = mask;
    while (todo) {
        // this lane's query: the slot-th remaining survivor
        uint64_t t = todo;
        for (uint32_t s = 0; s < slot && t; s++) t &= t - 1;
        const bool active = t != 0ull;
        const uint32_t qi = active ? (uint32_t)__builtin_ctzll(t) : 0u;
        for (uint32_t s = 0; s < per && todo; s++) todo &= todo - 1;
        const uint64_t q = (uint64_t)q0 + qi;
        Mat m{};
        if (active) m = load_mat(m6, sd.first + q);
        // transformed bbox (simple_polygon.rs:122-133), reduced over the group's lanes
        float4 bb = make_float4(3.40282347e+38f, 3.40282347e+38f, -3.40282347e+38f, -3.40282347e+38f);
        if (active)
            for (uint32_t i = sub; i < nv; i += 4u * G) {
                // four vertices in flight (past the end: vertex i again, which leaves the box as is)
                float2 v4[4];
#pragma unroll
                for (uint32_t u = 0; u < 4u; u++) v4[u] = ldg2(it.v + (i + u * G < nv ? i + u * G : i));
#pragma unroll
                for (uint32_t u = 0; u < 4u; u++) {
                    const float2 p = xform(m, v4[u].x, v4[u].y);
                    bb.x = mn(bb.x, p.x); bb.y = mn(bb.y, p.y); bb.z = mx(bb.z, p.x); bb.w = mx(bb.w, p.y);
                }
            }
        for (uint32_t o = G >> 1; o > 0; o >>= 1) {
            bb.x = fminf(bb.x, __shfl_xor(bb.x, (int)o, 64));
            bb.y = fminf(bb.y, __shfl_xor(bb.y, (int)o, 64));
            bb.z = fmaxf(bb.z, __shfl_xor(bb.z, (int)o, 64));
            bb.w = fmaxf(bb.w, __shfl_xor(bb.w, (int)o, 64));
        }
        uint8_t res = JG_Q_CLEAR;
        bool done = !active;
        if (active && box_relation(root_bb, bb) != REL_SURROUNDING) {
            res = JG_Q_COLLISION;
            done = true;
        }
        int vr = 0;
        uint32_t vr_depth = 0;
        if (G == 64u) {
            // one query per wavefront: its bbox is wave-uniform, so the descent's node reads go
            // through the scalar cache instead of a vector load per level
            const float4 ub = make_float4(uniform_f32(bb.x), uniform_f32(bb.y), uniform_f32(bb.z), uniform_f32(bb.w));
            if (__builtin_amdgcn_readfirstlane(active ? 1 : 0)) vr = virtual_root_cells(c, root_bb, ub, vr_depth);
        } else if (active) {
            vr = virtual_root_cells(c, root_bb, bb, vr_depth);
        }
        // the virtual root's box and cell, below which the edges' start nodes are followed
        float4 vbox = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        uint32_t vx = 0, vy = 0;
        const bool vstart = active && vr_depth < c.cell_level;
        if (vstart) {
            vbox = ldg_node(c.nodes + vr).bbox;
            node_cell(root_bb, vbox, vr_depth, vx, vy);
        }
        // edges of the pass's unfinished queries, descending from their virtual roots
        // (cd_engine.rs:170-175), each from the deepest node below its virtual root known to
        // hold it (start_node_below); a query slot is marked hit as soon as one edge collides
        wave_sync();
        if (sub == 0) {
            W.hit[slot] = done ? 1u : 0u;
            if (FILT) W.qidx[slot] = (uint32_t)q;
        }
        // round R: lane (slot, sub) sets up edges sub + G (2R + j), j < WL_ENT / 64
        constexpr uint32_t EPL = WL_ENT / WAVE;           // entities per lane per round
        for (uint32_t i0 = 0; i0 < nv; i0 += EPL * G) {
            wave_sync();
            if (__ballot(active && W.hit[slot] == 0u) == 0ull) break;
            uint32_t n = 0;
            for (uint32_t j = 0; j < EPL; j++) {
                const uint32_t i = i0 + j * G + sub;
                const uint32_t o = j * WAVE + lane;
                const bool want = active && i < nv && W.hit[slot] == 0u;
                uint32_t start = (uint32_t)vr;
                if (want) {
                    const uint32_t k = i + 1 == nv ? 0 : i + 1;
                    const float2 vi = ldg2(it.v + i), vk = ldg2(it.v + k);
                    const float2 a = xform(m, vi.x, vi.y);
                    const float2 b2 = xform(m, vk.x, vk.y);
                    W.geom[o] = make_float4(a.x, a.y, b2.x, b2.y);
                    W.ent_slot[o] = (uint8_t)slot;
                    const float4 eb = make_float4(mn(a.x, b2.x), mn(a.y, b2.y), mx(a.x, b2.x), mx(a.y, b2.y));
                    if (vstart) start = start_node_below(c, (uint32_t)vr, vr_depth, vbox, vx, vy, eb);
                }
                n = wave_push_roots(W, n, want, o, start, lane);
            }
            wave_traverse<FILT, KIND_EDGE, 1, POLY_EG>(c, f, W, n, lane);
        }
        // containment over the virtual root's hazards (cd_engine.rs:176-193)
        wave_sync();
        done = W.hit[slot] != 0u || done;
        if (active && W.hit[slot] != 0u && res == JG_Q_CLEAR) res = JG_Q_COLLISION;
        bool hit = false, heavy = false;
        uint32_t e0 = 0, cntv = 0;
        if (!done) {
            const DevNode vn = ldg_node(c.nodes + vr);
            cntv = vn.haz_info & COUNT_MASK;
            e0 = vn.haz_begin;
            if (cntv > 0 && (vn.haz_info & ENTIRE_BIT)) {
                res = JG_Q_PANIC_FILTERED_ENTIRE;
                done = true;
            } else if (G < 64u && cntv > WAVE_SCAN_MIN) {
                heavy = true;
            } else {
                const float2 qpoi = xform(m, it.poi.x, it.poi.y);
                hit = scan_containment<FILT>(c, f, q, e0, cntv, sub, G, slot * G, it, m, bb, qpoi);
            }
        }
        // the long candidate lists, one query at a time over all 64 lanes; the query's
        // transform, bbox, entry range and index come from its first lane bit for bit
        for (uint64_t hv = ballot(heavy && sub == 0); hv; hv &= hv - 1ull) {
            const int l0 = __builtin_ctzll(hv);
            Mat mq;
            mq.a = __shfl(m.a, l0, 64);
            mq.b = __shfl(m.b, l0, 64);
            mq.c = __shfl(m.c, l0, 64);
            mq.d = __shfl(m.d, l0, 64);
            mq.e = __shfl(m.e, l0, 64);
            mq.f = __shfl(m.f, l0, 64);
            const float4 bq = make_float4(__shfl(bb.x, l0, 64), __shfl(bb.y, l0, 64), __shfl(bb.z, l0, 64),
                                          __shfl(bb.w, l0, 64));
            const uint32_t e0q = (uint32_t)__shfl((int)e0, l0, 64), cq = (uint32_t)__shfl((int)cntv, l0, 64);
            const uint64_t qq = ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(q >> 32), l0, 64) << 32) |
                                (uint32_t)__shfl((int)(uint32_t)q, l0, 64);
            const float2 pq = xform(mq, it.poi.x, it.poi.y);
            const bool any = scan_containment_wave<FILT>(c, f, qq, e0q, cq, lane, it, mq, bq, pq);
            if ((int)(lane / G) == l0 / (int)G) hit = any;
        }
        const uint64_t bal = ballot(hit);
        if (!done) res = ((bal >> (slot * G)) & gmask) ? JG_Q_COLLISION : JG_Q_CLEAR;
        if (active && sub == 0) status[sd.first + q] = res;
    }
}

template <bool FILT>
__device__ __forceinline__ void poly_loop(const SetDesc* __restrict__ sets, uint32_t* __restrict__ counters,
                                          const Group* __restrict__ groups, const uint32_t* __restrict__ gbase,
                                          const float* __restrict__ m6, uint8_t* __restrict__ status, WaveLds& lds) {
    const uint32_t lane = threadIdx.x;
    const WorkList W = wave_list(lds);
    // the surrogate stage of this launch is done: clear its unit-claim counters for the next
    if (blockIdx.x == 0 && lane < N_STREAMS) counters[C_USTREAM0 + STREAM_STRIDE * lane] = 0u;
    // stream x holds the groups its units appended; a wave drains its own stream, then the next
    uint32_t x = blockIdx.x & (N_STREAMS - 1u);
    for (uint32_t tried = 0; tried < N_STREAMS;) {
        const uint32_t n_x = __atomic_load_n(&counters[C_GCOUNT0 + STREAM_STRIDE * x], __ATOMIC_RELAXED);
        uint32_t j = 0;
        if (lane == 0) j = atomicAdd(&counters[C_STREAM0 + STREAM_STRIDE * x], 1u);
        j = wave_bcast(j);
        if (j >= n_x) {
            x = (x + 1u) & (N_STREAMS - 1u);
            tried++;
            continue;
        }
        const Group gr = groups[gbase[x] + j];
        poly_queries<FILT>(sets[gr.set], gr.q0, gr.mask, W, m6, status, lane);
    }
    // every wave has read the group counts once the last one is out
    if (last_wave_out(counters, lane) && lane < N_STREAMS) counters[C_GCOUNT0 + STREAM_STRIDE * lane] = 0u;
}

// The two-stage kernels run 5 waves per SIMD (96 VGPRs, a few spilled; the LDS of a wave is
// within 8 KB).  The filtered polygon stage carries the filter state on top and spills more at
// 96, and still measured 4% faster on the filtered many-layout workload than at 4 waves; the
// fused kernels stay bounded to 4 (128 VGPRs: their pier-entity instance lands at 129), which
// measured 16% faster there than 3 waves without spills.
__global__ __launch_bounds__(WAVE, 5) void k_poly(const SetDesc* __restrict__ sets, uint32_t* __restrict__ counters,
                                              const Group* __restrict__ groups, const uint32_t* __restrict__ gbase,
                                              const float* __restrict__ m6, uint8_t* __restrict__ status) {
    __shared__ WaveLds lds;
    poly_loop<false>(sets, counters, groups, gbase, m6, status, lds);
}
__global__ __launch_bounds__(WAVE, 5) void k_poly_filtered(const SetDesc* __restrict__ sets, uint32_t* __restrict__ counters,
                                                          const Group* __restrict__ groups,
                                                          const uint32_t* __restrict__ gbase, const float* __restrict__ m6,
                                                          uint8_t* __restrict__ status) {
    __shared__ WaveLds lds;
    poly_loop<true>(sets, counters, groups, gbase, m6, status, lds);
}

// Both stages in one persistent kernel: a wave runs a work unit's surrogate stage and then the
// polygon stage of that unit's survivors itself.  No survivor queue is written, but the work of
// a unit cannot spread: a unit of a large item with 30 survivors keeps one wave busy for about
// 1.2 ms of a 1.34 ms configuration-2 step, so the two-kernel schedule is the default.
// Items per lane of the fused kernel's surrogate stage.  2 fits in 90 VGPRs (4 waves per SIMD
// kept) but measured 2% slower on the configuration-2 step than 1.
constexpr int FUSED_IPL = 1;
template <bool FILT, int KIND>
__device__ __forceinline__ void fused_loop(const SetDesc* __restrict__ sets, const Unit* __restrict__ units,
                                           uint32_t n_units, uint32_t* __restrict__ counters,
                                           const float* __restrict__ m6, uint8_t* __restrict__ status,
                                           WaveLds& lds) {
    const uint32_t lane = threadIdx.x;
    const WorkList W = wave_list(lds);
    uint32_t ux = blockIdx.x & (N_STREAMS - 1u), utried = 0;
    for (;;) {
        const uint32_t u = claim_next(counters, C_USTREAM0, ux, utried, n_units, lane);
        if (u >= n_units) break;
        const Unit un = units[u];
        const SetDesc& sd = sets[un.set];
        const uint64_t todo = surrogate_unit<FILT, KIND, FUSED_IPL>(sd, un, W, m6, status, lane);
        if (todo != 0ull) poly_queries<FILT>(sd, un.q0, todo, W, m6, status, lane);
        wave_sync();
    }
    // every wave has stopped claiming once the last one is out
    if (last_wave_out(counters, lane) && lane < N_STREAMS) counters[C_USTREAM0 + STREAM_STRIDE * lane] = 0u;
}

template <int KIND>
__global__ __launch_bounds__(WAVE, 4) void k_fused(const SetDesc* __restrict__ sets, const Unit* __restrict__ units,
                                               uint32_t n_units, uint32_t* __restrict__ counters,
                                               const float* __restrict__ m6, uint8_t* __restrict__ status) {
    __shared__ WaveLds lds;
    fused_loop<false, KIND>(sets, units, n_units, counters, m6, status, lds);
}
template <int KIND>
__global__ __launch_bounds__(WAVE, 4) void k_fused_filtered(const SetDesc* __restrict__ sets, const Unit* __restrict__ units,
                                                           uint32_t n_units, uint32_t* __restrict__ counters,
                                                           const float* __restrict__ m6, uint8_t* __restrict__ status) {
    __shared__ WaveLds lds;
    fused_loop<true, KIND>(sets, units, n_units, counters, m6, status, lds);
}

// ---------------------------------------------------------------------------------------
namespace {


// JG_FILTER_KEYS masks of the sets, one after another: set i's at words[off[i]..].
void filter_masks(const std::vector<jg_query_set>& sets, const std::vector<jg_hazard_filter>& filters,
                  const std::vector<std::vector<jg_hazkey>>& keys, std::vector<uint32_t>& words,
                  std::vector<size_t>& off) {
    words.clear();
    off.assign(sets.size(), 0);
    for (size_t i = 0; i < sets.size(); i++) {
        if (filters[i].kind != JG_FILTER_KEYS) continue;
        const jg_cde* c = sets[i].cde;
        const uint32_t n = (uint32_t)((c->map.slots.size() + 31) / 32);
        off[i] = words.size();
        words.resize(words.size() + n, 0u);
        for (const jg_hazkey& k : keys[i])
            if (c->map.valid(k)) words[off[i] + (k.idx >> 5)] |= 1u << (k.idx & 31u);
    }
}

FilterView set_filter(const jg_query_set& s, const jg_hazard_filter& f, const uint32_t* mask) {
    FilterView v{nullptr, 0, 0, 0, 0};
    if (f.kind == JG_FILTER_HASHED) {
        v.kind = 2;
        v.seed = f.seed;
        v.threshold = f.threshold;
    } else if (f.kind == JG_FILTER_KEYS) {
        v.kind = 1;
        v.mask = mask;
        v.n_words = (uint32_t)((s.cde->map.slots.size() + 31) / 32);
    }
    return v;
}

// The host bytes of the table of `sets` under schedule `cfg`, given each set's item view and
// filter view (device addresses), and its layout.
// seeds[i] (device address, count): the slots set i's collectors start with (may be empty).
jg_status table_image(const std::vector<jg_query_set>& sets, const std::vector<ItemView>& iv,
                      const std::vector<FilterView>& fv, const std::vector<std::pair<const uint32_t*, uint32_t>>& seeds,
                      const jg_plan_config& cfg, uint32_t n_cus, std::vector<uint8_t>& host, TableLayout& L) {
    for (auto& s : sets) {
        if (s.cde->view.n_nodes >= WL_MAX_NODES)
            return fail(JG_ERR_CAPACITY, "quadtree of more than 2^24 nodes in a batched plan");
        if (s.cde->img.cap_edges > WL_MAX_EDGES)
            return fail(JG_ERR_CAPACITY, "engine image of more than 2^29 hazard edges in a batched plan");
    }
    std::vector<SetDesc> sd(sets.size());
    bool any_filter = false, any_pole = false, any_pier = false;
    L.max_slot_words = 1;
    L.n_queries = 0;
    for (size_t i = 0; i < sets.size(); i++) {
        const jg_query_set& s = sets[i];
        std::memset(&sd[i], 0, sizeof(SetDesc));
        L.max_slot_words = std::max<uint32_t>(L.max_slot_words, (s.cde->view.n_slots + 31u) / 32u);
        L.n_queries += s.count;
        if (i < seeds.size()) {
            sd[i].seed = seeds[i].first;
            sd[i].n_seed = seeds[i].second;
        }
        sd[i].cde = s.cde->view;
        sd[i].item = iv[i];
        sd[i].flags = cfg.flags ? cfg.flags : (JG_QUERY_SURROGATE | JG_QUERY_POLY);
        if (!(sd[i].flags & JG_QUERY_SURROGATE)) { sd[i].item.n_poles = 0; sd[i].item.n_piers = 0; }
        sd[i].first = s.first;
        uint32_t lanes = cfg.lanes_per_query;
        if (lanes == 0) {
            // the vertex count rounded up to a power of two, 8..32: items above 32 vertices run
            // two queries per pass, which measured faster than one query on 64 lanes
            lanes = 8;
            while (lanes < s.item->poly.n() && lanes < 32u) lanes <<= 1;
        }
        uint32_t pow2 = 1;
        while (pow2 < lanes && pow2 < 64u) pow2 <<= 1;
        sd[i].lanes = pow2;
        sd[i].filter = fv[i];
        any_filter = any_filter || sd[i].filter.kind != 0;
        any_pole = any_pole || sd[i].item.n_poles > 0;
        any_pier = any_pier || sd[i].item.n_piers > 0;
    }
    L.filtered = any_filter;
    L.kind = any_pier ? (any_pole ? KIND_MIXED : KIND_EDGE) : KIND_CIRCLE;
    L.stages = cfg.stages == 1u ? 1u : 2u;
    // work units: sets in decreasing item size, `chunk` queries each (16, 32 or 64; 64 packs
    // 128 pole entities into one surrogate traversal; smaller units shorten the fused kernel's
    // tail)
    const uint32_t uq = cfg.chunk == 0 ? 64u : cfg.chunk <= 16u ? 16u : cfg.chunk <= 32u ? 32u : 64u;
    std::vector<size_t> order(sets.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return sets[a].item->poly.n() > sets[b].item->poly.n(); });
    std::vector<Unit> units;
    for (size_t oi : order) {
        const jg_query_set& s = sets[oi];
        for (uint64_t q = 0; q < s.count; q += uq)
            units.push_back(Unit{(uint32_t)oi, (uint32_t)q, (uint32_t)std::min<uint64_t>(uq, s.count - q), 0u});
    }
    if (units.size() >= 0xffffffffull) return fail(JG_ERR_CAPACITY, "too many work units in a batched plan");
    L.n_units = (uint32_t)units.size();
    L.o_units = (sizeof(SetDesc) * sd.size() + 255) & ~(size_t)255;
    L.o_ctr = (L.o_units + sizeof(Unit) * units.size() + 255) & ~(size_t)255;
    L.o_gbase = L.o_ctr + 4u * N_COUNTER_WORDS;
    // one group per polygon-stage pass at most; unit u's groups go to the region of stream u mod 8
    uint64_t stream_groups[N_STREAMS] = {};
    for (size_t u = 0; u < units.size(); u++) {
        const uint32_t per = 64u / sd[units[u].set].lanes;
        stream_groups[u & (N_STREAMS - 1u)] += (units[u].count + per - 1u) / per;
    }
    uint32_t gbase[N_STREAMS];
    uint64_t max_groups = 0;
    for (uint32_t x = 0; x < N_STREAMS; x++) {
        gbase[x] = (uint32_t)max_groups;
        max_groups += stream_groups[x];
        if (max_groups >= 0xffffffffull) return fail(JG_ERR_CAPACITY, "too many survivor groups in a batched plan");
    }
    L.o_grp = (L.o_gbase + sizeof(gbase) + 255) & ~(size_t)255;
    L.total = L.o_grp + sizeof(Group) * (max_groups + 1);
    host.assign(L.o_grp, 0);                 // the group array is written by the kernels
    std::memcpy(host.data(), sd.data(), sizeof(SetDesc) * sd.size());
    std::memcpy(host.data() + L.o_units, units.data(), sizeof(Unit) * units.size());
    std::memcpy(host.data() + L.o_gbase, gbase, sizeof(gbase));
    // one wavefront per workgroup, each taking one work unit (or survivor group) at a time
    const uint32_t want = cfg.workgroups ? cfg.workgroups : n_cus * 20u;
    L.grid_s = std::max<uint32_t>(1u, std::min<uint32_t>(want, L.n_units));
    L.grid_p = std::max<uint32_t>(1u, (uint32_t)std::min<uint64_t>(want, max_groups));
    return JG_OK;
}

// The launch sequence over a table at device address `base`: the two stages or the fused
// kernel.  Their claim counters start at zero (a new table's are; each sequence, detection or
// collection, leaves them so), so no clearing step precedes them.
jg_status table_launch(const TableLayout& L, uint8_t* base, const float* d_m6, uint8_t* d_status,
                       hipStream_t stream) {
    uint32_t* counters = (uint32_t*)(base + L.o_ctr);
    Group* groups = (Group*)(base + L.o_grp);
    const uint32_t* gbase = (const uint32_t*)(base + L.o_gbase);
    const SetDesc* sets = (const SetDesc*)base;
    const Unit* units = (const Unit*)(base + L.o_units);
    const dim3 gs(L.grid_s), gp(L.grid_p), blk(WAVE);
    if (L.stages == 1u) {
#define JG_FUSED(KER, K) hipLaunchKernelGGL((KER<K>), gs, blk, 0, stream, sets, units, L.n_units, counters, d_m6, d_status)
        if (L.filtered) {
            if (L.kind == KIND_CIRCLE) JG_FUSED(k_fused_filtered, KIND_CIRCLE);
            else if (L.kind == KIND_EDGE) JG_FUSED(k_fused_filtered, KIND_EDGE);
            else JG_FUSED(k_fused_filtered, KIND_MIXED);
        } else {
            if (L.kind == KIND_CIRCLE) JG_FUSED(k_fused, KIND_CIRCLE);
            else if (L.kind == KIND_EDGE) JG_FUSED(k_fused, KIND_EDGE);
            else JG_FUSED(k_fused, KIND_MIXED);
        }
#undef JG_FUSED
        return hip_check(hipGetLastError(), "batched query launch");
    }
#define JG_SURR(F, K) hipLaunchKernelGGL((k_surrogate<F, K>), gs, blk, 0, stream, sets, units, L.n_units, counters, \
                                         groups, gbase, d_m6, d_status)
    if (L.filtered) {
        if (L.kind == KIND_CIRCLE) JG_SURR(true, KIND_CIRCLE);
        else if (L.kind == KIND_EDGE) JG_SURR(true, KIND_EDGE);
        else JG_SURR(true, KIND_MIXED);
        hipLaunchKernelGGL(k_poly_filtered, gp, blk, 0, stream, sets, counters, (const Group*)groups, gbase, d_m6,
                           d_status);
    } else {
        if (L.kind == KIND_CIRCLE) JG_SURR(false, KIND_CIRCLE);
        else if (L.kind == KIND_EDGE) JG_SURR(false, KIND_EDGE);
        else JG_SURR(false, KIND_MIXED);
        hipLaunchKernelGGL(k_poly, gp, blk, 0, stream, sets, counters, (const Group*)groups, gbase, d_m6, d_status);
    }
#undef JG_SURR
    return hip_check(hipGetLastError(), "batched query launch");
}

uint32_t device_cus() {
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return (uint32_t)cus;
}

}  // namespace

struct jg_plan {
    std::vector<jg_query_set> sets;
    std::vector<jg_hazard_filter> filters;
    std::vector<std::vector<jg_hazkey>> filter_keys;
    jg_plan_config cfg{};
    DevBlob table;                  // SetDesc[], Unit[], counters, survivor groups
    DevBlob masks;                  // JG_FILTER_KEYS masks of every set, then their slots ascending
    DevBlob ovf;                    // collection: the queries handed to the per-query collector
    TableLayout L;
    // engine image generations and item upload generations the table was built from
    std::vector<uint64_t> generations;
    uint32_t n_cus = 0;
    // graph cache
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    const float* g_m6 = nullptr;
    uint8_t* g_status = nullptr;
    hipStream_t g_stream = nullptr;
    ~jg_plan() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace {

jg_status plan_build_table(jg_plan* p) {
    // engines and items resident and current
    for (auto& s : p->sets) {
        jg_status st = cde_upload(s.cde);
        if (st != JG_OK) return st;
        st = item_upload(const_cast<jg_spolygon*>(s.item));
        if (st != JG_OK) return st;
    }
    // the table holds engine views and item views: it is rebuilt when an engine image or an
    // item record (a regenerated surrogate, say) was uploaded anew
    std::vector<uint64_t> gens;
    for (auto& s : p->sets) {
        gens.push_back(s.cde->generation);
        gens.push_back(s.item->upload_gen);
    }
    if (gens == p->generations && p->table.ptr) return JG_OK;
    if (p->table.ptr) {             // launches in flight on any stream may still read the table
        jg_status e = hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        if (e != JG_OK) return e;
    }
    std::vector<uint32_t> mask_words;
    std::vector<size_t> mask_off;
    filter_masks(p->sets, p->filters, p->filter_keys, mask_words, mask_off);
    // the key filters' slots, ascending: the seeds of the sets' collectors
    std::vector<size_t> seed_off(p->sets.size(), 0);
    std::vector<uint32_t> seed_n(p->sets.size(), 0);
    for (size_t i = 0; i < p->sets.size(); i++) {
        if (p->filters[i].kind != JG_FILTER_KEYS) continue;
        const uint32_t words = (uint32_t)((p->sets[i].cde->map.slots.size() + 31) / 32);
        seed_off[i] = mask_words.size();
        std::vector<uint32_t> sl;
        for (uint32_t w = 0; w < words; w++)
            for (uint32_t b = 0; b < 32; b++)
                if ((mask_words[mask_off[i] + w] >> b) & 1u) sl.push_back(w * 32u + b);
        mask_words.insert(mask_words.end(), sl.begin(), sl.end());
        seed_n[i] = (uint32_t)sl.size();
    }
    if (!mask_words.empty()) {
        jg_status st = p->masks.reserve(mask_words.size() * 4);
        if (st != JG_OK) return st;
        st = hip_check(hipMemcpy(p->masks.ptr, mask_words.data(), mask_words.size() * 4, hipMemcpyHostToDevice),
                       "hipMemcpy masks");
        if (st != JG_OK) return st;
    }
    std::vector<ItemView> iv;
    std::vector<FilterView> fv;
    std::vector<std::pair<const uint32_t*, uint32_t>> seeds;
    for (size_t i = 0; i < p->sets.size(); i++) {
        iv.push_back(p->sets[i].item->view);
        fv.push_back(set_filter(p->sets[i], p->filters[i], (const uint32_t*)p->masks.ptr + mask_off[i]));
        seeds.emplace_back(seed_n[i] ? (const uint32_t*)p->masks.ptr + seed_off[i] : nullptr, seed_n[i]);
    }
    if (!p->n_cus) p->n_cus = device_cus();
    std::vector<uint8_t> host;
    jg_status st = table_image(p->sets, iv, fv, seeds, p->cfg, p->n_cus, host, p->L);
    if (st != JG_OK) return st;
    st = p->table.reserve(p->L.total);
    if (st != JG_OK) return st;
    st = hip_check(hipMemcpy(p->table.ptr, host.data(), host.size(), hipMemcpyHostToDevice), "hipMemcpy plan");
    if (st != JG_OK) return st;
    p->generations = gens;
    if (p->gexec) {                 // a captured graph refers to the old table
        (void)hipGraphExecDestroy(p->gexec);
        (void)hipGraphDestroy(p->graph);
        p->gexec = nullptr;
        p->graph = nullptr;
    }
    return JG_OK;
}

jg_status plan_launch(jg_plan* p, const float* d_m6, uint8_t* d_status, hipStream_t stream) {
    return table_launch(p->L, (uint8_t*)p->table.ptr, d_m6, d_status, stream);
}

}  // namespace

namespace jg {
// for other batched forms built on a plan (the placement search): table upload outside any
// graph capture, then launches only
jg_status plan_prepare(jg_plan* p) { return plan_build_table(p); }
jg_status plan_enqueue(jg_plan* p, const float* d_m6, uint8_t* d_status, hipStream_t stream) {
    return plan_launch(p, d_m6, d_status, stream);
}
}  // namespace jg

extern "C" jg_status jg_plan_create(const jg_query_set* sets, uint32_t n_sets, const jg_plan_config* cfg, jg_plan** out) {
    if (!out || (n_sets && !sets)) return fail(JG_ERR_INVALID_ARG, "null argument");
    auto p = std::make_unique<jg_plan>();
    if (cfg) p->cfg = *cfg;
    const uint32_t flags = p->cfg.flags ? p->cfg.flags : (JG_QUERY_SURROGATE | JG_QUERY_POLY);
    for (uint32_t i = 0; i < n_sets; i++) {
        const jg_query_set& s = sets[i];
        if (!s.cde || !s.item) return fail(JG_ERR_INVALID_ARG, "null engine or item in a query set");
        if (s.count > 0xffffffffull) return fail(JG_ERR_INVALID_ARG, "query set too large");
        const Polygon& poly = s.item->poly;
        if (flags & JG_QUERY_SURROGATE) {
            if (!poly.sur.present) return fail(JG_ERR_SURROGATE, "surrogate not generated");
            if (poly.sur.cfg.n_ff_poles > poly.sur.poles.size()) return fail(JG_ERR_SURROGATE, "ff_poles slice out of range");
        }
        if (s.filter.kind > JG_FILTER_HASHED) return fail(JG_ERR_INVALID_ARG, "unknown filter kind");
        p->sets.push_back(s);
        p->filters.push_back(s.filter);
        std::vector<jg_hazkey> keys;
        if (s.filter.kind == JG_FILTER_KEYS && s.filter.n_keys) keys.assign(s.filter.keys, s.filter.keys + s.filter.n_keys);
        p->filter_keys.push_back(std::move(keys));
        p->filters.back().keys = nullptr;
    }
    jg_status st = device_ready();
    if (st != JG_OK) return st;
    st = plan_build_table(p.get());
    if (st != JG_OK) return st;
    *out = p.release();
    return JG_OK;
}

extern "C" void jg_plan_free(jg_plan* p) { delete p; }

extern "C" jg_status jg_plan_detect(jg_plan* p, const float* d_m6, uint8_t* d_status, void* stream) {
    if (!p) return fail(JG_ERR_INVALID_ARG, "null plan");
    if (p->L.n_units == 0) return JG_OK;
    if (!d_m6 || !d_status) return fail(JG_ERR_INVALID_ARG, "null device buffer");
    jg_status st = plan_build_table(p);
    if (st != JG_OK) return st;
    return plan_launch(p, d_m6, d_status, (hipStream_t)stream);
}

extern "C" jg_status jg_plan_detect_graph(jg_plan* p, const float* d_m6, uint8_t* d_status, uint32_t replays, void* stream) {
    if (!p) return fail(JG_ERR_INVALID_ARG, "null plan");
    if (p->L.n_units == 0 || replays == 0) return JG_OK;
    if (!d_m6 || !d_status) return fail(JG_ERR_INVALID_ARG, "null device buffer");
    jg_status st = plan_build_table(p);
    if (st != JG_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    if (!p->gexec || p->g_m6 != d_m6 || p->g_status != d_status || p->g_stream != s) {
        if (p->gexec) {
            (void)hipGraphExecDestroy(p->gexec);
            (void)hipGraphDestroy(p->graph);
            p->gexec = nullptr;
            p->graph = nullptr;
        }
        hipStream_t cap = nullptr;
        st = hip_check(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate");
        if (st != JG_OK) return st;
        st = hip_check(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        if (st == JG_OK) {
            jg_status ls = plan_launch(p, d_m6, d_status, cap);
            hipGraph_t g = nullptr;
            jg_status es = hip_check(hipStreamEndCapture(cap, &g), "hipStreamEndCapture");
            st = ls != JG_OK ? ls : es;
            if (st == JG_OK) {
                p->graph = g;
                st = hip_check(hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
            } else if (g) {
                (void)hipGraphDestroy(g);
            }
        }
        (void)hipStreamDestroy(cap);
        if (st != JG_OK) return st;
        p->g_m6 = d_m6;
        p->g_status = d_status;
        p->g_stream = s;
    }
    for (uint32_t r = 0; r < replays; r++) {
        st = hip_check(hipGraphLaunch(p->gexec, s), "hipGraphLaunch");
        if (st != JG_OK) return st;
    }
    return JG_OK;
}

// ---------------------------------------------------------------------------------------
// The batched detection without any allocation (jg_detect_sets_noalloc): the launch table,
// the query items' records and the JG_FILTER_KEYS masks are staged into `scratch`, a fixed
// device array of JG_NOALLOC_SCRATCH_BYTES that every call on a device shares.  Host code
// never passes the array's name to the device: `scratch` names a host-side shadow, and its
// device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) uint8_t scratch[JG_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[MAX_DEVICES] = {};
uint32_t g_scratch_cus[MAX_DEVICES] = {};

// The device address of this device's `scratch` (looked up on the first call on the device,
// with that device current) and its compute-unit count.
jg_status scratch_of_device(uint8_t** out, uint32_t* cus) {
    int dev = 0;
    jg_status st = hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (st != JG_OK) return st;
    if (dev < 0 || dev >= MAX_DEVICES) return fail(JG_ERR_INTERNAL, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        st = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (st != JG_OK) return st;
        if (!p) return fail(JG_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
        g_scratch_ptr[dev] = (uint8_t*)p;
        g_scratch_cus[dev] = device_cus();
    }
    *out = g_scratch_ptr[dev];
    *cus = g_scratch_cus[dev];
    return JG_OK;
}
}  // namespace

extern "C" jg_status jg_detect_sets_noalloc(const jg_query_set* sets, uint32_t n_sets, const jg_plan_config* cfg,
                                            const float* d_m6, uint8_t* d_status, void* stream) {
    if (n_sets && !sets) return fail(JG_ERR_INVALID_ARG, "null argument");
    jg_plan_config pc{};
    if (cfg) pc = *cfg;
    const uint32_t flags = pc.flags ? pc.flags : (JG_QUERY_SURROGATE | JG_QUERY_POLY);
    std::vector<jg_query_set> vs;
    std::vector<jg_hazard_filter> filters;
    std::vector<std::vector<jg_hazkey>> keys;
    uint64_t total_q = 0;
    for (uint32_t i = 0; i < n_sets; i++) {
        const jg_query_set& s = sets[i];
        if (!s.cde || !s.item) return fail(JG_ERR_INVALID_ARG, "null engine or item in a query set");
        if (s.count > 0xffffffffull) return fail(JG_ERR_INVALID_ARG, "query set too large");
        const Polygon& poly = s.item->poly;
        if (flags & JG_QUERY_SURROGATE) {
            if (!poly.sur.present) return fail(JG_ERR_SURROGATE, "surrogate not generated");
            if (poly.sur.cfg.n_ff_poles > poly.sur.poles.size()) return fail(JG_ERR_SURROGATE, "ff_poles slice out of range");
        }
        if (s.filter.kind > JG_FILTER_HASHED) return fail(JG_ERR_INVALID_ARG, "unknown filter kind");
        // an engine image is uploaded (allocated) by jg_cde_sync, never here
        if (s.cde->dirty || !s.cde->dev_valid)
            return fail(JG_ERR_INVALID_ARG, "engine image not current on the device: call jg_cde_sync first");
        vs.push_back(s);
        filters.push_back(s.filter);
        keys.emplace_back();
        if (s.filter.kind == JG_FILTER_KEYS && s.filter.n_keys) {
            if (!s.filter.keys) return fail(JG_ERR_INVALID_ARG, "null filter keys");
            keys.back().assign(s.filter.keys, s.filter.keys + s.filter.n_keys);
        }
        total_q += s.count;
    }
    if (total_q == 0) return JG_OK;
    if (!d_m6 || !d_status) return fail(JG_ERR_INVALID_ARG, "null device buffer");
    uint8_t* base = nullptr;
    uint32_t cus = 0;
    jg_status st = scratch_of_device(&base, &cus);
    if (st != JG_OK) return st;
    // scratch layout: item records | filter masks | launch table
    std::map<const jg_spolygon*, size_t> item_off;
    size_t off = 0;
    for (auto& s : vs)
        if (item_off.emplace(s.item, off).second) off += item_image_bytes(s.item->poly);
    const size_t o_masks = off;
    std::vector<uint32_t> mask_words;
    std::vector<size_t> mask_off;
    filter_masks(vs, filters, keys, mask_words, mask_off);
    const size_t o_table = (o_masks + 4 * mask_words.size() + 255) & ~(size_t)255;
    std::vector<uint8_t> host(o_table, 0);
    std::vector<ItemView> iv(vs.size());
    for (auto& kv : item_off) {
        ItemView v{};
        item_image_write(kv.first->poly, host.data() + kv.second, base + kv.second, v);
        for (size_t i = 0; i < vs.size(); i++)
            if (vs[i].item == kv.first) iv[i] = v;
    }
    if (!mask_words.empty()) std::memcpy(host.data() + o_masks, mask_words.data(), 4 * mask_words.size());
    std::vector<FilterView> fv;
    for (size_t i = 0; i < vs.size(); i++)
        fv.push_back(set_filter(vs[i], filters[i], (const uint32_t*)(base + o_masks) + mask_off[i]));
    std::vector<uint8_t> table;
    TableLayout L;
    st = table_image(vs, iv, fv, {}, pc, cus, table, L);
    if (st != JG_OK) return st;
    const uint64_t need = (uint64_t)o_table + L.total;
    if (need > JG_NOALLOC_SCRATCH_BYTES)
        return fail(JG_ERR_CAPACITY, "jg_detect_sets_noalloc needs " + std::to_string(need) +
                                         " bytes of scratch; the fixed scratch holds " +
                                         std::to_string((uint64_t)JG_NOALLOC_SCRATCH_BYTES));
    host.resize(o_table + table.size());
    std::memcpy(host.data() + o_table, table.data(), table.size());
    // pageable source: the copy is complete when the call returns, and ordered on `stream`
    hipStream_t hs = (hipStream_t)stream;
    st = hip_check(hipMemcpyAsync(base, host.data(), host.size(), hipMemcpyHostToDevice, hs), "hipMemcpyAsync scratch");
    if (st != JG_OK) return st;
    return table_launch(L, base + o_table, d_m6, d_status, hs);
}

extern "C" jg_status jg_plan_collect(jg_plan* p, const float* d_m6, uint32_t capacity, uint32_t* d_count,
                                     uint64_t* d_hash, uint32_t* d_keys, void* stream) {
    if (!p) return fail(JG_ERR_INVALID_ARG, "null plan");
    for (const jg_hazard_filter& f : p->filters)
        if (f.kind == JG_FILTER_HASHED)
            return fail(JG_ERR_INVALID_ARG, "hazard collection takes no hashed filter (a collector is seeded with keys)");
    for (const jg_query_set& s : p->sets)
        if (!s.item->poly.sur.present || s.item->poly.sur.cfg.n_ff_poles > s.item->poly.sur.poles.size())
            return fail(JG_ERR_SURROGATE, "collection needs every item's surrogate (ff poles)");
    if (p->L.n_units == 0) return JG_OK;
    if (!d_m6 || !d_count || !d_hash) return fail(JG_ERR_INVALID_ARG, "null device buffer");
    jg_status st = plan_build_table(p);
    if (st != JG_OK) return st;
    if ((size_t)p->L.max_slot_words * 4 > 60000) return fail(JG_ERR_CAPACITY, "too many hazard slots for the LDS collector");
    st = p->ovf.reserve(sizeof(uint64_t) * p->L.n_queries);
    if (st != JG_OK) return st;
    return collect_launch(p->L, (uint8_t*)p->table.ptr, d_m6, capacity, d_count, d_hash, d_keys, (uint64_t*)p->ovf.ptr,
                          p->n_cus, (hipStream_t)stream);
}

extern "C" jg_status jg_plan_get_config(const jg_plan* p, jg_plan_config* out) {
    if (!p || !out) return fail(JG_ERR_INVALID_ARG, "null argument");
    *out = p->cfg;
    return JG_OK;
}

namespace {

// median of `reps` timed launches of the plan with its current schedule (after one warm-up)
jg_status plan_time(jg_plan* p, const float* d_m6, uint8_t* d_status, uint32_t reps, hipStream_t s,
                    hipEvent_t a, hipEvent_t b, float& ms) {
    p->generations.clear();                 // rebuild the table with the current schedule
    jg_status st = plan_build_table(p);
    if (st != JG_OK) return st;
    st = plan_launch(p, d_m6, d_status, s);
    if (st != JG_OK) return st;
    std::vector<float> t;
    for (uint32_t r = 0; r < reps; r++) {
        st = hip_check(hipEventRecord(a, s), "hipEventRecord");
        if (st == JG_OK) st = plan_launch(p, d_m6, d_status, s);
        if (st == JG_OK) st = hip_check(hipEventRecord(b, s), "hipEventRecord");
        if (st == JG_OK) st = hip_check(hipEventSynchronize(b), "hipEventSynchronize");
        float x = 0.0f;
        if (st == JG_OK) st = hip_check(hipEventElapsedTime(&x, a, b), "hipEventElapsedTime");
        if (st != JG_OK) return st;
        t.push_back(x);
    }
    std::sort(t.begin(), t.end());
    ms = t[t.size() / 2];
    return JG_OK;
}

}  // namespace

extern "C" jg_status jg_plan_autotune(jg_plan* p, const float* d_m6, uint8_t* d_status, uint32_t reps, void* stream,
                                      jg_plan_config* best, float* best_ms) {
    if (!p) return fail(JG_ERR_INVALID_ARG, "null plan");
    if (!d_m6 || !d_status) return fail(JG_ERR_INVALID_ARG, "null device buffer");
    if (reps == 0) reps = 3;
    if (p->L.n_units == 0) {
        if (best) *best = p->cfg;
        if (best_ms) *best_ms = 0.0f;
        return JG_OK;
    }
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t a = nullptr, b = nullptr;
    jg_status st = hip_check(hipEventCreate(&a), "hipEventCreate");
    if (st == JG_OK) st = hip_check(hipEventCreate(&b), "hipEventCreate");
    jg_plan_config cur = p->cfg;
    float cur_ms = 0.0f;
    if (st == JG_OK) st = plan_time(p, d_m6, d_status, reps, s, a, b, cur_ms);
    const uint32_t cus = p->n_cus ? p->n_cus : 256u;
    const std::vector<uint32_t> lanes = {0u, 8u, 16u, 32u, 64u};
    const std::vector<uint32_t> wgs = {4u * cus, 8u * cus, 16u * cus, 20u * cus, 32u * cus};
    const std::vector<uint32_t> chunks = {16u, 32u, 64u};
    // the whole grid (75 two-kernel and 75 fused schedules): timings are noisy enough that
    // coordinate descent can stop in a poor corner
    for (uint32_t sv : {2u, 1u})
    for (uint32_t lv : lanes)
        for (uint32_t wv : wgs)
            for (uint32_t cv : chunks) {
                if (st != JG_OK) break;
                jg_plan_config c = p->cfg;
                c.stages = sv;
                c.lanes_per_query = lv;
                c.workgroups = wv;
                c.chunk = cv;
                p->cfg = c;
                float ms = 0.0f;
                st = plan_time(p, d_m6, d_status, reps, s, a, b, ms);
                if (st == JG_OK && ms < cur_ms) {
                    cur = c;
                    cur_ms = ms;
                }
            }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    p->cfg = cur;
    p->generations.clear();
    jg_status st2 = plan_build_table(p);
    if (st == JG_OK) st = st2;
    if (st == JG_OK) st = plan_launch(p, d_m6, d_status, s);   // d_status from the kept schedule
    if (st != JG_OK) return st;
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return JG_OK;
}
