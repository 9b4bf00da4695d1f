// C ABI of the collision engine (include/jagua_amd.h): handle management, the HBM image of
// each engine, and the plain (one query set per launch) detect / collect kernels.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/jagua_amd.h"
#include "jg_device.hpp"
#include "jg_host.hpp"
#include "jg_internal.hpp"
#include "jg_plan.hpp"

using namespace jg;

// ------------------------------------------------------------------------------ errors
namespace jg {
thread_local std::string g_last_error;
jg_status fail(jg_status code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
jg_status hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return JG_OK;
    return fail(JG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace jg

#define HIP_TRY(expr)                                        \
    do {                                                     \
        jg_status _s = jg::hip_check((expr), #expr);         \
        if (_s != JG_OK) return _s;                          \
    } while (0)

extern "C" const char* jg_last_error(void) { return g_last_error.c_str(); }
extern "C" const char* jg_version(void) { return "jagua-rs_amd 0.1.0 (gfx950)"; }

// ------------------------------------------------------------------------------ device blob
namespace jg {

DevBlob::~DevBlob() {
    if (ptr) (void)hipFree(ptr);
}
jg_status DevBlob::reserve(size_t bytes) {
    if (bytes <= cap) return JG_OK;
    if (ptr) {
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipFree(ptr));
        ptr = nullptr;
        cap = 0;
    }
    size_t c = bytes < 4096 ? 4096 : bytes + bytes / 4;
    HIP_TRY(hipMalloc(&ptr, c));
    cap = c;
    return JG_OK;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

jg_status device_ready() {
    static int ok = -1;
    if (ok < 0) {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        ok = (e == hipSuccess && n > 0) ? 1 : 0;
    }
    if (!ok) return fail(JG_ERR_HIP, "no HIP device available: the collision engine runs on the GPU only");
    return JG_OK;
}

size_t item_image_bytes(const Polygon& p) {
    const uint32_t n_ff = p.sur.present ? std::min<uint32_t>(p.sur.cfg.n_ff_poles, (uint32_t)p.sur.poles.size()) : 0;
    const uint32_t n_pi = p.sur.present ? (uint32_t)p.sur.piers.size() : 0;
    return align256(align256(sizeof(float2) * p.n()) + sizeof(float4) * (n_ff + n_pi + 1));
}

// The device record of a query item (vertices, then the ff poles and piers) written into
// `host`, whose bytes are copied to device address `dev`; `view` points at `dev`.
void item_image_write(const Polygon& p, uint8_t* host, uint8_t* dev, ItemView& view) {
    const uint32_t n_ff = p.sur.present ? std::min<uint32_t>(p.sur.cfg.n_ff_poles, (uint32_t)p.sur.poles.size()) : 0;
    const uint32_t n_pi = p.sur.present ? (uint32_t)p.sur.piers.size() : 0;
    const size_t off_ff = align256(sizeof(float2) * p.n());
    std::memset(host, 0, item_image_bytes(p));
    std::memcpy(host, p.v.data(), sizeof(float2) * p.n());
    float4* ff = (float4*)(host + off_ff);
    for (uint32_t i = 0; i < n_ff; i++) ff[i] = p.sur.poles[i];
    for (uint32_t i = 0; i < n_pi; i++) ff[n_ff + i] = p.sur.piers[i];
    view.v = (const float2*)dev;
    view.ff = (const float4*)(dev + off_ff);
    view.poi = p.poi;
    view.n_v = p.n();
    view.n_poles = n_ff;
    view.n_piers = n_pi;
}

jg_status item_upload(jg_spolygon* it) {
    if (it->dev_valid) return JG_OK;
    jg_status s = device_ready();
    if (s != JG_OK) return s;
    const size_t total = item_image_bytes(it->poly);
    std::vector<uint8_t> host(total, 0);
    s = it->dev.reserve(total);
    if (s != JG_OK) return s;
    item_image_write(it->poly, host.data(), (uint8_t*)it->dev.ptr, it->view);
    HIP_TRY(hipMemcpy(it->dev.ptr, host.data(), total, hipMemcpyHostToDevice));
    it->dev_valid = true;
    it->upload_gen++;
    return JG_OK;
}

// Copies the patch segments of the staging buffer into the image, one workgroup per segment.
__global__ __launch_bounds__(256) void k_scatter(uint8_t* image, const ImageSeg* segs, uint32_t n_segs,
                                                 const uint8_t* data) {
    for (uint32_t s = blockIdx.x; s < n_segs; s += gridDim.x) {
        const ImageSeg g = segs[s];
        const uint32_t* src = (const uint32_t*)(data + g.src);
        uint32_t* dst = (uint32_t*)(image + g.dst);
        const uint64_t words = g.bytes >> 2;
        for (uint64_t w = threadIdx.x; w < words; w += blockDim.x) dst[w] = src[w];
    }
}

static void set_view(jg_cde* c) {
    const PoolImage& im = c->img;
    uint8_t* b = (uint8_t*)c->dev.ptr;
    CdeView& v = c->view;
    v.nodes = (const DevNode*)(b + im.o_nd);
    v.node_up = (const uint32_t*)(b + im.o_nu);
    v.ent = (const DevEntry*)(b + im.o_en);
    v.edges = (const float4*)(b + im.o_ed);
    v.haz_bbox = (const float4*)(b + im.o_hb);
    v.haz_poi = (const float4*)(b + im.o_hp);
    v.haz_vr = (const uint2*)(b + im.o_hv);
    v.haz_verts = (const float2*)(b + im.o_vv);
    v.cells = (const uint32_t*)(b + im.o_ce);
    v.cell_level = im.cell_level;
    v.n_slots = (uint32_t)im.slot_ver.size();
    v.exterior_slot = c->exterior.idx;
    v.cd_threshold = c->cfg.cd_threshold;
    v.n_nodes = (uint32_t)im.nodes.size();
    v.ext_inner = make_float4(1.0f, 1.0f, 0.0f, 0.0f);
    if (c->map.valid(c->exterior)) {
        const Polygon& ex = c->map.slots[c->exterior.idx].rec.shape;
        bool rect = ex.n() == 4;
        for (uint32_t i = 0; i < 4 && rect; i++) {
            const float4 e = ex.edge(i);
            rect = e.x == e.z || e.y == e.w;
        }
        if (rect) v.ext_inner = ex.bbox;
    }
    c->stats.n_nodes = im.nodes.size();
    c->stats.n_entries = im.ent.size();
    c->stats.n_edge_refs = im.edges.size();
    c->stats.n_hazard_vertices = im.verts.size();
    c->stats.n_slots = im.slot_ver.size();
    c->stats.device_bytes = im.total;
}

// Applies the changes since the last upload in place; false when a full upload is needed.
static jg_status cde_patch(jg_cde* c, bool& done) {
    done = false;
    if (!c->dev_valid || !c->img.valid || c->dev.cap < c->img.total) return JG_OK;
    auto t0 = std::chrono::steady_clock::now();
    if (!image_patch(c->img, c->qt, c->map, c->staging, c->segs)) {
        c->img.valid = false;
        return JG_OK;
    }
    c->img.valid = false;                  // until the device holds the patched image
    const uint64_t n_segs = c->segs.size();
    const uint64_t seg_bytes = (sizeof(ImageSeg) * n_segs + 255) & ~(uint64_t)255;
    const uint64_t data_bytes = c->staging.size();
    for (const ImageSeg& g : c->segs)
        if (g.dst + g.bytes > c->img.total || g.src + g.bytes > data_bytes || (g.dst | g.src | g.bytes) & 3)
            return fail(JG_ERR_INTERNAL, "image patch segment out of range");
    if (n_segs) {
        jg_status s = c->staging_dev.reserve(seg_bytes + data_bytes);
        if (s != JG_OK) return s;
        HIP_TRY(hipDeviceSynchronize());   // earlier launches may still read the image
        uint8_t* sd = (uint8_t*)c->staging_dev.ptr;
        HIP_TRY(hipMemcpy(sd, c->segs.data(), sizeof(ImageSeg) * n_segs, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(sd + seg_bytes, c->staging.data(), data_bytes, hipMemcpyHostToDevice));
        const uint32_t grid = (uint32_t)(n_segs < 8192 ? n_segs : 8192);
        hipLaunchKernelGGL(k_scatter, dim3(grid), dim3(256), 0, 0, (uint8_t*)c->dev.ptr, (const ImageSeg*)sd,
                           (uint32_t)n_segs, (const uint8_t*)(sd + seg_bytes));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(0));
    }
    c->img.valid = true;
    set_view(c);
    auto t1 = std::chrono::steady_clock::now();
    c->stats.patch_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    c->stats.patch_bytes = data_bytes;
    c->stats.n_patches++;
    done = true;
    return JG_OK;
}

jg_status cde_upload(jg_cde* c) {
    if (!c->dirty && c->dev_valid) return JG_OK;
    jg_status s = device_ready();
    if (s != JG_OK) return s;
    bool done = false;
    s = cde_patch(c, done);
    if (s != JG_OK) return s;
    if (!done) {
        auto t0 = std::chrono::steady_clock::now();
        PoolImage& im = c->img;
        image_build(im, c->qt, c->map);
        auto t1 = std::chrono::steady_clock::now();
        std::vector<uint8_t> host(im.total, 0);
        auto put = [&](uint64_t o, const void* src, uint64_t bytes) {
            if (bytes) std::memcpy(host.data() + o, src, bytes);
        };
        put(im.o_nd, im.nodes.data(), sizeof(FlatNode) * im.nodes.size());
        put(im.o_nu, im.up.data(), sizeof(uint32_t) * im.up.size());
        put(im.o_en, im.ent.data(), sizeof(FlatEntry) * im.ent.size());
        put(im.o_ed, im.edges.data(), sizeof(float4) * im.edges.size());
        put(im.o_hb, im.hbb.data(), sizeof(float4) * im.hbb.size());
        put(im.o_hp, im.hpoi.data(), sizeof(float4) * im.hpoi.size());
        put(im.o_hv, im.hvr.data(), sizeof(uint2) * im.hvr.size());
        put(im.o_vv, im.verts.data(), sizeof(float2) * im.verts.size());
        put(im.o_ce, im.cells.data(), sizeof(uint32_t) * im.cells.size());
        if (c->dev.ptr) HIP_TRY(hipDeviceSynchronize());   // earlier launches may still read the old image
        s = c->dev.reserve(im.total);
        if (s != JG_OK) return s;
        HIP_TRY(hipMemcpy(c->dev.ptr, host.data(), im.total, hipMemcpyHostToDevice));
        auto t2 = std::chrono::steady_clock::now();
        set_view(c);
        c->stats.flatten_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        c->stats.upload_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
        c->stats.n_full_uploads++;
    }
    c->dirty = false;
    c->dev_valid = true;
    c->generation++;
    return JG_OK;
}

// JG_FILTER_KEYS masks are uploaded into a per-engine buffer.  The launch that last read it may
// still run on another stream (a non-blocking one does not wait for the null stream), so the
// buffer is overwritten only after that launch's completion event; the copy is then ordered on
// the launch stream.  filter_read_done records the event after a launch that reads the mask.
jg_status filter_view(jg_cde* c, const jg_hazard_filter* f, FilterView& out, hipStream_t stream) {
    out = FilterView{nullptr, 0, 0, 0, 0};
    if (!f || f->kind == JG_FILTER_NONE) return JG_OK;
    if (f->kind == JG_FILTER_HASHED) {
        out.kind = 2;
        out.seed = f->seed;
        out.threshold = f->threshold;
        return JG_OK;
    }
    if (f->kind != JG_FILTER_KEYS) return fail(JG_ERR_INVALID_ARG, "unknown filter kind");
    const uint32_t words = (uint32_t)((c->map.slots.size() + 31) / 32);
    std::vector<uint32_t> mask(words, 0);
    for (uint32_t i = 0; i < f->n_keys; i++) {
        const jg_hazkey k = f->keys[i];
        if (c->map.valid(k)) mask[k.idx >> 5] |= 1u << (k.idx & 31u);
    }
    if (c->filter_read) HIP_TRY(hipEventSynchronize(c->filter_read));
    jg_status s = c->filter_dev.reserve(sizeof(uint32_t) * words);
    if (s != JG_OK) return s;
    // pageable source: the copy is complete when the call returns, and ordered on `stream`
    HIP_TRY(hipMemcpyAsync(c->filter_dev.ptr, mask.data(), sizeof(uint32_t) * words, hipMemcpyHostToDevice, stream));
    out.kind = 1;
    out.mask = (const uint32_t*)c->filter_dev.ptr;
    out.n_words = words;
    return JG_OK;
}

jg_status filter_read_done(jg_cde* c, hipStream_t stream) {
    if (!c->filter_read) HIP_TRY(hipEventCreateWithFlags(&c->filter_read, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(c->filter_read, stream));
    return JG_OK;
}

}  // namespace jg

// ------------------------------------------------------------------------------ kernels

// One collision query per thread: surrogate fail-fast (cd_engine.rs:204-223), then the
// polygon query (cd_engine.rs:160-197) from the virtual root, then containment.
__global__ __launch_bounds__(256) void k_detect_plain(CdeView c, ItemView it, const float* __restrict__ m6, uint64_t n,
                                                      FilterView f, uint32_t flags, uint8_t* __restrict__ status) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const Mat m = load_mat(m6, q);
    if (flags & JG_QUERY_SURROGATE) {
        for (uint32_t i = 0; i < it.n_poles; i++) {
            const float4 p = it.ff[i];
            const float2 cc = xform(m, p.x, p.y);
            if (collides_circle(c, 0, cc.x, cc.y, p.z, f, q)) { status[q] = JG_Q_COLLISION; return; }
        }
        for (uint32_t i = 0; i < it.n_piers; i++) {
            const float4 e = it.ff[it.n_poles + i];
            const float2 a = xform(m, e.x, e.y), b = xform(m, e.z, e.w);
            if (collides_edge(c, 0, make_float4(a.x, a.y, b.x, b.y), f, q)) { status[q] = JG_Q_COLLISION; return; }
        }
    }
    if (!(flags & JG_QUERY_POLY)) { status[q] = JG_Q_CLEAR; return; }
    const float4 qbb = query_bbox(it, m);
    if (box_relation(c.nodes[0].bbox, qbb) != REL_SURROUNDING) { status[q] = JG_Q_COLLISION; return; }
    const int vr = virtual_root(c, qbb);
    float2 a = xform(m, it.v[0].x, it.v[0].y);
    const float2 first = a;
    for (uint32_t i = 0; i < it.n_v; i++) {
        const float2 b = i + 1 == it.n_v ? first : xform(m, it.v[i + 1].x, it.v[i + 1].y);
        if (collides_edge(c, vr, make_float4(a.x, a.y, b.x, b.y), f, q)) { status[q] = JG_Q_COLLISION; return; }
        a = b;
    }
    const DevNode vn = c.nodes[vr];
    const float2 qpoi = xform(m, it.poi.x, it.poi.y);
    const int he = (int)(vn.haz_begin + (vn.haz_info & COUNT_MASK));
    for (int i = (int)vn.haz_begin; i < he; i++) {
        const uint32_t sf = c.ent[i].slot_flags;
        if (sf & ENTIRE_BIT) { status[q] = JG_Q_PANIC_FILTERED_ENTIRE; return; }
        const uint32_t slot = sf & SLOT_MASK;
        if (irrelevant(f, q, slot)) continue;
        if (containment(c, slot, it, m, qbb, qpoi)) { status[q] = JG_Q_COLLISION; return; }
    }
    status[q] = JG_Q_CLEAR;
}

namespace {

// The collector of one query: a bitmask over hazard slots (union semantics) plus the list of
// slots in insertion order, so that results and the reset cost O(set size), not O(slots).
constexpr uint32_t SET_LIST_CAP = 256;
struct SlotSet {
    uint32_t* bits;
    uint32_t* list;                 // [SET_LIST_CAP]
    uint32_t* n;                    // slots inserted (may exceed SET_LIST_CAP: then bits is scanned)
};
__device__ __forceinline__ bool set_has(const SlotSet& set, uint32_t slot) {
    return (set.bits[slot >> 5] >> (slot & 31u)) & 1u;
}
__device__ __forceinline__ void set_put(const SlotSet& set, uint32_t slot) {
    const uint32_t bit = 1u << (slot & 31u);
    if (atomicOr(&set.bits[slot >> 5], bit) & bit) return;
    const uint32_t at = atomicAdd(set.n, 1u);
    if (at < SET_LIST_CAP) set.list[at] = slot;
}

// Hazard tests of QTNode::collect_collisions at a node where detection is performed now
// (qt_node.rs:174-189), for a circle or an edge.
template <bool CIRCLE>
__device__ __forceinline__ void collect_now(const CdeView& c, const DevNode& nd, float4 ent_geom, float r,
                                            const SlotSet& set) {
    const int he = (int)(nd.haz_begin + (nd.haz_info & COUNT_MASK));
    for (int i = (int)nd.haz_begin; i < he; i++) {
        const DevEntry e = c.ent[i];
        const uint32_t slot = e.slot_flags & SLOT_MASK;
        if (set_has(set, slot)) continue;
        if (e.slot_flags & ENTIRE_BIT) { set_put(set, slot); continue; }
        bool hit = false;
        if (CIRCLE) {
            if (circle_box(ent_geom.x, ent_geom.y, r, e.ff))
                for (uint32_t j = 0; j < e.edge_count && !hit; j++) hit = circle_edge(ent_geom.x, ent_geom.y, r, c.edges[e.edge_begin + j]);
        } else {
            if (box_edge(e.ff, ent_geom))
                for (uint32_t j = 0; j < e.edge_count && !hit; j++) hit = edge_edge(ent_geom, c.edges[e.edge_begin + j]);
        }
        if (hit) set_put(set, slot);
    }
}

// QTNode::collect_collisions (qt_node.rs:151-191)
template <bool CIRCLE>
__device__ __forceinline__ void collect_from(const CdeView& c, int start, float4 g, float r, const SlotSet& set) {
    descend(c, start, [&](int n, int& fc) -> uint32_t {
        const DevNode nd = c.nodes[n];
        const bool now = nd.n_active_edges <= c.cd_threshold;
        if (nd.first_child >= 0 && !now) {
            fc = nd.first_child;
            const float4 nb = nd.bbox;
            if (CIRCLE) return circle_quadrants(g.x, g.y, r, nb);
            return edge_quadrants_node(g, nb);
        }
        collect_now<CIRCLE>(c, nd, g, r, set);
        return 0u;
    });
}

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, 64);
    return v;
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, o, 64);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), o, 64);
        v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}

// One hazard-collection query (ci_bench.rs:56-69: collect_surrogate_collisions then
// collect_poly_collisions, cd_engine.rs:261-310) by one wavefront, whose workgroup is that
// wavefront alone, so that the workgroup barrier orders the lanes' LDS accesses.  The collector
// is a SlotSet in LDS (union semantics of BasicHazardCollector, collector.rs:44-67), empty
// between queries; it starts with the n_seed slots of `seed` (a pre-seeded collector: the
// result is the seed together with the hits).  Lanes take the surrogate poles/piers and the
// query edges, each collecting its descent (collect_from); containment checks follow.  The
// count, the order-independent hash and the first `cap` slots ascending are written to
// count[out], hash[out], keys[out * cap ..].
__device__ __forceinline__ void collect_query(const CdeView& c, const ItemView& it, const Mat& m, uint64_t out,
                                              const uint32_t* seed, uint32_t n_seed, uint32_t words, uint32_t cap,
                                              uint32_t* lds, uint32_t* list, uint32_t* list_n,
                                              uint32_t* __restrict__ count, uint64_t* __restrict__ hash,
                                              uint32_t* __restrict__ keys, uint32_t lane) {
    const SlotSet set{lds, list, list_n};
    for (uint32_t i = lane; i < n_seed; i += 64) set_put(set, seed[i]);
    const uint32_t n_sur = it.n_poles + it.n_piers;
    for (uint32_t i = lane; i < n_sur; i += 64) {
        const float4 p = it.ff[i];
        if (i < it.n_poles) {
            const float2 cc = xform(m, p.x, p.y);
            collect_from<true>(c, 0, make_float4(cc.x, cc.y, 0, 0), p.z, set);
        } else {
            const float2 a = xform(m, p.x, p.y), b = xform(m, p.z, p.w);
            collect_from<false>(c, 0, make_float4(a.x, a.y, b.x, b.y), 0.0f, set);
        }
    }
    float4 bb = make_float4(3.40282347e+38f, 3.40282347e+38f, -3.40282347e+38f, -3.40282347e+38f);
    for (uint32_t i = lane; i < it.n_v; i += 64) {
        const float2 p = xform(m, it.v[i].x, it.v[i].y);
        bb.x = mn(bb.x, p.x); bb.y = mn(bb.y, p.y); bb.z = mx(bb.z, p.x); bb.w = mx(bb.w, p.y);
    }
    const float4 qbb = make_float4(wave_min(bb.x), wave_min(bb.y), wave_max(bb.z), wave_max(bb.w));
    if (box_relation(c.nodes[0].bbox, qbb) != REL_SURROUNDING && lane == 0) set_put(set, c.exterior_slot);
    const int vr = virtual_root(c, qbb);
    for (uint32_t i = lane; i < it.n_v; i += 64) {
        const float2 a = xform(m, it.v[i].x, it.v[i].y);
        const uint32_t j = i + 1 == it.n_v ? 0 : i + 1;
        const float2 b = xform(m, it.v[j].x, it.v[j].y);
        collect_from<false>(c, vr, make_float4(a.x, a.y, b.x, b.y), 0.0f, set);
    }
    __syncthreads();
    {
        // containment over the virtual root's Partial hazards not yet collected
        // (cd_engine.rs:276-287): each lane filters one candidate by bbox relation (no
        // relation: collides iff exterior scope); the point-in-polygon tests left run on
        // all 64 lanes, each polygon's edges split over the wave (containment()'s result)
        const DevNode vn = c.nodes[vr];
        const float2 qpoi = xform(m, it.poi.x, it.poi.y);
        const uint32_t e0 = vn.haz_begin, cnt = vn.haz_info & COUNT_MASK;
        for (uint32_t k0 = 0; k0 < cnt; k0 += 64u) {
            const uint32_t i = k0 + lane;
            uint32_t slot = 0;
            int rel = REL_DISJOINT;
            bool pend = false;
            if (i < cnt) {
                const uint32_t sf = c.ent[e0 + i].slot_flags;
                slot = sf & SLOT_MASK;
                if (!(sf & ENTIRE_BIT) && !set_has(set, slot)) {
                    rel = box_almost_relation(c.haz_bbox[slot], qbb);
                    if (rel == REL_SURROUNDING || rel == REL_ENCLOSED) pend = true;
                    else if (c.haz_poi[slot].w == 0.0f) set_put(set, slot);
                }
            }
            for (uint64_t pm = __ballot(pend); pm; pm &= pm - 1ull) {
                const int l = __builtin_ctzll(pm);
                const uint32_t s = (uint32_t)__shfl((int)slot, l, 64);
                const int rl = __shfl(rel, l, 64);
                bool contained;
                if (rl == REL_SURROUNDING) {
                    contained = pip_hazard_wave(c, s, qpoi.x, qpoi.y, lane);
                } else {
                    const float4 hp = c.haz_poi[s];
                    contained = pip_query_wave(it, m, qbb, hp.x, hp.y, lane);
                }
                if (lane == (uint32_t)l && (c.haz_poi[s].w != 0.0f ? contained : !contained)) set_put(set, s);
            }
        }
    }
    __syncthreads();
    // results: count, order-independent hash, the first `cap` slots ascending
    const uint32_t total = *list_n;
    uint32_t cnt = 0;
    uint64_t h = 0;
    if (total <= SET_LIST_CAP) {
        for (uint32_t i = lane; i < total; i += 64) {
            const uint32_t si = list[i];
            h += mix64((uint64_t)si);
            if (keys && cap) {
                uint32_t rank = 0;
                for (uint32_t j = 0; j < total; j++) rank += list[j] < si ? 1u : 0u;
                if (rank < cap) keys[out * cap + rank] = si;
            }
        }
        cnt = total;
        if (keys)
            for (uint32_t k = total + lane; k < cap; k += 64) keys[out * cap + k] = 0xffffffffu;
    } else {
        for (uint32_t w = lane; w < words; w += 64) {
            uint32_t bits = lds[w];
            cnt += (uint32_t)__builtin_popcount(bits);
            while (bits) {
                const uint32_t b = (uint32_t)__builtin_ctz(bits);
                bits &= bits - 1;
                h += mix64((uint64_t)(w * 32u + b));
            }
        }
        cnt = wave_sum_u32(cnt);
        if (lane == 0 && keys && cap) {
            uint32_t k = 0;
            for (uint32_t w = 0; w < words && k < cap; w++) {
                uint32_t bits = lds[w];
                while (bits && k < cap) {
                    const uint32_t b = (uint32_t)__builtin_ctz(bits);
                    bits &= bits - 1;
                    keys[out * cap + k++] = w * 32u + b;
                }
            }
            for (; k < cap; k++) keys[out * cap + k] = 0xffffffffu;
        }
    }
    h = wave_sum_u64(h);
    if (lane == 0) {
        count[out] = cnt;
        hash[out] = h;
    }
    __syncthreads();
    // reset the collector for the next query
    if (total <= SET_LIST_CAP) {
        for (uint32_t i = lane; i < total; i += 64) lds[list[i] >> 5] = 0u;
    } else {
        for (uint32_t w = lane; w < words; w += 64) lds[w] = 0u;
    }
    if (lane == 0) *list_n = 0u;
    __syncthreads();
}

}  // namespace

// Hazard-collection queries of one item, one at a time per wavefront (persistent one-wavefront
// workgroups).  A wave traversal with shared queues, as the batched forms use, measured slower
// here without the subtree skip of the batched collection: collection descents start at deep
// virtual roots.
__global__ __launch_bounds__(64) void k_collect(CdeView c, ItemView it, const float* __restrict__ m6, uint64_t n,
                                                uint32_t words, uint32_t cap, uint32_t* __restrict__ count,
                                                uint64_t* __restrict__ hash, uint32_t* __restrict__ keys) {
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t list[SET_LIST_CAP];
    __shared__ uint32_t list_n;
    const uint32_t lane = threadIdx.x;
    for (uint32_t w = lane; w < words; w += 64) lds[w] = 0u;
    if (lane == 0) list_n = 0u;
    __syncthreads();
    for (uint64_t q = blockIdx.x; q < n; q += gridDim.x)
        collect_query(c, it, load_mat(m6, q), q, nullptr, 0u, words, cap, lds, list, &list_n, count, hash, keys, lane);
}

// The queries the batched collection could not hold in its per-query collectors (more slots
// than it keeps, counted at counters[C_OVERFLOW]): (set << 32 | query in set) entries of ovf,
// each collected here by one wave with the per-query collector over every slot.
__global__ __launch_bounds__(64) void k_collect_overflow(const SetDesc* __restrict__ sets,
                                                         uint32_t* __restrict__ counters,
                                                         const uint64_t* __restrict__ ovf, const float* __restrict__ m6,
                                                         uint32_t words, uint32_t cap, uint32_t* __restrict__ count,
                                                         uint64_t* __restrict__ hash, uint32_t* __restrict__ keys) {
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t list[SET_LIST_CAP];
    __shared__ uint32_t list_n;
    const uint32_t lane = threadIdx.x;
    for (uint32_t w = lane; w < words; w += 64) lds[w] = 0u;
    if (lane == 0) list_n = 0u;
    __syncthreads();
    const uint32_t n = counters[C_OVERFLOW];
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const uint64_t e = ovf[i];
        const SetDesc& sd = sets[(uint32_t)(e >> 32)];
        const uint64_t out = sd.first + (uint32_t)e;
        const uint32_t sw = (sd.cde.n_slots + 31u) / 32u;
        collect_query(sd.cde, sd.item, load_mat(m6, out), out, sd.seed, sd.n_seed, sw, cap, lds, list, &list_n, count,
                      hash, keys, lane);
    }
    // the collection's last launch: its last wave leaves every claim counter at zero, as the
    // table's other launch sequences expect
    if (last_wave_out(counters, lane))
        for (uint32_t i = lane; i < N_COUNTER_WORDS; i += 64u)
            if (i != C_DONE) counters[i] = 0u;
}

namespace jg {
void collect_overflow_launch(const SetDesc* sets, uint32_t* counters, const uint64_t* ovf, const float* m6,
                             uint32_t words, uint32_t cap, uint32_t* count, uint64_t* hash, uint32_t* keys,
                             uint32_t grid, hipStream_t stream) {
    hipLaunchKernelGGL(k_collect_overflow, dim3(grid), dim3(64), (size_t)words * 4, stream, sets, counters, ovf, m6,
                       words, cap, count, hash, keys);
}
}  // namespace jg

// ------------------------------------------------------------------------------ ABI: graphs
struct jg_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~jg_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

extern "C" jg_status jg_graph_begin(void* stream) {
    if (!stream) return fail(JG_ERR_INVALID_ARG, "capture needs a stream other than the null stream");
    jg_status s = device_ready();
    if (s != JG_OK) return s;
    return hip_check(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal),
                     "hipStreamBeginCapture");
}

extern "C" jg_status jg_graph_end(void* stream, jg_graph** out) {
    if (!stream || !out) return fail(JG_ERR_INVALID_ARG, "null argument");
    hipGraph_t g = nullptr;
    jg_status s = hip_check(hipStreamEndCapture((hipStream_t)stream, &g), "hipStreamEndCapture");
    if (s != JG_OK) {
        if (g) (void)hipGraphDestroy(g);
        return s;
    }
    auto r = std::make_unique<jg_graph>();
    r->graph = g;
    s = hip_check(hipGraphInstantiate(&r->exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (s != JG_OK) return s;
    *out = r.release();
    return JG_OK;
}

extern "C" jg_status jg_graph_launch(const jg_graph* g, void* stream) {
    if (!g || !g->exec) return fail(JG_ERR_INVALID_ARG, "null graph");
    return hip_check(hipGraphLaunch(g->exec, (hipStream_t)stream), "hipGraphLaunch");
}

extern "C" void jg_graph_free(jg_graph* g) { delete g; }

// ------------------------------------------------------------------------------ ABI: polygons
using namespace jg;

extern "C" jg_status jg_spolygon_new(const float* xy, uint32_t n, jg_spolygon** out) {
    if (!xy || !out) return fail(JG_ERR_INVALID_ARG, "null argument");
    auto p = std::make_unique<jg_spolygon>();
    std::string err;
    if (!polygon_new(xy, n, p->poly, err)) return fail(JG_ERR_POLYGON, err);
    *out = p.release();
    return JG_OK;
}

extern "C" jg_status jg_spolygon_generate_surrogate(jg_spolygon* p, const jg_surrogate_config* cfg) {
    if (!p || !cfg) return fail(JG_ERR_INVALID_ARG, "null argument");
    std::string err;
    if (!polygon_generate_surrogate(p->poly, *cfg, err)) return fail(JG_ERR_SURROGATE, err);
    p->dev_valid = false;
    return JG_OK;
}

extern "C" jg_status jg_spolygon_transform_clone(const jg_spolygon* p, const float dt[3], jg_spolygon** out) {
    if (!p || !dt || !out) return fail(JG_ERR_INVALID_ARG, "null argument");
    if (dt[0] != dt[0] || dt[1] != dt[1] || dt[2] != dt[2]) return fail(JG_ERR_INVALID_ARG, "NaN transformation");
    auto r = std::make_unique<jg_spolygon>();
    r->poly = polygon_transform(p->poly, mat_from_dt(dt[0], dt[1], dt[2]));
    *out = r.release();
    return JG_OK;
}

extern "C" void jg_spolygon_free(jg_spolygon* p) { delete p; }

extern "C" jg_status jg_spolygon_get_info(const jg_spolygon* p, jg_spolygon_info* o) {
    if (!p || !o) return fail(JG_ERR_INVALID_ARG, "null argument");
    const Polygon& s = p->poly;
    o->n_vertices = s.n();
    o->n_poles = s.sur.present ? (uint32_t)s.sur.poles.size() : 0;
    o->n_piers = s.sur.present ? (uint32_t)s.sur.piers.size() : 0;
    o->n_ff_poles = s.sur.present ? s.sur.cfg.n_ff_poles : 0;
    o->area = s.area;
    o->diameter = s.diameter;
    o->bbox[0] = s.bbox.x; o->bbox[1] = s.bbox.y; o->bbox[2] = s.bbox.z; o->bbox[3] = s.bbox.w;
    o->poi[0] = s.poi.x; o->poi[1] = s.poi.y; o->poi[2] = s.poi.z;
    o->convex_hull_area = s.sur.hull_area;
    return JG_OK;
}
extern "C" jg_status jg_spolygon_get_vertices(const jg_spolygon* p, float* xy) {
    if (!p || !xy) return fail(JG_ERR_INVALID_ARG, "null argument");
    std::memcpy(xy, p->poly.v.data(), sizeof(float2) * p->poly.n());
    return JG_OK;
}
extern "C" jg_status jg_spolygon_get_poles(const jg_spolygon* p, float* out) {
    if (!p || !out) return fail(JG_ERR_INVALID_ARG, "null argument");
    for (size_t i = 0; i < p->poly.sur.poles.size(); i++) {
        out[3 * i] = p->poly.sur.poles[i].x; out[3 * i + 1] = p->poly.sur.poles[i].y; out[3 * i + 2] = p->poly.sur.poles[i].z;
    }
    return JG_OK;
}
extern "C" jg_status jg_spolygon_get_piers(const jg_spolygon* p, float* out) {
    if (!p || !out) return fail(JG_ERR_INVALID_ARG, "null argument");
    std::memcpy(out, p->poly.sur.piers.data(), sizeof(float4) * p->poly.sur.piers.size());
    return JG_OK;
}

extern "C" jg_status jg_transforms_from_dt(const float* dt, uint64_t n, float* out) {
    if ((!dt || !out) && n) return fail(JG_ERR_INVALID_ARG, "null argument");
    for (uint64_t i = 0; i < n; i++) {
        const Mat m = mat_from_dt(dt[3 * i], dt[3 * i + 1], dt[3 * i + 2]);
        float* o = out + 6 * i;
        o[0] = m.a; o[1] = m.b; o[2] = m.c; o[3] = m.d; o[4] = m.e; o[5] = m.f;
    }
    return JG_OK;
}

extern "C" jg_status jg_dt_chain(const float a[3], uint32_t invert_a, const float b[3], float out[3]) {
    if (!a || !b || !out) return fail(JG_ERR_INVALID_ARG, "null argument");
    if (!dt_chain(a, invert_a != 0, b, out)) return fail(JG_ERR_INVALID_ARG, "transformation is NaN");
    return JG_OK;
}

extern "C" jg_status jg_spolygon_centroid(const jg_spolygon* p, float out[2]) {
    if (!p || !out) return fail(JG_ERR_INVALID_ARG, "null argument");
    const float2 c = polygon_centroid(p->poly);
    out[0] = c.x;
    out[1] = c.y;
    return JG_OK;
}

// ------------------------------------------------------------------------------ ABI: engine
static jg_hazkey cde_register(jg_cde* c, const jg_hazard_entity& ent, const Polygon& shape, bool dynamic) {
    HazardRec r;
    r.entity = ent;
    r.shape = shape;
    r.dynamic = dynamic;
    const jg_hazkey k = c->map.insert(std::move(r));
    c->qt.register_hazard(0, root_entry(c->map.slots[k.idx].rec.shape, k), c->map);
    c->dirty = true;
    return k;
}

extern "C" jg_status jg_cde_new(const float bbox[4], const jg_cde_config* cfg, const jg_hazard_entity* ents,
                                const jg_spolygon* const* shapes, const uint8_t* dynamic, uint32_t n, jg_cde** out) {
    if (!bbox || !cfg || !out || (n && (!ents || !shapes))) return fail(JG_ERR_INVALID_ARG, "null argument");
    if (cfg->quadtree_depth > JG_MAX_QT_DEPTH) return fail(JG_ERR_DEPTH, "quadtree depth above JG_MAX_QT_DEPTH");
    if (cfg->cd_threshold > 255) return fail(JG_ERR_INVALID_ARG, "cd_threshold is a u8");
    auto c = std::make_unique<jg_cde>();
    c->cfg = *cfg;
    c->qt.init((uint8_t)cfg->quadtree_depth, make_float4(bbox[0], bbox[1], bbox[2], bbox[3]), (uint8_t)cfg->cd_threshold);
    for (uint32_t i = 0; i < n; i++) {
        if (!shapes[i]) return fail(JG_ERR_INVALID_ARG, "null shape");
        cde_register(c.get(), ents[i], shapes[i]->poly, dynamic ? dynamic[i] != 0 : false);
    }
    bool found = false;
    for (uint32_t s = 1; s < c->map.slots.size() && !found; s++)
        if (c->map.slots[s].occ && c->map.slots[s].rec.entity.kind == JG_ENTITY_EXTERIOR) {
            c->exterior = jg_hazkey{s, c->map.slots[s].version};
            found = true;
        }
    if (!found) return fail(JG_ERR_NO_EXTERIOR, "No exterior hazard registered in the CDE");
    *out = c.release();
    return JG_OK;
}

extern "C" void jg_cde_free(jg_cde* c) { delete c; }

extern "C" jg_status jg_cde_register_hazard(jg_cde* c, const jg_hazard_entity* e, const jg_spolygon* s, uint8_t dynamic,
                                            jg_hazkey* key_out) {
    if (!c || !e || !s) return fail(JG_ERR_INVALID_ARG, "null argument");
    const jg_hazkey k = cde_register(c, *e, s->poly, dynamic != 0);
    if (key_out) *key_out = k;
    return JG_OK;
}

extern "C" jg_status jg_cde_deregister_hazard_by_key(jg_cde* c, jg_hazkey k) {
    if (!c) return fail(JG_ERR_INVALID_ARG, "null argument");
    if (!c->map.valid(k)) return fail(JG_ERR_NOT_REGISTERED, "Cannot deregister hazard that is not registered");
    c->map.remove(k);
    c->qt.deregister_hazard(0, k.idx, k.version);
    c->dirty = true;
    return JG_OK;
}

extern "C" jg_status jg_cde_deregister_hazard_by_entity(jg_cde* c, const jg_hazard_entity* e) {
    if (!c || !e) return fail(JG_ERR_INVALID_ARG, "null argument");
    for (uint32_t s = 1; s < c->map.slots.size(); s++)
        if (c->map.slots[s].occ && entity_eq(c->map.slots[s].rec.entity, *e))
            return jg_cde_deregister_hazard_by_key(c, jg_hazkey{s, c->map.slots[s].version});
    return fail(JG_ERR_NOT_REGISTERED, "Cannot deregister hazard that is not registered");
}

extern "C" jg_status jg_cde_save(const jg_cde* c, jg_snapshot** out) {
    if (!c || !out) return fail(JG_ERR_INVALID_ARG, "null argument");
    auto s = std::make_unique<jg_snapshot>();
    for (uint32_t i = 1; i < c->map.slots.size(); i++)
        if (c->map.slots[i].occ && c->map.slots[i].rec.dynamic) s->dynamic.push_back(c->map.slots[i].rec);
    *out = s.release();
    return JG_OK;
}

// CDEngine::restore (cd_engine.rs:109-149); the swap_remove order decides which slots are reused.
extern "C" jg_status jg_cde_restore(jg_cde* c, const jg_snapshot* snap) {
    if (!c || !snap) return fail(JG_ERR_INVALID_ARG, "null argument");
    std::vector<std::pair<jg_hazkey, jg_hazard_entity>> rm;
    for (uint32_t i = 1; i < c->map.slots.size(); i++)
        if (c->map.slots[i].occ && c->map.slots[i].rec.dynamic) rm.push_back({jg_hazkey{i, c->map.slots[i].version}, c->map.slots[i].rec.entity});
    std::vector<const HazardRec*> add;
    for (const HazardRec& h : snap->dynamic) {
        size_t pos = rm.size();
        for (size_t i = 0; i < rm.size(); i++)
            if (entity_eq(rm[i].second, h.entity)) { pos = i; break; }
        if (pos < rm.size()) {
            rm[pos] = rm.back();
            rm.pop_back();
        } else {
            add.push_back(&h);
        }
    }
    for (auto& r : rm) {
        jg_status s = jg_cde_deregister_hazard_by_key(c, r.first);
        if (s != JG_OK) return s;
    }
    for (const HazardRec* h : add) cde_register(c, h->entity, h->shape, h->dynamic);
    return JG_OK;
}

extern "C" void jg_snapshot_free(jg_snapshot* s) { delete s; }

extern "C" jg_status jg_cde_bbox(const jg_cde* c, float out[4]) {
    if (!c || !out) return fail(JG_ERR_INVALID_ARG, "null argument");
    const float4 b = c->qt.pool[0].bbox;
    out[0] = b.x; out[1] = b.y; out[2] = b.z; out[3] = b.w;
    return JG_OK;
}
extern "C" jg_status jg_cde_hkey_exterior(const jg_cde* c, jg_hazkey* out) {
    if (!c || !out) return fail(JG_ERR_INVALID_ARG, "null argument");
    *out = c->exterior;
    return JG_OK;
}
extern "C" jg_status jg_cde_sync(jg_cde* c) {
    if (!c) return fail(JG_ERR_INVALID_ARG, "null argument");
    return cde_upload(c);
}
extern "C" jg_status jg_cde_get_stats(const jg_cde* c, jg_cde_stats* out) {
    if (!c || !out) return fail(JG_ERR_INVALID_ARG, "null argument");
    *out = c->stats;
    return JG_OK;
}

extern "C" jg_status jg_cde_dump(const jg_cde* c, uint32_t* nodes, uint64_t nodes_cap, uint32_t* haz, uint64_t haz_cap,
                                 uint64_t* n_nodes, uint64_t* n_haz) {
    if (!c || !n_nodes || !n_haz) return fail(JG_ERR_INVALID_ARG, "null argument");
    FlatCDE f;
    flatten(c->qt, c->map, f);
    *n_nodes = f.nodes.size();
    *n_haz = f.ent.size();
    for (size_t i = 0; nodes && i < f.nodes.size() && i < nodes_cap; i++) {
        uint32_t* o = nodes + 8 * i;
        // level is recovered from the breadth-first structure: root level = depth, children one less
        o[0] = 0;
        o[1] = f.nodes[i].first_child >= 0 ? 4 : 0;
        o[2] = f.nodes[i].haz_info & 0x3fffffffu;
        o[3] = f.nodes[i].n_active_edges;
        std::memcpy(o + 4, &f.nodes[i].bbox, 16);
    }
    if (nodes) {   // fill levels top-down
        for (size_t i = 0; i < f.nodes.size() && i < nodes_cap; i++) {
            if (i == 0) nodes[0] = c->cfg.quadtree_depth;
            const int fc = f.nodes[i].first_child;
            if (fc >= 0)
                for (int k = 0; k < 4; k++)
                    if ((uint64_t)(fc + k) < nodes_cap) nodes[8 * (size_t)(fc + k)] = nodes[8 * i] - 1;
        }
    }
    for (size_t i = 0; haz && i < f.ent.size() && i < haz_cap; i++) {
        uint32_t* o = haz + 8 * i;
        const FlatEntry& e = f.ent[i];
        o[0] = e.slot_flags & 0x7fffffffu;
        o[1] = (e.slot_flags & 0x80000000u) ? 2u : 1u;
        o[2] = e.n_edges_total;
        o[3] = 0;
        const float4 z = make_float4(0, 0, 0, 0);
        std::memcpy(o + 4, (e.slot_flags & 0x80000000u) ? &z : &e.ff, 16);
    }
    return JG_OK;
}

// ------------------------------------------------------------------------------ ABI: queries
// SPSurrogate::ff_poles slices poles[0..n_ff_poles] (sp_surrogate.rs:55-57): the reference
// panics when the item has fewer poles; here the call is rejected before any GPU work.
static jg_status check_item(const jg_spolygon* item, uint32_t flags) {
    if (flags == 0) flags = JG_QUERY_SURROGATE | JG_QUERY_POLY;
    const Polygon& p = item->poly;
    if (flags & JG_QUERY_SURROGATE) {
        if (!p.sur.present) return fail(JG_ERR_SURROGATE, "surrogate not generated");
        if (p.sur.cfg.n_ff_poles > p.sur.poles.size()) return fail(JG_ERR_SURROGATE, "ff_poles slice out of range");
    }
    return JG_OK;
}

extern "C" jg_status jg_detect_batch(jg_cde* c, const jg_spolygon* item, const float* d_m6, uint64_t n,
                                     const jg_hazard_filter* filter, uint32_t flags, uint8_t* d_status, void* stream) {
    if (!c || !item) return fail(JG_ERR_INVALID_ARG, "null argument");
    if (n == 0) return JG_OK;
    if (!d_m6 || !d_status) return fail(JG_ERR_INVALID_ARG, "null device buffer");
    if (flags == 0) flags = JG_QUERY_SURROGATE | JG_QUERY_POLY;
    jg_status s = check_item(item, flags);
    if (s != JG_OK) return s;
    s = cde_upload(c);
    if (s != JG_OK) return s;
    s = item_upload(const_cast<jg_spolygon*>(item));
    if (s != JG_OK) return s;
    FilterView fv;
    s = filter_view(c, filter, fv, (hipStream_t)stream);
    if (s != JG_OK) return s;
    ItemView iv = item->view;
    if (!(flags & JG_QUERY_SURROGATE)) { iv.n_poles = 0; iv.n_piers = 0; }
    const uint32_t block = 256;
    const uint64_t grid = (n + block - 1) / block;
    if (grid > 0x7fffffffull) return fail(JG_ERR_INVALID_ARG, "batch too large");
    hipLaunchKernelGGL(k_detect_plain, dim3((uint32_t)grid), dim3(block), 0, (hipStream_t)stream, c->view, iv, d_m6, n,
                       fv, flags, d_status);
    HIP_TRY(hipGetLastError());
    if (fv.kind == 1) return filter_read_done(c, (hipStream_t)stream);
    return JG_OK;
}

extern "C" jg_status jg_collect_batch(jg_cde* c, const jg_spolygon* item, const float* d_m6, uint64_t n,
                                      uint32_t capacity, uint32_t* d_count, uint64_t* d_hash, uint32_t* d_keys,
                                      void* stream) {
    if (!c || !item) return fail(JG_ERR_INVALID_ARG, "null argument");
    if (n == 0) return JG_OK;
    if (!d_m6 || !d_count || !d_hash) return fail(JG_ERR_INVALID_ARG, "null device buffer");
    jg_status s = check_item(item, JG_QUERY_SURROGATE | JG_QUERY_POLY);
    if (s != JG_OK) return s;
    s = cde_upload(c);
    if (s != JG_OK) return s;
    s = item_upload(const_cast<jg_spolygon*>(item));
    if (s != JG_OK) return s;
    const uint32_t words = (c->view.n_slots + 31u) / 32u;
    if ((size_t)words * 4 > 48000) return fail(JG_ERR_CAPACITY, "too many hazard slots for the LDS collector");
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint64_t grid = std::min<uint64_t>(n, (uint64_t)cus * 16u);     // persistent one-wave workgroups
    hipLaunchKernelGGL(k_collect, dim3((uint32_t)grid), dim3(64), (size_t)words * 4, (hipStream_t)stream,
                       c->view, item->view, d_m6, n, words, d_keys ? capacity : 0u, d_count, d_hash, d_keys);
    HIP_TRY(hipGetLastError());
    return JG_OK;
}

namespace {
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};
}  // namespace

extern "C" jg_status jg_detect_batch_host(jg_cde* c, const jg_spolygon* item, const float* m6, uint64_t n,
                                          const jg_hazard_filter* filter, uint32_t flags, uint8_t* status) {
    if (!c || !item || (n && (!m6 || !status))) return fail(JG_ERR_INVALID_ARG, "null argument");
    if (n == 0) return JG_OK;
    jg_status s = check_item(item, flags);
    if (s != JG_OK) return s;
    s = device_ready();
    if (s != JG_OK) return s;
    DevBuf dm, ds;
    HIP_TRY(hipMalloc(&dm.p, sizeof(float) * 6 * n));
    HIP_TRY(hipMalloc(&ds.p, n));
    HIP_TRY(hipMemcpy(dm.p, m6, sizeof(float) * 6 * n, hipMemcpyHostToDevice));
    s = jg_detect_batch(c, item, (const float*)dm.p, n, filter, flags, (uint8_t*)ds.p, nullptr);
    if (s != JG_OK) return s;
    HIP_TRY(hipMemcpy(status, ds.p, n, hipMemcpyDeviceToHost));
    return JG_OK;
}

extern "C" jg_status jg_collect_batch_host(jg_cde* c, const jg_spolygon* item, const float* m6, uint64_t n,
                                           uint32_t capacity, uint32_t* count, uint64_t* hash, uint32_t* keys) {
    if (!c || !item || (n && (!m6 || !count || !hash))) return fail(JG_ERR_INVALID_ARG, "null argument");
    if (n == 0) return JG_OK;
    jg_status s = check_item(item, 0);
    if (s != JG_OK) return s;
    s = device_ready();
    if (s != JG_OK) return s;
    DevBuf dm, dc, dh, dk;
    HIP_TRY(hipMalloc(&dm.p, sizeof(float) * 6 * n));
    HIP_TRY(hipMalloc(&dc.p, sizeof(uint32_t) * n));
    HIP_TRY(hipMalloc(&dh.p, sizeof(uint64_t) * n));
    if (keys && capacity) HIP_TRY(hipMalloc(&dk.p, sizeof(uint32_t) * n * capacity));
    HIP_TRY(hipMemcpy(dm.p, m6, sizeof(float) * 6 * n, hipMemcpyHostToDevice));
    s = jg_collect_batch(c, item, (const float*)dm.p, n, capacity, (uint32_t*)dc.p, (uint64_t*)dh.p, (uint32_t*)dk.p, nullptr);
    if (s != JG_OK) return s;
    HIP_TRY(hipMemcpy(count, dc.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hash, dh.p, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
    if (keys && capacity) HIP_TRY(hipMemcpy(keys, dk.p, sizeof(uint32_t) * n * capacity, hipMemcpyDeviceToHost));
    return JG_OK;
}
