// CPU check of the engine's in-place device image updates.  A byte buffer stands in for the
// HBM image: full builds copy the host mirror into it, patches apply only their segments.
// After every batch of random register / deregister operations the buffer, read as the
// kernels read it (from node 0 through first_child links), must hold exactly the tree and
// hazard records of a fresh breadth-first flatten of the host quadtree.
// Usage: image_patch_check <seed> <rounds> <depth>; prints "ok <patches> <builds>".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "../jagua-rs_amd/csrc/jg_host.hpp"

using namespace jg;

static uint64_t g_state;
static double urand() {                      // splitmix64 -> [0, 1)
    uint64_t z = (g_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static Polygon star(float cx, float cy, float r, int nv) {
    std::vector<float> xy;
    for (int i = 0; i < nv; i++) {
        const double a = 2.0 * M_PI * i / nv;
        const double rr = r * (0.55 + 0.45 * urand());
        xy.push_back((float)(cx + rr * std::cos(a)));
        xy.push_back((float)(cy + rr * std::sin(a)));
    }
    Polygon p;
    std::string err;
    if (!polygon_new(xy.data(), (uint32_t)nv, p, err)) {
        std::fprintf(stderr, "polygon_new: %s\n", err.c_str());
        std::exit(2);
    }
    return p;
}

static int fails = 0;
#define CHECK(cond, ...)                                  \
    do {                                                  \
        if (!(cond)) {                                    \
            if (fails++ < 10) {                           \
                std::fprintf(stderr, "mismatch: " __VA_ARGS__); \
                std::fprintf(stderr, "\n");               \
            }                                             \
        }                                                 \
    } while (0)

template <class T>
static T rd(const std::vector<uint8_t>& b, uint64_t off) {
    T t;
    if (off + sizeof(T) > b.size()) {
        std::fprintf(stderr, "read out of the image at %llu\n", (unsigned long long)off);
        std::exit(3);
    }
    std::memcpy(&t, b.data() + off, sizeof(T));
    return t;
}
static bool same_bits(const void* a, const void* b, size_t n) { return std::memcmp(a, b, n) == 0; }

static void full_copy(const PoolImage& im, std::vector<uint8_t>& dev) {
    dev.assign(im.total, 0);
    auto put = [&](uint64_t o, const void* src, uint64_t bytes) {
        if (bytes) std::memcpy(dev.data() + o, src, bytes);
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
}

static void verify(const PoolImage& im, const std::vector<uint8_t>& dev, const QuadTree& qt, const SlotMap& map) {
    FlatCDE f;
    flatten(qt, map, f);
    std::deque<uint32_t> q{0};
    size_t i = 0;
    for (; !q.empty(); i++) {
        const uint32_t d = q.front();
        q.pop_front();
        if (i >= f.nodes.size()) { CHECK(false, "image has more reachable nodes than the tree"); return; }
        const FlatNode dn = rd<FlatNode>(dev, im.o_nd + sizeof(FlatNode) * (uint64_t)d);
        const FlatNode& fn = f.nodes[i];
        CHECK(same_bits(&dn.bbox, &fn.bbox, 16), "node %zu bbox", i);
        CHECK(dn.haz_info == fn.haz_info, "node %zu haz_info %u vs %u", i, dn.haz_info, fn.haz_info);
        CHECK(dn.n_active_edges == fn.n_active_edges, "node %zu n_active_edges", i);
        CHECK((dn.first_child < 0) == (fn.first_child < 0), "node %zu leaf flag", i);
        if (dn.first_child >= 0)
            for (uint32_t k = 0; k < 4; k++) {
                const uint32_t c = (uint32_t)dn.first_child + k;
                CHECK(rd<uint32_t>(dev, im.o_nu + 4ull * c) == ((d << 2) | k), "node %zu child %u up link", i, k);
                q.push_back(c);
            }
        const uint32_t cnt = fn.haz_info & 0x3fffffffu;
        for (uint32_t k = 0; k < cnt; k++) {
            const FlatEntry de = rd<FlatEntry>(dev, im.o_en + sizeof(FlatEntry) * (uint64_t)(dn.haz_begin + k));
            const FlatEntry& fe = f.ent[fn.haz_begin + k];
            CHECK(de.slot_flags == fe.slot_flags, "node %zu entry %u slot", i, k);
            CHECK(de.n_edges_total == fe.n_edges_total, "node %zu entry %u n_edges_total", i, k);
            CHECK(de.edge_count == fe.edge_count, "node %zu entry %u edge_count", i, k);
            CHECK(same_bits(&de.ff, &fe.ff, 16), "node %zu entry %u ff", i, k);
            for (uint32_t e = 0; e < fe.edge_count && e < de.edge_count; e++) {
                const float4 ed = rd<float4>(dev, im.o_ed + 16ull * (de.edge_begin + e));
                CHECK(same_bits(&ed, &f.edges[fe.edge_begin + e], 16), "node %zu entry %u edge %u", i, k, e);
            }
        }
    }
    CHECK(i == f.nodes.size(), "reachable nodes %zu vs %zu", i, f.nodes.size());
    // collecting is inherited: a child's active edges are a subset of its parent's, so every
    // child of a node where collect_collisions tests entries tests them too (the collection's
    // start nodes and the cell tables' first collecting depth rely on it)
    for (size_t n = 0; n < qt.pool.size(); n++) {
        const QNode& q = qt.pool[n];
        if (q.child < 0 || !collects_at(qt, (int32_t)n)) continue;
        for (int k = 0; k < 4; k++)
            CHECK(collects_at(qt, q.child + k), "node %zu collects, its child %d does not", n, k);
    }
    // the cell tables: entry (ix, iy) of level l is cdepth << 28 | depth << 24 | node for the
    // deepest node on the tree path to that cell, cdepth the depth of the path's first node where
    // collect_collisions tests entries (a leaf or at most cd_threshold active edges), 15 if none
    const uint32_t L = im.cell_level;
    CHECK(L == (qt.pool[0].level < CELL_LEVELS_MAX ? qt.pool[0].level : CELL_LEVELS_MAX), "cell level %u", L);
    for (uint32_t l = 0; l <= L; l++) {
        const uint64_t off = ((1ull << (2 * l)) - 1) / 3;
        for (uint32_t iy = 0; iy < (1u << l); iy++)
            for (uint32_t ix = 0; ix < (1u << l); ix++) {
                int32_t n = 0;
                uint32_t d = 0;
                const auto collects = [&](int32_t k) {
                    return qt.pool[(size_t)k].child < 0 || qt.pool[(size_t)k].n_active_edges <= qt.cd_threshold;
                };
                uint32_t cd = collects(0) ? 0u : CELL_NO_CDEPTH;
                for (; d < l && qt.pool[(size_t)n].child >= 0; d++) {
                    const uint32_t bx = (ix >> (l - 1u - d)) & 1u, by = (iy >> (l - 1u - d)) & 1u;
                    n = qt.pool[(size_t)n].child + (bx ? (by ? 0 : 3) : (by ? 1 : 2));
                    if (cd == CELL_NO_CDEPTH && collects(n)) cd = d + 1u;
                }
                CHECK(rd<uint32_t>(dev, im.o_ce + 4ull * (off + (((uint64_t)iy << l) | ix))) ==
                          ((cd << 28) | (d << 24) | (uint32_t)n),
                      "level %u cell (%u, %u)", l, ix, iy);
            }
    }
    CHECK(im.slot_ver.size() == map.slots.size(), "slot count");
    for (size_t s = 1; s < map.slots.size(); s++) {
        if (!map.slots[s].occ) continue;
        const float4 hb = rd<float4>(dev, im.o_hb + 16 * s), hp = rd<float4>(dev, im.o_hp + 16 * s);
        const uint2 hv = rd<uint2>(dev, im.o_hv + 8 * s);
        CHECK(same_bits(&hb, &f.haz_bbox[s], 16), "slot %zu bbox", s);
        CHECK(same_bits(&hp, &f.haz_poi[s], 16), "slot %zu poi", s);
        CHECK(hv.y == f.haz_vr[s].y, "slot %zu vertex count", s);
        for (uint32_t v = 0; v < hv.y && v < f.haz_vr[s].y; v++) {
            const float2 a = rd<float2>(dev, im.o_vv + 8ull * (hv.x + v));
            CHECK(same_bits(&a, &f.haz_verts[f.haz_vr[s].x + v], 8), "slot %zu vertex %u", s, v);
        }
    }
}

int main(int argc, char** argv) {
    g_state = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1;
    const int rounds = argc > 2 ? std::atoi(argv[2]) : 200;
    const int depth = argc > 3 ? std::atoi(argv[3]) : 6;
    const float W = 1000.0f, H = 500.0f;
    QuadTree qt;
    SlotMap map;
    qt.init((uint8_t)depth, make_float4(0, 0, W, H), 16);
    auto reg = [&](const Polygon& p, uint32_t kind, uint64_t id) {
        HazardRec r;
        r.entity.kind = kind;
        r.entity.a = id;
        r.shape = p;
        r.dynamic = kind == JG_ENTITY_PLACED_ITEM;
        const jg_hazkey k = map.insert(std::move(r));
        qt.register_hazard(0, root_entry(map.slots[k.idx].rec.shape, k), map);
        return k;
    };
    {
        const float box[8] = {0, 0, W, 0, W, H, 0, H};
        Polygon c;
        std::string err;
        if (!polygon_new(box, 4, c, err)) return 2;
        reg(c, JG_ENTITY_EXTERIOR, 0);
    }
    std::vector<jg_hazkey> live;
    uint64_t next_id = 1;
    auto add_item = [&]() {
        const float r = (float)(5.0 + 30.0 * urand());
        live.push_back(reg(star((float)(W * urand()), (float)(H * urand()), r, 4 + (int)(40 * urand())),
                           JG_ENTITY_PLACED_ITEM, next_id++));
    };
    for (int i = 0; i < 60; i++) add_item();
    PoolImage im;
    std::vector<uint8_t> dev, staging;
    std::vector<ImageSeg> segs;
    image_build(im, qt, map);
    full_copy(im, dev);
    verify(im, dev, qt, map);
    int patches = 0, builds = 1;
    for (int r = 0; r < rounds && !fails; r++) {
        const int ops = 1 + (int)(6 * urand() * urand());
        for (int o = 0; o < ops; o++) {
            if (!live.empty() && urand() < 0.5) {
                const size_t j = (size_t)(urand() * live.size());
                const jg_hazkey k = live[j];
                live[j] = live.back();
                live.pop_back();
                map.remove(k);
                qt.deregister_hazard(0, k.idx, k.version);
            } else {
                add_item();
            }
        }
        if (image_patch(im, qt, map, staging, segs)) {
            for (const ImageSeg& g : segs) {
                if (g.dst + g.bytes > dev.size() || g.src + g.bytes > staging.size() || ((g.dst | g.src | g.bytes) & 3)) {
                    std::fprintf(stderr, "segment out of range\n");
                    return 1;
                }
                std::memcpy(dev.data() + g.dst, staging.data() + g.src, g.bytes);
            }
            patches++;
        } else {
            image_build(im, qt, map);
            full_copy(im, dev);
            builds++;
        }
        verify(im, dev, qt, map);
        if (fails) std::fprintf(stderr, "after round %d\n", r);
    }
    if (fails) return 1;
    std::printf("ok %d %d\n", patches, builds);
    return 0;
}
