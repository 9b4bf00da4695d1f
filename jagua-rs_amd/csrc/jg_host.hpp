// Host-side data structures of the MI355X collision engine: polygons (SPolygon + surrogate),
// the hazard slot map, the mutable quadtree (a node pool that keeps the reference's exact
// structure under register / deregister) and its device image: pool-ordered records with
// in-place updates, and the breadth-first flattening that the tree dump uses.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/jagua_amd.h"
#include "jg_geom.hpp"

namespace jg {

struct Surrogate {
    std::vector<float4> poles;          // cx, cy, r, 0
    std::vector<float4> piers;          // x1 y1 x2 y2
    std::vector<uint32_t> hull_idx;
    float hull_area = 0.0f;
    jg_surrogate_config cfg{};
    bool present = false;
};

struct Polygon {
    std::vector<float2> v;
    float4 bbox{};
    float area = 0.0f, diameter = 0.0f;
    float4 poi{};                       // cx, cy, r, 0
    Surrogate sur;
    uint32_t n() const { return (uint32_t)v.size(); }
    float4 edge(uint32_t i) const {
        uint32_t j = i + 1 == v.size() ? 0 : i + 1;
        return make_float4(v[i].x, v[i].y, v[j].x, v[j].y);
    }
};

// polygon construction (simple_polygon.rs, fail_fast/*.rs, convex_hull.rs)
bool polygon_new(const float* xy, uint32_t n, Polygon& out, std::string& err);
bool polygon_generate_surrogate(Polygon& p, const jg_surrogate_config& cfg, std::string& err);
// the two halves around the pole searches, for callers that search poles elsewhere:
// SPolygon::new without the pole of inaccessibility (poi left unset), and the surrogate from
// its poles (hull, piers)
bool polygon_prepare(const float* xy, uint32_t n, Polygon& out, std::string& err);
// piers, when given, replace the host pier search (they come from the device's)
bool surrogate_finish(Polygon& p, const jg_surrogate_config& cfg, std::vector<float4> poles, std::string& err,
                      const std::vector<float4>* piers = nullptr);
Polygon polygon_transform(const Polygon& p, const Mat& m);
// SPolygon::centroid (simple_polygon.rs:173-192)
float2 polygon_centroid(const Polygon& shape);
// decompose(rot_transl(b) . X . I) with X = rot_transl(a), inverted when invert_a
// (transformation.rs); false when the reference's NotNan arithmetic would panic.
bool dt_chain(const float a[3], bool invert_a, const float b[3], float out[3]);
// generate_piers' per-polygon constants and the (sin, cos) of its 90 ray angles (host libm)
void pier_setup(const Polygon& shape, PierSetup& out);
void pier_angles(float* sin_cos);
bool polygon_has_point(const Polygon& p, float x, float y);
Mat mat_from_dt(float rotation, float tx, float ty);

// ----------------------------------------------------------------------------- hazards
struct HazardRec {
    jg_hazard_entity entity{};
    Polygon shape;
    bool dynamic = false;
    bool interior() const { return entity.kind != JG_ENTITY_EXTERIOR; }
};

bool entity_eq(const jg_hazard_entity& a, const jg_hazard_entity& b);

// slotmap semantics: slot 0 is a sentinel, freed slots are reused last-in first-out.
struct SlotMap {
    struct Slot { uint32_t version = 0; bool occ = false; uint32_t next_free = 0; HazardRec rec; };
    std::vector<Slot> slots = std::vector<Slot>(1);
    uint32_t free_head = 1;
    jg_hazkey insert(HazardRec r);
    bool valid(jg_hazkey k) const { return k.idx < slots.size() && slots[k.idx].occ && slots[k.idx].version == k.version; }
    HazardRec remove(jg_hazkey k);
};

// ----------------------------------------------------------------------------- quadtree
enum : uint8_t { PRES_PARTIAL = 1, PRES_ENTIRE = 2 };

struct QEntry {                          // QTHazard in one node (qt_hazard.rs:13-33)
    uint32_t slot = 0, version = 0;
    uint8_t presence = 0;
    float4 ff{};                         // ff_bbox of the partial edge subset
    std::vector<uint32_t> edges;         // edge indices into the hazard polygon
};

// Levels of the cell tables: the tree's depth, at most this many (levels 0 .. 7: 87 KB).
constexpr uint32_t CELL_LEVELS_MAX = 7;
constexpr uint32_t CELL_NO_CDEPTH = 15;

struct QNode {                           // QTNode (qt_node.rs:14-25)
    uint8_t level = 0;
    float4 bbox{};
    int32_t child = -1;                  // pool index of the first of 4 children
    std::vector<QEntry> haz;             // Entire entries first, then Partial
    uint64_t n_active_edges = 0;
};

struct QuadTree {
    std::vector<QNode> pool;
    std::vector<int32_t> free_blocks;
    std::vector<int32_t> touched;        // nodes whose record, entries or children changed since the last image patch
    // nodes within the cell tables' levels whose cells changed since the last image build or
    // patch: children made or dropped, or the node started or stopped collecting
    std::vector<int32_t> cells_dirty;
    uint8_t cd_threshold = 16;
    void init(uint8_t depth, float4 bbox, uint8_t cdt);
    void register_hazard(int32_t node, QEntry e, const SlotMap& map);
    void deregister_hazard(int32_t node, uint32_t slot, uint32_t version);
    // Renumbers the pool in breadth-first order from the root and drops free blocks.
    void compact_bfs();
  private:
    int32_t alloc_children(int32_t parent);
    void free_subtree_children(int32_t node);
    void push_to_children(int32_t node, const QEntry& e, const SlotMap& map);
};

// Records of the device image (and of the breadth-first flattening).
struct FlatEntry {                       // 32 B
    float4 ff;
    uint32_t slot_flags;                 // slot | 0x80000000 when Entire
    uint32_t edge_begin;
    uint32_t edge_count;
    uint32_t n_edges_total;              // partial edge count (n_active_edges bookkeeping)
};

// One quadtree node, 32 B: everything a descent step reads in one load.
struct FlatNode {
    float4 bbox;
    int32_t first_child;                 // index of the first of 4 consecutive children, -1 for a leaf
    uint32_t haz_begin;                  // first hazard entry
    uint32_t haz_info;                   // entry count | 0x80000000 when the first entry is Entire
    uint32_t n_active_edges;             // partial edges over all entries (saturated)
};

struct FlatCDE {
    std::vector<FlatNode> nodes;
    std::vector<uint32_t> node_up;       // parent << 2 | index among siblings
    std::vector<FlatEntry> ent;
    std::vector<float4> edges;           // inline hazard edges of entries in leaves / threshold nodes
    std::vector<float4> haz_bbox;        // per slot
    std::vector<float4> haz_poi;         // cx, cy, r, interior ? 1 : 0
    std::vector<uint2> haz_vr;           // vertex offset, count
    std::vector<float2> haz_verts;
};

void flatten(const QuadTree& qt, const SlotMap& map, FlatCDE& out);

// The entry a new hazard registers at the root with: Partial, every edge, ff = its bbox.
QEntry root_entry(const Polygon& s, jg_hazkey k);

// The device image in pool order: node i of the image is node i of the host pool, whose
// children already come in blocks of 4 consecutive nodes.  Every node owns a range of entry
// records and one of inline edges; after an update only the touched nodes are rewritten, in
// place while their ranges are large enough, else at the end of the arena.  The image is
// rebuilt whole when a device array outgrows its capacity or garbage exceeds the live data.
struct ImageSeg {                         // one copy of the staging buffer into the device image
    uint64_t dst, src, bytes;             // byte offsets and size (multiples of 4)
};

struct PoolImage {
    bool valid = false;
    std::vector<FlatNode> nodes;
    std::vector<uint32_t> up;
    std::vector<uint32_t> ent_off, ent_cap, edge_off, edge_cap;
    std::vector<FlatEntry> ent;
    std::vector<float4> edges;
    std::vector<float4> hbb, hpoi;
    std::vector<uint2> hvr;
    std::vector<float2> verts;
    std::vector<uint32_t> slot_ver;       // version written per slot, 0 = none
    // cell tables of levels 0 .. cell_level, one after another (level l at (4^l - 1) / 3):
    // entry (ix, iy) of level l (at iy << l | ix) is cdepth << 28 | depth << 24 | node for the
    // deepest node on the tree path to cell (ix, iy) of the 2^l grid of the root box's
    // quadrant subdivision, cdepth the depth of the first node on that path where
    // QTNode::collect_collisions tests entries (a leaf, or at most cd_threshold active edges;
    // their descendants are such nodes too), CELL_NO_CDEPTH when there is none
    std::vector<uint32_t> cells;
    uint32_t cell_level = 0;
    uint64_t garbage = 0;                 // entry + edge records no node refers to
    bool building = false;                // inside image_build: ranges are packed tightly
    // device layout: record capacities and byte offsets
    uint64_t cap_nodes = 0, cap_ent = 0, cap_edges = 0, cap_slots = 0, cap_verts = 0;
    uint64_t o_nd = 0, o_nu = 0, o_en = 0, o_ed = 0, o_hb = 0, o_hp = 0, o_hv = 0, o_vv = 0, o_ce = 0, total = 0;
};

// Fills im.cells from the tree (quadrant k of a node: 0 = x high y high, 1 = x low y high,
// 2 = x low y low, 3 = x high y low, as box_quadrant orders them).
void image_cells(PoolImage& im, const QuadTree& qt);
// Whether QTNode::collect_collisions tests node n's entries (qt_node.rs:151-191).
bool collects_at(const QuadTree& qt, int32_t n);

// Full build of the host mirror and the device layout (capacities with headroom).
void image_build(PoolImage& im, QuadTree& qt, const SlotMap& map);
// Rewrites the nodes and slots changed since the last build or patch into the host mirror and
// lists the device copies; false when a full build is needed instead.
bool image_patch(PoolImage& im, QuadTree& qt, const SlotMap& map, std::vector<uint8_t>& staging,
                 std::vector<ImageSeg>& segs);

}  // namespace jg
