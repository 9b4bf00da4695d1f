// TEST INFRASTRUCTURE ONLY — the CPU oracle.
//
// A plain C++17 restatement of the jagua-rs collision-detection engine (CDE), written to
// follow the reference's algorithm and f32 operation order line by line.  It is the checker
// the MI355X path is compared against and the CPU baseline bench.py times.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the product library
// (jagua-rs_amd/) never links, includes or calls anything in this directory.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fno-fast-math).  No FMA contraction,
// IEEE division/sqrt, denormals kept: the same arithmetic Rust emits for the reference.
//
// Parity status of each part (how it is pinned) is listed in tests/test_oracle_*.py and in
// the header comment of jg_oracle.cpp.
#pragma once
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

namespace oracle {

// ---------------------------------------------------------------------------------------
// L1 primitives — geometry/primitives/{point,edge,rect,circle}.rs
// ---------------------------------------------------------------------------------------
struct Point { float x, y; };
struct Edge { Point start, end; };
struct Rect { float x_min, y_min, x_max, y_max; };
struct Circle { Point center; float radius; };

enum class GeoRelation { Intersecting, Enclosed, Surrounding, Disjoint };  // geo_enums.rs:10-19
enum class GeoPosition { Exterior, Interior };                              // geo_enums.rs:2-5

// f32::min / f32::max are IEEE minNum/maxNum (Rust core::f32::min).
inline float fmin32(float a, float b) { return std::fmin(a, b); }
inline float fmax32(float a, float b) { return std::fmax(a, b); }

// f32::to_bits
inline uint32_t bits(float a) {
    uint32_t u;
    std::memcpy(&u, &a, 4);
    return u;
}
// util/fpa.rs:18-31 — float_cmp::approx_eq!(f32, a, b) with the default F32Margin
// (epsilon = f32::EPSILON, ulps = 4): exact ==, then |a-b| <= EPSILON, then
// saturating_abs(bits(a) wrapping_sub bits(b)) <= 4.
inline bool fpa_eq(float a, float b) {
    if (a == b) return true;
    float eps = std::fabs(a - b);
    if (eps <= FLT_EPSILON) return true;
    uint32_t ua, ub;
    std::memcpy(&ua, &a, 4);
    std::memcpy(&ub, &b, 4);
    int32_t diff = (int32_t)(ua - ub);
    int32_t ad = diff < 0 ? (diff == INT32_MIN ? INT32_MAX : -diff) : diff;
    return ad <= 4;
}
// FPA::partial_cmp: Equal when approx-equal, else the plain f32 order.
inline bool fpa_lt(float a, float b) { return !fpa_eq(a, b) && a < b; }
inline bool fpa_gt(float a, float b) { return !fpa_eq(a, b) && a > b; }
inline bool fpa_le(float a, float b) { return fpa_eq(a, b) || a < b; }
inline bool fpa_ge(float a, float b) { return fpa_eq(a, b) || a > b; }

// point.rs:26-31  x' = m00*x + m01*y + m02*1.0 (left to right, no contraction)
struct Transformation { float m[2][3]; };
Transformation transformation_from_dt(float rotation, float tx, float ty);  // transformation.rs:131-139

inline Point transform_point(const Point& p, const Transformation& t) {
    float tx = t.m[0][0] * p.x + t.m[0][1] * p.y + t.m[0][2] * 1.0f;
    float ty = t.m[1][0] * p.x + t.m[1][1] * p.y + t.m[1][2] * 1.0f;
    return {tx, ty};
}

// edge.rs:89-103
inline float edge_x_min(const Edge& e) { return fmin32(e.start.x, e.end.x); }
inline float edge_y_min(const Edge& e) { return fmin32(e.start.y, e.end.y); }
inline float edge_x_max(const Edge& e) { return fmax32(e.start.x, e.end.x); }
inline float edge_y_max(const Edge& e) { return fmax32(e.start.y, e.end.y); }

// rect.rs:248-253
inline Point rect_centroid(const Rect& r) { return {(r.x_min + r.x_max) / 2.0f, (r.y_min + r.y_max) / 2.0f}; }
inline float rect_width(const Rect& r) { return r.x_max - r.x_min; }
inline float rect_height(const Rect& r) { return r.y_max - r.y_min; }
// rect.rs:259-263
inline float rect_diameter(const Rect& r) {
    float dx = r.x_max - r.x_min, dy = r.y_max - r.y_min;
    return std::sqrt(dx * dx + dy * dy);
}
// rect.rs:266-272
inline bool rect_collides_rect(const Rect& a, const Rect& b) {
    return fmax32(a.x_min, b.x_min) <= fmin32(a.x_max, b.x_max) &&
           fmax32(a.y_min, b.y_min) <= fmin32(a.y_max, b.y_max);
}
// rect.rs:282-288
inline bool rect_collides_point(const Rect& r, const Point& p) {
    return p.x >= r.x_min && p.x <= r.x_max && p.y >= r.y_min && p.y <= r.y_max;
}
// rect.rs:46-65
inline GeoRelation rect_relation_to(const Rect& s, const Rect& o) {
    if (!rect_collides_rect(s, o)) return GeoRelation::Disjoint;
    if (s.x_min <= o.x_min && s.y_min <= o.y_min && s.x_max >= o.x_max && s.y_max >= o.y_max)
        return GeoRelation::Surrounding;
    if (s.x_min >= o.x_min && s.y_min >= o.y_min && s.x_max <= o.x_max && s.y_max <= o.y_max)
        return GeoRelation::Enclosed;
    return GeoRelation::Intersecting;
}
// rect.rs:274-280
inline bool rect_almost_collides_rect(const Rect& a, const Rect& b) {
    return fpa_le(fmax32(a.x_min, b.x_min), fmin32(a.x_max, b.x_max)) &&
           fpa_le(fmax32(a.y_min, b.y_min), fmin32(a.y_max, b.y_max));
}
// rect.rs:70-89
inline GeoRelation rect_almost_relation_to(const Rect& s, const Rect& o) {
    if (!rect_almost_collides_rect(s, o)) return GeoRelation::Disjoint;
    if (fpa_le(s.x_min, o.x_min) && fpa_le(s.y_min, o.y_min) && fpa_ge(s.x_max, o.x_max) &&
        fpa_ge(s.y_max, o.y_max))
        return GeoRelation::Surrounding;
    if (fpa_ge(s.x_min, o.x_min) && fpa_ge(s.y_min, o.y_min) && fpa_le(s.x_max, o.x_max) &&
        fpa_le(s.y_max, o.y_max))
        return GeoRelation::Enclosed;
    return GeoRelation::Intersecting;
}
// rect.rs:93-109
inline Rect rect_inflate_to_square(const Rect& r) {
    float width = r.x_max - r.x_min;
    float height = r.y_max - r.y_min;
    float dx = 0.0f, dy = 0.0f;
    if (height < width) dy = (width - height) / 2.0f;
    else if (width < height) dx = (height - width) / 2.0f;
    return {r.x_min - dx, r.y_min - dy, r.x_max + dx, r.y_max + dy};
}
// rect.rs:154-161  [q1 TR, q2 TL, q3 BL, q4 BR]
inline std::array<Point, 4> rect_corners(const Rect& r) {
    return {Point{r.x_max, r.y_max}, Point{r.x_min, r.y_max}, Point{r.x_min, r.y_min},
            Point{r.x_max, r.y_min}};
}
// rect.rs:35-41 (from_diagonal_corners) and 141-151 (quadrants)
inline Rect rect_from_diagonal(Point c1, Point c2) {
    Rect r{fmin32(c1.x, c2.x), fmin32(c1.y, c2.y), fmax32(c1.x, c2.x), fmax32(c1.y, c2.y)};
    if (!(r.x_min < r.x_max && r.y_min < r.y_max)) throw std::runtime_error("invalid rectangle");
    return r;
}
inline std::array<Rect, 4> rect_quadrants(const Rect& r) {
    Point mid = rect_centroid(r);
    auto c = rect_corners(r);
    return {rect_from_diagonal(c[0], mid), rect_from_diagonal(c[1], mid),
            rect_from_diagonal(c[2], mid), rect_from_diagonal(c[3], mid)};
}
// rect.rs:164-184 [top, left, bottom, right]
inline std::array<Edge, 4> rect_sides(const Rect& r) {
    auto c = rect_corners(r);
    return {Edge{c[0], c[1]}, Edge{c[1], c[2]}, Edge{c[2], c[3]}, Edge{c[3], c[0]}};
}
// rect.rs:136
constexpr int QUADRANT_NEIGHBOR_LAYOUT[4][2] = {{1, 3}, {0, 2}, {1, 3}, {0, 2}};

// rect.rs:301-345 (Rect–Edge; edge.rs:172-177 delegates here)
inline bool rect_collides_edge(const Rect& r, const Edge& e) {
    float e_x_min = edge_x_min(e), e_x_max = edge_x_max(e);
    float e_y_min = edge_y_min(e), e_y_max = edge_y_max(e);
    bool x_no = fmax32(e_x_min, r.x_min) > fmin32(e_x_max, r.x_max);
    bool y_no = fmax32(e_y_min, r.y_min) > fmin32(e_y_max, r.y_max);
    if (x_no || y_no) return false;
    if (rect_collides_point(r, e.start) || rect_collides_point(r, e.end)) return true;
    float s_x = e.start.x, s_y = e.start.y, e_x = e.end.x, e_y = e.end.y;
    float edge_dx = e_x - s_x, edge_dy = e_y - s_y;
    auto c = rect_corners(r);
    float s[4];
    for (int i = 0; i < 4; i++) s[i] = (c[i].x - s_x) * edge_dy - (c[i].y - s_y) * edge_dx;
    bool all_pos = s[0] > 0.0f && s[1] > 0.0f && s[2] > 0.0f && s[3] > 0.0f;
    bool all_neg = s[0] < 0.0f && s[1] < 0.0f && s[2] < 0.0f && s[3] < 0.0f;
    return !(all_pos || all_neg);
}

// edge.rs:179-228 (calc_loc variant returns the location through *loc when requested)
inline bool edge_intersection(const Edge& e1, const Edge& e2, Point* loc = nullptr) {
    float x1 = e1.start.x, y1 = e1.start.y, x2 = e1.end.x, y2 = e1.end.y;
    float x3 = e2.start.x, y3 = e2.start.y, x4 = e2.end.x, y4 = e2.end.y;
    {
        float x_min_e1 = fmin32(x1, x2), x_max_e1 = fmax32(x1, x2);
        float y_min_e1 = fmin32(y1, y2), y_max_e1 = fmax32(y1, y2);
        float x_min_e2 = fmin32(x3, x4), x_max_e2 = fmax32(x3, x4);
        float y_min_e2 = fmin32(y3, y4), y_max_e2 = fmax32(y3, y4);
        bool x_no = fmax32(x_min_e1, x_min_e2) > fmin32(x_max_e1, x_max_e2);
        bool y_no = fmax32(y_min_e1, y_min_e2) > fmin32(y_max_e1, y_max_e2);
        if (x_no || y_no) return false;
    }
    float t_nom = (x2 - x4) * (y4 - y3) - (y2 - y4) * (x4 - x3);
    float t_denom = (x2 - x1) * (y4 - y3) - (y2 - y1) * (x4 - x3);
    float u_nom = (x2 - x4) * (y2 - y1) - (y2 - y4) * (x2 - x1);
    float u_denom = (x2 - x1) * (y4 - y3) - (y2 - y1) * (x4 - x3);
    if (t_denom == 0.0f || u_denom == 0.0f) return false;
    float t = t_nom / t_denom;
    float u = u_nom / u_denom;
    if (t >= 0.0f && t <= 1.0f && u >= 0.0f && u <= 1.0f) {
        if (loc) *loc = Point{x2 + t * (x1 - x2), y2 + t * (y1 - y2)};
        return true;
    }
    return false;
}

// edge.rs:63-87
inline Point closest_point_on_edge(const Edge& e, const Point& p) {
    float x1 = e.start.x, y1 = e.start.y, x2 = e.end.x, y2 = e.end.y;
    float a = p.x - x1, b = p.y - y1, c = x2 - x1, d = y2 - y1;
    float dot = a * c + b * d;
    float len_sq = c * c + d * d;
    float param = -1.0f;
    if (len_sq != 0.0f) param = dot / len_sq;
    if (param < 0.0f) return {x1, y1};
    if (param > 1.0f) return {x2, y2};
    return {x1 + param * c, y1 + param * d};
}
// edge.rs:153-159
inline float edge_sq_distance_to(const Edge& e, const Point& p) {
    Point q = closest_point_on_edge(e, p);
    float dx = p.x - q.x, dy = p.y - q.y;
    return dx * dx + dy * dy;
}
// edge.rs:148-150
inline float edge_distance_to(const Edge& e, const Point& p) { return std::sqrt(edge_sq_distance_to(e, p)); }

// point.rs:43-52
inline float point_sq_distance(const Point& a, const Point& b) {
    float dx = a.x - b.x, dy = a.y - b.y;
    return dx * dx + dy * dy;
}
inline float point_distance(const Point& a, const Point& b) { return std::sqrt(point_sq_distance(a, b)); }

// circle.rs:109-113
inline bool circle_collides_edge(const Circle& c, const Edge& e) {
    return edge_sq_distance_to(e, c.center) <= c.radius * c.radius;
}
// circle.rs:115-127
inline bool circle_collides_rect(const Circle& c, const Rect& r) {
    float cx = c.center.x, cy = c.center.y;
    float nx = fmax32(r.x_min, fmin32(cx, r.x_max));
    float ny = fmax32(r.y_min, fmin32(cy, r.y_max));
    float dx = nx - cx, dy = ny - cy;
    return dx * dx + dy * dy <= c.radius * c.radius;
}
// circle.rs:130-134
inline bool circle_collides_point(const Circle& c, const Point& p) {
    return point_sq_distance(p, c.center) <= c.radius * c.radius;
}
// circle.rs:154-163 (separation distance to a point)
inline float circle_separation_distance_point(const Circle& c, const Point& p) {
    float dx = p.x - c.center.x, dy = p.y - c.center.y;
    float d_center = std::sqrt(dx * dx + dy * dy);
    if (d_center <= c.radius) return c.radius - d_center;
    return d_center - c.radius;
}
inline float circle_area(const Circle& c) { return c.radius * c.radius * 3.14159265358979323846f; }  // circle.rs:60-62

// ---------------------------------------------------------------------------------------
// SPolygon — geometry/primitives/simple_polygon.rs
// ---------------------------------------------------------------------------------------
struct SPSurrogateConfig {                  // sp_surrogate.rs:120-131
    uint32_t n_pole_limits_n[3];
    float n_pole_limits_cov[3];
    uint32_t n_ff_poles;
    uint32_t n_ff_piers;
};

struct SPSurrogate {                        // sp_surrogate.rs:17-28
    std::vector<Circle> poles;
    std::vector<Edge> piers;
    std::vector<size_t> convex_hull_indices;
    float convex_hull_area = 0.0f;
    SPSurrogateConfig config{};
    bool present = false;
};

struct SPolygon {
    std::vector<Point> vertices;
    Rect bbox{};
    float area = 0.0f;
    float diameter = 0.0f;
    Circle poi{};
    SPSurrogate surrogate;

    size_t n() const { return vertices.size(); }
    Edge edge(size_t i) const {                         // simple_polygon.rs:86-93
        size_t j = (i == vertices.size() - 1) ? 0 : i + 1;
        return {vertices[i], vertices[j]};
    }
};

SPolygon spolygon_new(std::vector<Point> points);                      // simple_polygon.rs:41-71
float spolygon_calculate_area(const std::vector<Point>& points);        // :137-150
Rect spolygon_bbox(const std::vector<Point>& points);                   // :122-133
float spolygon_calculate_diameter(std::vector<Point> points);           // :107-120
Point spolygon_centroid(const SPolygon& s);                             // :173-192
bool spolygon_collides_point(const SPolygon& s, const Point& p);        // :255-292
SPolygon spolygon_transform_clone(const SPolygon& s, const Transformation& t);  // :195-224

// ---- instance import (io/import.rs)
// import_simple_polygon (:198-212) up to SPolygon::new: closing vertex stripped, consecutive
// near-duplicates eliminated (:235-260), non-consecutive duplicates rejected (throws).
std::vector<Point> import_simple_polygon_points(std::vector<Point> pts);
// Transformation as the full 3x3 matrix of transformation.rs
struct Mat3 { float m[3][3]; };
Mat3 mat3_rot_transl(float angle, float tx, float ty);   // rot_transl_m (:132-141)
Mat3 mat3_dot(const Mat3& l, const Mat3& r);             // dot_prod (:158-179)
Mat3 mat3_inverse(const Mat3& m);                        // inverse (:181-210)
// ext_to_int_transformation (io/import.rs:222-233) and int_to_ext_transformation
// (io/export.rs:34-45); dt = (rotation, tx, ty); false where NotNan arithmetic would panic
bool ext_to_int_transformation(const float ext[3], const float pre[3], float out[3]);
bool int_to_ext_transformation(const float in[3], const float pre[3], float out[3]);
std::vector<Point> convex_hull_from_points(std::vector<Point> points);  // convex_hull.rs:31-64
Circle compute_pole(const SPolygon& shape, const std::vector<Circle>& poles);  // pole.rs:52-77
std::vector<Circle> generate_surrogate_poles(const SPolygon& shape, const SPSurrogateConfig& cfg);  // pole.rs:12-47
std::vector<Edge> generate_piers(const SPolygon& shape, size_t n, const std::vector<Circle>& poles);  // piers.rs:25-94
void spolygon_generate_surrogate(SPolygon& s, const SPSurrogateConfig& cfg);  // simple_polygon.rs:73-80

// ---------------------------------------------------------------------------------------
// Hazards — collision_detection/hazards/*.rs
// ---------------------------------------------------------------------------------------
enum class EntityKind : uint32_t { PlacedItem = 0, Exterior = 1, Hole = 2, InferiorQualityZone = 3 };

struct HazardEntity {                       // hazard.rs:38-51
    EntityKind kind = EntityKind::Exterior;
    uint64_t a = 0;     // PlacedItem: item id; Hole: idx; InferiorQualityZone: quality
    uint64_t b = 0;     // PlacedItem: placed-item key; InferiorQualityZone: idx
    float dt[3] = {0, 0, 0};  // PlacedItem: DTransformation (rotation, tx, ty)
    bool operator==(const HazardEntity& o) const {
        return kind == o.kind && a == o.a && b == o.b && std::memcmp(dt, o.dt, sizeof dt) == 0;
    }
    GeoPosition scope() const {             // hazard.rs:55-62
        return kind == EntityKind::Exterior ? GeoPosition::Exterior : GeoPosition::Interior;
    }
};

struct Hazard {                             // hazard.rs:16-23
    HazardEntity entity;
    SPolygon shape;
    bool dynamic = false;
};

// HazKey: a slotmap key.  The slot index is what the CDE reports; the version tags reuse.
struct HazKey {
    uint32_t idx = 0, version = 0;
    bool operator==(const HazKey& o) const { return idx == o.idx && version == o.version; }
};

// slotmap::SlotMap<HazKey, Hazard> — index 0 is the sentinel, freed slots are reused LIFO.
struct HazardMap {
    struct Slot { uint32_t version = 0; bool occupied = false; uint32_t next_free = 0; Hazard value; };
    std::vector<Slot> slots;
    uint32_t free_head = 1;
    size_t count = 0;
    HazardMap() { slots.resize(1); }
    HazKey insert(Hazard h);
    bool contains(HazKey k) const {
        return k.idx < slots.size() && slots[k.idx].occupied && slots[k.idx].version == k.version;
    }
    const Hazard& at(HazKey k) const;
    Hazard remove(HazKey k);
    template <class F> void for_each(F f) const {
        for (uint32_t i = 1; i < slots.size(); i++)
            if (slots[i].occupied) f(HazKey{i, slots[i].version}, slots[i].value);
    }
};

// filter.rs:7-9 HazardFilter — the oracle represents every filter by the set of irrelevant
// slot indices (NoFilter = empty, HazKey self-filter = one key, HazKeyFilter = many).
struct HazardFilter {
    std::vector<uint8_t> irrelevant;   // indexed by slot idx
    bool is_irrelevant(HazKey k) const { return k.idx < irrelevant.size() && irrelevant[k.idx]; }
};

// collector.rs:44-67 BasicHazardCollector (SecondaryMap keyed by slot).
struct HazardCollector {
    std::vector<uint8_t> present;
    std::vector<uint32_t> keys;
    void reset(size_t n_slots) {
        for (uint32_t k : keys) present[k] = 0;
        keys.clear();
        if (present.size() < n_slots) present.resize(n_slots, 0);
    }
    bool contains_key(HazKey k) const { return k.idx < present.size() && present[k.idx]; }
    void insert(HazKey k) {
        if (k.idx >= present.size()) present.resize(k.idx + 1, 0);
        if (!present[k.idx]) { present[k.idx] = 1; keys.push_back(k.idx); }
    }
    size_t len() const { return keys.size(); }
};

// ---------------------------------------------------------------------------------------
// Quadtree — collision_detection/quadtree/*.rs
// ---------------------------------------------------------------------------------------
enum class Presence : uint8_t { None = 0, Partial = 1, Entire = 2 };

struct QTHazPartial {                       // qt_partial_hazard.rs:7-12
    std::vector<Edge> edges;
    Rect ff_bbox{};
};

struct QTHazard {                           // qt_hazard.rs:13-22
    Rect qt_bbox{};
    HazKey hkey;
    HazardEntity entity;
    Presence presence = Presence::None;
    QTHazPartial partial;                   // valid when presence == Partial
    size_t n_edges() const { return presence == Presence::Partial ? partial.edges.size() : 0; }
};

struct QTHazardVec {                        // qt_hazard_vec.rs:13-120
    std::vector<QTHazard> hazards;
    size_t n_active_edges = 0;
    void add(QTHazard h);
    bool remove(HazKey k);
    const QTHazard* strongest(const HazardFilter& f) const {
        for (const auto& h : hazards)
            if (!f.is_irrelevant(h.hkey)) return &h;
        return nullptr;
    }
    bool no_partial_hazards() const {
        for (const auto& h : hazards)
            if (h.presence == Presence::Partial) return false;
        return true;
    }
};

// Traversal counters: the per-query work the algorithmic-bytes figure is built from.
struct Counters {
    uint64_t nodes = 0;          // quadtree nodes visited
    uint64_t entries = 0;        // hazard entries examined
    uint64_t edges = 0;          // hazard edges tested
    uint64_t pip_vertices = 0;   // polygon vertices read by containment PIP
};

struct QTNode {                             // qt_node.rs:14-25
    uint8_t level;
    Rect bbox;
    std::unique_ptr<std::array<QTNode, 4>> children;
    QTHazardVec hazards;
    uint8_t cd_threshold;
    QTNode() : level(0), bbox{}, cd_threshold(0) {}
    QTNode(uint8_t lv, Rect b, uint8_t cdt) : level(lv), bbox(b), cd_threshold(cdt) {}
    void register_hazard(QTHazard h, const HazardMap& map);
    void deregister_hazard(HazKey k);
    // Circle / Edge queries (qt_node.rs:99-147 and 151-191)
    bool collides_circle(const Circle& c, const HazardFilter& f, Counters* ct) const;
    bool collides_edge(const Edge& e, const HazardFilter& f, Counters* ct) const;
    void collect_circle(const Circle& c, HazardCollector& col, Counters* ct) const;
    void collect_edge(const Edge& e, HazardCollector& col, Counters* ct) const;
};

// qt_traits.rs:19-106
std::array<bool, 4> edge_collides_with_quadrants(const Edge& e, const Rect& r, const std::array<Rect, 4>& qs);

struct CDEConfig {                          // cd_engine.rs:345-353
    uint8_t quadtree_depth;
    uint8_t cd_threshold;
    SPSurrogateConfig item_surrogate_config;
};

// Result of one detect query: 0 = clear, 1 = collision, 2 = the reference aborts
// (unreachable! at cd_engine.rs:179: a filtered Entire hazard in the virtual root).
enum : uint8_t { Q_CLEAR = 0, Q_COLLISION = 1, Q_PANIC_FILTERED_ENTIRE = 2 };

struct CDEngine {                           // cd_engine.rs:21-31
    QTNode quadtree;
    HazardMap hazards_map;
    CDEConfig config;
    HazKey hkey_exterior;

    CDEngine(Rect bbox, std::vector<Hazard> static_hazards, CDEConfig cfg);
    HazKey register_hazard(Hazard h);
    Hazard deregister_hazard_by_entity(const HazardEntity& e);
    Hazard deregister_hazard_by_key(HazKey k);
    std::vector<Hazard> save() const;                       // CDESnapshot
    void restore(const std::vector<Hazard>& dynamic_hazards);
    Rect bbox() const { return quadtree.bbox; }
    const QTNode& get_virtual_root(const Rect& b) const;

    uint8_t detect_poly_collision(const SPolygon& shape, const HazardFilter& f, Counters* ct = nullptr) const;
    bool detect_surrogate_collision(const SPSurrogate& s, const Transformation& t, const HazardFilter& f,
                                    Counters* ct = nullptr) const;
    bool detect_containment_collision(const SPolygon& shape, const SPolygon& haz_shape, const HazardEntity& e,
                                      Counters* ct = nullptr) const;
    void collect_poly_collisions(const SPolygon& shape, HazardCollector& col, Counters* ct = nullptr) const;
    void collect_surrogate_collisions(const SPSurrogate& s, const Transformation& t, HazardCollector& col,
                                      Counters* ct = nullptr) const;
};

// util/assertions.rs invariants (return true when they hold).
bool qt_contains_no_dangling_hazards(const CDEngine& cde);
bool qt_hazard_vecs_sorted_and_cached(const CDEngine& cde);
bool qt_nodes_match(const QTNode* a, const QTNode* b);

}  // namespace oracle
