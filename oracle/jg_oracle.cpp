// TEST INFRASTRUCTURE ONLY — CPU oracle of the jagua-rs CDE (see jg_oracle.hpp).
//
// Each function cites the reference file:line it restates (paths relative to the jagua-rs
// crate root, jagua-rs/src/).  Third-party arithmetic restated here from its published
// algorithm: float-cmp 0.10 approx_eq (fpa_eq in the header), ndarray 0.17 linspace
// (linspace_f32 below), slotmap 1.0 key allocation (HazardMap), Rust core binary_search_by
// (QTHazardVec::add).  libm sinf/cosf are called from glibc, which is what Rust std uses on
// x86_64-linux-gnu.
#include "jg_oracle.hpp"

#include <algorithm>
#include <deque>
#include <stdexcept>

namespace oracle {

// transformation.rs:131-139 rot_transl_m, with f32::sin_cos = (sinf, cosf)
Transformation transformation_from_dt(float rotation, float tx, float ty) {
    float s = sinf(rotation);
    float c = cosf(rotation);
    Transformation t;
    t.m[0][0] = c; t.m[0][1] = -s; t.m[0][2] = tx;
    t.m[1][0] = s; t.m[1][1] = c;  t.m[1][2] = ty;
    return t;
}

// ---------------------------------------------------------------------------------------
// SPolygon
// ---------------------------------------------------------------------------------------

// simple_polygon.rs:137-150 (shoelace, f32 accumulation in vertex order)
float spolygon_calculate_area(const std::vector<Point>& pts) {
    float sigma = 0.0f;
    size_t n = pts.size();
    for (size_t i = 0; i < n; i++) {
        size_t j = (i + 1) % n;
        sigma += (pts[i].y + pts[j].y) * (pts[i].x - pts[j].x);
    }
    return 0.5f * sigma;
}

// simple_polygon.rs:122-133
Rect spolygon_bbox(const std::vector<Point>& pts) {
    float x_min = FLT_MAX, y_min = FLT_MAX, x_max = -FLT_MAX, y_max = -FLT_MAX;
    for (const auto& p : pts) {
        x_min = fmin32(x_min, p.x);
        y_min = fmin32(y_min, p.y);
        x_max = fmax32(x_max, p.x);
        y_max = fmax32(y_max, p.y);
    }
    if (!(x_min < x_max && y_min < y_max)) throw std::runtime_error("invalid rectangle");
    return {x_min, y_min, x_max, y_max};
}

// convex_hull.rs:62-64
static float hull_cross(Point a, Point b, Point c) {
    return (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x);
}
// convex_hull.rs:53-60
static void grow_convex_hull(std::vector<Point>& h, Point next) {
    while (h.size() >= 2 && hull_cross(h[h.size() - 2], h[h.size() - 1], next) <= 0.0f) h.pop_back();
    h.push_back(next);
}
// convex_hull.rs:31-51 (monotone chain; the sort is stable, keyed on x)
std::vector<Point> convex_hull_from_points(std::vector<Point> pts) {
    std::stable_sort(pts.begin(), pts.end(), [](const Point& a, const Point& b) { return a.x < b.x; });
    std::vector<Point> lower, upper;
    for (const auto& p : pts) grow_convex_hull(lower, p);
    for (auto it = pts.rbegin(); it != pts.rend(); ++it) grow_convex_hull(upper, *it);
    upper.pop_back();
    lower.pop_back();
    lower.insert(lower.end(), upper.begin(), upper.end());
    return lower;
}

// simple_polygon.rs:107-120
float spolygon_calculate_diameter(std::vector<Point> pts) {
    auto ch = convex_hull_from_points(std::move(pts));
    if (ch.size() < 2) throw std::runtime_error("convex hull is empty");
    float best = 0.0f;
    bool first = true;
    for (size_t i = 0; i < ch.size(); i++)
        for (size_t j = i + 1; j < ch.size(); j++) {
            float d = point_sq_distance(ch[i], ch[j]);
            if (first || d >= best) { best = d; first = false; }
        }
    return std::sqrt(best);
}

// simple_polygon.rs:173-192
Point spolygon_centroid(const SPolygon& s) {
    float area = s.area;
    float c_x = 0.0f, c_y = 0.0f;
    size_t n = s.n();
    for (size_t i = 0; i < n; i++) {
        size_t j = (i == n - 1) ? 0 : i + 1;
        float x_i = s.vertices[i].x, y_i = s.vertices[i].y;
        float x_j = s.vertices[j].x, y_j = s.vertices[j].y;
        c_x += (x_i + x_j) * (x_i * y_j - x_j * y_i);
        c_y += (y_i + y_j) * (x_i * y_j - x_j * y_i);
    }
    c_x /= 6.0f * area;
    c_y /= 6.0f * area;
    return {c_x, c_y};
}

// simple_polygon.rs:255-292 (ray cast with FPA handling of rays through vertices)
bool spolygon_collides_point(const SPolygon& s, const Point& p) {
    if (!rect_collides_point(s.bbox, p)) return false;
    Point outside{s.bbox.x_max + rect_width(s.bbox), p.y};
    Edge ray{p, outside};
    int n_int = 0;
    size_t n = s.n();
    for (size_t i = 0; i < n; i++) {
        Edge e = s.edge(i);
        float s_x = e.start.x, s_y = e.start.y, e_x = e.end.x, e_y = e.end.y;
        if ((fpa_eq(s_y, p.y) && fpa_gt(s_x, p.x)) || (fpa_eq(e_y, p.y) && fpa_gt(e_x, p.x))) {
            if (fpa_lt(s_y, p.y) || fpa_lt(e_y, p.y)) n_int += 1;
        } else if (edge_intersection(ray, e)) {
            n_int += 1;
        }
    }
    return n_int % 2 == 1;
}

// pole.rs:81-135 — one node of the pole-of-inaccessibility search
namespace {
struct POINode {
    size_t level;
    Rect bbox;
    float radius;
    float distance;
};
POINode poi_node_new(Rect bbox, size_t level, const SPolygon& poly, const std::vector<Circle>& poles) {
    float radius = rect_diameter(bbox) / 2.0f;
    Point c = rect_centroid(bbox);
    bool inside = spolygon_collides_point(poly, c);
    if (inside)
        for (const auto& pc : poles)
            if (circle_collides_point(pc, c)) { inside = false; break; }
    float d = FLT_MAX;
    for (size_t i = 0; i < poly.n(); i++) d = fmin32(d, edge_distance_to(poly.edge(i), c));
    for (const auto& pc : poles) d = fmin32(d, circle_separation_distance_point(pc, c));
    return {level, bbox, radius, inside ? d : -d};
}
}  // namespace

constexpr size_t MAX_POI_TREE_DEPTH = 10;  // pole.rs:79

// pole.rs:52-77
Circle compute_pole(const SPolygon& shape, const std::vector<Circle>& poles) {
    Rect sq = rect_inflate_to_square(shape.bbox);
    std::deque<POINode> queue;
    queue.push_back(poi_node_new(sq, MAX_POI_TREE_DEPTH, shape, poles));
    bool have_best = false;
    Circle best{};
    while (!queue.empty()) {
        POINode node = queue.front();
        queue.pop_front();
        float best_d = have_best ? best.radius : 0.0f;
        if (node.distance > best_d) {
            best = Circle{rect_centroid(node.bbox), node.distance};
            have_best = true;
        }
        best_d = have_best ? best.radius : 0.0f;
        if (node.radius + node.distance > best_d && node.level > 0) {
            auto qs = rect_quadrants(node.bbox);
            for (int i = 0; i < 4; i++) queue.push_back(poi_node_new(qs[i], node.level - 1, shape, poles));
        }
    }
    if (!have_best) throw std::runtime_error("no pole found");
    return best;
}

// simple_polygon.rs:41-71
SPolygon spolygon_new(std::vector<Point> pts) {
    if (pts.size() < 3) throw std::runtime_error("simple polygon must have at least 3 points");
    for (size_t i = 0; i < pts.size(); i++)
        for (size_t j = i + 1; j < pts.size(); j++)
            if (bits(pts[i].x) == bits(pts[j].x) && bits(pts[i].y) == bits(pts[j].y))   // unique() hashes bits
                throw std::runtime_error("simple polygon should not contain duplicate points");
    float area = spolygon_calculate_area(pts);
    if (area == 0.0f) throw std::runtime_error("simple polygon has no area");
    if (area < 0.0f) {
        std::reverse(pts.begin(), pts.end());
        area = -area;
    }
    SPolygon s;
    s.diameter = spolygon_calculate_diameter(pts);
    s.bbox = spolygon_bbox(pts);
    s.vertices = std::move(pts);
    // calculate_poi (:152-171): a dummy polygon carrying the signed-area recomputation
    SPolygon dummy;
    dummy.vertices = s.vertices;
    dummy.bbox = s.bbox;
    dummy.area = spolygon_calculate_area(s.vertices);
    dummy.diameter = s.diameter;
    s.poi = compute_pole(dummy, {});
    s.area = area;
    return s;
}

// simple_polygon.rs:195-224 (transform: vertices, poi, surrogate, regenerated bbox)
SPolygon spolygon_transform_clone(const SPolygon& s, const Transformation& t) {
    SPolygon r = s;
    for (auto& p : r.vertices) p = transform_point(p, t);
    r.poi.center = transform_point(r.poi.center, t);
    if (r.surrogate.present) {
        for (auto& c : r.surrogate.poles) c.center = transform_point(c.center, t);
        for (auto& e : r.surrogate.piers) { e.start = transform_point(e.start, t); e.end = transform_point(e.end, t); }
    }
    r.bbox = spolygon_bbox(r.vertices);
    return r;
}

// pole.rs:12-47
std::vector<Circle> generate_surrogate_poles(const SPolygon& shape, const SPSurrogateConfig& cfg) {
    std::vector<Circle> all{shape.poi};
    float total = circle_area(shape.poi);
    for (;;) {
        Circle next = compute_pole(shape, all);
        total += circle_area(next);
        all.push_back(next);
        float cov = total / shape.area;
        bool have = false;
        uint32_t active = 0;
        for (int i = 0; i < 3; i++) {
            if (cov > cfg.n_pole_limits_cov[i]) {
                if (!have || cfg.n_pole_limits_n[i] < active) { active = cfg.n_pole_limits_n[i]; have = true; }
            }
        }
        if (have && all.size() >= active) break;
        if (all.size() >= 1000) throw std::runtime_error("more than 1000 poles");
    }
    return all;
}

// ndarray 0.17 Array::linspace: start + step * i, step = (end - start) / (n - 1)
static std::vector<float> linspace_f32(float a, float b, size_t n) {
    float step = n > 1 ? (b - a) / (float)(n - 1) : 0.0f;
    std::vector<float> v(n);
    for (size_t i = 0; i < n; i++) v[i] = a + step * (float)i;
    return v;
}

// transformation.rs:156-178 (3x3 product, left-to-right sums)
static void mat3_dot(const float l[3][3], const float r[3][3], float out[3][3]) {
    float o[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) o[i][j] = l[i][0] * r[0][j] + l[i][1] * r[1][j] + l[i][2] * r[2][j];
    std::memcpy(out, o, sizeof o);
}

// circle.rs:136-147
static float circle_distance_to_point(const Circle& c, const Point& p) {
    float dx = p.x - c.center.x, dy = p.y - c.center.y;
    float sq_d = dx * dx + dy * dy;
    if (sq_d < c.radius * c.radius) return 0.0f;
    return std::sqrt(sq_d) - c.radius;
}

// piers.rs:25-226 with the release-build constants (:16-20)
std::vector<Edge> generate_piers(const SPolygon& shape, size_t n, const std::vector<Circle>& poles) {
    if (n == 0) return {};
    const size_t RAYS_PER_ANGLE = 200, N_ANGLES = 90, N_POINTS = 100;
    const float CLIPPING_TRIM = 0.999f, ACTION_RADIUS_RATIO = 0.10f;
    Rect bbox = shape.bbox;
    Rect ex = rect_inflate_to_square(bbox);
    Point cen = spolygon_centroid(shape);
    Edge base{Point{cen.x, cen.y - 2.0f * rect_height(ex)}, Point{cen.x, cen.y + 2.0f * rect_height(ex)}};
    // generate_ray_transformations (:96-121)
    float dx = rect_width(ex) / (float)RAYS_PER_ANGLE;
    auto angles = linspace_f32(0.0f, 3.14159265358979323846f, N_ANGLES + 1);
    std::vector<Edge> clipped;
    for (size_t ai = 0; ai < N_ANGLES; ai++) {
        float s = sinf(angles[ai]), c = cosf(angles[ai]);
        float rot[3][3] = {{c, -s, 0.0f}, {s, c, 0.0f}, {0.0f, 0.0f, 1.0f}};
        for (size_t i = 0; i < RAYS_PER_ANGLE; i++) {
            float x = ex.x_min + dx * (float)i;
            float tr[3][3] = {{1.0f, 0.0f, x}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}};
            float m[3][3];
            mat3_dot(rot, tr, m);
            Transformation t;
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 3; b++) t.m[a][b] = m[a][b];
            Edge ray{transform_point(base.start, t), transform_point(base.end, t)};
            // clip (:124-151)
            if (spolygon_collides_point(shape, ray.start) || spolygon_collides_point(shape, ray.end))
                throw std::runtime_error("ray end inside shape");
            std::vector<Point> hits;
            for (size_t k = 0; k < shape.n(); k++) {
                Point loc;
                if (edge_intersection(shape.edge(k), ray, &loc)) hits.push_back(loc);
            }
            std::vector<float> keys(hits.size());
            std::vector<size_t> order(hits.size());
            for (size_t k = 0; k < hits.size(); k++) { keys[k] = point_distance(ray.start, hits[k]); order[k] = k; }
            std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return keys[a] < keys[b]; });
            for (size_t k = 0; k + 1 < order.size(); k += 2) {
                Point st = hits[order[k]], en = hits[order[k + 1]];
                if (st.x == en.x && st.y == en.y) continue;
                float edx = en.x - st.x, edy = en.y - st.y;
                Edge e{st, en};
                e.start.x -= edx * (CLIPPING_TRIM - 1.0f) / 2.0f;
                e.start.y -= edy * (CLIPPING_TRIM - 1.0f) / 2.0f;
                e.end.x += edx * (CLIPPING_TRIM - 1.0f) / 2.0f;
                e.end.y += edy * (CLIPPING_TRIM - 1.0f) / 2.0f;
                clipped.push_back(e);
            }
        }
    }
    // generate_unrepresented_point_grid (:153-172)
    auto xs = linspace_f32(ex.x_min, ex.x_max, N_POINTS);
    auto ys = linspace_f32(ex.y_min, ex.y_max, N_POINTS);
    std::vector<Point> grid;
    for (float x : xs)
        for (float y : ys) {
            Point p{x, y};
            if (!spolygon_collides_point(shape, p)) continue;
            bool free = true;
            for (const auto& pc : poles)
                if (circle_collides_point(pc, p)) { free = false; break; }
            if (free) grid.push_back(p);
        }
    float radius_infl = ACTION_RADIUS_RATIO * rect_width(ex);
    float bw = rect_width(bbox), bh = rect_height(bbox);
    float forfeit = std::sqrt((bw * bw) * (bh * bh));
    std::vector<Edge> selected;
    for (size_t it = 0; it < n; it++) {
        std::vector<float> md_rays(grid.size()), md_poles(grid.size());
        for (size_t g = 0; g < grid.size(); g++) {
            float a = forfeit;
            for (const auto& r : selected) a = fmin32(a, edge_distance_to(r, grid[g]));
            md_rays[g] = a;
            float b = forfeit;
            for (const auto& pc : poles) b = fmin32(b, circle_distance_to_point(pc, grid[g]));
            md_poles[g] = b;
        }
        if (clipped.empty()) throw std::runtime_error("no ray found");
        size_t best_i = 0;
        float best_l = 0.0f;
        for (size_t r = 0; r < clipped.size(); r++) {
            float loss = -0.0f;
            for (size_t g = 0; g < grid.size(); g++) {
                float d_new = edge_distance_to(clipped[r], grid[g]);
                float md_ray = fmin32(md_rays[g], d_new);
                float v = md_ray < radius_infl ? fmin32(md_poles[g], md_ray) : md_poles[g];
                loss += v * v;
            }
            if (r == 0 || loss < best_l) { best_l = loss; best_i = r; }
        }
        selected.push_back(clipped[best_i]);
    }
    return selected;
}

// sp_surrogate.rs:33-53 and simple_polygon.rs:73-80
void spolygon_generate_surrogate(SPolygon& s, const SPSurrogateConfig& cfg) {
    SPSurrogate sur;
    auto hull = convex_hull_from_points(s.vertices);
    for (const auto& p : hull)
        for (size_t i = 0; i < s.n(); i++)
            if (s.vertices[i].x == p.x && s.vertices[i].y == p.y) { sur.convex_hull_indices.push_back(i); break; }
    sur.convex_hull_area = spolygon_calculate_area(hull);
    sur.poles = generate_surrogate_poles(s, cfg);
    size_t n_ff = std::min<size_t>(cfg.n_ff_poles, sur.poles.size());
    std::vector<Circle> rel(sur.poles.begin(), sur.poles.begin() + n_ff);
    sur.piers = generate_piers(s, cfg.n_ff_piers, rel);
    sur.config = cfg;
    sur.present = true;
    s.surrogate = std::move(sur);
}

// ---------------------------------------------------------------------------------------
// HazardMap (slotmap 1.0 semantics)
// ---------------------------------------------------------------------------------------
HazKey HazardMap::insert(Hazard h) {
    uint32_t idx;
    if (free_head < slots.size()) {
        idx = free_head;
        free_head = slots[idx].next_free;
    } else {
        idx = (uint32_t)slots.size();
        slots.emplace_back();
        free_head = idx + 1;
    }
    Slot& s = slots[idx];
    s.version = s.version + 1;   // occupied slots carry odd versions in slotmap; only equality matters
    s.occupied = true;
    s.value = std::move(h);
    count++;
    return HazKey{idx, s.version};
}
const Hazard& HazardMap::at(HazKey k) const {
    if (!contains(k)) throw std::runtime_error("invalid HazKey");
    return slots[k.idx].value;
}
Hazard HazardMap::remove(HazKey k) {
    if (!contains(k)) throw std::runtime_error("Cannot deregister hazard that is not registered");
    Slot& s = slots[k.idx];
    Hazard h = std::move(s.value);
    s.value = Hazard{};
    s.occupied = false;
    s.version += 1;
    s.next_free = free_head;
    free_head = k.idx;
    count--;
    return h;
}

// ---------------------------------------------------------------------------------------
// QTHazardVec — qt_hazard_vec.rs
// ---------------------------------------------------------------------------------------
static int presence_key(Presence p) { return (int)p; }

// :24-43 — core::slice::binary_search_by with order_by_descending_strength (:89-100)
void QTHazardVec::add(QTHazard h) {
    if (h.presence == Presence::None) throw std::runtime_error("adding None hazard");
    auto cmp = [&](const QTHazard& probe) {   // -1 Less, 0 Equal, 1 Greater (already reversed)
        int a = presence_key(probe.presence), b = presence_key(h.presence);
        if (a == b) return 0;
        return a < b ? 1 : -1;
    };
    size_t pos;
    size_t size = hazards.size();
    if (size == 0) {
        pos = 0;
    } else {
        size_t base = 0;
        while (size > 1) {
            size_t half = size / 2;
            size_t mid = base + half;
            base = cmp(hazards[mid]) == 1 ? base : mid;
            size -= half;
        }
        int c = cmp(hazards[base]);
        pos = c == 0 ? base : base + (c == -1 ? 1 : 0);
    }
    n_active_edges += h.n_edges();
    hazards.insert(hazards.begin() + (long)pos, std::move(h));
}
// :45-55
bool QTHazardVec::remove(HazKey k) {
    for (size_t i = 0; i < hazards.size(); i++)
        if (hazards[i].hkey == k) {
            n_active_edges -= hazards[i].n_edges();
            hazards.erase(hazards.begin() + (long)i);
            return true;
        }
    return false;
}

// ---------------------------------------------------------------------------------------
// QTHazard::constrict — qt_hazard.rs:48-167 and QTHazPartial (qt_partial_hazard.rs:15-60)
// ---------------------------------------------------------------------------------------
static QTHazPartial partial_from_entire_shape(const SPolygon& s) {
    QTHazPartial p;
    for (size_t i = 0; i < s.n(); i++) p.edges.push_back(s.edge(i));
    p.ff_bbox = s.bbox;
    return p;
}
static QTHazPartial partial_from_parent(const QTHazPartial& parent, std::vector<Edge> edges) {
    QTHazPartial p;
    if (parent.edges.size() == edges.size()) {
        p.ff_bbox = parent.ff_bbox;
    } else {
        float x_min = INFINITY, y_min = INFINITY, x_max = -INFINITY, y_max = -INFINITY;
        for (const auto& e : edges) {
            x_min = fmin32(fmin32(x_min, e.start.x), e.end.x);
            y_min = fmin32(fmin32(y_min, e.start.y), e.end.y);
            x_max = fmax32(fmax32(x_max, e.start.x), e.end.x);
            y_max = fmax32(fmax32(y_max, e.start.y), e.end.y);
        }
        if (x_min < x_max && y_min < y_max) p.ff_bbox = Rect{x_min, y_min, x_max, y_max};
        else p.ff_bbox = parent.ff_bbox;
    }
    p.edges = std::move(edges);
    return p;
}

static std::array<QTHazard, 4> constrict(const QTHazard& h, const std::array<Rect, 4>& qs, const HazardMap& map) {
    std::array<QTHazard, 4> out;
    if (h.presence == Presence::None) throw std::runtime_error("Hazard presence cannot be None in a QTHazard");
    if (h.presence == Presence::Entire) {
        for (auto& o : out) o = h;
        return out;
    }
    const SPolygon& shape = map.at(h.hkey).shape;
    int enclosed = -1;
    for (int i = 0; i < 4; i++)
        if (rect_relation_to(shape.bbox, qs[i]) == GeoRelation::Enclosed) { enclosed = i; break; }
    if (enclosed >= 0) {
        for (int i = 0; i < 4; i++) {
            out[i].qt_bbox = qs[i];
            out[i].hkey = h.hkey;
            out[i].entity = h.entity;
            if (i == enclosed) { out[i].presence = h.presence; out[i].partial = h.partial; }
            else out[i].presence = Presence::None;
        }
        return out;
    }
    bool resolved[4] = {false, false, false, false};
    for (int i = 0; i < 4; i++) {
        std::vector<Edge> col;
        for (const auto& e : h.partial.edges)
            if (rect_collides_edge(qs[i], e)) col.push_back(e);
        if (!col.empty()) {
            out[i].qt_bbox = qs[i];
            out[i].hkey = h.hkey;
            out[i].entity = h.entity;
            out[i].presence = Presence::Partial;
            out[i].partial = partial_from_parent(h.partial, std::move(col));
            resolved[i] = true;
        }
    }
    for (int i = 0; i < 4; i++) {
        if (resolved[i]) continue;
        bool none_n = false, entire_n = false;
        for (int k = 0; k < 2; k++) {
            int idx = QUADRANT_NEIGHBOR_LAYOUT[i][k];
            if (!resolved[idx]) continue;
            if (out[idx].presence == Presence::None) none_n = true;
            if (out[idx].presence == Presence::Entire) entire_n = true;
        }
        Presence p;
        if (none_n && entire_n) throw std::runtime_error("quadrant with both None and Entire neighbours");
        if (none_n) p = Presence::None;
        else if (entire_n) p = Presence::Entire;
        else {
            bool colliding = spolygon_collides_point(shape, rect_centroid(qs[i]));
            GeoPosition sc = h.entity.scope();
            if (sc == GeoPosition::Interior && colliding) p = Presence::Entire;
            else if (sc == GeoPosition::Exterior && !colliding) p = Presence::Entire;
            else p = Presence::None;
        }
        out[i].qt_bbox = qs[i];
        out[i].hkey = h.hkey;
        out[i].entity = h.entity;
        out[i].presence = p;
        resolved[i] = true;
    }
    return out;
}

// ---------------------------------------------------------------------------------------
// QTNode — qt_node.rs
// ---------------------------------------------------------------------------------------
static void constrict_and_register(const QTHazard& h, std::array<QTNode, 4>& ch, const HazardMap& map) {
    std::array<Rect, 4> bbs{ch[0].bbox, ch[1].bbox, ch[2].bbox, ch[3].bbox};
    auto hs = constrict(h, bbs, map);
    for (int i = 0; i < 4; i++)
        if (hs[i].presence != Presence::None) ch[i].register_hazard(std::move(hs[i]), map);
}

// :38-81
void QTNode::register_hazard(QTHazard h, const HazardMap& map) {
    if (!children && level > 0 && h.presence == Presence::Partial) {
        auto qs = rect_quadrants(bbox);
        children = std::make_unique<std::array<QTNode, 4>>();
        for (int i = 0; i < 4; i++) {
            (*children)[i].level = (uint8_t)(level - 1);
            (*children)[i].bbox = qs[i];
            (*children)[i].cd_threshold = cd_threshold;
        }
        for (const auto& prev : hazards.hazards) constrict_and_register(prev, *children, map);
    }
    if (children) constrict_and_register(h, *children, map);
    hazards.add(std::move(h));
}

// :83-94
void QTNode::deregister_hazard(HazKey k) {
    bool modified = hazards.remove(k);
    if (modified) {
        if (hazards.no_partial_hazards()) children.reset();
        else if (children)
            for (auto& c : *children) c.deregister_hazard(k);
    }
}

// qt_traits.rs:144-177 — returns 0 none, 1 first half, 2 halfway, 3 second half (of e2)
static int edge_intersection_half(const Edge& e1, const Edge& e2) {
    float x1 = e1.start.x, y1 = e1.start.y, x2 = e1.end.x, y2 = e1.end.y;
    float x3 = e2.start.x, y3 = e2.start.y, x4 = e2.end.x, y4 = e2.end.y;
    float t_nom = (x2 - x4) * (y4 - y3) - (y2 - y4) * (x4 - x3);
    float t_denom = (x2 - x1) * (y4 - y3) - (y2 - y1) * (x4 - x3);
    float u_nom = (x2 - x4) * (y2 - y1) - (y2 - y4) * (x2 - x1);
    float u_denom = (x2 - x1) * (y4 - y3) - (y2 - y1) * (x4 - x3);
    if (t_denom == 0.0f || u_denom == 0.0f) return 0;
    float t = t_nom / t_denom;
    float u = u_nom / u_denom;
    if (t >= 0.0f && t <= 1.0f && u >= 0.0f && u <= 1.0f) {
        if (u > 0.5f) return 1;
        if (u < 0.5f) return 3;
        return 2;
    }
    return 0;
}

// qt_traits.rs:110-140; cq: -1 undetermined (None), 0 Some(false), 1 Some(true)
static void half_intersect(const Edge& e1, const Edge& e2, int* cq, std::initializer_list<int> fst,
                           std::initializer_list<int> sec) {
    bool any_none = false;
    for (int i : fst) any_none |= cq[i] < 0;
    for (int i : sec) any_none |= cq[i] < 0;
    if (!any_none) return;
    int loc = edge_intersection_half(e1, e2);
    if (loc == 1 || loc == 2)
        for (int i : fst) cq[i] = 1;
    if (loc == 2 || loc == 3)
        for (int i : sec) cq[i] = 1;
}

// qt_traits.rs:19-106
std::array<bool, 4> edge_collides_with_quadrants(const Edge& e, const Rect& r, const std::array<Rect, 4>& qs) {
    float ex0 = edge_x_min(e), ex1 = edge_x_max(e), ey0 = edge_y_min(e), ey1 = edge_y_max(e);
    int cq[4];
    for (int i = 0; i < 4; i++) {
        const Rect& q = qs[i];
        bool x_no = fmax32(ex0, q.x_min) > fmin32(ex1, q.x_max);
        bool y_no = fmax32(ey0, q.y_min) > fmin32(ey1, q.y_max);
        if (x_no || y_no) cq[i] = 0;
        else if (rect_collides_point(q, e.start) || rect_collides_point(q, e.end)) cq[i] = 1;
        else cq[i] = -1;
    }
    if (cq[0] >= 0 && cq[1] >= 0 && cq[2] >= 0 && cq[3] >= 0)
        return {cq[0] == 1, cq[1] == 1, cq[2] == 1, cq[3] == 1};
    Point c = rect_centroid(r);
    auto sides = rect_sides(r);
    const Edge& top = sides[0];
    const Edge& left = sides[1];
    const Edge& bottom = sides[2];
    const Edge& right = sides[3];
    Edge h_bisect{Point{r.x_min, c.y}, Point{r.x_max, c.y}};
    Edge v_bisect{Point{c.x, r.y_min}, Point{c.x, r.y_max}};
    half_intersect(e, left, cq, {1}, {2});
    half_intersect(e, right, cq, {3}, {0});
    half_intersect(e, top, cq, {0}, {1});
    half_intersect(e, bottom, cq, {2}, {3});
    half_intersect(e, h_bisect, cq, {1, 2}, {0, 3});
    half_intersect(e, v_bisect, cq, {2, 3}, {0, 1});
    return {cq[0] == 1, cq[1] == 1, cq[2] == 1, cq[3] == 1};
}

// qt_partial_hazard.rs:67-72 for Circle and Edge entities
static bool partial_collides_circle(const QTHazPartial& p, const Circle& c, Counters* ct) {
    if (!circle_collides_rect(c, p.ff_bbox)) return false;
    for (const auto& e : p.edges) {
        if (ct) ct->edges++;
        if (circle_collides_edge(c, e)) return true;
    }
    return false;
}
static bool partial_collides_edge(const QTHazPartial& p, const Edge& q, Counters* ct) {
    if (!rect_collides_edge(p.ff_bbox, q)) return false;
    for (const auto& e : p.edges) {
        if (ct) ct->edges++;
        if (edge_intersection(q, e)) return true;
    }
    return false;
}

// qt_node.rs:99-147 for a Circle (default collides_with_quadrants, qt_traits.rs:10-13)
bool QTNode::collides_circle(const Circle& c, const HazardFilter& f, Counters* ct) const {
    if (ct) ct->nodes++;
    const QTHazard* s = nullptr;
    for (const auto& h : hazards.hazards) {
        if (ct) ct->entries++;
        if (!f.is_irrelevant(h.hkey)) { s = &h; break; }
    }
    if (!s) return false;
    if (s->presence == Presence::Entire) return true;
    if (children) {
        bool cq[4];
        for (int i = 0; i < 4; i++) cq[i] = circle_collides_rect(c, (*children)[i].bbox);
        for (int i = 0; i < 4; i++)
            if (cq[i] && (*children)[i].collides_circle(c, f, ct)) return true;
        return false;
    }
    for (const auto& h : hazards.hazards) {
        if (f.is_irrelevant(h.hkey)) continue;
        if (ct) ct->entries++;
        if (h.presence == Presence::Entire) throw std::runtime_error("should have been handled above");
        if (h.presence == Presence::Partial && partial_collides_circle(h.partial, c, ct)) return true;
    }
    return false;
}

// qt_node.rs:99-147 for an Edge (specialised collides_with_quadrants, qt_traits.rs:19-106)
bool QTNode::collides_edge(const Edge& e, const HazardFilter& f, Counters* ct) const {
    if (ct) ct->nodes++;
    const QTHazard* s = nullptr;
    for (const auto& h : hazards.hazards) {
        if (ct) ct->entries++;
        if (!f.is_irrelevant(h.hkey)) { s = &h; break; }
    }
    if (!s) return false;
    if (s->presence == Presence::Entire) return true;
    if (children) {
        std::array<Rect, 4> qs{(*children)[0].bbox, (*children)[1].bbox, (*children)[2].bbox, (*children)[3].bbox};
        auto cq = edge_collides_with_quadrants(e, bbox, qs);
        for (int i = 0; i < 4; i++)
            if (cq[i] && (*children)[i].collides_edge(e, f, ct)) return true;
        return false;
    }
    for (const auto& h : hazards.hazards) {
        if (f.is_irrelevant(h.hkey)) continue;
        if (ct) ct->entries++;
        if (h.presence == Presence::Entire) throw std::runtime_error("should have been handled above");
        if (h.presence == Presence::Partial && partial_collides_edge(h.partial, e, ct)) return true;
    }
    return false;
}

// qt_node.rs:151-191
void QTNode::collect_circle(const Circle& c, HazardCollector& col, Counters* ct) const {
    if (ct) ct->nodes++;
    bool now = hazards.n_active_edges <= (size_t)cd_threshold;
    if (children && !now) {
        bool cq[4];
        for (int i = 0; i < 4; i++) cq[i] = circle_collides_rect(c, (*children)[i].bbox);
        for (int i = 0; i < 4; i++)
            if (cq[i]) (*children)[i].collect_circle(c, col, ct);
        return;
    }
    for (const auto& h : hazards.hazards) {
        if (ct) ct->entries++;
        if (col.contains_key(h.hkey)) continue;
        if (h.presence == Presence::Entire) col.insert(h.hkey);
        else if (h.presence == Presence::Partial && partial_collides_circle(h.partial, c, ct)) col.insert(h.hkey);
    }
}
void QTNode::collect_edge(const Edge& e, HazardCollector& col, Counters* ct) const {
    if (ct) ct->nodes++;
    bool now = hazards.n_active_edges <= (size_t)cd_threshold;
    if (children && !now) {
        std::array<Rect, 4> qs{(*children)[0].bbox, (*children)[1].bbox, (*children)[2].bbox, (*children)[3].bbox};
        auto cq = edge_collides_with_quadrants(e, bbox, qs);
        for (int i = 0; i < 4; i++)
            if (cq[i]) (*children)[i].collect_edge(e, col, ct);
        return;
    }
    for (const auto& h : hazards.hazards) {
        if (ct) ct->entries++;
        if (col.contains_key(h.hkey)) continue;
        if (h.presence == Presence::Entire) col.insert(h.hkey);
        else if (h.presence == Presence::Partial && partial_collides_edge(h.partial, e, ct)) col.insert(h.hkey);
    }
}

// ---------------------------------------------------------------------------------------
// CDEngine — cd_engine.rs
// ---------------------------------------------------------------------------------------
static QTHazard qt_hazard_from_root(Rect root_bbox, const Hazard& h, HazKey k) {  // qt_hazard.rs:36-43
    QTHazard q;
    q.qt_bbox = root_bbox;
    q.hkey = k;
    q.entity = h.entity;
    q.presence = Presence::Partial;
    q.partial = partial_from_entire_shape(h.shape);
    return q;
}

// :34-56
CDEngine::CDEngine(Rect bbox, std::vector<Hazard> static_hazards, CDEConfig cfg)
    : quadtree(cfg.quadtree_depth, bbox, cfg.cd_threshold), config(cfg) {
    for (auto& h : static_hazards) {
        HazKey k = hazards_map.insert(std::move(h));
        quadtree.register_hazard(qt_hazard_from_root(quadtree.bbox, hazards_map.at(k), k), hazards_map);
    }
    bool found = false;
    hazards_map.for_each([&](HazKey k, const Hazard& h) {
        if (!found && h.entity.kind == EntityKind::Exterior) { hkey_exterior = k; found = true; }
    });
    if (!found) throw std::runtime_error("No exterior hazard registered in the CDE");
}

// :59-69
HazKey CDEngine::register_hazard(Hazard h) {
    HazKey k = hazards_map.insert(std::move(h));
    quadtree.register_hazard(qt_hazard_from_root(bbox(), hazards_map.at(k), k), hazards_map);
    return k;
}
// :72-85
Hazard CDEngine::deregister_hazard_by_entity(const HazardEntity& e) {
    bool found = false;
    HazKey key;
    hazards_map.for_each([&](HazKey k, const Hazard& h) {
        if (!found && h.entity == e) { key = k; found = true; }
    });
    if (!found) throw std::runtime_error("Cannot deregister hazard that is not registered");
    quadtree.deregister_hazard(key);
    return hazards_map.remove(key);
}
// :87-96
Hazard CDEngine::deregister_hazard_by_key(HazKey k) {
    Hazard h = hazards_map.remove(k);
    quadtree.deregister_hazard(k);
    return h;
}
// :98-106
std::vector<Hazard> CDEngine::save() const {
    std::vector<Hazard> out;
    hazards_map.for_each([&](HazKey, const Hazard& h) {
        if (h.dynamic) out.push_back(h);
    });
    return out;
}
// :109-149 (swap_remove order preserved: it decides the slots reused afterwards)
void CDEngine::restore(const std::vector<Hazard>& snap) {
    std::vector<std::pair<HazKey, HazardEntity>> to_remove;
    hazards_map.for_each([&](HazKey k, const Hazard& h) {
        if (h.dynamic) to_remove.push_back({k, h.entity});
    });
    std::vector<Hazard> to_add;
    for (const auto& h : snap) {
        long pos = -1;
        for (size_t i = 0; i < to_remove.size(); i++)
            if (to_remove[i].second == h.entity) { pos = (long)i; break; }
        if (pos >= 0) {
            to_remove[(size_t)pos] = to_remove.back();
            to_remove.pop_back();
        } else {
            to_add.push_back(h);
        }
    }
    for (auto& r : to_remove) deregister_hazard_by_key(r.first);
    for (auto& h : to_add) register_hazard(h);
}

// :314-327
const QTNode& CDEngine::get_virtual_root(const Rect& b) const {
    const QTNode* v = &quadtree;
    while (v->children) {
        const QTNode* next = nullptr;
        for (const auto& c : *v->children)
            if (rect_relation_to(c.bbox, b) == GeoRelation::Surrounding) { next = &c; break; }
        if (!next) break;
        v = next;
    }
    return *v;
}

// :231-255
bool CDEngine::detect_containment_collision(const SPolygon& shape, const SPolygon& haz_shape,
                                            const HazardEntity& e, Counters* ct) const {
    GeoRelation rel = rect_almost_relation_to(haz_shape.bbox, shape.bbox);
    bool contained = false;
    if (rel == GeoRelation::Surrounding) {
        if (ct) ct->pip_vertices += haz_shape.n();
        contained = spolygon_collides_point(haz_shape, shape.poi.center);
    } else if (rel == GeoRelation::Enclosed) {
        if (ct) ct->pip_vertices += shape.n();
        contained = spolygon_collides_point(shape, haz_shape.poi.center);
    }
    GeoPosition sc = e.scope();
    return (sc == GeoPosition::Interior && contained) || (sc == GeoPosition::Exterior && !contained);
}

// :160-197 — returns Q_PANIC_FILTERED_ENTIRE where the reference reaches unreachable!()
uint8_t CDEngine::detect_poly_collision(const SPolygon& shape, const HazardFilter& f, Counters* ct) const {
    if (rect_relation_to(bbox(), shape.bbox) != GeoRelation::Surrounding) return Q_COLLISION;
    const QTNode& v = get_virtual_root(shape.bbox);
    for (size_t i = 0; i < shape.n(); i++)
        if (v.collides_edge(shape.edge(i), f, ct)) return Q_COLLISION;
    for (const auto& h : v.hazards.hazards) {
        if (h.presence == Presence::Entire) return Q_PANIC_FILTERED_ENTIRE;
        if (h.presence == Presence::Partial && !f.is_irrelevant(h.hkey)) {
            if (ct) ct->entries++;
            if (detect_containment_collision(shape, hazards_map.at(h.hkey).shape, h.entity, ct)) return Q_COLLISION;
        }
    }
    return Q_CLEAR;
}

// :204-223
bool CDEngine::detect_surrogate_collision(const SPSurrogate& s, const Transformation& t, const HazardFilter& f,
                                          Counters* ct) const {
    if (s.poles.size() < s.config.n_ff_poles) throw std::runtime_error("ff_poles slice out of range");
    for (uint32_t i = 0; i < s.config.n_ff_poles; i++) {
        Circle c{transform_point(s.poles[i].center, t), s.poles[i].radius};
        if (quadtree.collides_circle(c, f, ct)) return true;
    }
    for (const auto& p : s.piers) {
        Edge e{transform_point(p.start, t), transform_point(p.end, t)};
        if (quadtree.collides_edge(e, f, ct)) return true;
    }
    return false;
}

// :261-289
void CDEngine::collect_poly_collisions(const SPolygon& shape, HazardCollector& col, Counters* ct) const {
    if (rect_relation_to(bbox(), shape.bbox) != GeoRelation::Surrounding) col.insert(hkey_exterior);
    const QTNode& v = get_virtual_root(shape.bbox);
    for (size_t i = 0; i < shape.n(); i++) v.collect_edge(shape.edge(i), col, ct);
    for (const auto& h : v.hazards.hazards) {
        if (h.presence != Presence::Partial) continue;
        if (col.contains_key(h.hkey)) continue;
        if (ct) ct->entries++;
        if (detect_containment_collision(shape, hazards_map.at(h.hkey).shape, h.entity, ct)) col.insert(h.hkey);
    }
}

// :296-310
void CDEngine::collect_surrogate_collisions(const SPSurrogate& s, const Transformation& t, HazardCollector& col,
                                           Counters* ct) const {
    if (s.poles.size() < s.config.n_ff_poles) throw std::runtime_error("ff_poles slice out of range");
    for (uint32_t i = 0; i < s.config.n_ff_poles; i++) {
        Circle c{transform_point(s.poles[i].center, t), s.poles[i].radius};
        quadtree.collect_circle(c, col, ct);
    }
    for (const auto& p : s.piers) {
        Edge e{transform_point(p.start, t), transform_point(p.end, t)};
        quadtree.collect_edge(e, col, ct);
    }
}

// ---------------------------------------------------------------------------------------
// util/assertions.rs
// ---------------------------------------------------------------------------------------
static bool node_no_dangling(const QTNode& node, const QTNode& parent) {   // :63-89
    for (const auto& h : node.hazards.hazards) {
        bool found = false;
        for (const auto& ph : parent.hazards.hazards)
            if (ph.entity == h.entity) { found = true; break; }
        if (!found) return false;
    }
    if (node.children)
        for (const auto& c : *node.children)
            if (!node_no_dangling(c, node)) return false;
    return true;
}
bool qt_contains_no_dangling_hazards(const CDEngine& cde) {   // :52-61
    if (cde.quadtree.children)
        for (const auto& c : *cde.quadtree.children)
            if (!node_no_dangling(c, cde.quadtree)) return false;
    return true;
}

static bool vec_ok(const QTNode& n) {   // qt_hazard_vec.rs:102-120 and assertions.rs:32-50
    size_t sum = 0;
    for (size_t i = 0; i < n.hazards.hazards.size(); i++) {
        const auto& h = n.hazards.hazards[i];
        if (h.presence == Presence::None) return false;
        if (i > 0 && (int)n.hazards.hazards[i - 1].presence < (int)h.presence) return false;
        sum += h.n_edges();
    }
    if (sum != n.hazards.n_active_edges) return false;
    if (n.children)
        for (const auto& c : *n.children)
            if (!vec_ok(c)) return false;
    return true;
}
bool qt_hazard_vecs_sorted_and_cached(const CDEngine& cde) { return vec_ok(cde.quadtree); }

// :106-209 — per node, the set of (entity, presence) pairs must agree
static bool same_hazard_sets(const QTNode& a, const QTNode& b) {
    if (a.hazards.hazards.size() != b.hazards.hazards.size()) return false;
    for (const auto& ha : a.hazards.hazards) {
        bool f = false;
        for (const auto& hb : b.hazards.hazards)
            if (ha.entity == hb.entity && ha.presence == hb.presence) { f = true; break; }
        if (!f) return false;
    }
    return true;
}
static bool has_partial(const QTNode* n) {
    if (!n) return false;
    for (const auto& h : n->hazards.hazards)
        if (h.presence == Presence::Partial) return true;
    return false;
}
bool qt_nodes_match(const QTNode* a, const QTNode* b) {
    if (a && b) {
        if (!same_hazard_sets(*a, *b)) return false;
    } else if (a) {
        if (!a->hazards.hazards.empty()) return false;
    } else if (b) {
        if (!b->hazards.hazards.empty()) return false;
    } else {
        throw std::runtime_error("Both nodes are none");
    }
    const auto* ca = a ? a->children.get() : nullptr;
    const auto* cb = b ? b->children.get() : nullptr;
    if (!ca && !cb) return true;
    if (ca && !cb) {
        if (has_partial(a))
            for (const auto& c : *ca)
                if (!qt_nodes_match(&c, nullptr)) return false;
        return true;
    }
    if (!ca && cb) {
        if (has_partial(b))
            for (const auto& c : *cb)
                if (!qt_nodes_match(nullptr, &c)) return false;
        return true;
    }
    for (int i = 0; i < 4; i++)
        if (!qt_nodes_match(&(*ca)[i], &(*cb)[i])) return false;
    return true;
}

// ---------------------------------------------------------------------------------------------
// instance import (io/import.rs) and the transformation algebra it uses (transformation.rs)
// ---------------------------------------------------------------------------------------------
std::vector<Point> import_simple_polygon_points(std::vector<Point> pts) {
    if (pts.size() > 1 && pts.front().x == pts.back().x && pts.front().y == pts.back().y) pts.pop_back();
    // eliminate_degenerate_vertices (:235-260): mark i when point i ~ point i+1 (cyclic), remove
    // the marks from the highest index down
    const size_t n = pts.size();
    std::vector<size_t> marks;
    for (size_t i = 0; i < n; i++) {
        const size_t j = (i + 1) % n;
        if (fpa_eq(pts[i].x, pts[j].x) && fpa_eq(pts[i].y, pts[j].y)) marks.push_back(i);
    }
    std::sort(marks.rbegin(), marks.rend());
    for (size_t i : marks)
        if (i < pts.size()) pts.erase(pts.begin() + (long)i);
    for (size_t i = 0; i < pts.size(); i++)
        for (size_t j = i + 1; j < pts.size(); j++)
            if (bits(pts[i].x) == bits(pts[j].x) && bits(pts[i].y) == bits(pts[j].y))
                throw std::runtime_error("Simple polygon has non-consecutive duplicate vertices");
    return pts;
}

Mat3 mat3_rot_transl(float angle, float tx, float ty) {
    const float sin = std::sin(angle), cos = std::cos(angle);
    return Mat3{{{cos, -sin, tx}, {sin, cos, ty}, {0.0f, 0.0f, 1.0f}}};
}

Mat3 mat3_dot(const Mat3& l, const Mat3& r) {
    Mat3 o;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) o.m[i][j] = l.m[i][0] * r.m[0][j] + l.m[i][1] * r.m[1][j] + l.m[i][2] * r.m[2][j];
    return o;
}

Mat3 mat3_inverse(const Mat3& a) {
    const auto& m = a.m;
    const float det = m[0][0] * m[1][1] * m[2][2] + m[0][1] * m[1][2] * m[2][0] + m[0][2] * m[1][0] * m[2][1] -
                      m[0][2] * m[1][1] * m[2][0] - m[0][1] * m[1][0] * m[2][2] - m[0][0] * m[1][2] * m[2][1];
    return Mat3{{{(m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det, (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det,
                  (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det},
                 {(m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det, (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det,
                  (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det},
                 {(m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det, (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det,
                  (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det}}};
}

static const Mat3 MAT3_EMPTY{{{1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}}};

static bool decompose(const Mat3& t, float out[3]) {                 // transformation.rs:83-88
    for (auto& row : t.m)
        for (float v : row)
            if (v != v) return false;
    out[0] = std::atan2(t.m[1][0], t.m[0][0]);
    out[1] = t.m[0][2];
    out[2] = t.m[1][2];
    return true;
}

bool ext_to_int_transformation(const float ext[3], const float pre[3], float out[3]) {
    // Transformation::empty().transform(&pre.compose().inverse()).transform_from_decomposed(ext)
    const Mat3 t1 = mat3_dot(mat3_inverse(mat3_rot_transl(pre[0], pre[1], pre[2])), MAT3_EMPTY);
    return decompose(mat3_dot(mat3_rot_transl(ext[0], ext[1], ext[2]), t1), out);
}

bool int_to_ext_transformation(const float in[3], const float pre[3], float out[3]) {
    // Transformation::empty().transform_from_decomposed(pre).transform_from_decomposed(int)
    const Mat3 t1 = mat3_dot(mat3_rot_transl(pre[0], pre[1], pre[2]), MAT3_EMPTY);
    return decompose(mat3_dot(mat3_rot_transl(in[0], in[1], in[2]), t1), out);
}

}  // namespace oracle
