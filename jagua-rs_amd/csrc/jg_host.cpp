// Host build of the collision engine: SPolygon construction with its surrogate, the hazard
// slot map, the quadtree node pool under register/deregister, the breadth-first flattener
// (tree dumps) and the pool-ordered device image with its in-place patches.
// Arithmetic follows the reference's f32 operation order (compiled -ffp-contract=off, no
// fast-math, sinf/cosf as run-time libm calls).
#include "jg_host.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

namespace jg {

// ---------------------------------------------------------------------------------------
// polygons
// ---------------------------------------------------------------------------------------

// transformation.rs:131-139 (sin_cos = libm sinf, cosf)
Mat mat_from_dt(float rotation, float tx, float ty) {
    float s = sinf(rotation);
    float c = cosf(rotation);
    return Mat{c, -s, tx, s, c, ty};
}

// Transformation algebra (transformation.rs:104-210): 3x3 matrices in f32, every product and
// sum in the written order (products with the identity included: they decide zero signs).
using M3 = float[3][3];
static void rot_transl_m(float angle, float tx, float ty, M3 m) {     // :132-141
    const float s = sinf(angle), c = cosf(angle);
    const float r[3][3] = {{c, -s, tx}, {s, c, ty}, {0.0f, 0.0f, 1.0f}};
    std::memcpy(m, r, sizeof r);
}
static void dot_prod(const M3 l, const M3 r, M3 o) {                 // :158-179
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) o[i][j] = l[i][0] * r[0][j] + l[i][1] * r[1][j] + l[i][2] * r[2][j];
}
static void inverse3(const M3 m, M3 o) {                              // :181-210
    const float det = m[0][0] * m[1][1] * m[2][2] + m[0][1] * m[1][2] * m[2][0] + m[0][2] * m[1][0] * m[2][1] -
                      m[0][2] * m[1][1] * m[2][0] - m[0][1] * m[1][0] * m[2][2] - m[0][0] * m[1][2] * m[2][1];
    o[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det;
    o[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
    o[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
    o[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det;
    o[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
    o[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
    o[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det;
    o[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
    o[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
}

bool dt_chain(const float a[3], bool invert_a, const float b[3], float out[3]) {
    M3 A, X, T1, B, T2;
    const M3 I = {{1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}};
    rot_transl_m(a[0], a[1], a[2], A);
    if (invert_a) inverse3(A, X);
    else std::memcpy(X, A, sizeof A);
    dot_prod(X, I, T1);                      // Transformation::empty().transform(X) (:60-63)
    rot_transl_m(b[0], b[1], b[2], B);
    dot_prod(B, T1, T2);                     // .transform_from_decomposed(b) (:48-51, :65-67)
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (T2[i][j] != T2[i][j]) return false;   // NotNan arithmetic would panic
    out[0] = atan2f(T2[1][0], T2[0][0]);     // decompose (:83-88)
    out[1] = T2[0][2];
    out[2] = T2[1][2];
    return true;
}

static float signed_area(const std::vector<float2>& v) {           // simple_polygon.rs:137-150
    float sigma = 0.0f;
    const size_t n = v.size();
    for (size_t i = 0; i < n; i++) {
        const float2 a = v[i], b = v[(i + 1) % n];
        sigma += (a.y + b.y) * (a.x - b.x);
    }
    return 0.5f * sigma;
}

static bool bbox_of(const std::vector<float2>& v, float4& out) {   // simple_polygon.rs:122-133
    float4 b = make_float4(FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (const float2& p : v) {
        b.x = mn(b.x, p.x);
        b.y = mn(b.y, p.y);
        b.z = mx(b.z, p.x);
        b.w = mx(b.w, p.y);
    }
    out = b;
    return b.x < b.z && b.y < b.w;
}

bool polygon_has_point(const Polygon& p, float x, float y) {        // simple_polygon.rs:255-292
    return poly_has_point(p.v.data(), p.n(), p.bbox, x, y);
}

// Monotone-chain hull (convex_hull.rs:31-64): stable x-sort, lower then upper chain.
static std::vector<float2> hull_points(std::vector<float2> pts) {
    std::stable_sort(pts.begin(), pts.end(), [](const float2& a, const float2& b) { return a.x < b.x; });
    auto turn = [](float2 a, float2 b, float2 c) { return (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x); };
    auto chain = [&](auto begin, auto end) {
        std::vector<float2> h;
        for (auto it = begin; it != end; ++it) {
            while (h.size() >= 2 && turn(h[h.size() - 2], h.back(), *it) <= 0.0f) h.pop_back();
            h.push_back(*it);
        }
        h.pop_back();
        return h;
    };
    std::vector<float2> lo = chain(pts.begin(), pts.end());
    std::vector<float2> up = chain(pts.rbegin(), pts.rend());
    lo.insert(lo.end(), up.begin(), up.end());
    return lo;
}

static float sq_dist(float2 a, float2 b) {                          // point.rs:49-52
    const float dx = a.x - b.x, dy = a.y - b.y;
    return dx * dx + dy * dy;
}

// Pole of inaccessibility search (fail_fast/pole.rs:52-135): breadth-first over a quadtree of
// the square bbox, 10 levels, pruned by radius + distance.
static bool compute_pole(const Polygon& shape, const std::vector<float4>& poles, float4& out) {
    struct Cell { float4 box; int level; float radius, dist; };
    auto make_cell = [&](float4 box, int level) {
        Cell c{box, level, 0.0f, 0.0f};
        c.dist = pole_cell(shape.v.data(), shape.n(), shape.bbox, poles.data(), (uint32_t)poles.size(), box, c.radius);
        return c;
    };
    std::vector<Cell> fifo;
    fifo.push_back(make_cell(box_inflate_to_square(shape.bbox), POLE_LEVELS));
    bool have = false;
    float4 best{};
    for (size_t head = 0; head < fifo.size(); head++) {
        const Cell cell = fifo[head];
        float best_r = have ? best.z : 0.0f;
        if (cell.dist > best_r) {
            best = make_float4((cell.box.x + cell.box.z) / 2.0f, (cell.box.y + cell.box.w) / 2.0f, cell.dist, 0.0f);
            have = true;
            best_r = cell.dist;
        }
        if (cell.radius + cell.dist > best_r && cell.level > 0)
            for (int k = 0; k < 4; k++) fifo.push_back(make_cell(box_quadrant(cell.box, k), cell.level - 1));
        if (head > 4096 && head * 2 > fifo.size()) {             // keep the FIFO compact
            fifo.erase(fifo.begin(), fifo.begin() + (long)head + 1);
            head = (size_t)-1;
        }
    }
    out = best;
    return have;
}

// SPolygon::new (simple_polygon.rs:41-71) up to, not including, the pole of inaccessibility
bool polygon_prepare(const float* xy, uint32_t n, Polygon& out, std::string& err) {
    if (n < 3) { err = "simple polygon must have at least 3 points"; return false; }
    std::vector<float2> v(n);
    for (uint32_t i = 0; i < n; i++) v[i] = make_float2(xy[2 * i], xy[2 * i + 1]);
    {   // itertools unique() hashes Point by its bit patterns (point.rs:57-63): 0.0 and -0.0 differ
        std::vector<uint64_t> s(n);
        for (uint32_t i = 0; i < n; i++) s[i] = (uint64_t)f2u(v[i].x) << 32 | f2u(v[i].y);
        std::sort(s.begin(), s.end());
        for (uint32_t i = 1; i < n; i++)
            if (s[i] == s[i - 1]) { err = "simple polygon should not contain duplicate points"; return false; }
    }
    float area = signed_area(v);
    if (area == 0.0f) { err = "simple polygon has no area"; return false; }
    if (area < 0.0f) {
        std::reverse(v.begin(), v.end());
        area = -area;
    }
    Polygon p;
    // diameter: largest squared distance over convex-hull pairs (:107-120)
    std::vector<float2> h = hull_points(v);
    if (h.size() < 2) { err = "convex hull is empty"; return false; }
    float best = -1.0f;
    for (size_t i = 0; i < h.size(); i++)
        for (size_t j = i + 1; j < h.size(); j++) best = mx(best, sq_dist(h[i], h[j]));
    p.diameter = sqrtf(best);
    if (!bbox_of(v, p.bbox)) { err = "invalid rectangle"; return false; }
    p.v = std::move(v);
    p.area = area;
    out = std::move(p);
    return true;
}

bool polygon_new(const float* xy, uint32_t n, Polygon& out, std::string& err) {   // simple_polygon.rs:41-71
    Polygon p;
    if (!polygon_prepare(xy, n, p, err)) return false;
    float4 poi;
    if (!compute_pole(p, {}, poi)) { err = "no pole found"; return false; }
    p.poi = poi;
    out = std::move(p);
    return true;
}

// generate_surrogate_poles (pole.rs:12-47)
static bool surrogate_poles(const Polygon& shape, const jg_surrogate_config& cfg, std::vector<float4>& out) {
    const float PI_F = 3.14159265358979323846f;
    out.assign(1, shape.poi);
    float covered = shape.poi.z * shape.poi.z * PI_F;
    for (;;) {
        float4 next;
        if (!compute_pole(shape, out, next)) return false;
        covered += next.z * next.z * PI_F;
        out.push_back(next);
        const float coverage = covered / shape.area;
        bool any = false;
        uint32_t limit = 0;
        for (int i = 0; i < 3; i++)
            if (coverage > cfg.n_pole_limits_cov[i] && (!any || cfg.n_pole_limits_n[i] < limit)) {
                limit = cfg.n_pole_limits_n[i];
                any = true;
            }
        if (any && out.size() >= limit) return true;
        if (out.size() >= 1000) return false;
    }
}

// SPolygon::centroid (simple_polygon.rs:173-192)
float2 polygon_centroid(const Polygon& shape) {
    float gx = 0.0f, gy = 0.0f;
    for (uint32_t i = 0; i < shape.n(); i++) {
        const float2 a = shape.v[i], b = shape.v[i + 1 == shape.n() ? 0 : i + 1];
        gx += (a.x + b.x) * (a.x * b.y - b.x * a.y);
        gy += (a.y + b.y) * (a.x * b.y - b.x * a.y);
    }
    return make_float2(gx / (6.0f * shape.area), gy / (6.0f * shape.area));
}

void pier_setup(const Polygon& shape, PierSetup& o) {
    const float4 bb = shape.bbox;
    o.sq = box_inflate_to_square(bb);
    const float2 g = polygon_centroid(shape);
    const float sq_h = o.sq.w - o.sq.y;
    o.r0 = make_float2(g.x, g.y - 2.0f * sq_h);
    o.r1 = make_float2(g.x, g.y + 2.0f * sq_h);
    o.dx = (o.sq.z - o.sq.x) / (float)PIER_RAYS;
    o.infl = PIER_INFL * (o.sq.z - o.sq.x);
    const float bw = bb.z - bb.x, bh = bb.w - bb.y;
    o.forfeit = sqrtf((bw * bw) * (bh * bh));
}

void pier_angles(float* sc) {
    const float PI_F = 3.14159265358979323846f;
    for (uint32_t ai = 0; ai < PIER_ANGLES; ai++) {
        const float ang = linspace_at(0.0f, PI_F, PIER_ANGLES + 1, ai);
        sc[2 * ai] = sinf(ang);
        sc[2 * ai + 1] = cosf(ang);
    }
}

// generate_piers (fail_fast/piers.rs:25-226), release-build constants (:16-20)
static bool surrogate_piers(const Polygon& shape, uint32_t n, const std::vector<float4>& poles,
                            std::vector<float4>& out) {
    out.clear();
    if (n == 0) return true;
    PierSetup ps;
    pier_setup(shape, ps);
    float sc[2 * PIER_ANGLES];
    pier_angles(sc);
    std::vector<float4> clipped;
    for (uint32_t ai = 0; ai < PIER_ANGLES; ai++)
        for (uint32_t i = 0; i < PIER_RAYS; i++) {
            const float4 ray = pier_ray(sc[2 * ai], sc[2 * ai + 1], ps.sq.x + ps.dx * (float)i, ps.r0, ps.r1);
            if (polygon_has_point(shape, ray.x, ray.y) || polygon_has_point(shape, ray.z, ray.w)) return false;
            // clip (:124-151): intersections sorted by distance from the ray start
            std::vector<std::pair<float, float2>> hits;
            for (uint32_t k = 0; k < shape.n(); k++) {
                const float4 e = shape.edge(k);
                if (!edge_edge(e, ray)) continue;
                float d;
                const float2 loc = pier_hit(e, ray, d);
                hits.push_back({d, loc});
            }
            std::stable_sort(hits.begin(), hits.end(),
                             [](const std::pair<float, float2>& u, const std::pair<float, float2>& w) { return u.first < w.first; });
            for (size_t k = 0; k + 1 < hits.size(); k += 2) {
                const float2 st = hits[k].second, en = hits[k + 1].second;
                if (st.x == en.x && st.y == en.y) continue;
                clipped.push_back(pier_trim(st, en));
            }
        }
    std::vector<float2> grid;
    std::vector<float> d_pole;
    for (uint32_t i = 0; i < PIER_GRID; i++)
        for (uint32_t j = 0; j < PIER_GRID; j++) {
            float dp;
            float2 p;
            if (pier_grid_point(shape.v.data(), shape.n(), shape.bbox, poles.data(), (uint32_t)poles.size(), ps, i, j, p, dp)) {
                grid.push_back(p);
                d_pole.push_back(dp);
            }
        }
    std::vector<float> d_ray(grid.size(), ps.forfeit);
    for (uint32_t it = 0; it < n; it++) {
        if (clipped.empty()) return false;
        size_t best_i = 0;
        float best_loss = 0.0f;
        for (size_t r = 0; r < clipped.size(); r++) {
            float loss = -0.0f;
            for (size_t g = 0; g < grid.size(); g++)
                loss += pier_point_loss(d_ray[g], d_pole[g], edge_point_dist(clipped[r], grid[g].x, grid[g].y), ps.infl);
            if (r == 0 || loss < best_loss) { best_loss = loss; best_i = r; }
        }
        out.push_back(clipped[best_i]);
        // min_distances_to_rays (:215-225) folds min over the chosen piers; min is exact, so
        // folding in the new pier gives the same values
        for (size_t g = 0; g < grid.size(); g++)
            d_ray[g] = mn(d_ray[g], edge_point_dist(clipped[best_i], grid[g].x, grid[g].y));
    }
    return true;
}

bool polygon_generate_surrogate(Polygon& p, const jg_surrogate_config& cfg, std::string& err) {  // sp_surrogate.rs:33-53
    if (p.sur.present && std::memcmp(&p.sur.cfg, &cfg, sizeof cfg) == 0) return true;
    std::vector<float4> poles;
    if (!surrogate_poles(p, cfg, poles)) { err = "pole generation failed"; return false; }
    return surrogate_finish(p, cfg, std::move(poles), err);
}

bool surrogate_finish(Polygon& p, const jg_surrogate_config& cfg, std::vector<float4> poles, std::string& err,
                      const std::vector<float4>* piers) {
    Surrogate s;
    std::vector<float2> h = hull_points(p.v);
    for (const float2& q : h)
        for (uint32_t i = 0; i < p.n(); i++)
            if (p.v[i].x == q.x && p.v[i].y == q.y) { s.hull_idx.push_back(i); break; }
    s.hull_area = signed_area(h);
    s.poles = std::move(poles);
    const size_t n_ff = std::min<size_t>(cfg.n_ff_poles, s.poles.size());
    std::vector<float4> rel(s.poles.begin(), s.poles.begin() + (long)n_ff);
    if (piers) s.piers = *piers;
    else if (!surrogate_piers(p, cfg.n_ff_piers, rel, s.piers)) { err = "pier generation failed"; return false; }
    s.cfg = cfg;
    s.present = true;
    p.sur = std::move(s);
    return true;
}

Polygon polygon_transform(const Polygon& p, const Mat& m) {        // simple_polygon.rs:195-224
    Polygon r = p;
    for (float2& q : r.v) q = xform(m, q.x, q.y);
    const float2 c = xform(m, p.poi.x, p.poi.y);
    r.poi.x = c.x;
    r.poi.y = c.y;
    if (r.sur.present) {
        for (float4& c4 : r.sur.poles) {
            const float2 t = xform(m, c4.x, c4.y);
            c4.x = t.x;
            c4.y = t.y;
        }
        for (float4& e : r.sur.piers) {
            const float2 a = xform(m, e.x, e.y), b = xform(m, e.z, e.w);
            e = make_float4(a.x, a.y, b.x, b.y);
        }
    }
    bbox_of(r.v, r.bbox);
    return r;
}

// ---------------------------------------------------------------------------------------
// hazards
// ---------------------------------------------------------------------------------------
bool entity_eq(const jg_hazard_entity& a, const jg_hazard_entity& b) {
    return a.kind == b.kind && a.a == b.a && a.b == b.b && std::memcmp(a.dt, b.dt, sizeof a.dt) == 0;
}

jg_hazkey SlotMap::insert(HazardRec r) {
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
    s.version += 1;
    s.occ = true;
    s.rec = std::move(r);
    return jg_hazkey{idx, s.version};
}

HazardRec SlotMap::remove(jg_hazkey k) {
    Slot& s = slots[k.idx];
    HazardRec r = std::move(s.rec);
    s.rec = HazardRec{};
    s.occ = false;
    s.version += 1;
    s.next_free = free_head;
    free_head = k.idx;
    return r;
}

// ---------------------------------------------------------------------------------------
// quadtree (qt_node.rs, qt_hazard.rs, qt_partial_hazard.rs, qt_hazard_vec.rs)
// ---------------------------------------------------------------------------------------
void QuadTree::init(uint8_t depth, float4 bbox, uint8_t cdt) {
    pool.assign(1, QNode{});
    pool[0].level = depth;
    pool[0].bbox = bbox;
    free_blocks.clear();
    touched.assign(1, 0);
    cd_threshold = cdt;
}

int32_t QuadTree::alloc_children(int32_t parent) {
    int32_t base;
    if (!free_blocks.empty()) {
        base = free_blocks.back();
        free_blocks.pop_back();
    } else {
        base = (int32_t)pool.size();
        pool.resize(pool.size() + 4);
    }
    const float4 pb = pool[parent].bbox;
    const uint8_t lv = (uint8_t)(pool[parent].level - 1);
    for (int k = 0; k < 4; k++) {
        QNode& c = pool[base + k];
        c = QNode{};
        c.level = lv;
        c.bbox = box_quadrant(pb, k);
        if (!(c.bbox.x < c.bbox.z && c.bbox.y < c.bbox.w)) c.bbox = c.bbox;  // degenerate boxes are kept as is
    }
    pool[parent].child = base;
    if ((uint32_t)(pool[0].level - pool[parent].level) < CELL_LEVELS_MAX) cells_dirty.push_back(parent);
    touched.push_back(parent);
    for (int k = 0; k < 4; k++) touched.push_back(base + k);
    return base;
}

void QuadTree::free_subtree_children(int32_t node) {
    const int32_t base = pool[node].child;
    if (base < 0) return;
    for (int k = 0; k < 4; k++) free_subtree_children(base + k);
    for (int k = 0; k < 4; k++) pool[base + k] = QNode{};
    free_blocks.push_back(base);
    pool[node].child = -1;
    if ((uint32_t)(pool[0].level - pool[node].level) < CELL_LEVELS_MAX) cells_dirty.push_back(node);
    touched.push_back(node);
}

QEntry root_entry(const Polygon& s, jg_hazkey k) {     // QTHazard::from_root (qt_hazard.rs:36-43)
    QEntry e;
    e.slot = k.idx;
    e.version = k.version;
    e.presence = PRES_PARTIAL;
    e.ff = s.bbox;
    e.edges.resize(s.n());
    for (uint32_t i = 0; i < s.n(); i++) e.edges[i] = i;
    return e;
}

void QuadTree::compact_bfs() {
    std::vector<int32_t> order{0};
    for (size_t i = 0; i < order.size(); i++)
        if (pool[order[i]].child >= 0)
            for (int k = 0; k < 4; k++) order.push_back(pool[order[i]].child + k);
    std::vector<QNode> np(order.size());
    int32_t next_child = 1;
    for (size_t i = 0; i < order.size(); i++) {
        np[i] = std::move(pool[order[i]]);
        if (np[i].child >= 0) {
            np[i].child = next_child;
            next_child += 4;
        }
    }
    pool = std::move(np);
    free_blocks.clear();
    touched.clear();
}

static void add_entry(QNode& n, QEntry e) {                         // qt_hazard_vec.rs:24-43
    n.n_active_edges += e.presence == PRES_PARTIAL ? e.edges.size() : 0;
    if (e.presence == PRES_ENTIRE) {
        size_t pos = 0;
        while (pos < n.haz.size() && n.haz[pos].presence == PRES_ENTIRE) pos++;
        n.haz.insert(n.haz.begin() + (long)pos, std::move(e));
    } else {
        n.haz.push_back(std::move(e));
    }
}

// QTHazard::constrict (qt_hazard.rs:48-167) for one entry into the 4 quadrants `q`;
// present[k] = false means QTHazPresence::None.
static void constrict(const QEntry& e, const float4* q, const HazardRec& hz, QEntry* out, bool* present) {
    for (int k = 0; k < 4; k++) {
        out[k] = QEntry{};
        out[k].slot = e.slot;
        out[k].version = e.version;
        present[k] = false;
    }
    if (e.presence == PRES_ENTIRE) {
        for (int k = 0; k < 4; k++) { out[k].presence = PRES_ENTIRE; present[k] = true; }
        return;
    }
    const Polygon& shape = hz.shape;
    for (int k = 0; k < 4; k++)
        if (box_relation(shape.bbox, q[k]) == REL_ENCLOSED) {
            out[k] = e;
            present[k] = true;
            return;
        }
    bool resolved[4] = {false, false, false, false};
    for (int k = 0; k < 4; k++) {
        std::vector<uint32_t> sub;
        for (uint32_t ei : e.edges)
            if (box_edge(q[k], shape.edge(ei))) sub.push_back(ei);
        if (sub.empty()) continue;
        QEntry& o = out[k];
        o.presence = PRES_PARTIAL;
        if (sub.size() == e.edges.size()) {                       // qt_partial_hazard.rs:20-60
            o.ff = e.ff;
        } else {
            float4 b = make_float4(INFINITY, INFINITY, -INFINITY, -INFINITY);
            for (uint32_t ei : sub) {
                const float4 ed = shape.edge(ei);
                b.x = mn(mn(b.x, ed.x), ed.z);
                b.y = mn(mn(b.y, ed.y), ed.w);
                b.z = mx(mx(b.z, ed.x), ed.z);
                b.w = mx(mx(b.w, ed.y), ed.w);
            }
            o.ff = (b.x < b.z && b.y < b.w) ? b : e.ff;
        }
        o.edges = std::move(sub);
        present[k] = true;
        resolved[k] = true;
    }
    static const int NB[4][2] = {{1, 3}, {0, 2}, {1, 3}, {0, 2}};   // rect.rs:136
    for (int k = 0; k < 4; k++) {
        if (resolved[k]) continue;
        bool none_nb = false, entire_nb = false;
        for (int j = 0; j < 2; j++) {
            const int o = NB[k][j];
            if (!resolved[o]) continue;
            if (!present[o]) none_nb = true;
            else if (out[o].presence == PRES_ENTIRE) entire_nb = true;
        }
        bool entire;
        if (none_nb && !entire_nb) entire = false;
        else if (entire_nb && !none_nb) entire = true;
        else {
            const bool inside = polygon_has_point(shape, (q[k].x + q[k].z) / 2.0f, (q[k].y + q[k].w) / 2.0f);
            entire = hz.interior() ? inside : !inside;
        }
        if (entire) { out[k].presence = PRES_ENTIRE; present[k] = true; }
        resolved[k] = true;
    }
}

void QuadTree::push_to_children(int32_t node, const QEntry& e, const SlotMap& map) {
    const int32_t base = pool[node].child;
    float4 q[4];
    for (int k = 0; k < 4; k++) q[k] = pool[base + k].bbox;
    QEntry out[4];
    bool present[4];
    constrict(e, q, map.slots[e.slot].rec, out, present);
    for (int k = 0; k < 4; k++)
        if (present[k]) register_hazard(base + k, std::move(out[k]), map);
}

// QTNode::register_hazard (qt_node.rs:38-81)
void QuadTree::register_hazard(int32_t node, QEntry e, const SlotMap& map) {
    if (pool[node].child < 0 && pool[node].level > 0 && e.presence == PRES_PARTIAL) {
        alloc_children(node);
        const size_t n_prev = pool[node].haz.size();
        for (size_t i = 0; i < n_prev; i++) {
            QEntry prev = pool[node].haz[i];
            push_to_children(node, prev, map);
        }
    }
    if (pool[node].child >= 0) push_to_children(node, e, map);
    add_entry(pool[node], std::move(e));
    touched.push_back(node);
}

// QTNode::deregister_hazard (qt_node.rs:83-94)
void QuadTree::deregister_hazard(int32_t node, uint32_t slot, uint32_t version) {
    QNode& n = pool[node];
    size_t i = 0;
    while (i < n.haz.size() && !(n.haz[i].slot == slot && n.haz[i].version == version)) i++;
    if (i == n.haz.size()) return;
    n.n_active_edges -= n.haz[i].presence == PRES_PARTIAL ? n.haz[i].edges.size() : 0;
    n.haz.erase(n.haz.begin() + (long)i);
    touched.push_back(node);
    bool any_partial = false;
    for (const QEntry& h : n.haz) any_partial |= h.presence == PRES_PARTIAL;
    if (!any_partial) {
        free_subtree_children(node);
    } else if (n.child >= 0) {
        const int32_t base = n.child;
        for (int k = 0; k < 4; k++) deregister_hazard(base + k, slot, version);
    }
}

// ---------------------------------------------------------------------------------------
// breadth-first flattening (the tree dump's order)
// ---------------------------------------------------------------------------------------
void flatten(const QuadTree& qt, const SlotMap& map, FlatCDE& f) {
    f = FlatCDE{};
    std::vector<int32_t> order{0};
    std::vector<uint32_t> up{0};
    for (size_t i = 0; i < order.size(); i++) {
        const QNode& n = qt.pool[order[i]];
        if (n.child >= 0)
            for (int k = 0; k < 4; k++) {
                order.push_back(n.child + k);
                up.push_back((uint32_t)(i << 2) | (uint32_t)k);
            }
    }
    const size_t N = order.size();
    f.nodes.resize(N);
    f.node_up = up;
    size_t next_child = 1;
    for (size_t i = 0; i < N; i++) {
        const QNode& n = qt.pool[order[i]];
        FlatNode& fn = f.nodes[i];
        fn.bbox = n.bbox;
        fn.first_child = -1;
        if (n.child >= 0) {
            fn.first_child = (int32_t)next_child;
            next_child += 4;
        }
        fn.haz_begin = (uint32_t)f.ent.size();
        fn.haz_info = (uint32_t)n.haz.size() | (!n.haz.empty() && n.haz[0].presence == PRES_ENTIRE ? 0x80000000u : 0u);
        fn.n_active_edges = (uint32_t)std::min<uint64_t>(n.n_active_edges, 0xffffffffu);
        const bool tests_edges = n.child < 0 || n.n_active_edges <= qt.cd_threshold;
        for (const QEntry& h : n.haz) {
            FlatEntry fe{};
            fe.slot_flags = h.slot | (h.presence == PRES_ENTIRE ? 0x80000000u : 0u);
            fe.n_edges_total = h.presence == PRES_PARTIAL ? (uint32_t)h.edges.size() : 0u;
            if (h.presence == PRES_PARTIAL) {
                fe.ff = h.ff;
                if (tests_edges) {
                    fe.edge_begin = (uint32_t)f.edges.size();
                    fe.edge_count = (uint32_t)h.edges.size();
                    const Polygon& s = map.slots[h.slot].rec.shape;
                    for (uint32_t ei : h.edges) f.edges.push_back(s.edge(ei));
                }
            }
            f.ent.push_back(fe);
        }
    }
    const size_t S = map.slots.size();
    f.haz_bbox.assign(S, make_float4(0, 0, 0, 0));
    f.haz_poi.assign(S, make_float4(0, 0, 0, 0));
    f.haz_vr.assign(S, make_uint2(0, 0));
    for (size_t s = 1; s < S; s++) {
        if (!map.slots[s].occ) continue;
        const HazardRec& r = map.slots[s].rec;
        f.haz_bbox[s] = r.shape.bbox;
        f.haz_poi[s] = make_float4(r.shape.poi.x, r.shape.poi.y, r.shape.poi.z, r.interior() ? 1.0f : 0.0f);
        f.haz_vr[s] = make_uint2((uint32_t)f.haz_verts.size(), r.shape.n());
        f.haz_verts.insert(f.haz_verts.end(), r.shape.v.begin(), r.shape.v.end());
    }
}

// ---------------------------------------------------------------------------------------
// pool-ordered device image with in-place updates
// ---------------------------------------------------------------------------------------
static void image_write_node(PoolImage& im, const QuadTree& qt, const SlotMap& map, uint32_t i) {
    const QNode& n = qt.pool[i];
    const uint32_t count = (uint32_t)n.haz.size();
    // a full build packs the ranges tightly (the layout queries run on); a range that has to
    // grow later moves to the arena's end with room for half as much again
    if (count > im.ent_cap[i]) {
        im.garbage += im.ent_cap[i];
        im.ent_off[i] = (uint32_t)im.ent.size();
        im.ent_cap[i] = im.building ? count : count + count / 2 + 1;
        im.ent.resize(im.ent.size() + im.ent_cap[i]);
    }
    const bool tests_edges = n.child < 0 || n.n_active_edges <= qt.cd_threshold;
    uint32_t need = 0;
    if (tests_edges)
        for (const QEntry& h : n.haz)
            if (h.presence == PRES_PARTIAL) need += (uint32_t)h.edges.size();
    if (need > im.edge_cap[i]) {
        im.garbage += im.edge_cap[i];
        im.edge_off[i] = (uint32_t)im.edges.size();
        im.edge_cap[i] = im.building ? need : need + need / 2 + 2;
        im.edges.resize(im.edges.size() + im.edge_cap[i]);
    }
    FlatNode fn{};
    fn.bbox = n.bbox;
    fn.first_child = n.child;
    fn.haz_begin = im.ent_off[i];
    fn.haz_info = count | (count && n.haz[0].presence == PRES_ENTIRE ? 0x80000000u : 0u);
    fn.n_active_edges = (uint32_t)std::min<uint64_t>(n.n_active_edges, 0xffffffffu);
    uint32_t ec = im.edge_off[i];
    for (uint32_t k = 0; k < count; k++) {
        const QEntry& h = n.haz[k];
        FlatEntry fe{};
        fe.slot_flags = h.slot | (h.presence == PRES_ENTIRE ? 0x80000000u : 0u);
        fe.n_edges_total = h.presence == PRES_PARTIAL ? (uint32_t)h.edges.size() : 0u;
        if (h.presence == PRES_PARTIAL) {
            fe.ff = h.ff;
            if (tests_edges) {
                fe.edge_begin = ec;
                fe.edge_count = (uint32_t)h.edges.size();
                const Polygon& sh = map.slots[h.slot].rec.shape;
                for (uint32_t ei : h.edges) im.edges[ec++] = sh.edge(ei);
            }
        }
        im.ent[im.ent_off[i] + k] = fe;
    }
    im.nodes[i] = fn;
    if (n.child >= 0)
        for (uint32_t k = 0; k < 4; k++) im.up[(uint32_t)n.child + k] = (i << 2) | k;
}

static void image_write_slot(PoolImage& im, const SlotMap& map, uint32_t s) {
    const HazardRec& r = map.slots[s].rec;
    im.hbb[s] = r.shape.bbox;
    im.hpoi[s] = make_float4(r.shape.poi.x, r.shape.poi.y, r.shape.poi.z, r.interior() ? 1.0f : 0.0f);
    im.hvr[s] = make_uint2((uint32_t)im.verts.size(), r.shape.n());
    im.verts.insert(im.verts.end(), r.shape.v.begin(), r.shape.v.end());
    im.slot_ver[s] = map.slots[s].version;
}

static uint64_t with_headroom(uint64_t n) { return n + n / 2 + 1024; }

bool collects_at(const QuadTree& qt, int32_t n) {
    const QNode& q = qt.pool[(size_t)n];
    return q.child < 0 || q.n_active_edges <= qt.cd_threshold;
}

namespace {
// Writes cdepth << 28 | depth << 24 | n for node n (at depth `depth`, covering cell (ix, iy) of
// the 2^level grid, side cells per side at that level; cdepth the first collecting depth on its
// path so far) into the tables of levels level .. cell_level.
void fill_cells(PoolImage& im, const QuadTree& qt, int32_t n, uint32_t depth, uint32_t level, uint32_t ix, uint32_t iy,
                uint32_t cdepth) {
    const int32_t ch = qt.pool[(size_t)n].child;
    if (cdepth == CELL_NO_CDEPTH && collects_at(qt, n)) cdepth = depth;
    const uint32_t entry = (cdepth << 28) | (depth << 24) | (uint32_t)n;
    for (uint32_t l = level, s = 1; l <= im.cell_level; l++, s *= 2) {
        // the block of cells of level l under this level-`level` cell
        const size_t off = (((size_t)1 << (2 * l)) - 1) / 3;
        for (uint32_t y = iy * s; y < (iy + 1) * s; y++)
            for (uint32_t x = ix * s; x < (ix + 1) * s; x++) im.cells[off + (((size_t)y << l) | x)] = entry;
        if (ch >= 0 && l == level && level < im.cell_level) break;     // the children write deeper levels
    }
    if (ch < 0 || level == im.cell_level) return;
    fill_cells(im, qt, ch + 0, depth + 1, level + 1, 2 * ix + 1, 2 * iy + 1, cdepth);
    fill_cells(im, qt, ch + 1, depth + 1, level + 1, 2 * ix, 2 * iy + 1, cdepth);
    fill_cells(im, qt, ch + 2, depth + 1, level + 1, 2 * ix, 2 * iy, cdepth);
    fill_cells(im, qt, ch + 3, depth + 1, level + 1, 2 * ix + 1, 2 * iy, cdepth);
}
}  // namespace

void image_cells(PoolImage& im, const QuadTree& qt) {
    const float4 r = qt.pool[0].bbox;
    // a degenerate root box has no usable grid: level 0 only, the root
    const bool flat = !(r.z > r.x && r.w > r.y);
    im.cell_level = flat ? 0u : std::min<uint32_t>(qt.pool[0].level, CELL_LEVELS_MAX);
    im.cells.assign((((size_t)1 << (2 * (im.cell_level + 1))) - 1) / 3, 0u);
    fill_cells(im, qt, 0, 0, 0, 0, 0, CELL_NO_CDEPTH);
}

void image_build(PoolImage& im, QuadTree& qt, const SlotMap& map) {
    qt.compact_bfs();
    const uint32_t N = (uint32_t)qt.pool.size();
    im = PoolImage{};
    im.nodes.assign(N, FlatNode{});
    im.up.assign(N, 0u);
    im.ent_off.assign(N, 0u);
    im.ent_cap.assign(N, 0u);
    im.edge_off.assign(N, 0u);
    im.edge_cap.assign(N, 0u);
    im.building = true;
    for (uint32_t i = 0; i < N; i++) image_write_node(im, qt, map, i);
    im.building = false;
    const uint32_t S = (uint32_t)map.slots.size();
    im.hbb.assign(S, make_float4(0, 0, 0, 0));
    im.hpoi.assign(S, make_float4(0, 0, 0, 0));
    im.hvr.assign(S, make_uint2(0, 0));
    im.slot_ver.assign(S, 0u);
    for (uint32_t s = 1; s < S; s++)
        if (map.slots[s].occ) image_write_slot(im, map, s);
    image_cells(im, qt);
    qt.cells_dirty.clear();
    im.cap_nodes = with_headroom(N);
    im.cap_ent = with_headroom(im.ent.size());
    im.cap_edges = with_headroom(im.edges.size());
    im.cap_slots = with_headroom(S);
    im.cap_verts = with_headroom(im.verts.size());
    uint64_t off = 0;
    auto place = [&](uint64_t bytes) {
        const uint64_t o = off;
        off = (off + bytes + 255) & ~(uint64_t)255;
        return o;
    };
    im.o_nd = place(sizeof(FlatNode) * im.cap_nodes);
    im.o_nu = place(sizeof(uint32_t) * im.cap_nodes);
    im.o_en = place(sizeof(FlatEntry) * im.cap_ent);
    im.o_ed = place(sizeof(float4) * im.cap_edges);
    im.o_hb = place(sizeof(float4) * im.cap_slots);
    im.o_hp = place(sizeof(float4) * im.cap_slots);
    im.o_hv = place(sizeof(uint2) * im.cap_slots);
    im.o_vv = place(sizeof(float2) * im.cap_verts);
    im.o_ce = place(sizeof(uint32_t) * im.cells.size());
    im.total = off;
    im.valid = true;
    qt.touched.clear();
}

bool image_patch(PoolImage& im, QuadTree& qt, const SlotMap& map, std::vector<uint8_t>& staging,
                 std::vector<ImageSeg>& segs) {
    staging.clear();
    segs.clear();
    const uint32_t N = (uint32_t)qt.pool.size();
    if (!im.valid || N > im.cap_nodes || map.slots.size() > im.cap_slots) return false;
    if (N > im.nodes.size()) {
        im.nodes.resize(N, FlatNode{});
        im.up.resize(N, 0u);
        im.ent_off.resize(N, 0u);
        im.ent_cap.resize(N, 0u);
        im.edge_off.resize(N, 0u);
        im.edge_cap.resize(N, 0u);
    }
    std::vector<uint32_t> nodes;
    {
        std::vector<char> seen(N, 0);
        for (int32_t t : qt.touched)
            if (t >= 0 && (uint32_t)t < N && !seen[t]) {
                seen[t] = 1;
                nodes.push_back((uint32_t)t);
            }
    }
    for (uint32_t i : nodes) {
        // a node within the tables' levels that starts or stops collecting changes them
        const FlatNode& old = im.nodes[i];
        const bool was = old.first_child < 0 || old.n_active_edges <= qt.cd_threshold;
        if ((uint32_t)(qt.pool[0].level - qt.pool[i].level) <= CELL_LEVELS_MAX && was != collects_at(qt, (int32_t)i))
            qt.cells_dirty.push_back((int32_t)i);
        image_write_node(im, qt, map, i);
    }
    const uint32_t S = (uint32_t)map.slots.size();
    if (im.slot_ver.size() < S) {
        im.hbb.resize(S, make_float4(0, 0, 0, 0));
        im.hpoi.resize(S, make_float4(0, 0, 0, 0));
        im.hvr.resize(S, make_uint2(0, 0));
        im.slot_ver.resize(S, 0u);
    }
    std::vector<uint32_t> slots;
    const size_t verts0 = im.verts.size();
    for (uint32_t s = 1; s < S; s++)
        if (map.slots[s].occ && im.slot_ver[s] != map.slots[s].version) {
            image_write_slot(im, map, s);
            slots.push_back(s);
        }
    if (im.ent.size() > im.cap_ent || im.edges.size() > im.cap_edges || im.verts.size() > im.cap_verts) return false;
    if (im.garbage > (uint64_t)(im.ent.size() + im.edges.size()) / 2 + 4096) return false;
    auto add = [&](uint64_t dst, const void* src, uint64_t bytes) {
        if (!bytes) return;
        segs.push_back(ImageSeg{dst, staging.size(), bytes});
        const uint8_t* p = (const uint8_t*)src;
        staging.insert(staging.end(), p, p + bytes);
    };
    for (uint32_t i : nodes) {
        add(im.o_nd + sizeof(FlatNode) * (uint64_t)i, &im.nodes[i], sizeof(FlatNode));
        const uint32_t cnt = im.nodes[i].haz_info & 0x3fffffffu;
        if (cnt) add(im.o_en + sizeof(FlatEntry) * (uint64_t)im.ent_off[i], &im.ent[im.ent_off[i]], sizeof(FlatEntry) * (uint64_t)cnt);
        uint32_t ne = 0;
        for (uint32_t k = 0; k < cnt; k++) ne += im.ent[im.ent_off[i] + k].edge_count;
        if (ne) add(im.o_ed + sizeof(float4) * (uint64_t)im.edge_off[i], &im.edges[im.edge_off[i]], sizeof(float4) * (uint64_t)ne);
        if (im.nodes[i].first_child >= 0) {
            const uint32_t c = (uint32_t)im.nodes[i].first_child;
            add(im.o_nu + sizeof(uint32_t) * (uint64_t)c, &im.up[c], sizeof(uint32_t) * 4);
        }
    }
    for (uint32_t s : slots) {
        add(im.o_hb + sizeof(float4) * (uint64_t)s, &im.hbb[s], sizeof(float4));
        add(im.o_hp + sizeof(float4) * (uint64_t)s, &im.hpoi[s], sizeof(float4));
        add(im.o_hv + sizeof(uint2) * (uint64_t)s, &im.hvr[s], sizeof(uint2));
    }
    if (im.verts.size() > verts0)
        add(im.o_vv + sizeof(float2) * (uint64_t)verts0, im.verts.data() + verts0, sizeof(float2) * (im.verts.size() - verts0));
    // the cells under each dirty node: its depth, cell and the first collecting depth above it
    // from the parent links, then its block of each level from its own down, rewritten and
    // uploaded row by row (the whole tables when the root is dirty)
    for (int32_t x : qt.cells_dirty) {
        uint32_t d = 0, ix = 0, iy = 0, top = 0;
        bool above = false, live = true;
        for (uint32_t n = (uint32_t)x; n != 0u && live; d++) {
            const uint32_t up = im.up[n], k = up & 3u, p = up >> 2;
            // a node dropped since it was marked is no longer its parent's child: skip it
            live = d < CELL_LEVELS_MAX && qt.pool[p].child >= 0 && (uint32_t)qt.pool[p].child + k == n;
            ix |= (k == 0 || k == 3 ? 1u : 0u) << d;
            iy |= (k == 0 || k == 1 ? 1u : 0u) << d;
            n = p;
            if (collects_at(qt, (int32_t)n)) {
                above = true;
                top = d + 1u;                     // steps above x of the topmost collecting ancestor so far
            }
        }
        if (!live || d > im.cell_level) continue;
        fill_cells(im, qt, x, d, d, ix, iy, above ? d - top : CELL_NO_CDEPTH);
        if (d == 0) {
            add(im.o_ce, im.cells.data(), sizeof(uint32_t) * im.cells.size());
            break;
        }
        for (uint32_t l = d; l <= im.cell_level; l++) {
            const uint32_t s = l - d;
            const uint64_t off = (((uint64_t)1 << (2 * l)) - 1) / 3;
            for (uint32_t y = iy << s; y < (iy + 1u) << s; y++) {
                const uint64_t at = off + ((uint64_t)y << l) + ((uint64_t)ix << s);
                add(im.o_ce + sizeof(uint32_t) * at, &im.cells[at], sizeof(uint32_t) << s);
            }
        }
    }
    qt.touched.clear();
    qt.cells_dirty.clear();
    return true;
}

}  // namespace jg
