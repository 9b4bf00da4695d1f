// Geometric predicates of the collision query, compiled for both the host build (quadtree
// construction) and gfx950 device code.  Every function keeps the reference's f32 operation
// order; the library is compiled with -ffp-contract=off and IEEE division, so results are
// bit-identical on host and device.  Boxes are float4 (x_min, y_min, x_max, y_max); edges
// are float4 (x1, y1, x2, y2).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define JG_HD __host__ __device__ __forceinline__

namespace jg {

// f32::min / f32::max (IEEE minNum / maxNum)
JG_HD float mn(float a, float b) { return fminf(a, b); }
JG_HD float mx(float a, float b) { return fmaxf(a, b); }

JG_HD uint32_t f2u(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __float_as_uint(f);
#else
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return u;
#endif
}

// float-cmp approx_eq!(f32) default margin (util/fpa.rs:18-22): ==, |a-b| <= f32::EPSILON,
// or at most 4 ulps apart (wrapping i32 difference of the bit patterns, saturating abs).
JG_HD bool approx_eq(float a, float b) {
    if (a == b) return true;
    if (fabsf(a - b) <= 1.1920929e-07f) return true;
    int32_t d = (int32_t)(f2u(a) - f2u(b));
    int32_t ad = d < 0 ? (d == INT32_MIN ? INT32_MAX : -d) : d;
    return ad <= 4;
}
// FPA ordering (util/fpa.rs:24-31)
JG_HD bool fpa_lt(float a, float b) { return !approx_eq(a, b) && a < b; }
JG_HD bool fpa_gt(float a, float b) { return !approx_eq(a, b) && a > b; }
JG_HD bool fpa_le(float a, float b) { return approx_eq(a, b) || a < b; }
JG_HD bool fpa_ge(float a, float b) { return approx_eq(a, b) || a > b; }

enum Relation : int { REL_INTERSECTING = 0, REL_ENCLOSED = 1, REL_SURROUNDING = 2, REL_DISJOINT = 3 };

// rect.rs:266-272
JG_HD bool box_overlap(float4 a, float4 b) {
    return mx(a.x, b.x) <= mn(a.z, b.z) && mx(a.y, b.y) <= mn(a.w, b.w);
}
// rect.rs:282-288
JG_HD bool box_has_point(float4 r, float x, float y) {
    return x >= r.x && x <= r.z && y >= r.y && y <= r.w;
}
// rect.rs:46-65  (relation of a to b)
JG_HD int box_relation(float4 a, float4 b) {
    if (!box_overlap(a, b)) return REL_DISJOINT;
    if (a.x <= b.x && a.y <= b.y && a.z >= b.z && a.w >= b.w) return REL_SURROUNDING;
    if (a.x >= b.x && a.y >= b.y && a.z <= b.z && a.w <= b.w) return REL_ENCLOSED;
    return REL_INTERSECTING;
}
// rect.rs:70-89 and 274-280
JG_HD int box_almost_relation(float4 a, float4 b) {
    if (!(fpa_le(mx(a.x, b.x), mn(a.z, b.z)) && fpa_le(mx(a.y, b.y), mn(a.w, b.w)))) return REL_DISJOINT;
    if (fpa_le(a.x, b.x) && fpa_le(a.y, b.y) && fpa_ge(a.z, b.z) && fpa_ge(a.w, b.w)) return REL_SURROUNDING;
    if (fpa_ge(a.x, b.x) && fpa_ge(a.y, b.y) && fpa_le(a.z, b.z) && fpa_le(a.w, b.w)) return REL_ENCLOSED;
    return REL_INTERSECTING;
}
// rect.rs:93-109
JG_HD float4 box_inflate_to_square(float4 r) {
    float w = r.z - r.x, h = r.w - r.y;
    float dx = 0.0f, dy = 0.0f;
    if (h < w) dy = (w - h) / 2.0f;
    else if (w < h) dx = (h - w) / 2.0f;
    return make_float4(r.x - dx, r.y - dy, r.z + dx, r.w + dy);
}
// rect.rs:141-161 and 35-41: quadrant k of r, k in [q1 TR, q2 TL, q3 BL, q4 BR]
JG_HD float4 box_quadrant(float4 r, int k) {
    float mx_ = (r.x + r.z) / 2.0f, my_ = (r.y + r.w) / 2.0f;
    float cx = (k == 0 || k == 3) ? r.z : r.x;
    float cy = (k == 0 || k == 1) ? r.w : r.y;
    return make_float4(mn(cx, mx_), mn(cy, my_), mx(cx, mx_), mx(cy, my_));
}

// rect.rs:301-345  Rect–Edge
JG_HD bool box_edge(float4 r, float4 e) {
    float ex0 = mn(e.x, e.z), ex1 = mx(e.x, e.z), ey0 = mn(e.y, e.w), ey1 = mx(e.y, e.w);
    if (mx(ex0, r.x) > mn(ex1, r.z) || mx(ey0, r.y) > mn(ey1, r.w)) return false;
    if (box_has_point(r, e.x, e.y) || box_has_point(r, e.z, e.w)) return true;
    float dx = e.z - e.x, dy = e.w - e.y;
    // corners in quadrant order: (x_max,y_max) (x_min,y_max) (x_min,y_min) (x_max,y_min)
    float s0 = (r.z - e.x) * dy - (r.w - e.y) * dx;
    float s1 = (r.x - e.x) * dy - (r.w - e.y) * dx;
    float s2 = (r.x - e.x) * dy - (r.y - e.y) * dx;
    float s3 = (r.z - e.x) * dy - (r.y - e.y) * dx;
    bool pos = s0 > 0.0f && s1 > 0.0f && s2 > 0.0f && s3 > 0.0f;
    bool neg = s0 < 0.0f && s1 < 0.0f && s2 < 0.0f && s3 < 0.0f;
    return !(pos || neg);
}

// edge.rs:179-228  Edge–Edge (a = querying edge, b = hazard edge)
JG_HD bool edge_edge(float4 a, float4 b) {
    float x1 = a.x, y1 = a.y, x2 = a.z, y2 = a.w;
    float x3 = b.x, y3 = b.y, x4 = b.z, y4 = b.w;
    if (mx(mn(x1, x2), mn(x3, x4)) > mn(mx(x1, x2), mx(x3, x4))) return false;
    if (mx(mn(y1, y2), mn(y3, y4)) > mn(mx(y1, y2), mx(y3, y4))) return false;
    float t_nom = (x2 - x4) * (y4 - y3) - (y2 - y4) * (x4 - x3);
    float t_den = (x2 - x1) * (y4 - y3) - (y2 - y1) * (x4 - x3);
    float u_nom = (x2 - x4) * (y2 - y1) - (y2 - y4) * (x2 - x1);
    float u_den = (x2 - x1) * (y4 - y3) - (y2 - y1) * (x4 - x3);
    if (t_den == 0.0f || u_den == 0.0f) return false;
    float t = t_nom / t_den;
    float u = u_nom / u_den;
    return t >= 0.0f && t <= 1.0f && u >= 0.0f && u <= 1.0f;
}

// qt_traits.rs:144-177: where along b (the side / bisector) a crosses it.
// 0 = no crossing, 1 = first half (u > 0.5), 2 = exactly halfway, 3 = second half
JG_HD int edge_half(float4 a, float x3, float y3, float x4, float y4) {
    float x1 = a.x, y1 = a.y, x2 = a.z, y2 = a.w;
    float t_nom = (x2 - x4) * (y4 - y3) - (y2 - y4) * (x4 - x3);
    float t_den = (x2 - x1) * (y4 - y3) - (y2 - y1) * (x4 - x3);
    float u_nom = (x2 - x4) * (y2 - y1) - (y2 - y4) * (x2 - x1);
    float u_den = (x2 - x1) * (y4 - y3) - (y2 - y1) * (x4 - x3);
    if (t_den == 0.0f || u_den == 0.0f) return 0;
    float t = t_nom / t_den;
    float u = u_nom / u_den;
    if (!(t >= 0.0f && t <= 1.0f && u >= 0.0f && u <= 1.0f)) return 0;
    return u > 0.5f ? 1 : (u < 0.5f ? 3 : 2);
}

// Quadrant state bits: for each quadrant k, bit k = determined, bit k+4 = colliding.
// qt_traits.rs:110-140 half_intersect: a call runs only while one of its target quadrants is
// undetermined, and then marks the hit half's quadrants colliding (even if determined false).
JG_HD void half_update(uint32_t& st, float4 e, float x3, float y3, float x4, float y4, uint32_t fst, uint32_t sec) {
    uint32_t tgt = fst | sec;
    if ((st & tgt) == tgt) return;
    int loc = edge_half(e, x3, y3, x4, y4);
    uint32_t set = loc == 1 ? fst : (loc == 2 ? tgt : (loc == 3 ? sec : 0u));
    st |= set | (set << 4);
}

// qt_traits.rs:19-106  Edge::collides_with_quadrants; r = node box, q[k] = its children.
// Returns a 4-bit mask of colliding quadrants.
JG_HD uint32_t edge_quadrants_sides(float4 e, float4 r, uint32_t st);

// The same with the quadrant boxes derived from r on the fly (box_quadrant), keeping fewer
// values live in device code.
JG_HD uint32_t edge_quadrants(float4 e, float4 r) {
    float ex0 = mn(e.x, e.z), ex1 = mx(e.x, e.z), ey0 = mn(e.y, e.w), ey1 = mx(e.y, e.w);
    uint32_t st = 0;
    for (int k = 0; k < 4; k++) {
        const float4 b = box_quadrant(r, k);
        bool no = mx(ex0, b.x) > mn(ex1, b.z) || mx(ey0, b.y) > mn(ey1, b.w);
        if (no) st |= 1u << k;
        else if (box_has_point(b, e.x, e.y) || box_has_point(b, e.z, e.w)) st |= (1u << k) | (1u << (k + 4));
    }
    if ((st & 0xfu) == 0xfu) return st >> 4;
    return edge_quadrants_sides(e, r, st);
}

// Quadrant intervals of a box r with r.x <= r.z and r.y <= r.w (every quadtree node's box,
// each built by box_quadrant from its parent's).  The rounded midpoint m.x = (r.x + r.z) / 2
// lies in [r.x, r.z], so the min / max of box_quadrant(r, k) are plain bounds: quadrant k spans
// x in [m.x, r.z] for k = 0, 3 and [r.x, m.x] for k = 1, 2, and y in [m.y, r.w] for k = 0, 1
// and [r.y, m.y] for k = 2, 3.  The quadrant tests below evaluate the same comparisons and
// arithmetic on these bounds as the per-quadrant forms, sharing the two x and two y
// intervals instead of building four boxes: their results are identical for every entity.

// edge_quadrants for a node box (qt_traits.rs:19-106).
JG_HD uint32_t edge_quadrants_node(float4 e, float4 r) {
    const float ex0 = mn(e.x, e.z), ex1 = mx(e.x, e.z), ey0 = mn(e.y, e.w), ey1 = mx(e.y, e.w);
    const float mxm = (r.x + r.z) / 2.0f, mym = (r.y + r.w) / 2.0f;
    // box overlap per interval (rect.rs:301-345 first test) and endpoint containment (rect.rs:282-288)
    const bool oxr = !(mx(ex0, mxm) > mn(ex1, r.z)), oxl = !(mx(ex0, r.x) > mn(ex1, mxm));
    const bool oyt = !(mx(ey0, mym) > mn(ey1, r.w)), oyb = !(mx(ey0, r.y) > mn(ey1, mym));
    const bool axr = e.x >= mxm && e.x <= r.z, axl = e.x >= r.x && e.x <= mxm;
    const bool ayt = e.y >= mym && e.y <= r.w, ayb = e.y >= r.y && e.y <= mym;
    const bool bxr = e.z >= mxm && e.z <= r.z, bxl = e.z >= r.x && e.z <= mxm;
    const bool byt = e.w >= mym && e.w <= r.w, byb = e.w >= r.y && e.w <= mym;
    const bool ov[4] = {oxr && oyt, oxl && oyt, oxl && oyb, oxr && oyb};
    const bool pt[4] = {(axr && ayt) || (bxr && byt), (axl && ayt) || (bxl && byt), (axl && ayb) || (bxl && byb),
                        (axr && ayb) || (bxr && byb)};
    uint32_t st = 0;
    for (int k = 0; k < 4; k++) {
        if (!ov[k]) st |= 1u << k;
        else if (pt[k]) st |= (1u << k) | (1u << (k + 4));
    }
    if ((st & 0xfu) == 0xfu) return st >> 4;
    return edge_quadrants_sides(e, r, st);
}

// circle_box against the four quadrants of a node box (circle.rs:115-127): bit k = quadrant k.
JG_HD uint32_t circle_quadrants(float cx, float cy, float rad, float4 r) {
    const float mxm = (r.x + r.z) / 2.0f, mym = (r.y + r.w) / 2.0f;
    const float dxr = mx(mxm, mn(cx, r.z)) - cx, dxl = mx(r.x, mn(cx, mxm)) - cx;
    const float dyt = mx(mym, mn(cy, r.w)) - cy, dyb = mx(r.y, mn(cy, mym)) - cy;
    const float xr = dxr * dxr, xl = dxl * dxl, yt = dyt * dyt, yb = dyb * dyb;
    const float rr = rad * rad;
    return (xr + yt <= rr ? 1u : 0u) | (xl + yt <= rr ? 2u : 0u) | (xl + yb <= rr ? 4u : 0u) | (xr + yb <= rr ? 8u : 0u);
}

// get_virtual_root's step (cd_engine.rs:314-327): the first quadrant k of a node box r whose
// box_relation to bb is Surrounding (rect.rs:46-65), or -1.  For a box bb with bb.x <= bb.z
// and bb.y <= bb.w, surrounding implies overlap, so the test is four bound comparisons per
// quadrant; any other bb takes the per-quadrant relation.
JG_HD int quadrant_surrounding(float4 r, float4 bb) {
    if (!(bb.x <= bb.z && bb.y <= bb.w)) {
        for (int k = 0; k < 4; k++)
            if (box_relation(box_quadrant(r, k), bb) == REL_SURROUNDING) return k;
        return -1;
    }
    const float mxm = (r.x + r.z) / 2.0f, mym = (r.y + r.w) / 2.0f;
    const bool xr = mxm <= bb.x && r.z >= bb.z, xl = r.x <= bb.x && mxm >= bb.z;
    const bool yt = mym <= bb.y && r.w >= bb.w, yb = r.y <= bb.y && mym >= bb.w;
    return (xr && yt) ? 0 : (xl && yt) ? 1 : (xl && yb) ? 2 : (xr && yb) ? 3 : -1;
}

JG_HD uint32_t edge_quadrants(float4 e, float4 r, const float4* q) {
    float ex0 = mn(e.x, e.z), ex1 = mx(e.x, e.z), ey0 = mn(e.y, e.w), ey1 = mx(e.y, e.w);
    uint32_t st = 0;
    for (int k = 0; k < 4; k++) {
        float4 b = q[k];
        bool no = mx(ex0, b.x) > mn(ex1, b.z) || mx(ey0, b.y) > mn(ey1, b.w);
        if (no) st |= 1u << k;
        else if (box_has_point(b, e.x, e.y) || box_has_point(b, e.z, e.w)) st |= (1u << k) | (1u << (k + 4));
    }
    if ((st & 0xfu) == 0xfu) return st >> 4;
    return edge_quadrants_sides(e, r, st);
}

// The side and bisector half-intersections of collides_with_quadrants for the quadrants not
// yet determined in st.
JG_HD uint32_t edge_quadrants_sides(float4 e, float4 r, uint32_t st) {
    float cx = (r.x + r.z) / 2.0f, cy = (r.y + r.w) / 2.0f;
    // sides (rect.rs:164-184): top (x_max,y_max)->(x_min,y_max), left (x_min,y_max)->(x_min,y_min),
    // bottom (x_min,y_min)->(x_max,y_min), right (x_max,y_min)->(x_max,y_max)
    half_update(st, e, r.x, r.w, r.x, r.y, 1u << 1, 1u << 2);                       // left
    half_update(st, e, r.z, r.y, r.z, r.w, 1u << 3, 1u << 0);                       // right
    half_update(st, e, r.z, r.w, r.x, r.w, 1u << 0, 1u << 1);                       // top
    half_update(st, e, r.x, r.y, r.z, r.y, 1u << 2, 1u << 3);                       // bottom
    half_update(st, e, r.x, cy, r.z, cy, (1u << 1) | (1u << 2), (1u << 0) | (1u << 3));  // h bisector
    half_update(st, e, cx, r.y, cx, r.w, (1u << 2) | (1u << 3), (1u << 0) | (1u << 1));  // v bisector
    return st >> 4;
}

// circle.rs:115-127  Circle–Rect
JG_HD bool circle_box(float cx, float cy, float rad, float4 r) {
    float nx = mx(r.x, mn(cx, r.z));
    float ny = mx(r.y, mn(cy, r.w));
    float dx = nx - cx, dy = ny - cy;
    return dx * dx + dy * dy <= rad * rad;
}
// circle.rs:109-113 via edge.rs:63-87,153-159  Circle–Edge
JG_HD bool circle_edge(float cx, float cy, float rad, float4 e) {
    float a = cx - e.x, b = cy - e.y, c = e.z - e.x, d = e.w - e.y;
    float dot = a * c + b * d;
    float len_sq = c * c + d * d;
    float param = -1.0f;
    if (len_sq != 0.0f) param = dot / len_sq;
    float xx, yy;
    if (param < 0.0f) { xx = e.x; yy = e.y; }
    else if (param > 1.0f) { xx = e.z; yy = e.w; }
    else { xx = e.x + param * c; yy = e.y + param * d; }
    float dx = cx - xx, dy = cy - yy;
    return dx * dx + dy * dy <= rad * rad;
}

// point.rs:26-31 with a 2x3 matrix m = (m00 m01 m02 m10 m11 m12)
struct Mat { float a, b, c, d, e, f; };
JG_HD float2 xform(const Mat& m, float x, float y) {
    return make_float2(m.a * x + m.b * y + m.c * 1.0f, m.d * x + m.e * y + m.f * 1.0f);
}

// One edge of the ray-casting point-in-polygon test (simple_polygon.rs:270-287):
// returns 1 when edge (s -> e) adds an intersection for point p with ray end x = ray_x.
JG_HD int pip_edge(float px, float py, float ray_x, float sx, float sy, float ex, float ey) {
    if ((approx_eq(sy, py) && fpa_gt(sx, px)) || (approx_eq(ey, py) && fpa_gt(ex, px)))
        return (fpa_lt(sy, py) || fpa_lt(ey, py)) ? 1 : 0;
    return edge_edge(make_float4(px, py, ray_x, py), make_float4(sx, sy, ex, ey)) ? 1 : 0;
}

// SPolygon::collides_with(&Point) ray cast (simple_polygon.rs:255-292) over vertices v[0..n)
// with bounding box bb.
JG_HD bool poly_has_point(const float2* v, uint32_t n, float4 bb, float x, float y) {
    if (!box_has_point(bb, x, y)) return false;
    const float ray_x = bb.z + (bb.z - bb.x);
    int cnt = 0;
    for (uint32_t i = 0; i < n; i++) {
        const float2 s = v[i], e = v[i + 1 == n ? 0 : i + 1];
        cnt += pip_edge(x, y, ray_x, s.x, s.y, e.x, e.y);
    }
    return (cnt & 1) == 1;
}

// Edge::distance_to(Point) (edge.rs:146-159, closest_point_on_edge :63-87)
JG_HD float edge_point_dist(float4 e, float px, float py) {
    const float a = px - e.x, b = py - e.y, c = e.z - e.x, d = e.w - e.y;
    const float dot = a * c + b * d;
    const float len_sq = c * c + d * d;
    float param = -1.0f;
    if (len_sq != 0.0f) param = dot / len_sq;
    float xx, yy;
    if (param < 0.0f) { xx = e.x; yy = e.y; }
    else if (param > 1.0f) { xx = e.z; yy = e.w; }
    else { xx = e.x + param * c; yy = e.y + param * d; }
    const float dx = px - xx, dy = py - yy;
    return sqrtf(dx * dx + dy * dy);
}

// MAX_POI_TREE_DEPTH (fail_fast/pole.rs:82): levels below the root cell of the pole search
constexpr int POLE_LEVELS = 10;

// POINode::new (fail_fast/pole.rs:93-127): signed distance of the cell centre to the border
// of `shape` minus the existing poles (negative when the centre is outside); radius is half
// the cell's diameter.  Poles are (cx, cy, r, -).
JG_HD float pole_cell(const float2* v, uint32_t n, float4 bb, const float4* poles, uint32_t n_poles, float4 box,
                      float& radius) {
    const float dx = box.z - box.x, dy = box.w - box.y;
    radius = sqrtf(dx * dx + dy * dy) / 2.0f;
    const float cx = (box.x + box.z) / 2.0f, cy = (box.y + box.w) / 2.0f;
    bool inside = poly_has_point(v, n, bb, cx, cy);
    if (inside)
        for (uint32_t i = 0; i < n_poles; i++) {
            const float4 c = poles[i];
            const float ex = cx - c.x, ey = cy - c.y;
            if (ex * ex + ey * ey <= c.z * c.z) { inside = false; break; }     // circle.rs:95-107
        }
    float d = 3.40282347e+38f;
    for (uint32_t i = 0; i < n; i++) {
        const float2 a = v[i], b = v[i + 1 == n ? 0 : i + 1];
        d = mn(d, edge_point_dist(make_float4(a.x, a.y, b.x, b.y), cx, cy));
    }
    for (uint32_t i = 0; i < n_poles; i++) {
        const float4 c = poles[i];
        const float ex = cx - c.x, ey = cy - c.y;
        const float dc = sqrtf(ex * ex + ey * ey);
        d = mn(d, dc <= c.z ? c.z - dc : dc - c.z);                               // circle.rs:154-163
    }
    return inside ? d : -d;
}

// ---- generate_piers (fail_fast/piers.rs), release-build constants (:16-20)
constexpr uint32_t PIER_RAYS = 200, PIER_ANGLES = 90, PIER_GRID = 100;
constexpr uint32_t PIER_N_RAYS = PIER_RAYS * PIER_ANGLES, PIER_N_GRID = PIER_GRID * PIER_GRID;
constexpr float PIER_TRIM = 0.999f, PIER_INFL = 0.10f;

// ndarray 0.17 Array::linspace element i: a + step * i with step = (b - a) / (n - 1)
JG_HD float linspace_at(float a, float b, uint32_t n, uint32_t i) {
    const float step = n > 1 ? (b - a) / (float)(n - 1) : 0.0f;
    return a + step * (float)i;
}

// Ray i of the angle with (sin, cos) = (s, c) (generate_ray_transformations, piers.rs:100-123):
// the base ray r0-r1 through the centroid, translated by (tx, 0) then rotated,
// rot_m(angle) . transl_m(tx, 0) multiplied out term by term (transformation.rs:156-178).
JG_HD float4 pier_ray(float s, float c, float tx, float2 r0, float2 r1) {
    const Mat m{c * 1.0f + -s * 0.0f + 0.0f * 0.0f, c * 0.0f + -s * 1.0f + 0.0f * 0.0f,
                c * tx + -s * 0.0f + 0.0f * 1.0f,  s * 1.0f + c * 0.0f + 0.0f * 0.0f,
                s * 0.0f + c * 1.0f + 0.0f * 0.0f, s * tx + c * 0.0f + 0.0f * 1.0f};
    const float2 a = xform(m, r0.x, r0.y), b = xform(m, r1.x, r1.y);
    return make_float4(a.x, a.y, b.x, b.y);
}

// Edge::collides_at (edge.rs:179-228) of polygon edge e with ray r, once edge_edge(e, r)
// holds: the intersection point and its distance to the ray's start (clip, piers.rs:128-133).
JG_HD float2 pier_hit(float4 e, float4 r, float& dist) {
    const float x1 = e.x, y1 = e.y, x2 = e.z, y2 = e.w, x3 = r.x, y3 = r.y, x4 = r.z, y4 = r.w;
    const float t = ((x2 - x4) * (y4 - y3) - (y2 - y4) * (x4 - x3)) / ((x2 - x1) * (y4 - y3) - (y2 - y1) * (x4 - x3));
    const float2 loc = make_float2(x2 + t * (x1 - x2), y2 + t * (y1 - y2));
    const float dx = r.x - loc.x, dy = r.y - loc.y;
    dist = sqrtf(dx * dx + dy * dy);
    return loc;
}

// Edge::scale(CLIPPING_TRIM) of the clip from st to en (edge.rs:39-46)
JG_HD float4 pier_trim(float2 st, float2 en) {
    const float ddx = en.x - st.x, ddy = en.y - st.y;
    return make_float4(st.x - ddx * (PIER_TRIM - 1.0f) / 2.0f, st.y - ddy * (PIER_TRIM - 1.0f) / 2.0f,
                       en.x + ddx * (PIER_TRIM - 1.0f) / 2.0f, en.y + ddy * (PIER_TRIM - 1.0f) / 2.0f);
}

// Per-polygon constants of generate_piers: the inflated square bbox, the base ray through the
// centroid (piers.rs:34-41), the ray spacing, the ray influence radius (:63) and the forfeit
// distance (:64).
struct PierSetup {
    float4 sq;
    float2 r0, r1;
    float dx, infl, forfeit, pad;
};

// Circle::distance_to(Point) (circle.rs:136-147)
JG_HD float circle_point_dist(float4 c, float px, float py) {
    const float ddx = px - c.x, ddy = py - c.y;
    const float sqd = ddx * ddx + ddy * ddy;
    return sqd < c.z * c.z ? 0.0f : sqrtf(sqd) - c.z;
}

// Grid point (i, j) of generate_unrepresented_point_grid (piers.rs:153-176): kept when it lies
// in the shape and outside every pole; d_pole = min_distances_to_poles (:227-237).
JG_HD bool pier_grid_point(const float2* v, uint32_t n, float4 bb, const float4* poles, uint32_t n_poles,
                           const PierSetup& ps, uint32_t i, uint32_t j, float2& p, float& d_pole) {
    p = make_float2(linspace_at(ps.sq.x, ps.sq.z, PIER_GRID, i), linspace_at(ps.sq.y, ps.sq.w, PIER_GRID, j));
    if (!poly_has_point(v, n, bb, p.x, p.y)) return false;
    for (uint32_t k = 0; k < n_poles; k++) {
        const float dx = p.x - poles[k].x, dy = p.y - poles[k].y;
        if (dx * dx + dy * dy <= poles[k].z * poles[k].z) return false;
    }
    float d = ps.forfeit;
    for (uint32_t k = 0; k < n_poles; k++) d = mn(d, circle_point_dist(poles[k], p.x, p.y));
    d_pole = d;
    return true;
}

// The score of grid point g for a candidate ray at distance dn (loss_function, piers.rs:185-213)
JG_HD float pier_point_loss(float d_ray, float d_pole, float dn, float infl) {
    const float d = mn(d_ray, dn);
    const float v = d < infl ? mn(d_pole, d) : d_pole;
    return v * v;
}

// Hashed per-query hazard filter and set hash (specified in include/jagua_amd.h).
JG_HD uint32_t fmix32(uint32_t h) {
    h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
    return h;
}
JG_HD uint32_t filter_hash(uint32_t seed, uint64_t q, uint32_t slot) {
    return fmix32(seed ^ fmix32((uint32_t)q ^ fmix32((uint32_t)(q >> 32) + 0x9e3779b9u)) ^ (slot * 0x85ebca77u));
}
JG_HD uint64_t mix64(uint64_t z) {
    z ^= z >> 30; z *= 0xbf58476d1ce4e5b9ull; z ^= z >> 27; z *= 0x94d049bb133111ebull; z ^= z >> 31;
    return z;
}

}  // namespace jg
