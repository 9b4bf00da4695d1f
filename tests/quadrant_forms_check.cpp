// The shared-interval quadrant tests of jg_geom.hpp (edge_quadrants_node, circle_quadrants,
// quadrant_surrounding) against the per-quadrant forms they replace in device code
// (edge_quadrants over box_quadrant boxes, circle_box per quadrant, box_relation per quadrant),
// on node boxes built the way the quadtree builds them (box_quadrant of the parent, from a
// root box) and entities placed on and around the midlines and bounds, with signed zeros,
// zero extents, zero radii, inverted and NaN query boxes.
//
//   quadrant_forms_check <seed> <cases>   prints "ok <cases> <midline hits>" or the first mismatch
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../jagua-rs_amd/csrc/jg_geom.hpp"

using namespace jg;

namespace {

std::mt19937_64 rng;

float uni(float a, float b) { return std::uniform_real_distribution<float>(a, b)(rng); }
uint32_t pick(uint32_t n) { return (uint32_t)(rng() % n); }

// a coordinate near the box: one of its bounds or midlines (exactly, or one ulp off), or uniform
float coord(float lo, float mid, float hi) {
    const float span = hi - lo;
    switch (pick(8)) {
        case 0: return lo;
        case 1: return hi;
        case 2: return mid;
        case 3: return std::nextafter(mid, pick(2) ? INFINITY : -INFINITY);
        case 4: return std::nextafter(pick(2) ? lo : hi, pick(2) ? INFINITY : -INFINITY);
        case 5: return uni(lo - span, hi + span);
        default: return uni(lo, hi);
    }
}

uint32_t circle_loop(float cx, float cy, float r, float4 b) {
    uint32_t m = 0;
    for (int k = 0; k < 4; k++) m |= circle_box(cx, cy, r, box_quadrant(b, k)) ? (1u << k) : 0u;
    return m;
}

int surround_loop(float4 r, float4 bb) {
    for (int k = 0; k < 4; k++)
        if (box_relation(box_quadrant(r, k), bb) == REL_SURROUNDING) return k;
    return -1;
}

}  // namespace

int main(int argc, char** argv) {
    rng.seed(argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1);
    const long cases = argc > 2 ? std::atol(argv[2]) : 100000;
    long on_mid = 0;
    for (long i = 0; i < cases; i++) {
        // a node box: a root box subdivided 0..12 times
        const float cx0 = uni(-5000.0f, 5000.0f), cy0 = uni(-5000.0f, 5000.0f);
        const float w = std::ldexp(1.0f, (int)pick(24) - 10) * uni(1.0f, 2.0f);
        const float h = pick(10) == 0 ? 0.0f : std::ldexp(1.0f, (int)pick(24) - 10) * uni(1.0f, 2.0f);
        float4 r = make_float4(cx0, cy0, cx0 + (pick(10) == 0 ? 0.0f : w), cy0 + h);
        if (pick(20) == 0) r = make_float4(-0.0f, 0.0f, 0.0f, -0.0f);
        for (uint32_t d = pick(13); d > 0; d--) r = box_quadrant(r, (int)pick(4));
        const float mxm = (r.x + r.z) / 2.0f, mym = (r.y + r.w) / 2.0f;
        // an edge
        float4 e = make_float4(coord(r.x, mxm, r.z), coord(r.y, mym, r.w), coord(r.x, mxm, r.z), coord(r.y, mym, r.w));
        if (pick(50) == 0) e.x = NAN;
        if (pick(10) == 0) { e.z = e.x; e.w = e.y; }
        on_mid += (e.x == mxm || e.z == mxm || e.y == mym || e.w == mym) ? 1 : 0;
        const uint32_t a = edge_quadrants_node(e, r), b = edge_quadrants(e, r);
        if (a != b) {
            std::printf("edge mismatch %ld: r=(%a %a %a %a) e=(%a %a %a %a) %u %u\n", i, r.x, r.y, r.z, r.w, e.x, e.y, e.z, e.w, a, b);
            return 1;
        }
        // a circle
        const float ccx = coord(r.x, mxm, r.z), ccy = coord(r.y, mym, r.w);
        float rad = pick(8) == 0 ? 0.0f : uni(0.0f, 1.5f) * (r.z - r.x + r.w - r.y + 1e-3f);
        if (pick(8) == 0) rad = std::fabs(ccx - mxm);           // touching a midline
        const uint32_t ca = circle_quadrants(ccx, ccy, rad, r), cb = circle_loop(ccx, ccy, rad, r);
        if (ca != cb) {
            std::printf("circle mismatch %ld: r=(%a %a %a %a) c=(%a %a %a) %u %u\n", i, r.x, r.y, r.z, r.w, ccx, ccy, rad, ca, cb);
            return 1;
        }
        // a query box
        float x0 = coord(r.x, mxm, r.z), x1 = coord(r.x, mxm, r.z), y0 = coord(r.y, mym, r.w), y1 = coord(r.y, mym, r.w);
        if (pick(8) != 0) {
            if (x0 > x1) std::swap(x0, x1);
            if (y0 > y1) std::swap(y0, y1);
        }
        float4 bb = make_float4(x0, y0, x1, y1);
        if (pick(40) == 0) bb = make_float4(3.40282347e+38f, 3.40282347e+38f, -3.40282347e+38f, -3.40282347e+38f);
        if (pick(40) == 0) bb.z = NAN;
        const int sa = quadrant_surrounding(r, bb), sb = surround_loop(r, bb);
        if (sa != sb) {
            std::printf("surround mismatch %ld: r=(%a %a %a %a) bb=(%a %a %a %a) %d %d\n", i, r.x, r.y, r.z, r.w, bb.x, bb.y, bb.z,
                        bb.w, sa, sb);
            return 1;
        }
    }
    std::printf("ok %ld %ld\n", cases, on_mid);
    return 0;
}
