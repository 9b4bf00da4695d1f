// Synthetic nesting workloads: seeded polygons, containers (with holes), non-overlapping
// placed-item layouts and query transform streams.  Host-side data generation for the
// benchmark and tests; it carries no collision-detection semantics of its own (placement
// feasibility is checked here in f64 with plain segment/PIP geometry).
//
// PRNG: SplitMix64 seeding xoshiro256**; f32 uniforms are lo + (hi-lo)*u with
// u = (x >> 40) * 2^-24, f64 uniforms use the top 53 bits.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (auto& x : s) {
            z += 0x9e3779b97f4a7c15ull;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xbf58476d1ce4e5b9ull;
            t = (t ^ (t >> 27)) * 0x94d049bb133111ebull;
            x = t ^ (t >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    float uf(float lo, float hi) { return lo + (hi - lo) * ((float)(next() >> 40) * (1.0f / 16777216.0f)); }
    double ud(double lo, double hi) { return lo + (hi - lo) * ((double)(next() >> 11) * (1.0 / 9007199254740992.0)); }
    uint32_t ui(uint32_t n) { return (uint32_t)((next() >> 32) % n); }
    uint32_t log_uniform(uint32_t a, uint32_t b) {
        double v = std::exp(ud(std::log((double)a), std::log((double)b)));
        long r = std::lround(v);
        return (uint32_t)std::max<long>(a, std::min<long>(b, r));
    }
};

struct P2 { double x, y; };

// Shoelace area in f64 (sign tells orientation).
double area2(const std::vector<P2>& p) {
    double s = 0;
    for (size_t i = 0; i < p.size(); i++) {
        size_t j = (i + 1) % p.size();
        s += p[i].x * p[j].y - p[j].x * p[i].y;
    }
    return 0.5 * s;
}

// Star polygon around the origin: sorted uniform angles, radius from `rad(theta)`; vertices
// rounded to f32.  Re-drawn until the f32 vertices are distinct and the area is non-zero.
template <class F>
std::vector<float> star_polygon(Rng& rng, uint32_t n, F rad) {
    for (;;) {
        std::vector<double> th(n);
        for (auto& t : th) t = rng.ud(0.0, 2.0 * M_PI);
        std::sort(th.begin(), th.end());
        std::vector<float> v(2 * n);
        std::vector<P2> pd(n);
        for (uint32_t i = 0; i < n; i++) {
            double r = rad(rng, th[i]);
            v[2 * i] = (float)(r * std::cos(th[i]));
            v[2 * i + 1] = (float)(r * std::sin(th[i]));
            pd[i] = {v[2 * i], v[2 * i + 1]};
        }
        bool ok = true;
        for (uint32_t i = 0; i < n && ok; i++)
            for (uint32_t j = i + 1; j < n && ok; j++)
                if (v[2 * i] == v[2 * j] && v[2 * i + 1] == v[2 * j + 1]) ok = false;
        // minimum angular gap keeps the star simple after f32 rounding
        for (uint32_t i = 0; i + 1 < n && ok; i++)
            if (th[i + 1] - th[i] < 1e-6) ok = false;
        if (ok && std::fabs(area2(pd)) > 1e-6) return v;
    }
}

// Periodic value noise with `octaves` octaves in [-1, 1] (lattice 4*2^o points per octave).
struct RingNoise {
    std::vector<std::vector<double>> lat;
    RingNoise(Rng& rng, int octaves) {
        for (int o = 0; o < octaves; o++) {
            std::vector<double> l(4u << o);
            for (auto& x : l) x = rng.ud(-1.0, 1.0);
            lat.push_back(l);
        }
    }
    double operator()(double th) const {
        double sum = 0, norm = 0, amp = 1;
        for (const auto& l : lat) {
            double u = th / (2 * M_PI) * l.size();
            size_t i0 = (size_t)std::floor(u) % l.size();
            size_t i1 = (i0 + 1) % l.size();
            double f = u - std::floor(u);
            f = f * f * (3 - 2 * f);
            sum += amp * (l[i0] * (1 - f) + l[i1] * f);
            norm += amp;
            amp *= 0.5;
        }
        return sum / norm;
    }
};

std::vector<float> make_item(Rng& rng, uint32_t nv, float r_lo, float r_hi, uint32_t shape_kind, float* bound_r) {
    // shape_kind: 0 = star with radii U[r_lo, r_hi]; 1 = concave value-noise ring of radius r_hi;
    // 2 = star (radii U[0.6R, R], R ~ U[r_lo, r_hi]) or concave, 50/50
    uint32_t kind = shape_kind;
    if (kind == 2) kind = rng.ui(2) == 0 ? 3 : 1;
    std::vector<float> v;
    if (kind == 0) {
        v = star_polygon(rng, nv, [&](Rng& g, double) { return (double)g.uf(r_lo, r_hi); });
    } else if (kind == 3) {
        float R = rng.uf(r_lo, r_hi);
        v = star_polygon(rng, nv, [&](Rng& g, double) { return (double)g.uf(0.6f * R, R); });
    } else {
        float R = shape_kind == 1 ? rng.uf(r_lo, r_hi) : rng.uf(r_lo, r_hi);
        RingNoise nz(rng, 3);
        v = star_polygon(rng, nv, [&](Rng&, double th) { return R * (1.0 + 0.3 * nz(th)); });
    }
    float b = 0;
    for (size_t i = 0; i < v.size() / 2; i++) b = std::max(b, std::hypot(v[2 * i], v[2 * i + 1]));
    *bound_r = b;
    return v;
}

// --------------------------------------------------------------------------------------
// exact-enough f64 geometry for placement feasibility
// --------------------------------------------------------------------------------------
double seg_seg_dist2(P2 a, P2 b, P2 c, P2 d) {
    auto cross = [](P2 o, P2 p, P2 q) { return (p.x - o.x) * (q.y - o.y) - (p.y - o.y) * (q.x - o.x); };
    double d1 = cross(c, d, a), d2 = cross(c, d, b), d3 = cross(a, b, c), d4 = cross(a, b, d);
    if (((d1 > 0 && d2 < 0) || (d1 < 0 && d2 > 0)) && ((d3 > 0 && d4 < 0) || (d3 < 0 && d4 > 0))) return 0.0;
    auto pd = [](P2 p, P2 s, P2 e) {
        double dx = e.x - s.x, dy = e.y - s.y;
        double l = dx * dx + dy * dy;
        double t = l > 0 ? ((p.x - s.x) * dx + (p.y - s.y) * dy) / l : 0;
        t = std::max(0.0, std::min(1.0, t));
        double qx = s.x + t * dx - p.x, qy = s.y + t * dy - p.y;
        return qx * qx + qy * qy;
    };
    return std::min(std::min(pd(a, c, d), pd(b, c, d)), std::min(pd(c, a, b), pd(d, a, b)));
}

bool pip(const std::vector<P2>& poly, P2 p) {
    bool in = false;
    for (size_t i = 0, j = poly.size() - 1; i < poly.size(); j = i++) {
        if (((poly[i].y > p.y) != (poly[j].y > p.y)) &&
            (p.x < (poly[j].x - poly[i].x) * (p.y - poly[i].y) / (poly[j].y - poly[i].y) + poly[i].x))
            in = !in;
    }
    return in;
}

struct Placed {
    std::vector<P2> w;   // world vertices
    double bx0, by0, bx1, by1;
    double cx, cy, r;    // bounding circle
};

Placed make_placed(const std::vector<float>& local, float rot, float tx, float ty) {
    Placed p;
    // mirror the f32 transform the CDE will apply, so the checked shape is the registered one
    float s = sinf(rot), c = cosf(rot);
    size_t n = local.size() / 2;
    p.w.resize(n);
    p.bx0 = p.by0 = 1e300;
    p.bx1 = p.by1 = -1e300;
    for (size_t i = 0; i < n; i++) {
        float x = local[2 * i], y = local[2 * i + 1];
        float wx = c * x + (-s) * y + tx;
        float wy = s * x + c * y + ty;
        p.w[i] = {wx, wy};
        p.bx0 = std::min(p.bx0, (double)wx); p.by0 = std::min(p.by0, (double)wy);
        p.bx1 = std::max(p.bx1, (double)wx); p.by1 = std::max(p.by1, (double)wy);
    }
    p.cx = tx; p.cy = ty;
    p.r = 0;
    for (auto& q : p.w) p.r = std::max(p.r, std::hypot(q.x - tx, q.y - ty));
    return p;
}

Placed make_static(const std::vector<float>& world) {
    Placed p;
    size_t n = world.size() / 2;
    p.w.resize(n);
    p.bx0 = p.by0 = 1e300;
    p.bx1 = p.by1 = -1e300;
    for (size_t i = 0; i < n; i++) {
        p.w[i] = {world[2 * i], world[2 * i + 1]};
        p.bx0 = std::min(p.bx0, p.w[i].x); p.by0 = std::min(p.by0, p.w[i].y);
        p.bx1 = std::max(p.bx1, p.w[i].x); p.by1 = std::max(p.by1, p.w[i].y);
    }
    p.cx = 0.5 * (p.bx0 + p.bx1);
    p.cy = 0.5 * (p.by0 + p.by1);
    p.r = 0;
    for (auto& q : p.w) p.r = std::max(p.r, std::hypot(q.x - p.cx, q.y - p.cy));
    return p;
}

// true when A and B keep at least `sep` apart and neither contains the other
bool separated(const Placed& a, const Placed& b, double sep) {
    double dc = std::hypot(a.cx - b.cx, a.cy - b.cy);
    if (dc >= a.r + b.r + sep) return true;
    if (a.bx0 > b.bx1 + sep || b.bx0 > a.bx1 + sep || a.by0 > b.by1 + sep || b.by0 > a.by1 + sep) return true;
    double s2 = sep * sep;
    size_t na = a.w.size(), nb = b.w.size();
    for (size_t i = 0; i < na; i++) {
        P2 p = a.w[i], q = a.w[(i + 1) % na];
        double ex0 = std::min(p.x, q.x) - sep, ex1 = std::max(p.x, q.x) + sep;
        double ey0 = std::min(p.y, q.y) - sep, ey1 = std::max(p.y, q.y) + sep;
        if (ex1 < b.bx0 || ex0 > b.bx1 || ey1 < b.by0 || ey0 > b.by1) continue;
        for (size_t j = 0; j < nb; j++) {
            P2 r = b.w[j], t = b.w[(j + 1) % nb];
            if (std::max(r.x, t.x) < ex0 || std::min(r.x, t.x) > ex1 || std::max(r.y, t.y) < ey0 ||
                std::min(r.y, t.y) > ey1)
                continue;
            if (seg_seg_dist2(p, q, r, t) < s2 || (sep == 0 && seg_seg_dist2(p, q, r, t) == 0)) return false;
        }
    }
    if (pip(b.w, a.w[0]) || pip(a.w, b.w[0])) return false;
    return true;
}

// inside the container polygon, at least `sep` from its boundary
bool inside_container(const Placed& a, const Placed& cont, double sep) {
    if (a.bx0 < cont.bx0 + sep || a.by0 < cont.by0 + sep || a.bx1 > cont.bx1 - sep || a.by1 > cont.by1 - sep)
        return false;
    for (auto& q : a.w)
        if (!pip(cont.w, q)) return false;
    // no boundary edge closer than sep
    Placed tmp = a;
    tmp.r = a.r;
    double s2 = std::max(sep, 1e-9);
    s2 *= s2;
    size_t na = a.w.size(), nc = cont.w.size();
    for (size_t j = 0; j < nc; j++) {
        P2 r = cont.w[j], t = cont.w[(j + 1) % nc];
        if (std::max(r.x, t.x) < a.bx0 - sep || std::min(r.x, t.x) > a.bx1 + sep || std::max(r.y, t.y) < a.by0 - sep ||
            std::min(r.y, t.y) > a.by1 + sep)
            continue;
        for (size_t i = 0; i < na; i++)
            if (seg_seg_dist2(a.w[i], a.w[(i + 1) % na], r, t) < s2) return false;
    }
    return true;
}

struct Grid {
    double x0, y0, cell;
    int nx, ny;
    std::vector<std::vector<int>> cells;
    Grid(double bx0, double by0, double bx1, double by1, double c) : x0(bx0), y0(by0), cell(c) {
        nx = std::max(1, (int)std::ceil((bx1 - bx0) / c));
        ny = std::max(1, (int)std::ceil((by1 - by0) / c));
        cells.resize((size_t)nx * ny);
    }
    void range(double a0, double b0, double a1, double b1, int& i0, int& j0, int& i1, int& j1) const {
        i0 = std::max(0, std::min(nx - 1, (int)std::floor((a0 - x0) / cell)));
        i1 = std::max(0, std::min(nx - 1, (int)std::floor((a1 - x0) / cell)));
        j0 = std::max(0, std::min(ny - 1, (int)std::floor((b0 - y0) / cell)));
        j1 = std::max(0, std::min(ny - 1, (int)std::floor((b1 - y0) / cell)));
    }
    void insert(const Placed& p, int id) {
        int i0, j0, i1, j1;
        range(p.bx0, p.by0, p.bx1, p.by1, i0, j0, i1, j1);
        for (int j = j0; j <= j1; j++)
            for (int i = i0; i <= i1; i++) cells[(size_t)j * nx + i].push_back(id);
    }
};

}  // namespace

extern "C" {

struct js_layout_params {
    uint64_t seed;
    uint32_t container_kind;      // 0 rectangle cw x ch at the origin, 1 star polygon
    float cw, ch;                 // rectangle size; star: radius range [cw, ch] centred at (ch, ch)
    uint32_t container_nv_lo, container_nv_hi;
    uint32_t n_holes;
    uint32_t hole_kind;           // 0 rectangle with sides U[hole_lo, hole_hi], 1 star polygon
    float hole_lo, hole_hi;
    uint32_t hole_nv_lo, hole_nv_hi;
    uint32_t n_items;
    uint32_t item_nv_lo, item_nv_hi;
    float item_r_lo, item_r_hi;
    uint32_t item_shape;          // 0 star, 1 concave noise ring, 2 mixed
    float min_sep;
    uint32_t rot_kind;            // 0 continuous, 1 multiples of 45 degrees
    uint32_t placement;           // 0 uniform rejection, 1 left-bottom-fill (lowest 10*x_max + y_max of K feasible samples)
    uint32_t max_attempts;        // per item
    uint32_t lbf_samples;         // placement 1: samples per item
};

struct js_layout {
    std::vector<float> container;                 // world vertices
    std::vector<std::vector<float>> holes;        // world vertices
    std::vector<std::vector<float>> items;        // local vertices
    std::vector<float> dts;                       // per item rotation, tx, ty
};

static float rot_sample(Rng& rng, uint32_t kind) {
    if (kind == 1) return (float)((double)rng.ui(8) * (M_PI / 4.0));
    return rng.uf(0.0f, 6.28318530717958647692f);
}

void* js_layout_generate(const js_layout_params* pp) {
    const js_layout_params& P = *pp;
    Rng rng(P.seed);
    auto* L = new js_layout;
    // container
    if (P.container_kind == 0) {
        L->container = {0, 0, P.cw, 0, P.cw, P.ch, 0, P.ch};
    } else {
        uint32_t nv = rng.log_uniform(P.container_nv_lo, P.container_nv_hi);
        auto v = star_polygon(rng, nv, [&](Rng& g, double) { return (double)g.uf(P.cw, P.ch); });
        for (size_t i = 0; i < v.size() / 2; i++) { v[2 * i] += P.ch; v[2 * i + 1] += P.ch; }
        L->container = v;
    }
    Placed cont = make_static(L->container);
    std::vector<Placed> obst;
    Grid grid(cont.bx0, cont.by0, cont.bx1, cont.by1, std::max(8.0, 2.5 * (double)P.item_r_hi));
    // holes
    for (uint32_t h = 0; h < P.n_holes; h++) {
        for (uint32_t att = 0; att < 10000; att++) {
            std::vector<float> v;
            if (P.hole_kind == 0) {
                float w = rng.uf(P.hole_lo, P.hole_hi), hh = rng.uf(P.hole_lo, P.hole_hi);
                float x = rng.uf((float)cont.bx0, (float)cont.bx1 - w), y = rng.uf((float)cont.by0, (float)cont.by1 - hh);
                v = {x, y, x + w, y, x + w, y + hh, x, y + hh};
            } else {
                uint32_t nv = rng.log_uniform(P.hole_nv_lo, P.hole_nv_hi);
                float R = rng.uf(P.hole_lo, P.hole_hi);
                auto loc = star_polygon(rng, nv, [&](Rng& g, double) { return (double)g.uf(0.5f * R, R); });
                float x = rng.uf((float)cont.bx0, (float)cont.bx1), y = rng.uf((float)cont.by0, (float)cont.by1);
                for (size_t i = 0; i < loc.size() / 2; i++) { loc[2 * i] += x; loc[2 * i + 1] += y; }
                v = loc;
            }
            Placed ph = make_static(v);
            if (!inside_container(ph, cont, 2.0)) continue;
            bool ok = true;
            for (auto& o : obst) ok = ok && separated(ph, o, 2.0);
            if (!ok) continue;
            obst.push_back(ph);
            grid.insert(ph, (int)obst.size() - 1);
            L->holes.push_back(v);
            break;
        }
    }
    // items
    struct Cand { std::vector<float> local; float bound; };
    std::vector<Cand> cands(P.n_items);
    for (uint32_t i = 0; i < P.n_items; i++) {
        uint32_t nv = P.item_nv_lo == P.item_nv_hi ? P.item_nv_lo : rng.log_uniform(P.item_nv_lo, P.item_nv_hi);
        cands[i].local = make_item(rng, nv, P.item_r_lo, P.item_r_hi, P.item_shape, &cands[i].bound);
    }
    if (P.placement == 1) {
        // left-bottom-fill order: decreasing size (lbf/src/opt/search.rs:108-115 orders by diameter)
        std::stable_sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) { return a.bound > b.bound; });
    }
    auto feasible = [&](const Placed& p) {
        if (!inside_container(p, cont, P.min_sep)) return false;
        int i0, j0, i1, j1;
        grid.range(p.bx0 - P.min_sep, p.by0 - P.min_sep, p.bx1 + P.min_sep, p.by1 + P.min_sep, i0, j0, i1, j1);
        for (int j = j0; j <= j1; j++)
            for (int i = i0; i <= i1; i++)
                for (int id : grid.cells[(size_t)j * grid.nx + i])
                    if (!separated(p, obst[id], P.min_sep)) return false;
        return true;
    };
    for (uint32_t i = 0; i < P.n_items; i++) {
        const Cand& c = cands[i];
        bool placed = false;
        float best_cost = 0;
        float best_dt[3] = {0, 0, 0};
        uint32_t budget = P.placement == 1 ? P.lbf_samples : P.max_attempts;
        for (uint32_t att = 0; att < budget; att++) {
            float rot = rot_sample(rng, P.rot_kind);
            float tx = rng.uf((float)cont.bx0 + c.bound, (float)cont.bx1 - c.bound);
            float ty = rng.uf((float)cont.by0 + c.bound, (float)cont.by1 - c.bound);
            Placed p = make_placed(c.local, rot, tx, ty);
            if (P.placement == 1) {
                float cost = 10.0f * (float)p.bx1 + (float)p.by1;   // lbf/src/opt/loss.rs:29-31
                if (placed && cost >= best_cost) continue;
                if (!feasible(p)) continue;
                placed = true;
                best_cost = cost;
                best_dt[0] = rot; best_dt[1] = tx; best_dt[2] = ty;
            } else {
                if (!feasible(p)) continue;
                placed = true;
                best_dt[0] = rot; best_dt[1] = tx; best_dt[2] = ty;
                break;
            }
        }
        if (!placed) continue;
        Placed p = make_placed(c.local, best_dt[0], best_dt[1], best_dt[2]);
        obst.push_back(p);
        grid.insert(p, (int)obst.size() - 1);
        L->items.push_back(c.local);
        L->dts.insert(L->dts.end(), best_dt, best_dt + 3);
    }
    return L;
}

void js_layout_free(void* l) { delete (js_layout*)l; }

// sizes: [container n_verts, n_holes, n_items]
void js_layout_sizes(void* lp, uint32_t* out3) {
    auto* L = (js_layout*)lp;
    out3[0] = (uint32_t)(L->container.size() / 2);
    out3[1] = (uint32_t)L->holes.size();
    out3[2] = (uint32_t)L->items.size();
}
void js_layout_container(void* lp, float* xy) {
    auto* L = (js_layout*)lp;
    std::memcpy(xy, L->container.data(), L->container.size() * 4);
}
uint32_t js_layout_hole_n(void* lp, uint32_t i) { return (uint32_t)(((js_layout*)lp)->holes[i].size() / 2); }
void js_layout_hole(void* lp, uint32_t i, float* xy) {
    auto& h = ((js_layout*)lp)->holes[i];
    std::memcpy(xy, h.data(), h.size() * 4);
}
uint32_t js_layout_item_n(void* lp, uint32_t i) { return (uint32_t)(((js_layout*)lp)->items[i].size() / 2); }
void js_layout_item(void* lp, uint32_t i, float* xy, float* dt3) {
    auto* L = (js_layout*)lp;
    std::memcpy(xy, L->items[i].data(), L->items[i].size() * 4);
    std::memcpy(dt3, &L->dts[3 * i], 12);
}

// One polygon centred at the origin (query items).  Returns the vertex count; xy must hold
// 2*nv_hi floats.
uint32_t js_polygon(uint64_t seed, uint32_t nv_lo, uint32_t nv_hi, float r_lo, float r_hi, uint32_t shape, float* xy) {
    Rng rng(seed);
    uint32_t nv = nv_lo == nv_hi ? nv_lo : rng.log_uniform(nv_lo, nv_hi);
    float b;
    auto v = make_item(rng, nv, r_lo, r_hi, shape, &b);
    std::memcpy(xy, v.data(), v.size() * 4);
    return nv;
}

// n DTransformations: translation uniform in [x0,x1]x[y0,y1], rotation continuous (0) or
// multiples of 45 degrees (1).  Layout per transform: rotation, tx, ty.
void js_transforms(uint64_t seed, uint64_t n, float x0, float y0, float x1, float y1, uint32_t rot_kind, float* dt) {
    Rng rng(seed);
    for (uint64_t i = 0; i < n; i++) {
        dt[3 * i] = rot_sample(rng, rot_kind);
        dt[3 * i + 1] = rng.uf(x0, x1);
        dt[3 * i + 2] = rng.uf(y0, y1);
    }
}

}  // extern "C"
