// Piers of many polygons at once on the GPU: generate_piers (fail_fast/piers.rs:25-98) bit for
// bit with the host routine, for every polygon of a batch with n_ff_piers > 0.
//
// The reference casts 90 x 200 rays across each polygon, clips them to its interior, keeps
// the grid points of a 100 x 100 lattice that lie inside the polygon and outside the relevant
// poles, and then picks n piers one after another: each pick scores every clipped ray by a
// sum over the grid points (loss_function, :185-213) and keeps the lowest score, the first on
// ties.  Here:
//   k_pier_count     one thread per (polygon, ray): the number of crossings with the polygon's
//                    edges (and clip's assertion that both ray ends lie outside);
//   k_pier_layout    one workgroup per polygon: crossing offsets (prefix sum in ray order)
//                    and the grid points, compacted in lattice order;
//   k_pier_hits      per ray: the crossings written in edge order, then sorted in place by
//                    distance from the ray start (stable, as clip's sort), clips counted;
//   k_pier_seg_scan  per polygon: clip offsets;
//   k_pier_segs      per ray: the clips written at their offsets;
//   k_pier_loss      one thread per clip, the grid summed sequentially in lattice order (the
//                    reference's f32 sum order), every thread of a workgroup on the same
//                    polygon so the grid is read with scalar loads;
//   k_pier_select    one wave per polygon: argmin (first index on ties), the pier, and the
//                    grid's distance-to-piers update.
// A polygon for which the reference would panic or bail (a ray end inside it, no clip at all,
// a NaN distance or score) is handed back to the host routine, which reports the same.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/jagua_amd.h"
#include "jg_internal.hpp"

using namespace jg;

namespace {

constexpr uint32_t PIER_THREADS = 256;
constexpr uint32_t PIER_CHUNK = 1024;      // polygons per device pass (bounds device memory)

enum : uint32_t { PS_OK = 0, PS_HOST = 1 };

struct PierJob {
    uint32_t v_begin, n_v, p_begin, n_p;
    float4 bbox;
    PierSetup ps;
};

struct PierArgs {
    const PierJob* jobs;
    const float2* verts;
    const float4* poles;
    const float2* angles;      // (sin, cos) of the 90 ray angles
    uint32_t* status;          // [n] PS_*
    uint32_t* hit_off;         // [n][PIER_N_RAYS] crossings per ray, then their offsets
    uint32_t* n_hit;           // [n]
    const uint64_t* hit_base;  // [n]
    float4* hits;              // crossings of all polygons: distance, x, y, -
    uint32_t* seg_off;         // [n][PIER_N_RAYS] clips per ray, then their offsets
    uint32_t* n_seg;           // [n]
    const uint32_t* seg_base;  // [n]
    float4* grid;              // [n][PIER_N_GRID] x, y, distance to the chosen piers, to the poles
    uint32_t* n_grid;          // [n]
    float4* segs;              // clips of all polygons
    float* loss;               // one per clip
    const uint2* blocks;       // (polygon, first clip) of each k_pier_loss workgroup
    float4* piers;             // [n][n_piers]
    uint32_t n, n_piers, it;
};

__device__ float4 ray_of(const PierArgs& a, const PierJob& j, uint32_t r) {
    const float2 sc = a.angles[r / PIER_RAYS];
    return pier_ray(sc.x, sc.y, j.ps.sq.x + j.ps.dx * (float)(r % PIER_RAYS), j.ps.r0, j.ps.r1);
}

__global__ __launch_bounds__(PIER_THREADS) void k_pier_count(PierArgs a) {
    const uint64_t gid = (uint64_t)blockIdx.x * PIER_THREADS + threadIdx.x;
    const uint32_t p = (uint32_t)(gid / PIER_N_RAYS), r = (uint32_t)(gid % PIER_N_RAYS);
    if (p >= a.n) return;
    const PierJob j = a.jobs[p];
    const float4 ray = ray_of(a, j, r);
    const float2* v = a.verts + j.v_begin;
    uint32_t c = 0;
    if (poly_has_point(v, j.n_v, j.bbox, ray.x, ray.y) || poly_has_point(v, j.n_v, j.bbox, ray.z, ray.w)) {
        atomicOr(&a.status[p], PS_HOST);   // clip's assertion (piers.rs:126)
    } else {
        for (uint32_t k = 0; k < j.n_v; k++) {
            const float2 s = v[k], e = v[k + 1 == j.n_v ? 0 : k + 1];
            c += edge_edge(make_float4(s.x, s.y, e.x, e.y), ray) ? 1u : 0u;
        }
    }
    a.hit_off[gid] = c;
}

// Crossings of one ray in the order clip sorts them (by distance from the ray start, stable
// in edge order), then its clips counted.
__global__ __launch_bounds__(PIER_THREADS) void k_pier_hits(PierArgs a) {
    const uint64_t gid = (uint64_t)blockIdx.x * PIER_THREADS + threadIdx.x;
    const uint32_t p = (uint32_t)(gid / PIER_N_RAYS), r = (uint32_t)(gid % PIER_N_RAYS);
    if (p >= a.n) return;
    if (a.status[p] != PS_OK) {
        a.seg_off[gid] = 0;
        return;
    }
    const PierJob j = a.jobs[p];
    const float4 ray = ray_of(a, j, r);
    const float2* v = a.verts + j.v_begin;
    const uint32_t off = a.hit_off[gid];
    const uint32_t end = r + 1 < PIER_N_RAYS ? a.hit_off[gid + 1] : a.n_hit[p];
    float4* h = a.hits + a.hit_base[p] + off;
    const uint32_t cap = end - off;
    uint32_t n = 0;
    bool nan = false;
    for (uint32_t k = 0; k < j.n_v && n < cap; k++) {
        const float2 s = v[k], e = v[k + 1 == j.n_v ? 0 : k + 1];
        const float4 ed = make_float4(s.x, s.y, e.x, e.y);
        if (!edge_edge(ed, ray)) continue;
        float d;
        const float2 loc = pier_hit(ed, ray, d);
        nan |= d != d;
        const float4 x = make_float4(d, loc.x, loc.y, 0.0f);
        uint32_t q = n;                           // insertion after every equal distance
        while (q > 0 && h[q - 1].x > d) {
            h[q] = h[q - 1];
            q--;
        }
        h[q] = x;
        n++;
    }
    if (nan) atomicOr(&a.status[p], PS_HOST);
    uint32_t c = 0;
    for (uint32_t k = 0; k + 1 < n; k += 2)
        if (!(h[k].y == h[k + 1].y && h[k].z == h[k + 1].z)) c++;
    a.seg_off[gid] = c;
}

__global__ __launch_bounds__(PIER_THREADS) void k_pier_segs(PierArgs a) {
    const uint64_t gid = (uint64_t)blockIdx.x * PIER_THREADS + threadIdx.x;
    const uint32_t p = (uint32_t)(gid / PIER_N_RAYS), r = (uint32_t)(gid % PIER_N_RAYS);
    if (p >= a.n || a.status[p] != PS_OK) return;
    const uint32_t off = a.hit_off[gid];
    const uint32_t n = (r + 1 < PIER_N_RAYS ? a.hit_off[gid + 1] : a.n_hit[p]) - off;
    const float4* h = a.hits + a.hit_base[p] + off;
    const uint32_t so = a.seg_off[gid], se = a.n_seg[p];
    uint32_t c = 0;
    for (uint32_t k = 0; k + 1 < n; k += 2) {
        if (h[k].y == h[k + 1].y && h[k].z == h[k + 1].z) continue;
        if (so + c < se)
            a.segs[a.seg_base[p] + so + c] = pier_trim(make_float2(h[k].y, h[k].z), make_float2(h[k + 1].y, h[k + 1].z));
        c++;
    }
}

// Exclusive prefix sum over a 256-thread workgroup; total to every thread.
__device__ uint32_t block_scan(uint32_t x, uint32_t* sh, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t inc = x;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, 64);
        if ((int)lane >= o) inc += y;
    }
    if (lane == 63u) sh[w] = inc;
    __syncthreads();
    uint32_t pre = 0;
    for (uint32_t k = 0; k < w; k++) pre += sh[k];
    total = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return pre + inc - x;
}

__global__ __launch_bounds__(PIER_THREADS) void k_pier_layout(PierArgs a) {
    __shared__ uint32_t sh[4];
    const uint32_t p = blockIdx.x;
    if (p >= a.n) return;
    uint32_t* ro = a.hit_off + (uint64_t)p * PIER_N_RAYS;
    uint32_t base = 0;
    for (uint32_t c = 0; c < PIER_N_RAYS; c += PIER_THREADS) {
        const uint32_t r = c + threadIdx.x;
        const uint32_t x = r < PIER_N_RAYS ? ro[r] : 0u;
        uint32_t tot;
        const uint32_t ex = block_scan(x, sh, tot);
        if (r < PIER_N_RAYS) ro[r] = base + ex;
        base += tot;
    }
    if (threadIdx.x == 0) a.n_hit[p] = base;
    const PierJob j = a.jobs[p];
    float4* gr = a.grid + (uint64_t)p * PIER_N_GRID;
    base = 0;
    for (uint32_t c = 0; c < PIER_N_GRID; c += PIER_THREADS) {
        const uint32_t g = c + threadIdx.x;
        float2 pt = make_float2(0.0f, 0.0f);
        float dp = 0.0f;
        const bool keep = g < PIER_N_GRID && pier_grid_point(a.verts + j.v_begin, j.n_v, j.bbox, a.poles + j.p_begin,
                                                             j.n_p, j.ps, g / PIER_GRID, g % PIER_GRID, pt, dp);
        uint32_t tot;
        const uint32_t ex = block_scan(keep ? 1u : 0u, sh, tot);
        if (keep) gr[base + ex] = make_float4(pt.x, pt.y, j.ps.forfeit, dp);
        base += tot;
    }
    if (threadIdx.x == 0) a.n_grid[p] = base;
}

__global__ __launch_bounds__(PIER_THREADS) void k_pier_seg_scan(PierArgs a) {
    __shared__ uint32_t sh[4];
    const uint32_t p = blockIdx.x;
    if (p >= a.n) return;
    uint32_t* so = a.seg_off + (uint64_t)p * PIER_N_RAYS;
    uint32_t base = 0;
    for (uint32_t c = 0; c < PIER_N_RAYS; c += PIER_THREADS) {
        const uint32_t r = c + threadIdx.x;
        const uint32_t x = r < PIER_N_RAYS ? so[r] : 0u;
        uint32_t tot;
        const uint32_t ex = block_scan(x, sh, tot);
        if (r < PIER_N_RAYS) so[r] = base + ex;
        base += tot;
    }
    if (threadIdx.x == 0) a.n_seg[p] = base;
}

__global__ __launch_bounds__(PIER_THREADS) void k_pier_loss(PierArgs a) {
    const uint2 b = a.blocks[blockIdx.x];
    const uint32_t p = b.x, s = b.y + threadIdx.x;
    const uint32_t ns = a.n_seg[p], ng = a.n_grid[p];
    const float infl = a.jobs[p].ps.infl;
    const float4* gr = a.grid + (uint64_t)p * PIER_N_GRID;
    const float4 seg = a.segs[a.seg_base[p] + (s < ns ? s : 0u)];
    float loss = -0.0f;                       // f32 Sum's starting value
    for (uint32_t g = 0; g < ng; g++) {
        const float4 q = gr[g];
        loss += pier_point_loss(q.z, q.w, edge_point_dist(seg, q.x, q.y), infl);
    }
    if (s < ns) a.loss[a.seg_base[p] + s] = loss;
}

__global__ __launch_bounds__(64) void k_pier_select(PierArgs a) {
    const uint32_t p = blockIdx.x, lane = threadIdx.x;
    if (p >= a.n || a.status[p] != PS_OK) return;
    const uint32_t ns = a.n_seg[p], ng = a.n_grid[p];
    const float* L = a.loss + a.seg_base[p];
    float best = INFINITY;
    uint32_t bi = 0xffffffffu;
    bool nan = false;
    for (uint32_t s = lane; s < ns; s += 64) {
        const float l = L[s];
        nan |= l != l;
        if (bi == 0xffffffffu || l < best) {
            best = l;
            bi = s;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const float ob = __shfl_xor(best, o, 64);
        const uint32_t oi = __shfl_xor(bi, o, 64);
        if (oi != 0xffffffffu && (bi == 0xffffffffu || ob < best || (ob == best && oi < bi))) {
            best = ob;
            bi = oi;
        }
    }
    if (__any(nan) || bi == 0xffffffffu) {
        if (lane == 0) a.status[p] = PS_HOST;
        return;
    }
    const float4 pier = a.segs[a.seg_base[p] + bi];
    if (lane == 0) a.piers[(uint64_t)p * a.n_piers + a.it] = pier;
    float4* gr = a.grid + (uint64_t)p * PIER_N_GRID;
    for (uint32_t g = lane; g < ng; g += 64) {
        float4 q = gr[g];
        q.z = mn(q.z, edge_point_dist(pier, q.x, q.y));
        gr[g] = q;
    }
}

size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

namespace jg {

jg_status piers_batch(const std::vector<const Polygon*>& polys, const std::vector<std::vector<float4>>& poles,
                      uint32_t n_piers, std::vector<std::vector<float4>>& out, std::vector<uint8_t>& host) {
    const size_t N = polys.size();
    out.assign(N, {});
    host.assign(N, 0);
    if (N == 0 || n_piers == 0) return JG_OK;
    std::vector<float> sc(2 * PIER_ANGLES);
    pier_angles(sc.data());
    for (size_t c0 = 0; c0 < N; c0 += PIER_CHUNK) {
        const uint32_t n = (uint32_t)std::min<size_t>(PIER_CHUNK, N - c0);
        std::vector<PierJob> jobs(n);
        std::vector<float2> verts;
        std::vector<float4> pl;
        for (uint32_t i = 0; i < n; i++) {
            const Polygon& P = *polys[c0 + i];
            PierJob& j = jobs[i];
            j.v_begin = (uint32_t)verts.size();
            j.n_v = P.n();
            j.p_begin = (uint32_t)pl.size();
            j.n_p = (uint32_t)poles[c0 + i].size();
            j.bbox = P.bbox;
            pier_setup(P, j.ps);
            verts.insert(verts.end(), P.v.begin(), P.v.end());
            pl.insert(pl.end(), poles[c0 + i].begin(), poles[c0 + i].end());
        }
        // first device buffer: inputs, per-ray counts and offsets, the grid
        size_t off = 0;
        auto place = [&](size_t bytes) {
            const size_t o = off;
            off = up256(off + bytes);
            return o;
        };
        const size_t o_jobs = place(sizeof(PierJob) * n), o_v = place(sizeof(float2) * (verts.size() + 1)),
                     o_p = place(sizeof(float4) * (pl.size() + 1)), o_a = place(sizeof(float) * sc.size()),
                     o_st = place(4ull * n), o_ho = place(4ull * n * PIER_N_RAYS), o_nh = place(4ull * n),
                     o_hb = place(8ull * n), o_so = place(4ull * n * PIER_N_RAYS), o_ns = place(4ull * n),
                     o_sb = place(4ull * n), o_g = place(sizeof(float4) * (size_t)n * PIER_N_GRID),
                     o_ng = place(4ull * n), o_pi = place(sizeof(float4) * (size_t)n * n_piers);
        DevBlob blob;
        jg_status st = blob.reserve(off);
        if (st != JG_OK) return st;
        uint8_t* B = (uint8_t*)blob.ptr;
        auto h2d = [&](size_t o, const void* src, size_t bytes) {
            return bytes ? hip_check(hipMemcpy(B + o, src, bytes, hipMemcpyHostToDevice), "hipMemcpy pier inputs") : JG_OK;
        };
        st = h2d(o_jobs, jobs.data(), sizeof(PierJob) * n);
        if (st == JG_OK) st = h2d(o_v, verts.data(), sizeof(float2) * verts.size());
        if (st == JG_OK) st = h2d(o_p, pl.data(), sizeof(float4) * pl.size());
        if (st == JG_OK) st = h2d(o_a, sc.data(), sizeof(float) * sc.size());
        if (st == JG_OK) st = hip_check(hipMemset(B + o_st, 0, 4ull * n), "hipMemset pier status");
        if (st != JG_OK) return st;
        PierArgs a{};
        a.jobs = (const PierJob*)(B + o_jobs);
        a.verts = (const float2*)(B + o_v);
        a.poles = (const float4*)(B + o_p);
        a.angles = (const float2*)(B + o_a);
        a.status = (uint32_t*)(B + o_st);
        a.hit_off = (uint32_t*)(B + o_ho);
        a.n_hit = (uint32_t*)(B + o_nh);
        a.hit_base = (const uint64_t*)(B + o_hb);
        a.seg_off = (uint32_t*)(B + o_so);
        a.n_seg = (uint32_t*)(B + o_ns);
        a.seg_base = (const uint32_t*)(B + o_sb);
        a.grid = (float4*)(B + o_g);
        a.n_grid = (uint32_t*)(B + o_ng);
        a.piers = (float4*)(B + o_pi);
        a.n = n;
        a.n_piers = n_piers;
        const uint32_t ray_blocks = (uint32_t)(((uint64_t)n * PIER_N_RAYS + PIER_THREADS - 1) / PIER_THREADS);
        auto launched = [&](const char* what) { return hip_check(hipGetLastError(), what); };
        auto d2h = [&](std::vector<uint32_t>& v, size_t o) {
            return hip_check(hipMemcpy(v.data(), B + o, 4ull * n, hipMemcpyDeviceToHost), "hipMemcpy pier counts");
        };
        hipLaunchKernelGGL(k_pier_count, dim3(ray_blocks), dim3(PIER_THREADS), 0, nullptr, a);
        st = launched("pier count launch");
        if (st != JG_OK) return st;
        hipLaunchKernelGGL(k_pier_layout, dim3(n), dim3(PIER_THREADS), 0, nullptr, a);
        st = launched("pier layout launch");
        std::vector<uint32_t> nh(n), ns(n), status(n);
        if (st == JG_OK) st = d2h(nh, o_nh);
        if (st == JG_OK) st = d2h(status, o_st);
        if (st != JG_OK) return st;
        // crossing buffer
        std::vector<uint64_t> hbase(n, 0);
        uint64_t n_hits = 0;
        for (uint32_t i = 0; i < n; i++) {
            if (status[i] != PS_OK) continue;
            hbase[i] = n_hits;
            n_hits += nh[i];
        }
        DevBlob hblob;
        st = hblob.reserve(sizeof(float4) * (n_hits + 1));
        if (st == JG_OK) st = h2d(o_hb, hbase.data(), 8ull * n);
        if (st != JG_OK) return st;
        a.hits = (float4*)hblob.ptr;
        hipLaunchKernelGGL(k_pier_hits, dim3(ray_blocks), dim3(PIER_THREADS), 0, nullptr, a);
        st = launched("pier crossings launch");
        if (st != JG_OK) return st;
        hipLaunchKernelGGL(k_pier_seg_scan, dim3(n), dim3(PIER_THREADS), 0, nullptr, a);
        st = launched("pier clip scan launch");
        if (st == JG_OK) st = d2h(ns, o_ns);
        if (st == JG_OK) st = d2h(status, o_st);
        if (st != JG_OK) return st;
        // clip buffers: polygons with no clip (the reference bails) or a flagged ray go to the host
        std::vector<uint32_t> base(n, 0);
        std::vector<uint2> blocks;
        uint64_t total = 0;
        for (uint32_t i = 0; i < n; i++) {
            if (status[i] == PS_OK && ns[i] == 0) status[i] = PS_HOST;
            if (status[i] != PS_OK) continue;
            if (total + ns[i] > 0xffffffffull) return fail(JG_ERR_INTERNAL, "pier clip count overflow");
            base[i] = (uint32_t)total;
            for (uint32_t s = 0; s < ns[i]; s += PIER_THREADS) blocks.push_back(make_uint2(i, s));
            total += ns[i];
        }
        size_t off2 = 0;
        auto place2 = [&](size_t bytes) {
            const size_t o = off2;
            off2 = up256(off2 + bytes);
            return o;
        };
        const size_t o_sg = place2(sizeof(float4) * (total + 1)), o_l = place2(sizeof(float) * (total + 1)),
                     o_b = place2(sizeof(uint2) * (blocks.size() + 1));
        DevBlob blob2;
        st = blob2.reserve(off2);
        if (st != JG_OK) return st;
        uint8_t* B2 = (uint8_t*)blob2.ptr;
        a.segs = (float4*)(B2 + o_sg);
        a.loss = (float*)(B2 + o_l);
        a.blocks = (const uint2*)(B2 + o_b);
        st = h2d(o_sb, base.data(), 4ull * n);
        if (st == JG_OK) st = h2d(o_st, status.data(), 4ull * n);
        if (st == JG_OK && !blocks.empty())
            st = hip_check(hipMemcpy(B2 + o_b, blocks.data(), sizeof(uint2) * blocks.size(), hipMemcpyHostToDevice),
                           "hipMemcpy pier blocks");
        if (st != JG_OK) return st;
        if (!blocks.empty()) {
            hipLaunchKernelGGL(k_pier_segs, dim3(ray_blocks), dim3(PIER_THREADS), 0, nullptr, a);
            st = launched("pier clips launch");
            for (uint32_t it = 0; it < n_piers && st == JG_OK; it++) {
                a.it = it;
                hipLaunchKernelGGL(k_pier_loss, dim3((uint32_t)blocks.size()), dim3(PIER_THREADS), 0, nullptr, a);
                st = launched("pier loss launch");
                if (st != JG_OK) break;
                hipLaunchKernelGGL(k_pier_select, dim3(n), dim3(64), 0, nullptr, a);
                st = launched("pier select launch");
            }
            if (st == JG_OK) st = hip_check(hipDeviceSynchronize(), "pier search");
        }
        std::vector<float4> pi((size_t)n * n_piers);
        if (st == JG_OK) st = hip_check(hipMemcpy(status.data(), B + o_st, 4ull * n, hipMemcpyDeviceToHost), "hipMemcpy");
        if (st == JG_OK)
            st = hip_check(hipMemcpy(pi.data(), B + o_pi, sizeof(float4) * pi.size(), hipMemcpyDeviceToHost), "hipMemcpy");
        if (st != JG_OK) return st;
        for (uint32_t i = 0; i < n; i++) {
            if (status[i] != PS_OK) {
                host[c0 + i] = 1;
                continue;
            }
            out[c0 + i].assign(pi.begin() + (long)i * n_piers, pi.begin() + (long)(i + 1) * n_piers);
        }
    }
    return JG_OK;
}

}  // namespace jg
