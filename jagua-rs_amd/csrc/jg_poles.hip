// Pole searches of many polygons at once on the GPU: SPolygon::new's pole of inaccessibility
// (simple_polygon.rs:41-71 via fail_fast/pole.rs:52-80) and generate_surrogate_poles
// (fail_fast/pole.rs:12-47), bit for bit with the sequential host search.
//
// The reference search pops cells from a FIFO; every split pushes four children to the back,
// so cells are visited level by level, in order.  Within a level the running best radius
// after cell i is max(best before the level, d_0..d_i), a prefix maximum, and cell i splits
// iff radius_i + d_i exceeds it.  One workgroup per polygon therefore evaluates a level's
// cells in parallel (one cell per thread, chunks of 256), takes a block-wide prefix max and a
// prefix sum of the split flags, and writes the children of the split cells in order into the
// next level's frontier.  The cell arithmetic is the shared pole_cell (jg_geom.hpp), so the
// verdicts of every comparison equal the host's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/jagua_amd.h"
#include "jg_internal.hpp"

using namespace jg;

namespace {

constexpr uint32_t POLE_THREADS = 256;
constexpr uint32_t FRONTIER_CAP = 8192;       // cells per level held per workgroup
constexpr uint32_t MAX_POLES = 1000;          // generate_surrogate_poles' assertion (pole.rs:40-43)

struct PoleJob {
    uint32_t v_begin, n;
    float area, pad;
    float4 bbox;
};

enum : uint32_t { PJ_OK = 0, PJ_FAIL = 1, PJ_HOST = 2 };

struct PoleArgs {
    const PoleJob* jobs;
    const float2* verts;
    float4* poles;            // [n_jobs][pole_cap]: the POI, then the surrogate's further poles
    uint32_t* n_poles;
    uint32_t* status;
    float4* frontier;         // [grid][2][FRONTIER_CAP]
    uint32_t* counter;
    uint32_t n_jobs, pole_cap, surrogate;
    uint32_t lim_n[3];
    float lim_cov[3];
};

__device__ float block_incl_max(float x, float* sh, float& total) {
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        const float y = __shfl_up(x, o, 64);
        if ((int)lane >= o) x = fmaxf(x, y);
    }
    if (lane == 63u) sh[w] = x;
    __syncthreads();
    float pre = -INFINITY;
    for (uint32_t k = 0; k < w; k++) pre = fmaxf(pre, sh[k]);
    total = fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
    __syncthreads();
    return fmaxf(pre, x);
}

__device__ uint32_t block_excl_sum(uint32_t x, uint32_t* sh, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t incl = x;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o, 64);
        if ((int)lane >= o) incl += y;
    }
    if (lane == 63u) sh[w] = incl;
    __syncthreads();
    uint32_t pre = 0;
    for (uint32_t k = 0; k < w; k++) pre += sh[k];
    total = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return pre + incl - x;
}

// compute_pole (pole.rs:52-80) for one polygon against poles P[0..np); returns PJ_OK with the
// best circle, PJ_FAIL when no cell had a positive distance, PJ_HOST on frontier overflow.
__device__ uint32_t search_pole(const PoleJob& job, const float2* v, const float4* P, uint32_t np, float4* F,
                                float4* G, float* shf, uint32_t* shu, uint32_t* s_idx, float4& best) {
    const uint32_t tid = threadIdx.x;
    float best_r = 0.0f;
    bool have = false;
    best = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (tid == 0) F[0] = box_inflate_to_square(job.bbox);
    __syncthreads();
    uint32_t nF = 1;
    for (int level = POLE_LEVELS; level >= 0 && nF > 0; level--) {
        uint32_t nG = 0;                              // split cells of this level so far
        for (uint32_t c0 = 0; c0 < nF; c0 += POLE_THREADS) {
            const uint32_t i = c0 + tid;
            const bool valid = i < nF;
            float4 box = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float radius = 0.0f, d = -INFINITY;
            if (valid) {
                box = F[i];
                d = pole_cell(v, job.n, job.bbox, P, np, box, radius);
            }
            float chunk_max;
            const float incl = block_incl_max(d, shf, chunk_max);
            const bool split = valid && level > 0 && (radius + d > fmaxf(best_r, incl));
            if (chunk_max > best_r) {                 // the first cell reaching the chunk max becomes best
                if (tid == 0) *s_idx = 0xffffffffu;
                __syncthreads();
                if (valid && d == chunk_max) atomicMin(s_idx, i);
                __syncthreads();
                const float4 bb = F[*s_idx];
                best = make_float4((bb.x + bb.z) / 2.0f, (bb.y + bb.w) / 2.0f, chunk_max, 0.0f);
                best_r = chunk_max;
                have = true;
                __syncthreads();
            }
            uint32_t tot;
            const uint32_t ex = block_excl_sum(split ? 1u : 0u, shu, tot);
            if ((nG + tot) * 4u > FRONTIER_CAP) return PJ_HOST;
            if (split)
                for (int k = 0; k < 4; k++) G[(nG + ex) * 4u + (uint32_t)k] = box_quadrant(box, k);
            nG += tot;
        }
        __syncthreads();
        float4* t = F;
        F = G;
        G = t;
        nF = nG * 4u;
    }
    return have ? PJ_OK : PJ_FAIL;
}

__global__ __launch_bounds__(POLE_THREADS) void k_poles(PoleArgs a) {
    __shared__ float shf[4];
    __shared__ uint32_t shu[4];
    __shared__ uint32_t s_job, s_idx;
    const uint32_t tid = threadIdx.x;
    float4* F0 = a.frontier + (size_t)blockIdx.x * 2u * FRONTIER_CAP;
    float4* F1 = F0 + FRONTIER_CAP;
    const float PI_F = 3.14159265358979323846f;
    for (;;) {
        if (tid == 0) s_job = atomicAdd(a.counter, 1u);
        __syncthreads();
        const uint32_t j = s_job;
        __syncthreads();
        if (j >= a.n_jobs) return;
        const PoleJob job = a.jobs[j];
        const float2* v = a.verts + job.v_begin;
        float4* P = a.poles + (size_t)j * a.pole_cap;
        float4 best;
        uint32_t st = search_pole(job, v, P, 0, F0, F1, shf, shu, &s_idx, best);
        uint32_t np = 0;
        if (st == PJ_OK) {
            if (tid == 0) P[0] = best;
            np = 1;
            __syncthreads();
        }
        if (st == PJ_OK && a.surrogate) {
            // generate_surrogate_poles: the POI, then poles until a count/coverage limit holds
            float covered = best.z * best.z * PI_F;
            for (;;) {
                if (np >= a.pole_cap) { st = PJ_HOST; break; }
                st = search_pole(job, v, P, np, F0, F1, shf, shu, &s_idx, best);
                if (st != PJ_OK) break;
                covered += best.z * best.z * PI_F;
                if (tid == 0) P[np] = best;
                np++;
                __syncthreads();
                const float coverage = covered / job.area;
                bool any = false;
                uint32_t limit = 0;
                for (int i = 0; i < 3; i++)
                    if (coverage > a.lim_cov[i] && (!any || a.lim_n[i] < limit)) {
                        limit = a.lim_n[i];
                        any = true;
                    }
                if (any && np >= limit) break;
                if (np >= MAX_POLES) { st = PJ_FAIL; break; }
            }
        }
        if (tid == 0) {
            a.status[j] = st;
            a.n_poles[j] = np;
        }
        __syncthreads();
    }
}

template <class F>
void parallel_for(size_t n, F fn) {
    const size_t nt = std::max<size_t>(1, std::min<size_t>({(size_t)std::thread::hardware_concurrency(), 16, n}));
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            for (size_t i = t; i < n; i += nt) fn(i);
        });
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" jg_status jg_spolygon_new_batch(const float* xy, const uint32_t* offsets, uint32_t n_polys,
                                           const jg_surrogate_config* cfg, jg_spolygon** out,
                                           uint32_t* n_host_completed) {
    if (!out || (n_polys && (!xy || !offsets))) return fail(JG_ERR_INVALID_ARG, "null argument");
    if (n_host_completed) *n_host_completed = 0;
    if (n_polys == 0) return JG_OK;
    for (uint32_t i = 0; i < n_polys; i++)
        if (offsets[i + 1] < offsets[i]) return fail(JG_ERR_INVALID_ARG, "offsets must be non-decreasing");
    // SPolygon::new up to the pole search, on host threads
    std::vector<Polygon> polys(n_polys);
    std::vector<std::string> errs(n_polys);
    std::vector<char> ok(n_polys, 0);
    parallel_for(n_polys, [&](size_t i) {
        ok[i] = polygon_prepare(xy + 2 * (size_t)offsets[i], offsets[i + 1] - offsets[i], polys[i], errs[i]) ? 1 : 0;
    });
    for (uint32_t i = 0; i < n_polys; i++)
        if (!ok[i]) return fail(JG_ERR_POLYGON, "polygon " + std::to_string(i) + ": " + errs[i]);
    jg_status st = device_ready();
    if (st != JG_OK) return st;

    uint32_t pole_cap = 1;
    if (cfg) {
        pole_cap = 2;
        for (int i = 0; i < 3; i++) pole_cap = std::max(pole_cap, cfg->n_pole_limits_n[i]);
        pole_cap = std::min(pole_cap, MAX_POLES);
    }
    std::vector<PoleJob> jobs(n_polys);
    std::vector<float2> verts;
    verts.reserve(offsets[n_polys] - offsets[0]);
    for (uint32_t i = 0; i < n_polys; i++) {
        jobs[i] = PoleJob{(uint32_t)verts.size(), polys[i].n(), polys[i].area, 0.0f, polys[i].bbox};
        verts.insert(verts.end(), polys[i].v.begin(), polys[i].v.end());
    }
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint32_t grid = std::max<uint32_t>(1u, std::min<uint32_t>(n_polys, 4u * (uint32_t)cus));

    const size_t b_jobs = sizeof(PoleJob) * n_polys;
    const size_t o_verts = (b_jobs + 255) & ~(size_t)255;
    const size_t o_poles = (o_verts + sizeof(float2) * verts.size() + 255) & ~(size_t)255;
    const size_t o_np = (o_poles + sizeof(float4) * (size_t)pole_cap * n_polys + 255) & ~(size_t)255;
    const size_t o_st = (o_np + 4 * (size_t)n_polys + 255) & ~(size_t)255;
    const size_t o_ctr = (o_st + 4 * (size_t)n_polys + 255) & ~(size_t)255;
    const size_t o_fr = o_ctr + 256;
    const size_t total = o_fr + sizeof(float4) * 2 * FRONTIER_CAP * (size_t)grid;
    DevBlob blob;
    st = blob.reserve(total);
    if (st != JG_OK) return st;
    uint8_t* base = (uint8_t*)blob.ptr;
    st = hip_check(hipMemcpy(base, jobs.data(), b_jobs, hipMemcpyHostToDevice), "hipMemcpy pole jobs");
    if (st == JG_OK)
        st = hip_check(hipMemcpy(base + o_verts, verts.data(), sizeof(float2) * verts.size(), hipMemcpyHostToDevice),
                       "hipMemcpy pole vertices");
    if (st == JG_OK) st = hip_check(hipMemset(base + o_ctr, 0, 256), "hipMemset pole counter");
    if (st != JG_OK) return st;
    PoleArgs a{};
    a.jobs = (const PoleJob*)base;
    a.verts = (const float2*)(base + o_verts);
    a.poles = (float4*)(base + o_poles);
    a.n_poles = (uint32_t*)(base + o_np);
    a.status = (uint32_t*)(base + o_st);
    a.frontier = (float4*)(base + o_fr);
    a.counter = (uint32_t*)(base + o_ctr);
    a.n_jobs = n_polys;
    a.pole_cap = pole_cap;
    a.surrogate = cfg ? 1u : 0u;
    for (int i = 0; i < 3; i++) {
        a.lim_n[i] = cfg ? cfg->n_pole_limits_n[i] : 0u;
        a.lim_cov[i] = cfg ? cfg->n_pole_limits_cov[i] : 0.0f;
    }
    hipLaunchKernelGGL(k_poles, dim3(grid), dim3(POLE_THREADS), 0, nullptr, a);
    st = hip_check(hipGetLastError(), "pole search launch");
    if (st == JG_OK) st = hip_check(hipDeviceSynchronize(), "pole search");
    std::vector<float4> poles((size_t)pole_cap * n_polys);
    std::vector<uint32_t> np(n_polys), status(n_polys);
    if (st == JG_OK)
        st = hip_check(hipMemcpy(poles.data(), base + o_poles, sizeof(float4) * poles.size(), hipMemcpyDeviceToHost),
                       "hipMemcpy poles");
    if (st == JG_OK) st = hip_check(hipMemcpy(np.data(), base + o_np, 4 * (size_t)n_polys, hipMemcpyDeviceToHost), "hipMemcpy");
    if (st == JG_OK)
        st = hip_check(hipMemcpy(status.data(), base + o_st, 4 * (size_t)n_polys, hipMemcpyDeviceToHost), "hipMemcpy");
    if (st != JG_OK) return st;

    // finish: the POI, the surrogate from its poles; polygons whose search outgrew the device
    // buffers are searched by the host routine instead
    uint32_t n_host = 0;
    for (uint32_t i = 0; i < n_polys; i++) {
        if (status[i] == PJ_FAIL) return fail(JG_ERR_POLYGON, "polygon " + std::to_string(i) + ": no pole found");
        if (status[i] == PJ_HOST) n_host++;
    }
    // piers of the polygons whose poles came from the device, on the device as well
    std::vector<std::vector<float4>> piers;
    std::vector<uint8_t> pier_host;
    std::vector<uint32_t> pier_of(n_polys, 0xffffffffu);
    if (cfg && cfg->n_ff_piers > 0) {
        std::vector<const Polygon*> pp;
        std::vector<std::vector<float4>> rel;
        for (uint32_t i = 0; i < n_polys; i++) {
            if (status[i] != PJ_OK) continue;
            const float4* P = poles.data() + (size_t)i * pole_cap;
            const uint32_t n_ff = std::min(cfg->n_ff_poles, np[i]);
            polys[i].poi = P[0];
            pier_of[i] = (uint32_t)pp.size();
            pp.push_back(&polys[i]);
            rel.emplace_back(P, P + n_ff);
        }
        st = piers_batch(pp, rel, cfg->n_ff_piers, piers, pier_host);
        if (st != JG_OK) return st;
        for (uint8_t h : pier_host) n_host += h;
    }
    parallel_for(n_polys, [&](size_t i) {
        Polygon& p = polys[i];
        std::string& err = errs[i];
        ok[i] = 1;
        if (status[i] == PJ_HOST) {
            Polygon q;
            const uint32_t n = offsets[i + 1] - offsets[i];
            if (!polygon_new(xy + 2 * (size_t)offsets[i], n, q, err) ||
                (cfg && !polygon_generate_surrogate(q, *cfg, err)))
                ok[i] = 0;
            p = std::move(q);
            return;
        }
        const float4* P = poles.data() + (size_t)i * pole_cap;
        p.poi = P[0];
        const std::vector<float4>* dev_piers =
            pier_of[i] != 0xffffffffu && !pier_host[pier_of[i]] ? &piers[pier_of[i]] : nullptr;
        if (cfg && !surrogate_finish(p, *cfg, std::vector<float4>(P, P + np[i]), err, dev_piers)) ok[i] = 0;
    });
    for (uint32_t i = 0; i < n_polys; i++)
        if (!ok[i]) return fail(JG_ERR_SURROGATE, "polygon " + std::to_string(i) + ": " + errs[i]);
    for (uint32_t i = 0; i < n_polys; i++) {
        auto h = std::make_unique<jg_spolygon>();
        h->poly = std::move(polys[i]);
        out[i] = h.release();
    }
    if (n_host_completed) *n_host_completed = n_host;
    return JG_OK;
}
