// lbf's sampled placement search (lbf/src/opt/search.rs:17-106) resident on the GPU for many
// independent (engine, item) searches: per round, k_search_sample draws the samples of every
// search, the batched collision query (jg_batched.hip) evaluates them, k_search_cost prices
// the clear ones, and k_search_select keeps each search's best clear sample and tightens its
// sampling box.  A run is a fixed
// sequence of launches on one stream, optionally captured once in a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/jagua_amd.h"
#include "jg_internal.hpp"

using namespace jg;

namespace {

struct SearchDesc {                 // per search
    float4 cde_box;                 // first sampling box: the CDE bbox (UniformRectSampler::new(cde.bbox()))
    const float2* v;                // item vertices, local frame
    uint32_t n_v, rot_off, n_rot, continuous;
};

struct SearchState {                // per search, carried across rounds
    float4 box;                     // sampling box of the uniform rounds
    float best_cost, best_xmax, best_ymax, best_tx;
    float best_ty;
    uint32_t best_ridx, found, samples;
};

struct RotEntry {                   // one allowed rotation: angle and its libm sin / cos
    float angle, s, c, pad;
};

constexpr uint32_t SAMPLE_THREADS = 256;

__device__ __forceinline__ float unit_float(uint64_t h) {          // [0, 1), 24 bits
    return (float)(uint32_t)(h >> 40) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ float normal_sample(uint64_t h1, uint64_t h2) {   // Box-Muller
    const float u1 = fmaxf(unit_float(h1), 1.0f / 16777216.0f), u2 = unit_float(h2);
    return sqrtf(-2.0f * logf(u1)) * cosf(6.28318530717958647692f * u2);
}

__global__ void k_search_reset(const SearchDesc* __restrict__ d, SearchState* __restrict__ st, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    SearchState s{};
    s.box = d[i].cde_box;
    s.best_cost = INFINITY;
    st[i] = s;
}

// One thread per (search, sample): the DTransformation and its matrix, built exactly as
// DTransformation::compose builds it from the table's libm sin / cos.
__global__ __launch_bounds__(SAMPLE_THREADS) void k_search_sample(
    const SearchDesc* __restrict__ d, const SearchState* __restrict__ st, const RotEntry* __restrict__ rot,
    uint32_t n, uint32_t K, uint32_t round, uint32_t local, float sd_t, float sd_r, uint64_t seed,
    float* __restrict__ m6, uint32_t* __restrict__ ridx_out, float* __restrict__ dt_all) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n * K) return;
    const uint32_t s = (uint32_t)(i / K), j = (uint32_t)(i % K);
    const SearchDesc sd = d[s];
    const SearchState ss = st[s];
    const uint64_t key = mix64(seed ^ mix64(((uint64_t)s << 40) ^ ((uint64_t)round << 20) ^ (uint64_t)j));
    const uint64_t G = 0x9e3779b97f4a7c15ull;
    uint32_t ridx;
    float tx, ty;
    if (!local || !ss.found) {
        // UniformRectSampler::sample (uniform_rect_sampler.rs:31-38)
        ridx = (uint32_t)((mix64(key + G) >> 32) % sd.n_rot);
        const float4 b = ss.box;
        tx = b.x + (b.z - b.x) * unit_float(mix64(key + 2 * G));
        ty = b.y + (b.w - b.y) * unit_float(mix64(key + 3 * G));
    } else {
        // LSSampler::sample (ls_sampler.rs:97-106) around the best placement
        tx = ss.best_tx + sd_t * normal_sample(mix64(key + 2 * G), mix64(key + 3 * G));
        ty = ss.best_ty + sd_t * normal_sample(mix64(key + 4 * G), mix64(key + 5 * G));
        ridx = ss.best_ridx;
        if (sd.continuous) {
            const float step = 6.28318530717958647692f / (float)sd.n_rot;
            const int dk = (int)rintf(normal_sample(mix64(key + 6 * G), mix64(key + 7 * G)) * sd_r / step);
            const int64_t k = ((int64_t)ridx + dk) % (int64_t)sd.n_rot;
            ridx = (uint32_t)(k < 0 ? k + sd.n_rot : k);
        }
    }
    const RotEntry r = rot[sd.rot_off + ridx];
    float* m = m6 + 6 * i;
    m[0] = r.c; m[1] = -r.s; m[2] = tx;
    m[3] = r.s; m[4] = r.c; m[5] = ty;
    ridx_out[i] = ridx;
    float* o = dt_all + 3 * ((uint64_t)round * n * K + i);
    o[0] = r.angle; o[1] = tx; o[2] = ty;
}

// The LBFLoss (loss.rs:29-41) of every clear sample of an active search, one thread per
// sample: x_max * 10 + y_max of its transformed shape's bbox (simple_polygon.rs:122-133,
// 226-253), with x_max and y_max; INFINITY for the others.  Samples of one search are
// consecutive, so a wave's vertex loops mostly share their length.
__global__ __launch_bounds__(SAMPLE_THREADS) void k_search_cost(const SearchDesc* __restrict__ d,
                                                               const SearchState* __restrict__ st, uint32_t n,
                                                               uint32_t K, uint32_t local, const float* __restrict__ m6,
                                                               const uint8_t* __restrict__ status,
                                                               float4* __restrict__ cost) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n * K) return;
    const uint32_t s = (uint32_t)(i / K);
    float4 out = make_float4(INFINITY, 0.0f, 0.0f, 0.0f);
    if ((!local || st[s].found) && status[i] == JG_Q_CLEAR) {
        const Mat m = load_mat(m6, i);
        const float2* v = d[s].v;
        const uint32_t nv = d[s].n_v;
        float xm = -3.40282347e+38f, ym = -3.40282347e+38f;
        for (uint32_t k = 0; k < nv; k++) {
            const float2 vk = ldg2(v + k);
            const float2 p = xform(m, vk.x, vk.y);
            xm = mx(xm, p.x);
            ym = mx(ym, p.y);
        }
        out = make_float4(xm * 10.0f + ym, xm, ym, 0.0f);
    }
    cost[i] = out;
}

// One wavefront per search: the clear sample of lowest LBFLoss (k_search_cost), ties to the
// first, replaces the best when strictly lower (search.rs:53-66, 86-92); uniform rounds then
// tighten the sampling box (LBFLoss::tighten_sample_bbox, loss.rs:33-41).
__global__ __launch_bounds__(64) void k_search_select(const SearchDesc* __restrict__ d, SearchState* __restrict__ st,
                                                    uint32_t n, uint32_t K, uint32_t round, uint32_t local,
                                                    const float* __restrict__ m6, const uint32_t* __restrict__ ridx,
                                                    const uint8_t* __restrict__ status, const float4* __restrict__ cost,
                                                    uint8_t* __restrict__ status_all) {
    const uint32_t s = blockIdx.x, lane = threadIdx.x;
    if (s >= n) return;
    const SearchState ss0 = st[s];
    const bool active = !local || ss0.found;
    float best = INFINITY, bxm = 0.0f, bym = 0.0f;
    uint32_t bj = 0xffffffffu;
    for (uint32_t j = lane; j < K; j += 64) {
        const uint64_t i = (uint64_t)s * K + j;
        status_all[(uint64_t)round * n * K + i] = status[i];
        const float4 c = cost[i];
        if (c.x < best) { best = c.x; bj = j; bxm = c.y; bym = c.z; }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const float oc = __shfl_xor(best, o, 64), ox = __shfl_xor(bxm, o, 64), oy = __shfl_xor(bym, o, 64);
        const uint32_t oj = (uint32_t)__shfl_xor((int)bj, o, 64);
        if (oc < best || (oc == best && oj < bj)) { best = oc; bj = oj; bxm = ox; bym = oy; }
    }
    if (lane == 0) {
        SearchState ss = ss0;
        if (active) ss.samples += K;
        if (active && bj != 0xffffffffu && (!ss.found || best < ss.best_cost)) {
            const uint64_t i = (uint64_t)s * K + bj;
            ss.found = 1u;
            ss.best_cost = best;
            ss.best_xmax = bxm;
            ss.best_ymax = bym;
            ss.best_tx = m6[6 * i + 2];
            ss.best_ty = m6[6 * i + 5];
            ss.best_ridx = ridx[i];
            if (!local) ss.box.z = fminf(ss.box.z, best / 10.0f);
        }
        st[s] = ss;
    }
}

}  // namespace

struct jg_search {
    std::vector<jg_search_item> items;
    std::vector<RotEntry> rot;       // all searches' rotation tables
    std::vector<uint32_t> rot_off;
    jg_search_config cfg{};
    jg_plan* plan = nullptr;
    DevBlob blob;
    size_t o_desc = 0, o_state = 0, o_rot = 0, o_m6 = 0, o_ridx = 0, o_status = 0, o_status_all = 0, o_dt = 0,
           o_cost = 0;
    uint32_t rounds = 0;
    std::vector<uint64_t> generations;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipStream_t g_stream = nullptr;
    ~jg_search() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        jg_plan_free(plan);
    }
};

namespace {

jg_status search_prepare(jg_search* S) {
    jg_status st = plan_prepare(S->plan);              // engines, items, plan table current
    if (st != JG_OK) return st;
    std::vector<uint64_t> gens;                       // engine images and item uploads the buffers refer to
    for (auto& it : S->items) {
        gens.push_back(it.cde->generation);
        gens.push_back(it.item->upload_gen);
    }
    if (gens == S->generations && S->blob.ptr) return JG_OK;
    const uint32_t n = (uint32_t)S->items.size(), K = S->cfg.samples_per_round;
    const uint64_t nk = (uint64_t)n * K, all = nk * S->rounds;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    S->o_desc = 0;
    S->o_state = al(sizeof(SearchDesc) * n);
    S->o_rot = S->o_state + al(sizeof(SearchState) * n);
    S->o_m6 = S->o_rot + al(sizeof(RotEntry) * S->rot.size());
    S->o_ridx = S->o_m6 + al(24 * nk);
    S->o_status = S->o_ridx + al(4 * nk);
    S->o_status_all = S->o_status + al(nk);
    S->o_dt = S->o_status_all + al(all);
    S->o_cost = S->o_dt + al(12 * all);
    const size_t total = S->o_cost + 16 * nk;
    if (S->blob.ptr) {
        st = hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        if (st != JG_OK) return st;
    }
    st = S->blob.reserve(total);
    if (st != JG_OK) return st;
    std::vector<SearchDesc> desc(n);
    for (uint32_t i = 0; i < n; i++) {
        const jg_search_item& it = S->items[i];
        const float4 bb = it.cde->qt.pool[0].bbox;     // CDEngine::bbox (cd_engine.rs)
        desc[i] = SearchDesc{bb, it.item->view.v, it.item->view.n_v, S->rot_off[i], S->rot_off[i + 1] - S->rot_off[i],
                             it.continuous};
    }
    uint8_t* base = (uint8_t*)S->blob.ptr;
    st = hip_check(hipMemcpy(base + S->o_desc, desc.data(), sizeof(SearchDesc) * n, hipMemcpyHostToDevice), "hipMemcpy");
    if (st == JG_OK)
        st = hip_check(hipMemcpy(base + S->o_rot, S->rot.data(), sizeof(RotEntry) * S->rot.size(), hipMemcpyHostToDevice),
                       "hipMemcpy");
    if (st != JG_OK) return st;
    S->generations = gens;
    if (S->gexec) {                                  // a captured graph refers to the old buffers
        (void)hipGraphExecDestroy(S->gexec);
        (void)hipGraphDestroy(S->graph);
        S->gexec = nullptr;
        S->graph = nullptr;
    }
    return JG_OK;
}

jg_status search_enqueue(jg_search* S, hipStream_t stream) {
    const uint32_t n = (uint32_t)S->items.size(), K = S->cfg.samples_per_round;
    uint8_t* base = (uint8_t*)S->blob.ptr;
    const SearchDesc* desc = (const SearchDesc*)(base + S->o_desc);
    SearchState* state = (SearchState*)(base + S->o_state);
    const RotEntry* rot = (const RotEntry*)(base + S->o_rot);
    float* m6 = (float*)(base + S->o_m6);
    uint32_t* ridx = (uint32_t*)(base + S->o_ridx);
    uint8_t* status = base + S->o_status;
    uint8_t* status_all = base + S->o_status_all;
    float* dt = (float*)(base + S->o_dt);
    float4* cost = (float4*)(base + S->o_cost);
    hipLaunchKernelGGL(k_search_reset, dim3((n + 63) / 64), dim3(64), 0, stream, desc, state, n);
    const uint64_t nk = (uint64_t)n * K;
    const uint32_t blocks = (uint32_t)((nk + SAMPLE_THREADS - 1) / SAMPLE_THREADS);
    for (uint32_t r = 0; r < S->rounds; r++) {
        const bool local = r >= S->cfg.uniform_rounds;
        float sd_t = 0.0f, sd_r = 0.0f;
        if (local) {
            // LSSampler::decay_stddev (ls_sampler.rs:80-95): init * (end / init)^progress
            const uint32_t lr = r - S->cfg.uniform_rounds;
            const float pct = S->cfg.local_rounds > 1 ? (float)lr / (float)(S->cfg.local_rounds - 1) : 0.0f;
            sd_t = S->cfg.ls_sd_transl[0] * std::pow(S->cfg.ls_sd_transl[1] / S->cfg.ls_sd_transl[0], pct);
            sd_r = S->cfg.ls_sd_rot[0] * std::pow(S->cfg.ls_sd_rot[1] / S->cfg.ls_sd_rot[0], pct);
        }
        hipLaunchKernelGGL(k_search_sample, dim3(blocks), dim3(SAMPLE_THREADS), 0, stream, desc, (const SearchState*)state,
                           rot, n, K, r, local ? 1u : 0u, sd_t, sd_r, S->cfg.seed, m6, ridx, dt);
        jg_status st = plan_enqueue(S->plan, m6, status, stream);
        if (st != JG_OK) return st;
        hipLaunchKernelGGL(k_search_cost, dim3(blocks), dim3(SAMPLE_THREADS), 0, stream, desc,
                           (const SearchState*)state, n, K, local ? 1u : 0u, (const float*)m6, (const uint8_t*)status,
                           cost);
        hipLaunchKernelGGL(k_search_select, dim3(n), dim3(64), 0, stream, desc, state, n, K, r, local ? 1u : 0u,
                           (const float*)m6, (const uint32_t*)ridx, (const uint8_t*)status, (const float4*)cost,
                           status_all);
    }
    return hip_check(hipGetLastError(), "search launch");
}

}  // namespace

extern "C" jg_status jg_search_create(const jg_search_item* items, uint32_t n_items, const jg_search_config* cfg,
                                      jg_search** out) {
    if (!out || !cfg || (n_items && !items)) return fail(JG_ERR_INVALID_ARG, "null argument");
    if (n_items == 0) return fail(JG_ERR_INVALID_ARG, "no search items");
    if (cfg->samples_per_round == 0 || cfg->samples_per_round > 65536u)
        return fail(JG_ERR_INVALID_ARG, "samples_per_round must be in 1..65536");
    if (cfg->uniform_rounds + cfg->local_rounds == 0 || cfg->uniform_rounds + cfg->local_rounds > 4096u)
        return fail(JG_ERR_INVALID_ARG, "rounds must be in 1..4096");
    if (n_items >= (1u << 20)) return fail(JG_ERR_INVALID_ARG, "too many searches");
    if (cfg->local_rounds &&
        !(cfg->ls_sd_transl[0] > 0.0f && cfg->ls_sd_transl[1] > 0.0f && cfg->ls_sd_rot[0] > 0.0f && cfg->ls_sd_rot[1] > 0.0f))
        return fail(JG_ERR_INVALID_ARG, "local-search deviations must be positive");
    auto S = std::make_unique<jg_search>();
    S->cfg = *cfg;
    S->rounds = cfg->uniform_rounds + cfg->local_rounds;
    S->rot_off.push_back(0);
    std::vector<jg_query_set> sets(n_items);
    for (uint32_t i = 0; i < n_items; i++) {
        const jg_search_item& it = items[i];
        if (!it.cde || !it.item) return fail(JG_ERR_INVALID_ARG, "null engine or item in a search");
        if (it.n_rotations && !it.rotations) return fail(JG_ERR_INVALID_ARG, "null rotation list");
        S->items.push_back(it);
        S->items.back().rotations = nullptr;
        S->items.back().filter.keys = nullptr;
        // rotation table: DTransformation::compose's sin / cos on the host (transformation.rs:131-139)
        std::vector<float> ang;
        if (it.continuous)
            for (uint32_t k = 0; k < JG_SEARCH_CONT_ROTATIONS; k++)
                ang.push_back((float)(6.283185307179586476925 * k / JG_SEARCH_CONT_ROTATIONS));
        else if (it.n_rotations)
            ang.assign(it.rotations, it.rotations + it.n_rotations);
        else
            ang.push_back(0.0f);                      // RotationRange::None
        for (float a : ang) {
            if (a != a) return fail(JG_ERR_INVALID_ARG, "NaN rotation");
            const Mat m = mat_from_dt(a, 0.0f, 0.0f);
            S->rot.push_back(RotEntry{a, m.d, m.a, 0.0f});
        }
        S->rot_off.push_back((uint32_t)S->rot.size());
        sets[i] = jg_query_set{it.cde, it.item, (uint64_t)i * cfg->samples_per_round, cfg->samples_per_round, it.filter};
    }
    jg_plan_config pc{};
    pc.stages = 2;          // rounds are small launches: the survivor queue spreads the polygon stage
    jg_status st = jg_plan_create(sets.data(), n_items, &pc, &S->plan);
    if (st != JG_OK) return st;
    *out = S.release();
    return JG_OK;
}

extern "C" void jg_search_free(jg_search* s) { delete s; }

extern "C" jg_status jg_search_run(jg_search* S, uint32_t graph, void* stream) {
    if (!S) return fail(JG_ERR_INVALID_ARG, "null search");
    jg_status st = search_prepare(S);
    if (st != JG_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    if (!graph) return search_enqueue(S, s);
    if (!S->gexec || S->g_stream != s) {
        if (S->gexec) {
            (void)hipGraphExecDestroy(S->gexec);
            (void)hipGraphDestroy(S->graph);
            S->gexec = nullptr;
            S->graph = nullptr;
        }
        hipStream_t cap = nullptr;
        st = hip_check(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate");
        if (st != JG_OK) return st;
        st = hip_check(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        if (st == JG_OK) {
            const jg_status ls = search_enqueue(S, cap);
            hipGraph_t g = nullptr;
            const jg_status es = hip_check(hipStreamEndCapture(cap, &g), "hipStreamEndCapture");
            st = ls != JG_OK ? ls : es;
            if (st == JG_OK) {
                S->graph = g;
                st = hip_check(hipGraphInstantiate(&S->gexec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
            } else if (g) {
                (void)hipGraphDestroy(g);
            }
        }
        (void)hipStreamDestroy(cap);
        if (st != JG_OK) return st;
        S->g_stream = s;
    }
    return hip_check(hipGraphLaunch(S->gexec, s), "hipGraphLaunch");
}

extern "C" jg_status jg_search_results(jg_search* S, jg_search_result* out) {
    if (!S || !out) return fail(JG_ERR_INVALID_ARG, "null argument");
    if (!S->blob.ptr) return fail(JG_ERR_INVALID_ARG, "the search has not run");
    jg_status st = hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (st != JG_OK) return st;
    const uint32_t n = (uint32_t)S->items.size();
    std::vector<SearchState> ss(n);
    st = hip_check(hipMemcpy(ss.data(), (uint8_t*)S->blob.ptr + S->o_state, sizeof(SearchState) * n, hipMemcpyDeviceToHost),
                   "hipMemcpy");
    if (st != JG_OK) return st;
    for (uint32_t i = 0; i < n; i++) {
        jg_search_result r{};
        r.found = ss[i].found;
        r.samples = ss[i].samples;
        if (ss[i].found) {
            r.rotation = S->rot[S->rot_off[i] + ss[i].best_ridx].angle;
            r.tx = ss[i].best_tx;
            r.ty = ss[i].best_ty;
            r.cost = ss[i].best_cost;
            r.x_max = ss[i].best_xmax;
            r.y_max = ss[i].best_ymax;
        }
        out[i] = r;
    }
    return JG_OK;
}

extern "C" jg_status jg_search_samples(jg_search* S, float* dt, uint8_t* status) {
    if (!S || !dt || !status) return fail(JG_ERR_INVALID_ARG, "null argument");
    if (!S->blob.ptr) return fail(JG_ERR_INVALID_ARG, "the search has not run");
    jg_status st = hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (st != JG_OK) return st;
    const uint64_t all = (uint64_t)S->items.size() * S->cfg.samples_per_round * S->rounds;
    st = hip_check(hipMemcpy(dt, (uint8_t*)S->blob.ptr + S->o_dt, 12 * all, hipMemcpyDeviceToHost), "hipMemcpy");
    if (st == JG_OK)
        st = hip_check(hipMemcpy(status, (uint8_t*)S->blob.ptr + S->o_status_all, all, hipMemcpyDeviceToHost), "hipMemcpy");
    return st;
}
