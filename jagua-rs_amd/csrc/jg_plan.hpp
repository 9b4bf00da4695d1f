// Records of a batched plan's device table and the work-claiming helpers shared by the batched
// detection (jg_batched.hip) and the batched hazard collection (jg_collect.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/jagua_amd.h"
#include "jg_device.hpp"

namespace jg {

struct SetDesc {
    CdeView cde;
    ItemView item;
    FilterView filter;
    uint64_t first;
    uint32_t flags;
    uint32_t lanes;        // lanes per query in the polygon stage (power of two, 1..64)
    const uint32_t* seed;  // collection: slots every query's collector starts with (ascending)
    uint32_t n_seed;
    uint32_t pad;
};

struct Unit {              // 16 B
    uint32_t set;
    uint32_t q0;           // first query, relative to the set
    uint32_t count;        // <= 64
    uint32_t pad;
};

struct Group {             // 16 B: surviving queries q0 + i for the set bits i of mask
    uint32_t set;
    uint32_t q0;
    uint64_t mask;
};

// Work is claimed from 8 streams, each with its own counters on their own 128-byte lines:
// thousands of waves taking work with atomics on a single address queue behind each other
// (measured: 1.81 ms against 1.30 ms per configuration-2 step for the survivor groups alone).
//   unit stream x:  units x + 8 j; counter C_USTREAM0 + 32 x counts the units taken
//   group stream x: the survivor groups of the units of unit stream x, in their own region of
//                   the group array (base gbase[x]); C_GCOUNT0 + 32 x counts the groups
//                   appended, C_STREAM0 + 32 x the groups taken
// A wave starts on stream blockIdx.x mod 8 and moves to the next one as each runs dry.
// C_OVERFLOW counts the queries the batched collection hands to the per-query collector;
// C_DONE counts the waves of a launch that have finished (the last one clears counters).
// Every launch sequence leaves all counters at zero for the next, and a new table starts so.
constexpr uint32_t N_STREAMS = 8, STREAM_STRIDE = 32;
constexpr uint32_t C_STREAM0 = 0;
constexpr uint32_t C_USTREAM0 = C_STREAM0 + STREAM_STRIDE * N_STREAMS;
constexpr uint32_t C_GCOUNT0 = C_USTREAM0 + STREAM_STRIDE * N_STREAMS;
constexpr uint32_t C_OVERFLOW = C_GCOUNT0 + STREAM_STRIDE * N_STREAMS;
constexpr uint32_t C_DONE = C_OVERFLOW + STREAM_STRIDE;
constexpr uint32_t N_COUNTER_WORDS = C_DONE + STREAM_STRIDE;

// The calling wave's exit from a launch of gridDim.x one-wave workgroups: true in the last
// wave to leave, once every other wave has left (C_DONE is then back at zero).
__device__ __forceinline__ bool last_wave_out(uint32_t* counters, uint32_t lane) {
    uint32_t last = 0;
    if (lane == 0) {
        last = atomicAdd(&counters[C_DONE], 1u) == gridDim.x - 1u ? 1u : 0u;
        if (last) counters[C_DONE] = 0u;
    }
    return __builtin_amdgcn_readfirstlane(last) != 0u;
}

constexpr uint32_t WAVE = 64;

// The device table of a batched launch, in one buffer: set descriptors, work units, the claim
// counters, the survivor-group region bases of the 8 streams and the survivor groups.
struct TableLayout {
    size_t o_units = 0, o_ctr = 0, o_gbase = 0, o_grp = 0, total = 0;
    uint32_t n_units = 0, grid_s = 0, grid_p = 0;
    uint32_t stages = 2;
    bool filtered = false;            // some set has a filter (kernels with the filter test)
    int kind = KIND_CIRCLE;           // entity kinds of the surrogate stage over all sets
    uint32_t max_slot_words = 0;      // largest engine's slot bitmask, 32-bit words
    uint64_t n_queries = 0;           // over all sets
};

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Lane 0's f32 value, bit for bit, in a scalar register (readfirstlane moves 32-bit integers).
__device__ __forceinline__ float uniform_f32(float x) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

template <typename T>
__device__ __forceinline__ T wave_bcast(T v) {
    // lane 0's value in a scalar register: what is derived from it (work-unit records, the set
    // descriptor and the engine's pointers) stays wave-uniform and is read with scalar loads
    return (T)__builtin_amdgcn_readfirstlane((int)v);
}

// The next item of n from 8 interleaved streams (item x + 8 j in stream x, counters from word
// base on): the wave's stream x first, then the following ones as they drain; n when all are.
__device__ __forceinline__ uint32_t claim_next(uint32_t* counters, uint32_t base, uint32_t& x, uint32_t& tried,
                                               uint32_t n, uint32_t lane) {
    while (tried < N_STREAMS) {
        uint32_t j = 0;
        if (lane == 0) j = atomicAdd(&counters[base + STREAM_STRIDE * x], 1u);
        j = __builtin_amdgcn_readfirstlane(j);
        const uint64_t i = (uint64_t)x + (uint64_t)N_STREAMS * j;
        if (i < n) return (uint32_t)i;
        x = (x + 1u) & (N_STREAMS - 1u);
        tried++;
    }
    return n;
}

// Sum of v over the G lanes of this lane's aligned group (G a power of two <= 64).
__device__ __forceinline__ int group_sum_i32(int v, uint32_t G) {
    for (uint32_t o = G >> 1; o > 0; o >>= 1) v += __shfl_xor(v, (int)o, 64);
    return v;
}

// pip_hazard / pip_query with the G lanes of a group splitting the polygon's edges (lane sub
// takes edges sub, sub + G, ...).  Each edge's crossing count is the sequential loop's, and
// their sum is an integer, so the parity is the same.  The group's lanes call them together
// with the same arguments.
__device__ __forceinline__ bool pip_hazard_group(const CdeView& c, uint32_t slot, float px, float py, uint32_t sub,
                                                 uint32_t G) {
    const float4 hb = ldg4(c.haz_bbox + slot);
    if (!box_has_point(hb, px, py)) return false;
    const float ray_x = hb.z + (hb.z - hb.x);
    const uint2 vr = ldg2u(c.haz_vr + slot);
    const float2* v = c.haz_verts + vr.x;
    int cnt = 0;
    for (uint32_t i = sub; i < vr.y; i += G) {
        const float2 a = ldg2(v + i), b = ldg2(v + (i + 1 == vr.y ? 0 : i + 1));
        cnt += pip_edge(px, py, ray_x, a.x, a.y, b.x, b.y);
    }
    return (group_sum_i32(cnt, G) & 1) == 1;
}
__device__ __forceinline__ bool pip_query_group(const ItemView& it, const Mat& m, float4 qbb, float px, float py,
                                                uint32_t sub, uint32_t G) {
    if (!box_has_point(qbb, px, py)) return false;
    const float ray_x = qbb.z + (qbb.z - qbb.x);
    int cnt = 0;
    for (uint32_t i = sub; i < it.n_v; i += G) {
        const uint32_t k = i + 1 == it.n_v ? 0 : i + 1;
        const float2 vi = ldg2(it.v + i), vk = ldg2(it.v + k);
        const float2 a = xform(m, vi.x, vi.y), b = xform(m, vk.x, vk.y);
        cnt += pip_edge(px, py, ray_x, a.x, a.y, b.x, b.y);
    }
    return (group_sum_i32(cnt, G) & 1) == 1;
}

// The batched hazard collection over a plan's table (jg_collect.hip): verdict-free, one set
// per query.  ovf is a device list of n_queries (set, query) entries for the queries whose
// sets outgrow the batched kernel's collector.
jg_status collect_launch(const TableLayout& L, uint8_t* base, const float* d_m6, uint32_t cap, uint32_t* d_count,
                         uint64_t* d_hash, uint32_t* d_keys, uint64_t* ovf, uint32_t n_cus, hipStream_t stream);

// The per-query collector for the queries the batched collection handed over (jg_api.hip).
void collect_overflow_launch(const SetDesc* sets, uint32_t* counters, const uint64_t* ovf, const float* m6,
                             uint32_t words, uint32_t cap, uint32_t* count, uint64_t* hash, uint32_t* keys,
                             uint32_t grid, hipStream_t stream);

}  // namespace jg
