// Pole-pair overlap quantity and its gradient with respect to every item's (x, y, theta).
//
// Forward:  one wavefront per candidate item pair (lanes = poles of item a, item b's
//           rotated poles staged in LDS) and one per item for the container term (each
//           pole's signed distance to the container boundary); each
//           wave writes one partial; a two-pass fixed-order reduction (f64 accumulation)
//           produces Q.  Penetrations are those of Circle::separation_distance (circle.rs:
//           184-200, (r1 + r2) - d) and of the circle-edge distance (circle.rs:216-229).
// Backward: the same waves recompute the pole transforms, accumulate dQ/dc per pole and
//           chain it to (x, y, theta); per-pair and per-item partial gradients are gathered
//           per item through a CSR adjacency in a fixed order.
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

constexpr int OV_WAVES = 1;   // waves per workgroup: the LDS tile exchange uses the workgroup barrier

__device__ __forceinline__ float wsum(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Rotated local pole offset rp = R(theta) p.  All differences are taken in the items'
// translation frames: c_i - c_j = (t_a - t_b) + (rp_i - rp_j), where t_a - t_b is exact for
// nearby items (Sterbenz) and the rotated offsets are small, so no precision is lost to the
// magnitude of world coordinates.
__device__ __forceinline__ float2 rot(float s, float c, float4 p) {
    return make_float2(c * p.x - s * p.y, s * p.x + c * p.y);
}

// Offset from the closest point of segment e to point (px, py): the closest point is chosen
// as in edge.rs:63-87 (clamped projection), and in the interior case the offset is formed
// as the component perpendicular to the edge, ((a cy - b cx) / |c|^2) (cy, -cx), which is
// the same vector without the cancellation of subtracting a far-away projected point.
__device__ __forceinline__ float2 edge_offset(float4 e, float px, float py) {
    const float a = px - e.x, b = py - e.y, cx = e.z - e.x, cy = e.w - e.y;
    const float len_sq = cx * cx + cy * cy;
    float t = -1.0f;
    if (len_sq != 0.0f) t = (a * cx + b * cy) / len_sq;
    if (t < 0.0f) return make_float2(a, b);
    if (t > 1.0f) return make_float2(px - e.z, py - e.w);
    const float w = (a * cy - b * cx) / len_sq;
    return make_float2(w * cy, -(w * cx));
}

// Signed distance from point (px, py) to the container boundary (the v.n_cedges edges,
// translated by -t): the distance to the nearest edge, positive when the point is inside by
// ray-crossing parity, and the offset from that edge's closest point to the point.
__device__ __forceinline__ float container_distance(const float4* cedges, uint32_t n, float tx, float ty, float px,
                                                    float py, float2& off) {
    float best = INFINITY;
    off = make_float2(0.0f, 0.0f);
    uint32_t crossings = 0;
    for (uint32_t e = 0; e < n; e++) {
        const float4 ew = cedges[e];
        const float4 ed = make_float4(ew.x - tx, ew.y - ty, ew.z - tx, ew.w - ty);
        const float2 o = edge_offset(ed, px, py);
        const float d = sqrtf(o.x * o.x + o.y * o.y);
        if (d < best) {
            best = d;
            off = o;
        }
        if ((ed.y > py) != (ed.w > py) && px < ed.x + (py - ed.y) * (ed.z - ed.x) / (ed.w - ed.y)) crossings++;
    }
    return (crossings & 1u) ? best : -best;
}

struct OvView {
    const float4* poles;       // px, py, r, 0
    const uint32_t* off;       // n_items + 1
    const uint2* pairs;
    const float4* cedges;
    uint32_t n_items, n_cedges;
    uint64_t n_pairs;
};

}  // namespace

// ---- forward -------------------------------------------------------------------------
__global__ __launch_bounds__(64 * OV_WAVES) void k_ov_fwd_pairs(OvView v, const float* __restrict__ pose3,
                                                      float* __restrict__ partial) {
    __shared__ float4 lds[OV_WAVES][64];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint64_t pid = (uint64_t)blockIdx.x * OV_WAVES + wave;
    if (pid >= v.n_pairs) return;   // the workgroup is this one wave: it exits as a whole
    const uint2 pr = v.pairs[pid];
    const float3 pa = make_float3(pose3[3 * pr.x], pose3[3 * pr.x + 1], pose3[3 * pr.x + 2]);
    const float3 pb = make_float3(pose3[3 * pr.y], pose3[3 * pr.y + 1], pose3[3 * pr.y + 2]);
    const float dtx = pa.x - pb.x, dty = pa.y - pb.y;
    float sa, ca, sb, cb;
    sincosf(pa.z, &sa, &ca);
    sincosf(pb.z, &sb, &cb);
    const uint32_t a0 = v.off[pr.x], na = v.off[pr.x + 1] - a0;
    const uint32_t b0 = v.off[pr.y], nb = v.off[pr.y + 1] - b0;
    float acc = 0.0f;
    for (uint32_t jb = 0; jb < nb; jb += 64) {
        const uint32_t jn = min(64u, nb - jb);
        if (lane < jn) {
            const float4 p = v.poles[b0 + jb + lane];
            const float2 rp = rot(sb, cb, p);
            lds[wave][lane] = make_float4(rp.x, rp.y, p.z, 0.0f);
        }
        __syncthreads();
        for (uint32_t i = lane; i < na; i += 64) {
            const float4 p = v.poles[a0 + i];
            const float2 rpi = rot(sa, ca, p);
            for (uint32_t j = 0; j < jn; j++) {
                const float4 q = lds[wave][j];
                const float dx = (rpi.x - q.x) + dtx, dy = (rpi.y - q.y) + dty;
                const float sq = dx * dx + dy * dy;
                const float rs = p.z + q.z;
                if (sq < rs * rs) {
                    const float pen = rs - sqrtf(sq);
                    acc += pen * pen;
                }
            }
        }
        __syncthreads();
    }
    acc = wsum(acc);
    if (lane == 0) partial[pid] = acc;
}

// the container term: per pole, (r - D)^2 when r > D, D its signed distance to the container
__global__ __launch_bounds__(64 * OV_WAVES) void k_ov_fwd_container(OvView v, const float* __restrict__ pose3,
                                                          float* __restrict__ partial) {
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t item = blockIdx.x * OV_WAVES + wave;
    if (item >= v.n_items) return;
    const float3 pa = make_float3(pose3[3 * item], pose3[3 * item + 1], pose3[3 * item + 2]);
    float s, c;
    sincosf(pa.z, &s, &c);
    const uint32_t a0 = v.off[item], na = v.off[item + 1] - a0;
    float acc = 0.0f;
    if (v.n_cedges > 0)
        for (uint32_t i = lane; i < na; i += 64) {
            const float4 p = v.poles[a0 + i];
            const float2 rp = rot(s, c, p);
            float2 o;
            const float D = container_distance(v.cedges, v.n_cedges, pa.x, pa.y, rp.x, rp.y, o);
            const float pen = p.z - D;
            if (pen > 0.0f) acc += pen * pen;
        }
    acc = wsum(acc);
    if (lane == 0) partial[item] = acc;
}

// fixed-order two-pass reduction: each block sums a contiguous slice in f64
__global__ __launch_bounds__(256) void k_ov_reduce1(const float* __restrict__ x, uint64_t n, uint64_t per_block,
                                                    double* __restrict__ out) {
    __shared__ double sh[256];
    const uint64_t lo = (uint64_t)blockIdx.x * per_block;
    const uint64_t hi = min(n, lo + per_block);
    double acc = 0.0;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += 256) acc += (double)x[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_ov_reduce2(const double* __restrict__ x, uint32_t n, float* __restrict__ q) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 256) acc += x[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) q[0] = (float)sh[0];
}

// ---- backward ------------------------------------------------------------------------
// out[pid] = (dQ/dx_a, dQ/dy_a, dQ/dtheta_a, dQ/dx_b, dQ/dy_b, dQ/dtheta_b)
__global__ __launch_bounds__(64 * OV_WAVES) void k_ov_bwd_pairs(OvView v, const float* __restrict__ pose3,
                                                      float* __restrict__ out6) {
    __shared__ float4 lds[OV_WAVES][64];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint64_t pid = (uint64_t)blockIdx.x * OV_WAVES + wave;
    if (pid >= v.n_pairs) return;
    const uint2 pr = v.pairs[pid];
    const float3 pa = make_float3(pose3[3 * pr.x], pose3[3 * pr.x + 1], pose3[3 * pr.x + 2]);
    const float3 pb = make_float3(pose3[3 * pr.y], pose3[3 * pr.y + 1], pose3[3 * pr.y + 2]);
    const float dtx = pa.x - pb.x, dty = pa.y - pb.y;
    float sa, ca, sb, cb;
    sincosf(pa.z, &sa, &ca);
    sincosf(pb.z, &sb, &cb);
    const uint32_t a0 = v.off[pr.x], na = v.off[pr.x + 1] - a0;
    const uint32_t b0 = v.off[pr.y], nb = v.off[pr.y + 1] - b0;
    float gxa = 0, gya = 0, gta = 0, gxb = 0, gyb = 0, gtb = 0;
    for (uint32_t jb = 0; jb < nb; jb += 64) {
        const uint32_t jn = min(64u, nb - jb);
        if (lane < jn) {
            const float4 p = v.poles[b0 + jb + lane];
            const float2 rp = rot(sb, cb, p);
            lds[wave][lane] = make_float4(rp.x, rp.y, p.z, 0.0f);
        }
        __syncthreads();
        for (uint32_t i = lane; i < na; i += 64) {
            const float4 p = v.poles[a0 + i];
            const float2 rpi = rot(sa, ca, p);
            float gcx = 0.0f, gcy = 0.0f;
            for (uint32_t j = 0; j < jn; j++) {
                const float4 q = lds[wave][j];
                const float dx = (rpi.x - q.x) + dtx, dy = (rpi.y - q.y) + dty;
                const float sq = dx * dx + dy * dy;
                const float rs = p.z + q.z;
                if (sq < rs * rs && sq > 0.0f) {
                    const float d = sqrtf(sq);
                    const float k = -2.0f * (rs - d) / d;   // dQ/dc_i = k (c_i - c_j) = -dQ/dc_j
                    const float gx = k * dx, gy = k * dy;
                    gcx += gx;
                    gcy += gy;
                    gxb -= gx;                              // dc_j/dtheta_b = (-rp_j.y, rp_j.x)
                    gyb -= gy;
                    gtb += gx * q.y - gy * q.x;
                }
            }
            gxa += gcx;
            gya += gcy;
            gta += -gcx * rpi.y + gcy * rpi.x;
        }
        __syncthreads();
    }
    gxa = wsum(gxa); gya = wsum(gya); gta = wsum(gta);
    gxb = wsum(gxb); gyb = wsum(gyb); gtb = wsum(gtb);
    if (lane == 0) {
        float* o = out6 + 6 * pid;
        o[0] = gxa; o[1] = gya; o[2] = gta; o[3] = gxb; o[4] = gyb; o[5] = gtb;
    }
}

__global__ __launch_bounds__(64 * OV_WAVES) void k_ov_bwd_container(OvView v, const float* __restrict__ pose3,
                                                          float* __restrict__ out3) {
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t item = blockIdx.x * OV_WAVES + wave;
    if (item >= v.n_items) return;
    const float3 pa = make_float3(pose3[3 * item], pose3[3 * item + 1], pose3[3 * item + 2]);
    float s, c;
    sincosf(pa.z, &s, &c);
    const uint32_t a0 = v.off[item], na = v.off[item + 1] - a0;
    float gx = 0, gy = 0, gt = 0;
    if (v.n_cedges > 0)
        for (uint32_t i = lane; i < na; i += 64) {
            const float4 p = v.poles[a0 + i];
            const float2 rp = rot(s, c, p);
            float2 o;
            const float D = container_distance(v.cedges, v.n_cedges, pa.x, pa.y, rp.x, rp.y, o);
            const float pen = p.z - D;
            const float d = fabsf(D);
            if (pen > 0.0f && d > 0.0f) {
                // dQ/dc = 2 pen (-dD/dc), dD/dc = sign(D) o / d
                const float k = -2.0f * pen * (D > 0.0f ? 1.0f : -1.0f) / d;
                const float gcx = k * o.x, gcy = k * o.y;
                gx += gcx;
                gy += gcy;
                gt += -gcx * rp.y + gcy * rp.x;
            }
        }
    gx = wsum(gx); gy = wsum(gy); gt = wsum(gt);
    if (lane == 0) {
        out3[3 * item] = gx; out3[3 * item + 1] = gy; out3[3 * item + 2] = gt;
    }
}

// per item, in a fixed order: container term, then its pairs in CSR order (f64 accumulation)
__global__ __launch_bounds__(256) void k_ov_bwd_gather(uint32_t n_items, const uint32_t* __restrict__ adj_off,
                                                       const uint32_t* __restrict__ adj, const float* __restrict__ pg6,
                                                       const float* __restrict__ cg3, float* __restrict__ grad) {
    const uint32_t item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= n_items) return;
    double gx = cg3[3 * item], gy = cg3[3 * item + 1], gt = cg3[3 * item + 2];
    for (uint32_t k = adj_off[item]; k < adj_off[item + 1]; k++) {
        const uint32_t code = adj[k];               // pair << 1 | role (0 = a, 1 = b)
        const float* g = pg6 + 6 * (uint64_t)(code >> 1) + 3 * (code & 1u);
        gx += g[0]; gy += g[1]; gt += g[2];
    }
    grad[3 * item] = (float)gx;
    grad[3 * item + 1] = (float)gy;
    grad[3 * item + 2] = (float)gt;
}

// ---- candidate pairs from the engine's quadtree --------------------------------------
// One wave per item: the quadtree nodes whose box meets the box of the item's bounding circle
// are visited breadth first (a frontier in LDS; when a level outgrows it, the lanes walk the
// rest of their subtrees depth first), every hazard slot of the leaves reached is marked in an
// LDS bitmask, and the marked slots that are items k > i with an overlapping bounding circle
// are the item's pairs (i, k), counted (fill = 0) or written in slot order from off_a[i]
// (fill = 1).  The order depends on nothing but the tree, so the pairs are deterministic.
namespace {
constexpr uint32_t OVP_FCAP = 512;

__device__ __forceinline__ bool box_meets(float4 a, float4 b) {
    return a.x <= b.z && a.z >= b.x && a.y <= b.w && a.w >= b.y;
}
}  // namespace

__global__ __launch_bounds__(64) void k_ovp_items(CdeView c, const uint32_t* __restrict__ slot_item,
                                                  const float* __restrict__ pose3, const float* __restrict__ rad,
                                                  uint32_t n_items, uint32_t words, uint32_t fill,
                                                  uint32_t* __restrict__ cnt_a, const uint32_t* __restrict__ off_a,
                                                  uint2* __restrict__ pairs) {
    extern __shared__ uint32_t bits[];
    __shared__ uint32_t fr[2][OVP_FCAP];
    const uint32_t lane = threadIdx.x;
    for (uint32_t w = lane; w < words; w += 64) bits[w] = 0u;
    __syncthreads();
    for (uint32_t i = blockIdx.x; i < n_items; i += gridDim.x) {
        const float x = pose3[3 * i], y = pose3[3 * i + 1], r = rad[i];
        const float4 cb = make_float4(x - r, y - r, x + r, y + r);
        // the leaves' slots into the bitmask
        auto visit = [&](int n, int& fc) -> uint32_t {
            const DevNode nd = c.nodes[n];
            if (nd.first_child < 0) {
                const uint32_t cnt = nd.haz_info & COUNT_MASK;
                for (uint32_t k = 0; k < cnt; k++) {
                    const uint32_t sl = c.ent[nd.haz_begin + k].slot_flags & SLOT_MASK;
                    atomicOr(&bits[sl >> 5], 1u << (sl & 31u));
                }
                return 0u;
            }
            fc = nd.first_child;
            uint32_t m = 0;
            for (int k = 0; k < 4; k++) m |= box_meets(box_quadrant(nd.bbox, k), cb) ? (1u << k) : 0u;
            return m;
        };
        uint32_t nf = box_meets(c.nodes[0].bbox, cb) ? 1u : 0u, cur = 0;
        if (lane == 0) fr[0][0] = 0u;
        __syncthreads();
        while (nf > 0u) {
            uint32_t nb = 0;
            for (uint32_t b0 = 0; b0 < nf; b0 += 64u) {
                const bool mine = b0 + lane < nf;
                int fc = -1;
                uint32_t m = 0;
                int n = 0;
                if (mine) {
                    n = (int)fr[cur][b0 + lane];
                    m = visit(n, fc);
                }
                const uint32_t k = (uint32_t)__builtin_popcount(m);
                const uint32_t incl = wave_incl_scan(k, lane);
                const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
                if (nb + tot <= OVP_FCAP) {
                    uint32_t at = nb + incl - k;
                    for (uint32_t b = m; b; b &= b - 1u) fr[cur ^ 1u][at++] = (uint32_t)(fc + __builtin_ctz(b));
                    nb += tot;
                } else {
                    // the next level outgrows the frontier: this lane walks its node's subtrees
                    for (uint32_t b = m; b; b &= b - 1u) (void)descend(c, fc + __builtin_ctz(b), visit);
                }
            }
            __syncthreads();
            cur ^= 1u;
            nf = nb;
        }
        __syncthreads();
        // the marked slots, in slot order: items k > i whose bounding circles overlap item i's
        uint32_t base = fill ? off_a[i] : 0u;
        uint32_t count = 0;
        for (uint32_t w0 = 0; w0 < words; w0 += 64u) {
            const uint32_t w = w0 + lane;
            const uint32_t word = w < words ? bits[w] : 0u;
            // a neighbour: an item k > i whose bounding circle overlaps item i's
            auto neighbour = [&](uint32_t b, uint32_t& k) {
                k = slot_item[w * 32u + b];
                if (k == 0xffffffffu || k <= i) return false;
                const float dx = pose3[3 * k] - x, dy = pose3[3 * k + 1] - y, rr = rad[k] + r;
                return dx * dx + dy * dy <= rr * rr;
            };
            uint32_t nm = 0;
            for (uint32_t wb = word; wb; wb &= wb - 1u) {
                uint32_t k;
                nm += neighbour((uint32_t)__builtin_ctz(wb), k) ? 1u : 0u;
            }
            const uint32_t incl = wave_incl_scan(nm, lane);
            const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
            if (fill) {
                uint32_t at = base + incl - nm;
                for (uint32_t wb = word; wb; wb &= wb - 1u) {
                    uint32_t k;
                    if (neighbour((uint32_t)__builtin_ctz(wb), k)) pairs[at++] = make_uint2(i, k);
                }
            }
            base += tot;
            count += tot;
            if (w < words) bits[w] = 0u;
        }
        if (!fill && lane == 0) cnt_a[i] = count;
        __syncthreads();
    }
}

// Exclusive prefix sums of in[0..n) into out[0..n], out[n] the total: one workgroup of 1024.
__global__ __launch_bounds__(1024) void k_ovp_scan(const uint32_t* __restrict__ in, uint32_t n,
                                                   uint32_t* __restrict__ out) {
    __shared__ uint32_t sh[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t lo = min(n, t * per), hi = min(n, lo + per);
    uint32_t s = 0;
    for (uint32_t k = lo; k < hi; k++) s += in[k];
    sh[t] = s;
    __syncthreads();
    for (uint32_t o = 1; o < 1024u; o <<= 1) {
        const uint32_t v = t >= o ? sh[t - o] : 0u;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    uint32_t run = sh[t] - s;
    for (uint32_t k = lo; k < hi; k++) {
        out[k] = run;
        run += in[k];
    }
    if (t == 1023u) out[n] = sh[1023];
}

__global__ __launch_bounds__(256) void k_ovp_zero(uint32_t* __restrict__ x, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0u;
}

__global__ __launch_bounds__(256) void k_ovp_count_b(const uint2* __restrict__ pairs, uint32_t n_pairs,
                                                     uint32_t* __restrict__ cnt_b) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n_pairs) atomicAdd(&cnt_b[pairs[p].y], 1u);
}

__global__ __launch_bounds__(256) void k_ovp_fill_b(const uint2* __restrict__ pairs, uint32_t n_pairs,
                                                    uint32_t* __restrict__ cur_b, uint32_t* __restrict__ adj_b) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n_pairs) adj_b[atomicAdd(&cur_b[pairs[p].y], 1u)] = p;
}

// Per item: its adjacency in a fixed order, the pairs where it is b (ascending, sorted here
// from the order the atomics left), then those where it is a.
__global__ __launch_bounds__(256) void k_ovp_adj(uint32_t n_items, const uint32_t* __restrict__ off_a,
                                                 const uint32_t* __restrict__ off_b, uint32_t* __restrict__ adj_b,
                                                 uint32_t* __restrict__ adj_off, uint32_t* __restrict__ adj) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n_items) return;
    adj_off[i] = off_a[i] + off_b[i];
    if (i == n_items) return;
    const uint32_t b0 = off_b[i], b1 = off_b[i + 1];
    for (uint32_t k = b0 + 1; k < b1; k++) {          // insertion sort (a few entries)
        const uint32_t v = adj_b[k];
        uint32_t j = k;
        while (j > b0 && adj_b[j - 1] > v) {
            adj_b[j] = adj_b[j - 1];
            j--;
        }
        adj_b[j] = v;
    }
    uint32_t at = off_a[i] + off_b[i];
    for (uint32_t k = b0; k < b1; k++) adj[at++] = (adj_b[k] << 1) | 1u;
    for (uint32_t p = off_a[i]; p < off_a[i + 1]; p++) adj[at++] = p << 1;
}

// ---------------------------------------------------------------------------------------
struct jg_overlap {
    uint32_t n_items = 0, n_cedges = 0;
    uint64_t n_pairs = 0;
    DevBlob geo;                  // poles, offsets, container edges
    DevBlob pairs;                // pairs + CSR adjacency
    DevBlob scratch;              // partials, per-pair gradients, block sums
    DevBlob work;                 // pair search: slot -> item, radii, counts and offsets
    OvView view{};
    const uint32_t* adj_off = nullptr;
    const uint32_t* adj = nullptr;
};

namespace {

size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

jg_status ov_upload_pairs(jg_overlap* ov, const uint32_t* pairs, uint64_t n_pairs) {
    for (uint64_t i = 0; i < n_pairs; i++)
        if (pairs[2 * i] >= ov->n_items || pairs[2 * i + 1] >= ov->n_items || pairs[2 * i] == pairs[2 * i + 1])
            return fail(JG_ERR_INVALID_ARG, "pair index out of range or self pair");
    if (n_pairs >= (1ull << 31)) return fail(JG_ERR_INVALID_ARG, "too many pairs");
    // CSR adjacency: per item its (pair, role) codes in pair order
    std::vector<uint32_t> off(ov->n_items + 1, 0), adj(2 * n_pairs);
    for (uint64_t i = 0; i < n_pairs; i++) { off[pairs[2 * i] + 1]++; off[pairs[2 * i + 1] + 1]++; }
    for (uint32_t i = 0; i < ov->n_items; i++) off[i + 1] += off[i];
    std::vector<uint32_t> fill(off.begin(), off.end() - 1);
    for (uint64_t i = 0; i < n_pairs; i++) {
        adj[fill[pairs[2 * i]]++] = (uint32_t)(i << 1);
        adj[fill[pairs[2 * i + 1]]++] = (uint32_t)(i << 1) | 1u;
    }
    const size_t o_pairs = 0, o_off = al(8 * n_pairs + 8), o_adj = o_off + al(4 * off.size());
    const size_t total = o_adj + al(4 * adj.size() + 4);
    std::vector<uint8_t> host(total, 0);
    std::memcpy(host.data() + o_pairs, pairs, 8 * n_pairs);
    std::memcpy(host.data() + o_off, off.data(), 4 * off.size());
    std::memcpy(host.data() + o_adj, adj.data(), 4 * adj.size());
    jg_status s = ov->pairs.reserve(total);
    if (s != JG_OK) return s;
    s = hip_check(hipMemcpy(ov->pairs.ptr, host.data(), total, hipMemcpyHostToDevice), "hipMemcpy pairs");
    if (s != JG_OK) return s;
    uint8_t* b = (uint8_t*)ov->pairs.ptr;
    ov->view.pairs = (const uint2*)(b + o_pairs);
    ov->adj_off = (const uint32_t*)(b + o_off);
    ov->adj = (const uint32_t*)(b + o_adj);
    ov->n_pairs = n_pairs;
    ov->view.n_pairs = n_pairs;
    const size_t need = al(4 * (n_pairs + ov->n_items) + 4) + al(24 * n_pairs + 24) + al(12 * ov->n_items + 12) +
                        al(8 * 4096);
    return ov->scratch.reserve(need);
}

struct Scratch {
    float* partial;    // n_pairs + n_items
    float* pg6;        // n_pairs * 6
    float* cg3;        // n_items * 3
    double* blocks;    // up to 4096
};
Scratch scratch_of(jg_overlap* ov) {
    uint8_t* b = (uint8_t*)ov->scratch.ptr;
    Scratch s;
    size_t o = 0;
    s.partial = (float*)(b + o); o += al(4 * (ov->n_pairs + ov->n_items) + 4);
    s.pg6 = (float*)(b + o); o += al(24 * ov->n_pairs + 24);
    s.cg3 = (float*)(b + o); o += al(12 * ov->n_items + 12);
    s.blocks = (double*)(b + o);
    return s;
}

uint32_t blocks_for(uint64_t waves) { return (uint32_t)((waves + OV_WAVES - 1) / OV_WAVES); }

}  // namespace

extern "C" jg_status jg_overlap_create(const float* poles, const uint32_t* pole_off, uint32_t n_items,
                                       const uint32_t* pairs, uint64_t n_pairs, const float* cedges, uint32_t n_cedges,
                                       jg_overlap** out) {
    if (!out || !pole_off || (n_items && !poles) || (n_pairs && !pairs) || (n_cedges && !cedges))
        return fail(JG_ERR_INVALID_ARG, "null argument");
    jg_status s = device_ready();
    if (s != JG_OK) return s;
    for (uint32_t i = 0; i < n_items; i++)
        if (pole_off[i + 1] < pole_off[i]) return fail(JG_ERR_INVALID_ARG, "pole offsets must be non-decreasing");
    auto ov = std::make_unique<jg_overlap>();
    ov->n_items = n_items;
    ov->n_cedges = n_cedges;
    const uint32_t total = pole_off[n_items];
    const size_t o_p = 0, o_off = al(16 * (size_t)total + 16), o_e = o_off + al(4 * ((size_t)n_items + 1));
    const size_t bytes = o_e + al(16 * (size_t)n_cedges + 16);
    std::vector<uint8_t> host(bytes, 0);
    float4* p4 = (float4*)(host.data() + o_p);
    for (uint32_t i = 0; i < total; i++) p4[i] = make_float4(poles[3 * i], poles[3 * i + 1], poles[3 * i + 2], 0.0f);
    std::memcpy(host.data() + o_off, pole_off, 4 * ((size_t)n_items + 1));
    std::memcpy(host.data() + o_e, cedges, 16 * (size_t)n_cedges);
    s = ov->geo.reserve(bytes);
    if (s != JG_OK) return s;
    s = hip_check(hipMemcpy(ov->geo.ptr, host.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy overlap");
    if (s != JG_OK) return s;
    uint8_t* b = (uint8_t*)ov->geo.ptr;
    ov->view.poles = (const float4*)(b + o_p);
    ov->view.off = (const uint32_t*)(b + o_off);
    ov->view.cedges = (const float4*)(b + o_e);
    ov->view.n_items = n_items;
    ov->view.n_cedges = n_cedges;
    s = ov_upload_pairs(ov.get(), pairs, n_pairs);
    if (s != JG_OK) return s;
    *out = ov.release();
    return JG_OK;
}

extern "C" jg_status jg_overlap_set_pairs(jg_overlap* ov, const uint32_t* pairs, uint64_t n_pairs) {
    if (!ov || (n_pairs && !pairs)) return fail(JG_ERR_INVALID_ARG, "null argument");
    jg_status s = hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (s != JG_OK) return s;
    return ov_upload_pairs(ov, pairs, n_pairs);
}

// Candidate pairs from the engine's quadtree (k_ovp_*), then the CSR adjacency of the
// backward gather, all on the device; the pair count is read back once.
extern "C" jg_status jg_overlap_pairs_from_cde(jg_overlap* ov, jg_cde* cde, const uint32_t* item_slots,
                                               const float* d_pose, const float* rad, uint64_t* n_pairs_out,
                                               void* stream) {
    if (!ov || !cde || !d_pose || (ov->n_items && (!item_slots || !rad))) return fail(JG_ERR_INVALID_ARG, "null argument");
    jg_status s = cde_upload(cde);
    if (s != JG_OK) return s;
    const uint32_t n = ov->n_items;
    const uint32_t n_slots = cde->view.n_slots;
    const uint32_t words = (n_slots + 31u) / 32u;
    if ((size_t)words * 4 > 60000) return fail(JG_ERR_CAPACITY, "too many hazard slots for the LDS bitmask");
    std::vector<uint32_t> slot_item(n_slots, 0xffffffffu);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t sl = item_slots[i];
        if (sl >= n_slots || !cde->map.slots[sl].occ) return fail(JG_ERR_INVALID_ARG, "item slot is not a hazard of the engine");
        slot_item[sl] = i;
    }
    hipStream_t st = (hipStream_t)stream;
    // work arrays: slot_item, rad, cnt_a, off_a, cnt_b, off_b, cur_b
    const size_t o_si = 0, o_rad = al(4 * (size_t)n_slots), o_ca = o_rad + al(4 * (size_t)n + 4);
    const size_t o_oa = o_ca + al(4 * (size_t)n + 4), o_cb = o_oa + al(4 * (size_t)n + 4);
    const size_t o_ob = o_cb + al(4 * (size_t)n + 4), o_cu = o_ob + al(4 * (size_t)n + 4);
    const size_t work = o_cu + al(4 * (size_t)n + 4);
    s = hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");   // launches may still read the work buffer
    if (s != JG_OK) return s;
    s = ov->work.reserve(work);
    if (s != JG_OK) return s;
    uint8_t* w = (uint8_t*)ov->work.ptr;
    uint32_t* d_si = (uint32_t*)(w + o_si);
    float* d_rad = (float*)(w + o_rad);
    uint32_t *ca = (uint32_t*)(w + o_ca), *oa = (uint32_t*)(w + o_oa), *cb = (uint32_t*)(w + o_cb);
    uint32_t *ob = (uint32_t*)(w + o_ob), *cu = (uint32_t*)(w + o_cu);
    s = hip_check(hipMemcpy(d_si, slot_item.data(), 4 * (size_t)n_slots, hipMemcpyHostToDevice), "hipMemcpy slots");
    if (s == JG_OK) s = hip_check(hipMemcpy(d_rad, rad, 4 * (size_t)n, hipMemcpyHostToDevice), "hipMemcpy radii");
    if (s != JG_OK) return s;
    const uint32_t grid = std::max<uint32_t>(1u, std::min<uint32_t>(n, 4096u));
    const size_t lds = (size_t)words * 4;
    hipLaunchKernelGGL(k_ovp_items, dim3(grid), dim3(64), lds, st, cde->view, d_si, d_pose, d_rad, n, words, 0u, ca,
                       oa, (uint2*)nullptr);
    hipLaunchKernelGGL(k_ovp_scan, dim3(1), dim3(1024), 0, st, ca, n, oa);
    uint32_t total = 0;
    s = hip_check(hipMemcpyAsync(&total, oa + n, 4, hipMemcpyDeviceToHost, st), "hipMemcpy pair count");
    if (s == JG_OK) s = hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (s != JG_OK) return s;
    // pairs, adjacency offsets, adjacency, the b-role lists
    const size_t o_pairs = 0, o_off = al(8 * (size_t)total + 8), o_adj = o_off + al(4 * ((size_t)n + 1));
    const size_t o_adjb = o_adj + al(8 * (size_t)total + 4);
    const size_t need = o_adjb + al(4 * (size_t)total + 4);
    s = ov->pairs.reserve(need);
    if (s != JG_OK) return s;
    uint8_t* b = (uint8_t*)ov->pairs.ptr;
    uint2* pairs = (uint2*)(b + o_pairs);
    uint32_t* adj_off = (uint32_t*)(b + o_off);
    uint32_t* adj = (uint32_t*)(b + o_adj);
    uint32_t* adj_b = (uint32_t*)(b + o_adjb);
    hipLaunchKernelGGL(k_ovp_items, dim3(grid), dim3(64), lds, st, cde->view, d_si, d_pose, d_rad, n, words, 1u, ca,
                       (const uint32_t*)oa, pairs);
    hipLaunchKernelGGL(k_ovp_zero, dim3((n + 256) / 256), dim3(256), 0, st, cb, n + 1);
    const uint32_t pb = (total + 255u) / 256u;
    if (total) hipLaunchKernelGGL(k_ovp_count_b, dim3(pb), dim3(256), 0, st, (const uint2*)pairs, total, cb);
    hipLaunchKernelGGL(k_ovp_scan, dim3(1), dim3(1024), 0, st, cb, n, ob);
    s = hip_check(hipMemcpyAsync(cu, ob, 4 * (size_t)n, hipMemcpyDeviceToDevice, st), "hipMemcpy cursors");
    if (s != JG_OK) return s;
    if (total) hipLaunchKernelGGL(k_ovp_fill_b, dim3(pb), dim3(256), 0, st, (const uint2*)pairs, total, cu, adj_b);
    hipLaunchKernelGGL(k_ovp_adj, dim3((n + 1 + 255) / 256), dim3(256), 0, st, n, (const uint32_t*)oa,
                       (const uint32_t*)ob, adj_b, adj_off, adj);
    s = hip_check(hipGetLastError(), "pair launches");
    if (s == JG_OK) s = hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (s != JG_OK) return s;
    ov->view.pairs = pairs;
    ov->adj_off = adj_off;
    ov->adj = adj;
    ov->n_pairs = total;
    ov->view.n_pairs = total;
    if (n_pairs_out) *n_pairs_out = total;
    const size_t scr = al(4 * (total + (size_t)n) + 4) + al(24 * (size_t)total + 24) + al(12 * (size_t)n + 12) +
                       al(8 * 4096);
    return ov->scratch.reserve(scr);
}

// The current candidate pairs, copied to the host ([n_pairs][2]).
extern "C" jg_status jg_overlap_get_pairs(const jg_overlap* ov, uint32_t* pairs_out, uint64_t cap, uint64_t* n_out) {
    if (!ov || !n_out) return fail(JG_ERR_INVALID_ARG, "null argument");
    *n_out = ov->n_pairs;
    if (!pairs_out) return JG_OK;
    if (cap < ov->n_pairs) return fail(JG_ERR_CAPACITY, "pair buffer too small");
    if (!ov->n_pairs) return JG_OK;
    return hip_check(hipMemcpy(pairs_out, ov->view.pairs, 8 * ov->n_pairs, hipMemcpyDeviceToHost), "hipMemcpy pairs");
}

extern "C" void jg_overlap_free(jg_overlap* ov) { delete ov; }

extern "C" jg_status jg_overlap_fwd(jg_overlap* ov, const float* d_pose, float* d_q, void* stream) {
    if (!ov || !d_pose || !d_q) return fail(JG_ERR_INVALID_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    Scratch s = scratch_of(ov);
    if (ov->n_pairs)
        hipLaunchKernelGGL(k_ov_fwd_pairs, dim3(blocks_for(ov->n_pairs)), dim3(64 * OV_WAVES), 0, st, ov->view, d_pose,
                           s.partial);
    if (ov->n_items)
        hipLaunchKernelGGL(k_ov_fwd_container, dim3(blocks_for(ov->n_items)), dim3(64 * OV_WAVES), 0, st, ov->view,
                           d_pose, s.partial + ov->n_pairs);
    const uint64_t n = ov->n_pairs + ov->n_items;
    const uint64_t per = std::max<uint64_t>(1024, (n + 4095) / 4096);
    const uint32_t nb = (uint32_t)std::max<uint64_t>(1, (n + per - 1) / per);
    hipLaunchKernelGGL(k_ov_reduce1, dim3(nb), dim3(256), 0, st, s.partial, n, per, s.blocks);
    hipLaunchKernelGGL(k_ov_reduce2, dim3(1), dim3(256), 0, st, s.blocks, nb, d_q);
    return hip_check(hipGetLastError(), "overlap forward launch");
}

extern "C" jg_status jg_overlap_bwd(jg_overlap* ov, const float* d_pose, float* d_grad, void* stream) {
    if (!ov || !d_pose || !d_grad) return fail(JG_ERR_INVALID_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    Scratch s = scratch_of(ov);
    if (ov->n_pairs)
        hipLaunchKernelGGL(k_ov_bwd_pairs, dim3(blocks_for(ov->n_pairs)), dim3(64 * OV_WAVES), 0, st, ov->view, d_pose,
                           s.pg6);
    if (ov->n_items) {
        hipLaunchKernelGGL(k_ov_bwd_container, dim3(blocks_for(ov->n_items)), dim3(64 * OV_WAVES), 0, st, ov->view,
                           d_pose, s.cg3);
        hipLaunchKernelGGL(k_ov_bwd_gather, dim3((ov->n_items + 255) / 256), dim3(256), 0, st, ov->n_items, ov->adj_off,
                           ov->adj, s.pg6, s.cg3, d_grad);
    }
    return hip_check(hipGetLastError(), "overlap backward launch");
}

// uniform grid broad phase on bounding circles (host)
extern "C" jg_status jg_overlap_candidate_pairs(const float* pose, const float* rad, uint32_t n_items,
                                                const float bbox[4], uint32_t depth, uint32_t* pairs_out, uint64_t cap,
                                                uint64_t* n_out) {
    if (!pose || !rad || !bbox || !n_out) return fail(JG_ERR_INVALID_ARG, "null argument");
    if (depth > 12) return fail(JG_ERR_DEPTH, "grid depth above 12");
    const uint32_t side = 1u << depth;
    const float w = (bbox[2] - bbox[0]) / (float)side, h = (bbox[3] - bbox[1]) / (float)side;
    std::vector<std::vector<uint32_t>> cells((size_t)side * side);
    auto cell_range = [&](uint32_t i, int& x0, int& y0, int& x1, int& y1) {
        const float x = pose[3 * i], y = pose[3 * i + 1], r = rad[i];
        auto cl = [&](float v, float lo, float step) {
            int c = (int)std::floor((v - lo) / step);
            return std::max(0, std::min((int)side - 1, c));
        };
        x0 = cl(x - r, bbox[0], w); x1 = cl(x + r, bbox[0], w);
        y0 = cl(y - r, bbox[1], h); y1 = cl(y + r, bbox[1], h);
    };
    for (uint32_t i = 0; i < n_items; i++) {
        int x0, y0, x1, y1;
        cell_range(i, x0, y0, x1, y1);
        for (int yy = y0; yy <= y1; yy++)
            for (int xx = x0; xx <= x1; xx++) cells[(size_t)yy * side + xx].push_back(i);
    }
    uint64_t n = 0;
    std::vector<uint32_t> seen(n_items, 0xffffffffu);
    for (uint32_t i = 0; i < n_items; i++) {
        int x0, y0, x1, y1;
        cell_range(i, x0, y0, x1, y1);
        for (int yy = y0; yy <= y1; yy++)
            for (int xx = x0; xx <= x1; xx++)
                for (uint32_t j : cells[(size_t)yy * side + xx]) {
                    if (j <= i || seen[j] == i) continue;
                    seen[j] = i;
                    const float dx = pose[3 * i] - pose[3 * j], dy = pose[3 * i + 1] - pose[3 * j + 1];
                    const float rr = rad[i] + rad[j];
                    if (dx * dx + dy * dy > rr * rr) continue;
                    if (pairs_out && n < cap) { pairs_out[2 * n] = i; pairs_out[2 * n + 1] = j; }
                    n++;
                }
    }
    *n_out = n;
    return n > cap && pairs_out ? fail(JG_ERR_CAPACITY, "pair buffer too small") : JG_OK;
}
