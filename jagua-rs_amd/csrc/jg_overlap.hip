// Pole-pair overlap quantity and its gradient with respect to every item's (x, y, theta).
//
// Forward:  one wavefront per candidate item pair (both items' rotated poles staged in LDS,
//           the pole pairs spread over the lanes) and one per item for the container term
//           (each pole's signed distance to the container boundary), in one launch; each wave
//           writes one partial, and one workgroup of a second launch sums them in a fixed
//           order (f64 accumulation) into Q.  Penetrations are those of Circle::separation_distance
//           (circle.rs:184-200, (r1 + r2) - d) and of the circle-edge distance
//           (circle.rs:216-229).
// Backward: the same waves recompute the pole transforms, accumulate dQ/dc per pole pair and
//           chain it to (x, y, theta); per-pair and per-item partial gradients are gathered
//           per item through a CSR adjacency in a fixed order by the second launch.
// Both:     one pass computes Q's partials and the gradient's terms together (two launches
//           in all); Q and the gradient are bit for bit those of the separate passes.
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

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_f(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROW_MASK, 0xf, false));
}

// Sum of v over the 64 lanes of a wave with every lane active, in a fixed order (the same
// inputs give the same bits): lane pairs, quads, half rows and rows of 16 by DPP, then the
// four row sums by the row broadcasts of lanes 15 and 31 into lane 63 (GFX9 DPP); returned
// wave-uniform.
__device__ __forceinline__ float wsum(float v) {
    v += dpp_f<0xb1, 0xf>(v);    // quad_perm [1, 0, 3, 2]
    v += dpp_f<0x4e, 0xf>(v);    // quad_perm [2, 3, 0, 1]
    v += dpp_f<0x141, 0xf>(v);   // row_half_mirror
    v += dpp_f<0x140, 0xf>(v);   // row_mirror
    v += dpp_f<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp_f<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Rotated local pole offset rp = R(theta) p.  All differences are taken in the items'
// translation frames: c_i - c_j = (t_a - t_b) + (rp_i - rp_j), where t_a - t_b is exact for
// nearby items (Sterbenz) and the rotated offsets are small, so no precision is lost to the
// magnitude of world coordinates.
__device__ __forceinline__ float2 rot(float s, float c, float4 p) {
    return make_float2(c * p.x - s * p.y, s * p.x + c * p.y);
}

// Signed distance from point (px, py) to the container boundary (the n edges, translated by
// -t): the distance to the nearest edge, positive when the point is inside by ray-crossing
// parity, and the offset from that edge's closest point to the point.  Edge e is ce[2e] =
// (x1, y1, x2, y2) and ce[2e + 1] = (cx, cy, 1 / |c|^2, cx / cy) with c = (x2 - x1, y2 - y1)
// (0 for the last two when |c| or cy is 0).  The closest point is chosen as in edge.rs:63-87
// (clamped projection, parameter -1 for a zero-length edge), and in the interior case the
// offset is formed as the component perpendicular to the edge, ((a cy - b cx) / |c|^2)
// (cy, -cx), which is the same vector without the cancellation of subtracting a far-away
// projected point.  Squared distances are compared; one square root gives the nearest.
__device__ __forceinline__ float container_distance(const float4* ce, uint32_t n, float tx, float ty, float px,
                                                    float py, float2& off) {
    float best_sq = INFINITY;
    off = make_float2(0.0f, 0.0f);
    uint32_t crossings = 0;
    for (uint32_t e = 0; e < n; e++) {
        const float4 ew = ce[2 * e], ek = ce[2 * e + 1];
        const float x1 = ew.x - tx, y1 = ew.y - ty, x2 = ew.z - tx, y2 = ew.w - ty;
        const float a = px - x1, b = py - y1;
        const float t = ek.z != 0.0f ? (a * ek.x + b * ek.y) * ek.z : -1.0f;
        float2 o;
        if (t < 0.0f) {
            o = make_float2(a, b);
        } else if (t > 1.0f) {
            o = make_float2(px - x2, py - y2);
        } else {
            const float w = (a * ek.y - b * ek.x) * ek.z;
            o = make_float2(w * ek.y, -(w * ek.x));
        }
        const float dsq = o.x * o.x + o.y * o.y;
        if (dsq < best_sq) {
            best_sq = dsq;
            off = o;
        }
        if ((y1 > py) != (y2 > py) && px < x1 + (py - y1) * ek.w) crossings++;
    }
    const float best = sqrtf(best_sq);
    return (crossings & 1u) ? best : -best;
}

struct OvView {
    const float4* poles;       // px, py, r, 0
    const uint32_t* off;       // n_items + 1
    const uint2* pairs;
    const uint2* pos;          // per pair: the positions of its (a, b) terms in the items' adjacency order
    const float4* cedges;      // two float4 per container edge (container_distance)
    uint32_t n_items, n_cedges;
    uint64_t n_pairs;
};

}  // namespace

// ---- pair and container terms -------------------------------------------------------
namespace {

// what a pass computes: Q's partials, the gradient's terms, or both in one pass
constexpr int OV_Q = 1, OV_GRAD = 2;

// waves per workgroup of a pass; each wave evaluates its own pair or item in its own LDS
constexpr uint32_t OV_WAVES = 4;

// Orders a wave's LDS stores before its lanes' later LDS loads (and the loads before later
// stores): one wave's LDS operations are performed in order, so this only has to keep the
// compiler from moving them across.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// dQ/d(x, y, theta) of a pair's two items
struct PairGrad {
    float xa = 0.0f, ya = 0.0f, ta = 0.0f, xb = 0.0f, yb = 0.0f, tb = 0.0f;
};

// One wave evaluates one candidate pair.  The rotated poles of item a and item b are staged
// in LDS in chunks of 64 (la: rp_i; lb: rp_j - (t_a - t_b), so that c_i - c_j = rp_i - lb_j),
// each chunk padded with NaN poles (which overlap nothing) to a multiple of 8.  The lanes form
// an 8 x 8 grid: lane 8u + w takes the pole pairs (u + 8s, w + 8t) of the chunk pair, so the
// loops are the same for every lane; a lane holds its (at most 8) poles of b in registers and
// reads one pole of a from LDS per 8 pole pairs.  Pole pairs that overlap are rare, so their
// terms sit behind a wave-uniform test (a ballot) that skips them for the whole wave when no
// lane has one.  The two items' sin/cos come from one call: lane 0 takes theta_a, the other
// lanes theta_b.
template <int MODE>
__device__ __forceinline__ void pair_term(const OvView& v, const float* __restrict__ pose3, uint64_t pid,
                                          uint32_t lane, float4* la, float4* lb, float& acc, PairGrad& g) {
    const uint2 pr = v.pairs[pid];
    const float3 pa = make_float3(pose3[3 * pr.x], pose3[3 * pr.x + 1], pose3[3 * pr.x + 2]);
    const float3 pb = make_float3(pose3[3 * pr.y], pose3[3 * pr.y + 1], pose3[3 * pr.y + 2]);
    const float dtx = pa.x - pb.x, dty = pa.y - pb.y;
    float sn, cs;
    sincosf(lane == 0 ? pa.z : pb.z, &sn, &cs);
    const float sa = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sn), 0));
    const float ca = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cs), 0));
    const float sb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sn), 1));
    const float cb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cs), 1));
    const uint32_t a0 = v.off[pr.x], na = v.off[pr.x + 1] - a0;
    const uint32_t b0 = v.off[pr.y], nb = v.off[pr.y + 1] - b0;
    const float4 pad = make_float4(NAN, NAN, NAN, 0.0f);
    const uint32_t u = lane >> 3, w = lane & 7u;
    for (uint32_t ia = 0; ia < na; ia += 64) {
        const uint32_t ta = min(64u, na - ia), sa8 = (ta + 7u) >> 3;
        for (uint32_t jb = 0; jb < nb; jb += 64) {
            const uint32_t tb = min(64u, nb - jb), sb8 = (tb + 7u) >> 3;
            if (jb == 0 && lane < 8u * sa8) {
                float4 e = pad;
                if (lane < ta) {
                    const float4 p = v.poles[a0 + ia + lane];
                    const float2 rp = rot(sa, ca, p);
                    e = make_float4(rp.x, rp.y, p.z, 0.0f);
                }
                la[lane] = e;
            }
            if (lane < 8u * sb8) {
                float4 e = pad;
                if (lane < tb) {
                    const float4 p = v.poles[b0 + jb + lane];
                    const float2 rp = rot(sb, cb, p);
                    e = make_float4(rp.x - dtx, rp.y - dty, p.z, 0.0f);
                }
                lb[lane] = e;
            }
            wave_lds_sync();
            float4 Bv[8];   // the lane's poles of b, w + 8 sj for sj < sb8 (the rest unused)
#pragma unroll
            for (uint32_t sj = 0; sj < 8; sj++) Bv[sj] = lb[w + 8u * sj];
            for (uint32_t si = 0; si < sa8; si++) {
                const float4 A = la[u + 8u * si];
#pragma unroll
                for (uint32_t sj = 0; sj < 8; sj++) {
                    if (sj >= sb8) break;
                    const float4 B = Bv[sj];
                    const float dx = A.x - B.x, dy = A.y - B.y;
                    const float sq = dx * dx + dy * dy;
                    const float rs = A.z + B.z;
                    const bool hit = sq < rs * rs;
                    if (__builtin_amdgcn_ballot_w64(hit) != 0u) {
                        // keeps the block below a branch taken by the whole wave rather than
                        // selects evaluated on every iteration
                        __asm__ volatile("");
                        if (hit) {
                            const float d = sqrtf(sq);
                            if constexpr ((MODE & OV_Q) != 0) {
                                const float pen = rs - d;
                                acc += pen * pen;
                            }
                            if constexpr ((MODE & OV_GRAD) != 0) {
                                if (sq > 0.0f) {
                                    const float k = -2.0f * (rs - d) / d;   // dQ/dc_i = k (c_i - c_j) = -dQ/dc_j
                                    const float gx = k * dx, gy = k * dy;
                                    const float rjx = B.x + dtx, rjy = B.y + dty;
                                    g.xa += gx;
                                    g.ya += gy;
                                    g.ta += -gx * A.y + gy * A.x;      // dc_i/dtheta_a = (-rp_i.y, rp_i.x)
                                    g.xb -= gx;
                                    g.yb -= gy;
                                    g.tb += gx * rjy - gy * rjx;      // dc_j/dtheta_b = (-rp_j.y, rp_j.x)
                                }
                            }
                        }
                    }
                }
            }
            wave_lds_sync();
        }
    }
}

// the container term of one item: per pole, (r - D)^2 when r > D, D its signed distance to
// the container boundary; lanes take the item's poles
template <int MODE>
__device__ __forceinline__ void container_term(const OvView& v, const float* __restrict__ pose3, uint32_t item,
                                               uint32_t lane, float& acc, float& gx, float& gy, float& gt) {
    const float3 pa = make_float3(pose3[3 * item], pose3[3 * item + 1], pose3[3 * item + 2]);
    float s, c;
    sincosf(pa.z, &s, &c);
    const uint32_t a0 = v.off[item], na = v.off[item + 1] - a0;
    if (v.n_cedges == 0) return;
    for (uint32_t i = lane; i < na; i += 64) {
        const float4 p = v.poles[a0 + i];
        const float2 rp = rot(s, c, p);
        float2 o;
        const float D = container_distance(v.cedges, v.n_cedges, pa.x, pa.y, rp.x, rp.y, o);
        const float pen = p.z - D;
        if constexpr ((MODE & OV_Q) != 0) {
            if (pen > 0.0f) acc += pen * pen;
        }
        if constexpr ((MODE & OV_GRAD) != 0) {
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
    }
}

}  // namespace

// ---- the pass over pairs and items --------------------------------------------------
// Waves 0 .. n_pairs-1 evaluate the pairs, waves n_pairs .. n_pairs+n_items-1 the items'
// container terms.  With OV_Q each wave writes one partial of Q (partial); with OV_GRAD a pair
// wave writes its partial derivatives (dQ/dx, dQ/dy, dQ/dtheta) for item a and for item b at
// the pair's two positions in the items' adjacency order (pg), an item wave its container
// term's three (cg3).
template <int MODE>
__global__ __launch_bounds__(64 * OV_WAVES) void k_ov_pass(OvView v, const float* __restrict__ pose3,
                                                           float* __restrict__ partial, float4* __restrict__ pg,
                                                           float* __restrict__ cg3) {
    __shared__ float4 lds[OV_WAVES][2][64];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint64_t w = (uint64_t)blockIdx.x * OV_WAVES + wv;
    if (w >= v.n_pairs + v.n_items) return;   // the last workgroup's spare waves (no barrier follows)
    float4* la = lds[wv][0];
    float4* lb = lds[wv][1];
    float acc = 0.0f;
    if (w < v.n_pairs) {
        PairGrad g;
        pair_term<MODE>(v, pose3, w, lane, la, lb, acc, g);
        if constexpr ((MODE & OV_GRAD) != 0) {
            g.xa = wsum(g.xa); g.ya = wsum(g.ya); g.ta = wsum(g.ta);
            g.xb = wsum(g.xb); g.yb = wsum(g.yb); g.tb = wsum(g.tb);
            if (lane == 0) {
                const uint2 at = v.pos[w];
                pg[at.x] = make_float4(g.xa, g.ya, g.ta, 0.0f);
                pg[at.y] = make_float4(g.xb, g.yb, g.tb, 0.0f);
            }
        }
    } else {
        const uint32_t item = (uint32_t)(w - v.n_pairs);
        float gx = 0.0f, gy = 0.0f, gt = 0.0f;
        container_term<MODE>(v, pose3, item, lane, acc, gx, gy, gt);
        if constexpr ((MODE & OV_GRAD) != 0) {
            gx = wsum(gx); gy = wsum(gy); gt = wsum(gt);
            if (lane == 0) {
                cg3[3 * item] = gx; cg3[3 * item + 1] = gy; cg3[3 * item + 2] = gt;
            }
        }
    }
    if constexpr ((MODE & OV_Q) != 0) {
        acc = wsum(acc);
        if (lane == 0) partial[w] = acc;
    }
}

// The results of a pass.  With OV_GRAD, blocks 0 .. ceil(n_items / 1024)-1 gather each item's
// gradient in a fixed order: its container term, then its pair terms in adjacency order, which
// are contiguous in pg (f64 accumulation).  With OV_Q, the last block sums the n partials into
// Q in a fixed order in f64 (thread t takes t, t + 1024, ..., then a fixed butterfly per wave
// and the 16 wave sums in order).
template <int MODE>
__global__ __launch_bounds__(1024) void k_ov_finish(uint32_t n_items, const uint32_t* __restrict__ adj_off,
                                                    const float4* __restrict__ pg, const float* __restrict__ cg3,
                                                    float* __restrict__ grad, const float* __restrict__ partial,
                                                    uint64_t n, float* __restrict__ q) {
    if ((MODE & OV_Q) == 0 || blockIdx.x + 1u < gridDim.x) {
        const uint32_t item = blockIdx.x * 1024u + threadIdx.x;
        if (item >= n_items) return;
        double gx = cg3[3 * item], gy = cg3[3 * item + 1], gt = cg3[3 * item + 2];
        for (uint32_t k = adj_off[item]; k < adj_off[item + 1]; k++) {
            const float4 g = pg[k];
            gx += g.x; gy += g.y; gt += g.z;
        }
        grad[3 * item] = (float)gx;
        grad[3 * item + 1] = (float)gy;
        grad[3 * item + 2] = (float)gt;
        return;
    }
    __shared__ double sh[16];
    double acc = 0.0;
    for (uint64_t i = threadIdx.x; i < n; i += 4096) {   // four loads in flight, added in index order
        float e[4];
        for (int u = 0; u < 4; u++) e[u] = i + 1024u * u < n ? partial[i + 1024u * u] : 0.0f;
        for (int u = 0; u < 4; u++) acc += (double)e[u];
    }
    // a fixed butterfly within each wave, then the 16 wave sums in order by the first wave
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63u) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < 16; w++) t += sh[w];
        q[0] = (float)t;
    }
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
                                                 uint32_t* __restrict__ adj_off, uint2* __restrict__ pos) {
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
    // item i's terms: the pairs where it is b (by pair index), then those where it is a
    uint32_t at = off_a[i] + off_b[i];
    for (uint32_t k = b0; k < b1; k++) pos[adj_b[k]].y = at++;
    for (uint32_t p = off_a[i]; p < off_a[i + 1]; p++) pos[p].x = at++;
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
};

namespace {

size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

// One wave of 64 lanes per pair and per item in a launch: the grid's work-items must fit 32 bits.
constexpr uint64_t OV_MAX_WAVES = 1ull << 26;

struct Scratch {
    float* partial;    // n_pairs + n_items
    float4* pg;        // 2 n_pairs, in the items' adjacency order
    float* cg3;        // n_items * 3
};
size_t scratch_layout(uint64_t n_pairs, uint32_t n_items, uint8_t* b, Scratch* s) {
    size_t o = 0;
    if (s) s->partial = (float*)(b + o);
    o += al(4 * (n_pairs + n_items) + 4);
    if (s) s->pg = (float4*)(b + o);
    o += al(32 * n_pairs + 32);
    if (s) s->cg3 = (float*)(b + o);
    o += al(12 * (size_t)n_items + 12);
    return o;
}
Scratch scratch_of(jg_overlap* ov) {
    Scratch s;
    scratch_layout(ov->n_pairs, ov->n_items, (uint8_t*)ov->scratch.ptr, &s);
    return s;
}

// Scratch for n_pairs pairs; called before the pairs change.
jg_status ov_reserve_scratch(jg_overlap* ov, uint64_t n_pairs) {
    if (n_pairs + ov->n_items >= OV_MAX_WAVES) return fail(JG_ERR_CAPACITY, "pairs plus items above 2^26");
    return ov->scratch.reserve(scratch_layout(n_pairs, ov->n_items, nullptr, nullptr));
}

jg_status ov_upload_pairs(jg_overlap* ov, const uint32_t* pairs, uint64_t n_pairs) {
    for (uint64_t i = 0; i < n_pairs; i++)
        if (pairs[2 * i] >= ov->n_items || pairs[2 * i + 1] >= ov->n_items || pairs[2 * i] == pairs[2 * i + 1])
            return fail(JG_ERR_INVALID_ARG, "pair index out of range or self pair");
    jg_status s = ov_reserve_scratch(ov, n_pairs);
    if (s != JG_OK) return s;
    // adjacency: per item its pair terms in pair order; pos gives each pair's two positions
    std::vector<uint32_t> off(ov->n_items + 1, 0), pos(2 * n_pairs);
    for (uint64_t i = 0; i < n_pairs; i++) { off[pairs[2 * i] + 1]++; off[pairs[2 * i + 1] + 1]++; }
    for (uint32_t i = 0; i < ov->n_items; i++) off[i + 1] += off[i];
    std::vector<uint32_t> fill(off.begin(), off.end() - 1);
    for (uint64_t i = 0; i < n_pairs; i++) {
        pos[2 * i] = fill[pairs[2 * i]]++;
        pos[2 * i + 1] = fill[pairs[2 * i + 1]]++;
    }
    const size_t o_pairs = 0, o_off = al(8 * n_pairs + 8), o_pos = o_off + al(4 * off.size());
    const size_t total = o_pos + al(4 * pos.size() + 4);
    std::vector<uint8_t> host(total, 0);
    std::memcpy(host.data() + o_pairs, pairs, 8 * n_pairs);
    std::memcpy(host.data() + o_off, off.data(), 4 * off.size());
    std::memcpy(host.data() + o_pos, pos.data(), 4 * pos.size());
    s = ov->pairs.reserve(total);
    if (s != JG_OK) return s;
    s = hip_check(hipMemcpy(ov->pairs.ptr, host.data(), total, hipMemcpyHostToDevice), "hipMemcpy pairs");
    if (s != JG_OK) return s;
    uint8_t* b = (uint8_t*)ov->pairs.ptr;
    ov->view.pairs = (const uint2*)(b + o_pairs);
    ov->adj_off = (const uint32_t*)(b + o_off);
    ov->view.pos = (const uint2*)(b + o_pos);
    ov->n_pairs = n_pairs;
    ov->view.n_pairs = n_pairs;
    return JG_OK;
}

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
    const size_t bytes = o_e + al(32 * (size_t)n_cedges + 32);
    std::vector<uint8_t> host(bytes, 0);
    float4* p4 = (float4*)(host.data() + o_p);
    for (uint32_t i = 0; i < total; i++) p4[i] = make_float4(poles[3 * i], poles[3 * i + 1], poles[3 * i + 2], 0.0f);
    std::memcpy(host.data() + o_off, pole_off, 4 * ((size_t)n_items + 1));
    float4* e4 = (float4*)(host.data() + o_e);
    for (uint32_t e = 0; e < n_cedges; e++) {
        const float* c = cedges + 4 * e;
        const float cx = c[2] - c[0], cy = c[3] - c[1], len_sq = cx * cx + cy * cy;
        e4[2 * e] = make_float4(c[0], c[1], c[2], c[3]);
        e4[2 * e + 1] = make_float4(cx, cy, len_sq != 0.0f ? 1.0f / len_sq : 0.0f, cy != 0.0f ? cx / cy : 0.0f);
    }
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
    if (s == JG_OK) s = ov_reserve_scratch(ov, total);
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
    uint2* pos = (uint2*)(b + o_adj);
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
                       (const uint32_t*)ob, adj_b, adj_off, pos);
    s = hip_check(hipGetLastError(), "pair launches");
    if (s == JG_OK) s = hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (s != JG_OK) return s;
    ov->view.pairs = pairs;
    ov->adj_off = adj_off;
    ov->view.pos = pos;
    ov->n_pairs = total;
    ov->view.n_pairs = total;
    if (n_pairs_out) *n_pairs_out = total;
    return JG_OK;
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

namespace {
// One pass (k_ov_pass) and its finish (k_ov_finish) on `stream`: Q into d_q with OV_Q, the
// gradient into d_grad with OV_GRAD.
template <int MODE>
jg_status ov_launch(jg_overlap* ov, const float* d_pose, float* d_q, float* d_grad, hipStream_t st) {
    Scratch s = scratch_of(ov);
    const uint64_t waves = ov->n_pairs + ov->n_items;
    if (waves)
        hipLaunchKernelGGL(k_ov_pass<MODE>, dim3((uint32_t)((waves + OV_WAVES - 1) / OV_WAVES)), dim3(64 * OV_WAVES), 0,
                           st, ov->view, d_pose, s.partial, s.pg, s.cg3);
    const uint32_t gather_blocks = (MODE & OV_GRAD) != 0 ? (ov->n_items + 1023u) / 1024u : 0u;
    const uint32_t blocks = gather_blocks + ((MODE & OV_Q) != 0 ? 1u : 0u);
    if (blocks)
        hipLaunchKernelGGL(k_ov_finish<MODE>, dim3(blocks), dim3(1024), 0, st, ov->n_items, ov->adj_off,
                           (const float4*)s.pg, s.cg3, d_grad, (const float*)s.partial, waves, d_q);
    return hip_check(hipGetLastError(), "overlap launch");
}
}  // namespace

extern "C" jg_status jg_overlap_fwd(jg_overlap* ov, const float* d_pose, float* d_q, void* stream) {
    if (!ov || !d_q || (ov->n_items && !d_pose)) return fail(JG_ERR_INVALID_ARG, "null argument");
    return ov_launch<OV_Q>(ov, d_pose, d_q, nullptr, (hipStream_t)stream);
}

extern "C" jg_status jg_overlap_bwd(jg_overlap* ov, const float* d_pose, float* d_grad, void* stream) {
    if (!ov || (ov->n_items && (!d_pose || !d_grad))) return fail(JG_ERR_INVALID_ARG, "null argument");
    return ov_launch<OV_GRAD>(ov, d_pose, nullptr, d_grad, (hipStream_t)stream);
}

extern "C" jg_status jg_overlap_fwd_bwd(jg_overlap* ov, const float* d_pose, float* d_q, float* d_grad, void* stream) {
    if (!ov || !d_q || (ov->n_items && (!d_pose || !d_grad))) return fail(JG_ERR_INVALID_ARG, "null argument");
    return ov_launch<OV_Q | OV_GRAD>(ov, d_pose, d_q, d_grad, (hipStream_t)stream);
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
