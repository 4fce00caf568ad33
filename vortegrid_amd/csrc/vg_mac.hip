// vg_mac.hip -- near-field treecode at grid points (accurate mode; this
// build's own method, no reference counterpart).
//
// The reference treecode opens only the cell that contains the query
// (vortonSim.cpp:1819-1856) and is accurate to a few tens of percent.  This
// mode keeps the same NestedGrid cluster pyramid but:
//   * opens every cell within +-near_k cells of the query's brick on every
//     layer (the near box), applying the remaining children as clusters;
//   * corrects each applied cluster with its first moment about the centroid
//     (order 1): u -= s/|d|^3 [a - 3/|d|^2 (D d) x d], D[b][e] = sum Omega_b
//     (y - c)_e, a = antisymmetric part of D;
//   * sums the leaf near box directly over its vortons (cell-sorted lists).
//
// GPU mapping: one wave = one 4x4x4 brick of grid points.  Near boxes are
// per brick, so every loop bound, branch and cluster address is uniform
// across the wave: the tree walk runs in scalar registers and each cluster
// is fetched once per wave (one cache line, broadcast), while the 64 lanes
// evaluate it at their own points.
#include "vg_internal.h"

namespace vg {
vg_status tree_checked_view(const vg_tree *t, TreeView *v);

namespace {

struct MacParams {
    int k;
    float spread;
    float circ;
    uint32_t nb[3];
    uint32_t nbricks;
};

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Cumulative decimation of layer l relative to the leaf layer, per axis.
__device__ __forceinline__ uint32_t layer_scale(const TreeView &tv, int l, int a) {
    uint32_t s = 1;
    for (int j = 1; j <= l; ++j) s *= tv.dec[j][a];
    return s;
}

template <int POLICY>
__device__ __forceinline__ float kernel_law(float d2, float r2, float inv_r3) {
    float law;
    if (POLICY == VG_RSQRT_REFERENCE) {
        law = finvsqrtf(d2) * __builtin_amdgcn_rcpf(d2);
    } else {
        const float rs = __builtin_amdgcn_rsqf(d2);
        law = rs * rs * rs;
    }
    return d2 < r2 ? inv_r3 : law;
}

template <int POLICY, int ORDER, int COUNT>
__global__ __launch_bounds__(256) void k_mac_grid(const TreeView tv, vg_grid_desc g, MacParams mp,
                                                  float *__restrict__ out,
                                                  unsigned long long *__restrict__ counts) {
    const uint32_t wave = uni(threadIdx.x >> 6);
    const uint32_t brick = uni(blockIdx.x * 4 + wave);
    if (brick >= mp.nbricks) return;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t bb[3] = {brick % mp.nb[0], (brick / mp.nb[0]) % mp.nb[1],
                            brick / (mp.nb[0] * mp.nb[1])};
    const uint32_t lp[3] = {lane & 3u, (lane >> 2) & 3u, lane >> 4};
    uint32_t pt[3];
    bool valid = true;
    int cl0[3], cl1[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const uint32_t p = 4 * bb[a] + lp[a];
        valid = valid && p < g.np[a];
        pt[a] = p < g.np[a] ? p : g.np[a] - 1;
        const int nc = (int)g.np[a] - 1;
        const int first = (int)(4 * bb[a]);
        int last = first + 3;
        if (last > (int)g.np[a] - 1) last = (int)g.np[a] - 1;
        cl0[a] = first < nc - 1 ? first : nc - 1;
        cl1[a] = last < nc - 1 ? last : nc - 1;
    }
    const float3 q = grid_query_point(g, pt[0], pt[1], pt[2]);
    float3 sum = make_float3(0.f, 0.f, 0.f), comp = sum;
    unsigned long long n_clusters = 0, n_pairs = 0;
    const float spread = mp.spread;
    const int depth = (int)tv.depth;
    // per-size kernel factors, recomputed only when the cluster size changes
    float f_size = -1.0f, f_r2 = 0.f, f_s = 0.f, f_ir3 = 0.f;
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int l = depth - 2; l >= 0; --l) {
        const vg_grid_desc &gl = tv.layer[l];
        int c0[3], c1[3], n0[3], n1[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int dec = (int)tv.dec[l + 1][a];
            const int sc = (int)layer_scale(tv, l, a);
            c0[a] = lo[a] * dec;
            c1[a] = hi[a] * dec + dec - 1;
            const int ncl = (int)gl.np[a] - 1;
            int blo = cl0[a] / sc - mp.k, bhi = cl1[a] / sc + mp.k;
            blo = blo < 0 ? 0 : blo;
            bhi = bhi > ncl - 1 ? ncl - 1 : bhi;
            n0[a] = blo > c0[a] ? blo : c0[a];
            n1[a] = bhi < c1[a] ? bhi : c1[a];
        }
        const float4 *__restrict__ P = tv.P[l];
        const float4 *__restrict__ W = tv.W[l];
        const float4 *__restrict__ M = tv.M[l];
        float3 part = make_float3(0.f, 0.f, 0.f);
        for (int cz = c0[2]; cz <= c1[2]; ++cz)
            for (int cy = c0[1]; cy <= c1[1]; ++cy)
                for (int cx = c0[0]; cx <= c1[0]; ++cx) {
                    if (cx >= n0[0] && cx <= n1[0] && cy >= n0[1] && cy <= n1[1] &&
                        cz >= n0[2] && cz <= n1[2])
                        continue;
                    const size_t off = point_offset(gl, cx, cy, cz);
                    const float4 p = P[off];
                    if (p.w == 0.0f) continue;
                    if (COUNT) {
                        ++n_clusters;
                        continue;
                    }
                    const float4 w = W[off];
                    if (p.w != f_size) {  // uniform branch: sizes rarely change
                        f_size = p.w;
                        const float radius = p.w * 0.5f * spread;
                        f_r2 = radius * radius;
                        f_s = (2.0f / 3.0f) * f_r2 * radius;
                        f_ir3 = __builtin_amdgcn_rcpf(f_r2 * radius);
                    }
                    const float r2 = f_r2, s = f_s;
                    const float dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
                    const float d2 = dx * dx + dy * dy + dz * dz;
                    const float law = kernel_law<POLICY>(d2, r2, f_ir3);
                    const float wx = w.x * s, wy = w.y * s, wz = w.z * s;
                    part.x += (wy * dz - wz * dy) * law;
                    part.y += (wz * dx - wx * dz) * law;
                    part.z += (wx * dy - wy * dx) * law;
                    if (ORDER >= 1 && d2 >= r2) {
                        const float4 m0 = M[3 * off], m1 = M[3 * off + 1], m2 = M[3 * off + 2];
                        const float tx = m0.x * dx + m0.y * dy + m0.z * dz;
                        const float ty = m1.x * dx + m1.y * dy + m1.z * dz;
                        const float tz = m2.x * dx + m2.y * dy + m2.z * dz;
                        const float c3 = 3.0f * __builtin_amdgcn_rcpf(d2);
                        const float sl = s * law;
                        part.x -= (m0.w - c3 * (ty * dz - tz * dy)) * sl;
                        part.y -= (m1.w - c3 * (tz * dx - tx * dz)) * sl;
                        part.z -= (m2.w - c3 * (tx * dy - ty * dx)) * sl;
                    }
                }
        kahan_add(sum, comp, part);
#pragma unroll
        for (int a = 0; a < 3; ++a) { lo[a] = n0[a]; hi[a] = n1[a]; }
    }
    // Leaf near box (or every vorton when the tree has one layer).
    float3 part = make_float3(0.f, 0.f, 0.f);
    const vg_grid_desc &g0 = tv.layer[0];
    if (depth == 1) {
        lo[0] = lo[1] = lo[2] = 0;
        hi[0] = (int)g0.np[0] - 2;
        hi[1] = (int)g0.np[1] - 2;
        hi[2] = (int)g0.np[2] - 2;
    }
    for (int cz = lo[2]; cz <= hi[2]; ++cz)
        for (int cy = lo[1]; cy <= hi[1]; ++cy)
            for (int cx = lo[0]; cx <= hi[0]; ++cx) {
                const size_t key = point_offset(g0, cx, cy, cz);
                const uint32_t cnt = tv.cell_count[key];
                if (cnt == 0) continue;
                if (COUNT) {
                    n_pairs += cnt;
                    continue;
                }
                const uint32_t st = tv.cell_start[key];
                for (uint32_t j = 0; j < cnt; ++j) {
                    const float4 p = tv.sorted_p[st + j];
                    const float4 w = tv.sorted_w[st + j];
                    if (p.w != f_size) {
                        f_size = p.w;
                        const float radius = p.w * 0.5f * spread;
                        f_r2 = radius * radius;
                        f_s = (2.0f / 3.0f) * f_r2 * radius;
                        f_ir3 = __builtin_amdgcn_rcpf(f_r2 * radius);
                    }
                    const float4 A = make_float4(p.x, p.y, p.z, f_r2);
                    const float4 B = make_float4(w.x * f_s, w.y * f_s, w.z * f_s, f_ir3);
                    pair_velocity<POLICY>(part, q, A, B);
                }
            }
    kahan_add(sum, comp, part);
    if (COUNT) {
        // the ballot must run on every lane: inside the lane-0 branch only
        // lane 0 would vote
        const unsigned long long nv = (unsigned long long)__popcll(__ballot(valid));
        if (lane == 0) {
            atomicAdd(&counts[0], n_clusters * nv);
            atomicAdd(&counts[1], n_pairs * nv);
        }
        return;
    }
    if (valid) {
        const size_t o = point_offset(g, pt[0], pt[1], pt[2]);
        out[3 * o] = sum.x * mp.circ;
        out[3 * o + 1] = sum.y * mp.circ;
        out[3 * o + 2] = sum.z * mp.circ;
    }
}

template <int COUNT>
vg_status run_mac(const vg_tree *t, const vg_grid_desc *g, int near_k, int order, float spread,
                  float circ, int policy, float *out, unsigned long long *counts, hipStream_t s) {
    TreeView tv;
    VG_TRY(tree_checked_view(t, &tv));
    VG_REQUIRE(g, "grid is null");
    VG_REQUIRE(near_k >= 0 && near_k <= 16, "near_k must be in [0, 16]");
    VG_REQUIRE(order == 0 || order == 1, "order must be 0 (monopole) or 1 (dipole)");
    VG_REQUIRE(policy == VG_RSQRT_REFERENCE || policy == VG_RSQRT_ACCURATE, "bad rsqrt policy");
    for (int a = 0; a < 3; ++a)
        VG_REQUIRE(g->np[a] == tv.layer[0].np[a],
                   "grid must have the tree's leaf geometry (the grid template)");
    MacParams mp;
    mp.k = near_k;
    mp.spread = spread;
    mp.circ = circ;
    for (int a = 0; a < 3; ++a) mp.nb[a] = (g->np[a] + 3) / 4;
    mp.nbricks = mp.nb[0] * mp.nb[1] * mp.nb[2];
    const dim3 grid(ceil_div(mp.nbricks, 4));
    const char *name = COUNT ? "vg_tree_mac_count" : "vg_tree_velocity_grid_mac";
    using K = void (*)(const TreeView, vg_grid_desc, MacParams, float *, unsigned long long *);
    K k;
    if (policy == VG_RSQRT_REFERENCE)
        k = order ? k_mac_grid<VG_RSQRT_REFERENCE, 1, COUNT> : k_mac_grid<VG_RSQRT_REFERENCE, 0, COUNT>;
    else
        k = order ? k_mac_grid<VG_RSQRT_ACCURATE, 1, COUNT> : k_mac_grid<VG_RSQRT_ACCURATE, 0, COUNT>;
    return launch(name, k, grid, dim3(256), 0, s, tv, *g, mp, out, counts);
}

}  // namespace
}  // namespace vg

using namespace vg;

VG_API vg_status vg_tree_velocity_grid_mac(const vg_tree *t, const vg_grid_desc *g, int near_k,
                                           int order, float spread, float circ, int rsqrt_policy,
                                           float *out, vg_stream s) {
    VG_REQUIRE(out, "out is null");
    return run_mac<0>(t, g, near_k, order, spread, circ, rsqrt_policy, out, nullptr,
                      (hipStream_t)s);
}

VG_API vg_status vg_tree_mac_work(const vg_tree *t, const vg_grid_desc *g, int near_k,
                                  uint64_t *clusters, uint64_t *pairs, vg_stream stream) {
    VG_REQUIRE(clusters && pairs, "null output");
    hipStream_t s = (hipStream_t)stream;
    unsigned long long *d = nullptr;
    VG_TRY(salloc((void **)&d, 2 * sizeof(unsigned long long), s));
    vg_status st = zero_fill(d, 2 * sizeof(unsigned long long), s);
    if (st == VG_OK) st = run_mac<1>(t, g, near_k, 0, 1.0f, 1.0f, 0, nullptr, d, s);
    unsigned long long h[2] = {0, 0};
    if (st == VG_OK) {
        hipError_t e = hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = device_error(e, "mac work copy", __FILE__, __LINE__);
    }
    sfree(d, s);
    *clusters = h[0];
    *pairs = h[1];
    return st;
}
