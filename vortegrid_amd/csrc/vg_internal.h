// vg_internal.h -- shared internals of libvortegrid (host + device).
//
// Numerical conventions
//  * Index and position arithmetic that decides which cell a point falls in
//    (grid binning, tree opening tests, interpolation weights, cluster
//    reductions, SOR/curl stencils) is compiled with FP contraction OFF so it
//    rounds exactly like the reference's source order.
//  * The O(N^2) and treecode pair sums allow FMA contraction and accumulate
//    with tile-level Kahan compensation.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/vortegrid.h"

#define VG_API extern "C" __attribute__((visibility("default")))

namespace vg {

// ------------------------------------------------------------------ errors
void set_error(const std::string &msg);
vg_status device_error(hipError_t e, const char *what, const char *file, int line);

#define VG_HIP(call)                                                        \
    do {                                                                    \
        hipError_t _e = (call);                                             \
        if (_e != hipSuccess) return ::vg::device_error(_e, #call, __FILE__, __LINE__); \
    } while (0)

#define VG_REQUIRE(cond, msg)                                               \
    do {                                                                    \
        if (!(cond)) { ::vg::set_error(msg); return VG_EINVAL; }            \
    } while (0)

#define VG_TRY(expr)                                                        \
    do {                                                                    \
        vg_status _s = (expr);                                              \
        if (_s != VG_OK) return _s;                                         \
    } while (0)

// ------------------------------------------------------------------ profiling
// Kernel launches go through launch(); with profiling enabled each launch
// whose name is selected (vg_profile_select; all by default) is bracketed by
// two events on its stream and tallied under `name`.
bool profiling_enabled();
bool profile_wanted(const char *name);
void profile_begin(const char *name, hipStream_t s, void **token);
void profile_end(void *token, hipStream_t s);

template <typename Kernel, typename... Args>
vg_status launch(const char *name, Kernel k, dim3 grid, dim3 block, size_t shmem,
                 hipStream_t s, Args... args) {
    if (grid.x == 0 || grid.y == 0 || grid.z == 0) return VG_OK;
    void *tok = nullptr;
    const bool prof = profile_wanted(name);
    if (prof) profile_begin(name, s, &tok);
    hipLaunchKernelGGL(k, grid, block, shmem, s, args...);
    hipError_t e = hipGetLastError();
    if (prof) profile_end(tok, s);
    if (e != hipSuccess) return device_error(e, name, __FILE__, __LINE__);
    return VG_OK;
}

inline uint32_t ceil_div(size_t a, size_t b) { return (uint32_t)((a + b - 1) / b); }

// Transient device scratch, stream ordered (released with sfree on the same
// stream).
vg_status salloc(void **p, size_t bytes, hipStream_t s);
vg_status sfree(void *p, hipStream_t s);
// Zero device memory with a fill kernel (safe under graph capture, unlike
// a memset node).
vg_status zero_fill(void *p, size_t bytes, hipStream_t s);

// ------------------------------------------------------------------ geometry
// Shared by the host and the device (a batch defines its grids on the device
// inside a recorded frame), so both run the same float operations.
__host__ __device__ inline size_t capacity(const vg_grid_desc &g) {
    return (size_t)g.np[0] * g.np[1] * g.np[2];
}
constexpr uint32_t kMaxLayers = 16;
struct Nested {
    uint32_t depth = 0;
    vg_grid_desc layer[kMaxLayers];
    uint32_t dec[kMaxLayers][3];
};

// uniformGrid.h:794-809 PrecomputeSpacing
__host__ __device__ inline void precompute_spacing(vg_grid_desc *g) {
#pragma clang fp contract(off)
    for (int a = 0; a < 3; ++a) g->cell[a] = g->extent[a] / (float)(g->np[a] - 1);
    g->cpe[0] = (float)(g->np[0] - 1) / g->extent[0];
    g->cpe[1] = (float)(g->np[1] - 1) / g->extent[1];
    if (0.0f == g->extent[2])
        g->cpe[2] = 1.0f / FLT_MIN;
    else
        g->cpe[2] = (float)(g->np[2] - 1) / g->extent[2];
}

// Core/Math/math.h:149-175 NearestPowerOfTwo
__host__ __device__ inline size_t nearest_pow2(size_t v) {
    if (v == 0) return 1;
    --v;
    unsigned shift = 0;
    while ((v >> shift) != 0) ++shift;
    return (size_t)1 << shift;
}

__host__ __device__ inline void clear_desc(vg_grid_desc *g) {
    for (int a = 0; a < 3; ++a) {
        g->min[a] = 0.f;
        g->extent[a] = 0.f;
        g->cell[a] = 0.f;
        g->cpe[a] = 0.f;
        g->np[a] = 0u;
    }
}

// The reference's powf calls of the grid geometry (DefineShape's cells per
// extent, the auto-mollifier's cell size) as pow in double rounded once to
// float: within half an ulp of the exact power (up to double rounding), so
// the host's and the device's results agree bit for bit, which two powf
// implementations do not (each may be off by an ulp in different cases).
__host__ __device__ inline float pow_rounded(float x, float y) {
    return (float)pow((double)x, (double)y);
}

// uniformGrid.h:259-322 DefineShape
__host__ __device__ inline void define_shape(vg_grid_desc *g, size_t n, const float vmin[3],
                                             const float vmax[3], int pow2) {
#pragma clang fp contract(off)
    clear_desc(g);
    if (n == 0) return;
    const float nudge = 1.0f + FLT_EPSILON;
    float eff[3];
    int ndims = 3;
    for (int a = 0; a < 3; ++a) {
        g->min[a] = vmin[a];
        g->extent[a] = (vmax[a] - vmin[a]) * nudge;
        eff[a] = g->extent[a];
    }
    for (int a = 0; a < 3; ++a)
        if (0.0f == eff[a]) {
            eff[a] = 1.0f;
            g->extent[a] = 0.0f;
            --ndims;
        }
    const float volume = eff[0] * eff[1] * eff[2];
    const float icvcr = pow_rounded(volume / (float)n, -1.0f / (float)ndims);
    size_t nc[3];
    for (int a = 0; a < 3; ++a) {
        const size_t v = (size_t)(g->extent[a] * icvcr + 0.5f);
        nc[a] = v > 1 ? v : 1;
    }
    if (pow2)
        for (int a = 0; a < 3; ++a) nc[a] = nearest_pow2(nc[a]);
    while (nc[0] * nc[1] * nc[2] >= n * 8)
        for (int a = 0; a < 3; ++a) nc[a] = nc[a] / 2 > 1 ? nc[a] / 2 : 1;
    for (int a = 0; a < 3; ++a) g->np[a] = (uint32_t)(nc[a] + 1);
    precompute_spacing(g);
}

// uniformGrid.h:361-376 Decimate
__host__ __device__ inline void decimate(vg_grid_desc *dst, const vg_grid_desc *src,
                                         uint32_t dec) {
    vg_grid_desc d;
    clear_desc(&d);
    for (int a = 0; a < 3; ++a) {
        d.extent[a] = src->extent[a];
        d.min[a] = src->min[a];
        d.np[a] = (src->np[a] - 1) / dec + 1;
        if (dec > 1 && d.np[a] < 2) d.np[a] = 2;
    }
    precompute_spacing(&d);
    *dst = d;
}

// nestedGrid.h:69-85 Initialize, :371-391 PrecomputeDecimations
__host__ __device__ inline void nested_init(Nested *ng, const vg_grid_desc &src) {
    ng->depth = 0;
    for (uint32_t l = 0; l < kMaxLayers; ++l) {
        clear_desc(&ng->layer[l]);
        for (int a = 0; a < 3; ++a) ng->dec[l][a] = 0u;
    }
    decimate(&ng->layer[0], &src, 1);
    uint32_t idx = 1;
    while (capacity(ng->layer[idx - 1]) > 8 && idx < kMaxLayers) {
        decimate(&ng->layer[idx], &ng->layer[idx - 1], 2);
        ++idx;
    }
    ng->depth = idx;
    for (uint32_t l = 1; l < idx; ++l)
        for (int a = 0; a < 3; ++a)
            ng->dec[l][a] = (ng->layer[l - 1].np[a] - 1) / (ng->layer[l].np[a] - 1);
}

// True when a layer leaves a cell without a parent: on some axis the child
// layer has more cells than decimation x parent cells (an odd count above 3
// halved, e.g. 5 -> 2 parents covering 4).  AggregateClusters
// (vortonSim.cpp:1337-1569) never folds such a cell into the layer above,
// so a walk from the upper layers cannot reach its vortons.  The reference
// only builds power-of-two grids (UpdateBoundingBox calls DefineShape with
// bPowerOf2 = true, vortonSim.cpp:1089), where this cannot happen; the
// near-field walks refuse such trees instead of dropping vorticity.
template <class NG>
__host__ __device__ inline bool has_orphan_cells(const NG &ng) {
    for (uint32_t l = 1; l < ng.depth; ++l)
        for (int a = 0; a < 3; ++a)
            if (ng.layer[l - 1].np[a] - 1 != ng.dec[l][a] * (ng.layer[l].np[a] - 1)) return true;
    return false;
}

// ------------------------------------------------------------------ device view of a tree
struct TreeView {
    uint32_t depth;
    uint32_t dec[kMaxLayers][3];
    vg_grid_desc layer[kMaxLayers];
    const float4 *P[kMaxLayers];   // cluster (x, y, z, size)
    const float4 *W[kMaxLayers];   // cluster (Omega_x, Omega_y, Omega_z, sum|w|)
    const float4 *M[kMaxLayers];   // 3 per cluster: rows of the first moment
                                   // D[b][e] = sum Omega_b (y - c)_e, .w = antisymmetric part
    const uint32_t *C[kMaxLayers]; // layers >= 1, per cluster: bit cx + dx (cy + dy cz) set
                                   // when that child (dx dy dz per parent) is not empty
    // Leaf cell lists over the layer-0 points: vortons sorted by cell.
    const uint32_t *cell_start;    // per layer-0 point
    const uint32_t *cell_count;
    const float4 *sorted_p;        // (x, y, z, size) in sorted order
    const float4 *sorted_w;        // (Omega, 0)
    const uint32_t *sorted_idx;    // original vorton index
    uint32_t n;
};

// ------------------------------------------------------------------ batched trees
// `trees` influence trees of one shape (one Nested layout), stored back to
// back: layer l of tree k occupies points [k cap_l, (k+1) cap_l).  Vortons
// of all trees are concatenated; tree_of[v] names the owner, leaf_desc[k]
// (device) the tree's leaf geometry, size[k] (device) its vorton diameter.
struct BatchTreeArgs {
    uint32_t trees = 0;
    const Nested *shape = nullptr;
    size_t n = 0;
    const float *pos = nullptr, *angvel = nullptr;  // packed xyz
    const uint32_t *tree_of = nullptr;
    const vg_grid_desc *leaf_desc = nullptr;
    const float *size = nullptr;
    uint32_t *cell_count = nullptr, *cell_start = nullptr;  // trees * cap0
    uint32_t *keys = nullptr, *slots = nullptr, *sorted_idx = nullptr;  // n
    float4 *sorted_p = nullptr, *sorted_w = nullptr;                    // n
    uint32_t *scan_tmp = nullptr;  // scan_scratch_elems(trees * cap0)
    float4 *P[kMaxLayers] = {}, *W[kMaxLayers] = {}, *M[kMaxLayers] = {};
    uint32_t *C[kMaxLayers] = {};  // child masks, layers >= 1
    // non-null: the trees have their own shapes (device views), each within
    // `shape`, whose capacities are the storage strides
    const TreeView *views = nullptr;
    // write the clusters of empty leaf cells too (readers that scan leaf
    // clusters by position, such as the reference-rule walk)
    bool all_leaves = false;
};
size_t scan_scratch_elems(size_t n);
// Exclusive scan in -> out using tmp (scan_scratch_elems(n) elements).
vg_status exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, uint32_t *tmp,
                             hipStream_t s);
vg_status batch_tree_build(const BatchTreeArgs &a, hipStream_t s);
// Reference-rule velocity grids of batched trees (device views, each on its
// own leaf grid within np; mollify[tree] = (spread, circ)), tree k's grid at
// out + k * out_stride.
vg_status tree_velocity_grid_batch(const TreeView *views, const float2 *mollify, uint32_t trees,
                                   const uint32_t np[3], int rsqrt_policy, float *out,
                                   size_t out_stride, hipStream_t s);
// Work queue of a batched walk over trees of different shapes (see
// k_mac_grid_queue): trees in queue order, the first (tree, brick) item of
// each (trees + 1 entries), two counters (walk, expansions).
struct BatchQueue {
    const uint32_t *order = nullptr;
    const uint32_t *brick_start = nullptr;
    uint32_t *counters = nullptr;
};
// Stable counting sort of particles owned by batched trees into (owner,
// leaf cell of the owner's grid) order.  Owners are contiguous and ascending,
// so every particle stays in its owner's range: tree_of does not change.
struct BatchSortArgs {
    size_t n = 0, cap0 = 0, cells = 0;  // particles, points per tree, trees * cap0
    float *pos = nullptr;               // packed xyz, reordered in place
    const uint32_t *tree_of = nullptr;
    const vg_grid_desc *leaf_desc = nullptr;
    uint32_t *count = nullptr, *start = nullptr;                         // cells
    uint32_t *keys = nullptr, *slots = nullptr, *unsorted = nullptr;     // n
    uint32_t *sorted = nullptr;                                          // n
    uint32_t *scan_tmp = nullptr;  // scan_scratch_elems(cells)
    float *moved = nullptr;        // 3 n
};
vg_status batch_sort_particles(const BatchSortArgs &a, hipStream_t s);

// ------------------------------------------------------------------ device helpers
#ifdef __HIPCC__

// Core/Math/math.h:206-216: 0x5f3759df estimate + one Newton step.
__device__ __forceinline__ float finvsqrtf(float v) {
    int i = __float_as_int(v);
    i = 0x5f3759df - (i >> 1);
    float y = __int_as_float(i);
    const float x = v * 0.5f;
    y = y * (1.5f - x * y * y);
    return y;
}

// a + b*c with both roundings (no FMA): matches `min + float(i) * spacing`.
__device__ __forceinline__ float madd_exact(float a, float b, float c) {
#pragma clang fp contract(off)
    return a + b * c;
}

// IndicesOfPosition (uniformGrid.h:536-574) for one axis: truncation of
// (p - min) * cellsPerExtent, clamped to a valid cell (the reference leaves
// out-of-domain positions undefined).
__device__ __forceinline__ int cell_of(float p, float gmin, float cpe, int ncells) {
#pragma clang fp contract(off)
    const float f = (p - gmin) * cpe;
    int i = (f > -1.0f) ? (int)f : -1;   // trunc; values <= -1 clamp to 0 below
    i = i < 0 ? 0 : i;
    i = i > ncells - 1 ? ncells - 1 : i;
    return i;
}

__device__ __forceinline__ size_t point_offset(const vg_grid_desc &g, uint32_t ix, uint32_t iy,
                                               uint32_t iz) {
    return (size_t)ix + (size_t)g.np[0] * ((size_t)iy + (size_t)g.np[1] * iz);
}

// Grid point position with the (1 - 2 eps) spacing nudge of the grid-point
// query loops (vortonSim.cpp:2064-2065, 2229-2230).
__device__ __forceinline__ float3 grid_query_point(const vg_grid_desc &g, uint32_t ix,
                                                   uint32_t iy, uint32_t iz) {
#pragma clang fp contract(off)
    const float nudge = 1.0f - 2.0f * FLT_EPSILON;
    float3 p;
    p.x = g.min[0] + (float)ix * (g.cell[0] * nudge);
    p.y = g.min[1] + (float)iy * (g.cell[1] * nudge);
    p.z = g.min[2] + (float)iz * (g.cell[2] * nudge);
    return p;
}

// The cell containing p and the trilinear tweens of Interpolate
// (uniformGrid.h:1117-1135): t = (p - cell min corner) * cellsPerExtent,
// o = 1 - t.
struct CellWeights {
    int ix, iy, iz;
    float tx, ty, tz, ox, oy, oz;
};
__device__ __forceinline__ CellWeights cell_weights(const vg_grid_desc &g, float3 p) {
#pragma clang fp contract(off)
    CellWeights w;
    w.ix = cell_of(p.x, g.min[0], g.cpe[0], (int)g.np[0] - 1);
    w.iy = cell_of(p.y, g.min[1], g.cpe[1], (int)g.np[1] - 1);
    w.iz = cell_of(p.z, g.min[2], g.cpe[2], (int)g.np[2] - 1);
    w.tx = (p.x - (g.min[0] + (float)w.ix * g.cell[0])) * g.cpe[0];
    w.ty = (p.y - (g.min[1] + (float)w.iy * g.cell[1])) * g.cpe[1];
    w.tz = (p.z - (g.min[2] + (float)w.iz * g.cell[2])) * g.cpe[2];
    w.ox = 1.0f - w.tx;
    w.oy = 1.0f - w.ty;
    w.oz = 1.0f - w.tz;
    return w;
}

// Interpolate (uniformGrid.h:1117-1148, 1225-1258): nested lerp x, y, z.
__device__ __forceinline__ float3 interpolate(const vg_grid_desc &g, const float *__restrict__ f,
                                              float3 p) {
#pragma clang fp contract(off)
    const CellWeights w = cell_weights(g, p);
    const float tx = w.tx, ty = w.ty, tz = w.tz, ox = w.ox, oy = w.oy, oz = w.oz;
    const size_t nx = g.np[0], nxy = (size_t)g.np[0] * g.np[1];
    const size_t o000 = (size_t)w.ix + nx * ((size_t)w.iy + (size_t)g.np[1] * w.iz);
    const float *a000 = f + 3 * o000, *a100 = a000 + 3, *a010 = a000 + 3 * nx,
                *a110 = a010 + 3, *a001 = a000 + 3 * nxy, *a101 = a001 + 3,
                *a011 = a001 + 3 * nx, *a111 = a011 + 3;
    float r[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        r[c] = ((ox * a000[c] + tx * a100[c]) * oy + (ox * a010[c] + tx * a110[c]) * ty) * oz +
               ((ox * a001[c] + tx * a101[c]) * oy + (ox * a011[c] + tx * a111[c]) * ty) * tz;
    }
    return make_float3(r[0], r[1], r[2]);
}

// Kahan compensated add of a partial sum into (sum, comp).
__device__ __forceinline__ void kahan_add(float3 &sum, float3 &comp, float3 part) {
    float y, t;
    y = part.x - comp.x; t = sum.x + y; comp.x = (t - sum.x) - y; sum.x = t;
    y = part.y - comp.y; t = sum.y + y; comp.y = (t - sum.y) - y; sum.y = t;
    y = part.z - comp.z; t = sum.z + y; comp.z = (t - sum.z) - y; sum.z = t;
}

// Source record prepared for the pair kernels (vorton.h:118-179 with the
// per-source factors hoisted):  A = (x, y, z, radius^2),
// B = (Omega * (2/3) r^2 r, 1/(r^2 r)).  circ is applied by the caller.
__device__ __forceinline__ void prepare_source(float4 ps, float4 w, float spread, float4 &A,
                                               float4 &B) {
#pragma clang fp contract(off)
    const float radius = ps.w * 0.5f * spread;
    const float r2 = radius * radius;
    const float s = (2.0f / 3.0f) * r2 * radius;
    A = make_float4(ps.x, ps.y, ps.z, r2);
    B = make_float4(w.x * s, w.y * s, w.z * s, 1.0f / (r2 * radius));
}

// VORTON_ACCUMULATE_VELOCITY_private (vorton.h:161-179).
template <int POLICY>
__device__ __forceinline__ void pair_velocity(float3 &acc, float3 q, float4 A, float4 B) {
    const float dx = q.x - A.x, dy = q.y - A.y, dz = q.z - A.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    float law;
    if (POLICY == VG_RSQRT_REFERENCE) {
        law = finvsqrtf(d2) * __builtin_amdgcn_rcpf(d2);
    } else {
        const float rs = __builtin_amdgcn_rsqf(d2);
        law = rs * rs * rs;
    }
    law = d2 < A.w ? B.w : law;
    // (a x d) * law == a x (d * law)
    const float cx = B.y * dz - B.z * dy;
    const float cy = B.z * dx - B.x * dz;
    const float cz = B.x * dy - B.y * dx;
    acc.x += cx * law;
    acc.y += cy * law;
    acc.z += cz * law;
}

// VORTON_ACCUMULATE_VECTOR_POTENTIAL_private (vorton.h:118-136).
template <int POLICY>
__device__ __forceinline__ void pair_vecpot(float3 &acc, float3 q, float4 A, float4 B) {
    const float dx = q.x - A.x, dy = q.y - A.y, dz = q.z - A.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    float law = POLICY == VG_RSQRT_REFERENCE ? finvsqrtf(d2) : __builtin_amdgcn_rsqf(d2);
    law = d2 < A.w ? d2 * B.w : law;
    acc.x += B.x * law;
    acc.y += B.y * law;
    acc.z += B.z * law;
}

__device__ __forceinline__ float4 load_strided3(const float *p, size_t stride, size_t i,
                                                float w) {
    const float *r = p + i * stride;
    return make_float4(r[0], r[1], r[2], w);
}

#endif  // __HIPCC__

}  // namespace vg
