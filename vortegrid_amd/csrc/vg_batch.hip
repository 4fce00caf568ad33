// vg_batch.hip -- many independent small vorton simulations ("effects")
// advanced together (no reference counterpart: the reference runs one
// VortonSim per particle group, vortonSim.cpp:4117-4256, each on its own).
//
// Every effect owns a grid of a fixed shape (cells^3 cells over its own
// padded bounding box) and so a NestedGrid of one fixed layout: all effects'
// trees are stored back to back and every stage of the frame is ONE launch
// over all effects:
//
//   tracers   AssignVelocityFromField(previous grid) + Evolve   (per tracer)
//   bounds    vortons u tracers AABB per effect                 (per item)
//   geometry  pad 3 radii, grid + layer descriptors, mollifier  (per effect)
//   tree      bin / scan / scatter / cell sort / gather / leaf / aggregate
//   velocity  near-field treecode at grid points                (per brick)
//   vortons   AssignVelocityFromField + Evolve                  (per vorton)
//
// Nothing in the frame goes back to the host (geometry is computed on the
// device), so a frame can be recorded once into a hipGraph and replayed.
// Launch shapes of the walk and the tracer kernel are chosen by
// vg_batch_autotune from timed trials on the live batch.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "vg_internal.h"

namespace vg {
vg_status mac_velocity_grid_batch(const TreeView *views, const float2 *mollify, uint32_t trees,
                                  const uint32_t np[3], int near_k, int order, int expand_from,
                                  int policy, int waves_per_block, float *out, size_t out_stride,
                                  hipStream_t s, const BatchQueue *queue, bool own_shapes);
}

struct vg_batch {
    uint32_t effects = 0;
    vg_batch_config cfg{};
    vg::Nested shape{};  // the fixed shape, or with adaptive shapes the largest one
    bool adaptive = false;
    size_t cap[vg::kMaxLayers] = {};
    size_t cap0 = 0, nv = 0, nt = 0;
    // particles (packed xyz, effects back to back) and their owners
    float *vpos = nullptr, *vom = nullptr, *vvel = nullptr, *tpos = nullptr;
    uint32_t *vtree = nullptr, *ttree = nullptr;
    float *radius = nullptr, *size = nullptr;  // per effect
    // trees
    float4 *P[vg::kMaxLayers] = {}, *W[vg::kMaxLayers] = {}, *M[vg::kMaxLayers] = {};
    uint32_t *C[vg::kMaxLayers] = {};  // non-empty child masks (layers >= 1)
    uint32_t *cell_count = nullptr, *cell_start = nullptr, *keys = nullptr, *slots = nullptr,
             *sorted_idx = nullptr, *scan_tmp = nullptr;
    float4 *sorted_p = nullptr, *sorted_w = nullptr;
    vg::TreeView *views = nullptr;  // per effect, device
    vg_grid_desc *leaf_desc = nullptr;
    float2 *mollify = nullptr;
    uint32_t *bbox = nullptr;  // 6 ordered-encoded extrema per effect and slot
    float *grid = nullptr;     // effects * cap0 * 3
    // adaptive shapes: vorton count per effect, the walk's work queue (effects
    // by descending vorton count, first brick of each), and counters (walk,
    // expansions, effect-frames whose shape was clamped)
    uint32_t *vcount = nullptr, *qorder = nullptr, *brick_start = nullptr, *counters = nullptr;
    // tracer cell sort (allocated by the first vg_batch_sort_tracers)
    uint32_t *ts_count = nullptr, *ts_start = nullptr, *ts_keys = nullptr, *ts_slots = nullptr,
             *ts_unsorted = nullptr, *ts_sorted = nullptr, *ts_scan = nullptr;
    float *ts_moved = nullptr;
    // launch shapes
    int walk_waves = 4, adv_block = 256, adv_items = 1;
    // recorded frame
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool stepped = false;  // trees and grids hold a computed frame
    std::vector<void *> allocs;
};

namespace {
using namespace vg;

__device__ __forceinline__ uint32_t ord_enc(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord_dec(uint32_t u) {
    u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    return __uint_as_float(u);
}

// Each effect's extrema are folded into kBoxSlots copies, the copy chosen by
// the workgroup, so that the many workgroups of one large effect do not all
// contend for the same six words; k_geometry reduces the copies.
constexpr uint32_t kBoxSlots = 32;
static_assert(kBoxSlots <= 64, "k_geometry folds the copies across one wave");

__global__ void k_bbox_init(uint32_t *__restrict__ bbox, uint32_t effects) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 6 * kBoxSlots * effects) return;
    bbox[i] = (i % 6) < 3 ? 0xffffffffu : 0u;
}

// Extrema of items (positions packed xyz) per owner.  Owners are contiguous,
// so each wave folds a contiguous chunk of kBoxItems items: a lane keeps
// running extrema of its items while the owner stays the same and flushes
// them with atomics when it changes; at the end a wave whose lanes share one
// owner reduces before its 6 atomics.  Min / max are exact, so the result
// does not depend on the order.
constexpr int kBoxItems = 16;  // items per lane

__device__ __forceinline__ void box_flush(uint32_t *bbox, uint32_t k, const uint32_t *e) {
    uint32_t *box = bbox + 6 * ((size_t)k * kBoxSlots + (blockIdx.x & (kBoxSlots - 1)));
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        atomicMin(&box[a], e[a]);
        atomicMax(&box[3 + a], e[3 + a]);
    }
}

__global__ __launch_bounds__(256) void k_bbox_batch(const float *__restrict__ pos,
                                                    const uint32_t *__restrict__ owner, size_t n,
                                                    uint32_t *__restrict__ bbox) {
    const uint32_t lane = threadIdx.x & 63;
    const size_t wave = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const size_t base = wave * 64 * kBoxItems;
    uint32_t cur = 0xffffffffu;
    uint32_t e[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    auto fold = [&](uint32_t k, float x, float y, float z) {
        if (k != cur) {
            if (cur != 0xffffffffu) box_flush(bbox, cur, e);
            cur = k;
            e[0] = e[1] = e[2] = 0xffffffffu;
            e[3] = e[4] = e[5] = 0u;
        }
        const uint32_t v[3] = {ord_enc(x), ord_enc(y), ord_enc(z)};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            e[a] = min(e[a], v[a]);
            e[3 + a] = max(e[3 + a], v[a]);
        }
    };
    if (base + 64 * kBoxItems <= n) {
        // whole chunk in range: lane takes 4 consecutive items per group of
        // 256, read as one uint4 of owners and three float4 of positions
        // (the chunk starts at a multiple of 1024 items, so 16-byte aligned)
#pragma unroll
        for (int grp = 0; grp < kBoxItems / 4; ++grp) {
            const size_t j = base + 256 * (size_t)grp + 4 * lane;
            const uint4 k = *reinterpret_cast<const uint4 *>(owner + j);
            const float4 *p4 = reinterpret_cast<const float4 *>(pos + 3 * j);
            const float4 a = p4[0], b = p4[1], c = p4[2];
            fold(k.x, a.x, a.y, a.z);
            fold(k.y, a.w, b.x, b.y);
            fold(k.z, b.z, b.w, c.x);
            fold(k.w, c.y, c.z, c.w);
        }
    } else {
        for (int it = 0; it < kBoxItems; ++it) {
            const size_t i = base + (size_t)it * 64 + lane;
            if (i >= n) break;
            fold(owner[i], pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        }
    }
    const uint32_t k0 = __builtin_amdgcn_readfirstlane(cur);
    const bool uniform = __ballot(cur != 0xffffffffu && cur != k0) == 0ull;
    if (uniform) {
        if (k0 == 0xffffffffu) return;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                e[a] = min(e[a], (uint32_t)__shfl_xor((int)e[a], off, 64));
                e[3 + a] = max(e[3 + a], (uint32_t)__shfl_xor((int)e[3 + a], off, 64));
            }
        if (lane == 0) box_flush(bbox, k0, e);
    } else if (cur != 0xffffffffu) {
        box_flush(bbox, cur, e);
    }
}

// Tracers of all effects in one pass over their positions: AssignVelocity-
// FromField (gain 1) + Evolve with each owner's grid (k_advect_batch's
// arithmetic), and the extrema of the moved positions folded into the
// owners' boxes (the tracer half of the frame's bounding boxes).  A
// workgroup takes kAdvRound x 256 consecutive tracers, kAdvRound per thread
// with all their gathers in flight together.  Owners are contiguous, so a
// workgroup nearly always holds one owner: it then reduces its extrema across
// lanes and waves before one set of ordered-integer atomics; otherwise every
// lane flushes its own running extrema when its owner changes.
constexpr int kAdvRound = 4;

__global__ __launch_bounds__(256) void k_advect_box_batch(const TreeView *__restrict__ views,
                                                          const float *__restrict__ grid,
                                                          size_t cap0, float *__restrict__ pos,
                                                          const uint32_t *__restrict__ owner,
                                                          size_t n, float dt,
                                                          uint32_t *__restrict__ bbox) {
#pragma clang fp contract(off)
    __shared__ uint32_t wave_box[4][7];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t base = (size_t)blockIdx.x * (256 * kAdvRound) + threadIdx.x;
    uint32_t cur = 0xffffffffu;
    uint32_t e[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    uint32_t k[kAdvRound];
    float3 p[kAdvRound], v[kAdvRound];
    bool live[kAdvRound];
#pragma unroll
    for (int r = 0; r < kAdvRound; ++r) {
        const size_t i = base + (size_t)r * 256;
        live[r] = i < n;
        k[r] = live[r] ? owner[i] : 0u;
        p[r] = live[r] ? make_float3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2])
                       : make_float3(0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int r = 0; r < kAdvRound; ++r)
        v[r] = live[r] ? interpolate(views[k[r]].layer[0], grid + 3 * cap0 * k[r], p[r])
                       : make_float3(0.f, 0.f, 0.f);
#pragma unroll
    for (int r = 0; r < kAdvRound; ++r) {
        if (!live[r]) continue;
        const size_t i = base + (size_t)r * 256;
        const float x = p[r].x + v[r].x * dt, y = p[r].y + v[r].y * dt, z = p[r].z + v[r].z * dt;
        pos[3 * i] = x;
        pos[3 * i + 1] = y;
        pos[3 * i + 2] = z;
        if (k[r] != cur) {
            if (cur != 0xffffffffu) box_flush(bbox, cur, e);
            cur = k[r];
            e[0] = e[1] = e[2] = 0xffffffffu;
            e[3] = e[4] = e[5] = 0u;
        }
        const uint32_t c[3] = {ord_enc(x), ord_enc(y), ord_enc(z)};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            e[a] = min(e[a], c[a]);
            e[3 + a] = max(e[3 + a], c[a]);
        }
    }
    // one owner in the wave? then reduce it to lane 0
    const uint32_t k0 = __builtin_amdgcn_readfirstlane(cur);
    const bool uniform = __ballot(cur != 0xffffffffu && cur != k0) == 0ull;
    if (uniform) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                e[a] = min(e[a], (uint32_t)__shfl_xor((int)e[a], off, 64));
                e[3 + a] = max(e[3 + a], (uint32_t)__shfl_xor((int)e[3 + a], off, 64));
            }
    } else if (cur != 0xffffffffu) {
        box_flush(bbox, cur, e);  // mixed wave: every lane flushes its own
    }
    if (lane == 0) {
        wave_box[wave][6] = uniform ? k0 : 0xfffffffeu;  // 0xffffffff: nothing
#pragma unroll
        for (int a = 0; a < 6; ++a) wave_box[wave][a] = e[a];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // fold the uniform waves' extrema, one set of atomics per owner run
        uint32_t ko = 0xffffffffu, acc[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
        for (int w = 0; w < 4; ++w) {
            const uint32_t kw = wave_box[w][6];
            if (kw >= 0xfffffffeu) continue;
            if (kw != ko) {
                if (ko != 0xffffffffu) box_flush(bbox, ko, acc);
                ko = kw;
                acc[0] = acc[1] = acc[2] = 0xffffffffu;
                acc[3] = acc[4] = acc[5] = 0u;
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                acc[a] = min(acc[a], wave_box[w][a]);
                acc[3 + a] = max(acc[3 + a], wave_box[w][3 + a]);
            }
        }
        if (ko != 0xffffffffu) box_flush(bbox, ko, acc);
    }
}

struct Shape {
    uint32_t depth;
    uint32_t np[kMaxLayers][3];
    uint32_t adaptive;   // 1: each effect's shape from DefineShape
    uint32_t max_np;     // adaptive: at most this many points per axis
};

// Per effect: UpdateBoundingBox(final) padding (vortonSim.cpp:1027-1096),
// then either the fixed grid shape over the padded box, or (adaptive)
// DefineShape(vorton count, padded box, power-of-two cells) as
// UpdateBoundingBox calls it (vortonSim.cpp:1089) and the NestedGrid layers
// of that shape; every layer's descriptor; auto-mollification
// (vortonSim.cpp:2458-2477).  An effect with no items keeps a unit box (4
// cells per axis when adaptive).  An adaptive shape beyond max_np points on
// an axis is limited to max_np there, and adds one to *clamped.
__global__ void k_geometry(const uint32_t *__restrict__ bbox, const float *__restrict__ radius,
                           const uint32_t *__restrict__ vcount, uint32_t effects, Shape shape,
                           TreeView *__restrict__ views, vg_grid_desc *__restrict__ leaf_desc,
                           float2 *__restrict__ mollify, uint32_t *__restrict__ clamped) {
#pragma clang fp contract(off)
    // one wave per effect: lane c < kBoxSlots reads copy c of the extrema and
    // the wave folds them (ordered encodings: min / max commute, so the
    // result does not depend on the order); lane 0 then defines the grid
    const uint32_t e = blockIdx.x, lane = threadIdx.x;
    uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
    if (lane < kBoxSlots) {
        const uint32_t *box = bbox + 6 * ((size_t)e * kBoxSlots + lane);
        for (int a = 0; a < 3; ++a) {
            lo[a] = box[a];
            hi[a] = box[3 + a];
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = min(lo[a], (uint32_t)__shfl_xor((int)lo[a], m, 64));
            hi[a] = max(hi[a], (uint32_t)__shfl_xor((int)hi[a], m, 64));
        }
    if (lane != 0) return;
    float mn[3], mx[3];
    bool empty = false;
    for (int a = 0; a < 3; ++a) {
        mn[a] = ord_dec(lo[a]);
        mx[a] = ord_dec(hi[a]);
        empty = empty || !(mn[a] <= mx[a]);
    }
    const float r = radius[e];
    const float pad = (1.0f + 2.0f) * r;
    for (int a = 0; a < 3; ++a) {
        if (empty) {
            mn[a] = 0.f;
            mx[a] = 1.f;
        } else {
            mn[a] -= pad;
            mx[a] += pad;
        }
    }
    TreeView &tv = views[e];
    vg_grid_desc leaf;  // the leaf layer, for the mollifier
    if (shape.adaptive) {
        vg_grid_desc g0;
        bool clamp = false;
        if (empty || vcount[e] == 0) {
            for (int a = 0; a < 3; ++a) {
                g0.min[a] = mn[a];
                g0.extent[a] = mx[a] - mn[a];
                g0.np[a] = 5;
            }
            precompute_spacing(&g0);
        } else {
            define_shape(&g0, vcount[e], mn, mx, 1);
            for (int a = 0; a < 3; ++a)
                if (g0.np[a] > shape.max_np) {
                    g0.np[a] = shape.max_np;
                    clamp = true;
                }
            if (clamp) precompute_spacing(&g0);
        }
        if (clamp) atomicAdd(clamped, 1u);
        // NestedGrid layers of g0 (nested_init's steps), each decimated from
        // the previous one held in registers and written into the view:
        // layers past the depth keep stale descriptors, which nothing reads
        vg_grid_desc cur;
        decimate(&cur, &g0, 1);
        tv.layer[0] = cur;
        leaf_desc[e] = cur;
        leaf = cur;
        uint32_t depth = 1;
        while (capacity(cur) > 8 && depth < kMaxLayers) {
            vg_grid_desc next;
            decimate(&next, &cur, 2);
            tv.layer[depth] = next;
            for (int a = 0; a < 3; ++a) tv.dec[depth][a] = (cur.np[a] - 1) / (next.np[a] - 1);
            cur = next;
            ++depth;
        }
        tv.depth = depth;
    } else {
        for (uint32_t l = 0; l < shape.depth; ++l) {
            vg_grid_desc g;
            for (int a = 0; a < 3; ++a) {
                g.min[a] = mn[a];
                g.extent[a] = mx[a] - mn[a];
                g.np[a] = shape.np[l][a];
            }
            precompute_spacing(&g);
            tv.layer[l] = g;
            if (l == 0) {
                leaf_desc[e] = g;
                leaf = g;
            }
        }
    }
    const float cell_volume = leaf.cell[0] * leaf.cell[1] * leaf.cell[2];
    float s = pow_rounded(cell_volume, 1.0f / 3.0f) / r;
    if (s < 1.0f) s = 1.0f;
    mollify[e] = make_float2(s, 1.0f / (s * s * s));
}

// First (tree, brick) item of every effect in queue order, and the total:
// one workgroup, each thread summing a contiguous run of effects.
__global__ __launch_bounds__(256) void k_brick_starts(const TreeView *__restrict__ views,
                                                      const uint32_t *__restrict__ order,
                                                      uint32_t effects,
                                                      uint32_t *__restrict__ brick_start) {
    __shared__ uint32_t part[256];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (effects + 255) / 256;
    const uint32_t q0 = t * per < effects ? t * per : effects;
    const uint32_t q1 = q0 + per < effects ? q0 + per : effects;
    auto bricks = [&](uint32_t q) {
        const vg_grid_desc &g = views[order[q]].layer[0];
        return ((g.np[0] + 3) / 4) * ((g.np[1] + 3) / 4) * ((g.np[2] + 3) / 4);
    };
    uint32_t sum = 0;
    for (uint32_t q = q0; q < q1; ++q) sum += bricks(q);
    part[t] = sum;
    __syncthreads();
    if (t == 0) {
        uint32_t run = 0;
        for (int i = 0; i < 256; ++i) {
            const uint32_t v = part[i];
            part[i] = run;
            run += v;
        }
    }
    __syncthreads();
    uint32_t run = part[t];
    for (uint32_t q = q0; q < q1; ++q) {
        brick_start[q] = run;
        run += bricks(q);
    }
    if (q1 == effects && q0 < q1) brick_start[effects] = run;
    if (effects == 0 && t == 0) brick_start[0] = 0;
}

// AssignVelocityFromField (gain 1) + Evolve for items owned by effects, each
// interpolating its owner's grid; vel (may be null) receives the velocity.
template <int ITEMS>
__global__ void k_advect_batch(const TreeView *__restrict__ views, const float *__restrict__ grid,
                               size_t cap0, float *__restrict__ pos, float *__restrict__ vel,
                               const uint32_t *__restrict__ owner, size_t n, float dt) {
#pragma clang fp contract(off)
    const size_t base = ((size_t)blockIdx.x * ITEMS) * blockDim.x + threadIdx.x;
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        const size_t i = base + (size_t)it * blockDim.x;
        if (i >= n) return;
        const uint32_t k = owner[i];
        const vg_grid_desc &g = views[k].layer[0];
        float *p = pos + 3 * i;
        const float3 v = interpolate(g, grid + 3 * cap0 * k, make_float3(p[0], p[1], p[2]));
        if (vel) {
            vel[3 * i] = v.x;
            vel[3 * i + 1] = v.y;
            vel[3 * i + 2] = v.z;
        }
        p[0] = p[0] + v.x * dt;
        p[1] = p[1] + v.y * dt;
        p[2] = p[2] + v.z * dt;
    }
}

vg_status advect(vg_batch *b, float *pos, float *vel, const uint32_t *owner, size_t n, float dt,
                 int block, int items, hipStream_t s) {
    if (n == 0) return VG_OK;
    const dim3 grid(ceil_div(n, (size_t)block * items));
    const char *name = vel ? "vg_batch_vortons" : "vg_batch_tracers";
    switch (items) {
        case 1:
            return launch(name, k_advect_batch<1>, grid, dim3(block), 0, s,
                          (const TreeView *)b->views, (const float *)b->grid, b->cap0, pos, vel,
                          owner, n, dt);
        case 2:
            return launch(name, k_advect_batch<2>, grid, dim3(block), 0, s,
                          (const TreeView *)b->views, (const float *)b->grid, b->cap0, pos, vel,
                          owner, n, dt);
        default:
            return launch(name, k_advect_batch<4>, grid, dim3(block), 0, s,
                          (const TreeView *)b->views, (const float *)b->grid, b->cap0, pos, vel,
                          owner, n, dt);
    }
}

template <typename T>
vg_status dalloc(vg_batch *b, T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    VG_HIP(hipMalloc((void **)p, count * sizeof(T)));
    b->allocs.push_back(*p);
    return VG_OK;
}

__global__ void k_zero_f(float *p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// The near-field walk of every effect: plain launches over the fixed
// shape's bricks; with device-defined shapes, work-queue launches over each
// effect's own bricks, except for a single effect, whose plain launch over
// the largest shape's bricks wastes nothing when max_cells fits its grid
// (the workgroups past its own bricks leave at once).
vg_status walk(vg_batch *b, int waves, hipStream_t s) {
    BatchQueue q;
    q.order = b->qorder;
    q.brick_start = b->brick_start;
    q.counters = b->counters;
    if (b->cfg.near_k < 0)
        return tree_velocity_grid_batch(b->views, b->mollify, b->effects, b->shape.layer[0].np,
                                        b->cfg.rsqrt_policy, b->grid, 3 * b->cap0, s);
    const bool queue = b->adaptive && b->effects > 1;
    return mac_velocity_grid_batch(b->views, b->mollify, b->effects, b->shape.layer[0].np,
                                   b->cfg.near_k, b->cfg.order, b->cfg.expand_from,
                                   b->cfg.rsqrt_policy, waves, b->grid, 3 * b->cap0, s,
                                   queue ? &q : nullptr, b->adaptive);
}

// Bounds, geometry, trees and the velocity grids of the current particles.
// tracers_boxed: the tracers' extrema are already folded into b->bbox (by
// the fused tracer pass), which then must not be cleared here.
vg_status solve(vg_batch *b, bool tracers_boxed, hipStream_t s) {
    if (!tracers_boxed) {
        VG_TRY(launch("vg_batch_bbox", k_bbox_init,
                      dim3(ceil_div(6 * kBoxSlots * b->effects, 256)), dim3(256),
                      0, s, b->bbox, b->effects));
        if (b->nt)
            VG_TRY(launch("vg_batch_bbox", k_bbox_batch, dim3(ceil_div(b->nt, 256 * kBoxItems)),
                          dim3(256), 0, s, (const float *)b->tpos, (const uint32_t *)b->ttree,
                          b->nt, b->bbox));
    }
    if (b->nv)
        VG_TRY(launch("vg_batch_bbox", k_bbox_batch, dim3(ceil_div(b->nv, 256 * kBoxItems)),
                      dim3(256), 0, s, (const float *)b->vpos, (const uint32_t *)b->vtree, b->nv,
                      b->bbox));
    Shape sh;
    sh.depth = b->shape.depth;
    for (uint32_t l = 0; l < kMaxLayers; ++l)
        for (int a = 0; a < 3; ++a) sh.np[l][a] = l < sh.depth ? b->shape.layer[l].np[a] : 0;
    sh.adaptive = b->adaptive ? 1u : 0u;
    sh.max_np = b->shape.layer[0].np[0];
    VG_TRY(launch("vg_batch_geometry", k_geometry, dim3(b->effects), dim3(64), 0, s,
                  (const uint32_t *)b->bbox, (const float *)b->radius,
                  (const uint32_t *)b->vcount, b->effects, sh, b->views, b->leaf_desc, b->mollify,
                  b->counters + 2));
    if (b->adaptive)
        VG_TRY(launch("vg_batch_geometry", k_brick_starts, dim3(1), dim3(256), 0, s,
                      (const TreeView *)b->views, (const uint32_t *)b->qorder, b->effects,
                      b->brick_start));
    BatchTreeArgs ta;
    ta.trees = b->effects;
    ta.shape = &b->shape;
    ta.n = b->nv;
    ta.pos = b->vpos;
    ta.angvel = b->vom;
    ta.tree_of = b->vtree;
    ta.leaf_desc = b->leaf_desc;
    ta.size = b->size;
    ta.cell_count = b->cell_count;
    ta.cell_start = b->cell_start;
    ta.keys = b->keys;
    ta.slots = b->slots;
    ta.sorted_idx = b->sorted_idx;
    ta.sorted_p = b->sorted_p;
    ta.sorted_w = b->sorted_w;
    ta.scan_tmp = b->scan_tmp;
    for (uint32_t l = 0; l < b->shape.depth; ++l) {
        ta.P[l] = b->P[l];
        ta.W[l] = b->W[l];
        ta.M[l] = b->M[l];
        ta.C[l] = b->C[l];
    }
    if (b->adaptive) ta.views = b->views;
    ta.all_leaves = b->cfg.near_k < 0;
    VG_TRY(batch_tree_build(ta, s));
    return walk(b, b->walk_waves, s);
}

vg_status frame(vg_batch *b, float dt, hipStream_t s) {
    // tracers first, with the grids of the previous frame, folding their
    // moved positions into the cleared boxes on the way
    VG_TRY(launch("vg_batch_bbox", k_bbox_init, dim3(ceil_div(6 * kBoxSlots * b->effects, 256)),
                  dim3(256), 0, s,
                  b->bbox, b->effects));
    if (b->nt)
        VG_TRY(launch("vg_batch_tracers", k_advect_box_batch,
                      dim3(ceil_div(b->nt, 256 * kAdvRound)), dim3(256), 0, s,
                      (const TreeView *)b->views, (const float *)b->grid, b->cap0, b->tpos,
                      (const uint32_t *)b->ttree, b->nt, dt, b->bbox));
    VG_TRY(solve(b, true, s));
    return advect(b, b->vpos, b->vvel, b->vtree, b->nv, dt, b->adv_block, b->adv_items, s);
}

void drop_graph(vg_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

}  // namespace

using namespace vg;

VG_API vg_status vg_batch_config_default(vg_batch_config *cfg) {
    VG_REQUIRE(cfg, "null argument");
    cfg->cells = 32;
    cfg->near_k = 2;
    cfg->order = 1;
    cfg->rsqrt_policy = VG_RSQRT_ACCURATE;
    cfg->expand_from = -1;
    cfg->max_cells = 64;
    return VG_OK;
}

VG_API vg_status vg_batch_create(vg_batch **out, uint32_t effects, const uint32_t *vorton_counts,
                                 const uint32_t *tracer_counts, const float *vorton_radii,
                                 const vg_batch_config *cfg) {
    VG_REQUIRE(out && vorton_counts && vorton_radii, "null argument");
    VG_REQUIRE(effects > 0, "no effects");
    vg_batch_config c;
    if (cfg) c = *cfg;
    else vg_batch_config_default(&c);
    const bool adaptive = c.cells == 0;
    VG_REQUIRE(adaptive || (c.cells >= 4 && c.cells <= 256), "cells must be 0 or in [4, 256]");
    VG_REQUIRE(!adaptive || (c.max_cells >= 4 && c.max_cells <= 256 &&
                             (c.max_cells & (c.max_cells - 1)) == 0),
               "max_cells must be a power of two in [4, 256]");
    VG_REQUIRE(c.near_k >= -1 && c.near_k <= 16, "near_k must be in [-1, 16]");
    VG_REQUIRE(c.order == 0 || c.order == 1, "order must be 0 or 1");
    VG_REQUIRE(c.rsqrt_policy == VG_RSQRT_REFERENCE || c.rsqrt_policy == VG_RSQRT_ACCURATE,
               "bad rsqrt policy");
    *out = nullptr;
    vg_batch *b = new vg_batch();
    b->effects = effects;
    b->cfg = c;
    auto fail = [&](vg_status st) {
        vg_batch_destroy(b);
        return st;
    };
    // the fixed shape, or the largest adaptive one: (cells + 1)^3 points
    // over a unit box
    b->adaptive = adaptive;
    const uint32_t cells_axis = adaptive ? c.max_cells : c.cells;
    vg_grid_desc tmpl;
    for (int a = 0; a < 3; ++a) {
        tmpl.min[a] = 0.f;
        tmpl.extent[a] = 1.f;
        tmpl.np[a] = cells_axis + 1;
        tmpl.cell[a] = 1.f / (float)cells_axis;
        tmpl.cpe[a] = (float)cells_axis;
    }
    nested_init(&b->shape, tmpl);
    if (has_orphan_cells(b->shape))
        return fail((set_error("cells leave a cell without a parent cluster on some tree layer "
                               "(an odd count above 3 after halving): use a power of two, or "
                               "3 * 2^k"),
                     VG_EINVAL));
    b->cap0 = capacity(b->shape.layer[0]);
    std::vector<uint32_t> vt, tt;
    std::vector<float> rad(effects), sz(effects);
    for (uint32_t e = 0; e < effects; ++e) {
        if (!(vorton_radii[e] > 0.f)) return fail((set_error("vorton radii must be positive"), VG_EINVAL));
        rad[e] = vorton_radii[e];
        sz[e] = 2.0f * vorton_radii[e];
        vt.insert(vt.end(), vorton_counts[e], e);
        if (tracer_counts) tt.insert(tt.end(), tracer_counts[e], e);
    }
    b->nv = vt.size();
    b->nt = tt.size();
    if (b->nv >= (1ull << 31) || b->nt >= (1ull << 34))
        return fail((set_error("too many particles"), VG_EINVAL));
    if ((size_t)effects * b->cap0 >= (1ull << 32))
        return fail((set_error("effects x grid points too large"), VG_EINVAL));
    vg_status st = VG_OK;
#define VG_B(expr)                                   \
    do {                                             \
        st = (expr);                                 \
        if (st != VG_OK) return fail(st);            \
    } while (0)
    VG_B(dalloc(b, &b->vpos, 3 * b->nv));
    VG_B(dalloc(b, &b->vom, 3 * b->nv));
    VG_B(dalloc(b, &b->vvel, 3 * b->nv));
    VG_B(dalloc(b, &b->tpos, 3 * b->nt));
    VG_B(dalloc(b, &b->vtree, b->nv));
    VG_B(dalloc(b, &b->ttree, b->nt));
    VG_B(dalloc(b, &b->radius, effects));
    VG_B(dalloc(b, &b->size, effects));
    for (uint32_t l = 0; l < b->shape.depth; ++l) {
        b->cap[l] = capacity(b->shape.layer[l]);
        VG_B(dalloc(b, &b->P[l], effects * b->cap[l]));
        VG_B(dalloc(b, &b->W[l], effects * b->cap[l]));
        VG_B(dalloc(b, &b->M[l], 3 * effects * b->cap[l]));
        if (l >= 1) VG_B(dalloc(b, &b->C[l], effects * b->cap[l]));
    }
    const size_t cells = (size_t)effects * b->cap0;
    VG_B(dalloc(b, &b->cell_count, cells));
    VG_B(dalloc(b, &b->cell_start, cells));
    VG_B(dalloc(b, &b->keys, b->nv));
    VG_B(dalloc(b, &b->slots, b->nv));
    VG_B(dalloc(b, &b->sorted_idx, b->nv));
    VG_B(dalloc(b, &b->sorted_p, b->nv));
    VG_B(dalloc(b, &b->sorted_w, b->nv));
    VG_B(dalloc(b, &b->scan_tmp, scan_scratch_elems(cells)));
    VG_B(dalloc(b, &b->views, effects));
    VG_B(dalloc(b, &b->leaf_desc, effects));
    VG_B(dalloc(b, &b->mollify, effects));
    VG_B(dalloc(b, &b->bbox, 6 * (size_t)kBoxSlots * effects));
    VG_B(dalloc(b, &b->grid, 3 * cells));
    VG_B(dalloc(b, &b->counters, 4));
    if (adaptive) {
        VG_B(dalloc(b, &b->vcount, effects));
        VG_B(dalloc(b, &b->qorder, effects));
        VG_B(dalloc(b, &b->brick_start, (size_t)effects + 1));
    }
#undef VG_B
    auto up = [&](void *dst, const void *src, size_t bytes) -> vg_status {
        if (bytes) VG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        return VG_OK;
    };
    if ((st = up(b->vtree, vt.data(), vt.size() * 4)) != VG_OK) return fail(st);
    if ((st = up(b->ttree, tt.data(), tt.size() * 4)) != VG_OK) return fail(st);
    if ((st = up(b->radius, rad.data(), effects * 4)) != VG_OK) return fail(st);
    if ((st = up(b->size, sz.data(), effects * 4)) != VG_OK) return fail(st);
    if (adaptive) {
        // the walk's queue takes the effects with the most vortons first
        std::vector<uint32_t> vc(vorton_counts, vorton_counts + effects), order(effects);
        for (uint32_t e = 0; e < effects; ++e) order[e] = e;
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t x, uint32_t y) { return vc[x] > vc[y]; });
        if ((st = up(b->vcount, vc.data(), effects * 4)) != VG_OK) return fail(st);
        if ((st = up(b->qorder, order.data(), effects * 4)) != VG_OK) return fail(st);
    }
    // per-effect views: fixed pointers and decimations, unit-box geometry
    // until the first frame computes the real one
    std::vector<TreeView> views(effects);
    for (uint32_t e = 0; e < effects; ++e) {
        TreeView &v = views[e];
        memset(&v, 0, sizeof v);
        v.depth = b->shape.depth;
        for (uint32_t l = 0; l < kMaxLayers; ++l)
            for (int a = 0; a < 3; ++a) v.dec[l][a] = b->shape.dec[l][a];
        for (uint32_t l = 0; l < b->shape.depth; ++l) {
            v.layer[l] = b->shape.layer[l];
            v.P[l] = b->P[l] + e * b->cap[l];
            v.W[l] = b->W[l] + e * b->cap[l];
            v.M[l] = b->M[l] + 3 * e * b->cap[l];
            v.C[l] = l >= 1 ? b->C[l] + e * b->cap[l] : nullptr;
        }
        v.cell_start = b->cell_start + e * b->cap0;
        v.cell_count = b->cell_count + e * b->cap0;
        v.sorted_p = b->sorted_p;
        v.sorted_w = b->sorted_w;
        v.sorted_idx = b->sorted_idx;
        v.n = (uint32_t)b->nv;
    }
    if ((st = up(b->views, views.data(), effects * sizeof(TreeView))) != VG_OK) return fail(st);
    hipError_t e = hipMemset(b->grid, 0, 3 * cells * sizeof(float));
    if (e == hipSuccess) e = hipMemset(b->counters, 0, 4 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b->vvel, 0, 3 * (b->nv ? b->nv : 1) * sizeof(float));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(device_error(e, "batch init", __FILE__, __LINE__));
    *out = b;
    return VG_OK;
}

VG_API vg_status vg_batch_destroy(vg_batch *b) {
    if (!b) return VG_OK;
    drop_graph(b);
    for (void *p : b->allocs) (void)hipFree(p);
    delete b;
    return VG_OK;
}

VG_API vg_status vg_batch_buffers(vg_batch *b, float **vorton_pos, float **vorton_angvel,
                                  float **vorton_vel, float **tracer_pos, size_t *n_vortons,
                                  size_t *n_tracers) {
    VG_REQUIRE(b, "null batch");
    if (vorton_pos) *vorton_pos = b->vpos;
    if (vorton_angvel) *vorton_angvel = b->vom;
    if (vorton_vel) *vorton_vel = b->vvel;
    if (tracer_pos) *tracer_pos = b->tpos;
    if (n_vortons) *n_vortons = b->nv;
    if (n_tracers) *n_tracers = b->nt;
    return VG_OK;
}

VG_API vg_status vg_batch_set_method(vg_batch *b, int near_k, int order, int expand_from) {
    VG_REQUIRE(b, "null batch");
    VG_REQUIRE(near_k >= -1 && near_k <= 16, "near_k must be in [-1, 16]");
    VG_REQUIRE(order == 0 || order == 1, "order must be 0 or 1");
    drop_graph(b);
    b->cfg.near_k = near_k;
    b->cfg.order = order;
    b->cfg.expand_from = expand_from;
    return VG_OK;
}

VG_API vg_status vg_batch_solve(vg_batch *b, vg_stream s) {
    VG_REQUIRE(b, "null batch");
    VG_TRY(solve(b, false, (hipStream_t)s));
    b->stepped = true;
    return VG_OK;
}

VG_API vg_status vg_batch_sort_tracers(vg_batch *b, vg_stream stream) {
    VG_REQUIRE(b, "null batch");
    if (!b->stepped) {
        set_error("no velocity grids yet (step or solve the batch first)");
        return VG_ESTATE;
    }
    if (b->nt == 0) return VG_OK;
    const size_t cells = (size_t)b->effects * b->cap0;
    if (!b->ts_count) {
        vg_status st = VG_OK;
        if (st == VG_OK) st = dalloc(b, &b->ts_count, cells);
        if (st == VG_OK) st = dalloc(b, &b->ts_start, cells);
        if (st == VG_OK) st = dalloc(b, &b->ts_keys, b->nt);
        if (st == VG_OK) st = dalloc(b, &b->ts_slots, b->nt);
        if (st == VG_OK) st = dalloc(b, &b->ts_unsorted, b->nt);
        if (st == VG_OK) st = dalloc(b, &b->ts_sorted, b->nt);
        if (st == VG_OK) st = dalloc(b, &b->ts_scan, scan_scratch_elems(cells));
        if (st == VG_OK) st = dalloc(b, &b->ts_moved, 3 * b->nt);
        if (st != VG_OK) {
            b->ts_count = nullptr;  // the partial buffers stay in allocs until destroy
            return st;
        }
    }
    BatchSortArgs a;
    a.n = b->nt;
    a.cap0 = b->cap0;
    a.cells = cells;
    a.pos = b->tpos;
    a.tree_of = b->ttree;
    a.leaf_desc = b->leaf_desc;
    a.count = b->ts_count;
    a.start = b->ts_start;
    a.keys = b->ts_keys;
    a.slots = b->ts_slots;
    a.unsorted = b->ts_unsorted;
    a.sorted = b->ts_sorted;
    a.scan_tmp = b->ts_scan;
    a.moved = b->ts_moved;
    return batch_sort_particles(a, (hipStream_t)stream);
}

VG_API vg_status vg_batch_step(vg_batch *b, float dt, vg_stream s) {
    VG_REQUIRE(b, "null batch");
    VG_TRY(frame(b, dt, (hipStream_t)s));
    b->stepped = true;
    return VG_OK;
}

VG_API vg_status vg_batch_capture(vg_batch *b, float dt, vg_stream stream) {
    VG_REQUIRE(b, "null batch");
    VG_REQUIRE(stream, "capture needs a created stream (not the legacy stream)");
    hipStream_t s = (hipStream_t)stream;
    drop_graph(b);
    const bool prof = profiling_enabled();
    if (prof) vg_profile_enable(0);
    VG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    vg_status st = frame(b, dt, s);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (prof) vg_profile_enable(1);
    if (st != VG_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return device_error(e, "hipStreamEndCapture", __FILE__, __LINE__);
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(b);
        return device_error(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    return VG_OK;
}

VG_API vg_status vg_batch_launch(vg_batch *b, vg_stream s) {
    VG_REQUIRE(b, "null batch");
    if (!b->exec) {
        set_error("no recorded frame (call vg_batch_capture first)");
        return VG_ESTATE;
    }
    VG_HIP(hipGraphLaunch(b->exec, (hipStream_t)s));
    b->stepped = true;
    return VG_OK;
}

VG_API vg_status vg_batch_grid(const vg_batch *b, uint32_t effect, vg_grid_desc *g,
                               const float **field) {
    VG_REQUIRE(b && g, "null argument");
    VG_REQUIRE(effect < b->effects, "effect out of range");
    VG_HIP(hipDeviceSynchronize());
    VG_HIP(hipMemcpy(g, b->leaf_desc + effect, sizeof *g, hipMemcpyDeviceToHost));
    if (field) *field = b->grid + 3 * b->cap0 * effect;
    return VG_OK;
}

VG_API vg_status vg_batch_clamped(const vg_batch *b, uint32_t *count) {
    VG_REQUIRE(b && count, "null argument");
    VG_HIP(hipDeviceSynchronize());
    VG_HIP(hipMemcpy(count, b->counters + 2, sizeof(uint32_t), hipMemcpyDeviceToHost));
    return VG_OK;
}

VG_API vg_status vg_batch_mollify(const vg_batch *b, uint32_t effect, float *spread, float *circ) {
    VG_REQUIRE(b && spread && circ, "null argument");
    VG_REQUIRE(effect < b->effects, "effect out of range");
    float2 m;
    VG_HIP(hipDeviceSynchronize());
    VG_HIP(hipMemcpy(&m, b->mollify + effect, sizeof m, hipMemcpyDeviceToHost));
    *spread = m.x;
    *circ = m.y;
    return VG_OK;
}

VG_API vg_status vg_batch_set_launch(vg_batch *b, int walk_waves, int tracer_block,
                                     int tracer_items) {
    VG_REQUIRE(b, "null batch");
    VG_REQUIRE(walk_waves == 1 || walk_waves == 2 || walk_waves == 4, "walk_waves in {1, 2, 4}");
    VG_REQUIRE(tracer_block >= 64 && tracer_block <= 1024 && tracer_block % 64 == 0,
               "tracer_block must be a multiple of 64 in [64, 1024]");
    VG_REQUIRE(tracer_items == 1 || tracer_items == 2 || tracer_items == 4,
               "tracer_items in {1, 2, 4}");
    b->walk_waves = walk_waves;
    b->adv_block = tracer_block;
    b->adv_items = tracer_items;
    drop_graph(b);
    return VG_OK;
}

VG_API vg_status vg_batch_get_launch(const vg_batch *b, int *walk_waves, int *tracer_block,
                                     int *tracer_items) {
    VG_REQUIRE(b && walk_waves && tracer_block && tracer_items, "null argument");
    *walk_waves = b->walk_waves;
    *tracer_block = b->adv_block;
    *tracer_items = b->adv_items;
    return VG_OK;
}

// Times each candidate launch shape of the two dominant kernels on the live
// batch.  The walk trials rewrite the grids with the values they already
// hold (same trees, same arithmetic); the tracer trials move a scratch copy
// of the positions.
VG_API vg_status vg_batch_autotune(vg_batch *b, vg_stream stream, char *report, size_t cap) {
    VG_REQUIRE(b, "null batch");
    if (!b->stepped) {
        set_error("autotune needs a computed frame (call vg_batch_step first)");
        return VG_ESTATE;
    }
    hipStream_t s = (hipStream_t)stream;
    drop_graph(b);
    hipEvent_t e0, e1;
    VG_HIP(hipEventCreate(&e0));
    VG_HIP(hipEventCreate(&e1));
    std::string rep;
    auto timed = [&](auto &&fn, float *ms) -> vg_status {
        VG_TRY(fn());  // warm
        VG_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < 3; ++r) VG_TRY(fn());
        VG_HIP(hipEventRecord(e1, s));
        VG_HIP(hipEventSynchronize(e1));
        VG_HIP(hipEventElapsedTime(ms, e0, e1));
        *ms /= 3.0f;
        return VG_OK;
    };
    vg_status st = VG_OK;
    // walk: waves per block
    float best = FLT_MAX;
    int best_w = b->walk_waves;
    for (int w : {1, 2, 4}) {
        float ms = 0.f;
        st = timed([&]() { return walk(b, w, s); }, &ms);
        if (st != VG_OK) break;
        char line[96];
        snprintf(line, sizeof line, "walk waves_per_block=%d: %.3f ms\n", w, ms);
        rep += line;
        if (ms < best) { best = ms; best_w = w; }
    }
    // vorton advection: block size x items per thread, on scratch copies of
    // positions and velocities (the tracers go through the fused
    // advection + bounds pass, which has one shape)
    float *scratch = nullptr;
    int best_b = b->adv_block, best_i = b->adv_items;
    if (st == VG_OK && b->nv) {
        st = salloc((void **)&scratch, 6 * b->nv * sizeof(float), s);
        best = FLT_MAX;
        for (int blk : {128, 256, 512, 1024}) {
            for (int it : {1, 2, 4}) {
                if (st != VG_OK) break;
                float ms = 0.f;
                hipError_t he = hipMemcpyAsync(scratch, b->vpos, 3 * b->nv * sizeof(float),
                                               hipMemcpyDeviceToDevice, s);
                if (he != hipSuccess) { st = device_error(he, "autotune copy", __FILE__, __LINE__); break; }
                st = timed([&]() {
                    return advect(b, scratch, scratch + 3 * b->nv, b->vtree, b->nv, 0.0f, blk, it, s);
                }, &ms);
                char line[96];
                snprintf(line, sizeof line, "vortons block=%d items=%d: %.3f ms\n", blk, it, ms);
                rep += line;
                if (ms < best) { best = ms; best_b = blk; best_i = it; }
            }
        }
        if (scratch) sfree(scratch, s);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != VG_OK) return st;
    b->walk_waves = best_w;
    b->adv_block = best_b;
    b->adv_items = best_i;
    char line[128];
    snprintf(line, sizeof line, "chosen: walk waves_per_block=%d, vortons block=%d items=%d\n",
             best_w, best_b, best_i);
    rep += line;
    if (report && cap) {
        snprintf(report, cap, "%s", rep.c_str());
    }
    return VG_OK;
}
