// vg_tree.hip -- the NestedGrid influence tree (VortonSim::CreateInfluenceTree,
// vortonSim.cpp:1192-1624) and everything that walks it: treecode velocity
// and vector potential (:1650-1905), boundary-face vector potential
// (:2055-2138), and the vorton -> grid splats that reuse its leaf cell lists
// (PopulateVorticityGridFromVortons :2372-2436, Particles::
// PopulateVelocityGrid pclOpPopulateVelocityGrid.cpp:33-87).
//
// Build (all deterministic, no float atomics):
//   bin     vorton -> leaf cell key (truncation, as OffsetOfPosition),
//           slot = atomicAdd(count[key]), one atomic per run of equal keys
//           in a wave
//   scan    count -> cell start (two-level exclusive scan)
//   scatter start + slot = i into a cell-grouped list, then each vorton
//           moves to its rank among its cell's indices (k_rank_cells), so
//           cells hold vortons in index order
//   leaf    per leaf point: sequential |w|-weighted sums in index order
//           (bit-identical to MakeBaseVortonGrid's serial loop)
//   levels  per parent point: AggregateClusters' 2x2x2 sums in z,y,x order
// Layers are dense over grid points (SoA float4 pairs), like the reference's
// UniformGrid<Vorton> layers, so traversal is pure index arithmetic.
#include <cstring>

#include "vg_internal.h"

struct vg_tree {
    vg::Nested ng;
    float4 *P[vg::kMaxLayers] = {};
    float4 *W[vg::kMaxLayers] = {};
    float4 *M[vg::kMaxLayers] = {};  // 3 float4 per cluster: first moment rows + antisymmetric part
    uint32_t *C[vg::kMaxLayers] = {};  // non-empty child masks (layers >= 1)
    size_t c_cap[vg::kMaxLayers] = {};
    size_t m_cap[vg::kMaxLayers] = {};
    size_t layer_cap[vg::kMaxLayers] = {};
    uint32_t *cell_start = nullptr, *cell_count = nullptr;
    size_t cell_cap = 0;
    uint32_t *keys = nullptr, *slots = nullptr, *sorted_idx = nullptr;
    float4 *sorted_p = nullptr, *sorted_w = nullptr;
    size_t n_cap = 0;
    uint32_t *scan_tmp = nullptr;
    size_t scan_cap = 0;
    uint32_t n = 0;
    bool built = false;
};

namespace vg {
namespace {

// ------------------------------------------------------------------ scan
constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;  // elements per thread
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// Exclusive scan of one tile per block; block totals to sums (if non-null).
__global__ __launch_bounds__(kScanBlock) void k_scan_tile(const uint32_t *__restrict__ in,
                                                          uint32_t *__restrict__ out, size_t n,
                                                          uint32_t *__restrict__ sums) {
    __shared__ uint32_t wsum[kScanBlock / 64];
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t tot = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        tot += v[k];
    }
    const uint32_t incl = wave_incl_scan(tot);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t woff = 0;
    for (int w = 0; w < wave; ++w) woff += wsum[w];
    uint32_t run = woff + incl - tot;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (sums && threadIdx.x == kScanBlock - 1) sums[blockIdx.x] = run;
}

__global__ void k_scan_add(uint32_t *__restrict__ out, size_t n,
                           const uint32_t *__restrict__ offs) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += offs[i / kScanTile];
}

size_t scan_elems(size_t n) {
    size_t total = 0;
    while (n > (size_t)kScanTile) {
        n = (n + kScanTile - 1) / kScanTile;
        total += 2 * n;
    }
    return total + 1;
}

// Exclusive scan in -> out using tmp (scan_scratch_elems(n) elements).
vg_status exclusive_scan(const uint32_t *in, uint32_t *out, size_t n, uint32_t *tmp,
                         hipStream_t s) {
    if (n == 0) return VG_OK;
    const uint32_t nb = ceil_div(n, kScanTile);
    if (nb == 1)
        return launch("vg_scan_tile", k_scan_tile, dim3(1), dim3(kScanBlock), 0, s, in, out, n,
                      (uint32_t *)nullptr);
    uint32_t *sums = tmp, *sums_scan = tmp + nb;
    VG_TRY(launch("vg_scan_tile", k_scan_tile, dim3(nb), dim3(kScanBlock), 0, s, in, out, n, sums));
    VG_TRY(exclusive_scan(sums, sums_scan, nb, tmp + 2 * nb, s));
    return launch("vg_scan_add", k_scan_add, dim3(ceil_div(n, 256)), dim3(256), 0, s, out, n,
                  (const uint32_t *)sums_scan);
}

// ------------------------------------------------------------------ build
struct VortonIn {
    const float *pos, *angvel, *size;
    size_t stride;
    float usize;
};

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// A slot in cell `key` for each live lane of the wave.  Lanes holding the
// same key next to each other (particles already in or near cell order) form
// a run that takes its slots with ONE atomic by the run's first lane; the
// others add their distance from it.  Slots inside a cell are then unique
// but arbitrary, as with one atomic per particle: the cell sort puts them in
// index order afterwards.  Every lane of the wave must call this (dead lanes
// with live = false, after all live lanes).
__device__ __forceinline__ uint32_t bin_slot(uint32_t *__restrict__ count, uint32_t key,
                                             bool live) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t k = live ? key : 0xffffffffu;
    const uint32_t prev = (uint32_t)__shfl_up((int)k, 1, 64);
    const bool head = lane == 0 || prev != k;
    const uint64_t heads = __ballot(head);
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const uint32_t first = 63u - (uint32_t)__clzll(heads & upto);
    const uint64_t after = heads & ~upto;
    const uint32_t next = after ? (uint32_t)__ffsll((long long)after) - 1u : 64u;
    uint32_t base = 0;
    if (head && live) base = atomicAdd(&count[key], next - lane);
    return (uint32_t)__shfl((int)base, (int)first, 64) + (lane - first);
}

// MakeBaseVortonGrid binning: OffsetOfPosition (uniformGrid.h:645-651).
__global__ void k_bin(VortonIn vin, size_t n, vg_grid_desc g, uint32_t *__restrict__ count,
                      uint32_t *__restrict__ keys, uint32_t *__restrict__ slots) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    uint32_t key = 0;
    if (live) {
        const float *p = vin.pos + i * vin.stride;
        const int ix = cell_of(p[0], g.min[0], g.cpe[0], (int)g.np[0] - 1);
        const int iy = cell_of(p[1], g.min[1], g.cpe[1], (int)g.np[1] - 1);
        const int iz = cell_of(p[2], g.min[2], g.cpe[2], (int)g.np[2] - 1);
        key = (uint32_t)point_offset(g, ix, iy, iz);
    }
    const uint32_t slot = bin_slot(count, key, live);
    if (!live) return;
    keys[i] = key;
    slots[i] = slot;
}

__global__ void k_scatter(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ slots,
                          const uint32_t *__restrict__ start, size_t n,
                          uint32_t *__restrict__ sorted_idx) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) sorted_idx[start[keys[i]] + slots[i]] = (uint32_t)i;
}

// Restore index order inside each cell (the atomic slots are arbitrary):
// vorton v, found at position p of its cell's segment of `unsorted`, goes to
// the position given by its rank, the number of smaller indices in the cell.
// One thread per vorton; the indices are distinct, so the ranks are too.
__global__ void k_rank_cells(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ start,
                             const uint32_t *__restrict__ count,
                             const uint32_t *__restrict__ unsorted, size_t n,
                             uint32_t *__restrict__ sorted_idx) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t v = unsorted[p];
    const uint32_t c = keys[v];
    const uint32_t b = start[c], e = b + count[c];
    uint32_t r = 0;
    for (uint32_t q = b; q < e; ++q) r += unsorted[q] < v ? 1u : 0u;
    sorted_idx[b + r] = v;
}

__global__ void k_gather_sorted(VortonIn vin, const uint32_t *__restrict__ sorted_idx, size_t n,
                                float4 *__restrict__ sp, float4 *__restrict__ sw) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const size_t i = sorted_idx[j];
    const float *p = vin.pos + i * vin.stride;
    const float *w = vin.angvel + i * vin.stride;
    const float sz = vin.size ? vin.size[i * vin.stride] : vin.usize;
    sp[j] = make_float4(p[0], p[1], p[2], sz);
    sw[j] = make_float4(w[0], w[1], w[2], 0.f);
}

// MakeBaseVortonGrid accumulation + normalisation (vortonSim.cpp:1237-1310).
// views (may be null): trees of their own shapes, blockIdx.y = tree, cap
// the per-tree stride; only each tree's own leaf points are written.
// skip_empty: cells without vortons are not written (their previous values
// stay); only for trees whose readers reach leaf clusters through the cell
// counts or the parents' child masks (the batched walk), never by scanning.
// kLeafSkipUnlessRead: per tree of `views`, skip them unless the
// reference-rule walk of that tree scans leaf clusters (reads_empty_leaves).
constexpr int kLeafSkipUnlessRead = 2;

// Whether the reference-rule walk (tree_eval_split) reads the clusters of
// empty leaf cells.  It reads layer 0 only as the children of layer-1
// parents: through the parents' child masks when layer 1 decimates by 2 on
// every axis (at most 2 parents per axis are ever open, so layer_masked
// applies and visits non-empty children only), by enumeration otherwise.  A
// single-layer tree sums the sorted vortons and reads no clusters.
__host__ __device__ inline bool reads_empty_leaves(uint32_t depth, const uint32_t *dec1) {
    return depth >= 2 && !(dec1[0] == 2 && dec1[1] == 2 && dec1[2] == 2);
}

__global__ void k_leaf(const uint32_t *__restrict__ start, const uint32_t *__restrict__ count,
                       const float4 *__restrict__ sp, const float4 *__restrict__ sw, size_t cap,
                       float4 *__restrict__ P, float4 *__restrict__ W, float4 *__restrict__ M,
                       const TreeView *__restrict__ views, int skip_empty) {
#pragma clang fp contract(off)
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (views) {
        const TreeView &tv = views[blockIdx.y];
        if (c >= capacity(tv.layer[0])) return;
        if (skip_empty == kLeafSkipUnlessRead)
            skip_empty = reads_empty_leaves(tv.depth, tv.dec[1]) ? 0 : 1;
        c += (size_t)blockIdx.y * cap;
    } else if (c >= cap) {
        return;
    }
    const uint32_t cnt = count[c];
    if (skip_empty && cnt == 0) return;
    float px = 0.f, py = 0.f, pz = 0.f, wx = 0.f, wy = 0.f, wz = 0.f, sz = 0.f;
    float aux = FLT_MIN;
    const uint32_t b = cnt ? start[c] : 0u;
    for (uint32_t k = 0; k < cnt; ++k) {
        const float4 p = sp[b + k], w = sw[b + k];
        const float vx = w.x * 2.0f, vy = w.y * 2.0f, vz = w.z * 2.0f;
        const float mag = sqrtf(vx * vx + vy * vy + vz * vz);
        px += p.x * mag; py += p.y * mag; pz += p.z * mag;
        wx += w.x; wy += w.y; wz += w.z;
        sz = p.w;
        aux += mag;
    }
    if (aux != FLT_MIN) {
        const float inv = 1.0f / aux;
        px *= inv; py *= inv; pz *= inv;
    }
    P[c] = make_float4(px, py, pz, sz);
    W[c] = make_float4(wx, wy, wz, aux);
    // First moment about the centroid, D[b][e] = sum Omega_b (y - c)_e, in
    // index order (no reference counterpart; used by the near-field mode).
    float D[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (uint32_t k = 0; k < cnt; ++k) {
        const float4 p = sp[b + k], w = sw[b + k];
        const float e[3] = {p.x - px, p.y - py, p.z - pz};
        const float om[3] = {w.x, w.y, w.z};
#pragma unroll
        for (int bb = 0; bb < 3; ++bb)
#pragma unroll
            for (int a = 0; a < 3; ++a) D[3 * bb + a] += om[bb] * e[a];
    }
    M[3 * c] = make_float4(D[0], D[1], D[2], D[5] - D[7]);
    M[3 * c + 1] = make_float4(D[3], D[4], D[5], D[6] - D[2]);
    M[3 * c + 2] = make_float4(D[6], D[7], D[8], D[1] - D[3]);
}

// Running sums of one parent cluster.  Both steps are shared by the general
// child loop and the 2x2x2 form below, so the two give identical bits.
struct ParentSums {
    float px = 0.f, py = 0.f, pz = 0.f, wx = 0.f, wy = 0.f, wz = 0.f, sz = 0.f;
    float aux = FLT_MIN;
    uint32_t cmask = 0;
    float D[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
};

// Child with position p and vorticity w adds to the magnitude-weighted
// centroid, the summed vorticity and the child mask (bit = child number).
__device__ __forceinline__ void parent_add_child(ParentSums &a, float4 p, float4 w, uint32_t bit) {
#pragma clang fp contract(off)
    const float vx = w.x * 2.0f, vy = w.y * 2.0f, vz = w.z * 2.0f;
    const float mag = sqrtf(vx * vx + vy * vy + vz * vz);
    a.px += p.x * mag; a.py += p.y * mag; a.pz += p.z * mag;
    a.wx += w.x; a.wy += w.y; a.wz += w.z;
    a.aux += mag;
    if (p.w != 0.0f) {
        a.sz = p.w;
        a.cmask |= 1u << bit;
    }
}

// First moment about the parent centroid cp: D_p += D_c + Omega_c (c_c - c_p).
__device__ __forceinline__ void parent_add_moment(ParentSums &a, float4 p, float4 w, float4 m0,
                                                  float4 m1, float4 m2, const float cp[3]) {
#pragma clang fp contract(off)
    const float cd[9] = {m0.x, m0.y, m0.z, m1.x, m1.y, m1.z, m2.x, m2.y, m2.z};
    const float cw[3] = {w.x, w.y, w.z}, cc[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int a3 = 0; a3 < 3; ++a3)
            a.D[3 * bb + a3] += cd[3 * bb + a3] + cw[bb] * (cc[a3] - cp[a3]);
}

// AggregateClusters (vortonSim.cpp:1337-1569): parent point o of layer gp
// from its dx x dy x dz children in layer gc.
// cc (may be null): vorton counts of the child cells; children with none are
// not added (they would add exact zeros; the 2x2x2 form loads their clusters
// with the others' and drops them), so layer-0 clusters of empty cells need
// not be written (k_leaf skip_empty).
__device__ __forceinline__ void aggregate_point(const vg_grid_desc &gp, const vg_grid_desc &gc,
                                                uint32_t dx, uint32_t dy, uint32_t dz, size_t o,
                                                const float4 *__restrict__ Pc,
                                                const float4 *__restrict__ Wc,
                                                const float4 *__restrict__ Mc,
                                                const uint32_t *__restrict__ cc,
                                                float4 *__restrict__ Pp, float4 *__restrict__ Wp,
                                                float4 *__restrict__ Mp,
                                                uint32_t *__restrict__ Cp) {
#pragma clang fp contract(off)
    const uint32_t ix = (uint32_t)(o % gp.np[0]);
    const uint32_t iy = (uint32_t)((o / gp.np[0]) % gp.np[1]);
    const uint32_t iz = (uint32_t)(o / ((size_t)gp.np[0] * gp.np[1]));
    ParentSums a;
    if (ix < gp.np[0] - 1 && iy < gp.np[1] - 1 && iz < gp.np[2] - 1) {
        if (dx == 2 && dy == 2 && dz == 2) {
            // Eight children: their loads are all issued before the sums
            // start (the general loop waits on one child's loads per trip,
            // which sets the time of every small layer).  Children of an
            // empty cell are loaded (in bounds, possibly never written) but
            // not added.
            const size_t base = point_offset(gc, ix * 2, iy * 2, iz * 2);
            const size_t sy = gc.np[0], sz2 = (size_t)gc.np[0] * gc.np[1];
            size_t co[8];
            bool live[8];
            float4 p[8], w[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) co[k] = base + (k & 1) + ((k >> 1) & 1) * sy + (k >> 2) * sz2;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                live[k] = !cc || cc[co[k]] != 0;
                p[k] = Pc[co[k]];
                w[k] = Wc[co[k]];
            }
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (live[k]) parent_add_child(a, p[k], w[k], (uint32_t)k);
            const float inv = 1.0f / a.aux;
            a.px *= inv; a.py *= inv; a.pz *= inv;
            const float cp[3] = {a.px, a.py, a.pz};
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const float4 m0 = Mc[3 * co[k]], m1 = Mc[3 * co[k] + 1], m2 = Mc[3 * co[k] + 2];
                if (live[k]) parent_add_moment(a, p[k], w[k], m0, m1, m2, cp);
            }
        } else {
            for (uint32_t cz = 0; cz < dz; ++cz)
                for (uint32_t cy = 0; cy < dy; ++cy)
                    for (uint32_t cx = 0; cx < dx; ++cx) {
                        const size_t co = point_offset(gc, ix * dx + cx, iy * dy + cy, iz * dz + cz);
                        if (cc && cc[co] == 0) continue;
                        parent_add_child(a, Pc[co], Wc[co], cx + dx * (cy + dy * cz));
                    }
            const float inv = 1.0f / a.aux;
            a.px *= inv; a.py *= inv; a.pz *= inv;
            const float cp[3] = {a.px, a.py, a.pz};
            for (uint32_t cz = 0; cz < dz; ++cz)
                for (uint32_t cy = 0; cy < dy; ++cy)
                    for (uint32_t cx = 0; cx < dx; ++cx) {
                        const size_t co = point_offset(gc, ix * dx + cx, iy * dy + cy, iz * dz + cz);
                        if (cc && cc[co] == 0) continue;
                        parent_add_moment(a, Pc[co], Wc[co], Mc[3 * co], Mc[3 * co + 1],
                                          Mc[3 * co + 2], cp);
                    }
        }
    } else {
        a.aux = 0.f;
    }
    const float *D = a.D;
    Pp[o] = make_float4(a.px, a.py, a.pz, a.sz);
    Wp[o] = make_float4(a.wx, a.wy, a.wz, a.aux);
    Cp[o] = a.cmask;
    Mp[3 * o] = make_float4(D[0], D[1], D[2], D[5] - D[7]);
    Mp[3 * o + 1] = make_float4(D[3], D[4], D[5], D[6] - D[2]);
    Mp[3 * o + 2] = make_float4(D[6], D[7], D[8], D[1] - D[3]);
}

// One parent layer of one tree, or (blockIdx.y) of several trees of this
// shape stored back to back.
__global__ void k_aggregate(vg_grid_desc gp, vg_grid_desc gc, uint32_t dx, uint32_t dy,
                            uint32_t dz, const float4 *__restrict__ Pc,
                            const float4 *__restrict__ Wc, const float4 *__restrict__ Mc,
                            const uint32_t *__restrict__ cc, float4 *__restrict__ Pp,
                            float4 *__restrict__ Wp, float4 *__restrict__ Mp,
                            uint32_t *__restrict__ Cp) {
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t cap = capacity(gp);
    if (o >= cap) return;
    const size_t tp = (size_t)blockIdx.y * cap, tc = (size_t)blockIdx.y * capacity(gc);
    aggregate_point(gp, gc, dx, dy, dz, o, Pc + tc, Wc + tc, Mc + 3 * tc, cc ? cc + tc : nullptr,
                    Pp + tp, Wp + tp, Mp + 3 * tp, Cp + tp);
}

// Parent layer l of trees of different shapes (blockIdx.y = tree), each
// read from its view; storage strides are the per-layer capacities cp, cc.
__global__ void k_aggregate_views(const TreeView *__restrict__ views, uint32_t l, size_t cp,
                                  size_t ccap, const float4 *__restrict__ Pc,
                                  const float4 *__restrict__ Wc, const float4 *__restrict__ Mc,
                                  const uint32_t *__restrict__ cc, float4 *__restrict__ Pp,
                                  float4 *__restrict__ Wp, float4 *__restrict__ Mp,
                                  uint32_t *__restrict__ Cp) {
    const TreeView &tv = views[blockIdx.y];
    if (l >= tv.depth) return;
    const vg_grid_desc &gp = tv.layer[l];
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= capacity(gp)) return;
    const size_t tp = (size_t)blockIdx.y * cp, tc = (size_t)blockIdx.y * ccap;
    aggregate_point(gp, tv.layer[l - 1], tv.dec[l][0], tv.dec[l][1], tv.dec[l][2], o, Pc + tc,
                    Wc + tc, Mc + 3 * tc, cc ? cc + tc : nullptr, Pp + tp, Wp + tp, Mp + 3 * tp,
                    Cp + tp);
}

template <typename T>
vg_status grow(T **p, size_t *cap, size_t need) {
    if (need <= *cap && *p) return VG_OK;
    if (*p) VG_HIP(hipFree(*p));
    *p = nullptr;
    *cap = 0;
    size_t alloc = need + need / 8 + 64;
    VG_HIP(hipMalloc((void **)p, alloc * sizeof(T)));
    *cap = alloc;
    return VG_OK;
}

// ------------------------------------------------------------------ traversal
struct QuerySrc {
    const float *q;
    size_t stride;
    size_t nq;
    vg_grid_desc g;  // grid/face queries
};

// Reference opening rule (vortonSim.cpp:1819-1856): a child cell is opened
// when the query lies in [min - m, max + m) on every axis, m = 1e-4 * child
// spacing; all other children are applied as supervortons; the children of
// opened layer-1 cells (the leaves) are always applied.  The opened set on a
// layer is a box of at most 2 cells per axis, so the walk is a loop over
// layers with no stack.  Sum order differs from the recursion (rounding
// only); opening decisions are bit-identical.
template <int POLICY, int VECPOT, int COUNT = 0>
__device__ float3 tree_eval(const TreeView &tv, float3 q, float spread, uint32_t *count = nullptr) {
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    float3 sum = make_float3(0.f, 0.f, 0.f), comp = sum;
    const float qa[3] = {q.x, q.y, q.z};
    for (int il = (int)tv.depth - 1; il >= 1; --il) {
        const vg_grid_desc &gc = tv.layer[il - 1];
        const uint32_t *dims = tv.dec[il];
        int c0[3], c1[3], olo[3], ohi[3];
        bool open_any = il > 1;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            c0[a] = lo[a] * (int)dims[a];
            c1[a] = hi[a] * (int)dims[a] + (int)dims[a] - 1;
            olo[a] = 1 << 30;
            ohi[a] = -1;
        }
        if (open_any) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float cell = gc.cell[a];
                float m;
                {
#pragma clang fp contract(off)
                    m = 0.0001f * cell;
                    if (a == 2 && 0.0f == cell) m = m + FLT_MIN;
                }
                for (int c = c0[a]; c <= c1[a]; ++c) {
                    const float mn = madd_exact(gc.min[a], (float)c, cell);
                    const float mx = madd_exact(gc.min[a], (float)(c + 1), cell);
                    bool in;
                    {
#pragma clang fp contract(off)
                        in = qa[a] >= mn - m && qa[a] < mx + m;
                    }
                    if (in) {
                        olo[a] = c < olo[a] ? c : olo[a];
                        ohi[a] = c > ohi[a] ? c : ohi[a];
                    }
                }
                if (ohi[a] < olo[a]) open_any = false;
            }
        }
        const float4 *__restrict__ P = tv.P[il - 1];
        const float4 *__restrict__ W = tv.W[il - 1];
        float3 part = make_float3(0.f, 0.f, 0.f);
        for (int cz = c0[2]; cz <= c1[2]; ++cz)
            for (int cy = c0[1]; cy <= c1[1]; ++cy)
                for (int cx = c0[0]; cx <= c1[0]; ++cx) {
                    if (open_any && cx >= olo[0] && cx <= ohi[0] && cy >= olo[1] &&
                        cy <= ohi[1] && cz >= olo[2] && cz <= ohi[2])
                        continue;
                    const size_t off = point_offset(gc, cx, cy, cz);
                    const float4 p = P[off];
                    if (p.w == 0.0f) continue;  // empty cluster: zero circulation
                    if (COUNT) {
                        ++*count;
                        continue;
                    }
                    const float4 w = W[off];
                    const float radius = p.w * 0.5f * spread;
                    const float r2 = radius * radius;
                    const float s = (2.0f / 3.0f) * r2 * radius;
                    const float4 A = make_float4(p.x, p.y, p.z, r2);
                    const float4 B =
                        make_float4(w.x * s, w.y * s, w.z * s, __builtin_amdgcn_rcpf(r2 * radius));
                    if (VECPOT) pair_vecpot<POLICY>(part, q, A, B);
                    else pair_velocity<POLICY>(part, q, A, B);
                }
        kahan_add(sum, comp, part);
        if (!open_any) break;
#pragma unroll
        for (int a = 0; a < 3; ++a) { lo[a] = olo[a]; hi[a] = ohi[a]; }
    }
    return sum;
}

// Direct fallback when the tree has a single layer (vortonSim.cpp:2265-2277),
// over the cell-sorted vortons [begin, end).
template <int POLICY, int VECPOT>
__device__ float3 sorted_direct_range(const TreeView &tv, uint32_t begin, uint32_t end, float3 q,
                                      float spread) {
    float3 sum = make_float3(0.f, 0.f, 0.f), comp = sum;
    float3 part = sum;
    for (uint32_t j = begin; j < end; ++j) {
        float4 A, B;
        prepare_source(tv.sorted_p[j], tv.sorted_w[j], spread, A, B);
        if (VECPOT) pair_vecpot<POLICY>(part, q, A, B);
        else pair_velocity<POLICY>(part, q, A, B);
        if (((j - begin) & 255u) == 255u) {
            kahan_add(sum, comp, part);
            part = make_float3(0.f, 0.f, 0.f);
        }
    }
    kahan_add(sum, comp, part);
    return sum;
}

template <int POLICY, int VECPOT>
__device__ float3 sorted_direct(const TreeView &tv, float3 q, float spread) {
    return sorted_direct_range<POLICY, VECPOT>(tv, 0u, tv.n, q, spread);
}

template <int POLICY, int VECPOT, int SPLIT>
__device__ float3 tree_eval_split(const TreeView &tv, float3 q, float spread, uint32_t sub);

// One lane per query point: the walk of tree_eval_split with SPLIT = 1.
template <int POLICY, int VECPOT>
__device__ __forceinline__ float3 eval_query(const TreeView &tv, float3 q, float spread) {
    return tv.depth > 1 ? tree_eval_split<POLICY, VECPOT, 1>(tv, q, spread, 0u)
                        : sorted_direct<POLICY, VECPOT>(tv, q, spread);
}

template <int POLICY, int VECPOT>
__global__ __launch_bounds__(256) void k_tree_points(const TreeView tv, QuerySrc qs, float spread,
                                                     float circ, float *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= qs.nq) return;
    const float *p = qs.q + i * qs.stride;
    const float3 r = eval_query<POLICY, VECPOT>(tv, make_float3(p[0], p[1], p[2]), spread);
    out[3 * i] = r.x * circ;
    out[3 * i + 1] = r.y * circ;
    out[3 * i + 2] = r.z * circ;
}

// One cluster applied to the query as a supervorton (vortonSim.cpp:1888-1892,
// VORTON_ACCUMULATE_VELOCITY on a child of the influence tree).
template <int POLICY, int VECPOT>
__device__ __forceinline__ void apply_cluster(float3 &part, float3 q, float4 p, float4 w,
                                              float spread) {
    const float radius = p.w * 0.5f * spread;
    const float r2 = radius * radius;
    const float s = (2.0f / 3.0f) * r2 * radius;
    const float4 A = make_float4(p.x, p.y, p.z, r2);
    const float4 B = make_float4(w.x * s, w.y * s, w.z * s, __builtin_amdgcn_rcpf(r2 * radius));
    if (VECPOT) pair_vecpot<POLICY>(part, q, A, B);
    else pair_velocity<POLICY>(part, q, A, B);
}

// A layer's candidates by enumeration: every child in [c0, c1] (x fastest)
// outside the opened box, dealt round robin to the SPLIT lanes; empty
// clusters are read and skipped.  Any decimation.
template <int POLICY, int VECPOT, int SPLIT>
__device__ __forceinline__ void layer_enumerated(float3 &part, const TreeView &tv, int il, float3 q,
                                                 float spread, const int *c0, const int *c1,
                                                 bool open_any, const int *olo, const int *ohi,
                                                 uint32_t sub) {
    const vg_grid_desc &gc = tv.layer[il - 1];
    const float4 *__restrict__ P = tv.P[il - 1];
    const float4 *__restrict__ W = tv.W[il - 1];
    const int bx = c1[0] - c0[0] + 1, by = c1[1] - c0[1] + 1, bz = c1[2] - c0[2] + 1;
    const int total = bx * by * bz;
    for (int t = (int)sub; t < total; t += SPLIT) {
        const int rx = t % bx, r = t / bx;
        const int cx = c0[0] + rx, cy = c0[1] + r % by, cz = c0[2] + r / by;
        if (open_any && cx >= olo[0] && cx <= ohi[0] && cy >= olo[1] && cy <= ohi[1] &&
            cz >= olo[2] && cz <= ohi[2])
            continue;
        const size_t off = point_offset(gc, cx, cy, cz);
        const float4 p = P[off];
        const float4 w = W[off];
        if (p.w == 0.0f) continue;  // empty cluster: zero circulation
        apply_cluster<POLICY, VECPOT>(part, q, p, w, spread);
    }
}

// A layer's candidates through the parents' masks of non-empty children
// (bit cx + 2 (cy + 2 cz), decimation 2 on every axis, at most 2 parents per
// axis).  The candidates form a mask over the (at most) 4x4x4 children box,
// bit x + 4y + 16z standing for child c0 + (x, y, z), less the opened box;
// only its set bits are visited, in bit order (z, y, x: the order of the
// enumeration), lane `sub` taking the bits of rank sub, sub + SPLIT, ...
// Empty clusters cost nothing here, and most candidates are empty (about 70 %
// on a 64-cell grid around two rings).
template <int POLICY, int VECPOT, int SPLIT>
__device__ __forceinline__ void layer_masked(float3 &part, const TreeView &tv, int il, float3 q,
                                             float spread, const int *lo, const int *hi,
                                             const int *c0, bool open_any, const int *olo,
                                             const int *ohi, uint32_t sub) {
    const vg_grid_desc &gp = tv.layer[il];
    const vg_grid_desc &gc = tv.layer[il - 1];
    const uint32_t *__restrict__ Cm = tv.C[il];
    const float4 *__restrict__ P = tv.P[il - 1];
    const float4 *__restrict__ W = tv.W[il - 1];
    uint64_t cand = 0;
    for (int pz = lo[2]; pz <= hi[2]; ++pz)
        for (int py = lo[1]; py <= hi[1]; ++py)
            for (int px = lo[0]; px <= hi[0]; ++px) {
                const uint64_t cm = Cm[point_offset(gp, px, py, pz)];
                const uint64_t e = (cm & 3u) | ((cm & 0xCu) << 2) | ((cm & 0x30u) << 12) |
                                   ((cm & 0xC0u) << 14);
                cand |= e << (2 * (px - lo[0]) + 8 * (py - lo[1]) + 32 * (pz - lo[2]));
            }
    if (open_any) {
        // opened box as bit rows (x), row sets (y) and planes (z)
        const uint32_t xm = ((1u << (ohi[0] - olo[0] + 1)) - 1u) << (olo[0] - c0[0]);
        const uint32_t ym = ((1u << (ohi[1] - olo[1] + 1)) - 1u) << (olo[1] - c0[1]);
        const uint32_t zm = ((1u << (ohi[2] - olo[2] + 1)) - 1u) << (olo[2] - c0[2]);
        const uint32_t sy = (ym & 1u) | ((ym & 2u) << 3) | ((ym & 4u) << 6) | ((ym & 8u) << 9);
        const uint64_t sz = (uint64_t)(zm & 1u) | ((uint64_t)(zm & 2u) << 15) |
                            ((uint64_t)(zm & 4u) << 30) | ((uint64_t)(zm & 8u) << 45);
        cand &= ~((uint64_t)(xm * sy) * sz);
    }
    for (uint32_t k = 0; k < sub; ++k) cand &= cand - 1;
    while (cand) {
        const int b = __builtin_ctzll(cand);
#pragma unroll
        for (int k = 0; k < SPLIT; ++k) cand &= cand - 1;
        const size_t off =
            point_offset(gc, c0[0] + (b & 3), c0[1] + ((b >> 2) & 3), c0[2] + (b >> 4));
        apply_cluster<POLICY, VECPOT>(part, q, P[off], W[off], spread);
    }
}

// The reference-rule walk of one query split over SPLIT lanes (lanes of one
// query are adjacent and follow the same control flow): on every layer the
// candidates (children of the parents' box outside the opened box) are dealt
// to the lanes, which sum their share; the shares are added across the lanes
// (the same value on each) before the layer's Kahan step.  Opening decisions
// are those of tree_eval; only the summation order differs.  The walk is
// bound by instruction issue, not by its loads (the loads hit the caches), so
// what counts is the work per candidate: layers of decimation 2 visit only the
// non-empty children (layer_masked), others enumerate (layer_enumerated).
template <int POLICY, int VECPOT, int SPLIT>
__device__ float3 tree_eval_split(const TreeView &tv, float3 q, float spread, uint32_t sub) {
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    float3 sum = make_float3(0.f, 0.f, 0.f), comp = sum;
    const float qa[3] = {q.x, q.y, q.z};
    for (int il = (int)tv.depth - 1; il >= 1; --il) {
        const vg_grid_desc &gc = tv.layer[il - 1];
        const uint32_t *dims = tv.dec[il];
        int c0[3], c1[3], olo[3], ohi[3];
        bool open_any = il > 1;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            c0[a] = lo[a] * (int)dims[a];
            c1[a] = hi[a] * (int)dims[a] + (int)dims[a] - 1;
            olo[a] = 1 << 30;
            ohi[a] = -1;
        }
        if (open_any) {
            if (SPLIT >= 4) {
                // lane `sub` of the point's group decides axis `sub` (the
                // fourth lane idles) and the group shares the box
                const int a = (int)sub < 3 ? (int)sub : 2;
                const float cell = a == 0 ? gc.cell[0] : a == 1 ? gc.cell[1] : gc.cell[2];
                const float gmin = a == 0 ? gc.min[0] : a == 1 ? gc.min[1] : gc.min[2];
                const float qv = a == 0 ? qa[0] : a == 1 ? qa[1] : qa[2];
                const int ca = a == 0 ? c0[0] : a == 1 ? c0[1] : c0[2];
                const int cb = a == 0 ? c1[0] : a == 1 ? c1[1] : c1[2];
                float m;
                {
#pragma clang fp contract(off)
                    m = 0.0001f * cell;
                    if (a == 2 && 0.0f == cell) m = m + FLT_MIN;
                }
                int mylo = 1 << 30, myhi = -1;
                for (int c = ca; c <= cb; ++c) {
                    const float mn = madd_exact(gmin, (float)c, cell);
                    const float mx = madd_exact(gmin, (float)(c + 1), cell);
                    bool in;
                    {
#pragma clang fp contract(off)
                        in = qv >= mn - m && qv < mx + m;
                    }
                    if (in) {
                        mylo = c < mylo ? c : mylo;
                        myhi = c > myhi ? c : myhi;
                    }
                }
                const int base = (int)(threadIdx.x & 63u) & ~(SPLIT - 1);
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    olo[b] = __shfl(mylo, base + b, 64);
                    ohi[b] = __shfl(myhi, base + b, 64);
                    if (ohi[b] < olo[b]) open_any = false;
                }
            } else {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float cell = gc.cell[a];
                    float m;
                    {
#pragma clang fp contract(off)
                        m = 0.0001f * cell;
                        if (a == 2 && 0.0f == cell) m = m + FLT_MIN;
                    }
                    for (int c = c0[a]; c <= c1[a]; ++c) {
                        const float mn = madd_exact(gc.min[a], (float)c, cell);
                        const float mx = madd_exact(gc.min[a], (float)(c + 1), cell);
                        bool in;
                        {
#pragma clang fp contract(off)
                            in = qa[a] >= mn - m && qa[a] < mx + m;
                        }
                        if (in) {
                            olo[a] = c < olo[a] ? c : olo[a];
                            ohi[a] = c > ohi[a] ? c : ohi[a];
                        }
                    }
                    if (ohi[a] < olo[a]) open_any = false;
                }
            }
        }
        float3 part = make_float3(0.f, 0.f, 0.f);
        if (dims[0] == 2 && dims[1] == 2 && dims[2] == 2 && hi[0] - lo[0] <= 1 &&
            hi[1] - lo[1] <= 1 && hi[2] - lo[2] <= 1)
            layer_masked<POLICY, VECPOT, SPLIT>(part, tv, il, q, spread, lo, hi, c0, open_any, olo,
                                                ohi, sub);
        else
            layer_enumerated<POLICY, VECPOT, SPLIT>(part, tv, il, q, spread, c0, c1, open_any, olo,
                                                    ohi, sub);
#pragma unroll
        for (int m = 1; m < SPLIT; m <<= 1) {
            part.x += __shfl_xor(part.x, m, 64);
            part.y += __shfl_xor(part.y, m, 64);
            part.z += __shfl_xor(part.z, m, 64);
        }
        kahan_add(sum, comp, part);
        if (!open_any) break;
#pragma unroll
        for (int a = 0; a < 3; ++a) { lo[a] = olo[a]; hi[a] = ohi[a]; }
    }
    return sum;
}

// One workgroup per 4x4x4 brick of grid points, kTreeGridSplit lanes per
// point (lane t: point t / kTreeGridSplit of the brick, x fastest, share
// t % kTreeGridSplit of its walk).  Measured at 65^3 around two rings: 2
// lanes per point 0.055 ms, 4 lanes 0.060 ms, 1 lane 0.059 ms.
constexpr int kTreeGridSplit = 2;
// Face points: 4 lanes each (0.0195 ms for the faces of a 65^3 grid, against
// 0.023 ms with 2 lanes and 0.032 ms with 1).
constexpr int kTreeFaceSplit = 4;
// Brick `brick` of grid g (bricks nbx x nby x ..., x fastest) from tree tv.
// own_cells: a single-layer tree sums the vortons of its one leaf cell
// (trees of a batch share the sorted arrays) instead of all tv.n.
template <int POLICY>
__device__ __forceinline__ void tree_grid_brick(const TreeView &tv, const vg_grid_desc &g,
                                                uint32_t nbx, uint32_t nby, uint32_t brick,
                                                float spread, float circ, bool own_cells,
                                                float *__restrict__ out) {
    const uint32_t sub = threadIdx.x % kTreeGridSplit, pt = threadIdx.x / kTreeGridSplit;
    uint32_t ix = 4u * (brick % nbx) + (pt & 3u);
    uint32_t iy = 4u * ((brick / nbx) % nby) + ((pt >> 2) & 3u);
    uint32_t iz = 4u * (brick / (nbx * nby)) + (pt >> 4);
    // points past the grid's end walk the last point (their lanes take part
    // in the cross-lane sums) and write nothing
    const bool valid = ix < g.np[0] && iy < g.np[1] && iz < g.np[2];
    ix = ix < g.np[0] ? ix : g.np[0] - 1;
    iy = iy < g.np[1] ? iy : g.np[1] - 1;
    iz = iz < g.np[2] ? iz : g.np[2] - 1;
    const size_t i = point_offset(g, ix, iy, iz);
    const float3 q = grid_query_point(g, ix, iy, iz);
    float3 r;
    if (tv.depth > 1)
        r = tree_eval_split<POLICY, 0, kTreeGridSplit>(tv, q, spread, sub);
    else if (own_cells)
        r = sorted_direct_range<POLICY, 0>(tv, tv.cell_start[0],
                                           tv.cell_start[0] + tv.cell_count[0], q, spread);
    else
        r = sorted_direct<POLICY, 0>(tv, q, spread);
    if (!valid || sub != 0) return;
    out[3 * i] = r.x * circ;
    out[3 * i + 1] = r.y * circ;
    out[3 * i + 2] = r.z * circ;
}

template <int POLICY>
__global__ __launch_bounds__(64 * kTreeGridSplit) void k_tree_grid(const TreeView tv, vg_grid_desc g,
                                                                   uint32_t nbx, uint32_t nby,
                                                                   uint32_t nbricks, float spread,
                                                                   float circ,
                                                                   float *__restrict__ out) {
    const uint32_t brick = blockIdx.x;
    if (brick >= nbricks) return;
    tree_grid_brick<POLICY>(tv, g, nbx, nby, brick, spread, circ, false, out);
}

// Trees of a batch (blockIdx.y = tree), each on its own leaf grid with its
// own mollification; workgroups past a tree's brick count leave at once.
template <int POLICY>
__global__ __launch_bounds__(64 * kTreeGridSplit) void k_tree_grid_views(
    const TreeView *__restrict__ views, const float2 *__restrict__ mollify,
    float *__restrict__ out, size_t out_stride) {
    const uint32_t tree = blockIdx.y, brick = blockIdx.x;
    const TreeView &tv = views[tree];
    const vg_grid_desc &g = tv.layer[0];
    const uint32_t nbx = (g.np[0] + 3) / 4, nby = (g.np[1] + 3) / 4, nbz = (g.np[2] + 3) / 4;
    if (brick >= nbx * nby * nbz) return;
    const float2 sc = mollify[tree];
    tree_grid_brick<POLICY>(tv, g, nbx, nby, brick, sc.x, sc.y, true, out + tree * out_stride);
}

__global__ __launch_bounds__(256) void k_tree_count(const TreeView tv, vg_grid_desc g,
                                                    unsigned long long *__restrict__ total) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c = 0;
    if (i < capacity(g)) {
        const uint32_t ix = (uint32_t)(i % g.np[0]);
        const uint32_t iy = (uint32_t)((i / g.np[0]) % g.np[1]);
        const uint32_t iz = (uint32_t)(i / ((size_t)g.np[0] * g.np[1]));
        if (tv.depth > 1) tree_eval<0, 0, 1>(tv, grid_query_point(g, ix, iy, iz), 1.0f, &c);
        else c = tv.n;
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(total, (unsigned long long)c);
}

// Face point enumeration: z=0 plane, z=nz-1 plane, then for each interior
// z the y=0 row, the y=ny-1 row and the x=0 / x=nx-1 points of interior rows.
__device__ __forceinline__ bool face_point(const vg_grid_desc &g, size_t t, uint32_t &ix,
                                           uint32_t &iy, uint32_t &iz) {
    const size_t nx = g.np[0], ny = g.np[1], nz = g.np[2];
    const size_t plane = nx * ny;
    if (t < plane) {
        iz = 0; iy = (uint32_t)(t / nx); ix = (uint32_t)(t % nx);
        return true;
    }
    t -= plane;
    if (nz > 1) {
        if (t < plane) {
            iz = (uint32_t)(nz - 1); iy = (uint32_t)(t / nx); ix = (uint32_t)(t % nx);
            return true;
        }
        t -= plane;
    }
    if (nz <= 2) return false;
    const size_t side = ny > 2 ? 2 * (ny - 2) : 0;
    const size_t slab = 2 * nx + side;
    const size_t zz = t / slab;
    if (zz >= nz - 2) return false;
    size_t r = t % slab;
    iz = (uint32_t)(1 + zz);
    if (r < nx) { iy = 0; ix = (uint32_t)r; return true; }
    r -= nx;
    if (r < nx) { iy = (uint32_t)(ny - 1); ix = (uint32_t)r; return true; }
    r -= nx;
    iy = (uint32_t)(1 + r / 2);
    ix = (r & 1) ? (uint32_t)(nx - 1) : 0u;
    return true;
}

size_t face_count(const vg_grid_desc &g) {
    const size_t nx = g.np[0], ny = g.np[1], nz = g.np[2];
    size_t c = nx * ny * (nz > 1 ? 2 : 1);
    if (nz > 2) c += (nz - 2) * (2 * nx + (ny > 2 ? 2 * (ny - 2) : 0));
    return c;
}

template <int POLICY>
__global__ __launch_bounds__(256) void k_tree_faces(const TreeView tv, vg_grid_desc g,
                                                    size_t nface, int copy_odd, float spread,
                                                    float circ, float *__restrict__ out) {
    // kTreeFaceSplit lanes per face point, which leave together
    const size_t gt = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t t = gt / kTreeFaceSplit;
    const uint32_t sub = (uint32_t)(gt % kTreeFaceSplit);
    if (t >= nface) return;
    uint32_t ix, iy, iz;
    if (!face_point(g, t, ix, iy, iz)) return;
    if (copy_odd && (ix & 1u)) return;
    const float3 q = grid_query_point(g, ix, iy, iz);
    const float3 r = tv.depth > 1 ? tree_eval_split<POLICY, 1, kTreeFaceSplit>(tv, q, spread, sub)
                                  : sorted_direct<POLICY, 1>(tv, q, spread);
    if (sub != 0) return;
    const size_t o = point_offset(g, ix, iy, iz);
    out[3 * o] = r.x * circ;
    out[3 * o + 1] = r.y * circ;
    out[3 * o + 2] = r.z * circ;
}

// Odd-x face points copy x-1 (vortonSim.cpp:2124-2129).
__global__ void k_faces_copy_odd(vg_grid_desc g, size_t nface, float *__restrict__ out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nface) return;
    uint32_t ix, iy, iz;
    if (!face_point(g, t, ix, iy, iz) || !(ix & 1u)) return;
    const size_t o = point_offset(g, ix, iy, iz);
    out[3 * o] = out[3 * (o - 1)];
    out[3 * o + 1] = out[3 * (o - 1) + 1];
    out[3 * o + 2] = out[3 * (o - 1) + 2];
}

// ------------------------------------------------------------------ splats
// UniformGrid::Accumulate (uniformGrid.h:1278-1342) as a gather: grid point
// (ix,iy,iz) receives, from each vorton in the <=8 cells it is a corner of,
// weight ((wx*wy)*wz) with wx = tween.x when the point is the cell's +x
// corner, 1-tween.x otherwise, tween clamped to [0,1].
__device__ __forceinline__ float splat_weight(const vg_grid_desc &g, float4 p, int cx, int cy,
                                              int cz, int px, int py, int pz) {
#pragma clang fp contract(off)
    float t[3];
    const float pp[3] = {p.x, p.y, p.z};
    const int c[3] = {cx, cy, cz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float v = (pp[a] - (g.min[a] + (float)c[a] * g.cell[a])) * g.cpe[a];
        v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
        t[a] = v;
    }
    const float wx = px ? t[0] : 1.0f - t[0];
    const float wy = py ? t[1] : 1.0f - t[1];
    const float wz = pz ? t[2] : 1.0f - t[2];
    return wx * wy * wz;
}

// mode 0: vorticity item (Omega*2)*vc; mode 1: velocity average.
template <int MODE>
__global__ __launch_bounds__(256) void k_splat_gather(const TreeView tv, float vc,
                                                      const float *__restrict__ vel,
                                                      size_t vel_stride, float *__restrict__ out) {
#pragma clang fp contract(off)
    const vg_grid_desc &g = tv.layer[0];
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= capacity(g)) return;
    const int ix = (int)(o % g.np[0]);
    const int iy = (int)((o / g.np[0]) % g.np[1]);
    const int iz = (int)(o / ((size_t)g.np[0] * g.np[1]));
    const int ncx = (int)g.np[0] - 1, ncy = (int)g.np[1] - 1, ncz = (int)g.np[2] - 1;
    float ax = 0.f, ay = 0.f, az = 0.f, wsum = 0.f;
    for (int cz = iz - 1; cz <= iz; ++cz) {
        if (cz < 0 || cz >= ncz) continue;
        for (int cy = iy - 1; cy <= iy; ++cy) {
            if (cy < 0 || cy >= ncy) continue;
            for (int cx = ix - 1; cx <= ix; ++cx) {
                if (cx < 0 || cx >= ncx) continue;
                const size_t key = point_offset(g, cx, cy, cz);
                const uint32_t cnt = tv.cell_count[key];
                if (!cnt) continue;
                const uint32_t b = tv.cell_start[key];
                for (uint32_t k = 0; k < cnt; ++k) {
                    const float4 p = tv.sorted_p[b + k];
                    const float w = splat_weight(g, p, cx, cy, cz, cx != ix, cy != iy, cz != iz);
                    if (MODE == 0) {
                        const float4 om = tv.sorted_w[b + k];
                        ax += ((om.x * 2.0f) * vc) * w;
                        ay += ((om.y * 2.0f) * vc) * w;
                        az += ((om.z * 2.0f) * vc) * w;
                    } else {
                        const float *v = vel + (size_t)tv.sorted_idx[b + k] * vel_stride;
                        ax += v[0] * w;
                        ay += v[1] * w;
                        az += v[2] * w;
                        wsum += 1.0f * w;
                    }
                }
            }
        }
    }
    if (MODE == 1 && wsum > 0.0f) {
        const float inv = 1.0f / wsum;
        ax *= inv; ay *= inv; az *= inv;
    }
    out[3 * o] = ax;
    out[3 * o + 1] = ay;
    out[3 * o + 2] = az;
}

TreeView make_view(const vg_tree *t) {
    TreeView v;
    memset(&v, 0, sizeof v);
    v.depth = t->ng.depth;
    for (uint32_t l = 0; l < kMaxLayers; ++l) {
        v.layer[l] = t->ng.layer[l];
        for (int a = 0; a < 3; ++a) v.dec[l][a] = t->ng.dec[l][a];
        v.P[l] = t->P[l];
        v.W[l] = t->W[l];
        v.M[l] = t->M[l];
        v.C[l] = t->C[l];
    }
    v.cell_start = t->cell_start;
    v.cell_count = t->cell_count;
    v.sorted_p = t->sorted_p;
    v.sorted_w = t->sorted_w;
    v.sorted_idx = t->sorted_idx;
    v.n = t->n;
    return v;
}

// Batched binning: key = tree * cap0 + leaf cell of the owner's grid.
__global__ void k_bin_batch(const float *__restrict__ pos, size_t n,
                            const uint32_t *__restrict__ tree_of,
                            const vg_grid_desc *__restrict__ leaf_desc, size_t cap0,
                            uint32_t *__restrict__ count, uint32_t *__restrict__ keys,
                            uint32_t *__restrict__ slots) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    uint32_t key = 0;
    if (live) {
        const uint32_t k = tree_of[i];
        const vg_grid_desc g = leaf_desc[k];
        const float *p = pos + 3 * i;
        const int ix = cell_of(p[0], g.min[0], g.cpe[0], (int)g.np[0] - 1);
        const int iy = cell_of(p[1], g.min[1], g.cpe[1], (int)g.np[1] - 1);
        const int iz = cell_of(p[2], g.min[2], g.cpe[2], (int)g.np[2] - 1);
        key = (uint32_t)((size_t)k * cap0 + point_offset(g, ix, iy, iz));
    }
    const uint32_t slot = bin_slot(count, key, live);
    if (!live) return;
    keys[i] = key;
    slots[i] = slot;
}

__global__ void k_gather_batch(const float *__restrict__ pos, const float *__restrict__ angvel,
                               const uint32_t *__restrict__ tree_of,
                               const float *__restrict__ size,
                               const uint32_t *__restrict__ sorted_idx, size_t n,
                               float4 *__restrict__ sp, float4 *__restrict__ sw) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const size_t i = sorted_idx[j];
    const float *p = pos + 3 * i, *w = angvel + 3 * i;
    sp[j] = make_float4(p[0], p[1], p[2], size[tree_of[i]]);
    sw[j] = make_float4(w[0], w[1], w[2], 0.f);
}

// Record gather for the cell sort: record j of dst is record sorted_idx[j]
// of src (stride floats each).
__global__ void k_gather_records(const float *__restrict__ src, size_t stride,
                                 const uint32_t *__restrict__ sorted_idx, size_t n,
                                 float *__restrict__ dst) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const float *a = src + (size_t)sorted_idx[j] * stride;
    float *b = dst + j * stride;
    for (size_t c = 0; c < stride; ++c) b[c] = a[c];
}

}  // namespace

size_t scan_scratch_elems(size_t n) { return scan_elems(n); }

vg_status exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, uint32_t *tmp,
                             hipStream_t s) {
    return exclusive_scan(in, out, n, tmp, s);
}

vg_status batch_tree_build(const BatchTreeArgs &a, hipStream_t s) {
    const Nested &ng = *a.shape;
    const size_t cap0 = capacity(ng.layer[0]);
    const size_t cells = (size_t)a.trees * cap0;
    VG_REQUIRE(cells < (1ull << 32), "batched trees too large");
    const size_t n = a.n;
    VG_TRY(launch("vg_batch_fill", k_fill_u32, dim3(ceil_div(cells, 256)), dim3(256), 0, s,
                  a.cell_count, cells, 0u));
    VG_TRY(launch("vg_batch_bin", k_bin_batch, dim3(ceil_div(n, 256)), dim3(256), 0, s, a.pos, n,
                  a.tree_of, a.leaf_desc, cap0, a.cell_count, a.keys, a.slots));
    VG_TRY(exclusive_scan(a.cell_count, a.cell_start, cells, a.scan_tmp, s));
    uint32_t *unsorted = (uint32_t *)a.sorted_p;  // the gather fills sorted_p afterwards
    VG_TRY(launch("vg_batch_scatter", k_scatter, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                  (const uint32_t *)a.keys, (const uint32_t *)a.slots,
                  (const uint32_t *)a.cell_start, n, unsorted));
    VG_TRY(launch("vg_batch_sort_cells", k_rank_cells, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                  (const uint32_t *)a.keys, (const uint32_t *)a.cell_start,
                  (const uint32_t *)a.cell_count, (const uint32_t *)unsorted, n, a.sorted_idx));
    VG_TRY(launch("vg_batch_gather", k_gather_batch, dim3(ceil_div(n, 256)), dim3(256), 0, s, a.pos,
                  a.angvel, a.tree_of, a.size, (const uint32_t *)a.sorted_idx, n, a.sorted_p,
                  a.sorted_w));
    // The near-field walk reaches leaf clusters only through the layer-1
    // child masks and the cell lists, so empty leaf cells are neither
    // written nor, in layer 1's aggregation, read; the reference-rule walk
    // reads leaf clusters by position on layers it enumerates (all_leaves).
    if (a.views)
        VG_TRY(launch("vg_batch_leaf", k_leaf, dim3(ceil_div(cap0, 256), a.trees), dim3(256), 0, s,
                      (const uint32_t *)a.cell_start, (const uint32_t *)a.cell_count,
                      (const float4 *)a.sorted_p, (const float4 *)a.sorted_w, cap0, a.P[0],
                      a.W[0], a.M[0], a.views, a.all_leaves ? kLeafSkipUnlessRead : 1));
    else
        VG_TRY(launch("vg_batch_leaf", k_leaf, dim3(ceil_div(cells, 256)), dim3(256), 0, s,
                      (const uint32_t *)a.cell_start, (const uint32_t *)a.cell_count,
                      (const float4 *)a.sorted_p, (const float4 *)a.sorted_w, cells, a.P[0],
                      a.W[0], a.M[0], (const TreeView *)nullptr,
                      a.all_leaves && reads_empty_leaves(ng.depth, ng.dec[1]) ? 0 : 1));
    for (uint32_t l = 1; l < ng.depth; ++l) {
        const size_t cap = capacity(ng.layer[l]);
        const uint32_t *cc = l == 1 ? (const uint32_t *)a.cell_count : nullptr;
        if (a.views) {
            // per-tree shapes within ng (the largest shape): storage strides
            // are ng's capacities, geometry comes from each view
            VG_TRY(launch("vg_batch_aggregate", k_aggregate_views, dim3(ceil_div(cap, 256), a.trees),
                          dim3(256), 0, s, a.views, l, cap, capacity(ng.layer[l - 1]),
                          (const float4 *)a.P[l - 1], (const float4 *)a.W[l - 1],
                          (const float4 *)a.M[l - 1], cc, a.P[l], a.W[l], a.M[l], a.C[l]));
            continue;
        }
        VG_TRY(launch("vg_batch_aggregate", k_aggregate, dim3(ceil_div(cap, 256), a.trees), dim3(256),
                      0, s, ng.layer[l], ng.layer[l - 1], ng.dec[l][0], ng.dec[l][1], ng.dec[l][2],
                      (const float4 *)a.P[l - 1], (const float4 *)a.W[l - 1],
                      (const float4 *)a.M[l - 1], cc, a.P[l], a.W[l], a.M[l], a.C[l]));
    }
    return VG_OK;
}

vg_status batch_sort_particles(const BatchSortArgs &a, hipStream_t s) {
    VG_REQUIRE(a.cells < (1ull << 32) && a.n < (1ull << 32), "batched sort too large");
    const size_t n = a.n;
    if (n == 0) return VG_OK;
    VG_TRY(launch("vg_batch_sort_fill", k_fill_u32, dim3(ceil_div(a.cells, 256)), dim3(256), 0, s,
                  a.count, a.cells, 0u));
    VG_TRY(launch("vg_batch_sort_bin", k_bin_batch, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                  (const float *)a.pos, n, a.tree_of, a.leaf_desc, a.cap0, a.count, a.keys,
                  a.slots));
    VG_TRY(exclusive_scan(a.count, a.start, a.cells, a.scan_tmp, s));
    VG_TRY(launch("vg_batch_sort_scatter", k_scatter, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                  (const uint32_t *)a.keys, (const uint32_t *)a.slots, (const uint32_t *)a.start,
                  n, a.unsorted));
    VG_TRY(launch("vg_batch_sort_rank", k_rank_cells, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                  (const uint32_t *)a.keys, (const uint32_t *)a.start, (const uint32_t *)a.count,
                  (const uint32_t *)a.unsorted, n, a.sorted));
    VG_TRY(launch("vg_batch_sort_gather", k_gather_records, dim3(ceil_div(n, 256)), dim3(256), 0,
                  s, (const float *)a.pos, (size_t)3, (const uint32_t *)a.sorted, n, a.moved));
    VG_HIP(hipMemcpyAsync(a.pos, a.moved, 3 * n * sizeof(float), hipMemcpyDeviceToDevice, s));
    return VG_OK;
}

// Used by the simulation driver.
TreeView tree_view(const vg_tree *t) { return make_view(t); }

vg_status tree_checked_view(const vg_tree *t, TreeView *v) {
    VG_REQUIRE(t && v, "tree is null");
    if (!t->built) {
        set_error("tree not built (call vg_tree_build first)");
        return VG_ESTATE;
    }
    *v = make_view(t);
    return VG_OK;
}

vg_status tree_velocity_grid_batch(const TreeView *views, const float2 *mollify, uint32_t trees,
                                   const uint32_t np[3], int rsqrt_policy, float *out,
                                   size_t out_stride, hipStream_t s) {
    VG_REQUIRE(rsqrt_policy == VG_RSQRT_REFERENCE || rsqrt_policy == VG_RSQRT_ACCURATE,
               "bad rsqrt policy");
    auto k = rsqrt_policy == VG_RSQRT_REFERENCE ? k_tree_grid_views<VG_RSQRT_REFERENCE>
                                                : k_tree_grid_views<VG_RSQRT_ACCURATE>;
    const uint32_t nbricks = ((np[0] + 3) / 4) * ((np[1] + 3) / 4) * ((np[2] + 3) / 4);
    return launch("vg_batch_velocity_grid", k, dim3(nbricks, trees), dim3(64 * kTreeGridSplit), 0,
                  s, views, mollify, out, out_stride);
}

}  // namespace vg

using namespace vg;

VG_API vg_status vg_tree_create(vg_tree **t) {
    VG_REQUIRE(t, "tree out is null");
    *t = new vg_tree();
    return VG_OK;
}

VG_API vg_status vg_tree_destroy(vg_tree *t) {
    if (!t) return VG_OK;
    for (uint32_t l = 0; l < kMaxLayers; ++l) {
        (void)hipFree(t->P[l]);
        (void)hipFree(t->W[l]);
        (void)hipFree(t->M[l]);
        (void)hipFree(t->C[l]);
    }
    (void)hipFree(t->cell_start);
    (void)hipFree(t->cell_count);
    (void)hipFree(t->keys);
    (void)hipFree(t->slots);
    (void)hipFree(t->sorted_idx);
    (void)hipFree(t->sorted_p);
    (void)hipFree(t->sorted_w);
    (void)hipFree(t->scan_tmp);
    delete t;
    return VG_OK;
}

VG_API vg_status vg_tree_build(vg_tree *t, const vg_grid_desc *tmpl, const vg_vortons *v,
                               vg_stream stream) {
    VG_REQUIRE(t && tmpl && v, "null argument");
    VG_REQUIRE(v->n == 0 || (v->pos && v->angvel && v->stride >= 3), "bad vorton view");
    VG_REQUIRE(v->n < (1ull << 31), "too many vortons");
    VG_REQUIRE(tmpl->np[0] >= 2 && tmpl->np[1] >= 2 && tmpl->np[2] >= 2,
               "grid template is empty (define its shape first)");
    VG_REQUIRE(capacity(*tmpl) < (1ull << 32), "grid template too large");
    hipStream_t s = (hipStream_t)stream;
    t->built = false;
    nested_init(&t->ng, *tmpl);
    const size_t n = v->n;
    const size_t cap0 = capacity(t->ng.layer[0]);
    for (uint32_t l = 0; l < t->ng.depth; ++l) {
        const size_t cap = capacity(t->ng.layer[l]);
        size_t c1 = t->layer_cap[l], c2 = t->layer_cap[l];
        VG_TRY(grow(&t->P[l], &c1, cap));
        VG_TRY(grow(&t->W[l], &c2, cap));
        t->layer_cap[l] = c1 < c2 ? c1 : c2;
        VG_TRY(grow(&t->M[l], &t->m_cap[l], 3 * cap));
        if (l >= 1) VG_TRY(grow(&t->C[l], &t->c_cap[l], cap));
    }
    size_t cc1 = t->cell_cap, cc2 = t->cell_cap;
    VG_TRY(grow(&t->cell_start, &cc1, cap0));
    VG_TRY(grow(&t->cell_count, &cc2, cap0));
    t->cell_cap = cc1 < cc2 ? cc1 : cc2;
    const size_t nn = n ? n : 1;
    size_t n1 = t->n_cap, n2 = t->n_cap, n3 = t->n_cap, n4 = t->n_cap, n5 = t->n_cap;
    VG_TRY(grow(&t->keys, &n1, nn));
    VG_TRY(grow(&t->slots, &n2, nn));
    VG_TRY(grow(&t->sorted_idx, &n3, nn));
    VG_TRY(grow(&t->sorted_p, &n4, nn));
    VG_TRY(grow(&t->sorted_w, &n5, nn));
    t->n_cap = n1;
    if (n2 < t->n_cap) t->n_cap = n2;
    if (n3 < t->n_cap) t->n_cap = n3;
    if (n4 < t->n_cap) t->n_cap = n4;
    if (n5 < t->n_cap) t->n_cap = n5;
    VG_TRY(grow(&t->scan_tmp, &t->scan_cap, scan_elems(cap0)));
    t->n = (uint32_t)n;

    const VortonIn vin{v->pos, v->angvel, v->size, v->stride, v->uniform_size};
    VG_TRY(launch("vg_tree_fill", k_fill_u32, dim3(ceil_div(cap0, 256)), dim3(256), 0, s,
                  t->cell_count, cap0, 0u));
    VG_TRY(launch("vg_tree_bin", k_bin, dim3(ceil_div(n, 256)), dim3(256), 0, s, vin, n,
                  t->ng.layer[0], t->cell_count, t->keys, t->slots));
    VG_TRY(exclusive_scan(t->cell_count, t->cell_start, cap0, t->scan_tmp, s));
    // the cell-grouped (not yet index-ordered) list goes through sorted_p,
    // which the gather fills afterwards
    uint32_t *unsorted = (uint32_t *)t->sorted_p;
    VG_TRY(launch("vg_tree_scatter", k_scatter, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                  (const uint32_t *)t->keys, (const uint32_t *)t->slots,
                  (const uint32_t *)t->cell_start, n, unsorted));
    VG_TRY(launch("vg_tree_sort_cells", k_rank_cells, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                  (const uint32_t *)t->keys, (const uint32_t *)t->cell_start,
                  (const uint32_t *)t->cell_count, (const uint32_t *)unsorted, n,
                  t->sorted_idx));
    VG_TRY(launch("vg_tree_gather", k_gather_sorted, dim3(ceil_div(n, 256)), dim3(256), 0, s, vin,
                  (const uint32_t *)t->sorted_idx, n, t->sorted_p, t->sorted_w));
    VG_TRY(launch("vg_tree_leaf", k_leaf, dim3(ceil_div(cap0, 256)), dim3(256), 0, s,
                  (const uint32_t *)t->cell_start, (const uint32_t *)t->cell_count,
                  (const float4 *)t->sorted_p, (const float4 *)t->sorted_w, cap0, t->P[0],
                  t->W[0], t->M[0], (const TreeView *)nullptr, 0));
    for (uint32_t l = 1; l < t->ng.depth; ++l) {
        const size_t cap = capacity(t->ng.layer[l]);
        VG_TRY(launch("vg_tree_aggregate", k_aggregate, dim3(ceil_div(cap, 256)), dim3(256), 0, s,
                      t->ng.layer[l], t->ng.layer[l - 1], t->ng.dec[l][0], t->ng.dec[l][1],
                      t->ng.dec[l][2], (const float4 *)t->P[l - 1], (const float4 *)t->W[l - 1],
                      (const float4 *)t->M[l - 1], (const uint32_t *)nullptr, t->P[l], t->W[l],
                      t->M[l], t->C[l]));
    }
    t->built = true;
    return VG_OK;
}

VG_API vg_status vg_tree_depth(const vg_tree *t, uint32_t *depth) {
    VG_REQUIRE(t && depth, "null argument");
    VG_REQUIRE(t->built, "tree not built");
    *depth = t->ng.depth;
    return VG_OK;
}

VG_API vg_status vg_tree_layer(const vg_tree *t, uint32_t l, vg_grid_desc *g, float *pos,
                               float *angvel, float *size) {
    VG_REQUIRE(t && g, "null argument");
    VG_REQUIRE(t->built, "tree not built");
    VG_REQUIRE(l < t->ng.depth, "layer out of range");
    *g = t->ng.layer[l];
    const size_t cap = capacity(*g);
    if (!pos && !angvel && !size) return VG_OK;
    VG_HIP(hipDeviceSynchronize());
    float4 *hp = new float4[cap], *hw = new float4[cap];
    hipError_t e1 = hipMemcpy(hp, t->P[l], cap * sizeof(float4), hipMemcpyDeviceToHost);
    hipError_t e2 = hipMemcpy(hw, t->W[l], cap * sizeof(float4), hipMemcpyDeviceToHost);
    if (e1 == hipSuccess && e2 == hipSuccess) {
        for (size_t i = 0; i < cap; ++i) {
            if (pos) { pos[3 * i] = hp[i].x; pos[3 * i + 1] = hp[i].y; pos[3 * i + 2] = hp[i].z; }
            if (angvel) { angvel[3 * i] = hw[i].x; angvel[3 * i + 1] = hw[i].y; angvel[3 * i + 2] = hw[i].z; }
            if (size) size[i] = hp[i].w;
        }
    }
    delete[] hp;
    delete[] hw;
    VG_HIP(e1);
    VG_HIP(e2);
    return VG_OK;
}

VG_API vg_status vg_tree_layer_moments(const vg_tree *t, uint32_t l, float *dip) {
    VG_REQUIRE(t && dip, "null argument");
    VG_REQUIRE(t->built, "tree not built");
    VG_REQUIRE(l < t->ng.depth, "layer out of range");
    const size_t cap = capacity(t->ng.layer[l]);
    VG_HIP(hipDeviceSynchronize());
    float4 *h = new float4[3 * cap];
    hipError_t e = hipMemcpy(h, t->M[l], 3 * cap * sizeof(float4), hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        for (size_t i = 0; i < cap; ++i)
            for (int b = 0; b < 3; ++b) {
                dip[9 * i + 3 * b] = h[3 * i + b].x;
                dip[9 * i + 3 * b + 1] = h[3 * i + b].y;
                dip[9 * i + 3 * b + 2] = h[3 * i + b].z;
            }
    delete[] h;
    VG_HIP(e);
    return VG_OK;
}

static vg_status tree_ready(const vg_tree *t, int policy) {
    VG_REQUIRE(t, "tree is null");
    if (!t->built) {
        set_error("tree not built (call vg_tree_build first)");
        return VG_ESTATE;
    }
    VG_REQUIRE(policy == VG_RSQRT_REFERENCE || policy == VG_RSQRT_ACCURATE, "bad rsqrt policy");
    return VG_OK;
}

template <int VECPOT>
static vg_status tree_points(const vg_tree *t, const float *q, size_t q_stride, size_t nq,
                             float spread, float circ, int policy, float *out, vg_stream s,
                             const char *name) {
    VG_TRY(tree_ready(t, policy));
    VG_REQUIRE(nq == 0 || (q && q_stride >= 3 && out), "bad query array");
    if (nq == 0) return VG_OK;
    if (t->n == 0) return zero_fill(out, 3 * nq * sizeof(float), (hipStream_t)s);
    QuerySrc qs{q, q_stride, nq, {}};
    auto k = policy == VG_RSQRT_REFERENCE ? k_tree_points<VG_RSQRT_REFERENCE, VECPOT>
                                          : k_tree_points<VG_RSQRT_ACCURATE, VECPOT>;
    return launch(name, k, dim3(ceil_div(nq, 256)), dim3(256), 0, (hipStream_t)s, make_view(t),
                  qs, spread, circ, out);
}

VG_API vg_status vg_tree_velocity(const vg_tree *t, const float *q, size_t q_stride, size_t nq,
                                  float spread, float circ, int rsqrt_policy, float *out,
                                  vg_stream s) {
    return tree_points<0>(t, q, q_stride, nq, spread, circ, rsqrt_policy, out, s,
                          "vg_tree_velocity");
}

VG_API vg_status vg_tree_vecpot(const vg_tree *t, const float *q, size_t q_stride, size_t nq,
                                float spread, float circ, int rsqrt_policy, float *out,
                                vg_stream s) {
    return tree_points<1>(t, q, q_stride, nq, spread, circ, rsqrt_policy, out, s,
                          "vg_tree_vecpot");
}

VG_API vg_status vg_tree_velocity_grid(const vg_tree *t, const vg_grid_desc *g, float spread,
                                       float circ, int rsqrt_policy, float *out, vg_stream s) {
    VG_TRY(tree_ready(t, rsqrt_policy));
    VG_REQUIRE(g && out, "null argument");
    const size_t cap = capacity(*g);
    if (t->n == 0) return zero_fill(out, 3 * cap * sizeof(float), (hipStream_t)s);
    auto k = rsqrt_policy == VG_RSQRT_REFERENCE ? k_tree_grid<VG_RSQRT_REFERENCE>
                                                : k_tree_grid<VG_RSQRT_ACCURATE>;
    const uint32_t nbx = (g->np[0] + 3) / 4, nby = (g->np[1] + 3) / 4, nbz = (g->np[2] + 3) / 4;
    const uint32_t nbricks = nbx * nby * nbz;
    return launch("vg_tree_velocity_grid", k, dim3(nbricks), dim3(64 * kTreeGridSplit), 0,
                  (hipStream_t)s, make_view(t), *g, nbx, nby, nbricks, spread, circ, out);
}

VG_API vg_status vg_tree_interaction_count(const vg_tree *t, const vg_grid_desc *g,
                                           uint64_t *count, vg_stream stream) {
    VG_TRY(tree_ready(t, VG_RSQRT_REFERENCE));
    VG_REQUIRE(g && count, "null argument");
    hipStream_t s = (hipStream_t)stream;
    unsigned long long *d = nullptr;
    VG_TRY(salloc((void **)&d, sizeof(unsigned long long), s));
    vg_status st = zero_fill(d, sizeof(unsigned long long), s);
    if (st == VG_OK)
        st = launch("vg_tree_count", k_tree_count, dim3(ceil_div(capacity(*g), 256)), dim3(256), 0,
                    s, make_view(t), *g, d);
    unsigned long long h = 0;
    if (st == VG_OK) {
        hipError_t e = hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = device_error(e, "count copy", __FILE__, __LINE__);
    }
    sfree(d, s);
    *count = h;
    return st;
}

VG_API vg_status vg_tree_vecpot_boundary(const vg_tree *t, const vg_grid_desc *g, float spread,
                                         float circ, int rsqrt_policy, int copy_odd, float *out,
                                         vg_stream stream) {
    VG_TRY(tree_ready(t, rsqrt_policy));
    VG_REQUIRE(g && out, "null argument");
    hipStream_t s = (hipStream_t)stream;
    const size_t nf = face_count(*g);
    // With no vortons every cluster is empty and the faces evaluate to zero.
    auto k = rsqrt_policy == VG_RSQRT_REFERENCE ? k_tree_faces<VG_RSQRT_REFERENCE>
                                                : k_tree_faces<VG_RSQRT_ACCURATE>;
    VG_TRY(launch("vg_tree_vecpot_faces", k, dim3(ceil_div(nf * kTreeFaceSplit, 256)), dim3(256), 0, s,
                  make_view(t), *g, nf, copy_odd, spread, circ, out));
    if (copy_odd)
        VG_TRY(launch("vg_tree_faces_copy_odd", k_faces_copy_odd, dim3(ceil_div(nf, 256)),
                      dim3(256), 0, s, *g, nf, out));
    return VG_OK;
}

VG_API vg_status vg_tree_splat_vorticity(const vg_tree *t, float vorton_radius, float scale,
                                         float *out, vg_stream s) {
#pragma clang fp contract(off)
    VG_TRY(tree_ready(t, VG_RSQRT_REFERENCE));
    VG_REQUIRE(out, "out is null");
    const vg_grid_desc &g = t->ng.layer[0];
    const size_t cap = capacity(g);
    const float kPi = 3.1415926535897932384626433832795f;
    const float cell_volume = g.cell[0] * g.cell[1] * g.cell[2];
    const float four_pi_over3 = (4.0f * kPi) / 3.0f;
    const float vorton_volume = four_pi_over3 * (vorton_radius * vorton_radius * vorton_radius);
    const float vc = scale * vorton_volume / cell_volume;
    return launch("vg_splat_vorticity", k_splat_gather<0>, dim3(ceil_div(cap, 256)), dim3(256),
                  0, (hipStream_t)s, make_view(t), vc, (const float *)nullptr, (size_t)0, out);
}

VG_API vg_status vg_tree_populate_velocity_grid(const vg_tree *t, const float *vel,
                                                size_t vel_stride, float *out, vg_stream s) {
    VG_TRY(tree_ready(t, VG_RSQRT_REFERENCE));
    VG_REQUIRE(out && (vel || t->n == 0) && vel_stride >= 3, "bad argument");
    const size_t cap = capacity(t->ng.layer[0]);
    return launch("vg_populate_velocity_grid", k_splat_gather<1>, dim3(ceil_div(cap, 256)),
                  dim3(256), 0, (hipStream_t)s, make_view(t), 0.0f, vel, vel_stride, out);
}

VG_API vg_status vg_sort_particles_by_cell(const vg_grid_desc *g, float *records, size_t stride,
                                          size_t n, uint32_t *perm, vg_stream stream) {
    VG_REQUIRE(g, "grid is null");
    VG_REQUIRE(n == 0 || (records && stride >= 3), "bad particle records");
    VG_REQUIRE(n < (1ull << 32), "too many particles");
    VG_REQUIRE(g->np[0] >= 2 && g->np[1] >= 2 && g->np[2] >= 2, "grid is empty");
    const size_t cap = capacity(*g);
    VG_REQUIRE(cap < (1ull << 32), "grid too large");
    if (n == 0) return VG_OK;
    hipStream_t s = (hipStream_t)stream;
    // one scratch block: count, start (cells), keys, slots, unsorted, sorted
    // (particles), scan partials, then the reordered records
    const size_t scan_n = scan_elems(cap);
    const size_t u32s = 2 * cap + 4 * n + scan_n;
    const size_t rec_off = (u32s * sizeof(uint32_t) + 255) & ~size_t(255);
    unsigned char *base = nullptr;
    VG_TRY(salloc((void **)&base, rec_off + n * stride * sizeof(float), s));
    uint32_t *count = reinterpret_cast<uint32_t *>(base), *start = count + cap;
    uint32_t *keys = start + cap, *slots = keys + n, *unsorted = slots + n;
    uint32_t *sorted = unsorted + n, *tmp = sorted + n;
    float *moved = reinterpret_cast<float *>(base + rec_off);
    VortonIn vin{records, nullptr, nullptr, stride, 0.0f};
    vg_status st = launch("vg_sort_fill", k_fill_u32, dim3(ceil_div(cap, 256)), dim3(256), 0, s,
                          count, cap, 0u);
    if (st == VG_OK)
        st = launch("vg_sort_bin", k_bin, dim3(ceil_div(n, 256)), dim3(256), 0, s, vin, n, *g,
                    count, keys, slots);
    if (st == VG_OK) st = exclusive_scan(count, start, cap, tmp, s);
    if (st == VG_OK)
        st = launch("vg_sort_scatter", k_scatter, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                    (const uint32_t *)keys, (const uint32_t *)slots, (const uint32_t *)start, n,
                    unsorted);
    if (st == VG_OK)
        st = launch("vg_sort_rank", k_rank_cells, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                    (const uint32_t *)keys, (const uint32_t *)start, (const uint32_t *)count,
                    (const uint32_t *)unsorted, n, sorted);
    if (st == VG_OK)
        st = launch("vg_sort_gather", k_gather_records, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                    (const float *)records, stride, (const uint32_t *)sorted, n, moved);
    if (st == VG_OK) {
        hipError_t e = hipMemcpyAsync(records, moved, n * stride * sizeof(float),
                                      hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess && perm)
            e = hipMemcpyAsync(perm, sorted, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) st = device_error(e, "sort copy", __FILE__, __LINE__);
    }
    sfree(base, s);
    return st;
}
