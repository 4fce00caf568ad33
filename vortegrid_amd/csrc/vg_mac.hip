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
// per brick, so every loop bound and branch is uniform across the wave.  The
// clusters of a layer are enumerated from the parent layer's near box: each
// lane reads one parent's mask of non-empty children (written by the tree
// build), and the children outside this layer's near box are queued in LDS.
// Every 64 queued clusters are staged in LDS, and every lane then reads each
// one back (a broadcast) to evaluate it at its own point.  Batches whose
// clusters all have the current size skip the per-cluster size test, and
// clusters that lie beyond the core radius by construction skip the core
// test.
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
    int expand_from;  // layers >= expand_from go through brick-centred local expansions
};

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Cumulative decimation of layer l relative to the leaf layer, per axis.
__device__ __forceinline__ uint32_t layer_scale(const TreeView &tv, int l, int a) {
    uint32_t s = 1;
    for (int j = 1; j <= l; ++j) s *= tv.dec[j][a];
    return s;
}

// Per-wave LDS staging area: 64 fetched clusters (or vortons) at a time.
struct WaveStage {
    float4 P[64], W[64], M0[64], M1[64], M2[64];
    // leaf row tables (offsets in [0, 64), starts in [64, 128)), or the
    // cluster queue of the layers above
    uint32_t rows[128];
};

// Lane-strided walk over the cells of a bx x by x bz box: lane `lane` visits
// cells t = lane, lane + 64, ... as (x, y, z) = (t % bx, t / bx % by,
// t / (bx by)), stepped without divisions (64 = qz bx by + ry bx + rx, one
// carry per axis).  t < bx by bz exactly when z < bz.
struct BoxIter {
    int x, y, z, bx, by, rx, ry, qz;
};

__device__ __forceinline__ BoxIter box_begin(int bx, int by, uint32_t lane) {
    BoxIter it;
    const int bxy = bx * by;
    it.bx = bx;
    it.by = by;
    it.z = (int)lane / bxy;
    const int r = (int)lane - it.z * bxy;
    it.y = r / bx;
    it.x = r - it.y * bx;
    it.qz = 64 / bxy;
    const int r64 = 64 - it.qz * bxy;
    it.ry = r64 / bx;
    it.rx = r64 - it.ry * bx;
    return it;
}

__device__ __forceinline__ void box_next(BoxIter &it) {
    it.x += it.rx;
    int c = it.x >= it.bx ? 1 : 0;
    it.x -= c ? it.bx : 0;
    it.y += it.ry + c;
    c = it.y >= it.by ? 1 : 0;
    it.y -= c ? it.by : 0;
    it.z += it.qz + c;
}

// Offset of a cell of a layer (the tree's grids hold fewer than 2^32 points).
__device__ __forceinline__ uint32_t cell_offset(const vg_grid_desc &g, int x, int y, int z) {
    return (uint32_t)x + g.np[0] * ((uint32_t)y + g.np[1] * (uint32_t)z);
}

// Kernel factors for one cluster size; contributions are accumulated
// unscaled and multiplied by s = (2/3) r^3 when the size changes or the
// layer ends (every cluster of a tree normally has the vorton size).
struct SizeFactors {
    float size = -1.0f, r2 = 0.f, s = 0.f, ir3 = 0.f;
};

__device__ __forceinline__ void set_size(SizeFactors &f, float size, float spread, float3 &part,
                                         float3 &scaled) {
    scaled.x += part.x * f.s;
    scaled.y += part.y * f.s;
    scaled.z += part.z * f.s;
    part = make_float3(0.f, 0.f, 0.f);
    f.size = size;
    const float radius = size * 0.5f * spread;
    f.r2 = radius * radius;
    f.s = (2.0f / 3.0f) * f.r2 * radius;
    f.ir3 = __builtin_amdgcn_rcpf(f.r2 * radius);
}

// Monopole (+ first-moment) contribution of one cluster, unscaled by s.
// m0..m2 carry the first moment's rows times 3 (.xyz, staged so by the
// walk) and its antisymmetric part (.w, unscaled).
// FAR: the caller knows the cluster lies outside the core radius (the law
// inside the core never applies); d2 carries 1e-24 so that a cluster of
// zero vorticity (whose centroid defaults to the origin) adds 0 even at a
// query point on the origin (1e-24 is below half an ulp of any squared
// distance above 1e-16, so elsewhere it can only settle a rounding tie).
template <int POLICY, int ORDER, bool FAR = false>
__device__ __forceinline__ void cluster_term(float3 &part, float3 q, float4 p, float4 w, float4 m0,
                                             float4 m1, float4 m2, const SizeFactors &f) {
    const float dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
    const float d2 = FAR ? fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, 1e-24f)))
                         : dx * dx + dy * dy + dz * dz;
    float law, rs2;
    if (POLICY == VG_RSQRT_REFERENCE) {
        rs2 = __builtin_amdgcn_rcpf(d2);
        law = finvsqrtf(d2) * rs2;
    } else {
        const float rs = __builtin_amdgcn_rsqf(d2);
        rs2 = rs * rs;
        law = rs2 * rs;
    }
    const bool inside = !FAR && d2 < f.r2;
    const float lm = inside ? f.ir3 : law;
    if (ORDER >= 1) {
        // (w x d) lm - (a - 3 rs2 (t x d)) ld with t = D d folds into
        // ((w + c3 t3) x d) lm - a ld with t3 = 3 D d, c3 = rs2 outside the
        // core, 0 inside (where ld = 0): one cross product instead of two
        const float tx = m0.x * dx + m0.y * dy + m0.z * dz;
        const float ty = m1.x * dx + m1.y * dy + m1.z * dz;
        const float tz = m2.x * dx + m2.y * dy + m2.z * dz;
        const float c3 = inside ? 0.0f : rs2;
        const float vx = w.x + c3 * tx, vy = w.y + c3 * ty, vz = w.z + c3 * tz;
        if (FAR) {
            // lm = ld = law: ((v x d) - a) law, three fused operations a row
            part.x = fmaf(fmaf(vy, dz, -fmaf(vz, dy, m0.w)), law, part.x);
            part.y = fmaf(fmaf(vz, dx, -fmaf(vx, dz, m1.w)), law, part.y);
            part.z = fmaf(fmaf(vx, dy, -fmaf(vy, dx, m2.w)), law, part.z);
        } else {
            const float ld = inside ? 0.0f : law;
            part.x += (vy * dz - vz * dy) * lm - m0.w * ld;
            part.y += (vz * dx - vx * dz) * lm - m1.w * ld;
            part.z += (vx * dy - vy * dx) * lm - m2.w * ld;
        }
    } else {
        part.x += (w.y * dz - w.z * dy) * lm;
        part.y += (w.z * dx - w.x * dz) * lm;
        part.z += (w.x * dy - w.y * dx) * lm;
    }
}

// Local (Taylor) expansion about the brick centre of one cluster's field:
// monopole to second order, first moment to first order.  With d = c_b - c,
// K = d/|d|^3, u = s [Om x K(d) - sum_b e_b x (D_b . grad) K(d)]:
//   c0     = u(c_b)
//   c1[i][j] = du_i/dx_j                                 (9)
//   c2[i][p] = d2u_i/dx_j dx_k, p = 00 11 22 01 02 12   (18)
// Each lane expands its own cluster; the brick sums the lanes.
struct Expansion {
    float c0[3], c1[3][3], c2[3][6];
};

__device__ __forceinline__ void expansion_zero(Expansion &E) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        E.c0[i] = 0.f;
#pragma unroll
        for (int j = 0; j < 3; ++j) E.c1[i][j] = 0.f;
#pragma unroll
        for (int p = 0; p < 6; ++p) E.c2[i][p] = 0.f;
    }
}

__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

template <int ORDER>
__device__ __forceinline__ void expansion_add(Expansion &E, float3 cb, float4 p, float4 w, float4 m0,
                                              float4 m1, float4 m2, float spread) {
    const float3 d = make_float3(cb.x - p.x, cb.y - p.y, cb.z - p.z);
    const float r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float ir2 = ir * ir, ir3 = ir2 * ir, ir5 = ir3 * ir2, ir7 = ir5 * ir2;
    const float radius = p.w * 0.5f * spread;
    const float s = (2.0f / 3.0f) * radius * radius * radius;
    const float3 om = make_float3(s * w.x, s * w.y, s * w.z);
    const float dv[3] = {d.x, d.y, d.z};
    // Om x e_j and Om x d
    const float3 oe[3] = {make_float3(0.f, om.z, -om.y), make_float3(-om.z, 0.f, om.x),
                          make_float3(om.y, -om.x, 0.f)};
    const float3 od = cross3(om, d);
    E.c0[0] += od.x * ir3;
    E.c0[1] += od.y * ir3;
    E.c0[2] += od.z * ir3;
    // J[:, j] = e_j ir3 - 3 d d_j ir5
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const float f = -3.0f * ir5 * dv[j];
        E.c1[0][j] += oe[j].x * ir3 + od.x * f;
        E.c1[1][j] += oe[j].y * ir3 + od.y * f;
        E.c1[2][j] += oe[j].z * ir3 + od.z * f;
    }
    // c2[:, jk] = Om x T_jk. with T_jkb = d_j d_k K_b = 15 ir7 d_j d_k d_b
    // - 3 ir5 (delta_jb d_k + delta_kb d_j + delta_jk d_b), fully symmetric:
    // its 10 distinct components once, then one cross product per jk
    {
        const float a = -3.0f * ir5, g = 15.0f * ir7;
        const float gx = g * d.x, gy = g * d.y;
        const float xx = gx * d.x, yy = gy * d.y, zz = g * d.z * d.z, xy = gx * d.y;
        const float xxa = xx + a, yya = yy + a, zza = zz + a;
        const float txxx = d.x * (xxa + 2.0f * a), tyyy = d.y * (yya + 2.0f * a),
                    tzzz = d.z * (zza + 2.0f * a);
        const float txxy = d.y * xxa, txxz = d.z * xxa, tyyx = d.x * yya, tyyz = d.z * yya,
                    tzzx = d.x * zza, tzzy = d.y * zza, txyz = xy * d.z;
        const float3 T[6] = {make_float3(txxx, txxy, txxz), make_float3(tyyx, tyyy, tyyz),
                             make_float3(tzzx, tzzy, tzzz), make_float3(txxy, tyyx, txyz),
                             make_float3(txxz, txyz, tzzx), make_float3(txyz, tyyz, tzzy)};
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            E.c2[0][q] = fmaf(om.y, T[q].z, fmaf(-om.z, T[q].y, E.c2[0][q]));
            E.c2[1][q] = fmaf(om.z, T[q].x, fmaf(-om.x, T[q].z, E.c2[1][q]));
            E.c2[2][q] = fmaf(om.x, T[q].y, fmaf(-om.y, T[q].x, E.c2[2][q]));
        }
    }
    if (ORDER >= 1) {
        // D rows (scaled), antisymmetric part a, t_b = D_b . d
        const float3 D0 = make_float3(s * m0.x, s * m0.y, s * m0.z);
        const float3 D1 = make_float3(s * m1.x, s * m1.y, s * m1.z);
        const float3 D2 = make_float3(s * m2.x, s * m2.y, s * m2.z);
        const float3 av = make_float3(s * m0.w, s * m1.w, s * m2.w);
        const float3 t = make_float3(D0.x * d.x + D0.y * d.y + D0.z * d.z,
                                     D1.x * d.x + D1.y * d.y + D1.z * d.z,
                                     D2.x * d.x + D2.y * d.y + D2.z * d.z);
        const float3 td = cross3(t, d);
        // c0 += -a ir3 + 3 ir5 (t x d)
        E.c0[0] += -av.x * ir3 + 3.0f * ir5 * td.x;
        E.c0[1] += -av.y * ir3 + 3.0f * ir5 * td.y;
        E.c0[2] += -av.z * ir3 + 3.0f * ir5 * td.z;
        // c1[:, j] += 3 ir5 [t x e_j + d_j a + col_j x d] - 15 ir7 d_j (t x d)
        const float3 col[3] = {make_float3(D0.x, D1.x, D2.x), make_float3(D0.y, D1.y, D2.y),
                               make_float3(D0.z, D1.z, D2.z)};
        const float3 te[3] = {make_float3(0.f, t.z, -t.y), make_float3(-t.z, 0.f, t.x),
                              make_float3(t.y, -t.x, 0.f)};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float3 cd = cross3(col[j], d);
            const float a3 = 3.0f * ir5, b15 = 15.0f * ir7 * dv[j];
            E.c1[0][j] += a3 * (te[j].x + dv[j] * av.x + cd.x) - b15 * td.x;
            E.c1[1][j] += a3 * (te[j].y + dv[j] * av.y + cd.y) - b15 * td.y;
            E.c1[2][j] += a3 * (te[j].z + dv[j] * av.z + cd.z) - b15 * td.z;
        }
    }
}

// Sum of the lanes' expansions (xor butterfly: every lane ends with the same
// bits) evaluated at offset x from the brick centre.
__device__ __forceinline__ float3 expansion_eval(Expansion &E, float3 x) {
    float *v = &E.c0[0];
#pragma unroll
    for (int i = 0; i < 30; ++i) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off, 64);
    }
    const float xv[3] = {x.x, x.y, x.z};
    float u[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        u[i] = E.c0[i] + E.c1[i][0] * xv[0] + E.c1[i][1] * xv[1] + E.c1[i][2] * xv[2] +
               0.5f * (E.c2[i][0] * xv[0] * xv[0] + E.c2[i][1] * xv[1] * xv[1] +
                       E.c2[i][2] * xv[2] * xv[2]) +
               E.c2[i][3] * xv[0] * xv[1] + E.c2[i][4] * xv[0] * xv[2] +
               E.c2[i][5] * xv[1] * xv[2];
    }
    return make_float3(u[0], u[1], u[2]);
}

// One wave evaluates one brick of grid points; S is the wave's LDS slot.
// Appends to the LDS queue Q (128 entries, qn < 64 on entry) the non-empty
// cells of layer l that are children of the cells lo..hi of layer l + 1 and
// lie outside the near box n0..n1, as tag | offset.  Each lane takes one
// parent and reads its mask of non-empty children (so empty and near cells
// are never loaded); the wave then appends one child per lane per round.
// Whenever 64 entries are queued, flush() consumes Q[0, 64) and the rest
// moves down.
template <class Flush>
__device__ __forceinline__ void queue_far_children(const TreeView &tv, int l, const int lo[3],
                                                   const int hi[3], const int n0[3],
                                                   const int n1[3], uint32_t tag,
                                                   uint32_t *Q, uint32_t &qn, uint32_t lane,
                                                   Flush &&flush) {
    const vg_grid_desc &gl = tv.layer[l], &gp = tv.layer[l + 1];
    const uint32_t *__restrict__ Cm = tv.C[l + 1];
    const int dx = (int)tv.dec[l + 1][0], dy = (int)tv.dec[l + 1][1];
    const int dz = (int)tv.dec[l + 1][2];
    const int px = hi[0] - lo[0] + 1, py = hi[1] - lo[1] + 1, pz = hi[2] - lo[2] + 1;
    const int ptotal = px * py * pz;
    BoxIter it = box_begin(px, py, lane);
    for (int base = 0; base < ptotal; base += 64, box_next(it)) {
        uint32_t m = 0u, cbase = 0u;
        if (it.z < pz) {
            const int ax = lo[0] + it.x, ay = lo[1] + it.y, az = lo[2] + it.z;
            m = Cm[cell_offset(gp, ax, ay, az)];
            const int cx = ax * dx, cy = ay * dy, cz = az * dz;
            cbase = cell_offset(gl, cx, cy, cz);
            // children of this parent inside the near box, per axis (a
            // decimation is at most 3)
            uint32_t nx = 0u, ny = 0u, nz = 0u;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                if (i < dx && cx + i >= n0[0] && cx + i <= n1[0]) nx |= 1u << i;
                if (i < dy && cy + i >= n0[1] && cy + i <= n1[1]) ny |= 1u << i;
                if (i < dz && cz + i >= n0[2] && cz + i <= n1[2]) nz |= 1u << i;
            }
            if (nx && ny && nz) {
                uint32_t nm = 0u;
#pragma unroll
                for (int k = 0; k < 3; ++k)
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        if (((nz >> k) & (ny >> j) & 1u) != 0u) nm |= nx << (dx * (j + dy * k));
                m &= ~nm;
            }
        }
        for (;;) {
            const bool take = m != 0u;
            const unsigned long long b = __ballot(take);
            if (b == 0ull) break;
            if (take) {
                const uint32_t bit = (uint32_t)__builtin_ctz(m);
                m &= m - 1u;
                const uint32_t i = bit % (uint32_t)dx, jk = bit / (uint32_t)dx;
                const uint32_t j = jk % (uint32_t)dy, k = jk / (uint32_t)dy;
                const uint32_t below = (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
                Q[qn + below] = tag | (cbase + i + gl.np[0] * (j + gl.np[1] * k));
            }
            qn += (uint32_t)__popcll(b);
            if (qn >= 64u) {
                __builtin_amdgcn_wave_barrier();
                flush();
                const uint32_t rest = qn - 64u;
                const uint32_t moved = lane < rest ? Q[64 + lane] : 0u;
                __builtin_amdgcn_wave_barrier();
                if (lane < rest) Q[lane] = moved;
                __builtin_amdgcn_wave_barrier();
                qn = rest;
            }
        }
    }
}

template <int POLICY, int ORDER, int COUNT>
__device__ __forceinline__ void mac_brick(const TreeView &tv, const vg_grid_desc &g,
                                          const MacParams &mp, uint32_t brick, WaveStage &S,
                                          float *__restrict__ out,
                                          unsigned long long *__restrict__ counts) {
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
    unsigned long long n_clusters = 0, n_pairs = 0, n_expand = 0;
    const bool expanding = mp.expand_from >= 1 && mp.expand_from <= (int)tv.depth - 2;
    const float spread = mp.spread;
    const int depth = (int)tv.depth;
    SizeFactors f;
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
        // A cluster of this layer outside the near box lies at least k of
        // its cells from every point of the brick (its centroid is inside
        // its cell; the 0.99 absorbs the grid points' inward nudge): core
        // radii below that never apply.
        float far2;
        {
            const float h = fminf(gl.cell[0], fminf(gl.cell[1], gl.cell[2]));
            const float dmin = 0.99f * (float)mp.k * h;
            far2 = dmin * dmin;
        }
        const int bx = c1[0] - c0[0] + 1, by = c1[1] - c0[1] + 1, bz = c1[2] - c0[2] + 1;
        const int total = bx * by * bz;
        if (expanding && l >= mp.expand_from) {
            // these layers were applied by k_mac_expand, which left the raw
            // (unscaled by circ) expansion value in out
            if (COUNT) {
                BoxIter it = box_begin(bx, by, lane);
                for (int base = 0; base < total; base += 64, box_next(it)) {
                    bool use = false;
                    if (it.z < bz) {
                        const int cx = c0[0] + it.x, cy = c0[1] + it.y, cz = c0[2] + it.z;
                        const bool near = cx >= n0[0] && cx <= n1[0] && cy >= n0[1] &&
                                          cy <= n1[1] && cz >= n0[2] && cz <= n1[2];
                        if (!near) use = P[cell_offset(gl, cx, cy, cz)].w != 0.0f;
                    }
                    n_expand += __popcll(__ballot(use));
                }
            } else if (l == mp.expand_from && valid) {
                const size_t o = point_offset(g, pt[0], pt[1], pt[2]);
                kahan_add(sum, comp, make_float3(out[3 * o], out[3 * o + 1], out[3 * o + 2]));
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) { lo[a] = n0[a]; hi[a] = n1[a]; }
            continue;
        }
        float3 part = make_float3(0.f, 0.f, 0.f), scaled = part;
        if (COUNT) {
            BoxIter it = box_begin(bx, by, lane);
            for (int base = 0; base < total; base += 64, box_next(it)) {
                bool use = false;
                if (it.z < bz) {
                    const int cx = c0[0] + it.x, cy = c0[1] + it.y, cz = c0[2] + it.z;
                    const bool near = cx >= n0[0] && cx <= n1[0] && cy >= n0[1] &&
                                      cy <= n1[1] && cz >= n0[2] && cz <= n1[2];
                    if (!near) use = P[cell_offset(gl, cx, cy, cz)].w != 0.0f;
                }
                n_clusters += __popcll(__ballot(use));
            }
        } else {
            // Far clusters of the layer come through the LDS queue, 64 at a
            // time (fewer than 64 wait between rounds).
            uint32_t *Q = S.rows;
            uint32_t qn = 0;
            auto consume = [&](uint32_t cnt) {
                const bool use = lane < cnt;
                const uint32_t off = use ? Q[lane] : 0u;
                float4 p = make_float4(0.f, 0.f, 0.f, 0.f), w = p, m0 = p, m1 = p, m2 = p;
                if (use) {
                    p = P[off];
                    w = W[off];
                    if (ORDER) {
                        m0 = M[3 * (size_t)off];
                        m1 = M[3 * (size_t)off + 1];
                        m2 = M[3 * (size_t)off + 2];
                    }
                }
                unsigned long long mask = __ballot(use);
                S.P[lane] = p;
                S.W[lane] = w;
                if (ORDER) {
                    // rows times 3 (cluster_term's convention), once per cluster
                    S.M0[lane] = make_float4(3.0f * m0.x, 3.0f * m0.y, 3.0f * m0.z, m0.w);
                    S.M1[lane] = make_float4(3.0f * m1.x, 3.0f * m1.y, 3.0f * m1.z, m1.w);
                    S.M2[lane] = make_float4(3.0f * m2.x, 3.0f * m2.y, 3.0f * m2.z, m2.w);
                }
                // every cluster of the batch has the current size (the usual
                // case: one vorton size per tree)?
                const bool one_size = __ballot(use && p.w != f.size) == 0ull;
                __builtin_amdgcn_wave_barrier();
                if (one_size && f.r2 < far2) {
                    // no size test (the five LDS reads of a cluster go out
                    // together, one wait per cluster), and no core test
                    while (mask) {
                        const int j = __builtin_ctzll(mask);
                        mask &= mask - 1;
                        const float4 cp = S.P[j], cw = S.W[j];
                        float4 a0 = cp, a1 = cp, a2 = cp;
                        if (ORDER) {
                            a0 = S.M0[j];
                            a1 = S.M1[j];
                            a2 = S.M2[j];
                        }
                        cluster_term<POLICY, ORDER, true>(part, q, cp, cw, a0, a1, a2, f);
                    }
                } else if (one_size) {
                    while (mask) {
                        const int j = __builtin_ctzll(mask);
                        mask &= mask - 1;
                        const float4 cp = S.P[j], cw = S.W[j];
                        float4 a0 = cp, a1 = cp, a2 = cp;
                        if (ORDER) {
                            a0 = S.M0[j];
                            a1 = S.M1[j];
                            a2 = S.M2[j];
                        }
                        cluster_term<POLICY, ORDER>(part, q, cp, cw, a0, a1, a2, f);
                    }
                } else {
                    while (mask) {
                        const int j = __builtin_ctzll(mask);
                        mask &= mask - 1;
                        const float4 cp = S.P[j];
                        if (cp.w != f.size) set_size(f, cp.w, spread, part, scaled);
                        const float4 cw = S.W[j];
                        float4 a0 = cp, a1 = cp, a2 = cp;
                        if (ORDER) {
                            a0 = S.M0[j];
                            a1 = S.M1[j];
                            a2 = S.M2[j];
                        }
                        cluster_term<POLICY, ORDER>(part, q, cp, cw, a0, a1, a2, f);
                    }
                }
                __builtin_amdgcn_wave_barrier();
            };
            queue_far_children(tv, l, lo, hi, n0, n1, 0u, Q, qn, lane, [&] { consume(64u); });
            __builtin_amdgcn_wave_barrier();
            if (qn) consume(qn);
        }
        scaled.x += part.x * f.s;
        scaled.y += part.y * f.s;
        scaled.z += part.z * f.s;
        kahan_add(sum, comp, scaled);
#pragma unroll
        for (int a = 0; a < 3; ++a) { lo[a] = n0[a]; hi[a] = n1[a]; }
    }
    // Leaf near box (or every vorton when the tree has one layer), fetched by
    // rows: the vortons of cells x0..x1 of one (y, z) row are contiguous in
    // the cell-sorted arrays.
    const vg_grid_desc &g0 = tv.layer[0];
    if (depth == 1) {
        lo[0] = lo[1] = lo[2] = 0;
        hi[0] = (int)g0.np[0] - 2;
        hi[1] = (int)g0.np[1] - 2;
        hi[2] = (int)g0.np[2] - 2;
    }
    {
        float3 part = make_float3(0.f, 0.f, 0.f), scaled = part;
        const int ry = hi[1] - lo[1] + 1, rz = hi[2] - lo[2] + 1;
        const int nrows = ry * rz;
        for (int rbase = 0; rbase < nrows; rbase += 64) {
            const int r = rbase + (int)lane;
            uint32_t rs = 0, rc = 0;
            if (r < nrows) {
                const int cy = lo[1] + r % ry, cz = lo[2] + r / ry;
                const size_t k0 = point_offset(g0, lo[0], cy, cz);
                const size_t k1 = point_offset(g0, hi[0], cy, cz);
                rs = tv.cell_start[k0];
                rc = tv.cell_start[k1] + tv.cell_count[k1] - rs;
            }
            // inclusive scan of the row counts across the wave
            uint32_t incl = rc;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t tt = __shfl_up(incl, off, 64);
                if ((int)lane >= off) incl += tt;
            }
            const uint32_t total = uni(__shfl(incl, 63, 64));
            if (COUNT) {
                n_pairs += total;
                continue;
            }
            S.rows[lane] = incl - rc;
            S.rows[64 + lane] = rs;
            __builtin_amdgcn_wave_barrier();
            for (uint32_t vbase = 0; vbase < total; vbase += 64) {
                const uint32_t t = vbase + lane;
                float4 p = make_float4(0.f, 0.f, 0.f, 0.f), w = p;
                if (t < total) {
                    // last row whose offset is <= t
                    int lo_i = 0, hi_i = 63;
                    while (lo_i < hi_i) {
                        const int mid = (lo_i + hi_i + 1) >> 1;
                        if (S.rows[mid] <= t) lo_i = mid;
                        else hi_i = mid - 1;
                    }
                    const uint32_t idx = S.rows[64 + lo_i] + (t - S.rows[lo_i]);
                    p = tv.sorted_p[idx];
                    w = tv.sorted_w[idx];
                }
                const bool one_size = __ballot(t < total && p.w != f.size) == 0ull;
                __builtin_amdgcn_wave_barrier();
                S.P[lane] = p;
                S.W[lane] = w;
                __builtin_amdgcn_wave_barrier();
                const uint32_t cnt = total - vbase < 64u ? total - vbase : 64u;
                if (one_size) {
                    for (uint32_t j = 0; j < cnt; ++j) {
                        const float4 vp = S.P[j], vw = S.W[j];
                        cluster_term<POLICY, 0>(part, q, vp, vw, vp, vp, vp, f);
                    }
                } else {
                    for (uint32_t j = 0; j < cnt; ++j) {
                        const float4 vp = S.P[j];
                        if (vp.w != f.size) set_size(f, vp.w, spread, part, scaled);
                        cluster_term<POLICY, 0>(part, q, vp, S.W[j], vp, vp, vp, f);
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        scaled.x += part.x * f.s;
        scaled.y += part.y * f.s;
        scaled.z += part.z * f.s;
        kahan_add(sum, comp, scaled);
    }
    if (COUNT) {
        // the ballot must run on every lane: inside the lane-0 branch only
        // lane 0 would vote
        const unsigned long long nv = (unsigned long long)__popcll(__ballot(valid));
        if (lane == 0) {
            atomicAdd(&counts[0], n_clusters * nv);
            atomicAdd(&counts[1], n_pairs * nv);
            atomicAdd(&counts[2], n_expand);
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

// Layers >= expand_from of the near-field walk for one brick: every lane
// expands its fetched clusters about the brick centre, the wave sums the
// expansions and each point writes their raw value (not yet scaled by circ)
// to out, where the per-point walk picks it up.  Kept apart from mac_brick
// so the per-point walk keeps its small register footprint.
template <int ORDER>
__device__ __forceinline__ void expand_brick(const TreeView &tv, const vg_grid_desc &g,
                                             const MacParams &mp, uint32_t brick,
                                             uint32_t *__restrict__ Q, float *__restrict__ out) {
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
    float3 cb;
    {
        const float nudge = 1.0f - 2.0f * FLT_EPSILON;
        cb.x = g.min[0] + ((float)(4 * bb[0]) + 1.5f) * (g.cell[0] * nudge);
        cb.y = g.min[1] + ((float)(4 * bb[1]) + 1.5f) * (g.cell[1] * nudge);
        cb.z = g.min[2] + ((float)(4 * bb[2]) + 1.5f) * (g.cell[2] * nudge);
    }
    Expansion E;
    expansion_zero(E);
    // Accepted clusters (layer << 28 | offset) queue up in Q and are expanded
    // 64 at a time, one per lane, so that no lane idles on an empty or near
    // candidate cell while the others expand.
    uint32_t qn = 0;  // wave-uniform queue length, < 128
    auto expand_one = [&](uint32_t e) {
        const int l = (int)(e >> 28);
        const uint32_t off = e & 0x0fffffffu;
        const float4 p = tv.P[l][off];
        const float4 w = tv.W[l][off];
        float4 m0 = p, m1 = p, m2 = p;
        if (ORDER) {
            m0 = tv.M[l][3 * (size_t)off];
            m1 = tv.M[l][3 * (size_t)off + 1];
            m2 = tv.M[l][3 * (size_t)off + 2];
        }
        expansion_add<ORDER>(E, cb, p, w, m0, m1, m2, mp.spread);
    };
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int l = (int)tv.depth - 2; l >= mp.expand_from; --l) {
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
        const int bx = c1[0] - c0[0] + 1, by = c1[1] - c0[1] + 1, bz = c1[2] - c0[2] + 1;
        const int total = bx * by * bz;
        BoxIter it = box_begin(bx, by, lane);
        for (int base = 0; base < total; base += 64, box_next(it)) {
            bool take = false;
            uint32_t off = 0;
            if (it.z < bz) {
                const int cx = c0[0] + it.x, cy = c0[1] + it.y, cz = c0[2] + it.z;
                const bool near = cx >= n0[0] && cx <= n1[0] && cy >= n0[1] && cy <= n1[1] &&
                                  cz >= n0[2] && cz <= n1[2];
                if (!near) {
                    off = cell_offset(gl, cx, cy, cz);
                    take = tv.P[l][off].w != 0.0f;
                }
            }
            const unsigned long long m = __ballot(take);
            if (take) {
                const uint32_t below = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                Q[qn + below] = ((uint32_t)l << 28) | off;
            }
            qn += (uint32_t)__popcll(m);
            if (qn >= 64) {
                __builtin_amdgcn_wave_barrier();
                expand_one(Q[lane]);
                const uint32_t rest = qn - 64;
                const uint32_t moved = lane < rest ? Q[64 + lane] : 0u;
                __builtin_amdgcn_wave_barrier();
                if (lane < rest) Q[lane] = moved;
                __builtin_amdgcn_wave_barrier();
                qn = rest;
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) { lo[a] = n0[a]; hi[a] = n1[a]; }
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < qn) expand_one(Q[lane]);
    const float3 u = expansion_eval(E, make_float3(q.x - cb.x, q.y - cb.y, q.z - cb.z));
    if (valid) {
        const size_t o = point_offset(g, pt[0], pt[1], pt[2]);
        out[3 * o] = u.x;
        out[3 * o + 1] = u.y;
        out[3 * o + 2] = u.z;
    }
}

template <int ORDER>
__global__ __launch_bounds__(256) void k_mac_expand(const TreeView tv, vg_grid_desc g, MacParams mp,
                                                    float *__restrict__ out) {
    __shared__ uint32_t queue[4][128];
    const uint32_t wave = uni(threadIdx.x >> 6);
    const uint32_t brick = uni(blockIdx.x * 4 + wave);
    if (brick >= mp.nbricks) return;
    expand_brick<ORDER>(tv, g, mp, brick, queue[wave], out);
}

// Brick counts of a tree's own leaf layer.
__device__ __forceinline__ void tree_bricks(const vg_grid_desc &g, MacParams &p) {
    for (int a = 0; a < 3; ++a) p.nb[a] = (g.np[a] + 3) / 4;
    p.nbricks = p.nb[0] * p.nb[1] * p.nb[2];
}

template <int ORDER, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_mac_expand_batch(const TreeView *__restrict__ views,
                                                               const float2 *__restrict__ mollify,
                                                               MacParams mp,
                                                               float *__restrict__ out,
                                                               size_t out_stride) {
    __shared__ uint32_t queue[WPB][128];
    const uint32_t wave = uni(threadIdx.x >> 6);
    const uint32_t brick = uni(blockIdx.x * WPB + wave);
    const uint32_t tree = blockIdx.y;
    const TreeView &tv = views[tree];
    MacParams p = mp;
    if (p.nbricks == 0) tree_bricks(tv.layer[0], p);  // trees of their own shapes
    if (brick >= p.nbricks) return;
    p.spread = mollify[tree].x;
    expand_brick<ORDER>(tv, tv.layer[0], p, brick, queue[wave], out + tree * out_stride);
}

template <int POLICY, int ORDER, int COUNT>
__global__ __launch_bounds__(256) void k_mac_grid(const TreeView tv, vg_grid_desc g, MacParams mp,
                                                  float *__restrict__ out,
                                                  unsigned long long *__restrict__ counts) {
    __shared__ WaveStage stage[4];
    const uint32_t wave = uni(threadIdx.x >> 6);
    const uint32_t brick = uni(blockIdx.x * 4 + wave);
    if (brick >= mp.nbricks) return;
    mac_brick<POLICY, ORDER, COUNT>(tv, g, mp, brick, stage[wave], out, counts);
}

// Many trees of one shape at once: blockIdx.y selects the tree (its view,
// mollification and output slab), WPB waves per block cover WPB bricks.
template <int POLICY, int ORDER, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_mac_grid_batch(const TreeView *__restrict__ views,
                                                             const float2 *__restrict__ mollify,
                                                             MacParams mp, float *__restrict__ out,
                                                             size_t out_stride) {
    __shared__ WaveStage stage[WPB];
    const uint32_t wave = uni(threadIdx.x >> 6);
    const uint32_t brick = uni(blockIdx.x * WPB + wave);
    const uint32_t tree = blockIdx.y;
    const TreeView &tv = views[tree];
    MacParams p = mp;
    if (p.nbricks == 0) tree_bricks(tv.layer[0], p);  // trees of their own shapes
    if (brick >= p.nbricks) return;
    const float2 sc = mollify[tree];
    p.spread = sc.x;
    p.circ = sc.y;
    mac_brick<POLICY, ORDER, 0>(tv, tv.layer[0], p, brick, stage[wave], out + tree * out_stride,
                                nullptr);
}

// Work-queue forms for trees of different shapes: a grid of resident
// workgroups that take (tree, brick) items from a device counter until none
// is left, so every wave reaches the exit and uneven trees balance across
// the chip.  Items are numbered tree by tree in queue order (the
// largest trees first): brick_start[q] is the first item of the q-th tree
// (tree id order[q]), brick_start[trees] the total; each tree's bricks come
// from its own leaf layer.
// The walk's waves take one item at a time: its bricks are long and uneven,
// so a wave that finishes early goes on at once.
__device__ __forceinline__ uint32_t take_item(uint32_t *__restrict__ counter) {
    uint32_t item = 0;
    if ((threadIdx.x & 63) == 0) item = atomicAdd(counter, 1u);
    return uni((uint32_t)__shfl((int)item, 0, 64));
}

// The expansion's workgroups take WPB consecutive items at a time: its bricks
// are short and even, and adjacent bricks share their clusters in the CU's
// caches, as with a plain launch of WPB bricks per workgroup.  grab_items
// returns the first item, the same in every wave; every wave of the
// workgroup calls it (it holds two barriers).
template <int WPB>
__device__ __forceinline__ uint32_t grab_items(uint32_t *__restrict__ counter) {
    __shared__ uint32_t first;
    if (threadIdx.x == 0) first = atomicAdd(counter, (uint32_t)WPB);
    __syncthreads();
    const uint32_t f = uni(first);
    __syncthreads();  // read by every wave before the next grab overwrites it
    return f;
}

// Tree and brick of work item `item` (< brick_start[trees]).
__device__ __forceinline__ void item_brick(const uint32_t *__restrict__ order,
                                           const uint32_t *__restrict__ brick_start,
                                           uint32_t trees, uint32_t item, uint32_t &tree,
                                           uint32_t &brick) {
    uint32_t lo = 0, hi = trees;  // brick_start[lo] <= item < brick_start[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (brick_start[mid] <= item) lo = mid;
        else hi = mid;
    }
    tree = uni(order[lo]);
    brick = item - brick_start[lo];
}

template <int ORDER, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_mac_expand_queue(
    const TreeView *__restrict__ views, const float2 *__restrict__ mollify, MacParams mp,
    float *__restrict__ out, size_t out_stride, const uint32_t *__restrict__ order,
    const uint32_t *__restrict__ brick_start, uint32_t trees, uint32_t *__restrict__ counter) {
    __shared__ uint32_t queue[WPB][128];
    const uint32_t wave = uni(threadIdx.x >> 6);
    const uint32_t total = brick_start[trees];
    for (;;) {
        const uint32_t first = grab_items<WPB>(counter);
        if (first >= total) break;  // the same in every wave of the workgroup
        const uint32_t item = first + wave;
        if (item < total) {
            uint32_t tree, brick;
            item_brick(order, brick_start, trees, item, tree, brick);
            const TreeView &tv = views[tree];
            MacParams p = mp;
            p.spread = mollify[tree].x;
            tree_bricks(tv.layer[0], p);
            expand_brick<ORDER>(tv, tv.layer[0], p, brick, queue[wave], out + tree * out_stride);
        }
    }
}

template <int POLICY, int ORDER, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_mac_grid_queue(
    const TreeView *__restrict__ views, const float2 *__restrict__ mollify, MacParams mp,
    float *__restrict__ out, size_t out_stride, const uint32_t *__restrict__ order,
    const uint32_t *__restrict__ brick_start, uint32_t trees, uint32_t *__restrict__ counter) {
    __shared__ WaveStage stage[WPB];
    const uint32_t wave = uni(threadIdx.x >> 6);
    const uint32_t total = brick_start[trees];
    for (;;) {
        const uint32_t item = take_item(counter);
        if (item >= total) break;
        uint32_t tree, brick;
        item_brick(order, brick_start, trees, item, tree, brick);
        const TreeView &tv = views[tree];
        const float2 sc = mollify[tree];
        MacParams p = mp;
        p.spread = sc.x;
        p.circ = sc.y;
        tree_bricks(tv.layer[0], p);
        mac_brick<POLICY, ORDER, 0>(tv, tv.layer[0], p, brick, stage[wave],
                                    out + tree * out_stride, nullptr);
        __builtin_amdgcn_wave_barrier();
    }
}

__global__ void k_clear_counters(uint32_t *__restrict__ c) {
    if (threadIdx.x < 2) c[threadIdx.x] = 0u;
}

// Workgroups of `block` threads of kernel k that fit on the whole chip at once.
template <typename K>
vg_status resident_blocks(K k, int block, uint32_t *out) {
    int dev = 0, cus = 0, per_cu = 0;
    VG_HIP(hipGetDevice(&dev));
    VG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    VG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, block, 0));
    *out = (uint32_t)((per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1));
    return VG_OK;
}

template <int COUNT>
vg_status run_mac(const vg_tree *t, const vg_grid_desc *g, int near_k, int order, int expand_from,
                  float spread, float circ, int policy, float *out, unsigned long long *counts,
                  hipStream_t s) {
    TreeView tv;
    VG_TRY(tree_checked_view(t, &tv));
    VG_REQUIRE(g, "grid is null");
    VG_REQUIRE(near_k >= 0 && near_k <= 16, "near_k must be in [0, 16]");
    VG_REQUIRE(order == 0 || order == 1, "order must be 0 (monopole) or 1 (dipole)");
    VG_REQUIRE(policy == VG_RSQRT_REFERENCE || policy == VG_RSQRT_ACCURATE, "bad rsqrt policy");
    for (int a = 0; a < 3; ++a)
        VG_REQUIRE(g->np[a] == tv.layer[0].np[a],
                   "grid must have the tree's leaf geometry (the grid template)");
    VG_REQUIRE(!has_orphan_cells(tv),
               "the tree's grid leaves cells without a parent cluster (a cell count that is odd "
               "and above 3 on some layer); define the grid with power-of-two cells");
    MacParams mp;
    mp.k = near_k;
    mp.spread = spread;
    mp.circ = circ;
    for (int a = 0; a < 3; ++a) mp.nb[a] = (g->np[a] + 3) / 4;
    mp.nbricks = mp.nb[0] * mp.nb[1] * mp.nb[2];
    mp.expand_from = expand_from;
    const dim3 grid(ceil_div(mp.nbricks, 4));
    const char *name = COUNT ? "vg_tree_mac_count" : "vg_tree_velocity_grid_mac";
    using K = void (*)(const TreeView, vg_grid_desc, MacParams, float *, unsigned long long *);
    K k;
    if (policy == VG_RSQRT_REFERENCE)
        k = order ? k_mac_grid<VG_RSQRT_REFERENCE, 1, COUNT> : k_mac_grid<VG_RSQRT_REFERENCE, 0, COUNT>;
    else
        k = order ? k_mac_grid<VG_RSQRT_ACCURATE, 1, COUNT> : k_mac_grid<VG_RSQRT_ACCURATE, 0, COUNT>;
    if (!COUNT && expand_from >= 1 && expand_from <= (int)tv.depth - 2) {
        // the expansion queue packs (layer << 28 | offset)
        VG_REQUIRE(tv.depth <= 17 && capacity(tv.layer[expand_from]) < (1u << 28),
                   "grid too large for local expansions");
        VG_TRY(launch("vg_tree_mac_expand", order ? k_mac_expand<1> : k_mac_expand<0>, grid,
                      dim3(256), 0, s, tv, *g, mp, out));
    }
    return launch(name, k, grid, dim3(256), 0, s, tv, *g, mp, out, counts);
}

}  // namespace

vg_status mac_velocity_grid(const vg_tree *t, const vg_grid_desc *g, int near_k, int order,
                            float spread, float circ, int policy, float *out, hipStream_t s) {
    return run_mac<0>(t, g, near_k, order, -1, spread, circ, policy, out, nullptr, s);
}

// Batched walk over `trees` trees of one shape (views and per-tree
// (spread, circ) in device memory); out holds trees slabs of out_stride
// floats.  waves_per_block in {1, 2, 4}.
vg_status mac_velocity_grid_batch(const TreeView *views, const float2 *mollify, uint32_t trees,
                                  const uint32_t np[3], int near_k, int order, int expand_from,
                                  int policy, int waves_per_block, float *out, size_t out_stride,
                                  hipStream_t s, const BatchQueue *queue, bool own_shapes) {
    VG_REQUIRE(near_k >= 0 && near_k <= 16, "near_k must be in [0, 16]");
    VG_REQUIRE(order == 0 || order == 1, "order must be 0 or 1");
    VG_REQUIRE(policy == VG_RSQRT_REFERENCE || policy == VG_RSQRT_ACCURATE, "bad rsqrt policy");
    VG_REQUIRE(waves_per_block == 1 || waves_per_block == 2 || waves_per_block == 4,
               "waves_per_block must be 1, 2 or 4");
    MacParams mp;
    mp.k = near_k;
    mp.spread = 1.0f;
    mp.circ = 1.0f;
    for (int a = 0; a < 3; ++a) mp.nb[a] = (np[a] + 3) / 4;
    mp.nbricks = mp.nb[0] * mp.nb[1] * mp.nb[2];
    mp.expand_from = expand_from;
    // the plain launch covers np's bricks; with own_shapes each tree's
    // workgroups past its own brick count leave at once
    const uint32_t launch_bricks = mp.nbricks;
    if (own_shapes) mp.nbricks = 0;
    using K = void (*)(const TreeView *, const float2 *, MacParams, float *, size_t);
    K tab[2][2][3] = {
        {{k_mac_grid_batch<0, 0, 1>, k_mac_grid_batch<0, 0, 2>, k_mac_grid_batch<0, 0, 4>},
         {k_mac_grid_batch<0, 1, 1>, k_mac_grid_batch<0, 1, 2>, k_mac_grid_batch<0, 1, 4>}},
        {{k_mac_grid_batch<1, 0, 1>, k_mac_grid_batch<1, 0, 2>, k_mac_grid_batch<1, 0, 4>},
         {k_mac_grid_batch<1, 1, 1>, k_mac_grid_batch<1, 1, 2>, k_mac_grid_batch<1, 1, 4>}}};
    const int wi = waves_per_block == 1 ? 0 : waves_per_block == 2 ? 1 : 2;
    if (queue) {
        // trees of different shapes, each within np: work-queue kernels
        using KQ = void (*)(const TreeView *, const float2 *, MacParams, float *, size_t,
                            const uint32_t *, const uint32_t *, uint32_t, uint32_t *);
        VG_TRY(launch("vg_batch_queue_clear", k_clear_counters, dim3(1), dim3(64), 0, s,
                      queue->counters));
        if (expand_from >= 1) {
            VG_REQUIRE((size_t)np[0] * np[1] * np[2] < (1u << 28),
                       "grid too large for local expansions");
            KQ te[2][3] = {{k_mac_expand_queue<0, 1>, k_mac_expand_queue<0, 2>,
                            k_mac_expand_queue<0, 4>},
                           {k_mac_expand_queue<1, 1>, k_mac_expand_queue<1, 2>,
                            k_mac_expand_queue<1, 4>}};
            uint32_t nb = 0;
            VG_TRY(resident_blocks(te[order][wi], 64 * waves_per_block, &nb));
            VG_TRY(launch("vg_batch_expand", te[order][wi], dim3(nb), dim3(64 * waves_per_block),
                          0, s, views, mollify, mp, out, out_stride, queue->order,
                          queue->brick_start, trees, queue->counters + 1));
        }
        KQ tq[2][2][3] = {
            {{k_mac_grid_queue<0, 0, 1>, k_mac_grid_queue<0, 0, 2>, k_mac_grid_queue<0, 0, 4>},
             {k_mac_grid_queue<0, 1, 1>, k_mac_grid_queue<0, 1, 2>, k_mac_grid_queue<0, 1, 4>}},
            {{k_mac_grid_queue<1, 0, 1>, k_mac_grid_queue<1, 0, 2>, k_mac_grid_queue<1, 0, 4>},
             {k_mac_grid_queue<1, 1, 1>, k_mac_grid_queue<1, 1, 2>, k_mac_grid_queue<1, 1, 4>}}};
        uint32_t nb = 0;
        VG_TRY(resident_blocks(tq[policy][order][wi], 64 * waves_per_block, &nb));
        return launch("vg_batch_velocity_grid", tq[policy][order][wi], dim3(nb),
                      dim3(64 * waves_per_block), 0, s, views, mollify, mp, out, out_stride,
                      queue->order, queue->brick_start, trees, queue->counters);
    }
    if (expand_from >= 1) {
        // trees of the batch share one shape; the walk checks the depth.
        // The expansion queue packs (layer << 28 | offset).
        VG_REQUIRE((size_t)np[0] * np[1] * np[2] < (1u << 28),
                   "grid too large for local expansions");
        using KE = void (*)(const TreeView *, const float2 *, MacParams, float *, size_t);
        KE te[2][3] = {{k_mac_expand_batch<0, 1>, k_mac_expand_batch<0, 2>, k_mac_expand_batch<0, 4>},
                       {k_mac_expand_batch<1, 1>, k_mac_expand_batch<1, 2>, k_mac_expand_batch<1, 4>}};
        VG_TRY(launch("vg_batch_expand", te[order][wi],
                      dim3(ceil_div(launch_bricks, waves_per_block), trees),
                      dim3(64 * waves_per_block), 0, s, views, mollify, mp, out, out_stride));
    }
    return launch("vg_batch_velocity_grid", tab[policy][order][wi],
                  dim3(ceil_div(launch_bricks, waves_per_block), trees), dim3(64 * waves_per_block),
                  0, s, views, mollify, mp, out, out_stride);
}
}  // namespace vg

using namespace vg;

VG_API vg_status vg_tree_velocity_grid_mac(const vg_tree *t, const vg_grid_desc *g, int near_k,
                                           int order, int expand_from, float spread, float circ,
                                           int rsqrt_policy, float *out, vg_stream s) {
    VG_REQUIRE(out, "out is null");
    return run_mac<0>(t, g, near_k, order, expand_from, spread, circ, rsqrt_policy, out, nullptr,
                      (hipStream_t)s);
}

VG_API vg_status vg_tree_mac_work(const vg_tree *t, const vg_grid_desc *g, int near_k,
                                  int expand_from, uint64_t *clusters, uint64_t *pairs,
                                  uint64_t *expansions, vg_stream stream) {
    VG_REQUIRE(clusters && pairs, "null output");
    hipStream_t s = (hipStream_t)stream;
    unsigned long long *d = nullptr;
    VG_TRY(salloc((void **)&d, 3 * sizeof(unsigned long long), s));
    vg_status st = zero_fill(d, 3 * sizeof(unsigned long long), s);
    if (st == VG_OK) st = run_mac<1>(t, g, near_k, 0, expand_from, 1.0f, 1.0f, 0, nullptr, d, s);
    unsigned long long h[3] = {0, 0, 0};
    if (st == VG_OK) {
        hipError_t e = hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = device_error(e, "mac work copy", __FILE__, __LINE__);
    }
    sfree(d, s);
    *clusters = h[0];
    *pairs = h[1];
    if (expansions) *expansions = h[2];
    return st;
}
