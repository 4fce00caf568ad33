// vg_pse.hip -- viscous diffusion of vorticity by particle strength exchange
// (VortonSim::DiffuseAndDissipateVorticityPSE, vortonSim.cpp:3616-3922), on
// the partition of PartitionVortons (:3578-3607, particle.cpp:398-435),
// which is the influence tree's leaf cell lists.
//
// The reference's result depends on the order of its in-place exchanges.
// This follows its threaded build (USE_TBB, :3855-3860): every even z slice
// of cells, then every odd one; inside a slice, cells in (y, x) order,
// inside a cell vortons in index order.  Each live vorton h exchanges with
// the later vortons of its cell and then with those of 7 forward cells
// (+x, +y, +x+y, +z, +x+z, +y+z, +x+y+z; :3711-3728):
//   f = clamp(2 nu dt exp(-d^2 / (4 size_h)^2), -0.5, 0.5)
//   e = f (Omega_h - Omega_t),  Omega_h -= e,  Omega_t += e
// and then dissipates, Omega_h -= nu dt Omega_h.
//
// Schedule.  Cell (x, y) of a slice writes the vortons of cells
// {x, x+1} x {y, y+1} of that slice and the next, so two cells conflict only
// when |dx| <= 1 and |dy| <= 1.  Stepping T = x + 2y visits conflicting
// cells in the reference's (y, x) order (T grows by 1 for (+1, 0) and by
// 1..3 for (-1..+1, +1)) and never puts two of them in one step, so the
// cells of a step run in parallel and the result equals the serial
// traversal bit for bit.  Slices of one parity touch disjoint vortons
// (z, z+1), so each gets a workgroup; a barrier separates steps and a
// launch separates the phases.  A step lasts as long as its busiest wave, so
// k_pse_schedule first deals the non-empty cells of every step to the
// waves, largest first, each to the least loaded wave.  The exchange
// fraction f of a pair depends on positions and sizes only, so
// k_pse_fractions computes those of the cells of 4+ vortons over the whole
// chip beforehand, and their serial chains only move Omega.  The vorton
// state lives in the tree's cell-sorted order (positions and sizes from the
// tree, angular velocities gathered from the caller and scattered back), so
// a cell's vortons are contiguous.  exp is vg_pse_exp, the same function as
// the oracle's.
#include <cstdlib>

#include "vg_internal.h"

namespace vg {
vg_status tree_checked_view(const vg_tree *t, TreeView *v);

namespace {

constexpr int kPseThreads = 512;

// Cody-Waite reduction with the 16-bit split of ln 2, degree-7 Taylor
// polynomial, 2^k from the exponent bits; 0 below -87.
__device__ __forceinline__ float pse_exp(float x) {
#pragma clang fp contract(off)
    if (x != x) return x;
    if (x < -87.0f) return 0.0f;
    const float k = __builtin_rintf(x * 1.44269504f);
    const float r = (x - k * 0.693145751953125f) - k * 1.42860677e-06f;
    const float p =
        1.0f + r * (1.0f + r * (0.5f + r * (0.166666672f + r * (0.0416666679f +
        r * (0.00833333377f + r * (0.00138888892f + r * 0.000198412701f))))));
    return p * __int_as_float(((int)k + 127) << 23);
}

// Exchange fraction of the pair (here h, partner t): depends on positions
// and the here-vorton's size only.
__device__ __forceinline__ float pse_fraction(const float4 ph, float r2, const float4 pt,
                                              float nu_dt2) {
#pragma clang fp contract(off)
    const float sx = ph.x - pt.x, sy = ph.y - pt.y, sz = ph.z - pt.z;
    const float d2 = sx * sx + sy * sy + sz * sz;
    const float f = nu_dt2 * pse_exp(-d2 / r2);
    return f < -0.5f ? -0.5f : (f > 0.5f ? 0.5f : f);
}

__device__ __forceinline__ void exchange_f(float4 &wh, float4 &w, float f) {
#pragma clang fp contract(off)
    const float ex = (wh.x - w.x) * f, ey = (wh.y - w.y) * f, ez = (wh.z - w.z) * f;
    wh.x -= ex;
    wh.y -= ey;
    wh.z -= ez;
    w.x += ex;
    w.y += ey;
    w.z += ez;
}

__device__ __forceinline__ void exchange(const float4 ph, float r2, float4 &wh, const float4 pt,
                                         float4 *__restrict__ wt, float nu_dt2) {
    float4 w = *wt;
    exchange_f(wh, w, pse_fraction(ph, r2, pt, nu_dt2));
    *wt = w;
}

struct PseArgs {
    uint32_t nx, ny, nz;
    const uint32_t *start, *count;  // per leaf point (cell = its min corner)
    const float4 *p;                // sorted (x, y, z, size)
    float4 *w;                      // sorted (Omega, alive)
    float nu_dt2, dissipation;
    // Exchange fractions of the systolic cells, precomputed (or null): the
    // cell with leaf point c holds frac[frac_off[c] + i n + j] for here-vorton
    // i < k and partner j < n (0 where j <= i).
    const float *frac;
    const uint32_t *frac_off;
};

__device__ __forceinline__ void pse_cell(const PseArgs &a, uint32_t x, uint32_t y, uint32_t z) {
#pragma clang fp contract(off)
    const uint32_t nx = a.nx, nxy = a.nx * a.ny;
    const uint32_t c = x + nx * y + nxy * z;
    const uint32_t fwd[7] = {1, nx, 1 + nx, nxy, 1 + nxy, nx + nxy, 1 + nx + nxy};
    const uint32_t b = a.start[c], e = b + a.count[c];
    for (uint32_t i = b; i < e; ++i) {
        float4 wh = a.w[i];
        if (wh.w == 0.f) continue;  // dead: neither initiates nor dissipates
        const float4 ph = a.p[i];
        const float range = 4.0f * ph.w;
        const float r2 = range * range;
        for (uint32_t j = i + 1; j < e; ++j) exchange(ph, r2, wh, a.p[j], a.w + j, a.nu_dt2);
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const uint32_t cn = c + fwd[k];
            const uint32_t bn = a.start[cn], en = bn + a.count[cn];
            for (uint32_t j = bn; j < en; ++j) exchange(ph, r2, wh, a.p[j], a.w + j, a.nu_dt2);
        }
        wh.x -= wh.x * a.dissipation;
        wh.y -= wh.y * a.dissipation;
        wh.z -= wh.z * a.dissipation;
        a.w[i] = wh;
    }
}

// Wave-wide form of pse_cell: the cell's partner list (its own vortons, then
// those of the 7 forward cells, in order) is held in registers, partner j in
// lane j % 64, slot j / 64.  For each live here-vorton the lanes compute the
// exchange fractions of all its partners at once (positions only), then the
// wave walks the partners in order: the running Omega_h is wave-uniform and
// each step reads one partner through readlane and selects its new value
// into the owning lane, with the same operations, in the same order, as
// pse_cell.  Cells with more than kSlots * 64 partners fall back to
// pse_cell on lane 0.
constexpr int kSlots = 9;

__device__ __forceinline__ float lane_get(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
// The value v (wave-uniform) into lane `lane` of old.
__device__ __forceinline__ float lane_set(float v, float old, int lane) {
    return (int)(threadIdx.x & 63) == lane ? v : old;
}

// Systolic form for cells with several vortons.  Pair (i, j) needs Omega_h
// after pairs (i, j' < j) and Omega_t after pairs (i' < i, j); lane l holds
// here-vorton i = B + l while the partners stream down the lanes, one lane
// per step: partner j enters lane 0 at step j - B and meets lane l at step
// j - B + l, after lanes 0 .. l-1 have exchanged with it.  At j == i the
// incoming value is the here-vorton itself, already updated by the earlier
// here-vortons, and the lane adopts it.  Every variable sees the same
// operations in the same order as in pse_cell, with the serial chain
// shortened from k * n exchanges to k + n steps.  Here-vortons go in chunks
// of 64 lanes. Partners are read from memory in windows of 64 and leave the
// last lane back into memory, with a fence before the next chunk reads them.
constexpr uint32_t kSystolicMin = 4;

__device__ __forceinline__ uint32_t partner_index(const uint32_t base[8], const uint32_t pre[9],
                                                  uint32_t j) {
    uint32_t g = base[0] + j;
#pragma unroll
    for (int q = 1; q < 8; ++q) g = j >= pre[q] ? base[q] + (j - pre[q]) : g;
    return g;
}

// Lane l takes v from lane l - 1 and lane 0 takes `in`: the whole-wave DPP
// shift (wave_shr:1; lane 0 has no source lane and keeps the old value,
// here `in`), a register move instead of an LDS permute on the serial chain.
__device__ __forceinline__ float wave_shr(float v, float in) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(in), __float_as_int(v), 0x138, 0xf, 0xf, false));
}

__device__ __forceinline__ float4 wave_shr4(float4 v, float4 in) {
    return make_float4(wave_shr(v.x, in.x), wave_shr(v.y, in.y), wave_shr(v.z, in.z),
                       wave_shr(v.w, in.w));
}

__device__ __forceinline__ float4 lane4(float4 v, int l) {
    return make_float4(lane_get(v.x, l), lane_get(v.y, l), lane_get(v.z, l), lane_get(v.w, l));
}

__device__ __forceinline__ void pse_cell_systolic(const PseArgs &a, const uint32_t base[8],
                                                  const uint32_t pre[9], uint32_t k, uint32_t n) {
#pragma clang fp contract(off)
    const int lane = (int)(threadIdx.x & 63);
    for (uint32_t B = 0; B < k; B += 64) {
        const uint32_t L = k - B < 64u ? k - B : 64u;
        const uint32_t i = B + (uint32_t)lane;
        float4 wh = make_float4(0.f, 0.f, 0.f, 0.f), ph = wh, sp = wh, sw = wh, bp = wh, bw = wh;
        float r2 = 0.f;
        bool live = false;
        const uint32_t steps = (n - B) + L - 1;
        for (uint32_t t = 0; t < steps; ++t) {
            if ((t & 63u) == 0) {  // the next 64 partners to enter, one per lane
                const uint32_t jj = B + t + (uint32_t)lane;
                if (jj < n) {
                    const uint32_t g = partner_index(base, pre, jj);
                    bp = a.p[g];
                    bw = a.w[g];
                }
            }
            sp = wave_shr4(sp, lane4(bp, (int)(t & 63u)));
            sw = wave_shr4(sw, lane4(bw, (int)(t & 63u)));
            const int jl = (int)t - lane;  // partner j - B at this lane
            const uint32_t j = B + (uint32_t)jl;
            if ((uint32_t)lane < L && jl >= 0 && j < n) {
                if (j == i) {
                    wh = sw;
                    ph = sp;
                    live = sw.w != 0.f;
                    const float range = 4.0f * ph.w;
                    r2 = range * range;
                } else if (j > i && live) {
                    exchange(ph, r2, wh, sp, &sw, a.nu_dt2);
                }
                if (j == n - 1) {  // the lane's last partner: done
                    if (live) {
                        wh.x -= wh.x * a.dissipation;
                        wh.y -= wh.y * a.dissipation;
                        wh.z -= wh.z * a.dissipation;
                    }
                    a.w[partner_index(base, pre, i)] = wh;
                }
                // a partner that is not a here-vorton of this chunk leaves
                // the last lane with its value for the next chunk
                if ((uint32_t)lane == L - 1 && j >= B + L) a.w[partner_index(base, pre, j)] = sw;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
}

// Lane l takes lane l + 1 (wrapping): the whole-wave DPP rotation
// (wave_rol:1), so the window's next partner is always in lane 0.
__device__ __forceinline__ float wave_rol(float v) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x134, 0xf, 0xf, false));
}
__device__ __forceinline__ float4 wave_rol4(float4 v) {
    return make_float4(wave_rol(v.x), wave_rol(v.y), wave_rol(v.z), wave_rol(v.w));
}
__device__ __forceinline__ uint32_t wave_rol_u(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x134, 0xf, 0xf, false);
}
__device__ __forceinline__ uint32_t wave_shr_u(uint32_t v, uint32_t in) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)in, (int)v, 0x138, 0xf, 0xf, false);
}

// The systolic cell with its fractions read from a.frac: only Omega (and
// its vorton's index, for the write-back of partners leaving the last lane)
// streams down the lanes.  The window of 64 loaded partners rotates by one
// lane per step, so lane 0 always holds the next partner to enter; each
// lane's fractions are loaded kFracAhead steps before use (a lane walks its
// row of the fraction block left to right).
constexpr int kFracAhead = 8;

__device__ __forceinline__ void pse_cell_systolic_frac(const PseArgs &a, uint32_t c,
                                                       const uint32_t base[8],
                                                       const uint32_t pre[9], uint32_t k,
                                                       uint32_t n) {
#pragma clang fp contract(off)
    const int lane = (int)(threadIdx.x & 63);
    const float *__restrict__ fc = a.frac + a.frac_off[c];
    for (uint32_t B = 0; B < k; B += 64) {
        const uint32_t L = k - B < 64u ? k - B : 64u;
        const uint32_t i = B + (uint32_t)lane;
        const uint32_t gi = base[0] + i;  // here-vortons are the cell's own
        float4 wh = make_float4(0.f, 0.f, 0.f, 0.f), sw = wh, bw = wh;
        uint32_t sg = 0, bg = 0;
        bool live = false;
        const uint32_t steps = (n - B) + L - 1;
        // fraction of this lane at step t (0 unless it exchanges then)
        auto fetch = [&](uint32_t t) -> float {
            const int jl = (int)t - lane;
            const uint32_t j = B + (uint32_t)jl;
            return ((uint32_t)lane < L && jl > lane && j < n) ? fc[(size_t)i * n + j] : 0.f;
        };
        float cur[kFracAhead], nxt[kFracAhead];
#pragma unroll
        for (int d = 0; d < kFracAhead; ++d) cur[d] = fetch((uint32_t)d);
        for (uint32_t t0 = 0; t0 < steps; t0 += kFracAhead) {
#pragma unroll
            for (int d = 0; d < kFracAhead; ++d) nxt[d] = fetch(t0 + kFracAhead + (uint32_t)d);
#pragma unroll
            for (int d = 0; d < kFracAhead; ++d) {
                const uint32_t t = t0 + (uint32_t)d;
                if (t >= steps) break;
                if ((t & 63u) == 0) {  // the next 64 partners to enter, one per lane
                    const uint32_t jj = B + t + (uint32_t)lane;
                    if (jj < n) {
                        bg = partner_index(base, pre, jj);
                        bw = a.w[bg];
                    }
                }
                sw = wave_shr4(sw, bw);  // lane 0 takes the window's lane 0
                sg = wave_shr_u(sg, bg);
                bw = wave_rol4(bw);
                bg = wave_rol_u(bg);
                const int jl = (int)t - lane;  // partner j - B at this lane
                const uint32_t j = B + (uint32_t)jl;
                if ((uint32_t)lane < L && jl >= 0 && j < n) {
                    if (jl == lane) {  // j == i: the here-vorton itself
                        wh = sw;
                        live = sw.w != 0.f;
                    } else if (jl > lane && live) {
                        exchange_f(wh, sw, cur[d]);
                    }
                    if (j == n - 1) {  // the lane's last partner: done
                        if (live) {
                            wh.x -= wh.x * a.dissipation;
                            wh.y -= wh.y * a.dissipation;
                            wh.z -= wh.z * a.dissipation;
                        }
                        a.w[gi] = wh;
                    }
                    // a partner that is not a here-vorton of this chunk leaves
                    // the last lane with its value for the next chunk
                    if ((uint32_t)lane == L - 1 && j >= B + L) a.w[sg] = sw;
                }
            }
#pragma unroll
            for (int d = 0; d < kFracAhead; ++d) cur[d] = nxt[d];
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
}

__device__ __forceinline__ void pse_cell_wave(const PseArgs &a, uint32_t x, uint32_t y, uint32_t z) {
#pragma clang fp contract(off)
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t nx = a.nx, nxy = a.nx * a.ny;
    const uint32_t c = x + nx * y + nxy * z;
    const uint32_t offs[8] = {0, 1, nx, 1 + nx, nxy, 1 + nxy, nx + nxy, 1 + nx + nxy};
    uint32_t base[8], pre[9];
    pre[0] = 0;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        base[r] = a.start[c + offs[r]];
        pre[r + 1] = pre[r] + a.count[c + offs[r]];
    }
    const uint32_t k = pre[1], n = pre[8];
    if (k == 0) return;
    if (k >= kSystolicMin) {
        if (a.frac)
            pse_cell_systolic_frac(a, c, base, pre, k, n);
        else
            pse_cell_systolic(a, base, pre, k, n);
        return;
    }
    if (n > 64u * kSlots) {
        if (lane == 0) pse_cell(a, x, y, z);
        return;
    }
    float4 P[kSlots], W[kSlots];
    int G[kSlots];
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
        const uint32_t j = (uint32_t)lane + 64u * s;
        G[s] = -1;
        P[s] = W[s] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (j < n) {
            uint32_t g = base[0] + j;
#pragma unroll
            for (int q = 1; q < 8; ++q) g = j >= pre[q] ? base[q] + (j - pre[q]) : g;
            G[s] = (int)g;
            P[s] = a.p[g];
            W[s] = a.w[g];
        }
    }
    for (uint32_t i = 0; i < k; ++i) {
        const int si = (int)(i >> 6), li = (int)(i & 63);
        float4 wh = make_float4(0.f, 0.f, 0.f, 0.f), ph = wh;
#pragma unroll
        for (int s = 0; s < kSlots; ++s)
            if (s == si) {
                wh = make_float4(lane_get(W[s].x, li), lane_get(W[s].y, li), lane_get(W[s].z, li),
                                 lane_get(W[s].w, li));
                ph = make_float4(lane_get(P[s].x, li), lane_get(P[s].y, li), lane_get(P[s].z, li),
                                 lane_get(P[s].w, li));
            }
        if (wh.w == 0.f) continue;  // dead: neither initiates nor dissipates
        const float range = 4.0f * ph.w;
        const float r2 = range * range;
        float f[kSlots];
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            const float sx = ph.x - P[s].x, sy = ph.y - P[s].y, sz = ph.z - P[s].z;
            const float d2 = sx * sx + sy * sy + sz * sz;
            float fs = a.nu_dt2 * pse_exp(-d2 / r2);
            f[s] = fs < -0.5f ? -0.5f : (fs > 0.5f ? 0.5f : fs);
        }
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            const uint32_t j0 = 64u * s;
            const uint32_t lo = i + 1 > j0 ? i + 1 - j0 : 0;
            const uint32_t hi = n > j0 ? (n - j0 < 64u ? n - j0 : 64u) : 0;
            for (uint32_t L = lo; L < hi; ++L) {
                const int l = (int)L;
                const float fj = lane_get(f[s], l);
                const float tx = lane_get(W[s].x, l), ty = lane_get(W[s].y, l),
                            tz = lane_get(W[s].z, l);
                const float ex = (wh.x - tx) * fj, ey = (wh.y - ty) * fj, ez = (wh.z - tz) * fj;
                wh.x -= ex;
                wh.y -= ey;
                wh.z -= ez;
                W[s].x = lane_set(tx + ex, W[s].x, l);
                W[s].y = lane_set(ty + ey, W[s].y, l);
                W[s].z = lane_set(tz + ez, W[s].z, l);
            }
        }
        wh.x -= wh.x * a.dissipation;
        wh.y -= wh.y * a.dissipation;
        wh.z -= wh.z * a.dissipation;
#pragma unroll
        for (int s = 0; s < kSlots; ++s)
            if (s == si) {
                W[s].x = lane_set(wh.x, W[s].x, li);
                W[s].y = lane_set(wh.y, W[s].y, li);
                W[s].z = lane_set(wh.z, W[s].z, li);
            }
    }
#pragma unroll
    for (int s = 0; s < kSlots; ++s)
        if (G[s] >= 0) a.w[G[s]] = W[s];
}

// One workgroup per slice z = phase + 2 blockIdx.x (z < nz - 1); each wave
// takes whole cells of the current wavefront step.
constexpr int kPseWaves = kPseThreads / 64;

__global__ void k_pse_zero64(unsigned long long *v) { *v = 0ull; }

// Fraction block size of each leaf point's cell (k n for systolic cells,
// else 0), and the 64-bit total.
__global__ __launch_bounds__(256) void k_pse_frac_size(PseArgs a, uint32_t cells,
                                                       uint32_t *__restrict__ size,
                                                       unsigned long long *__restrict__ total) {
    __shared__ unsigned long long part[256];
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long kn = 0;
    if (c < cells) {
        const uint32_t k = a.count[c];
        if (k >= kSystolicMin) {
            const uint32_t nx = a.nx, nxy = a.nx * a.ny;
            const uint32_t offs[8] = {0, 1, nx, 1 + nx, nxy, 1 + nxy, nx + nxy, 1 + nx + nxy};
            uint32_t n = 0;
#pragma unroll
            for (int r = 0; r < 8; ++r) n += a.count[c + offs[r]];
            kn = (unsigned long long)k * n;
        }
        size[c] = kn < 0x80000000ull ? (uint32_t)kn : 0x80000000u;
    }
    part[threadIdx.x] = kn;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0 && part[0]) atomicAdd(total, part[0]);
}

// Exchange fractions of the systolic cells, one wave per leaf point's cell:
// the same operations as the in-chain exchange, so the same bits.
__global__ __launch_bounds__(256) void k_pse_fractions(PseArgs a, uint32_t cells,
                                                       float *__restrict__ frac) {
#pragma clang fp contract(off)
    const uint32_t c = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (c >= cells) return;
    const uint32_t k = a.count[c];
    if (k < kSystolicMin) return;
    const uint32_t nx = a.nx, nxy = a.nx * a.ny;
    const uint32_t offs[8] = {0, 1, nx, 1 + nx, nxy, 1 + nxy, nx + nxy, 1 + nx + nxy};
    uint32_t base[8], pre[9];
    pre[0] = 0;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        base[r] = a.start[c + offs[r]];
        pre[r + 1] = pre[r] + a.count[c + offs[r]];
    }
    const uint32_t n = pre[8];
    float *__restrict__ fc = frac + a.frac_off[c];
    const uint32_t total = k * n;
    for (uint32_t q = lane; q < total; q += 64) {
        const uint32_t i = q / n, j = q - i * n;
        float f = 0.f;
        if (j > i) {
            const float4 ph = a.p[base[0] + i];
            const float range = 4.0f * ph.w;
            f = pse_fraction(ph, range * range, a.p[partner_index(base, pre, j)], a.nu_dt2);
        }
        fc[q] = f;
    }
}

// Cells of wavefront step t of a slice: x + 2y = t, y from
// max(0, ceil((t - cx + 1) / 2)) to min(cy - 1, t / 2).
__device__ __forceinline__ void step_rows(int t, int cx, int cy, int &ylo, int &yhi) {
    ylo = t - (cx - 1) > 0 ? (t - (cx - 1) + 1) / 2 : 0;
    yhi = t / 2 < cy - 1 ? t / 2 : cy - 1;
}

// Work of a cell in steps of pse_cell_wave: the partner stream of each chunk
// of 64 here-vortons in the systolic form, k n in the small-cell form; 0 for
// an empty cell (nothing to do), at least 1 otherwise.
__device__ __forceinline__ uint32_t cell_cost(const PseArgs &a, uint32_t c) {
    const uint32_t nx = a.nx, nxy = a.nx * a.ny;
    const uint32_t k = a.count[c];
    if (k == 0) return 0u;
    const uint32_t offs[8] = {0, 1, nx, 1 + nx, nxy, 1 + nxy, nx + nxy, 1 + nx + nxy};
    uint32_t n = 0;
#pragma unroll
    for (int r = 0; r < 8; ++r) n += a.count[c + offs[r]];
    uint64_t cost = 0;
    if (k < kSystolicMin) {
        cost = (uint64_t)k * n;
    } else {
        for (uint32_t b = 0; b < k; b += 64) cost += (n - b) + (k - b < 64u ? k - b : 64u) - 1;
    }
    // a non-empty cell must never look empty: at least 1, saturated
    return cost == 0 ? 1u : (cost > 0xffffffffull ? 0xffffffffu : (uint32_t)cost);
}

// Cells of one wavefront step share no vortons, so any assignment of them to
// the waves gives the same result; the step lasts as long as the busiest
// wave.  One wave per (step t, slice z) assigns the step's non-empty cells,
// largest first, each to the wave with the least work so far (greedy LPT),
// and writes them grouped by wave: perm[sched * M + i] for i in
// [wstart[9 sched + w], wstart[9 sched + w + 1]), sched = z steps + t.
// Up to kSchedSlots * 64 cells per step.
constexpr int kSchedSlots = 4;

__global__ __launch_bounds__(64) void k_pse_schedule(PseArgs a, uint32_t steps, uint32_t M,
                                                     uint32_t *__restrict__ perm,
                                                     uint32_t *__restrict__ wstart) {
    const int t = (int)blockIdx.x;
    const uint32_t z = blockIdx.y;
    const int lane = (int)threadIdx.x;
    const int cx = (int)a.nx - 1, cy = (int)a.ny - 1;
    int ylo, yhi;
    step_rows(t, cx, cy, ylo, yhi);
    const uint32_t nx = a.nx, nxy = a.nx * a.ny;
    uint32_t rem[kSchedSlots];  // cost still to assign (0 once assigned or empty)
    int slot_w[kSchedSlots], slot_pos[kSchedSlots];
#pragma unroll
    for (int q = 0; q < kSchedSlots; ++q) {
        const int y = ylo + lane + 64 * q;
        rem[q] = y <= yhi ? cell_cost(a, (uint32_t)(t - 2 * y) + nx * (uint32_t)y + nxy * z) : 0u;
        slot_w[q] = -1;
        slot_pos[q] = 0;
    }
    uint32_t load[kPseWaves], cnt[kPseWaves];
#pragma unroll
    for (int w = 0; w < kPseWaves; ++w) load[w] = cnt[w] = 0u;
    for (;;) {
        uint32_t best = rem[0];
#pragma unroll
        for (int q = 1; q < kSchedSlots; ++q) best = rem[q] > best ? rem[q] : best;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const uint32_t o = (uint32_t)__shfl_xor((int)best, off, 64);
            best = o > best ? o : best;
        }
        best = __builtin_amdgcn_readfirstlane(best);
        if (best == 0u) break;
        // the first (slot, lane) holding it, and the least loaded wave
        int sel_q = 0, sel_lane = 0;
#pragma unroll
        for (int q = kSchedSlots - 1; q >= 0; --q) {
            const unsigned long long m = __ballot(rem[q] == best);
            if (m) {
                sel_q = q;
                sel_lane = (int)__builtin_ctzll(m);
            }
        }
        int w = 0;
#pragma unroll
        for (int v = 1; v < kPseWaves; ++v) w = load[v] < load[w] ? v : w;
        uint32_t pos = 0;
#pragma unroll
        for (int v = 0; v < kPseWaves; ++v) {
            if (v == w) {
                pos = cnt[v];
                cnt[v] += 1u;
                load[v] += best;
            }
        }
#pragma unroll
        for (int q = 0; q < kSchedSlots; ++q) {
            if (q == sel_q && lane == sel_lane) {
                slot_w[q] = w;
                slot_pos[q] = (int)pos;
                rem[q] = 0u;
            }
        }
    }
    uint32_t ws[kPseWaves + 1];
    ws[0] = 0;
#pragma unroll
    for (int v = 0; v < kPseWaves; ++v) ws[v + 1] = ws[v] + cnt[v];
    const size_t sched = (size_t)z * steps + (size_t)t;
    if (lane <= kPseWaves) {
        uint32_t v = 0;
#pragma unroll
        for (int u = 0; u <= kPseWaves; ++u) v = lane == u ? ws[u] : v;
        wstart[(kPseWaves + 1) * sched + (size_t)lane] = v;
    }
#pragma unroll
    for (int q = 0; q < kSchedSlots; ++q) {
        if (slot_w[q] >= 0) {
            uint32_t base = 0;
#pragma unroll
            for (int u = 0; u < kPseWaves; ++u) base = slot_w[q] == u ? ws[u] : base;
            perm[sched * M + base + (uint32_t)slot_pos[q]] = (uint32_t)(ylo + lane + 64 * q);
        }
    }
}

// perm / wstart: the schedule of k_pse_schedule, or null for round robin by
// row (steps of more than kSchedSlots * 64 cells).
__global__ __launch_bounds__(kPseThreads) void k_pse_phase(PseArgs a, uint32_t phase,
                                                           const uint32_t *__restrict__ perm,
                                                           const uint32_t *__restrict__ wstart,
                                                           uint32_t M) {
    const uint32_t z = phase + 2 * blockIdx.x;
    if (z + 1 >= a.nz) return;  // uniform per workgroup
    const int wave = (int)(threadIdx.x >> 6);
    const int cx = (int)a.nx - 1, cy = (int)a.ny - 1;  // cells per axis
    const int steps = (cx - 1) + 2 * (cy - 1) + 1;
    for (int t = 0; t < steps; ++t) {
        int ylo, yhi;
        step_rows(t, cx, cy, ylo, yhi);
        if (perm) {
            const size_t sched = (size_t)z * (size_t)steps + (size_t)t;
            const uint32_t b = wstart[(kPseWaves + 1) * sched + (size_t)wave];
            const uint32_t e = wstart[(kPseWaves + 1) * sched + (size_t)wave + 1];
            for (uint32_t i = b; i < e; ++i) {
                const int y = (int)perm[sched * M + i];
                pse_cell_wave(a, (uint32_t)(t - 2 * y), (uint32_t)y, z);
            }
        } else {
            for (int y = ylo + wave; y <= yhi; y += kPseWaves)
                pse_cell_wave(a, (uint32_t)(t - 2 * y), (uint32_t)y, z);
        }
        __syncthreads();
    }
}

// Simultaneous exchange (no reference counterpart): one thread per vorton
// in cell-sorted order, every exchange from the starting values w0, summed
// in the order of vgo_pse_diffuse_simultaneous: received from the live
// vortons of the 7 backward cells, from the live earlier vortons of the own
// cell, then (if alive) given to the later vortons of the own cell and the
// vortons of the 7 forward cells.
__device__ __forceinline__ void pse_pair(const float4 ph, float r2h, const float4 wh,
                                         const float4 pt, const float4 wt, float nu_dt2,
                                         float &ex, float &ey, float &ez) {
#pragma clang fp contract(off)
    const float sx = ph.x - pt.x, sy = ph.y - pt.y, sz = ph.z - pt.z;
    const float d2 = sx * sx + sy * sy + sz * sz;
    float f = nu_dt2 * pse_exp(-d2 / r2h);
    f = f < -0.5f ? -0.5f : (f > 0.5f ? 0.5f : f);
    ex = (wh.x - wt.x) * f;
    ey = (wh.y - wt.y) * f;
    ez = (wh.z - wt.z) * f;
}

__global__ void k_pse_simultaneous(PseArgs a, const vg_grid_desc g, const float4 *__restrict__ w0,
                                   float4 *__restrict__ w1, uint32_t n) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 pi = a.p[i], wi = w0[i];
    const int ix = cell_of(pi.x, g.min[0], g.cpe[0], (int)g.np[0] - 1);
    const int iy = cell_of(pi.y, g.min[1], g.cpe[1], (int)g.np[1] - 1);
    const int iz = cell_of(pi.z, g.min[2], g.cpe[2], (int)g.np[2] - 1);
    const uint32_t nx = a.nx, nxy = a.nx * a.ny;
    const uint32_t c = (uint32_t)ix + nx * (uint32_t)iy + nxy * (uint32_t)iz;
    const int d[7][3] = {{1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    float ax = 0.f, ay = 0.f, az = 0.f, ex, ey, ez;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        if (ix < d[k][0] || iy < d[k][1] || iz < d[k][2]) continue;
        const uint32_t cb = c - ((uint32_t)d[k][0] + nx * d[k][1] + nxy * d[k][2]);
        const uint32_t b = a.start[cb], e = b + a.count[cb];
        for (uint32_t j = b; j < e; ++j) {
            const float4 wh = w0[j];
            if (wh.w == 0.f) continue;
            const float4 ph = a.p[j];
            const float sh = 4.0f * ph.w;
            pse_pair(ph, sh * sh, wh, pi, wi, a.nu_dt2, ex, ey, ez);
            ax += ex;
            ay += ey;
            az += ez;
        }
    }
    const uint32_t b = a.start[c], e = b + a.count[c];
    for (uint32_t j = b; j < i; ++j) {
        const float4 wh = w0[j];
        if (wh.w == 0.f) continue;
        const float4 ph = a.p[j];
        const float sh = 4.0f * ph.w;
        pse_pair(ph, sh * sh, wh, pi, wi, a.nu_dt2, ex, ey, ez);
        ax += ex;
        ay += ey;
        az += ez;
    }
    const bool live = wi.w != 0.f;
    if (live) {
        const float si = 4.0f * pi.w;
        const float r2 = si * si;
        for (uint32_t j = i + 1; j < e; ++j) {
            pse_pair(pi, r2, wi, a.p[j], w0[j], a.nu_dt2, ex, ey, ez);
            ax -= ex;
            ay -= ey;
            az -= ez;
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const uint32_t cf = c + ((uint32_t)d[k][0] + nx * d[k][1] + nxy * d[k][2]);
            const uint32_t bf = a.start[cf], ef = bf + a.count[cf];
            for (uint32_t j = bf; j < ef; ++j) {
                pse_pair(pi, r2, wi, a.p[j], w0[j], a.nu_dt2, ex, ey, ez);
                ax -= ex;
                ay -= ey;
                az -= ez;
            }
        }
    }
    float4 o = make_float4(wi.x + ax, wi.y + ay, wi.z + az, wi.w);
    if (live) {
        o.x -= o.x * a.dissipation;
        o.y -= o.y * a.dissipation;
        o.z -= o.z * a.dissipation;
    }
    w1[i] = o;
}

__global__ void k_pse_gather(const uint32_t *__restrict__ idx, const float *__restrict__ angvel,
                             size_t ang_stride, const int32_t *__restrict__ birth,
                             size_t birth_stride, size_t n, float4 *__restrict__ w) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const size_t i = idx[j];
    const float *o = angvel + i * ang_stride;
    const bool alive = birth ? birth[i * birth_stride] >= 0 : true;
    w[j] = make_float4(o[0], o[1], o[2], alive ? 1.f : 0.f);
}

__global__ void k_pse_scatter(const uint32_t *__restrict__ idx, const float4 *__restrict__ w,
                              size_t n, float *__restrict__ angvel, size_t ang_stride) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float *o = angvel + (size_t)idx[j] * ang_stride;
    const float4 v = w[j];
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
}

}  // namespace

vg_status pse_diffuse(const vg_tree *t, float *angvel, size_t ang_stride, const int32_t *birth,
                      size_t birth_stride, float viscosity, float dt, int order, hipStream_t s) {
    TreeView tv;
    VG_TRY(tree_checked_view(t, &tv));
    const size_t n = tv.n;
    if (n == 0) return VG_OK;
    const vg_grid_desc &g = tv.layer[0];
    VG_REQUIRE(g.np[0] >= 2 && g.np[1] >= 2 && g.np[2] >= 2, "leaf grid is degenerate");
    VG_REQUIRE(capacity(g) < (size_t(1) << 32), "leaf grid has 2^32 points or more");
    VG_REQUIRE(order == VG_PSE_REFERENCE_ORDER || order == VG_PSE_SIMULTANEOUS, "bad PSE order");
    const bool sim = order == VG_PSE_SIMULTANEOUS;
    float4 *w = nullptr;
    VG_TRY(salloc((void **)&w, (sim ? 2 : 1) * n * sizeof(float4), s));
    PseArgs a;
    a.nx = g.np[0];
    a.ny = g.np[1];
    a.nz = g.np[2];
    a.start = tv.cell_start;
    a.count = tv.cell_count;
    a.p = tv.sorted_p;
    a.w = w;
    a.nu_dt2 = 2.0f * viscosity * dt;   // 2.0f * mViscosity * timeStep
    a.dissipation = viscosity * dt;     // mViscosity * timeStep
    a.frac = nullptr;
    a.frac_off = nullptr;
    vg_status st = launch("vg_pse_gather", k_pse_gather, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                          tv.sorted_idx, (const float *)angvel, ang_stride, birth, birth_stride,
                          n, w);
    const float4 *result = w;
    if (sim) {
        if (st == VG_OK)
            st = launch("vg_pse_simultaneous", k_pse_simultaneous, dim3(ceil_div(n, 256)),
                        dim3(256), 0, s, a, g, (const float4 *)w, w + n, (uint32_t)n);
        result = w + n;
    } else {
        const uint32_t slices = a.nz - 1;  // cell slices 0 .. nz-2
        const uint32_t cx = a.nx - 1, cy = a.ny - 1;
        const uint32_t steps = (cx - 1) + 2 * (cy - 1) + 1;
        // cells per wavefront step are at most min(cy, (cx + 1) / 2)
        const uint32_t M = cy < (cx + 1) / 2 ? cy : (cx + 1) / 2;
        uint32_t *perm = nullptr, *wstart = nullptr;
        const size_t nsched = (size_t)slices * steps;
        // exchange fractions of the systolic cells, computed over the whole
        // chip before the serial phases (the block sizes are read back to
        // size the buffer: one stream synchronisation)
        const size_t cells = capacity(g);
        uint32_t *fsize = nullptr, *foff = nullptr, *stmp = nullptr;
        unsigned long long *dtot = nullptr;
        float *frac = nullptr;
        if (st == VG_OK) st = salloc((void **)&fsize, cells * sizeof(uint32_t), s);
        if (st == VG_OK) st = salloc((void **)&foff, cells * sizeof(uint32_t), s);
        if (st == VG_OK) st = salloc((void **)&stmp, scan_scratch_elems(cells) * sizeof(uint32_t), s);
        if (st == VG_OK) st = salloc((void **)&dtot, sizeof(unsigned long long), s);
        if (st == VG_OK) st = launch("vg_pse_frac_size", k_pse_zero64, dim3(1), dim3(1), 0, s, dtot);
        if (st == VG_OK)
            st = launch("vg_pse_frac_size", k_pse_frac_size, dim3(ceil_div(cells, 256)), dim3(256),
                        0, s, a, (uint32_t)cells, fsize, dtot);
        if (st == VG_OK) st = exclusive_scan_u32(fsize, foff, cells, stmp, s);
        unsigned long long total = 0;
        if (st == VG_OK) {
            hipError_t e = hipMemcpyAsync(&total, dtot, sizeof total, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) {
                set_error(std::string("PSE fraction sizes: ") + hipGetErrorString(e));
                st = VG_EDEVICE;
            }
        }
        // blocks of 2^31 floats or more keep the in-chain fractions, and so
        // does VG_PSE_FRACTIONS=0 in the environment (the same bits either
        // way; the switch keeps the in-chain path testable)
        const char *env = getenv("VG_PSE_FRACTIONS");
        const bool precompute = !(env && env[0] == '0');
        if (st == VG_OK && precompute && total > 0 && total < (1ull << 31)) {
            st = salloc((void **)&frac, total * sizeof(float), s);
            if (st == VG_OK) {
                a.frac = frac;
                a.frac_off = foff;
                st = launch("vg_pse_fractions", k_pse_fractions, dim3(ceil_div(cells, 4)),
                            dim3(256), 0, s, a, (uint32_t)cells, frac);
            }
        }
        if (st == VG_OK && M <= 64u * kSchedSlots && nsched * M < (size_t(1) << 28)) {
            st = salloc((void **)&perm, nsched * M * sizeof(uint32_t), s);
            if (st == VG_OK)
                st = salloc((void **)&wstart, nsched * (kPseWaves + 1) * sizeof(uint32_t), s);
            if (st == VG_OK)
                st = launch("vg_pse_schedule", k_pse_schedule, dim3(steps, slices), dim3(64), 0, s,
                            a, steps, M, perm, wstart);
        }
        for (uint32_t phase = 0; phase < 2 && st == VG_OK; ++phase) {
            const uint32_t groups = (slices - phase + 1) / 2;
            if (groups)
                st = launch("vg_pse_phase", k_pse_phase, dim3(groups), dim3(kPseThreads), 0, s,
                            a, phase, (const uint32_t *)perm, (const uint32_t *)wstart, M);
        }
        if (perm) sfree(perm, s);
        if (wstart) sfree(wstart, s);
        if (frac) sfree(frac, s);
        if (fsize) sfree(fsize, s);
        if (foff) sfree(foff, s);
        if (stmp) sfree(stmp, s);
        if (dtot) sfree(dtot, s);
    }
    if (st == VG_OK)
        st = launch("vg_pse_scatter", k_pse_scatter, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                    tv.sorted_idx, result, n, angvel, ang_stride);
    sfree(w, s);
    return st;
}

}  // namespace vg

using namespace vg;

VG_API vg_status vg_tree_partition(const vg_tree *t, const uint32_t **cell_start,
                                   const uint32_t **cell_count, const uint32_t **vorton_index,
                                   size_t *cells) {
    VG_REQUIRE(cell_start && cell_count && vorton_index && cells, "null argument");
    TreeView tv;
    VG_TRY(tree_checked_view(t, &tv));
    *cell_start = tv.cell_start;
    *cell_count = tv.cell_count;
    *vorton_index = tv.sorted_idx;
    *cells = capacity(tv.layer[0]);
    return VG_OK;
}

VG_API vg_status vg_tree_pse_diffuse(const vg_tree *t, float *angvel, size_t ang_stride,
                                     const int32_t *birth, size_t birth_stride, float viscosity,
                                     float dt, int order, vg_stream s) {
    VG_REQUIRE(angvel && ang_stride >= 3, "bad angular velocity array");
    VG_REQUIRE(!birth || birth_stride >= 1, "bad birth stride");
    VG_REQUIRE(viscosity * dt <= 1.0f, "viscosity * dt exceeds 1 (vorticity would flip)");
    return pse_diffuse(t, angvel, ang_stride, birth, birth_stride, viscosity, dt, order,
                       (hipStream_t)s);
}
