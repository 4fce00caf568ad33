// vg_grid.hip -- grid operators of the differential path and the particle
// operators: restrict / prolong (uniformGrid.h:1453-1718), the zebra SOR
// smoother (uniformGridMath.cpp:1116-1405), the fused Jacobian+curl
// (:207-451), FP64 residual norms, tracer interpolation + Euler advection
// (pclOpAdvect.cpp:42-152, pclOpEvolve.cpp:36-123) and the AABB reduction
// (pclOpFindBoundingBox.cpp:20-148).
//
// All stencil arithmetic here is HBM/MALL bound and compiled with FP
// contraction off, so results match the reference's operation order bit for
// bit.
#include <cstring>
#include <vector>

#include "vg_internal.h"

namespace vg {
namespace {

__global__ void k_zero(uint32_t *__restrict__ p, size_t nwords) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nwords) p[i] = 0u;
}

__global__ void k_zero_bytes(uint8_t *__restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ------------------------------------------------------------------ restrict / prolong
__global__ void k_inject(vg_grid_desc lo, float *__restrict__ lf, vg_grid_desc hi,
                         const float *__restrict__ hf, uint32_t dx, uint32_t dy, uint32_t dz) {
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= capacity(lo)) return;
    const uint32_t ix = (uint32_t)(o % lo.np[0]);
    const uint32_t iy = (uint32_t)((o / lo.np[0]) % lo.np[1]);
    const uint32_t iz = (uint32_t)(o / ((size_t)lo.np[0] * lo.np[1]));
    const size_t h = point_offset(hi, ix * dx, iy * dy, iz * dz);
    lf[3 * o] = hf[3 * h];
    lf[3 * o + 1] = hf[3 * h + 1];
    lf[3 * o + 2] = hf[3 * h + 2];
}

// UpSampleSlice INTERIOR_ONLY with the sampling nudges of uniformGrid.h:1647-1652.
__global__ void k_prolong_interior(vg_grid_desc hi, float *__restrict__ hf, vg_grid_desc lo,
                                   const float *__restrict__ lf, float3 pmin, float3 sp) {
#pragma clang fp contract(off)
    const uint32_t ix = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t iy = 1 + blockIdx.y;
    const uint32_t iz = 1 + blockIdx.z;
    if (ix + 1 >= hi.np[0]) return;
    float3 p;
    p.x = pmin.x + (float)ix * sp.x;
    p.y = pmin.y + (float)iy * sp.y;
    p.z = pmin.z + (float)iz * sp.z;
    const float3 v = interpolate(lo, lf, p);
    const size_t o = point_offset(hi, ix, iy, iz);
    hf[3 * o] = v.x;
    hf[3 * o + 1] = v.y;
    hf[3 * o + 2] = v.z;
}

// ------------------------------------------------------------------ zebra SOR
struct SorCoef {
    float r2x, r2y, r2z, h, relax, omr;
};

SorCoef sor_coef(const vg_grid_desc &g) {
#pragma clang fp contract(off)
    SorCoef c;
    const float rsx = 1.0f / g.cell[0], rsy = 1.0f / g.cell[1];
    const float rsz = g.cell[2] > FLT_EPSILON ? 1.0f / g.cell[2] : 0.0f;
    c.r2x = rsx * rsx;
    c.r2y = rsy * rsy;
    c.r2z = rsz * rsz;
    c.h = 0.5f / (c.r2x + c.r2y + c.r2z);
    c.relax = 1.72f;
    c.omr = 1.0f - c.relax;
    return c;
}

// One colour pass of StepTowardVectorPoissonSolution: one thread per row
// (iy, iz) of the colour, x swept sequentially in place so each point sees
// its already-updated x-1 neighbour exactly as the reference does.  Only
// that neighbour chains the steps: the other inputs of a step (the old x+1
// value, the four rows of the other colour, f) are loaded kSorAhead steps at
// a time before those steps run, so the row's loads are in flight together
// instead of one round trip per point.  Loading x+1..x+kSorAhead before
// writing x..x+kSorAhead-1 reads the same old values the sequential loop
// does, so the result is unchanged bit for bit.
constexpr int kSorAhead = 8;

__global__ __launch_bounds__(64) void k_sor_rows(vg_grid_desc g, float *__restrict__ s,
                                                 const float *__restrict__ lap, int color,
                                                 SorCoef c) {
#pragma clang fp contract(off)
    const uint32_t iz = 1 + blockIdx.y;
    const uint32_t shift = ((iz % 2u) == (uint32_t)color) ? 1u : 0u;
    const uint32_t iy = 1 + shift + 2 * (blockIdx.x * blockDim.x + threadIdx.x);
    const size_t nx = g.np[0], ny = g.np[1];
    if (iy + 1 >= ny) return;
    const size_t nxy = nx * ny;
    const size_t row = (size_t)iy * nx + (size_t)iz * nxy;
    float lx = s[3 * row], ly = s[3 * row + 1], lz = s[3 * row + 2];          // x-1 (new)
    float cx = s[3 * (row + 1)], cy = s[3 * (row + 1) + 1], cz = s[3 * (row + 1) + 2];  // x
    for (size_t ix0 = 1; ix0 + 1 < nx; ix0 += kSorAhead) {
        // per step: old x+1 (3), y+1 + y-1 (3), z+1 + z-1 (3) as the sums the
        // update uses, f (3)
        float xp[kSorAhead][3], ys[kSorAhead][3], zs[kSorAhead][3], ff[kSorAhead][3];
#pragma unroll
        for (int u = 0; u < kSorAhead; ++u) {
            const size_t ix = ix0 + u;
            if (ix + 1 >= nx) break;
            const size_t o = row + ix;
            const float *a = s + 3 * (o + 1);
            const float *ym = s + 3 * (o - nx), *yp = s + 3 * (o + nx);
            const float *zm = s + 3 * (o - nxy), *zp = s + 3 * (o + nxy);
            const float *f = lap + 3 * o;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                xp[u][k] = a[k];
                ys[u][k] = yp[k] + ym[k];
                zs[u][k] = zp[k] + zm[k];
                ff[u][k] = f[k];
            }
        }
#pragma unroll
        for (int u = 0; u < kSorAhead; ++u) {
            const size_t ix = ix0 + u;
            if (ix + 1 >= nx) break;
            const size_t o = row + ix;
            const float vx = ((xp[u][0] + lx) * c.r2x + ys[u][0] * c.r2y + zs[u][0] * c.r2z - ff[u][0]) * c.h;
            const float vy = ((xp[u][1] + ly) * c.r2x + ys[u][1] * c.r2y + zs[u][1] * c.r2z - ff[u][1]) * c.h;
            const float vz = ((xp[u][2] + lz) * c.r2x + ys[u][2] * c.r2y + zs[u][2] * c.r2z - ff[u][2]) * c.h;
            const float nxv = c.omr * cx + c.relax * vx;
            const float nyv = c.omr * cy + c.relax * vy;
            const float nzv = c.omr * cz + c.relax * vz;
            s[3 * o] = nxv;
            s[3 * o + 1] = nyv;
            s[3 * o + 2] = nzv;
            lx = nxv; ly = nyv; lz = nzv;
            cx = xp[u][0]; cy = xp[u][1]; cz = xp[u][2];
        }
    }
}

// The same pass with one lane per row AND vector component: the three
// components of a row are independent chains, so they run on three adjacent
// lanes (21 rows per wave; their loads of one point stay one 12-byte span)
// and a pass puts three times as many waves on the chip.  That pays on large
// grids (129x129x257: 86 against 127 us per colour) and costs on small ones
// (65x65x129: 28 against 19 us), where k_sor_rows's three chains per lane
// hide each other's latency better; sor_solve picks by the row count.  The
// arithmetic per point is k_sor_rows's, bit for bit.
constexpr uint32_t kSorRowsPerWave = 21;
constexpr size_t kSorSplitRows = 8192;  // rows per colour from which to split

__global__ __launch_bounds__(64) void k_sor_rows_split(vg_grid_desc g, float *__restrict__ s,
                                                 const float *__restrict__ lap, int color,
                                                 SorCoef c) {
#pragma clang fp contract(off)
    const uint32_t r = threadIdx.x / 3u, k = threadIdx.x - 3u * r;
    if (r >= kSorRowsPerWave) return;
    const uint32_t iz = 1 + blockIdx.y;
    const uint32_t shift = ((iz % 2u) == (uint32_t)color) ? 1u : 0u;
    const uint32_t iy = 1 + shift + 2 * (blockIdx.x * kSorRowsPerWave + r);
    const size_t nx = g.np[0], ny = g.np[1];
    if (iy + 1 >= ny) return;
    const size_t nxy = nx * ny;
    const size_t row = (size_t)iy * nx + (size_t)iz * nxy;
    float *sk = s + k;
    const float *fk = lap + k;
    float l = sk[3 * row];         // x-1 (new)
    float cur = sk[3 * (row + 1)];  // x (old)
    for (size_t ix0 = 1; ix0 + 1 < nx; ix0 += kSorAhead) {
        // per step: old x+1, y+1 + y-1 and z+1 + z-1 as the sums the update
        // uses, f
        float xp[kSorAhead], ys[kSorAhead], zs[kSorAhead], ff[kSorAhead];
#pragma unroll
        for (int u = 0; u < kSorAhead; ++u) {
            const size_t ix = ix0 + u;
            if (ix + 1 >= nx) break;
            const size_t o = row + ix;
            xp[u] = sk[3 * (o + 1)];
            ys[u] = sk[3 * (o + nx)] + sk[3 * (o - nx)];
            zs[u] = sk[3 * (o + nxy)] + sk[3 * (o - nxy)];
            ff[u] = fk[3 * o];
        }
#pragma unroll
        for (int u = 0; u < kSorAhead; ++u) {
            const size_t ix = ix0 + u;
            if (ix + 1 >= nx) break;
            const size_t o = row + ix;
            const float v = ((xp[u] + l) * c.r2x + ys[u] * c.r2y + zs[u] * c.r2z - ff[u]) * c.h;
            const float nv = c.omr * cur + c.relax * v;
            sk[3 * o] = nv;
            l = nv;
            cur = xp[u];
        }
    }
}

// ------------------------------------------------------------------ curl
// ComputeJacobian (one-sided 1/h on faces, central 1/(2h) inside) fused with
// ComputeCurlFromJacobian; j[a][b] = d v_b / d a.
__global__ void k_curl(vg_grid_desc g, const float *__restrict__ v, float *__restrict__ out,
                       float3 r, float3 hr) {
#pragma clang fp contract(off)
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= capacity(g)) return;
    const size_t n[3] = {g.np[0], g.np[1], g.np[2]};
    const size_t idx[3] = {o % n[0], (o / n[0]) % n[1], o / (n[0] * n[1])};
    const size_t stride[3] = {1, n[0], n[0] * n[1]};
    const float rr[3] = {r.x, r.y, r.z}, hh[3] = {hr.x, hr.y, hr.z};
    float j[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const size_t st = stride[a];
        const float *lo, *hi;
        float scale;
        if (idx[a] == 0) { lo = v + 3 * o; hi = v + 3 * (o + st); scale = rr[a]; }
        else if (idx[a] == n[a] - 1) { lo = v + 3 * (o - st); hi = v + 3 * o; scale = rr[a]; }
        else { lo = v + 3 * (o - st); hi = v + 3 * (o + st); scale = hh[a]; }
#pragma unroll
        for (int b = 0; b < 3; ++b) j[a][b] = (hi[b] - lo[b]) * scale;
    }
    out[3 * o] = j[1][2] - j[2][1];
    out[3 * o + 1] = j[2][0] - j[0][2];
    out[3 * o + 2] = j[0][1] - j[1][0];
}

// ------------------------------------------------------------------ FP64 residual
constexpr int kRedBlock = 256;

__global__ __launch_bounds__(kRedBlock) void k_residual_partial(vg_grid_desc g,
                                                                const float *__restrict__ u,
                                                                const float *__restrict__ f,
                                                                double r2x, double r2y,
                                                                double r2z,
                                                                double *__restrict__ part) {
    __shared__ double sn[kRedBlock], sd[kRedBlock];
    const size_t nx = g.np[0], ny = g.np[1], nz = g.np[2], nxy = nx * ny;
    const size_t total = capacity(g);
    double num = 0.0, den = 0.0;
    for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < total;
         o += (size_t)gridDim.x * blockDim.x) {
        const size_t ix = o % nx, iy = (o / nx) % ny, iz = o / nxy;
        if (ix == 0 || iy == 0 || iz == 0 || ix + 1 == nx || iy + 1 == ny || iz + 1 == nz) continue;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double uc = u[3 * o + c];
            const double lap = ((double)u[3 * (o + 1) + c] + (double)u[3 * (o - 1) + c] - 2.0 * uc) * r2x +
                               ((double)u[3 * (o + nx) + c] + (double)u[3 * (o - nx) + c] - 2.0 * uc) * r2y +
                               ((double)u[3 * (o + nxy) + c] + (double)u[3 * (o - nxy) + c] - 2.0 * uc) * r2z;
            const double fc = f[3 * o + c];
            const double res = fc - lap;
            num += res * res;
            den += fc * fc;
        }
    }
    sn[threadIdx.x] = num;
    sd[threadIdx.x] = den;
    __syncthreads();
    for (int w = kRedBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sn[threadIdx.x] += sn[threadIdx.x + w];
            sd[threadIdx.x] += sd[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sn[0];
        part[2 * blockIdx.x + 1] = sd[0];
    }
}

// ------------------------------------------------------------------ particles
__global__ void k_advect(vg_grid_desc g, const float *__restrict__ field, float *__restrict__ pos,
                         float *__restrict__ vel, size_t stride, size_t n, float dt, float gain,
                         int zero_extent) {
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float *p = pos + i * stride;
    float3 q = make_float3(p[0], p[1], p[2]);
    float3 v;
    if (zero_extent) {
        if (!vel) return;
        const float *pv = vel + i * stride;
        v = make_float3(pv[0], pv[1], pv[2]);
    } else {
        v = interpolate(g, field, q);
        if (vel) {
            float *pv = vel + i * stride;
            if (gain != 1.0f) {
                v.x = pv[0] + gain * (v.x - pv[0]);
                v.y = pv[1] + gain * (v.y - pv[1]);
                v.z = pv[2] + gain * (v.z - pv[2]);
            }
            pv[0] = v.x;
            pv[1] = v.y;
            pv[2] = v.z;
        }
    }
    p[0] = q.x + v.x * dt;
    p[1] = q.y + v.y * dt;
    p[2] = q.z + v.z * dt;
}

__global__ void k_assign_velocity(vg_grid_desc g, const float *__restrict__ field,
                                  const float *__restrict__ pos, float *__restrict__ vel,
                                  size_t stride, size_t n, float gain) {
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *p = pos + i * stride;
    float3 v = interpolate(g, field, make_float3(p[0], p[1], p[2]));
    float *pv = vel + i * stride;
    if (gain != 1.0f) {
        v.x = pv[0] + gain * (v.x - pv[0]);
        v.y = pv[1] + gain * (v.y - pv[1]);
        v.z = pv[2] + gain * (v.z - pv[2]);
    }
    pv[0] = v.x;
    pv[1] = v.y;
    pv[2] = v.z;
}

__global__ void k_evolve(float *__restrict__ pos, const float *__restrict__ vel, size_t stride,
                         size_t n, float dt) {
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float *p = pos + i * stride;
    const float *v = vel + i * stride;
    p[0] = p[0] + v[0] * dt;
    p[1] = p[1] + v[1] * dt;
    p[2] = p[2] + v[2] * dt;
}

// ------------------------------------------------------------------ AABB
__device__ __forceinline__ uint32_t ord_enc(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
float ord_dec(uint32_t u) {
    u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Pass 1: each block reduces a grid-stride share of the particles and writes
// its 6 extrema (ordered-integer encoded) to part[6 * block].  Pass 2 folds
// the block results.  Min/max are exact, so the result is order-independent.
__global__ __launch_bounds__(256) void k_bbox(const float *__restrict__ pos, size_t stride,
                                              size_t n, uint32_t *__restrict__ part) {
    __shared__ uint32_t s[4][6];
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const float *p = pos + i * stride;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            mn[a] = fminf(mn[a], p[a]);
            mx[a] = fmaxf(mx[a], p[a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        for (int off = 32; off > 0; off >>= 1) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], off, 64));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off, 64));
        }
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            s[wave][a] = ord_enc(mn[a]);
            s[wave][3 + a] = ord_enc(mx[a]);
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        uint32_t v = s[0][threadIdx.x];
        for (int w = 1; w < 4; ++w)
            v = threadIdx.x < 3 ? min(v, s[w][threadIdx.x]) : max(v, s[w][threadIdx.x]);
        part[6 * blockIdx.x + threadIdx.x] = v;
    }
}

__global__ __launch_bounds__(256) void k_bbox_final(const uint32_t *__restrict__ part,
                                                    uint32_t nblocks, uint32_t *__restrict__ res) {
    __shared__ uint32_t s[256][6];
    uint32_t v[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    for (uint32_t b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int k = 0; k < 6; ++k)
            v[k] = k < 3 ? min(v[k], part[6 * b + k]) : max(v[k], part[6 * b + k]);
    for (int k = 0; k < 6; ++k) s[threadIdx.x][k] = v[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; ++k)
                s[threadIdx.x][k] = k < 3 ? min(s[threadIdx.x][k], s[threadIdx.x + w][k])
                                          : max(s[threadIdx.x][k], s[threadIdx.x + w][k]);
        __syncthreads();
    }
    if (threadIdx.x < 6) res[threadIdx.x] = s[0][threadIdx.x];
}

}  // namespace

vg_status zero_fill(void *p, size_t bytes, hipStream_t s) {
    if (bytes == 0) return VG_OK;
    if (bytes % 4 == 0 && ((uintptr_t)p & 3) == 0)
        return launch("vg_zero", k_zero, dim3(ceil_div(bytes / 4, 256)), dim3(256), 0, s,
                      (uint32_t *)p, bytes / 4);
    return launch("vg_zero", k_zero_bytes, dim3(ceil_div(bytes, 256)), dim3(256), 0, s,
                  (uint8_t *)p, bytes);
}

// Exposed to the multigrid driver.
vg_status sor_solve(const vg_grid_desc &g, float *soln, const float *lap, uint32_t steps,
                    hipStream_t s) {
    if (g.np[0] < 3 || g.np[1] < 3 || g.np[2] < 3) return VG_OK;  // no interior
    const SorCoef c = sor_coef(g);
    const uint32_t rows_half = (g.np[1] - 1) / 2 + 1;
    const bool split = (size_t)rows_half * (g.np[2] - 2) >= kSorSplitRows;
    const dim3 grid(ceil_div(rows_half, split ? kSorRowsPerWave : 64u), g.np[2] - 2);
    auto k = split ? k_sor_rows_split : k_sor_rows;
    for (uint32_t it = 0; it < steps; ++it) {
        VG_TRY(launch("vg_sor_rows", k, grid, dim3(64), 0, s, g, soln, lap, 0, c));
        VG_TRY(launch("vg_sor_rows", k, grid, dim3(64), 0, s, g, soln, lap, 1, c));
    }
    return VG_OK;
}

vg_status inject(const vg_grid_desc &lo, float *lf, const vg_grid_desc &hi, const float *hf,
                 hipStream_t s) {
    const uint32_t dx = (hi.np[0] - 1) / (lo.np[0] - 1), dy = (hi.np[1] - 1) / (lo.np[1] - 1),
                   dz = (hi.np[2] - 1) / (lo.np[2] - 1);
    return launch("vg_restrict_inject", k_inject, dim3(ceil_div(capacity(lo), 256)), dim3(256), 0,
                  s, lo, lf, hi, hf, dx, dy, dz);
}

vg_status prolong_interior(const vg_grid_desc &hi, float *hf, const vg_grid_desc &lo,
                           const float *lf, hipStream_t s) {
#pragma clang fp contract(off)
    if (hi.np[0] < 3 || hi.np[1] < 3 || hi.np[2] < 3) return VG_OK;
    float3 pmin, sp;
    const float e = FLT_EPSILON;
    pmin.x = hi.min[0] + (fabsf(hi.min[0]) * 2.0f) * e;
    pmin.y = hi.min[1] + (fabsf(hi.min[1]) * 2.0f) * e;
    pmin.z = hi.min[2] + (fabsf(hi.min[2]) * 2.0f) * e;
    const float shrink = 1.0f - 4.0f * FLT_EPSILON;
    sp.x = hi.cell[0] * shrink;
    sp.y = hi.cell[1] * shrink;
    sp.z = hi.cell[2] * shrink;
    const dim3 grid(ceil_div(hi.np[0] - 2, 64), hi.np[1] - 2, hi.np[2] - 2);
    return launch("vg_prolong_interior", k_prolong_interior, grid, dim3(64), 0, s, hi, hf, lo, lf,
                  pmin, sp);
}

vg_status curl(const vg_grid_desc &g, const float *a, float *out, hipStream_t s) {
#pragma clang fp contract(off)
    const float rx = 1.0f / g.cell[0], ry = 1.0f / g.cell[1];
    const float rz = g.cell[2] > FLT_EPSILON ? 1.0f / g.cell[2] : 0.0f;
    const float3 r = make_float3(rx, ry, rz);
    const float3 hr = make_float3(0.5f * rx, 0.5f * ry, 0.5f * rz);
    return launch("vg_curl", k_curl, dim3(ceil_div(capacity(g), 256)), dim3(256), 0, s, g, a, out,
                  r, hr);
}

vg_status residual_norm(const vg_grid_desc &g, const float *u, const float *f, double *rel,
                        hipStream_t s) {
    const double rx = 1.0 / (double)g.cell[0], ry = 1.0 / (double)g.cell[1];
    const double rz = g.cell[2] > FLT_EPSILON ? 1.0 / (double)g.cell[2] : 0.0;
    const uint32_t nb = 1024;
    double *part = nullptr;
    VG_TRY(salloc((void **)&part, 2 * nb * sizeof(double), s));
    vg_status st = launch("vg_residual_norm", k_residual_partial, dim3(nb), dim3(kRedBlock), 0, s,
                          g, u, f, rx * rx, ry * ry, rz * rz, part);
    std::vector<double> h(2 * nb);
    if (st == VG_OK) {
        hipError_t e = hipMemcpyAsync(h.data(), part, 2 * nb * sizeof(double),
                                      hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = device_error(e, "residual copy", __FILE__, __LINE__);
    }
    sfree(part, s);
    VG_TRY(st);
    double num = 0.0, den = 0.0;
    for (uint32_t b = 0; b < nb; ++b) {
        num += h[2 * b];
        den += h[2 * b + 1];
    }
    *rel = den > 0.0 ? sqrt(num / den) : sqrt(num);
    return VG_OK;
}

}  // namespace vg

using namespace vg;

static vg_status check_grid(const vg_grid_desc *g) {
    VG_REQUIRE(g, "grid is null");
    VG_REQUIRE(g->np[0] >= 2 && g->np[1] >= 2 && g->np[2] >= 2, "grid has fewer than 2 points per axis");
    return VG_OK;
}

VG_API vg_status vg_downsample_inject(const vg_grid_desc *lo, float *lo_field,
                                      const vg_grid_desc *hi, const float *hi_field,
                                      vg_stream s) {
    VG_TRY(check_grid(lo));
    VG_TRY(check_grid(hi));
    VG_REQUIRE(lo_field && hi_field, "null field");
    for (int a = 0; a < 3; ++a)
        VG_REQUIRE((hi->np[a] - 1) % (lo->np[a] - 1) == 0, "grids are not nested");
    return inject(*lo, lo_field, *hi, hi_field, (hipStream_t)s);
}

VG_API vg_status vg_upsample_interior(const vg_grid_desc *hi, float *hi_field,
                                      const vg_grid_desc *lo, const float *lo_field,
                                      vg_stream s) {
    VG_TRY(check_grid(lo));
    VG_TRY(check_grid(hi));
    VG_REQUIRE(lo_field && hi_field, "null field");
    return prolong_interior(*hi, hi_field, *lo, lo_field, (hipStream_t)s);
}

VG_API vg_status vg_sor_solve(const vg_grid_desc *g, float *soln, const float *lap,
                              uint32_t steps, vg_stream s) {
    VG_TRY(check_grid(g));
    VG_REQUIRE(soln && lap, "null field");
    uint32_t n = steps;
    if (n == 0) {  // numStepsAuto = 2 * max dimension (uniformGridMath.cpp:1383)
        uint32_t m = g->np[0];
        if (g->np[1] > m) m = g->np[1];
        if (g->np[2] > m) m = g->np[2];
        n = 2 * m;
    }
    return sor_solve(*g, soln, lap, n, (hipStream_t)s);
}

VG_API vg_status vg_curl(const vg_grid_desc *g, const float *a, float *out, vg_stream s) {
    VG_TRY(check_grid(g));
    VG_REQUIRE(a && out, "null field");
    return curl(*g, a, out, (hipStream_t)s);
}

VG_API vg_status vg_residual_norm(const vg_grid_desc *g, const float *u, const float *f,
                                  double *rel, vg_stream s) {
    VG_TRY(check_grid(g));
    VG_REQUIRE(u && f && rel, "null argument");
    return residual_norm(*g, u, f, rel, (hipStream_t)s);
}

VG_API vg_status vg_assign_velocity_from_field(const vg_grid_desc *g, const float *field,
                                               const float *pos, float *vel, size_t stride,
                                               size_t n, float gain, vg_stream s) {
    VG_REQUIRE(g, "grid is null");
    if (n == 0) return VG_OK;
    VG_REQUIRE(pos && vel && stride >= 3, "bad particle arrays");
    if (g->extent[0] == 0.f && g->extent[1] == 0.f && g->extent[2] == 0.f) return VG_OK;
    VG_TRY(check_grid(g));
    VG_REQUIRE(field, "field is null");
    return launch("vg_assign_velocity", k_assign_velocity, dim3(ceil_div(n, 256)), dim3(256), 0,
                  (hipStream_t)s, *g, field, pos, vel, stride, n, gain);
}

VG_API vg_status vg_evolve(float *pos, const float *vel, size_t stride, size_t n, float dt,
                           vg_stream s) {
    if (n == 0) return VG_OK;
    VG_REQUIRE(pos && vel && stride >= 3, "bad particle arrays");
    return launch("vg_evolve", k_evolve, dim3(ceil_div(n, 256)), dim3(256), 0, (hipStream_t)s,
                  pos, vel, stride, n, dt);
}

VG_API vg_status vg_advect_tracers(const vg_grid_desc *g, const float *field, float *pos,
                                   float *vel, size_t stride, size_t n, float dt, float gain,
                                   vg_stream s) {
    VG_REQUIRE(g, "grid is null");
    if (n == 0) return VG_OK;
    VG_REQUIRE(pos && stride >= 3, "bad particle arrays");
    VG_REQUIRE(vel || gain == 1.0f, "gain != 1 needs the stored velocity");
    const int zero = g->extent[0] == 0.f && g->extent[1] == 0.f && g->extent[2] == 0.f;
    if (!zero) {
        VG_TRY(check_grid(g));
        VG_REQUIRE(field, "field is null");
    }
    return launch("vg_advect_tracers", k_advect, dim3(ceil_div(n, 256)), dim3(256), 0,
                  (hipStream_t)s, *g, field, pos, vel, stride, n, dt, gain, zero);
}

namespace vg {
// Boxes of `count` particle arrays with one host wait: per array one
// reduction pass and one fold, then a single readback of all extrema.
vg_status find_bounding_boxes(uint32_t count, const float *const *pos, const size_t *stride,
                              const size_t *n, float *mn, float *mx, hipStream_t s) {
    VG_REQUIRE(count <= 8, "at most 8 arrays per call");
    uint32_t nb[8] = {};
    size_t words = 6 * (size_t)count;
    for (uint32_t k = 0; k < count; ++k) {
        if (n[k] == 0) continue;
        VG_REQUIRE(pos[k] && stride[k] >= 3, "bad particle array");
        nb[k] = ceil_div(n[k], 256 * 8);
        if (nb[k] > 1024) nb[k] = 1024;
        if (nb[k] < 1) nb[k] = 1;
        words += 6 * (size_t)nb[k];
    }
    uint32_t *res = nullptr;
    VG_TRY(salloc((void **)&res, words * sizeof(uint32_t), s));
    vg_status st = VG_OK;
    uint32_t *part = res + 6 * count;
    for (uint32_t k = 0; k < count && st == VG_OK; ++k) {
        if (n[k] == 0) continue;
        st = launch("vg_bbox", k_bbox, dim3(nb[k]), dim3(256), 0, s, pos[k], stride[k], n[k], part);
        if (st == VG_OK)
            st = launch("vg_bbox_final", k_bbox_final, dim3(1), dim3(256), 0, s,
                        (const uint32_t *)part, nb[k], res + 6 * k);
        part += 6 * (size_t)nb[k];
    }
    uint32_t h[48];
    if (st == VG_OK) {
        hipError_t e = hipMemcpyAsync(h, res, 6 * count * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = device_error(e, "bbox copy", __FILE__, __LINE__);
    }
    sfree(res, s);
    VG_TRY(st);
    for (uint32_t k = 0; k < count; ++k) {
        if (n[k] == 0) continue;
        for (int a = 0; a < 3; ++a) {
            const float lo = ord_dec(h[6 * k + a]), hi = ord_dec(h[6 * k + 3 + a]);
            if (lo < mn[3 * k + a]) mn[3 * k + a] = lo;
            if (hi > mx[3 * k + a]) mx[3 * k + a] = hi;
        }
    }
    return VG_OK;
}
}  // namespace vg

VG_API vg_status vg_find_bounding_box(const float *pos, size_t stride, size_t n, float mn[3],
                                      float mx[3], vg_stream stream) {
    VG_REQUIRE(mn && mx, "null bounds");
    if (n == 0) return VG_OK;
    VG_REQUIRE(pos && stride >= 3, "bad particle array");
    hipStream_t s = (hipStream_t)stream;
    uint32_t nb = ceil_div(n, 256 * 8);
    if (nb > 1024) nb = 1024;
    if (nb < 1) nb = 1;
    uint32_t *res = nullptr;
    VG_TRY(salloc((void **)&res, (6 + 6 * (size_t)nb) * sizeof(uint32_t), s));
    vg_status st = launch("vg_bbox", k_bbox, dim3(nb), dim3(256), 0, s, pos, stride, n, res + 6);
    if (st == VG_OK)
        st = launch("vg_bbox_final", k_bbox_final, dim3(1), dim3(256), 0, s,
                    (const uint32_t *)(res + 6), nb, res);
    uint32_t h[6] = {0, 0, 0, 0, 0, 0};
    if (st == VG_OK) {
        hipError_t e = hipMemcpyAsync(h, res, sizeof h, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = device_error(e, "bbox copy", __FILE__, __LINE__);
    }
    sfree(res, s);
    VG_TRY(st);
    for (int a = 0; a < 3; ++a) {
        const float lo = ord_dec(h[a]), hi = ord_dec(h[3 + a]);
        if (lo < mn[a]) mn[a] = lo;
        if (hi > mx[a]) mx[a] = hi;
    }
    return VG_OK;
}
