// vg_multigrid.hip -- vector-Poisson solves for the vector potential,
// nabla^2 A = -w, with Dirichlet values on the 6 faces.
//
// 1. Reference cascade (VortonSim::ComputeVectorPotential, vortonSim.cpp:
//    2523-2571): SOR on the finest grid, injection-restrict both A and -w down
//    the NestedGrid pyramid with SOR on each layer, then prolong-overwrite the
//    interiors back up with SOR after each.  Kept for parity.
// 2. Residual-correction V-cycles for a converged solve: red-black point
//    Gauss-Seidel smoothing, full-weighting restriction, trilinear
//    correction, FP64 Gauss-Seidel on the coarsest layer.  The finest-layer
//    solution is carried as a compensated FP32 pair (hi + lo) and refined
//    against FP64 residuals (mixed-precision iterative refinement): an FP32
//    field alone cannot reach a 1e-6 relative residual on large grids, since
//    its rounding noise amplified by the Laplacian sits near 1e-4.
#include <cmath>
#include <cstdlib>
#include <vector>

#include "vg_internal.h"
#include "vg_rbgs.h"

namespace vg {
vg_status sor_solve(const vg_grid_desc &g, float *soln, const float *lap, uint32_t steps,
                    hipStream_t s);
vg_status inject(const vg_grid_desc &lo, float *lf, const vg_grid_desc &hi, const float *hf,
                 hipStream_t s);
vg_status prolong_interior(const vg_grid_desc &hi, float *hf, const vg_grid_desc &lo,
                           const float *lf, hipStream_t s);
}  // namespace vg

struct vg_multigrid {
    // cascade pyramid (NestedGrid layers)
    vg::Nested ng;
    float *A[vg::kMaxLayers] = {};
    float *W[vg::kMaxLayers] = {};
    size_t cap[vg::kMaxLayers] = {};
    // V-cycle hierarchy
    uint32_t vdepth = 0;
    vg_grid_desc vg[vg::kMaxLayers];
    float *u[vg::kMaxLayers] = {};
    float *f[vg::kMaxLayers] = {};
    float *r[vg::kMaxLayers] = {};
    size_t vcap[vg::kMaxLayers] = {};
    float *lo = nullptr;       // compensation term of the finest solution
    size_t lo_cap = 0;
    double *part = nullptr;    // FP64 norm partials
    uint32_t bottom = 0xffffffffu;  // first level of the one-workgroup bottom solve
};

namespace vg {
namespace {

template <typename T>
vg_status grow(T **p, size_t *cap, size_t need) {
    if (need <= *cap && *p) return VG_OK;
    if (*p) VG_HIP(hipFree(*p));
    *p = nullptr;
    *cap = 0;
    VG_HIP(hipMalloc((void **)p, need * sizeof(T)));
    *cap = need;
    return VG_OK;
}

// Point o of a grid as (ix, iy, iz), in 32-bit arithmetic (solves require
// fewer than 2^32 points): 64-bit division costs several times more.
struct Pt {
    uint32_t x, y, z;
};
__device__ inline Pt split(uint32_t o, uint32_t nx, uint32_t ny) {
    const uint32_t row = o / nx;
    return Pt{o - row * nx, row % ny, row / ny};
}
__device__ inline bool on_boundary(const Pt &p, const vg_grid_desc &g) {
    return p.x == 0 || p.y == 0 || p.z == 0 || p.x + 1 == g.np[0] || p.y + 1 == g.np[1] ||
           p.z + 1 == g.np[2];
}

__global__ void k_zero_interior(vg_grid_desc g, float *__restrict__ a) {
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= capacity(g)) return;
    const size_t nx = g.np[0], ny = g.np[1], nz = g.np[2];
    const size_t ix = o % nx, iy = (o / nx) % ny, iz = o / (nx * ny);
    if (ix == 0 || iy == 0 || iz == 0 || ix + 1 == nx || iy + 1 == ny || iz + 1 == nz) return;
    a[3 * o] = 0.f;
    a[3 * o + 1] = 0.f;
    a[3 * o + 2] = 0.f;
}

__global__ void k_zero_all(float *__restrict__ a, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.f;
}

struct Lap {
    float r2x, r2y, r2z, diag;  // diag = 2 (r2x + r2y + r2z)
    float inv;                  // 1 / diag
};

Lap lap_coef(const vg_grid_desc &g) {
    Lap c;
    const float rx = 1.0f / g.cell[0], ry = 1.0f / g.cell[1], rz = 1.0f / g.cell[2];
    c.r2x = rx * rx;
    c.r2y = ry * ry;
    c.r2z = rz * rz;
    c.diag = 2.0f * (c.r2x + c.r2y + c.r2z);
    c.inv = 1.0f / c.diag;
    return c;
}

// Per-point operations of the V-cycle, shared by the per-level kernels and
// the one-workgroup bottom solve (k_vcycle_bottom) so that both give the
// same bits: contraction off, fixed operation order.  Arrays are 3 floats
// per point, o the point's offset.

// Red-black Gauss-Seidel update of interior point o.
__device__ __forceinline__ void rbgs_point(float *u, const float *f, uint32_t o, uint32_t nx,
                                           uint32_t nxy, const Lap &c) {
#pragma clang fp contract(off)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float s = (u[3 * (o + 1) + k] + u[3 * (o - 1) + k]) * c.r2x +
                        (u[3 * (o + nx) + k] + u[3 * (o - nx) + k]) * c.r2y +
                        (u[3 * (o + nxy) + k] + u[3 * (o - nxy) + k]) * c.r2z;
        u[3 * o + k] = (s - f[3 * o + k]) * c.inv;
    }
}

// r = f - L u at interior point o.
__device__ __forceinline__ void residual_point(const float *u, const float *f, float *r,
                                               uint32_t o, uint32_t nx, uint32_t nxy,
                                               const Lap &c) {
#pragma clang fp contract(off)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float uc = u[3 * o + k];
        const float lap = (u[3 * (o + 1) + k] + u[3 * (o - 1) + k]) * c.r2x +
                          (u[3 * (o + nx) + k] + u[3 * (o - nx) + k]) * c.r2y +
                          (u[3 * (o + nxy) + k] + u[3 * (o - nxy) + k]) * c.r2z - c.diag * uc;
        r[3 * o + k] = f[3 * o + k] - lap;
    }
}

// Full weighting of rhi onto coarse interior point p (weights 8/4/2/1 by
// Manhattan distance, /64).
__device__ __forceinline__ void restrict_point(const Pt &p, const vg_grid_desc &hi,
                                               const float *rhi, float out[3]) {
#pragma clang fp contract(off)
    const uint32_t hx = hi.np[0], hxy = hi.np[0] * hi.np[1];
    const uint32_t c = 2 * p.x + 2 * p.y * hx + 2 * p.z * hxy;
    float acc[3] = {0.f, 0.f, 0.f};
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int md = abs(dx) + abs(dy) + abs(dz);
                const float w = (float)(8 >> md);
                const uint32_t h = c + dx + dy * (int)hx + dz * (int)hxy;
#pragma unroll
                for (int k = 0; k < 3; ++k) acc[k] += w * rhi[3 * h + k];
            }
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = acc[k] * (1.0f / 64.0f);
}

// u_hi at interior point p (offset o) += trilinear(u_lo).
__device__ __forceinline__ void prolong_point(const Pt &p, uint32_t o, float *uhi,
                                              const vg_grid_desc &lo, const float *ulo) {
#pragma clang fp contract(off)
    const uint32_t lx = lo.np[0], lxy = lo.np[0] * lo.np[1];
    const uint32_t x0 = p.x / 2, y0 = p.y / 2, z0 = p.z / 2;
    const uint32_t xs[2] = {x0, x0 + (p.x & 1)}, ys[2] = {y0, y0 + (p.y & 1)},
                   zs[2] = {z0, z0 + (p.z & 1)};
    const float wx = (p.x & 1) ? 0.5f : 1.0f, wy = (p.y & 1) ? 0.5f : 1.0f,
                wz = (p.z & 1) ? 0.5f : 1.0f;
    const int nxs = (p.x & 1) ? 2 : 1, nys = (p.y & 1) ? 2 : 1, nzs = (p.z & 1) ? 2 : 1;
    float acc[3] = {0.f, 0.f, 0.f};
    for (int a = 0; a < nzs; ++a)
        for (int b = 0; b < nys; ++b)
            for (int d = 0; d < nxs; ++d) {
                const uint32_t l = xs[d] + ys[b] * lx + zs[a] * lxy;
                const float w = wx * wy * wz;
#pragma unroll
                for (int k = 0; k < 3; ++k) acc[k] += w * ulo[3 * l + k];
            }
#pragma unroll
    for (int k = 0; k < 3; ++k) uhi[3 * o + k] += acc[k];
}

// FP64 red-black update of interior point o of the coarsest level.
struct Lap64 {
    double r2x, r2y, r2z, inv;
};
__device__ __forceinline__ Lap64 lap64(const vg_grid_desc &g) {
    Lap64 c;
    const double rx = 1.0 / g.cell[0], ry = 1.0 / g.cell[1], rz = 1.0 / g.cell[2];
    c.r2x = rx * rx;
    c.r2y = ry * ry;
    c.r2z = rz * rz;
    c.inv = 1.0 / (2.0 * (c.r2x + c.r2y + c.r2z));
    return c;
}
__device__ __forceinline__ void coarse_point(double *w, const float *f, uint32_t o, uint32_t nx,
                                             uint32_t nxy, const Lap64 &c) {
#pragma clang fp contract(off)
    for (int k = 0; k < 3; ++k) {
        const double s = (w[3 * (o + 1) + k] + w[3 * (o - 1) + k]) * c.r2x +
                         (w[3 * (o + nx) + k] + w[3 * (o - nx) + k]) * c.r2y +
                         (w[3 * (o + nxy) + k] + w[3 * (o - nxy) + k]) * c.r2z;
        w[3 * o + k] = (s - (double)f[3 * o + k]) * c.inv;
    }
}

// `sweeps` FP64 red-black sweeps of the coarsest level on w (in LDS or
// memory).  Up to 64 points the first wave does them alone, one point per
// lane, ordered by wave barriers instead of workgroup barriers (a wave's LDS
// operations complete in order); the caller synchronises the workgroup
// afterwards.
__device__ __forceinline__ void coarse_sweeps_f64(const vg_grid_desc &g, double *w,
                                                  const float *f, int sweeps) {
    const uint32_t nx = g.np[0], nxy = nx * g.np[1], n = nxy * g.np[2];
    const Lap64 c = lap64(g);
    if (n <= 64) {
        if (threadIdx.x >= 64) return;
        const uint32_t o = threadIdx.x;
        bool inner = false;
        uint32_t parity = 0;
        if (o < n) {
            const Pt p = split(o, g.np[0], g.np[1]);
            inner = !on_boundary(p, g);
            parity = (p.x + p.y + p.z) & 1u;
        }
        for (int it = 0; it < sweeps; ++it) {
            for (uint32_t color = 0; color < 2; ++color) {
                if (inner && parity == color) coarse_point(w, f, o, nx, nxy, c);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        return;
    }
    for (int it = 0; it < sweeps; ++it) {
        for (uint32_t color = 0; color < 2; ++color) {
            for (uint32_t o = threadIdx.x; o < n; o += blockDim.x) {
                const Pt p = split(o, g.np[0], g.np[1]);
                if (on_boundary(p, g) || ((p.x + p.y + p.z) & 1u) != color) continue;
                coarse_point(w, f, o, nx, nxy, c);
            }
            __syncthreads();
        }
    }
}

// Red-black point Gauss-Seidel: colour = (ix + iy + iz) & 1.  Threads map to
// the colour's points of one x row: ix = 1 + ((iy + iz + color + 1) & 1) + 2k.
__global__ void k_rbgs(vg_grid_desc g, float *__restrict__ u, const float *__restrict__ f,
                       int color, Lap c) {
    const uint32_t iy = 1 + blockIdx.y, iz = 1 + blockIdx.z;
    const uint32_t first = 1 + ((iy + iz + 1 + (uint32_t)color) & 1u);
    const uint32_t ix = first + 2 * (blockIdx.x * blockDim.x + threadIdx.x);
    const uint32_t nx = g.np[0], nxy = g.np[0] * g.np[1];
    if (ix + 1 >= nx) return;
    rbgs_point(u, f, ix + iy * nx + iz * nxy, nx, nxy, c);
}

// r = f - L u on interior points, 0 on the boundary.
__global__ void k_residual32(vg_grid_desc g, const float *__restrict__ u,
                             const float *__restrict__ f, float *__restrict__ r, Lap c) {
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= capacity(g)) return;
    if (on_boundary(split((uint32_t)o, g.np[0], g.np[1]), g)) {
        r[3 * o] = r[3 * o + 1] = r[3 * o + 2] = 0.f;
        return;
    }
    residual_point(u, f, r, (uint32_t)o, g.np[0], g.np[0] * g.np[1], c);
}

// Full weighting onto coarse interior points; coarse boundary = 0
// (homogeneous Dirichlet correction).  Also clears the coarse correction ulo
// (the coarse level's initial guess).
__global__ void k_restrict_fw(vg_grid_desc lo, float *__restrict__ flo, float *__restrict__ ulo,
                              vg_grid_desc hi, const float *__restrict__ rhi) {
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= capacity(lo)) return;
    ulo[3 * o] = ulo[3 * o + 1] = ulo[3 * o + 2] = 0.f;
    const Pt p = split((uint32_t)o, lo.np[0], lo.np[1]);
    float v[3] = {0.f, 0.f, 0.f};
    if (!on_boundary(p, lo)) restrict_point(p, hi, rhi, v);
#pragma unroll
    for (int k = 0; k < 3; ++k) flo[3 * o + k] = v[k];
}

// u_hi(interior) += trilinear(u_lo).
__global__ void k_prolong_add(vg_grid_desc hi, float *__restrict__ uhi, vg_grid_desc lo,
                              const float *__restrict__ ulo) {
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= capacity(hi)) return;
    const Pt p = split((uint32_t)o, hi.np[0], hi.np[1]);
    if (on_boundary(p, hi)) return;
    prolong_point(p, (uint32_t)o, uhi, lo, ulo);
}

// FP64 coarsest solve: one workgroup, red-black Gauss-Seidel in double, on
// a copy of u in LDS (IN_LDS, grids up to kCoarseLdsPoints) or in `work`.
constexpr uint32_t kCoarseLdsPoints = 2048;  // 48 KB of doubles

template <bool IN_LDS>
__global__ __launch_bounds__(256) void k_coarse_f64(vg_grid_desc g, float *__restrict__ u,
                                                    const float *__restrict__ f, int sweeps,
                                                    double *__restrict__ work) {
    __shared__ double lds[IN_LDS ? 3 * kCoarseLdsPoints : 1];
    double *w = IN_LDS ? lds : work;
    const uint32_t n = g.np[0] * g.np[1] * g.np[2];
    for (uint32_t i = threadIdx.x; i < 3 * n; i += blockDim.x) w[i] = u[i];
    __syncthreads();
    coarse_sweeps_f64(g, w, f, sweeps);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 3 * n; i += blockDim.x) u[i] = (float)w[i];
}

// The bottom of a V-cycle in one workgroup with every array in LDS: from
// level 0 of `b` (u and f read from memory, u written back) down to the
// coarsest, the same steps as the per-level launches (2 red-black sweeps,
// residual, full weighting, FP64 coarsest solve, trilinear correction, 2
// sweeps) with the same per-point arithmetic, so the same bits, without a
// launch per step.
constexpr int kBottomLevels = 4;
constexpr uint32_t kBottomLdsFloats = 16384;  // 64 KB
constexpr int kBottomThreads = 1024;

struct Bottom {
    uint32_t levels;  // 2 .. kBottomLevels; the last is the coarsest
    vg_grid_desc g[kBottomLevels];
    Lap c[kBottomLevels];
    float *u0;
    const float *f0;
    int coarse_sweeps;
};

// LDS floats of the bottom solve: u, f, r per level, plus the coarsest
// level's double copy.
inline size_t bottom_floats(const vg_grid_desc *g, uint32_t levels) {
    size_t n = 0;
    for (uint32_t i = 0; i < levels; ++i) n += 9 * capacity(g[i]);
    return n + 6 * capacity(g[levels - 1]);
}

__device__ __forceinline__ void bottom_smooth(const vg_grid_desc &g, const Lap &c, float *u,
                                              const float *f, int sweeps) {
    const uint32_t nx = g.np[0], nxy = g.np[0] * g.np[1], n = nxy * g.np[2];
    for (int it = 0; it < sweeps; ++it) {
        for (uint32_t color = 0; color < 2; ++color) {
            for (uint32_t o = threadIdx.x; o < n; o += blockDim.x) {
                const Pt p = split(o, g.np[0], g.np[1]);
                if (on_boundary(p, g) || ((p.x + p.y + p.z) & 1u) != color) continue;
                rbgs_point(u, f, o, nx, nxy, c);
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(kBottomThreads) void k_vcycle_bottom(Bottom b) {
    __shared__ float lds[kBottomLdsFloats];
    float *u[kBottomLevels], *f[kBottomLevels], *r[kBottomLevels];
    float *p = lds;
    for (uint32_t i = 0; i < b.levels; ++i) {
        const uint32_t n3 = 3 * b.g[i].np[0] * b.g[i].np[1] * b.g[i].np[2];
        u[i] = p;
        f[i] = p + n3;
        r[i] = p + 2 * n3;
        p += 3 * n3;
    }
    double *w = (double *)(((uintptr_t)p + 7) & ~(uintptr_t)7);
    const uint32_t n0 = 3 * b.g[0].np[0] * b.g[0].np[1] * b.g[0].np[2];
    for (uint32_t i = threadIdx.x; i < n0; i += blockDim.x) {
        u[0][i] = b.u0[i];
        f[0][i] = b.f0[i];
    }
    __syncthreads();
    const uint32_t top = b.levels - 1;  // the coarsest
    for (uint32_t l = 0; l < top; ++l) {
        const vg_grid_desc &g = b.g[l], &gc = b.g[l + 1];
        const uint32_t nx = g.np[0], nxy = g.np[0] * g.np[1], n = nxy * g.np[2];
        bottom_smooth(g, b.c[l], u[l], f[l], 2);
        for (uint32_t o = threadIdx.x; o < n; o += blockDim.x) {
            if (on_boundary(split(o, g.np[0], g.np[1]), g)) {
                r[l][3 * o] = r[l][3 * o + 1] = r[l][3 * o + 2] = 0.f;
            } else {
                residual_point(u[l], f[l], r[l], o, nx, nxy, b.c[l]);
            }
        }
        __syncthreads();
        const uint32_t nc = gc.np[0] * gc.np[1] * gc.np[2];
        for (uint32_t o = threadIdx.x; o < nc; o += blockDim.x) {
            u[l + 1][3 * o] = u[l + 1][3 * o + 1] = u[l + 1][3 * o + 2] = 0.f;
            const Pt q = split(o, gc.np[0], gc.np[1]);
            float v[3] = {0.f, 0.f, 0.f};
            if (!on_boundary(q, gc)) restrict_point(q, g, r[l], v);
#pragma unroll
            for (int k = 0; k < 3; ++k) f[l + 1][3 * o + k] = v[k];
        }
        __syncthreads();
    }
    {  // coarsest: FP64 red-black on a double copy of u
        const vg_grid_desc &g = b.g[top];
        const uint32_t n = g.np[0] * g.np[1] * g.np[2];
        for (uint32_t i = threadIdx.x; i < 3 * n; i += blockDim.x) w[i] = u[top][i];
        __syncthreads();
        coarse_sweeps_f64(g, w, f[top], b.coarse_sweeps);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < 3 * n; i += blockDim.x) u[top][i] = (float)w[i];
        __syncthreads();
    }
    for (int l = (int)top - 1; l >= 0; --l) {
        const vg_grid_desc &g = b.g[l];
        const uint32_t n = g.np[0] * g.np[1] * g.np[2];
        for (uint32_t o = threadIdx.x; o < n; o += blockDim.x) {
            const Pt q = split(o, g.np[0], g.np[1]);
            if (!on_boundary(q, g)) prolong_point(q, o, u[l], b.g[l + 1], u[l + 1]);
        }
        __syncthreads();
        bottom_smooth(g, b.c[l], u[l], f[l], 2);
    }
    for (uint32_t i = threadIdx.x; i < n0; i += blockDim.x) b.u0[i] = u[0][i];
}

// FP64 residual of the compensated solution (hi + lo): writes the FP32
// residual for the next correction and block partials of ||r||^2, ||f||^2.
__global__ __launch_bounds__(256) void k_residual_pair(vg_grid_desc g, const float *__restrict__ hi,
                                                       const float *__restrict__ lo,
                                                       const float *__restrict__ f,
                                                       float *__restrict__ r32, double r2x,
                                                       double r2y, double r2z,
                                                       double *__restrict__ part) {
    __shared__ double sn[256], sd[256];
    const size_t nx = g.np[0], nxy = nx * g.np[1];
    const size_t total = capacity(g);
    double num = 0.0, den = 0.0;
    for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < total;
         o += (size_t)gridDim.x * blockDim.x) {
        if (on_boundary(split((uint32_t)o, g.np[0], g.np[1]), g)) {
            r32[3 * o] = r32[3 * o + 1] = r32[3 * o + 2] = 0.f;
            continue;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
#define U(off) ((double)hi[3 * (off) + k] + (double)lo[3 * (off) + k])
            const double uc = U(o);
            const double lap = (U(o + 1) + U(o - 1) - 2.0 * uc) * r2x +
                               (U(o + nx) + U(o - nx) - 2.0 * uc) * r2y +
                               (U(o + nxy) + U(o - nxy) - 2.0 * uc) * r2z;
#undef U
            const double fc = f[3 * o + k];
            const double res = fc - lap;
            r32[3 * o + k] = (float)res;
            num += res * res;
            den += fc * fc;
        }
    }
    sn[threadIdx.x] = num;
    sd[threadIdx.x] = den;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
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

// (hi, lo) += e with an error-free TwoSum per component; e is cleared (it
// is the next cycle's zero initial correction).
__global__ void k_twosum_add(float *__restrict__ hi, float *__restrict__ lo,
                             float *__restrict__ e, size_t n) {
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float a = hi[i], b = e[i] + lo[i];
    const float s = a + b;
    const float bb = s - a;
    const float err = (a - (s - bb)) + (b - bb);
    hi[i] = s;
    lo[i] = err;
    e[i] = 0.f;
}

// Workgroups of the FP64 residual: enough to keep every CU's loads in
// flight (512 left it latency-bound at 110 us on a 129x129x257 grid).
constexpr uint32_t kNormBlocks = 2048;

vg_status fp64_residual(vg_multigrid *m, const vg_grid_desc &g, const float *hi, const float *lo,
                        const float *f, float *r32, double *rel, hipStream_t s) {
    const double rx = 1.0 / (double)g.cell[0], ry = 1.0 / (double)g.cell[1],
                 rz = 1.0 / (double)g.cell[2];
    VG_TRY(launch("vg_mg_residual_f64", k_residual_pair, dim3(kNormBlocks), dim3(256), 0, s, g, hi,
                  lo, f, r32, rx * rx, ry * ry, rz * rz, m->part));
    std::vector<double> h(2 * kNormBlocks);
    VG_HIP(hipMemcpyAsync(h.data(), m->part, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    VG_HIP(hipStreamSynchronize(s));
    double num = 0.0, den = 0.0;
    for (uint32_t b = 0; b < kNormBlocks; ++b) {
        num += h[2 * b];
        den += h[2 * b + 1];
    }
    *rel = den > 0.0 ? sqrt(num / den) : sqrt(num);
    return VG_OK;
}

// Workgroups of the fused sweep for tile depth zc.
size_t sweep_blocks(const vg_grid_desc &g, int zc) {
    return (size_t)ceil_div(g.np[0] - 2, rbgs::kSX) * ceil_div(g.np[1] - 2, rbgs::kSY) *
           ceil_div(g.np[2] - 2, zc);
}

// Tile depth of the fused sweep: 16 planes, halved while fewer than 512
// workgroups would result (measured on MI355X: 129x129x257 sweeps in 58 us
// at 16 planes and 512 workgroups, 70 us at 8 planes; 65x65x129 in 14 us at
// 2 planes).  0 = one launch per colour is faster (grids too small to fill
// the chip with tiles: 33x33x65 sweeps in 6.8 us that way, 12 us fused).
int sweep_depth(const vg_grid_desc &g) {
    if (g.np[0] < 33 || sweep_blocks(g, 2) < 256) return 0;
    int zc = 16;
    while (zc > 2 && sweep_blocks(g, zc) < 512) zc /= 2;
    return zc;
}

// `sweeps` red-black sweeps of u.  Even counts on grids large enough run the
// fused out-of-place sweep, ping-ponging through tmp (whose boundary must
// equal u's); otherwise one launch per colour.  Both give the same values
// bit for bit.
vg_status smooth(const vg_grid_desc &g, float *u, float *tmp, const float *f, int sweeps,
                 hipStream_t s) {
    if (g.np[0] < 3 || g.np[1] < 3 || g.np[2] < 3) return VG_OK;
    const Lap c = lap_coef(g);
    const int zc = sweep_depth(g);
    if (tmp && sweeps % 2 == 0 && zc > 0) {
        rbgs::Tile t{(int)g.np[0], (int)g.np[1], (int)g.np[2], 0, 0, 3 * (size_t)g.np[0],
                     3 * (size_t)g.np[0] * g.np[1], c.r2x, c.r2y, c.r2z, c.inv};
        const dim3 grid(ceil_div(g.np[0] - 2, rbgs::kSX), ceil_div(g.np[1] - 2, rbgs::kSY),
                        ceil_div(g.np[2] - 2, zc));
        for (int it = 0; it < sweeps; it += 2) {
            VG_TRY(launch("vg_mg_rbgs_sweep", rbgs::k_rbgs_sweep, grid, dim3(rbgs::kThreads), 0, s,
                          t, (const float *)u, tmp, f, zc));
            VG_TRY(launch("vg_mg_rbgs_sweep", rbgs::k_rbgs_sweep, grid, dim3(rbgs::kThreads), 0, s,
                          t, (const float *)tmp, u, f, zc));
        }
        return VG_OK;
    }
    const dim3 grid(ceil_div((g.np[0] - 1) / 2 + 1, 64), g.np[1] - 2, g.np[2] - 2);
    for (int it = 0; it < sweeps; ++it) {
        VG_TRY(launch("vg_mg_rbgs", k_rbgs, grid, dim3(64), 0, s, g, u, f, 0, c));
        VG_TRY(launch("vg_mg_rbgs", k_rbgs, grid, dim3(64), 0, s, g, u, f, 1, c));
    }
    return VG_OK;
}

// Red-black sweeps of the coarsest level: red-black Gauss-Seidel shrinks
// the slowest error mode by about 1 - (pi / (dmax - 1))^2 per sweep, so
// dmax^2 + 16 sweeps leave below 1e-4 of it for every coarsest size the
// hierarchy allows (dmax <= 33): more than one V-cycle needs.
int coarse_sweeps(const vg_grid_desc &g) {
    uint32_t dmax = g.np[0];
    if (g.np[1] > dmax) dmax = g.np[1];
    if (g.np[2] > dmax) dmax = g.np[2];
    const int sweeps = (int)(dmax * dmax + 16);
    return sweeps > 20000 ? 20000 : sweeps;
}

// One V-cycle on level l (u[l] holds the initial guess).
vg_status vcycle(vg_multigrid *m, uint32_t l, hipStream_t s) {
    const vg_grid_desc &g = m->vg[l];
    const size_t cap = capacity(g);
    if (l == m->bottom) {
        Bottom b;
        b.levels = m->vdepth - l;
        for (uint32_t i = 0; i < b.levels; ++i) {
            b.g[i] = m->vg[l + i];
            b.c[i] = lap_coef(b.g[i]);
        }
        b.u0 = m->u[l];
        b.f0 = m->f[l];
        b.coarse_sweeps = coarse_sweeps(m->vg[m->vdepth - 1]);
        return launch("vg_mg_bottom", k_vcycle_bottom, dim3(1), dim3(kBottomThreads), 0, s, b);
    }
    if (l + 1 == m->vdepth) {
        const bool in_lds = cap <= kCoarseLdsPoints;
        double *work = nullptr;
        if (!in_lds) VG_TRY(salloc((void **)&work, 3 * cap * sizeof(double), s));
        const int sweeps = coarse_sweeps(g);
        vg_status st =
            in_lds ? launch("vg_mg_coarse_f64", k_coarse_f64<true>, dim3(1), dim3(256), 0, s, g,
                            m->u[l], (const float *)m->f[l], sweeps, work)
                   : launch("vg_mg_coarse_f64", k_coarse_f64<false>, dim3(1), dim3(256), 0, s, g,
                            m->u[l], (const float *)m->f[l], sweeps, work);
        if (work) sfree(work, s);
        return st;
    }
    const int nu1 = 2, nu2 = 2;
    VG_TRY(smooth(g, m->u[l], m->r[l], m->f[l], nu1, s));
    VG_TRY(launch("vg_mg_residual", k_residual32, dim3(ceil_div(cap, 256)), dim3(256), 0, s, g,
                  (const float *)m->u[l], (const float *)m->f[l], m->r[l], lap_coef(g)));
    const vg_grid_desc &gc = m->vg[l + 1];
    const size_t capc = capacity(gc);
    VG_TRY(launch("vg_mg_restrict_fw", k_restrict_fw, dim3(ceil_div(capc, 256)), dim3(256), 0, s,
                  gc, m->f[l + 1], m->u[l + 1], g, (const float *)m->r[l]));
    VG_TRY(vcycle(m, l + 1, s));
    VG_TRY(launch("vg_mg_prolong_add", k_prolong_add, dim3(ceil_div(cap, 256)), dim3(256), 0, s, g,
                  m->u[l], gc, (const float *)m->u[l + 1]));
    return smooth(g, m->u[l], m->r[l], m->f[l], nu2, s);
}

}  // namespace
}  // namespace vg

using namespace vg;

VG_API vg_status vg_mg_create(vg_multigrid **m) {
    VG_REQUIRE(m, "out is null");
    *m = new vg_multigrid();
    return VG_OK;
}

VG_API vg_status vg_mg_destroy(vg_multigrid *m) {
    if (!m) return VG_OK;
    for (uint32_t l = 0; l < kMaxLayers; ++l) {
        hipFree(m->A[l]);
        hipFree(m->W[l]);
        hipFree(m->u[l]);
        hipFree(m->f[l]);
        hipFree(m->r[l]);
    }
    hipFree(m->lo);
    hipFree(m->part);
    delete m;
    return VG_OK;
}

VG_API vg_status vg_mg_cascade(vg_multigrid *m, const vg_grid_desc *g, float *a0,
                               const float *negw0, uint32_t steps, vg_stream stream) {
    VG_REQUIRE(m && g && a0 && negw0, "null argument");
    VG_REQUIRE(g->np[0] >= 2 && g->np[1] >= 2 && g->np[2] >= 2, "grid is empty");
    hipStream_t s = (hipStream_t)stream;
    nested_init(&m->ng, *g);
    const Nested &ng = m->ng;
    for (uint32_t l = 1; l < ng.depth; ++l) {
        size_t c1 = m->cap[l], c2 = m->cap[l];
        VG_TRY(grow(&m->A[l], &c1, 3 * capacity(ng.layer[l])));
        VG_TRY(grow(&m->W[l], &c2, 3 * capacity(ng.layer[l])));
        m->cap[l] = c1 < c2 ? c1 : c2;
    }
    float *A[kMaxLayers];
    const float *W[kMaxLayers];
    A[0] = a0;
    W[0] = negw0;
    for (uint32_t l = 1; l < ng.depth; ++l) {
        A[l] = m->A[l];
        W[l] = m->W[l];
    }
    // Cold start: interior of the finest layer is zero (vortonSim.cpp:2531).
    VG_TRY(launch("vg_mg_zero_interior", k_zero_interior, dim3(ceil_div(capacity(*g), 256)),
                  dim3(256), 0, s, *g, a0));
    VG_TRY(sor_solve(ng.layer[0], A[0], W[0], steps, s));
    uint32_t max_valid = 0;
    for (uint32_t l = 1; l < ng.depth; ++l) {
        const uint32_t *np = ng.layer[l].np;
        uint32_t md = np[0];
        if (np[1] < md) md = np[1];
        if (np[2] < md) md = np[2];
        if (md > 2) {
            VG_TRY(inject(ng.layer[l], m->W[l], ng.layer[l - 1], W[l - 1], s));
            VG_TRY(inject(ng.layer[l], A[l], ng.layer[l - 1], A[l - 1], s));
            VG_TRY(sor_solve(ng.layer[l], A[l], W[l], steps, s));
        } else {
            max_valid = l - 1;
            break;
        }
    }
    for (uint32_t l = max_valid; l >= 1; --l) {
        VG_TRY(prolong_interior(ng.layer[l - 1], A[l - 1], ng.layer[l], A[l], s));
        VG_TRY(sor_solve(ng.layer[l - 1], A[l - 1], W[l - 1], steps, s));
    }
    return VG_OK;
}

VG_API vg_status vg_mg_vcycle_solve(vg_multigrid *m, const vg_grid_desc *g, float *a0,
                                    const float *negw0, double tol, uint32_t max_cycles,
                                    double *rel_residual, uint32_t *cycles, vg_stream stream) {
    VG_REQUIRE(m && g && a0 && negw0, "null argument");
    VG_REQUIRE(g->np[0] >= 3 && g->np[1] >= 3 && g->np[2] >= 3, "grid has no interior");
    VG_REQUIRE(g->cell[0] > 0.f && g->cell[1] > 0.f && g->cell[2] > 0.f, "degenerate spacing");
    VG_REQUIRE(capacity(*g) < (size_t(1) << 32), "grid has 2^32 points or more");
    hipStream_t s = (hipStream_t)stream;
    // Hierarchy: halve every axis while all have >= 4 even cells.
    m->vg[0] = *g;
    uint32_t d = 1;
    while (d < kMaxLayers) {
        const vg_grid_desc &p = m->vg[d - 1];
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            const uint32_t c = p.np[a] - 1;
            if (c < 4 || (c & 1u)) ok = false;
        }
        if (!ok) break;
        decimate(&m->vg[d], &p, 2);
        ++d;
    }
    m->vdepth = d;
    // the deepest levels that fit in LDS go to k_vcycle_bottom (one launch
    // per cycle instead of one per step); VG_MG_BOTTOM=0 keeps the
    // per-level launches (the same bits)
    m->bottom = 0xffffffffu;
    {
        const char *env = getenv("VG_MG_BOTTOM");
        if (!(env && env[0] == '0')) {
            for (uint32_t l = 0; l + 1 < d; ++l) {
                const uint32_t levels = d - l;
                if (levels <= (uint32_t)kBottomLevels &&
                    bottom_floats(&m->vg[l], levels) + 1 <= kBottomLdsFloats) {
                    m->bottom = l;
                    break;
                }
            }
        }
    }
    // The coarsest layer is solved by one workgroup; refuse hierarchies that
    // cannot coarsen far enough for that to stay cheap.
    VG_REQUIRE(capacity(m->vg[d - 1]) <= 32768,
               "grid too anisotropic for the V-cycle hierarchy (coarsest layer > 32768 points)");
    const size_t cap0 = capacity(*g);
    for (uint32_t l = 0; l < d; ++l) {
        const size_t c = 3 * capacity(m->vg[l]);
        size_t a = m->vcap[l], b = m->vcap[l], e = m->vcap[l];
        VG_TRY(grow(&m->u[l], &a, c));
        VG_TRY(grow(&m->f[l], &b, c));
        VG_TRY(grow(&m->r[l], &e, c));
        m->vcap[l] = a < b ? (a < e ? a : e) : (b < e ? b : e);
    }
    // r[l] is the smoother's ping-pong buffer: its boundary must be zero, as
    // u[l]'s is (the residual kernel keeps it so once set)
    for (uint32_t l = 0; l + 1 < d; ++l) {
        const size_t c = 3 * capacity(m->vg[l]);
        VG_TRY(launch("vg_mg_zero", k_zero_all, dim3(ceil_div(c, 256)), dim3(256), 0, s, m->r[l], c));
    }
    VG_TRY(grow(&m->lo, &m->lo_cap, 3 * cap0));
    if (!m->part) VG_HIP(hipMalloc((void **)&m->part, 2 * kNormBlocks * sizeof(double)));
    VG_TRY(launch("vg_mg_zero", k_zero_all, dim3(ceil_div(3 * cap0, 256)), dim3(256), 0, s, m->lo,
                  3 * cap0));
    double rel = 0.0;
    uint32_t it = 0;
    // f[0] holds the FP32 residual r = -w - L(hi + lo); u[0] the correction.
    VG_TRY(fp64_residual(m, *g, a0, m->lo, negw0, m->f[0], &rel, s));
    VG_TRY(launch("vg_mg_zero", k_zero_all, dim3(ceil_div(3 * cap0, 256)), dim3(256), 0, s, m->u[0],
                  3 * cap0));
    while (rel > tol && it < max_cycles) {
        // u[0] is zero here: cleared above, then by each two-sum
        VG_TRY(vcycle(m, 0, s));
        VG_TRY(launch("vg_mg_twosum", k_twosum_add, dim3(ceil_div(3 * cap0, 256)), dim3(256), 0, s,
                      a0, m->lo, m->u[0], 3 * cap0));
        ++it;
        VG_TRY(fp64_residual(m, *g, a0, m->lo, negw0, m->f[0], &rel, s));
    }
    if (rel_residual) *rel_residual = rel;
    if (cycles) *cycles = it;
    return VG_OK;
}
