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
    uint32_t cascade_layers = 0;  // layers holding values after the last cascade
    // V-cycle hierarchy
    uint32_t vdepth = 0;
    vg_grid_desc vg[vg::kMaxLayers];
    float *u[vg::kMaxLayers] = {};
    float *f[vg::kMaxLayers] = {};
    float *r[vg::kMaxLayers] = {};
    size_t vcap[vg::kMaxLayers] = {};
    float *lo = nullptr;       // compensation term of the finest solution
    size_t lo_cap = 0;
    size_t lo_valid = 0;       // points of the last solve's finest grid
    double *part = nullptr;    // FP64 norm partials, one pair per workgroup, then the sums
    size_t part_cap = 0;
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

// Red-black point Gauss-Seidel: colour = (ix + iy + iz) & 1.  Threads map to
// the colour's points of one x row: ix = 1 + ((iy + iz + color + 1) & 1) + 2k.
__global__ void k_rbgs(vg_grid_desc g, float *__restrict__ u, const float *__restrict__ f,
                       int color, Lap c) {
#pragma clang fp contract(off)
    const uint32_t iy = 1 + blockIdx.y, iz = 1 + blockIdx.z;
    const uint32_t first = 1 + ((iy + iz + 1 + (uint32_t)color) & 1u);
    const uint32_t ix = first + 2 * (blockIdx.x * blockDim.x + threadIdx.x);
    const size_t nx = g.np[0], nxy = (size_t)g.np[0] * g.np[1];
    if (ix + 1 >= nx) return;
    const size_t o = ix + (size_t)iy * nx + (size_t)iz * nxy;
    const float inv = c.inv;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float s = (u[3 * (o + 1) + k] + u[3 * (o - 1) + k]) * c.r2x +
                        (u[3 * (o + nx) + k] + u[3 * (o - nx) + k]) * c.r2y +
                        (u[3 * (o + nxy) + k] + u[3 * (o - nxy) + k]) * c.r2z;
        u[3 * o + k] = (s - f[3 * o + k]) * inv;
    }
}

// r = f - L u on interior points, 0 on the boundary.
__global__ void k_residual32(vg_grid_desc g, const float *__restrict__ u,
                             const float *__restrict__ f, float *__restrict__ r, Lap c) {
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= capacity(g)) return;
    const size_t nx = g.np[0], nxy = nx * g.np[1];
    if (on_boundary(split((uint32_t)o, g.np[0], g.np[1]), g)) {
        r[3 * o] = r[3 * o + 1] = r[3 * o + 2] = 0.f;
        return;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float uc = u[3 * o + k];
        const float lap = (u[3 * (o + 1) + k] + u[3 * (o - 1) + k]) * c.r2x +
                          (u[3 * (o + nx) + k] + u[3 * (o - nx) + k]) * c.r2y +
                          (u[3 * (o + nxy) + k] + u[3 * (o - nxy) + k]) * c.r2z - c.diag * uc;
        r[3 * o + k] = f[3 * o + k] - lap;
    }
}

// Full weighting (weights 8/4/2/1 by Manhattan distance, /64) onto coarse
// interior points; coarse boundary = 0 (homogeneous Dirichlet correction).
// Also clears the coarse correction ulo (the coarse level's initial guess).
__global__ void k_restrict_fw(vg_grid_desc lo, float *__restrict__ flo, float *__restrict__ ulo,
                              vg_grid_desc hi, const float *__restrict__ rhi) {
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= capacity(lo)) return;
    ulo[3 * o] = ulo[3 * o + 1] = ulo[3 * o + 2] = 0.f;
    const Pt p = split((uint32_t)o, lo.np[0], lo.np[1]);
    const size_t ix = p.x, iy = p.y, iz = p.z;
    if (on_boundary(p, lo)) {
        flo[3 * o] = flo[3 * o + 1] = flo[3 * o + 2] = 0.f;
        return;
    }
    const size_t hx = hi.np[0], hxy = (size_t)hi.np[0] * hi.np[1];
    const size_t c = 2 * ix + 2 * iy * hx + 2 * iz * hxy;
    float acc[3] = {0.f, 0.f, 0.f};
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int md = abs(dx) + abs(dy) + abs(dz);
                const float w = (float)(8 >> md);
                const size_t h = c + dx + (long long)dy * (long long)hx + (long long)dz * (long long)hxy;
#pragma unroll
                for (int k = 0; k < 3; ++k) acc[k] += w * rhi[3 * h + k];
            }
#pragma unroll
    for (int k = 0; k < 3; ++k) flo[3 * o + k] = acc[k] * (1.0f / 64.0f);
}

// u_hi(interior) += trilinear(u_lo).
__global__ void k_prolong_add(vg_grid_desc hi, float *__restrict__ uhi, vg_grid_desc lo,
                              const float *__restrict__ ulo) {
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= capacity(hi)) return;
    const Pt p = split((uint32_t)o, hi.np[0], hi.np[1]);
    if (on_boundary(p, hi)) return;
    const size_t ix = p.x, iy = p.y, iz = p.z;
    const size_t lx = lo.np[0], lxy = (size_t)lo.np[0] * lo.np[1];
    const size_t x0 = ix / 2, y0 = iy / 2, z0 = iz / 2;
    const size_t x1 = x0 + (ix & 1), y1 = y0 + (iy & 1), z1 = z0 + (iz & 1);
    const float wx = (ix & 1) ? 0.5f : 1.0f, wy = (iy & 1) ? 0.5f : 1.0f, wz = (iz & 1) ? 0.5f : 1.0f;
    float acc[3] = {0.f, 0.f, 0.f};
    const size_t xs[2] = {x0, x1}, ys[2] = {y0, y1}, zs[2] = {z0, z1};
    const int nxs = (ix & 1) ? 2 : 1, nys = (iy & 1) ? 2 : 1, nzs = (iz & 1) ? 2 : 1;
    for (int a = 0; a < nzs; ++a)
        for (int b = 0; b < nys; ++b)
            for (int d = 0; d < nxs; ++d) {
                const size_t l = xs[d] + ys[b] * lx + zs[a] * lxy;
                const float w = wx * wy * wz;
#pragma unroll
                for (int k = 0; k < 3; ++k) acc[k] += w * ulo[3 * l + k];
            }
#pragma unroll
    for (int k = 0; k < 3; ++k) uhi[3 * o + k] += acc[k];
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
    const uint32_t nx = g.np[0], nxy = nx * g.np[1], n = nxy * g.np[2];
    const double rx = 1.0 / g.cell[0], ry = 1.0 / g.cell[1], rz = 1.0 / g.cell[2];
    const double r2x = rx * rx, r2y = ry * ry, r2z = rz * rz;
    const double inv = 1.0 / (2.0 * (r2x + r2y + r2z));
    for (uint32_t i = threadIdx.x; i < 3 * n; i += blockDim.x) w[i] = u[i];
    __syncthreads();
    auto update = [&](uint32_t o) {
        for (int k = 0; k < 3; ++k) {
            const double s = (w[3 * (o + 1) + k] + w[3 * (o - 1) + k]) * r2x +
                             (w[3 * (o + nx) + k] + w[3 * (o - nx) + k]) * r2y +
                             (w[3 * (o + nxy) + k] + w[3 * (o - nxy) + k]) * r2z;
            w[3 * o + k] = (s - (double)f[3 * o + k]) * inv;
        }
    };
    if (n <= 64) {
        // one point per lane of the first wave, colours ordered by wave
        // barriers (a wave's LDS operations complete in order) instead of
        // workgroup barriers
        if (threadIdx.x < 64) {
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
                    if (inner && parity == color) update(o);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            }
        }
        __syncthreads();
    } else {
        for (int it = 0; it < sweeps; ++it) {
            for (uint32_t color = 0; color < 2; ++color) {
                for (uint32_t o = threadIdx.x; o < n; o += blockDim.x) {
                    const Pt p = split(o, g.np[0], g.np[1]);
                    if (on_boundary(p, g) || ((p.x + p.y + p.z) & 1u) != color) continue;
                    update(o);
                }
                __syncthreads();
            }
        }
    }
    for (uint32_t i = threadIdx.x; i < 3 * n; i += blockDim.x) u[i] = (float)w[i];
}

// FP64 residual of the compensated solution (hi + lo): writes the FP32
// residual for the next correction and per-workgroup partials of ||r||^2,
// ||f||^2.  One point per thread: a grid-stride loop left each wave walking
// eight points one after another, each waiting on its fourteen stencil loads.
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
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o < total) {
        if (on_boundary(split((uint32_t)o, g.np[0], g.np[1]), g)) {
            r32[3 * o] = r32[3 * o + 1] = r32[3 * o + 2] = 0.f;
            goto reduce;
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
reduce:
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

// The workgroups' partial pairs summed in a fixed order by one workgroup of
// 1024 threads: thread t takes pairs t + 1024 j into accumulator j % 4 (four
// loads in flight), adds the four in order, then a fixed tree; the sums go
// to part[2 * nb], part[2 * nb + 1].
constexpr int kFinishThreads = 1024;
__global__ __launch_bounds__(kFinishThreads) void k_norm_finish(double *__restrict__ part,
                                                                uint32_t nb) {
    __shared__ double sn[kFinishThreads], sd[kFinishThreads];
    double n4[4] = {0.0, 0.0, 0.0, 0.0}, d4[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t b = threadIdx.x; b < nb; b += 4 * kFinishThreads) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t i = b + (uint32_t)j * kFinishThreads;
            if (i < nb) {
                n4[j] += part[2 * i];
                d4[j] += part[2 * i + 1];
            }
        }
    }
    sn[threadIdx.x] = (n4[0] + n4[1]) + (n4[2] + n4[3]);
    sd[threadIdx.x] = (d4[0] + d4[1]) + (d4[2] + d4[3]);
    __syncthreads();
    for (int w = kFinishThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sn[threadIdx.x] += sn[threadIdx.x + w];
            sd[threadIdx.x] += sd[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * nb] = sn[0];
        part[2 * nb + 1] = sd[0];
    }
}

// Pairs of doubles m->part must hold for grid g: one per workgroup of the
// residual, plus the sums.
size_t norm_part_pairs(const vg_grid_desc &g) { return ceil_div(capacity(g), 256) + 1; }

vg_status fp64_residual(vg_multigrid *m, const vg_grid_desc &g, const float *hi, const float *lo,
                        const float *f, float *r32, double *rel, hipStream_t s) {
    const double rx = 1.0 / (double)g.cell[0], ry = 1.0 / (double)g.cell[1],
                 rz = 1.0 / (double)g.cell[2];
    const uint32_t nb = ceil_div(capacity(g), 256);
    VG_REQUIRE(m->part && m->part_cap >= 2 * norm_part_pairs(g), "norm partials not allocated");
    VG_TRY(launch("vg_mg_residual_f64", k_residual_pair, dim3(nb), dim3(256), 0, s, g, hi, lo, f,
                  r32, rx * rx, ry * ry, rz * rz, m->part));
    VG_TRY(launch("vg_mg_residual_f64", k_norm_finish, dim3(1), dim3(kFinishThreads), 0, s, m->part,
                  nb));
    double h[2] = {0.0, 0.0};
    VG_HIP(hipMemcpyAsync(h, m->part + 2 * (size_t)nb, sizeof h, hipMemcpyDeviceToHost, s));
    VG_HIP(hipStreamSynchronize(s));
    *rel = h[1] > 0.0 ? sqrt(h[0] / h[1]) : sqrt(h[0]);
    return VG_OK;
}

// Workgroups of the fused sweep for tile depth zc.
size_t sweep_blocks(const vg_grid_desc &g, int zc) {
    return (size_t)ceil_div(g.np[0] - 2, rbgs::kSX) * ceil_div(g.np[1] - 2, rbgs::kSY) *
           ceil_div(g.np[2] - 2, zc);
}

// Workgroups of the fused sweep resident at once: three per CU (48,960 B of
// LDS each) on the 256 CUs of an MI355X.
constexpr size_t kSweepResident = 3 * 256;

// Tile depth of the fused sweep.  A workgroup marches its planes for the
// whole launch, so the deepest tile that still gives every resident slot a
// workgroup is best (129x129x257: 11 planes, 768 workgroups, 49 us; 16
// planes and 512 workgroups 54 us), down to 8 planes (a tile stages four
// halo planes besides its own).  Grids too small for that: 16 planes,
// halved while fewer than 512 workgroups would result (65x65x129 in 15 us at
// 2 planes).  0 = one launch per colour is faster (grids too small to fill
// the chip with tiles: 33x33x65 sweeps in 6.8 us that way, 12 us fused).
int sweep_depth(const vg_grid_desc &g) {
    if (g.np[0] < 33 || sweep_blocks(g, 2) < 256) return 0;
    for (int zc = 16; zc >= 8; --zc)
        if (sweep_blocks(g, zc) >= kSweepResident) return zc;
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

// One V-cycle on level l (u[l] holds the initial guess).
vg_status vcycle(vg_multigrid *m, uint32_t l, hipStream_t s) {
    const vg_grid_desc &g = m->vg[l];
    const size_t cap = capacity(g);
    if (l + 1 == m->vdepth) {
        const bool in_lds = cap <= kCoarseLdsPoints;
        double *work = nullptr;
        if (!in_lds) VG_TRY(salloc((void **)&work, 3 * cap * sizeof(double), s));
        uint32_t dmax = g.np[0];
        if (g.np[1] > dmax) dmax = g.np[1];
        if (g.np[2] > dmax) dmax = g.np[2];
        // red-black Gauss-Seidel shrinks the slowest error mode by about
        // 1 - (pi / (dmax - 1))^2 per sweep, so dmax^2 + 16 sweeps leave
        // below 1e-4 of it for every coarsest size the hierarchy allows
        // (dmax <= 33): more than one V-cycle needs
        int sweeps = (int)(dmax * dmax + 16);
        if (sweeps > 20000) sweeps = 20000;
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
        (void)hipFree(m->A[l]);
        (void)hipFree(m->W[l]);
        (void)hipFree(m->u[l]);
        (void)hipFree(m->f[l]);
        (void)hipFree(m->r[l]);
    }
    (void)hipFree(m->lo);
    (void)hipFree(m->part);
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
    m->cascade_layers = max_valid + 1;
    return VG_OK;
}

namespace vg {
uint32_t mg_cascade_depth(const vg_multigrid *m) { return m ? m->cascade_layers : 0; }
}  // namespace vg

VG_API vg_status vg_mg_layer(const vg_multigrid *m, int field, uint32_t layer, vg_grid_desc *g,
                             const float **data, uint32_t *depth) {
    VG_REQUIRE(m && g && data, "null argument");
    VG_REQUIRE(field == VG_MG_VECTOR_POTENTIAL || field == VG_MG_NEGATIVE_VORTICITY,
               "field must be VG_MG_VECTOR_POTENTIAL or VG_MG_NEGATIVE_VORTICITY");
    if (depth) *depth = m->cascade_layers;
    VG_REQUIRE(layer >= 1 && layer < m->cascade_layers,
               "layer must be a coarse layer the last cascade filled (1 <= layer < depth)");
    *g = m->ng.layer[layer];
    *data = field == VG_MG_VECTOR_POTENTIAL ? m->A[layer] : m->W[layer];
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
    VG_TRY(grow(&m->part, &m->part_cap, 2 * norm_part_pairs(*g)));
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
    m->lo_valid = cap0;
    return VG_OK;
}

VG_API vg_status vg_mg_vcycle_compensation(const vg_multigrid *m, const float **lo,
                                           size_t *count) {
    VG_REQUIRE(m && lo, "null argument");
    VG_REQUIRE(m->lo && m->lo_valid, "no V-cycle solve has run on this object");
    *lo = m->lo;
    if (count) *count = 3 * m->lo_valid;
    return VG_OK;
}
