// vg_stretch.hip -- vortex stretching and tilting (VortonSim::
// StretchAndTiltVortons, vortonSim.cpp:2748-2820).
//
//   J      = velocity Jacobian grid, J.a.b = d v_b / d a (ComputeJacobian,
//            uniformGridMath.cpp:304-451): central 1/(2h) inside, one-sided
//            1/h on the faces, 1/h_z = 0 when h_z <= FLT_EPSILON;
//   per vorton: J(p) = trilinear interpolation of the Mat33 grid
//            (uniformGrid.h:1117-1148, the Vec3 formula on each of the nine
//            components), s_a = J.a . Omega (the transpose formulation,
//            vortonSim.cpp:2800-2801 with mat33.h:131-145), and
//            Omega += 0.5 s dt.
//
// Vortons may be processed in any order (`order`, a permutation of
// 0..n-1, or natural order): each update is independent, so the values do
// not depend on it, but the cell order of the influence tree turns the
// grid gathers of neighbouring threads into cache hits.
//
// Two routes give the same values bit for bit:
//   k_jacobian + k_stretch_tilt: the reference's two steps, a 36 B/point
//     Jacobian grid then a 72-value gather per vorton;
//   k_stretch_tilt_fused: the Jacobian of the 8 corners of the vorton's cell
//     computed on the fly from the velocity grid (cache-resident), so the
//     Jacobian grid is never written or read.
// Every expression keeps the reference's operation order with contraction
// off.
#include "vg_internal.h"

namespace vg {
namespace {

struct JacScale {
    float r[3];   // 1 / h (0 for a flat z axis)
    float hr[3];  // 0.5 / h
};

JacScale jac_scale(const vg_grid_desc &g) {
    JacScale s;
    s.r[0] = 1.0f / g.cell[0];
    s.r[1] = 1.0f / g.cell[1];
    s.r[2] = g.cell[2] > FLT_EPSILON ? 1.0f / g.cell[2] : 0.0f;
    for (int a = 0; a < 3; ++a) s.hr[a] = 0.5f * s.r[a];
    return s;
}

// j[3 a + b] = d v_b / d a at grid point (ix, iy, iz).
__device__ __forceinline__ void jacobian_at(const vg_grid_desc &g, const float *__restrict__ v,
                                            const JacScale &sc, uint32_t ix, uint32_t iy,
                                            uint32_t iz, float j[9]) {
#pragma clang fp contract(off)
    const size_t o = point_offset(g, ix, iy, iz);
    const uint32_t idx[3] = {ix, iy, iz};
    const size_t st[3] = {1, g.np[0], (size_t)g.np[0] * g.np[1]};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        size_t lo = o - st[a], hi = o + st[a];
        float scale = sc.hr[a];
        if (idx[a] == 0) {
            lo = o;
            scale = sc.r[a];
        } else if (idx[a] == g.np[a] - 1) {
            hi = o;
            scale = sc.r[a];
        }
#pragma unroll
        for (int b = 0; b < 3; ++b) j[3 * a + b] = (v[3 * hi + b] - v[3 * lo + b]) * scale;
    }
}

__global__ void k_jacobian(vg_grid_desc g, const float *__restrict__ v, float *__restrict__ jac,
                           JacScale sc) {
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= capacity(g)) return;
    const uint32_t nx = g.np[0], ny = g.np[1];
    const uint32_t row = (uint32_t)o / nx;
    float j[9];
    jacobian_at(g, v, sc, (uint32_t)o - row * nx, row % ny, row / ny, j);
#pragma unroll
    for (int k = 0; k < 9; ++k) jac[9 * o + k] = j[k];
}

// Omega += 0.5 (Omega * J) dt with J interpolated (9 components).
__device__ __forceinline__ void stretch_update(const float j[9], float *w, float dt) {
#pragma clang fp contract(off)
    const float w0 = w[0], w1 = w[1], w2 = w[2];
    float s[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) s[a] = j[3 * a] * w0 + j[3 * a + 1] * w1 + j[3 * a + 2] * w2;
#pragma unroll
    for (int a = 0; a < 3; ++a) w[a] = w[a] + s[a] * 0.5f * dt;
}

__global__ void k_stretch_tilt(vg_grid_desc g, const float *__restrict__ jac,
                               const float *__restrict__ pos, size_t pos_stride,
                               float *__restrict__ angvel, size_t ang_stride, size_t n, float dt,
                               const uint32_t *__restrict__ order) {
#pragma clang fp contract(off)
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const size_t i = order ? order[t] : t;
    const float *p = pos + i * pos_stride;
    const CellWeights w = cell_weights(g, make_float3(p[0], p[1], p[2]));
    const size_t nx = g.np[0], nxy = (size_t)g.np[0] * g.np[1];
    const size_t o000 = (size_t)w.ix + nx * ((size_t)w.iy + (size_t)g.np[1] * w.iz);
    const float *a000 = jac + 9 * o000, *a100 = a000 + 9, *a010 = a000 + 9 * nx,
                *a110 = a010 + 9, *a001 = a000 + 9 * nxy, *a101 = a001 + 9,
                *a011 = a001 + 9 * nx, *a111 = a011 + 9;
    float j[9];
#pragma unroll
    for (int k = 0; k < 9; ++k)
        j[k] = ((w.ox * a000[k] + w.tx * a100[k]) * w.oy + (w.ox * a010[k] + w.tx * a110[k]) * w.ty) *
                   w.oz +
               ((w.ox * a001[k] + w.tx * a101[k]) * w.oy + (w.ox * a011[k] + w.tx * a111[k]) * w.ty) *
                   w.tz;
    stretch_update(j, angvel + i * ang_stride, dt);
}

// x-lerp of the corner Jacobians at (ix, y, z) and (ix+1, y, z).
__device__ __forceinline__ void lerp_x(const vg_grid_desc &g, const float *__restrict__ v,
                                       const JacScale &sc, const CellWeights &w, uint32_t y,
                                       uint32_t z, float out[9]) {
#pragma clang fp contract(off)
    float j0[9], j1[9];
    jacobian_at(g, v, sc, (uint32_t)w.ix, y, z, j0);
    jacobian_at(g, v, sc, (uint32_t)w.ix + 1, y, z, j1);
#pragma unroll
    for (int k = 0; k < 9; ++k) out[k] = w.ox * j0[k] + w.tx * j1[k];
}

__global__ void k_stretch_tilt_fused(vg_grid_desc g, const float *__restrict__ vel, JacScale sc,
                                     const float *__restrict__ pos, size_t pos_stride,
                                     float *__restrict__ angvel, size_t ang_stride, size_t n,
                                     float dt, const uint32_t *__restrict__ order) {
#pragma clang fp contract(off)
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const size_t i = order ? order[t] : t;
    const float *p = pos + i * pos_stride;
    const CellWeights w = cell_weights(g, make_float3(p[0], p[1], p[2]));
    const uint32_t y0 = (uint32_t)w.iy, z0 = (uint32_t)w.iz;
    float a[9], b[9], lo[9], j[9];
    lerp_x(g, vel, sc, w, y0, z0, a);
    lerp_x(g, vel, sc, w, y0 + 1, z0, b);
#pragma unroll
    for (int k = 0; k < 9; ++k) lo[k] = a[k] * w.oy + b[k] * w.ty;
    lerp_x(g, vel, sc, w, y0, z0 + 1, a);
    lerp_x(g, vel, sc, w, y0 + 1, z0 + 1, b);
#pragma unroll
    for (int k = 0; k < 9; ++k) j[k] = lo[k] * w.oz + (a[k] * w.oy + b[k] * w.ty) * w.tz;
    stretch_update(j, angvel + i * ang_stride, dt);
}

vg_status check_grid(const vg_grid_desc *g) {
    VG_REQUIRE(g, "grid is null");
    VG_REQUIRE(g->np[0] >= 2 && g->np[1] >= 2 && g->np[2] >= 2,
               "grid has fewer than 2 points per axis");
    return VG_OK;
}

bool flat(const vg_grid_desc &g) {
    return g.extent[0] == 0.f || g.extent[1] == 0.f || g.extent[2] == 0.f;
}

}  // namespace

vg_status stretch_tilt_fused(const vg_grid_desc &g, const float *vel, const float *pos,
                             size_t pos_stride, float *angvel, size_t ang_stride, size_t n,
                             float dt, const uint32_t *order, hipStream_t s) {
    if (n == 0 || flat(g)) return VG_OK;
    return launch("vg_stretch_tilt_fused", k_stretch_tilt_fused, dim3(ceil_div(n, 256)),
                  dim3(256), 0, s, g, vel, jac_scale(g), pos, pos_stride, angvel, ang_stride, n,
                  dt, order);
}

}  // namespace vg

using namespace vg;

VG_API vg_status vg_jacobian(const vg_grid_desc *g, const float *vel, float *jac, vg_stream s) {
    VG_TRY(check_grid(g));
    VG_REQUIRE(vel && jac, "null field");
    return launch("vg_jacobian", k_jacobian, dim3(ceil_div(capacity(*g), 256)), dim3(256), 0,
                  (hipStream_t)s, *g, vel, jac, jac_scale(*g));
}

VG_API vg_status vg_stretch_tilt(const vg_grid_desc *g, const float *jac, const float *pos,
                                 size_t pos_stride, float *angvel, size_t ang_stride, size_t n,
                                 float dt, const uint32_t *order, vg_stream s) {
    VG_TRY(check_grid(g));
    if (n == 0 || flat(*g)) return VG_OK;
    VG_REQUIRE(jac && pos && angvel && pos_stride >= 3 && ang_stride >= 3, "bad arguments");
    return launch("vg_stretch_tilt", k_stretch_tilt, dim3(ceil_div(n, 256)), dim3(256), 0,
                  (hipStream_t)s, *g, jac, pos, pos_stride, angvel, ang_stride, n, dt, order);
}

VG_API vg_status vg_stretch_tilt_from_velocity(const vg_grid_desc *g, const float *vel,
                                               const float *pos, size_t pos_stride, float *angvel,
                                               size_t ang_stride, size_t n, float dt,
                                               const uint32_t *order, vg_stream s) {
    VG_TRY(check_grid(g));
    if (n == 0 || flat(*g)) return VG_OK;
    VG_REQUIRE(vel && pos && angvel && pos_stride >= 3 && ang_stride >= 3, "bad arguments");
    return stretch_tilt_fused(*g, vel, pos, pos_stride, angvel, ang_stride, n, dt, order,
                              (hipStream_t)s);
}
