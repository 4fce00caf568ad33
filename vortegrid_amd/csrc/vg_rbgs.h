// vg_rbgs.h -- one full red-black Gauss-Seidel sweep of the 7-point vector
// Laplacian, out of place and LDS-tiled (2.5-D).
//
//   colour 0 = points with (x + y + z) even, updated first from ui;
//   colour 1 = the others, updated from the new colour-0 values;
//   uo = the swept field on interior points (uo's boundary must already
//   hold ui's boundary: the sweep writes interior points only).
//
// A workgroup of 256 threads owns kSX x kSY interior points of x and y and
// marches over zc planes of z.  It stages ui over the tile plus a two-point
// halo in x and y into a ring of kSlots planes in LDS, computes the new
// colour-0 values in place over the tile plus a one-point halo (they need
// only colour-1 values of ui, which nothing in the ring overwrites), then
// the new colour-1 values of the tile from them.  Halo colour-0 values are
// recomputed by each workgroup that needs them, so no workgroup reads
// another's output, and one launch does what two colour launches do with
// half the traffic: ui, f read once and uo written once per point.
//
// Global loads run two planes ahead in registers (the staged plane and the f
// values both passes of a plane need), in two sets that alternate between
// planes: a workgroup's march is a chain of dependent plane steps, and with
// two workgroups per CU a one-plane lead left each step waiting on memory.
//
// The phases are __host__ __device__ functions of the thread index so the
// same code runs in a serial host model of the workgroup.
#pragma once

#include <hip/hip_runtime.h>

namespace vg {
namespace rbgs {

constexpr int kSX = 64, kSY = 8;         // interior points per tile in x, y
constexpr int kSW = kSX + 4;             // staged points per row (2-point halo)
constexpr int kSH = kSY + 4;             // staged rows
constexpr int kSlotF = 3 * kSW * kSH;    // floats per staged plane
constexpr int kSlots = 5;                // planes z-1 .. z+2 in use, one being staged
constexpr int kThreads = 256;
constexpr int kStageN = (kSlotF + kThreads - 1) / kThreads;  // staged floats per thread
constexpr int kRed = (kSX + 2) / 2;      // colour-0 points per row of tile + 1 halo
constexpr int kRedN = (kRed * (kSY + 2) + kThreads - 1) / kThreads;  // colour-0 points per thread

struct Tile {
    int nx, ny, nz;
    int x0, y0;           // first interior point of the tile
    size_t sy, sz;        // float strides of a row and a plane
    float r2x, r2y, r2z, inv;
};

__host__ __device__ inline int slot(int q) { return (q + kSlots) % kSlots; }

// The staged region is rows y0-2 .. y0+kSY+1, points x0-2 .. x0+kSX+1, 3
// components; thread tid stages elements tid + j kThreads.  off[j] is the
// element's offset within a plane of ui, or -1 outside the grid (or past
// the region's end).
__host__ __device__ inline void stage_map(const Tile &t, int tid, int (&off)[kStageN]) {
    for (int j = 0; j < kStageN; ++j) {
        const int e = tid + j * kThreads;
        const int row = e / (3 * kSW), col = e - row * (3 * kSW);
        const int x = t.x0 - 2 + col / 3, y = t.y0 - 2 + row;
        const bool in = e < kSlotF && x >= 0 && x < t.nx && y >= 0 && y < t.ny;
        off[j] = in ? (int)((size_t)y * t.sy + 3 * (size_t)x + col % 3) : -1;
    }
}

// Plane q's staged values into registers; zero outside the grid.
__host__ __device__ inline void stage_load(const Tile &t, const float *__restrict__ ui,
                                           const int (&off)[kStageN], int q,
                                           float (&v)[kStageN]) {
    const bool inq = q >= 0 && q < t.nz;
    const float *plane = ui + (size_t)(inq ? q : 0) * t.sz;
    for (int j = 0; j < kStageN; ++j) v[j] = (inq && off[j] >= 0) ? plane[off[j]] : 0.f;
}

__host__ __device__ inline void stage_store(const float (&v)[kStageN], float *dst, int tid) {
    for (int j = 0; j < kStageN; ++j) {
        const int e = tid + j * kThreads;
        if (e < kSlotF) dst[e] = v[j];
    }
}

// Colour-0 point k of thread tid on plane q (tile plus one-point halo):
// false when it is not an interior point.
__host__ __device__ inline bool red_point(const Tile &t, int q, int tid, int k, int &x, int &y) {
    const int e = tid + k * kThreads;
    if (e >= kRed * (kSY + 2) || q < 1 || q > t.nz - 2) return false;
    const int row = e / kRed, i = e - row * kRed;
    const int xs = t.x0 - 1;
    y = t.y0 - 1 + row;
    x = xs + ((xs + y + q) & 1) + 2 * i;
    return x >= 1 && x <= t.nx - 2 && y >= 1 && y <= t.ny - 2;
}

__host__ __device__ inline void red_load(const Tile &t, const float *__restrict__ f, int q,
                                         int tid, float (&fr)[3 * kRedN]) {
    for (int k = 0; k < kRedN; ++k) {
        int x = 0, y = 0;
        const bool ok = red_point(t, q, tid, k, x, y);
        const size_t o = ok ? (size_t)q * t.sz + (size_t)y * t.sy + 3 * (size_t)x : 0;
        for (int c = 0; c < 3; ++c) fr[3 * k + c] = ok ? f[o + c] : 0.f;
    }
}

// New colour-0 values of plane q over the tile plus a one-point halo,
// written in place into pc; boundary and outside points keep their staged
// values.  fr = red_load of the same plane.
__host__ __device__ inline void red(const Tile &t, const float (&fr)[3 * kRedN], const float *pm,
                                    float *pc, const float *pp, int q, int tid) {
#pragma clang fp contract(off)
    for (int k = 0; k < kRedN; ++k) {
        int x = 0, y = 0;
        if (!red_point(t, q, tid, k, x, y)) continue;
        const int i = 3 * ((y - t.y0 + 2) * kSW + (x - t.x0 + 2));
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float s = (pc[i + 3 + c] + pc[i - 3 + c]) * t.r2x +
                            (pc[i + 3 * kSW + c] + pc[i - 3 * kSW + c]) * t.r2y +
                            (pp[i + c] + pm[i + c]) * t.r2z;
            pc[i + c] = (s - fr[3 * k + c]) * t.inv;
        }
    }
}

// Thread tid owns the adjacent points x0 + 2p and x0 + 2p + 1 of row
// y0 + tid / 32 of the tile (p = tid % 32); the colour-1 one is
// x0 + 2p + black_h.
__host__ __device__ inline int black_h(const Tile &t, int z, int tid) {
    const int y = t.y0 + tid / 32, x = t.x0 + 2 * (tid % 32);
    return ((x + y + z) & 1) ? 0 : 1;
}

__host__ __device__ inline void black_load(const Tile &t, const float *__restrict__ f, int z,
                                           int tid, float (&fb)[3]) {
    const int y = t.y0 + tid / 32, x = t.x0 + 2 * (tid % 32) + black_h(t, z, tid);
    const bool ok = z >= 1 && z <= t.nz - 2 && y <= t.ny - 2 && x <= t.nx - 2;
    const size_t o = ok ? (size_t)z * t.sz + (size_t)y * t.sy + 3 * (size_t)x : 0;
    for (int c = 0; c < 3; ++c) fb[c] = ok ? f[o + c] : 0.f;
}

// Plane z of the tile into uo: colour-0 points copy their new value,
// colour-1 points are updated from the new colour-0 neighbours.
// fb = black_load of the same plane.
__host__ __device__ inline void black(const Tile &t, const float (&fb)[3], float *__restrict__ uo,
                                      const float *pm, const float *pc, const float *pp, int z,
                                      int tid) {
#pragma clang fp contract(off)
    const int y = t.y0 + tid / 32, p = tid % 32;
    if (y > t.ny - 2) return;
    const int hb = black_h(t, z, tid);
    for (int h = 0; h < 2; ++h) {
        const int x = t.x0 + 2 * p + h;
        if (x > t.nx - 2) continue;
        const int i = 3 * ((y - t.y0 + 2) * kSW + (x - t.x0 + 2));
        const size_t o = (size_t)z * t.sz + (size_t)y * t.sy + 3 * (size_t)x;
        if (h != hb) {
#pragma unroll
            for (int c = 0; c < 3; ++c) uo[o + c] = pc[i + c];
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float s = (pc[i + 3 + c] + pc[i - 3 + c]) * t.r2x +
                                (pc[i + 3 * kSW + c] + pc[i - 3 * kSW + c]) * t.r2y +
                                (pp[i + c] + pm[i + c]) * t.r2z;
                uo[o + c] = (s - fb[c]) * t.inv;
            }
        }
    }
}

static_assert(kThreads == 32 * kSY && kSX == 64, "black() maps 32 pairs x kSY rows to threads");

// The tile origin of workgroup (bx, by).
__host__ __device__ inline Tile at(Tile t, int bx, int by) {
    t.x0 = 1 + bx * kSX;
    t.y0 = 1 + by * kSY;
    return t;
}

// One plane step of the march: store the prefetched plane z+2 into its
// ring slot (that of z-3, which no pass still in flight reads: the colour-1
// pass of z-1 reads z-2 .. z), barrier, colour 0 of plane z+1 (reads
// z .. z+2), barrier, colour 1 of plane z (reads z-1 .. z+1).  The register
// set it uses is refilled for the step two planes on: plane z+4 to stage,
// f of plane z+3 for its colour-0 pass and of plane z+2 for its colour-1
// pass, so every load has two steps to arrive.
__device__ __forceinline__ void march_step(const Tile &t, const float *__restrict__ ui,
                                           float *__restrict__ uo, const float *__restrict__ f,
                                           float (*ring)[kSlotF], const int (&off)[kStageN], int z,
                                           int tid, float (&sv)[kStageN], float (&fr)[3 * kRedN],
                                           float (&fb)[3]) {
    stage_store(sv, ring[slot(z + 2)], tid);
    stage_load(t, ui, off, z + 4, sv);
    __syncthreads();
    red(t, fr, ring[slot(z)], ring[slot(z + 1)], ring[slot(z + 2)], z + 1, tid);
    __syncthreads();
    black(t, fb, uo, ring[slot(z - 1)], ring[slot(z)], ring[slot(z + 1)], z, tid);
    red_load(t, f, z + 3, tid, fr);
    black_load(t, f, z + 2, tid, fb);
}

// The march of one workgroup over planes z0 .. z1-1 (z1 exclusive); ring
// slot of plane q: slot(q).  Two register sets alternate between the steps
// (the loop is unrolled by two so that each step names its set).
__global__ __launch_bounds__(kThreads) void k_rbgs_sweep(Tile base, const float *__restrict__ ui,
                                                         float *__restrict__ uo,
                                                         const float *__restrict__ f, int zc) {
    __shared__ float ring[kSlots][kSlotF];
    const Tile t = at(base, (int)blockIdx.x, (int)blockIdx.y);
    const int tid = (int)threadIdx.x;
    const int z0 = 1 + (int)blockIdx.z * zc;
    const int z1 = min(z0 + zc, t.nz - 1);
    int off[kStageN];
    float sv0[kStageN], sv1[kStageN], fr0[3 * kRedN], fr1[3 * kRedN], fb0[3], fb1[3];
    stage_map(t, tid, off);
    for (int q = z0 - 2; q <= z0 + 1; ++q) {
        stage_load(t, ui, off, q, sv0);
        stage_store(sv0, ring[slot(q)], tid);
    }
    red_load(t, f, z0 - 1, tid, fr0);
    red_load(t, f, z0, tid, fr1);
    __syncthreads();
    red(t, fr0, ring[slot(z0 - 2)], ring[slot(z0 - 1)], ring[slot(z0)], z0 - 1, tid);
    red(t, fr1, ring[slot(z0 - 1)], ring[slot(z0)], ring[slot(z0 + 1)], z0, tid);
    // set 0 serves steps z0, z0 + 2, ...; set 1 steps z0 + 1, z0 + 3, ...
    stage_load(t, ui, off, z0 + 2, sv0);
    red_load(t, f, z0 + 1, tid, fr0);
    black_load(t, f, z0, tid, fb0);
    stage_load(t, ui, off, z0 + 3, sv1);
    red_load(t, f, z0 + 2, tid, fr1);
    black_load(t, f, z0 + 1, tid, fb1);
    for (int z = z0; z < z1; z += 2) {
        march_step(t, ui, uo, f, ring, off, z, tid, sv0, fr0, fb0);
        if (z + 1 < z1) march_step(t, ui, uo, f, ring, off, z + 1, tid, sv1, fr1, fb1);
    }
}

}  // namespace rbgs
}  // namespace vg
