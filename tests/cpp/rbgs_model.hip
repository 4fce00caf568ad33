// Host model of k_rbgs_sweep (vortegrid_amd/csrc/vg_rbgs.h).  It runs the
// kernel's own phase functions serially, workgroup by workgroup, in the
// order of the kernel's barriers, and compares the result with a plain
// two-colour red-black sweep, bit for bit, over grid sizes that exercise
// partial tiles, one-plane tiles and boundary-only halos.
//
// Per phase the 256 threads run forward, or in reverse with the next
// plane's stores moved ahead of the passes they may overlap on the GPU
// (no barrier separates them), so an aliasing ring slot shows as a
// mismatch.  Built with AddressSanitizer on the host: every global and LDS
// access of the phase functions lands in an exactly sized heap block.
// Prints "ALL MATCH" and exits 0 on success.
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <random>
#include <cstring>
#include "../../vortegrid_amd/csrc/vg_rbgs.h"
using namespace vg::rbgs;

static void plain(int nx, int ny, int nz, std::vector<float> &u, const std::vector<float> &f,
                  float r2x, float r2y, float r2z, float inv) {
#pragma clang fp contract(off)
    const size_t sy = 3 * (size_t)nx, sz = sy * ny;
    for (int color = 0; color < 2; ++color)
        for (int z = 1; z < nz - 1; ++z)
            for (int y = 1; y < ny - 1; ++y)
                for (int x = 1; x < nx - 1; ++x) {
                    if (((x + y + z) & 1) != color) continue;
                    const size_t o = z * sz + y * sy + 3 * (size_t)x;
                    for (int c = 0; c < 3; ++c) {
                        const float s = (u[o + 3 + c] + u[o - 3 + c]) * r2x +
                                        (u[o + sy + c] + u[o - sy + c]) * r2y +
                                        (u[o + sz + c] + u[o - sz + c]) * r2z;
                        u[o + c] = (s - f[o + c]) * inv;
                    }
                }
}

int main() {
    const int sizes[][3] = {{5, 5, 5}, {9, 9, 17}, {17, 17, 33}, {33, 33, 65}, {67, 19, 23},
                            {66, 10, 12}, {3, 3, 3}, {130, 11, 7}, {65, 65, 129}};
    const int zcs[] = {1, 2, 3, 8, 11, 16};
    int bad = 0;
    for (auto &sz3 : sizes) {
        const int nx = sz3[0], ny = sz3[1], nz = sz3[2];
        const size_t n = 3 * (size_t)nx * ny * nz;
        std::mt19937 rng(nx * 1000 + ny * 10 + nz);
        std::uniform_real_distribution<float> U(-1.f, 1.f);
        float *ui = (float *)malloc(n * 4), *f = (float *)malloc(n * 4), *uo = (float *)malloc(n * 4);
        for (size_t i = 0; i < n; ++i) { ui[i] = U(rng); f[i] = U(rng); }
        const float r2x = 1.3f, r2y = 0.7f, r2z = 2.1f, inv = 1.0f / (2.0f * (r2x + r2y + r2z));
        std::vector<float> ref(ui, ui + n), fv(f, f + n);
        plain(nx, ny, nz, ref, fv, r2x, r2y, r2z, inv);
        for (int zc : zcs) for (int rev = 0; rev < 2; ++rev) {
            memcpy(uo, ui, n * 4);
            // poison uo interior so missing writes show
            for (int z = 1; z < nz - 1; ++z) for (int y = 1; y < ny - 1; ++y) for (int x = 1; x < nx - 1; ++x)
                for (int c = 0; c < 3; ++c) uo[((size_t)z * ny * nx + (size_t)y * nx + x) * 3 + c] = 1e30f;
            Tile base{nx, ny, nz, 0, 0, 3 * (size_t)nx, 3 * (size_t)nx * ny, r2x, r2y, r2z, inv};
            const int gx = (nx - 2 + kSX - 1) / kSX, gy = (ny - 2 + kSY - 1) / kSY, gz = (nz - 2 + zc - 1) / zc;
            float *ring[kSlots];
            for (int k = 0; k < kSlots; ++k) ring[k] = (float *)malloc(kSlotF * 4);
            // two register sets alternate between plane steps, as in the kernel:
            // set s = (z - z0) & 1 serves step z and is refilled for step z + 2
            struct Th { int off[kStageN]; float sv[2][kStageN], fr[2][3 * kRedN], fb[2][3]; };
            std::vector<Th> th(kThreads);
            for (int bz = 0; bz < gz; ++bz) for (int by = 0; by < gy; ++by) for (int bx = 0; bx < gx; ++bx) {
                const Tile t = at(base, bx, by);
                const int z0 = 1 + bz * zc, z1 = std::min(z0 + zc, nz - 1);
                auto each = [&](auto fn) { if (!rev) for (int i = 0; i < kThreads; ++i) fn(i, th[i]); else for (int i = kThreads - 1; i >= 0; --i) fn(i, th[i]); };
                for (int k = 0; k < kSlots; ++k) for (int e = 0; e < kSlotF; ++e) ring[k][e] = 7e30f;
                each([&](int tid, Th &r) {
                    stage_map(t, tid, r.off);
                    for (int q = z0 - 2; q <= z0 + 1; ++q) { stage_load(t, ui, r.off, q, r.sv[0]); stage_store(r.sv[0], ring[slot(q)], tid); }
                    red_load(t, f, z0 - 1, tid, r.fr[0]);
                    red_load(t, f, z0, tid, r.fr[1]);
                });
                // barrier
                if (rev) {  // the store of plane z0+2 racing ahead of the colour-0 passes
                    each([&](int tid, Th &r) { float v[kStageN]; stage_load(t, ui, r.off, z0 + 2, v); stage_store(v, ring[slot(z0 + 2)], tid); });
                }
                each([&](int tid, Th &r) {
                    red(t, r.fr[0], ring[slot(z0 - 2)], ring[slot(z0 - 1)], ring[slot(z0)], z0 - 1, tid);
                    red(t, r.fr[1], ring[slot(z0 - 1)], ring[slot(z0)], ring[slot(z0 + 1)], z0, tid);
                    stage_load(t, ui, r.off, z0 + 2, r.sv[0]);
                    red_load(t, f, z0 + 1, tid, r.fr[0]);
                    black_load(t, f, z0, tid, r.fb[0]);
                    stage_load(t, ui, r.off, z0 + 3, r.sv[1]);
                    red_load(t, f, z0 + 2, tid, r.fr[1]);
                    black_load(t, f, z0 + 1, tid, r.fb[1]);
                });
                for (int z = z0; z < z1; ++z) {
                    const int st = (z - z0) & 1;
                    each([&](int tid, Th &r) {
                        if (!rev || z == z0) stage_store(r.sv[st], ring[slot(z + 2)], tid);
                        stage_load(t, ui, r.off, z + 4, r.sv[st]);
                    });
                    each([&](int tid, Th &r) { red(t, r.fr[st], ring[slot(z)], ring[slot(z + 1)], ring[slot(z + 2)], z + 1, tid); });
                    if (rev && z + 1 < z1)  // the next plane's store racing ahead of this colour-1 pass
                        each([&](int tid, Th &r) { stage_store(r.sv[st ^ 1], ring[slot(z + 3)], tid); });
                    each([&](int tid, Th &r) {
                        black(t, r.fb[st], uo, ring[slot(z - 1)], ring[slot(z)], ring[slot(z + 1)], z, tid);
                        red_load(t, f, z + 3, tid, r.fr[st]);
                        black_load(t, f, z + 2, tid, r.fb[st]);
                    });
                }
            }
            for (int k = 0; k < kSlots; ++k) free(ring[k]);
            size_t mism = 0;
            for (size_t i = 0; i < n; ++i) if (memcmp(&uo[i], &ref[i], 4)) ++mism;
            if (mism) { ++bad; printf("MISMATCH %dx%dx%d zc=%d rev=%d: %zu\n", nx, ny, nz, zc, rev, mism); }
        }
        printf("%dx%dx%d ok\n", nx, ny, nz);
        free(ui); free(f); free(uo);
    }
    printf(bad ? "FAILED\n" : "ALL MATCH\n");
    return bad != 0;
}
