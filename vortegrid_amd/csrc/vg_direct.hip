// vg_direct.hip -- O(N^2) direct Biot-Savart sums (VortonSim::
// ComputeVelocity_Direct / ComputeVectorPotential_Direct, vortonSim.cpp:
// 1113-1158) at arbitrary query points and at grid points.
//
// Layout: sources are first packed into two float4 streams with the
// per-source factors hoisted (A = pos, r^2; B = Omega*(2/3)r^3, 1/r^3).
// A workgroup of 256 threads owns 256*QPT queries; source tiles of 256 are
// staged through LDS and read as wave-wide broadcasts.  Each thread keeps a
// plain FP32 partial per tile and folds it into a Kahan-compensated total
// once per tile.  Small problems split the source range over blockIdx.y and
// reduce the partials in fixed order (deterministic).
#include <mutex>
#include <string>

#include "vg_internal.h"

// Fixed device scratch of vg_direct_velocity_noalloc (32 MiB per GPU).
constexpr size_t kFixedScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kFixedScratchBytes];

namespace vg {
namespace {

constexpr int kBlock = 256;
constexpr int kTile = 256;
constexpr int kQPT = 2;

struct SourceView {
    const float *pos, *angvel, *size;
    size_t stride;
    float usize;
};

struct QueryView {
    const float *q;      // strided query positions (grid == 0)
    size_t stride;
    vg_grid_desc g;      // grid geometry (grid == 1)
    size_t nq;
};

template <int GRID>
__device__ __forceinline__ float3 query_at(const QueryView &qv, size_t i) {
    if (GRID) {
        const uint32_t nx = qv.g.np[0], ny = qv.g.np[1];
        const uint32_t ix = (uint32_t)(i % nx);
        const size_t r = i / nx;
        const uint32_t iy = (uint32_t)(r % ny), iz = (uint32_t)(r / ny);
        return grid_query_point(qv.g, ix, iy, iz);
    } else {
        const float *p = qv.q + i * qv.stride;
        return make_float3(p[0], p[1], p[2]);
    }
}

__global__ void k_pack_sources(SourceView sv, size_t n, float spread, float4 *__restrict__ A,
                               float4 *__restrict__ B) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *p = sv.pos + i * sv.stride;
    const float *w = sv.angvel + i * sv.stride;
    const float sz = sv.size ? sv.size[i * sv.stride] : sv.usize;
    float4 a, b;
    prepare_source(make_float4(p[0], p[1], p[2], sz), make_float4(w[0], w[1], w[2], 0.f), spread,
                   a, b);
    A[i] = a;
    B[i] = b;
}

template <int POLICY, int VECPOT, int GRID>
__global__ __launch_bounds__(kBlock) void k_direct(const float4 *__restrict__ A,
                                                   const float4 *__restrict__ B, size_t n,
                                                   size_t chunk, QueryView qv, float circ,
                                                   float *__restrict__ out,
                                                   float *__restrict__ partial) {
    __shared__ float4 sA[kTile];
    __shared__ float4 sB[kTile];
    const size_t q0 = (size_t)blockIdx.x * (kBlock * kQPT) + threadIdx.x;
    float3 q[kQPT], sum[kQPT], comp[kQPT];
#pragma unroll
    for (int k = 0; k < kQPT; ++k) {
        const size_t qi = q0 + (size_t)k * kBlock;
        q[k] = qi < qv.nq ? query_at<GRID>(qv, qi) : make_float3(0.f, 0.f, 0.f);
        sum[k] = comp[k] = make_float3(0.f, 0.f, 0.f);
    }
    const size_t s0 = (size_t)blockIdx.y * chunk;
    const size_t s1 = s0 + chunk < n ? s0 + chunk : n;
    for (size_t t0 = s0; t0 < s1; t0 += kTile) {
        const size_t cnt = s1 - t0 < (size_t)kTile ? s1 - t0 : (size_t)kTile;
        __syncthreads();
        if (threadIdx.x < cnt) {
            sA[threadIdx.x] = A[t0 + threadIdx.x];
            sB[threadIdx.x] = B[t0 + threadIdx.x];
        }
        __syncthreads();
        float3 part[kQPT];
#pragma unroll
        for (int k = 0; k < kQPT; ++k) part[k] = make_float3(0.f, 0.f, 0.f);
        if (cnt == (size_t)kTile) {
#pragma unroll 4
            for (int j = 0; j < kTile; ++j) {
                const float4 a = sA[j], b = sB[j];
#pragma unroll
                for (int k = 0; k < kQPT; ++k) {
                    if (VECPOT) pair_vecpot<POLICY>(part[k], q[k], a, b);
                    else pair_velocity<POLICY>(part[k], q[k], a, b);
                }
            }
        } else {
            for (int j = 0; j < (int)cnt; ++j) {
                const float4 a = sA[j], b = sB[j];
#pragma unroll
                for (int k = 0; k < kQPT; ++k) {
                    if (VECPOT) pair_vecpot<POLICY>(part[k], q[k], a, b);
                    else pair_velocity<POLICY>(part[k], q[k], a, b);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kQPT; ++k) kahan_add(sum[k], comp[k], part[k]);
    }
#pragma unroll
    for (int k = 0; k < kQPT; ++k) {
        const size_t qi = q0 + (size_t)k * kBlock;
        if (qi >= qv.nq) continue;
        if (partial) {
            float *o = partial + 3 * ((size_t)blockIdx.y * qv.nq + qi);
            o[0] = sum[k].x; o[1] = sum[k].y; o[2] = sum[k].z;
        } else {
            float *o = out + 3 * qi;
            o[0] = sum[k].x * circ; o[1] = sum[k].y * circ; o[2] = sum[k].z * circ;
        }
    }
}

// Fixed-order Kahan reduction of the source-split partials.
__global__ void k_reduce_split(const float *__restrict__ partial, uint32_t nsplit, size_t nq,
                               float circ, float *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    float3 sum = make_float3(0.f, 0.f, 0.f), comp = sum;
    for (uint32_t s = 0; s < nsplit; ++s) {
        const float *p = partial + 3 * ((size_t)s * nq + i);
        kahan_add(sum, comp, make_float3(p[0], p[1], p[2]));
    }
    out[3 * i] = sum.x * circ;
    out[3 * i + 1] = sum.y * circ;
    out[3 * i + 2] = sum.z * circ;
}

// ---------------------------------------------------------------- FP64
struct Src64 {
    double x, y, z, r2, sx, sy, sz, ir3;
};

__global__ __launch_bounds__(kBlock) void k_direct_f64(SourceView sv, size_t n, QueryView qv,
                                                       float spread, float circ,
                                                       double *__restrict__ out) {
    __shared__ Src64 sS[kTile];
    const size_t qi = (size_t)blockIdx.x * kBlock + threadIdx.x;
    double qx = 0, qy = 0, qz = 0;
    if (qi < qv.nq) {
        const float3 q = query_at<0>(qv, qi);
        qx = q.x; qy = q.y; qz = q.z;
    }
    double ax = 0, ay = 0, az = 0;
    for (size_t t0 = 0; t0 < n; t0 += kTile) {
        const size_t cnt = n - t0 < (size_t)kTile ? n - t0 : (size_t)kTile;
        __syncthreads();
        if (threadIdx.x < cnt) {
            const size_t i = t0 + threadIdx.x;
            const float *p = sv.pos + i * sv.stride;
            const float *w = sv.angvel + i * sv.stride;
            const double sz = sv.size ? (double)sv.size[i * sv.stride] : (double)sv.usize;
            const double r = sz * 0.5 * (double)spread;
            const double r2 = r * r;
            const double s = (2.0 / 3.0) * r2 * r * (double)circ;
            Src64 e;
            e.x = p[0]; e.y = p[1]; e.z = p[2]; e.r2 = r2;
            e.sx = w[0] * s; e.sy = w[1] * s; e.sz = w[2] * s;
            e.ir3 = 1.0 / (r2 * r);
            sS[threadIdx.x] = e;
        }
        __syncthreads();
        for (int j = 0; j < (int)cnt; ++j) {
            const Src64 e = sS[j];
            const double dx = qx - e.x, dy = qy - e.y, dz = qz - e.z;
            const double d2 = dx * dx + dy * dy + dz * dz;
            double law;
            if (d2 < e.r2) {
                law = e.ir3;
            } else {
                double y = __builtin_amdgcn_rsq(d2);
                y = y * (1.5 - 0.5 * d2 * y * y);
                y = y * (1.5 - 0.5 * d2 * y * y);
                law = y * y * y;
            }
            ax += (e.sy * dz - e.sz * dy) * law;
            ay += (e.sz * dx - e.sx * dz) * law;
            az += (e.sx * dy - e.sy * dx) * law;
        }
    }
    if (qi < qv.nq) {
        out[3 * qi] = ax;
        out[3 * qi + 1] = ay;
        out[3 * qi + 2] = az;
    }
}

vg_status check_vortons(const vg_vortons *v) {
    VG_REQUIRE(v, "vortons is null");
    VG_REQUIRE(v->n == 0 || (v->pos && v->angvel), "vorton arrays are null");
    VG_REQUIRE(v->n == 0 || v->stride >= 3, "vorton stride must be >= 3 floats");
    VG_REQUIRE(v->size || v->uniform_size > 0.f || v->n == 0,
               "vortons need a size array or a positive uniform_size");
    return VG_OK;
}

SourceView source_view(const vg_vortons *v) {
    return SourceView{v->pos, v->angvel, v->size, v->stride, v->uniform_size};
}

// Caller-provided scratch for the transient buffers (null: stream-ordered
// allocations).
struct Scratch {
    unsigned char *base = nullptr;
    size_t bytes = 0;
};

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

template <int VECPOT, int GRID>
vg_status run_direct(const vg_vortons *v, QueryView qv, float spread, float circ, int policy,
                     float *out, hipStream_t s, const char *name, Scratch fixed = Scratch()) {
    VG_TRY(check_vortons(v));
    VG_REQUIRE(out || qv.nq == 0, "out is null");
    VG_REQUIRE(policy == VG_RSQRT_REFERENCE || policy == VG_RSQRT_ACCURATE, "bad rsqrt policy");
    if (qv.nq == 0) return VG_OK;
    if (v->n == 0) {
        VG_TRY(zero_fill(out, 3 * qv.nq * sizeof(float), s));
        return VG_OK;
    }
    const size_t n = v->n;
    // Split the source range when the query count alone cannot fill 256 CUs.
    const uint32_t qblocks = ceil_div(qv.nq, (size_t)kBlock * kQPT);
    uint32_t nsplit = 1;
    const uint32_t target_blocks = 2048;
    if (qblocks < target_blocks) {
        nsplit = (target_blocks + qblocks - 1) / qblocks;
        const uint32_t max_split = ceil_div(n, 4 * (size_t)kTile);
        if (nsplit > max_split) nsplit = max_split;
        if (nsplit < 1) nsplit = 1;
    }
    size_t chunk = (n + nsplit - 1) / nsplit;
    chunk = (chunk + kTile - 1) / kTile * kTile;
    nsplit = (uint32_t)((n + chunk - 1) / chunk);
    const size_t src_bytes = n * sizeof(float4);
    const size_t part_bytes = nsplit > 1 ? (size_t)nsplit * qv.nq * 3 * sizeof(float) : 0;
    float4 *A = nullptr, *B = nullptr;
    float *partial = nullptr;
    if (fixed.base) {
        const size_t need = 2 * align256(src_bytes) + align256(part_bytes);
        if (need > fixed.bytes) {
            set_error("direct sum needs " + std::to_string(need) + " bytes of scratch (" +
                      std::to_string(n) + " vortons, " + std::to_string(qv.nq) + " queries, " +
                      std::to_string(nsplit) + " source splits); the fixed scratch holds " +
                      std::to_string(fixed.bytes));
            return VG_EINVAL;
        }
        A = reinterpret_cast<float4 *>(fixed.base);
        B = reinterpret_cast<float4 *>(fixed.base + align256(src_bytes));
        if (part_bytes)
            partial = reinterpret_cast<float *>(fixed.base + 2 * align256(src_bytes));
    } else {
        VG_TRY(salloc((void **)&A, src_bytes, s));
        VG_TRY(salloc((void **)&B, src_bytes, s));
        if (part_bytes) VG_TRY(salloc((void **)&partial, part_bytes, s));
    }
    VG_TRY(launch("vg_pack_sources", k_pack_sources, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                  source_view(v), n, spread, A, B));
    vg_status st;
    auto kern = policy == VG_RSQRT_REFERENCE ? k_direct<VG_RSQRT_REFERENCE, VECPOT, GRID>
                                             : k_direct<VG_RSQRT_ACCURATE, VECPOT, GRID>;
    st = launch(name, kern, dim3(qblocks, nsplit), dim3(kBlock), 0, s, (const float4 *)A,
                (const float4 *)B, n, chunk, qv, circ, out, partial);
    if (st == VG_OK && nsplit > 1)
        st = launch("vg_direct_reduce", k_reduce_split, dim3(ceil_div(qv.nq, 256)), dim3(256), 0,
                    s, (const float *)partial, nsplit, qv.nq, circ, out);
    if (!fixed.base) {
        sfree(partial, s);
        sfree(A, s);
        sfree(B, s);
    }
    return st;
}

// ---------------------------------------------------------------- fixed scratch
// The transient buffers of vg_direct_velocity_noalloc: one array per GPU in
// the code object, kept until the process ends.  In host code `scratch` only
// names the symbol; its device address on the current GPU comes from
// hipGetSymbolAddress, taken once per device and kept in g_scratch_ptr.
constexpr int kMaxDevices = 64;
unsigned char *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

vg_status fixed_scratch(Scratch *out) {
    int dev = 0;
    VG_HIP(hipGetDevice(&dev));
    VG_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index beyond the fixed-scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        VG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        VG_REQUIRE(p, "no device address for the fixed scratch array");
        g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
    }
    out->base = g_scratch_ptr[dev];
    out->bytes = kFixedScratchBytes;
    return VG_OK;
}

}  // namespace
}  // namespace vg

using namespace vg;

VG_API vg_status vg_direct_velocity(const vg_vortons *v, const float *q, size_t q_stride,
                                    size_t nq, float spread, float circ, int rsqrt_policy,
                                    float *out, vg_stream s) {
    VG_REQUIRE(nq == 0 || (q && q_stride >= 3), "bad query array");
    QueryView qv{q, q_stride, {}, nq};
    return run_direct<0, 0>(v, qv, spread, circ, rsqrt_policy, out, (hipStream_t)s,
                            "vg_direct_velocity");
}

VG_API vg_status vg_direct_velocity_noalloc(const vg_vortons *v, const float *q,
                                            size_t q_stride, size_t nq, float spread, float circ,
                                            int rsqrt_policy, float *out, vg_stream s) {
    VG_REQUIRE(nq == 0 || (q && q_stride >= 3), "bad query array");
    Scratch fixed;
    VG_TRY(fixed_scratch(&fixed));
    QueryView qv{q, q_stride, {}, nq};
    return run_direct<0, 0>(v, qv, spread, circ, rsqrt_policy, out, (hipStream_t)s,
                            "vg_direct_velocity", fixed);
}

VG_API vg_status vg_direct_vecpot(const vg_vortons *v, const float *q, size_t q_stride, size_t nq,
                                  float spread, float circ, int rsqrt_policy, float *out,
                                  vg_stream s) {
    VG_REQUIRE(nq == 0 || (q && q_stride >= 3), "bad query array");
    QueryView qv{q, q_stride, {}, nq};
    return run_direct<1, 0>(v, qv, spread, circ, rsqrt_policy, out, (hipStream_t)s,
                            "vg_direct_vecpot");
}

VG_API vg_status vg_direct_velocity_grid(const vg_vortons *v, const vg_grid_desc *g, float spread,
                                         float circ, int rsqrt_policy, float *out, vg_stream s) {
    VG_REQUIRE(g, "grid is null");
    QueryView qv{nullptr, 0, *g, capacity(*g)};
    return run_direct<0, 1>(v, qv, spread, circ, rsqrt_policy, out, (hipStream_t)s,
                            "vg_direct_velocity_grid");
}

VG_API vg_status vg_direct_velocity_f64(const vg_vortons *v, const float *q, size_t q_stride,
                                        size_t nq, float spread, float circ, double *out,
                                        vg_stream s) {
    VG_TRY(check_vortons(v));
    VG_REQUIRE(nq == 0 || (q && q_stride >= 3 && out), "bad query array");
    if (nq == 0) return VG_OK;
    hipStream_t st = (hipStream_t)s;
    if (v->n == 0) {
        VG_TRY(zero_fill(out, 3 * nq * sizeof(double), st));
        return VG_OK;
    }
    QueryView qv{q, q_stride, {}, nq};
    return launch("vg_direct_velocity_f64", k_direct_f64, dim3(ceil_div(nq, kBlock)),
                  dim3(kBlock), 0, st, source_view(v), v->n, qv, spread, circ, out);
}
