// vg_mlp.hip -- learned far-field correction (no reference counterpart: the
// reference applies every cluster as a monopole supervorton, vorton.h:161-179
// on clusters from vortonSim.cpp:1337-1569).
//
// A 16 -> 64 -> 64 -> 64 -> 3 network with GELU (tanh form) maps a
// (query, cluster) pair to the difference between the cluster's exact
// velocity (the direct sum over its member vortons) and its monopole plus
// analytic first-moment velocity: the part of the cluster's field its
// moments do not capture.  With h the cube root of the layer's cell volume,
// S the cluster's weight sum (W.w), x = (q - c)/h, r = |x|, e = x/r,
// O = Omega/S, Dt = D/(h S), at = its antisymmetric vector, u = Dt e,
// v = Dt^T e, the inputs are 16 rotation invariants that are even in the
// vorticity and the 3 outputs c weight three vectors that rotate with the
// field and are odd in the vorticity:
//     correction = (2/3)(size/2)^3 S/h^2 * r^-3 [O x e, at, u x e] c
// so the correction is exactly rotation-equivariant and reverses with the
// vorticity (c = (r, 0, 0) is the monopole, (0, -1, 3) the first moment).
//
// Everything that multiplies matrices runs on BF16 MFMA with FP32
// accumulation:
//   * training: k_gemm (v_mfma_f32_16x16x32_bf16, 64x64 block tiles) for the
//     forward and backward products, elementwise GELU / loss / bias-gradient
//     kernels, Adam;
//   * inference at grid points: k_mlp_grid, one wave per 4x4x4 brick, one
//     accepted cluster at a time: the 64 (point, cluster) pairs of the brick
//     form the batch, all four layers run on v_mfma_f32_32x32x16_bf16 with
//     the weights in LDS, and the hidden activations stay in registers
//     between layers (the output layout of one product is consumed directly
//     as the input layout of the next; the weights are packed in that order).
#include <cmath>
#include <cstring>
#include <vector>

#include "vg_internal.h"

namespace vg {
vg_status tree_checked_view(const vg_tree *t, TreeView *v);
vg_status mac_velocity_grid(const vg_tree *t, const vg_grid_desc *g, int near_k, int order,
                            float spread, float circ, int policy, float *out, hipStream_t s);
}  // namespace vg

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kIn = VG_MLP_INPUTS, kHid = VG_MLP_HIDDEN, kOut = VG_MLP_OUTPUTS;
constexpr int kOutPad = 16;  // output layer padded to 16 columns (zero weights)
// Padded device parameter layout (floats).
constexpr int kOffW1 = 0;                       // [16][64]
constexpr int kOffB1 = kOffW1 + kIn * kHid;     // [64]
constexpr int kOffW2 = kOffB1 + kHid;           // [64][64]
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid * kHid;
constexpr int kOffW4 = kOffB3 + kHid;           // [64][16], columns 3..15 zero
constexpr int kOffB4 = kOffW4 + kHid * kOutPad; // [16], entries 3..15 zero
constexpr int kPadParams = kOffB4 + kOutPad;
static_assert(VG_MLP_PARAMS == kIn * kHid + kHid + 2 * (kHid * kHid + kHid) + kHid * kOut + kOut,
              "parameter count");

// Packed inference weights: 22 operand blocks of 64 lanes x 8 bf16.
constexpr int kBlkL1 = 0, kBlkL2 = 2, kBlkL3 = 10, kBlkL4 = 18, kBlocks = 22;
constexpr int kPackedElems = kBlocks * 64 * 8;

constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2/pi)
constexpr float kGeluA = 0.044715f;

__device__ __forceinline__ float tanh_fast(float u) {
    // 1 - 2/(exp(2u) + 1): exact limits +-1, absolute error ~1e-7
    const float e = __expf(2.0f * u);
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

// GELU(z) = z (1 + tanh u) / 2 with u = kGeluC (z + kGeluA z^3), written as
// z - z / (1 + e^(2u)) and e^(2u) = 2^(z (g1 + g3 z^2)): six VALU operations
// besides the exp and the reciprocal (against thirteen for the tanh form);
// the limits stay exact (z for large z, where e^(2u) overflows; 0 for very
// negative z).
constexpr float kGeluG1 = 2.0f * kGeluC * 1.4426950408889634f;  // 2 sqrt(2/pi) log2(e)
constexpr float kGeluG3 = kGeluG1 * kGeluA;
__device__ __forceinline__ float gelu(float z) {
    const float q = z * fmaf(kGeluG3, z * z, kGeluG1);
    const float r = __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(q) + 1.0f);
    return fmaf(-z, r, z);
}

__device__ __forceinline__ float gelu_grad(float z) {
    const float t = tanh_fast(kGeluC * (z + kGeluA * z * z * z));
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * kGeluC * (1.0f + 3.0f * kGeluA * z * z);
}

// ------------------------------------------------------------------ GEMM
// C[M][N] (ldc) = sum_k opA(m, k) opB(k, n) with opA(m, k) = TA ? A[k][m] :
// A[m][k] and opB(k, n) = TB ? B[n][k] : B[k][n].  FP32 in memory, operands
// rounded to bf16 when staged to LDS, FP32 accumulation.  Block = 4 waves,
// 64x64 tile, each wave 32x32 = 2x2 MFMA 16x16x32 tiles, K step 32.
// blockIdx.z splits K into chunks of kc: chunk z writes its partial product
// to C + z * M * ldc (reduced afterwards by k_reduce_split).
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const float *__restrict__ A,
                                              int lda, const float *__restrict__ B, int ldb,
                                              float *__restrict__ C, int ldc, int kc) {
    __shared__ __attribute__((aligned(16))) __bf16 As[64][40];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[64][40];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int kb = blockIdx.z * kc, ke = min(K, kb + kc);
    C += (size_t)blockIdx.z * M * ldc;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += 32) {
        for (int e = tid; e < 64 * 32; e += 256) {
            const int mm = TA ? (e & 63) : (e >> 5), kk = TA ? (e >> 6) : (e & 31);
            const int gm = m0 + mm, gk = k0 + kk;
            float v = 0.f;
            if (gm < M && gk < ke) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
            As[mm][kk] = (__bf16)v;
        }
        for (int e = tid; e < 64 * 32; e += 256) {
            const int nn = TB ? (e >> 5) : (e & 63), kk = TB ? (e & 31) : (e >> 6);
            const int gn = n0 + nn, gk = k0 + kk;
            float v = 0.f;
            if (gn < N && gk < ke) v = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
            Bs[nn][kk] = (__bf16)v;
        }
        __syncthreads();
        const int r = lane & 15, g = lane >> 4;
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = *(const bf16x8 *)&As[wm + i * 16 + r][g * 8];
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = *(const bf16x8 *)&Bs[wn + j * 16 + r][g * 8];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gm = m0 + wm + i * 16 + (lane >> 4) * 4 + v;
                const int gn = n0 + wn + j * 16 + (lane & 15);
                if (gm < M && gn < N) C[(size_t)gm * ldc + gn] = acc[i][j][v];
            }
}

// C[m][n] = sum of the nsplit partial slabs, in slab order.
__global__ void k_reduce_split(const float *__restrict__ part, int nsplit, int M, int N, int ldp,
                               float *__restrict__ C, int ldc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * N) return;
    const int m = i / N, n = i % N;
    float s = 0.f;
    for (int z = 0; z < nsplit; ++z) s += part[(size_t)z * M * ldp + (size_t)m * ldp + n];
    C[(size_t)m * ldc + n] = s;
}

constexpr int kSplitChunk = 64;  // K per block when the product is split
constexpr int kMaxSplit = 64;

// Products with a small output and a long K (the weight gradients, K = the
// batch) are split over K into kSplitChunk chunks reduced in fixed order;
// `split` (kMaxSplit * 64 * 64 floats) holds the partials.
template <bool TA, bool TB>
vg_status gemm(int M, int N, int K, const float *A, int lda, const float *B, int ldb, float *C,
               int ldc, hipStream_t s, float *split = nullptr) {
    const dim3 tiles(vg::ceil_div(M, 64), vg::ceil_div(N, 64));
    int nsplit = split && M <= 64 && N <= 64 ? (int)vg::ceil_div(K, kSplitChunk) : 1;
    if (nsplit > kMaxSplit) nsplit = kMaxSplit;
    if (nsplit <= 1)
        return vg::launch("vg_mlp_gemm", k_gemm<TA, TB>, tiles, dim3(256), 0, s, M, N, K, A, lda, B,
                          ldb, C, ldc, K);
    const int kc = (int)(vg::ceil_div(vg::ceil_div(K, nsplit), 32) * 32);
    nsplit = (int)vg::ceil_div(K, kc);
    VG_TRY(vg::launch("vg_mlp_gemm", k_gemm<TA, TB>, dim3(tiles.x, tiles.y, nsplit), dim3(256), 0, s,
                      M, N, K, A, lda, B, ldb, split, N, kc));
    return vg::launch("vg_mlp_gemm_reduce", k_reduce_split, dim3(vg::ceil_div(M * N, 256)), dim3(256),
                      0, s, (const float *)split, nsplit, M, N, N, C, ldc);
}
// C = A B, C = A^T B, C = A B^T
vg_status gemm_nn(int M, int N, int K, const float *A, int lda, const float *B, int ldb, float *C,
                  int ldc, hipStream_t s) {
    return gemm<false, false>(M, N, K, A, lda, B, ldb, C, ldc, s);
}
vg_status gemm_tn(int M, int N, int K, const float *A, int lda, const float *B, int ldb, float *C,
                  int ldc, hipStream_t s, float *split) {
    return gemm<true, false>(M, N, K, A, lda, B, ldb, C, ldc, s, split);
}
vg_status gemm_nt(int M, int N, int K, const float *A, int lda, const float *B, int ldb, float *C,
                  int ldc, hipStream_t s) {
    return gemm<false, true>(M, N, K, A, lda, B, ldb, C, ldc, s);
}

// ------------------------------------------------------------------ elementwise
__global__ void k_bias_gelu(float *__restrict__ Z, float *__restrict__ A,
                            const float *__restrict__ b, size_t n, int cols) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float z = Z[i] + b[i % cols];
    Z[i] = z;
    A[i] = gelu(z);
}

__global__ void k_bias_add(float *__restrict__ Y, const float *__restrict__ b, size_t n, int cols) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Y[i] += b[i % cols];
}

__global__ void k_gelu_back(float *__restrict__ dA, const float *__restrict__ Z, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dA[i] *= gelu_grad(Z[i]);
}

// Loss of the correction vectors B c (B = identity when null) against T:
// dY = 2 B^T (B c - T) / (3 rows) on the 3 live columns, 0 on the padding;
// one block partial of sum |B c - T|^2 per block.
__global__ __launch_bounds__(256) void k_loss_grad(const float *__restrict__ Y,
                                                   const float *__restrict__ Bm,
                                                   const float *__restrict__ T,
                                                   float *__restrict__ dY, uint32_t rows,
                                                   float scale, float *__restrict__ partial) {
    __shared__ float red[256];
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    float s = 0.f;
    if (r < rows) {
        const float *y = Y + (size_t)r * kOutPad;
        float e[3], g[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            float v = y[i];
            if (Bm) {
                const float *b = Bm + (size_t)r * 9 + 3 * i;
                v = b[0] * y[0] + b[1] * y[1] + b[2] * y[2];
            }
            e[i] = v - T[(size_t)r * kOut + i];
            s += e[i] * e[i];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            g[k] = e[k];
            if (Bm) {
                const float *b = Bm + (size_t)r * 9;
                g[k] = b[k] * e[0] + b[3 + k] * e[1] + b[6 + k] * e[2];
            }
        }
#pragma unroll
        for (int c = 0; c < kOutPad; ++c) dY[(size_t)r * kOutPad + c] = c < kOut ? scale * g[c] : 0.f;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Column sums of D [rows][cols] (cols divides 256) in a fixed order, in two
// stages: workgroup b sums rows [b chunk, (b + 1) chunk) into part[b][cols],
// then k_colsum_final adds the partials in workgroup order.
constexpr uint32_t kColRows = 128;  // rows per workgroup unless the scratch needs more
__global__ __launch_bounds__(256) void k_colsum_part(const float *__restrict__ D, uint32_t rows,
                                                     int cols, uint32_t chunk,
                                                     float *__restrict__ part) {
    __shared__ float red[256];
    const int t = threadIdx.x, c = t % cols, g = t / cols, ng = 256 / cols;
    const uint32_t r0 = blockIdx.x * chunk;
    const uint32_t r1 = rows - r0 < chunk ? rows : r0 + chunk;
    float s = 0.f;
    for (uint32_t r = r0 + (uint32_t)g; r < r1; r += (uint32_t)ng) s += D[(size_t)r * cols + c];
    red[t] = s;
    __syncthreads();
    if (t < cols) {
        float a = 0.f;
        for (int j = 0; j < ng; ++j) a += red[j * cols + t];
        part[(size_t)blockIdx.x * cols + t] = a;
    }
}

__global__ void k_colsum_final(const float *__restrict__ part, uint32_t nparts, int cols,
                               float *__restrict__ out) {
    const int t = threadIdx.x;
    if (t >= cols) return;
    float a = 0.f;
    for (uint32_t b = 0; b < nparts; ++b) a += part[(size_t)b * cols + t];
    out[t] = a;
}

// Column sums of D [rows][cols] into out, with `part` (kMaxSplit * 64 * 64
// floats, the split products' scratch, free between products) holding the
// per-workgroup partials.
vg_status colsum(const float *D, uint32_t rows, int cols, float *out, float *part,
                 hipStream_t s) {
    VG_REQUIRE(cols > 0 && cols <= 64 && 256 % cols == 0, "column sum: cols must divide 256");
    const size_t max_parts = (size_t)kMaxSplit * 64 * 64 / cols;
    uint32_t chunk = kColRows;
    if (vg::ceil_div(rows, chunk) > max_parts) chunk = vg::ceil_div(rows, max_parts);
    const uint32_t nparts = vg::ceil_div(rows, chunk);
    VG_TRY(vg::launch("vg_mlp_colsum", k_colsum_part, dim3(nparts), dim3(256), 0, s, D, rows, cols,
                      chunk, part));
    return vg::launch("vg_mlp_colsum", k_colsum_final, dim3(1), dim3(64), 0, s, (const float *)part,
                      nparts, cols, out);
}

__global__ void k_sum_partials(const float *__restrict__ partial, uint32_t n, float *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        float s = 0.f;
        for (uint32_t i = 0; i < n; ++i) s += partial[i];
        out[0] = s;
    }
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, int n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float v = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    p[i] -= lr * (a / bc1) / (sqrtf(v / bc2) + eps);
}

__global__ void k_copy_cols(const float *__restrict__ src, int lds, float *__restrict__ dst, int ldd,
                            int cols, uint32_t rows) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * cols) return;
    const size_t r = i / cols, c = i % cols;
    dst[r * ldd + c] = src[r * lds + c];
}

// ------------------------------------------------------------------ inference packing
// Feature (input slot) of k-block b, lane group g, element t, as laid out by
// the previous layer's 32x32 accumulator: rows 8 (v/4) + 4 g + v%4 with
// v = 8 (b%2) + t of row tile b/2.
__host__ __device__ inline int hidden_slot(int b, int g, int t) {
    return 16 * b + 8 * (t >> 2) + 4 * g + (t & 3);
}

__global__ void k_mlp_pack(const float *__restrict__ p, __bf16 *__restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kPackedElems) return;
    const int blk = e / 512, lane = (e / 8) % 64, t = e % 8;
    const int i = lane & 31, g = lane >> 5;
    float w = 0.f;
    if (blk < kBlkL2) {  // layer 1, row tile r, inputs 8g + t
        const int r = blk - kBlkL1;
        w = p[kOffW1 + (8 * g + t) * kHid + r * 32 + i];
    } else if (blk < kBlkL4) {
        const int base = blk < kBlkL3 ? kOffW2 : kOffW3;
        const int rb = blk - (blk < kBlkL3 ? kBlkL2 : kBlkL3);
        const int r = rb / 4, b = rb % 4;
        w = p[base + hidden_slot(b, g, t) * kHid + r * 32 + i];
    } else {
        const int b = blk - kBlkL4;
        w = i < kOut ? p[kOffW4 + hidden_slot(b, g, t) * kOutPad + i] : 0.f;
    }
    out[e] = (__bf16)w;
}

// ------------------------------------------------------------------ features
struct Feat {
    float v[kIn];
};

// 16 invariants of one (point, cluster) pair and the three output vectors
// (times r^-3), B[i][k] = component i of vector k.
__device__ __forceinline__ void cluster_features(float *f, float *B, float3 q, float4 p, float4 w,
                                                 float4 m0, float4 m1, float4 m2, float inv_h) {
    const float x = (q.x - p.x) * inv_h, y = (q.y - p.y) * inv_h, z = (q.z - p.z) * inv_h;
    const float ir = __builtin_amdgcn_rsqf(x * x + y * y + z * z);
    const float ex = x * ir, ey = y * ir, ez = z * ir;
    const float iS = 1.0f / w.w, iSh = iS * inv_h;
    const float ox = w.x * iS, oy = w.y * iS, oz = w.z * iS;
    const float D[3][3] = {{m0.x * iSh, m0.y * iSh, m0.z * iSh},
                           {m1.x * iSh, m1.y * iSh, m1.z * iSh},
                           {m2.x * iSh, m2.y * iSh, m2.z * iSh}};
    const float ax = m0.w * iSh, ay = m1.w * iSh, az = m2.w * iSh;
    const float ux = D[0][0] * ex + D[0][1] * ey + D[0][2] * ez;
    const float uy = D[1][0] * ex + D[1][1] * ey + D[1][2] * ez;
    const float uz = D[2][0] * ex + D[2][1] * ey + D[2][2] * ez;
    const float vx = D[0][0] * ex + D[1][0] * ey + D[2][0] * ez;
    const float vy = D[0][1] * ex + D[1][1] * ey + D[2][1] * ez;
    const float vz = D[0][2] * ex + D[1][2] * ey + D[2][2] * ez;
    const float oe = ox * ex + oy * ey + oz * ez;
    const float ae = ax * ex + ay * ey + az * ez;
    const float eu = ex * ux + ey * uy + ez * uz;
    // u x e
    const float wx = uy * ez - uz * ey, wy = uz * ex - ux * ez, wz = ux * ey - uy * ex;
    float fro = 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) fro += D[i][j] * D[i][j];
    const float tr = D[0][0] + D[1][1] + D[2][2];
    f[0] = ir;
    f[1] = ir * ir;
    f[2] = ox * ox + oy * oy + oz * oz;
    f[3] = oe * oe;
    f[4] = ax * ax + ay * ay + az * az;
    f[5] = ae * ae;
    f[6] = ux * ux + uy * uy + uz * uz;
    f[7] = vx * vx + vy * vy + vz * vz;
    f[8] = eu * eu;
    f[9] = fro;
    f[10] = tr * tr;
    f[11] = ox * ax + oy * ay + oz * az;
    f[12] = oe * eu;
    f[13] = ox * ux + oy * uy + oz * uz;
    f[14] = ox * vx + oy * vy + oz * vz;
    f[15] = ax * wx + ay * wy + az * wz;
    if (B) {
        const float ir3 = ir * ir * ir;
        // columns O x e, at, u x e
        B[0] = (oy * ez - oz * ey) * ir3; B[1] = ax * ir3; B[2] = wx * ir3;
        B[3] = (oz * ex - ox * ez) * ir3; B[4] = ay * ir3; B[5] = wy * ir3;
        B[6] = (ox * ey - oy * ex) * ir3; B[7] = az * ir3; B[8] = wz * ir3;
    }
}

__device__ __forceinline__ float correction_scale(float4 p, float4 w, float inv_h) {
    const float r = 0.5f * p.w;
    return (2.0f / 3.0f) * r * r * r * w.w * inv_h * inv_h;
}

// a cluster the correction applies to: non-empty with a positive weight sum
__device__ __forceinline__ bool mlp_cluster(float4 p, float4 w) {
    return p.w != 0.0f && w.w > 1e-30f;
}

struct MlpGridParams {
    int k;
    uint32_t layer_mask;
    uint32_t nb[3];
    uint32_t nbricks;
    float inv_h[vg::kMaxLayers];
};

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint32_t layer_scale(const vg::TreeView &tv, int l, int a) {
    uint32_t s = 1;
    for (int j = 1; j <= l; ++j) s *= tv.dec[j][a];
    return s;
}

// Children box [c0, c1] of the parent near box [lo, hi] on layer l, and the
// layer's near box [n0, n1] for leaf cells [cl0, cl1] widened by k.
__device__ __forceinline__ void layer_boxes(const vg::TreeView &tv, int l, const int *cl0,
                                            const int *cl1, int k, const int *lo, const int *hi,
                                            int *c0, int *c1, int *n0, int *n1) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int dec = (int)tv.dec[l + 1][a];
        const int sc = (int)layer_scale(tv, l, a);
        c0[a] = lo[a] * dec;
        c1[a] = hi[a] * dec + dec - 1;
        const int ncl = (int)tv.layer[l].np[a] - 1;
        int blo = cl0[a] / sc - k, bhi = cl1[a] / sc + k;
        blo = blo < 0 ? 0 : blo;
        bhi = bhi > ncl - 1 ? ncl - 1 : bhi;
        n0[a] = blo > c0[a] ? blo : c0[a];
        n1[a] = bhi < c1[a] ? bhi : c1[a];
    }
}

__device__ __forceinline__ void brick_cells(const vg_grid_desc &g, const uint32_t *bb, int *cl0,
                                            int *cl1) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int nc = (int)g.np[a] - 1;
        const int first = (int)(4 * bb[a]);
        int last = first + 3;
        if (last > (int)g.np[a] - 1) last = (int)g.np[a] - 1;
        cl0[a] = first < nc - 1 ? first : nc - 1;
        cl1[a] = last < nc - 1 ? last : nc - 1;
    }
}

__device__ __forceinline__ float3 brick_point(const vg_grid_desc &g, const uint32_t *bb, uint32_t p,
                                              bool *valid) {
    const uint32_t lp[3] = {p & 3u, (p >> 2) & 3u, p >> 4};
    uint32_t pt[3];
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const uint32_t v = 4 * bb[a] + lp[a];
        ok = ok && v < g.np[a];
        pt[a] = v < g.np[a] ? v : g.np[a] - 1;
    }
    if (valid) *valid = ok;
    return vg::grid_query_point(g, pt[0], pt[1], pt[2]);
}

// ------------------------------------------------------------------ fused inference
struct ClusterStage {
    float4 P[64], W[64], M0[64], M1[64], M2[64];
};

__device__ __forceinline__ bf16x8 pack8(const float *v) {
    bf16x8 r;
#pragma unroll
    for (int t = 0; t < 8; ++t) r[t] = (__bf16)v[t];
    return r;
}

// One wave = one brick; for every accepted cluster of the masked layers the
// brick's 64 points form one batch of 64 pairs (two 32-column tiles).
__global__ __launch_bounds__(256, 2) void k_mlp_grid(const vg::TreeView tv, vg_grid_desc g,
                                                  MlpGridParams mp,
                                                  const __bf16 *__restrict__ packed,
                                                  const float *__restrict__ params,
                                                  float *__restrict__ out) {
    __shared__ __attribute__((aligned(16))) __bf16 wts[kPackedElems];
    __shared__ float bias[3 * kHid + kOutPad];
    // hidden-layer biases in accumulator order: binit[layer][r][grp] holds
    // the 16 rows 32 r + 8 (v >> 2) + 4 grp + (v & 3) that element v of a
    // 32x32 tile's accumulator covers on lane group grp; the products start
    // from them instead of adding them afterwards
    __shared__ f32x4 binit[3][2][2][4];
    __shared__ ClusterStage stage[4];
    {
        const uint4 *src = (const uint4 *)packed;
        uint4 *dst = (uint4 *)wts;
        for (int i = threadIdx.x; i < kPackedElems / 8; i += 256) dst[i] = src[i];
        for (int i = threadIdx.x; i < 3 * kHid + kOutPad; i += 256) {
            const int layer = i / kHid;
            bias[i] = layer == 0 ? params[kOffB1 + i]
                    : layer == 1 ? params[kOffB2 + i - kHid]
                    : layer == 2 ? params[kOffB3 + i - 2 * kHid]
                                 : params[kOffB4 + i - 3 * kHid];
        }
        for (int i = threadIdx.x; i < 3 * 2 * 2 * 16; i += 256) {
            const int v = i & 15, grp_i = (i >> 4) & 1, r = (i >> 5) & 1, layer = i >> 6;
            const int row = 32 * r + 8 * (v >> 2) + 4 * grp_i + (v & 3);
            const int off = layer == 0 ? kOffB1 : layer == 1 ? kOffB2 : kOffB3;
            binit[layer][r][grp_i][v >> 2][v & 3] = params[off + row];
        }
    }
    __syncthreads();
    const uint32_t wave = uni(threadIdx.x >> 6);
    const uint32_t brick = uni(blockIdx.x * 4 + wave);
    if (brick >= mp.nbricks) return;
    ClusterStage &S = stage[wave];
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t col = lane & 31, grp = lane >> 5;
    const uint32_t bb[3] = {brick % mp.nb[0], (brick / mp.nb[0]) % mp.nb[1],
                            brick / (mp.nb[0] * mp.nb[1])};
    bool valid;
    const float3 qo = brick_point(g, bb, lane, &valid);  // this lane's own point
    const float3 qa = brick_point(g, bb, col, nullptr);       // pair column tile 0
    const float3 qb = brick_point(g, bb, col + 32, nullptr);  // pair column tile 1
    int cl0[3], cl1[3];
    brick_cells(g, bb, cl0, cl1);
    const bf16x8 *W8 = (const bf16x8 *)wts;
    float3 corr = make_float3(0.f, 0.f, 0.f);
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int l = (int)tv.depth - 2; l >= 0; --l) {
        int c0[3], c1[3], n0[3], n1[3];
        layer_boxes(tv, l, cl0, cl1, mp.k, lo, hi, c0, c1, n0, n1);
        if ((mp.layer_mask >> l) & 1u) {
            const vg_grid_desc &gl = tv.layer[l];
            const float inv_h = mp.inv_h[l];
            const int bx = c1[0] - c0[0] + 1, by = c1[1] - c0[1] + 1, bz = c1[2] - c0[2] + 1;
            const int total = bx * by * bz;
            for (int base = 0; base < total; base += 64) {
                const int t = base + (int)lane;
                bool use = false;
                float4 p = make_float4(0.f, 0.f, 0.f, 0.f), w = p, m0 = p, m1 = p, m2 = p;
                if (t < total) {
                    const int cx = c0[0] + t % bx, cy = c0[1] + (t / bx) % by, cz = c0[2] + t / (bx * by);
                    const bool near = cx >= n0[0] && cx <= n1[0] && cy >= n0[1] && cy <= n1[1] &&
                                      cz >= n0[2] && cz <= n1[2];
                    if (!near) {
                        const size_t off = vg::point_offset(gl, cx, cy, cz);
                        p = tv.P[l][off];
                        w = tv.W[l][off];
                        use = mlp_cluster(p, w);
                        if (use) {
                            m0 = tv.M[l][3 * off];
                            m1 = tv.M[l][3 * off + 1];
                            m2 = tv.M[l][3 * off + 2];
                        }
                    }
                }
                unsigned long long mask = __ballot(use);
                S.P[lane] = p;
                S.W[lane] = w;
                S.M0[lane] = m0;
                S.M1[lane] = m1;
                S.M2[lane] = m2;
                __builtin_amdgcn_wave_barrier();
                while (mask) {
                    // The weights and biases are re-read from LDS for every
                    // cluster: hoisted out of this loop they hold ~190
                    // registers and leave one wave per SIMD to hide the
                    // MFMA, LDS and exp latencies alone.  The empty asm
                    // (no instruction) tells the compiler that LDS may
                    // change here, so it does not hoist them.
                    __asm__ volatile("" ::: "memory");
                    const int j = __builtin_ctzll(mask);
                    mask &= mask - 1;
                    const float4 cp = S.P[j], cw = S.W[j];
                    const float4 a0 = S.M0[j], a1 = S.M1[j], a2 = S.M2[j];
                    // layer-1 inputs: lane group 0 holds features 0..7, group 1 8..15
                    float fa[kIn], fb[kIn];
                    cluster_features(fa, nullptr, qa, cp, cw, a0, a1, a2, inv_h);
                    cluster_features(fb, nullptr, qb, cp, cw, a0, a1, a2, inv_h);
                    const bf16x8 xa = pack8(fa + 8 * grp), xb = pack8(fb + 8 * grp);
                    auto bias16 = [&](int layer, int r) {
                        f32x16 b;
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const f32x4 t = binit[layer][r][grp][q];
                            b[4 * q] = t[0];
                            b[4 * q + 1] = t[1];
                            b[4 * q + 2] = t[2];
                            b[4 * q + 3] = t[3];
                        }
                        return b;
                    };
                    f32x16 acc[2][2];
#pragma unroll
                    for (int r = 0; r < 2; ++r) {
                        const bf16x8 wv = W8[(kBlkL1 + r) * 64 + lane];
                        const f32x16 b0 = bias16(0, r);
                        acc[r][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wv, xa, b0, 0, 0, 0);
                        acc[r][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wv, xb, b0, 0, 0, 0);
                    }
                    bf16x8 h[2][4];
#pragma unroll
                    for (int layer = 0; layer < 3; ++layer) {
                        // GELU on the accumulators (biases included), repacked as the next inputs
#pragma unroll
                        for (int c = 0; c < 2; ++c)
#pragma unroll
                            for (int b = 0; b < 4; ++b) {
                                float v8[8];
#pragma unroll
                                for (int t8 = 0; t8 < 8; ++t8) {
                                    v8[t8] = gelu(acc[b >> 1][c][8 * (b & 1) + t8]);
                                }
                                h[c][b] = pack8(v8);
                            }
                        if (layer == 2) break;
                        const int blk0 = layer == 0 ? kBlkL2 : kBlkL3;
#pragma unroll
                        for (int r = 0; r < 2; ++r) {
                            f32x16 s0 = bias16(layer + 1, r), s1 = s0;
#pragma unroll
                            for (int b = 0; b < 4; ++b) {
                                const bf16x8 wv = W8[(blk0 + r * 4 + b) * 64 + lane];
                                s0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wv, h[0][b], s0, 0, 0, 0);
                                s1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wv, h[1][b], s1, 0, 0, 0);
                            }
                            acc[r][0] = s0;
                            acc[r][1] = s1;
                        }
                    }
                    f32x16 y0 = {}, y1 = {};
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const bf16x8 wv = W8[(kBlkL4 + b) * 64 + lane];
                        y0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wv, h[0][b], y0, 0, 0, 0);
                        y1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wv, h[1][b], y1, 0, 0, 0);
                    }
                    // outputs 0..2 sit in lane group 0 (rows v = 0..2); pair
                    // 32 + c of tile 1 moves to lane 32 + c
                    float o[3];
#pragma unroll
                    for (int v = 0; v < 3; ++v) {
                        const float other = __shfl(y1[v], (int)col, 64);
                        o[v] = (lane < 32 ? y0[v] : other) + bias[3 * kHid + v];
                    }
                    // correction = scale * B c at this lane's own point
                    float fo[kIn], Bo[9];
                    cluster_features(fo, Bo, qo, cp, cw, a0, a1, a2, inv_h);
                    const float sc = correction_scale(cp, cw, inv_h);
                    corr.x += (Bo[0] * o[0] + Bo[1] * o[1] + Bo[2] * o[2]) * sc;
                    corr.y += (Bo[3] * o[0] + Bo[4] * o[1] + Bo[5] * o[2]) * sc;
                    corr.z += (Bo[6] * o[0] + Bo[7] * o[1] + Bo[8] * o[2]) * sc;
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) { lo[a] = n0[a]; hi[a] = n1[a]; }
    }
    if (valid) {
        bool dummy;
        (void)dummy;
        uint32_t pt[3];
        const uint32_t lp[3] = {lane & 3u, (lane >> 2) & 3u, lane >> 4};
#pragma unroll
        for (int a = 0; a < 3; ++a) pt[a] = 4 * bb[a] + lp[a];
        const size_t o = vg::point_offset(g, pt[0], pt[1], pt[2]);
        out[3 * o] += corr.x;
        out[3 * o + 1] += corr.y;
        out[3 * o + 2] += corr.z;
    }
}

// ------------------------------------------------------------------ training samples
__device__ __forceinline__ uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct SampleParams {
    int k;
    uint32_t layer_mask;  // only layers 0..depth-2
    uint32_t nb[3];
    uint32_t nbricks;
    float inv_h[vg::kMaxLayers];
    float spread, circ;
    uint32_t n;
    uint64_t seed;
};

// One thread per sample: a random grid point, a random masked layer and a
// random accepted cluster of that layer for the point's brick; target =
// (direct sum over the cluster's member vortons - its monopole) / scale.
__global__ __launch_bounds__(256) void k_mlp_samples(const vg::TreeView tv, vg_grid_desc g,
                                                     SampleParams sp, float *__restrict__ X,
                                                     float *__restrict__ Bout,
                                                     float *__restrict__ T,
                                                     uint32_t *__restrict__ meta,
                                                     uint32_t *__restrict__ failed) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sp.n) return;
    uint64_t st = sp.seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(i + 1));
    const int nl = __popc(sp.layer_mask);
    const vg_grid_desc &g0 = tv.layer[0];
    for (int attempt = 0; attempt < 256; ++attempt) {
        const uint32_t brick = (uint32_t)(splitmix(st) % sp.nbricks);
        const uint32_t pidx = (uint32_t)(splitmix(st) & 63u);
        const uint32_t bb[3] = {brick % sp.nb[0], (brick / sp.nb[0]) % sp.nb[1],
                                brick / (sp.nb[0] * sp.nb[1])};
        const float3 q = brick_point(g, bb, pidx, nullptr);
        int pick = (int)(splitmix(st) % (uint64_t)nl), layer = 0;
        for (int l = 0; l < 32; ++l)
            if ((sp.layer_mask >> l) & 1u) {
                if (pick == 0) { layer = l; break; }
                --pick;
            }
        int cl0[3], cl1[3];
        brick_cells(g, bb, cl0, cl1);
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, c0[3], c1[3], n0[3], n1[3];
        for (int l = (int)tv.depth - 2; l >= layer; --l) {
            layer_boxes(tv, l, cl0, cl1, sp.k, lo, hi, c0, c1, n0, n1);
            if (l > layer)
                for (int a = 0; a < 3; ++a) { lo[a] = n0[a]; hi[a] = n1[a]; }
        }
        int cc[3];
        for (int a = 0; a < 3; ++a) cc[a] = c0[a] + (int)(splitmix(st) % (uint64_t)(c1[a] - c0[a] + 1));
        if (cc[0] >= n0[0] && cc[0] <= n1[0] && cc[1] >= n0[1] && cc[1] <= n1[1] && cc[2] >= n0[2] &&
            cc[2] <= n1[2])
            continue;
        const vg_grid_desc &gl = tv.layer[layer];
        const size_t off = vg::point_offset(gl, cc[0], cc[1], cc[2]);
        const float4 p = tv.P[layer][off], w = tv.W[layer][off];
        if (!mlp_cluster(p, w)) continue;
        const float4 m0 = tv.M[layer][3 * off], m1 = tv.M[layer][3 * off + 1],
                     m2 = tv.M[layer][3 * off + 2];
        float f[kIn], Bm[9];
        const float inv_h = sp.inv_h[layer];
        cluster_features(f, Bm, q, p, w, m0, m1, m2, inv_h);
        // member vortons: leaf cells of the cluster, row by row
        int x0[3], x1[3];
        for (int a = 0; a < 3; ++a) {
            const int sc = (int)layer_scale(tv, layer, a);
            x0[a] = cc[a] * sc;
            x1[a] = x0[a] + sc - 1;
            const int nc0 = (int)g0.np[a] - 2;
            if (x1[a] > nc0) x1[a] = nc0;
        }
        float3 sum = make_float3(0.f, 0.f, 0.f), comp = sum;
        for (int z = x0[2]; z <= x1[2]; ++z)
            for (int y = x0[1]; y <= x1[1]; ++y) {
                const size_t k0 = vg::point_offset(g0, x0[0], y, z);
                const size_t k1 = vg::point_offset(g0, x1[0], y, z);
                const uint32_t b0 = tv.cell_start[k0], e0 = tv.cell_start[k1] + tv.cell_count[k1];
                float3 part = make_float3(0.f, 0.f, 0.f);
                for (uint32_t v = b0; v < e0; ++v) {
                    float4 A, B;
                    vg::prepare_source(tv.sorted_p[v], tv.sorted_w[v], sp.spread, A, B);
                    vg::pair_velocity<VG_RSQRT_ACCURATE>(part, q, A, B);
                }
                vg::kahan_add(sum, comp, part);
            }
        float3 mono = make_float3(0.f, 0.f, 0.f);
        {
            float4 A, B;
            vg::prepare_source(p, make_float4(w.x, w.y, w.z, 0.f), sp.spread, A, B);
            vg::pair_velocity<VG_RSQRT_ACCURATE>(mono, q, A, B);
            // analytic first-moment term: -s/|d|^3 [a - 3/|d|^2 (D d) x d],
            // zero inside the core
            const float dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
            const float d2 = dx * dx + dy * dy + dz * dz;
            if (d2 >= A.w) {
                const float rs = __builtin_amdgcn_rsqf(d2);
                const float law = rs * rs * rs, c3 = 3.0f * rs * rs;
                const float rad = p.w * 0.5f * sp.spread;
                const float s = (2.0f / 3.0f) * rad * rad * rad;
                const float tx = m0.x * dx + m0.y * dy + m0.z * dz;
                const float ty = m1.x * dx + m1.y * dy + m1.z * dz;
                const float tz = m2.x * dx + m2.y * dy + m2.z * dz;
                mono.x -= (m0.w - c3 * (ty * dz - tz * dy)) * law * s;
                mono.y -= (m1.w - c3 * (tz * dx - tx * dz)) * law * s;
                mono.z -= (m2.w - c3 * (tx * dy - ty * dx)) * law * s;
            }
        }
        const float isc = 1.0f / correction_scale(p, w, inv_h);
        for (int c = 0; c < kIn; ++c) X[(size_t)i * kIn + c] = f[c];
        if (Bout)
            for (int c = 0; c < 9; ++c) Bout[(size_t)i * 9 + c] = Bm[c];
        T[(size_t)i * kOut + 0] = (sum.x - mono.x) * sp.circ * isc;
        T[(size_t)i * kOut + 1] = (sum.y - mono.y) * sp.circ * isc;
        T[(size_t)i * kOut + 2] = (sum.z - mono.z) * sp.circ * isc;
        if (meta) {
            uint32_t pt[3];
            const uint32_t lp[3] = {pidx & 3u, (pidx >> 2) & 3u, pidx >> 4};
            for (int a = 0; a < 3; ++a) {
                pt[a] = 4 * bb[a] + lp[a];
                if (pt[a] > g.np[a] - 1) pt[a] = g.np[a] - 1;
            }
            meta[2 * (size_t)i] = (uint32_t)vg::point_offset(g, pt[0], pt[1], pt[2]);
            meta[2 * (size_t)i + 1] = ((uint32_t)layer << 28) | (uint32_t)off;
        }
        return;
    }
    for (int c = 0; c < kIn; ++c) X[(size_t)i * kIn + c] = 0.f;
    for (int c = 0; c < kOut; ++c) T[(size_t)i * kOut + c] = 0.f;
    if (Bout)
        for (int c = 0; c < 9; ++c) Bout[(size_t)i * 9 + c] = 0.f;
    if (meta) meta[2 * (size_t)i] = meta[2 * (size_t)i + 1] = 0xffffffffu;
    atomicAdd(failed, 1u);
}

}  // namespace

// ------------------------------------------------------------------ handle
struct vg_mlp {
    float *params = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    __bf16 *packed = nullptr;
    bool packed_ok = false;
    uint32_t step = 0;
};

namespace {
using namespace vg;

void compact_to_padded(const float *c, std::vector<float> &p) {
    p.assign(kPadParams, 0.f);
    size_t o = 0;
    for (int i = 0; i < kIn * kHid; ++i) p[kOffW1 + i] = c[o++];
    for (int i = 0; i < kHid; ++i) p[kOffB1 + i] = c[o++];
    for (int i = 0; i < kHid * kHid; ++i) p[kOffW2 + i] = c[o++];
    for (int i = 0; i < kHid; ++i) p[kOffB2 + i] = c[o++];
    for (int i = 0; i < kHid * kHid; ++i) p[kOffW3 + i] = c[o++];
    for (int i = 0; i < kHid; ++i) p[kOffB3 + i] = c[o++];
    for (int r = 0; r < kHid; ++r)
        for (int j = 0; j < kOut; ++j) p[kOffW4 + r * kOutPad + j] = c[o++];
    for (int j = 0; j < kOut; ++j) p[kOffB4 + j] = c[o++];
}

void padded_to_compact(const std::vector<float> &p, float *c) {
    size_t o = 0;
    for (int i = 0; i < kIn * kHid; ++i) c[o++] = p[kOffW1 + i];
    for (int i = 0; i < kHid; ++i) c[o++] = p[kOffB1 + i];
    for (int i = 0; i < kHid * kHid; ++i) c[o++] = p[kOffW2 + i];
    for (int i = 0; i < kHid; ++i) c[o++] = p[kOffB2 + i];
    for (int i = 0; i < kHid * kHid; ++i) c[o++] = p[kOffW3 + i];
    for (int i = 0; i < kHid; ++i) c[o++] = p[kOffB3 + i];
    for (int r = 0; r < kHid; ++r)
        for (int j = 0; j < kOut; ++j) c[o++] = p[kOffW4 + r * kOutPad + j];
    for (int j = 0; j < kOut; ++j) c[o++] = p[kOffB4 + j];
}

vg_status upload_params(vg_mlp *m, const float *compact) {
    std::vector<float> p;
    compact_to_padded(compact, p);
    VG_HIP(hipMemcpy(m->params, p.data(), kPadParams * sizeof(float), hipMemcpyHostToDevice));
    m->packed_ok = false;
    return VG_OK;
}

vg_status ensure_packed(vg_mlp *m, hipStream_t s) {
    if (m->packed_ok) return VG_OK;
    VG_TRY(launch("vg_mlp_pack", k_mlp_pack, dim3(ceil_div(kPackedElems, 256)), dim3(256), 0, s,
                  (const float *)m->params, m->packed));
    m->packed_ok = true;
    return VG_OK;
}

// Activations of one batch.
struct Acts {
    float *Z1, *A1, *Z2, *A2, *Z3, *A3, *Y, *dY, *dA, *partial, *loss, *split;
    void *base = nullptr;
};

vg_status acts_alloc(Acts *a, uint32_t rows, hipStream_t s) {
    const size_t h = (size_t)rows * kHid, y = (size_t)rows * kOutPad;
    const size_t nparts = ceil_div(rows, 256);
    const size_t nsplit = (size_t)kMaxSplit * 64 * 64;
    const size_t total = 7 * h + 2 * y + nparts + 1 + nsplit;
    VG_TRY(salloc(&a->base, total * sizeof(float), s));
    float *p = (float *)a->base;
    a->Z1 = p; p += h; a->A1 = p; p += h; a->Z2 = p; p += h; a->A2 = p; p += h;
    a->Z3 = p; p += h; a->A3 = p; p += h; a->dA = p; p += h;
    a->Y = p; p += y; a->dY = p; p += y;
    a->partial = p; p += nparts; a->loss = p; p += 1;
    a->split = p;
    return VG_OK;
}

vg_status forward(vg_mlp *m, const float *X, uint32_t rows, Acts &a, hipStream_t s) {
    const float *P = m->params;
    const size_t h = (size_t)rows * kHid;
    const dim3 eg(ceil_div(h, 256));
    VG_TRY(gemm_nn(rows, kHid, kIn, X, kIn, P + kOffW1, kHid, a.Z1, kHid, s));
    VG_TRY(launch("vg_mlp_gelu", k_bias_gelu, eg, dim3(256), 0, s, a.Z1, a.A1, P + kOffB1, h, kHid));
    VG_TRY(gemm_nn(rows, kHid, kHid, a.A1, kHid, P + kOffW2, kHid, a.Z2, kHid, s));
    VG_TRY(launch("vg_mlp_gelu", k_bias_gelu, eg, dim3(256), 0, s, a.Z2, a.A2, P + kOffB2, h, kHid));
    VG_TRY(gemm_nn(rows, kHid, kHid, a.A2, kHid, P + kOffW3, kHid, a.Z3, kHid, s));
    VG_TRY(launch("vg_mlp_gelu", k_bias_gelu, eg, dim3(256), 0, s, a.Z3, a.A3, P + kOffB3, h, kHid));
    VG_TRY(gemm_nn(rows, kOutPad, kHid, a.A3, kHid, P + kOffW4, kOutPad, a.Y, kOutPad, s));
    const size_t ny = (size_t)rows * kOutPad;
    VG_TRY(launch("vg_mlp_bias", k_bias_add, dim3(ceil_div(ny, 256)), dim3(256), 0, s, a.Y,
                  (const float *)(P + kOffB4), ny, kOutPad));
    return VG_OK;
}

// Loss gradient and parameter gradients into m->grad (padded layout).
vg_status backward(vg_mlp *m, const float *X, const float *Bm, const float *T, uint32_t rows,
                   Acts &a, hipStream_t s) {
    const float *P = m->params;
    float *G = m->grad;
    const size_t h = (size_t)rows * kHid;
    const dim3 eg(ceil_div(h, 256));
    const uint32_t nparts = ceil_div(rows, 256);
    VG_TRY(launch("vg_mlp_loss", k_loss_grad, dim3(nparts), dim3(256), 0, s, (const float *)a.Y, Bm,
                  T, a.dY, rows, 2.0f / (3.0f * (float)rows), a.partial));
    VG_TRY(launch("vg_mlp_loss", k_sum_partials, dim3(1), dim3(64), 0, s, (const float *)a.partial,
                  nparts, a.loss));
    // output layer
    VG_TRY(gemm_tn(kHid, kOutPad, rows, a.A3, kHid, a.dY, kOutPad, G + kOffW4, kOutPad, s, a.split));
    VG_TRY(colsum(a.dY, rows, kOutPad, G + kOffB4, a.split, s));
    VG_TRY(gemm_nt(rows, kHid, kOutPad, a.dY, kOutPad, P + kOffW4, kOutPad, a.dA, kHid, s));
    // hidden layers 3, 2, 1: dZ = dA * gelu'(Z), dW = A_prev^T dZ, dA_prev = dZ W^T
    struct L { float *Z; const float *Aprev; int lda_prev, kprev, offW, offB; };
    const L layers[3] = {{a.Z3, a.A2, kHid, kHid, kOffW3, kOffB3},
                         {a.Z2, a.A1, kHid, kHid, kOffW2, kOffB2},
                         {a.Z1, X, kIn, kIn, kOffW1, kOffB1}};
    for (int li = 0; li < 3; ++li) {
        const L &L_ = layers[li];
        VG_TRY(launch("vg_mlp_gelu_back", k_gelu_back, eg, dim3(256), 0, s, a.dA,
                      (const float *)L_.Z, h));
        VG_TRY(gemm_tn(L_.kprev, kHid, rows, L_.Aprev, L_.lda_prev, a.dA, kHid,
                       G + L_.offW, kHid, s, a.split));
        VG_TRY(colsum(a.dA, rows, kHid, G + L_.offB, a.split, s));
        if (li < 2) {
            // dA_prev = dZ W^T, written over the previous layer's Z (its
            // pre-activation is still needed: use A_prev's slot instead)
            float *dprev = (li == 0) ? a.A3 : a.A2;  // A3/A2 are no longer needed
            VG_TRY(gemm_nt(rows, kHid, kHid, a.dA, kHid, P + L_.offW, kHid, dprev, kHid, s));
            VG_HIP(hipMemcpyAsync(a.dA, dprev, h * sizeof(float), hipMemcpyDeviceToDevice, s));
        }
    }
    return VG_OK;
}

vg_status adam_step(vg_mlp *m, float lr, hipStream_t s) {
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.0f - std::pow(b1, (float)m->step), bc2 = 1.0f - std::pow(b2, (float)m->step);
    m->packed_ok = false;
    return launch("vg_mlp_adam", k_adam, dim3(ceil_div(kPadParams, 256)), dim3(256), 0, s, m->params,
                  (const float *)m->grad, m->m1, m->m2, kPadParams, lr, b1, b2, eps, bc1, bc2);
}

vg_status layer_inv_h(const TreeView &tv, float *inv_h) {
    for (uint32_t l = 0; l < kMaxLayers; ++l) inv_h[l] = 0.f;
    for (uint32_t l = 0; l < tv.depth; ++l) {
        const vg_grid_desc &gl = tv.layer[l];
        const float h = std::cbrt(gl.cell[0] * gl.cell[1] * gl.cell[2]);
        VG_REQUIRE(h > 0.f, "degenerate tree layer");
        inv_h[l] = 1.0f / h;
    }
    return VG_OK;
}

}  // namespace

using namespace vg;

VG_API vg_status vg_mlp_create(vg_mlp **out, uint64_t seed) {
    VG_REQUIRE(out, "null handle");
    *out = nullptr;
    vg_mlp *m = new vg_mlp();
    auto fail = [&](vg_status st) {
        vg_mlp_destroy(m);
        return st;
    };
    hipError_t e = hipMalloc((void **)&m->params, 4 * kPadParams * sizeof(float));
    if (e != hipSuccess) return fail(device_error(e, "hipMalloc mlp", __FILE__, __LINE__));
    m->grad = m->params + kPadParams;
    m->m1 = m->grad + kPadParams;
    m->m2 = m->m1 + kPadParams;
    e = hipMalloc((void **)&m->packed, kPackedElems * sizeof(__bf16));
    if (e != hipSuccess) return fail(device_error(e, "hipMalloc mlp", __FILE__, __LINE__));
    e = hipMemset(m->params, 0, 4 * kPadParams * sizeof(float));
    if (e != hipSuccess) return fail(device_error(e, "hipMemset mlp", __FILE__, __LINE__));
    // uniform +-sqrt(6 / (fan_in + fan_out)) weights from splitmix64(seed),
    // zero biases
    std::vector<float> c(VG_MLP_PARAMS, 0.f);
    uint64_t st = seed;
    auto uniform = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (float)((double)(z >> 11) * (1.0 / 9007199254740992.0)) * 2.0f - 1.0f;
    };
    size_t o = 0;
    const int fan[4][2] = {{kIn, kHid}, {kHid, kHid}, {kHid, kHid}, {kHid, kOut}};
    for (int l = 0; l < 4; ++l) {
        const float lim = std::sqrt(6.0f / (float)(fan[l][0] + fan[l][1]));
        for (int i = 0; i < fan[l][0] * fan[l][1]; ++i) c[o++] = lim * uniform();
        o += fan[l][1];
    }
    vg_status st2 = upload_params(m, c.data());
    if (st2 != VG_OK) return fail(st2);
    *out = m;
    return VG_OK;
}

VG_API vg_status vg_mlp_destroy(vg_mlp *m) {
    if (!m) return VG_OK;
    if (m->params) (void)hipFree(m->params);
    if (m->packed) (void)hipFree(m->packed);
    delete m;
    return VG_OK;
}

VG_API vg_status vg_mlp_set_params(vg_mlp *m, const float *params) {
    VG_REQUIRE(m && params, "null argument");
    VG_HIP(hipDeviceSynchronize());
    VG_TRY(upload_params(m, params));
    m->step = 0;
    VG_HIP(hipMemset(m->m1, 0, 2 * kPadParams * sizeof(float)));
    return VG_OK;
}

VG_API vg_status vg_mlp_get_params(vg_mlp *m, float *params) {
    VG_REQUIRE(m && params, "null argument");
    std::vector<float> p(kPadParams);
    VG_HIP(hipDeviceSynchronize());
    VG_HIP(hipMemcpy(p.data(), m->params, kPadParams * sizeof(float), hipMemcpyDeviceToHost));
    padded_to_compact(p, params);
    return VG_OK;
}

VG_API vg_status vg_mlp_forward(vg_mlp *m, const float *x, uint32_t n, float *y, vg_stream stream) {
    VG_REQUIRE(m, "null handle");
    if (n == 0) return VG_OK;
    VG_REQUIRE(x && y, "null buffer");
    hipStream_t s = (hipStream_t)stream;
    Acts a;
    VG_TRY(acts_alloc(&a, n, s));
    vg_status st = forward(m, x, n, a, s);
    if (st == VG_OK)
        st = launch("vg_mlp_copy", k_copy_cols, dim3(ceil_div((size_t)n * kOut, 256)), dim3(256), 0,
                    s, (const float *)a.Y, kOutPad, y, kOut, kOut, n);
    sfree(a.base, s);
    return st;
}

VG_API vg_status vg_mlp_gradients(vg_mlp *m, const float *x, const float *basis, const float *t,
                                  uint32_t n, float *grad, float *loss, vg_stream stream) {
    VG_REQUIRE(m && x && t && grad, "null argument");
    VG_REQUIRE(n > 0, "empty batch");
    hipStream_t s = (hipStream_t)stream;
    Acts a;
    VG_TRY(acts_alloc(&a, n, s));
    vg_status st = forward(m, x, n, a, s);
    if (st == VG_OK) st = backward(m, x, basis, t, n, a, s);
    float l = 0.f;
    std::vector<float> g(kPadParams);
    if (st == VG_OK) {
        hipError_t e = hipMemcpyAsync(&l, a.loss, sizeof(float), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(g.data(), m->grad, kPadParams * sizeof(float), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = device_error(e, "mlp gradient copy", __FILE__, __LINE__);
    }
    sfree(a.base, s);
    if (st != VG_OK) return st;
    padded_to_compact(g, grad);
    if (loss) *loss = l / (3.0f * (float)n);
    return VG_OK;
}

VG_API vg_status vg_mlp_train(vg_mlp *m, const float *x, const float *basis, const float *t,
                              uint32_t n, uint32_t batch, uint32_t steps, float lr,
                              float *final_loss, vg_stream stream) {
    VG_REQUIRE(m && x && t, "null argument");
    VG_REQUIRE(batch > 0 && batch <= n, "batch must be in [1, n]");
    hipStream_t s = (hipStream_t)stream;
    Acts a;
    VG_TRY(acts_alloc(&a, batch, s));
    vg_status st = VG_OK;
    const uint64_t span = (uint64_t)n - batch + 1;
    for (uint32_t it = 0; it < steps && st == VG_OK; ++it) {
        const uint64_t off = ((uint64_t)it * batch) % span;
        const float *xb = x + off * kIn, *tb = t + off * kOut;
        const float *bb = basis ? basis + off * 9 : nullptr;
        st = forward(m, xb, batch, a, s);
        if (st == VG_OK) st = backward(m, xb, bb, tb, batch, a, s);
        if (st == VG_OK) st = adam_step(m, lr, s);
    }
    float l = 0.f;
    if (st == VG_OK && final_loss) {
        hipError_t e = hipMemcpyAsync(&l, a.loss, sizeof(float), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = device_error(e, "mlp loss copy", __FILE__, __LINE__);
        *final_loss = l / (3.0f * (float)batch);
    }
    sfree(a.base, s);
    return st;
}

VG_API vg_status vg_tree_mlp_samples(const vg_tree *t, const vg_grid_desc *g, int near_k,
                                     uint32_t layer_mask, float spread, float circ, uint32_t n,
                                     uint64_t seed, float *x, float *basis, float *target,
                                     uint32_t *meta, uint32_t *n_failed, vg_stream stream) {
    TreeView tv;
    VG_TRY(tree_checked_view(t, &tv));
    VG_REQUIRE(g && x && target, "null argument");
    VG_REQUIRE(near_k >= 0 && near_k <= 16, "near_k must be in [0, 16]");
    for (int a = 0; a < 3; ++a)
        VG_REQUIRE(g->np[a] == tv.layer[0].np[a], "grid must have the tree's leaf geometry");
    VG_REQUIRE(tv.depth >= 2, "the tree has no cluster layers");
    VG_REQUIRE(!has_orphan_cells(tv),
               "the tree's grid leaves cells without a parent cluster; define the grid with "
               "power-of-two cells");
    const uint32_t valid = (tv.depth - 1 >= 32) ? 0xffffffffu : ((1u << (tv.depth - 1)) - 1u);
    SampleParams sp;
    sp.layer_mask = layer_mask & valid;
    VG_REQUIRE(sp.layer_mask != 0, "layer_mask selects no cluster layer of this tree");
    sp.k = near_k;
    for (int a = 0; a < 3; ++a) sp.nb[a] = (g->np[a] + 3) / 4;
    sp.nbricks = sp.nb[0] * sp.nb[1] * sp.nb[2];
    VG_TRY(layer_inv_h(tv, sp.inv_h));
    VG_REQUIRE(capacity(tv.layer[0]) < (1u << 28), "grid too large for the sample metadata");
    sp.spread = spread;
    sp.circ = circ;
    sp.n = n;
    sp.seed = seed;
    if (n == 0) {
        if (n_failed) *n_failed = 0;
        return VG_OK;
    }
    hipStream_t s = (hipStream_t)stream;
    uint32_t *d = nullptr;
    VG_TRY(salloc((void **)&d, sizeof(uint32_t), s));
    vg_status st = zero_fill(d, sizeof(uint32_t), s);
    if (st == VG_OK)
        st = launch("vg_mlp_samples", k_mlp_samples, dim3(ceil_div(n, 256)), dim3(256), 0, s, tv, *g,
                    sp, x, basis, target, meta, d);
    uint32_t h = 0;
    if (st == VG_OK) {
        hipError_t e = hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = device_error(e, "mlp sample count", __FILE__, __LINE__);
    }
    sfree(d, s);
    if (n_failed) *n_failed = h;
    return st;
}

VG_API vg_status vg_tree_velocity_grid_mlp(const vg_tree *t, const vg_grid_desc *g, vg_mlp *m,
                                           int near_k, uint32_t layer_mask, float spread,
                                           float circ, int rsqrt_policy, float *out,
                                           vg_stream stream) {
    VG_REQUIRE(m && out, "null argument");
    hipStream_t s = (hipStream_t)stream;
    // near-field treecode with the analytic first moments (the fallback),
    // then the learned residual of the clusters on the masked layers
    VG_TRY(mac_velocity_grid(t, g, near_k, 1, spread, circ, rsqrt_policy, out, s));
    TreeView tv;
    VG_TRY(tree_checked_view(t, &tv));
    if (tv.depth < 2) return VG_OK;
    MlpGridParams mp;
    mp.k = near_k;
    mp.layer_mask = layer_mask;
    for (int a = 0; a < 3; ++a) mp.nb[a] = (g->np[a] + 3) / 4;
    mp.nbricks = mp.nb[0] * mp.nb[1] * mp.nb[2];
    VG_TRY(layer_inv_h(tv, mp.inv_h));
    if ((layer_mask & ((1u << (tv.depth - 1)) - 1u)) == 0) return VG_OK;
    VG_TRY(ensure_packed(m, s));
    return launch("vg_tree_velocity_grid_mlp", k_mlp_grid, dim3(ceil_div(mp.nbricks, 4)), dim3(256), 0,
                  s, tv, *g, mp, (const __bf16 *)m->packed, (const float *)m->params, out);
}
