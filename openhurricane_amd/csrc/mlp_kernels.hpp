// mlp_kernels.hpp -- BF16 MFMA kernels of the chemistry surrogate (an MLP
// with GELU hidden layers): a tiled GEMM with fused epilogues for training,
// element-wise helpers (normalisation, loss gradient, bias gradient, Adam),
// and the fused inference + gate kernel that advances the cell states.
//
// Matrix-core layout (gfx950, v_mfma_f32_32x32x16_bf16): lane l holds
//   A[i][8*(l/32) + 0..7] with i = l % 32 (a 32x16 tile, K contiguous),
//   B[8*(l/32) + 0..7][j] with j = l % 32 (a 16x32 tile),
//   D[8*(v/4) + 4*(l/32) + v%4][l % 32] in accumulator v = 0..15.
// Activations are row-major [cells][features], weights [out][in] (so both
// GEMM operands of the forward pass have the reduction dimension
// contiguous), all BF16 with FP32 accumulation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ohcm {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
typedef unsigned short bf16_t;  // storage type (raw bits)

__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }
// round-to-nearest-even FP32 -> BF16 (NaN kept quiet)
__device__ __forceinline__ bf16_t f2bf(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (bf16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}

// GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))
__device__ __forceinline__ float gelu(float x) {
    const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
    const float t = 1.0f - 2.0f / (1.0f + __expf(2.0f * u));
    return 0.5f * x * (1.0f + t);
}
__device__ __forceinline__ float gelu_grad(float x) {
    const float x2 = x * x;
    const float u = 0.7978845608028654f * (x + 0.044715f * x2 * x);
    const float t = 1.0f - 2.0f / (1.0f + __expf(2.0f * u));
    const float du = 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * x2);
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * du;
}

enum Epi : int {
    EPI_F32 = 0,        // C32 = acc
    EPI_BIAS_GELU = 1,  // Z32 = acc + bias, C16 = bf16(gelu(Z))
    EPI_BIAS_F32 = 2,   // C32 = acc + bias
    EPI_GELU_BWD = 3,   // C16 = bf16(acc * gelu'(Z32))
};

struct GemmArgs {
    int M, N, K;
    const bf16_t *A;  // A(m,k) = A_KC ? A[m*lda + k] : A[k*lda + m]
    int64_t lda;
    const bf16_t *B;  // B(k,n) = B_KC ? B[n*ldb + k] : B[k*ldb + n]
    int64_t ldb;
    float *C32;       // EPI_F32 / EPI_BIAS_F32 output, EPI_BIAS_GELU pre-activation
    bf16_t *C16;      // EPI_BIAS_GELU / EPI_GELU_BWD output
    int64_t ldc;
    const float *bias;  // [N]
    const float *Z;     // EPI_GELU_BWD: saved pre-activation, ld ldc
    // split-K (EPI_F32 only): block z reduces k in [z kc, min(K, (z+1) kc))
    // into C32 + z zstride; splitk_sum then adds the partial tiles in order
    int kc;
    int64_t zstride;
};

constexpr int kTM = 64, kTN = 64, kTK = 32, kLdsK = kTK + 8;  // +8: LDS row padding

// C[M x N] = A[M x K] B[K x N] with a fused epilogue.  A 256-thread block
// computes a 64 x 64 tile (each wavefront a 32 x 32 MFMA tile), staging
// 64 x 32 slices of A and B through LDS as [row][k]; operands whose rows are
// not K-contiguous are transposed on the way in.  Edges are zero-padded.
template <bool A_KC, bool B_KC, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16(const GemmArgs g) {
    __shared__ bf16_t sA[kTM][kLdsK], sB[kTN][kLdsK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int kb = (int)blockIdx.z * g.kc, ke = min(g.K, kb + g.kc);
    f32x16 acc = {};
    for (int k0 = kb; k0 < ke; k0 += kTK) {
        // ---- stage A and B slices (256 threads x 8 elements each) ----
        {
            bf16_t va[8], vb[8];
            if (A_KC) {
                const int r = tid >> 2, kk = (tid & 3) * 8;
                const int m = m0 + r;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int k = k0 + kk + e;
                    va[e] = (m < g.M && k < ke) ? g.A[(int64_t)m * g.lda + k] : (bf16_t)0;
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) sA[r][kk + e] = va[e];
            } else {
                const int kk = tid >> 3, r = (tid & 7) * 8;
                const int k = k0 + kk;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int m = m0 + r + e;
                    va[e] = (m < g.M && k < ke) ? g.A[(int64_t)k * g.lda + m] : (bf16_t)0;
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) sA[r + e][kk] = va[e];
            }
            if (B_KC) {
                const int r = tid >> 2, kk = (tid & 3) * 8;
                const int n = n0 + r;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int k = k0 + kk + e;
                    vb[e] = (n < g.N && k < ke) ? g.B[(int64_t)n * g.ldb + k] : (bf16_t)0;
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) sB[r][kk + e] = vb[e];
            } else {
                const int kk = tid >> 3, r = (tid & 7) * 8;
                const int k = k0 + kk;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int n = n0 + r + e;
                    vb[e] = (n < g.N && k < ke) ? g.B[(int64_t)k * g.ldb + n] : (bf16_t)0;
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) sB[r + e][kk] = vb[e];
            }
        }
        __syncthreads();
        // ---- two 32x32x16 MFMA steps per slice ----
#pragma unroll
        for (int ks = 0; ks < kTK; ks += 16) {
            const int i = lane & 31, kb = ks + 8 * (lane >> 5);
            bf16x8 a, b;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                a[e] = __builtin_bit_cast(__bf16, sA[wm + i][kb + e]);
                b[e] = __builtin_bit_cast(__bf16, sB[wn + i][kb + e]);
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // ---- epilogue ----
    const int j = n0 + wn + (lane & 31);
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int i = m0 + wm + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
        if (i >= g.M || j >= g.N) continue;
        const int64_t o = (int64_t)i * g.ldc + j;
        const float x = acc[v];
        if (EPI == EPI_F32) {
            g.C32[o + (int64_t)blockIdx.z * g.zstride] = x;
        } else if (EPI == EPI_BIAS_F32) {
            g.C32[o] = x + g.bias[j];
        } else if (EPI == EPI_BIAS_GELU) {
            const float z = x + g.bias[j];
            if (g.C32) g.C32[o] = z;
            g.C16[o] = f2bf(gelu(z));
        } else {
            g.C16[o] = f2bf(x * gelu_grad(g.Z[o]));
        }
    }
}

// C[i][j] = sum over z = 0..S-1, in that order, of the split-K partial tiles.
__global__ void splitk_sum(int S, int M, int N, int64_t ldc, int64_t zstride, const float *part,
                           float *C) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)M * N) return;
    const int64_t o = (e / N) * ldc + e % N;
    float v = part[o];
    for (int z = 1; z < S; ++z) v += part[z * zstride + o];
    C[o] = v;
}

// ---- element-wise training helpers ----

// X16[i][k] = bf16((X[i][k] - mean[k]) / std[k]) for k < nin, 0 for padding
__global__ void normalize_in(int64_t n, int nin, int kpad, const float *X, const float *mean,
                             const float *rstd, bf16_t *X16) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * kpad) return;
    const int64_t i = e / kpad;
    const int k = (int)(e % kpad);
    X16[e] = k < nin ? f2bf((X[i * nin + k] - mean[k]) * rstd[k]) : (bf16_t)0;
}

// Weighted mean-square loss on normalised targets t = (Yt - mean) / std:
// loss = sum w_k (out - t)^2 / n, dOut = 2 w (out - t) / n (BF16 for the
// backward GEMMs); per-block partial losses into part[].
__global__ void loss_grad(int64_t n, int nout, const float *out, const float *Yt,
                          const float *mean, const float *rstd, const float *w, bf16_t *dOut,
                          float *part) {
    __shared__ float red[256];
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (e < n * nout) {
        const int k = (int)(e % nout);
        const float r = out[e] - (Yt[e] - mean[k]) * rstd[k];
        l = w[k] * r * r / (float)n;
        dOut[e] = f2bf(2.0f * w[k] * r / (float)n);
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Column sums db[j] = sum_i dZ[i][j] in two deterministic passes: block
// (x, y) sums rows [y*kRowChunk, ...) of columns [64x, 64x+64) into
// part[y][j]; colsum_finish adds the partials in chunk order.
constexpr int kRowChunk = 256;
__global__ __launch_bounds__(256) void colsum_partial(int64_t n, int N, int64_t ld,
                                                      const bf16_t *dZ, float *part) {
    __shared__ float red[4][64];
    const int c = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + c;
    const int64_t r0 = (int64_t)blockIdx.y * kRowChunk;
    float s = 0.0f;
    if (j < N)
        for (int64_t i = r0 + q; i < r0 + kRowChunk && i < n; i += 4) s += bf2f(dZ[i * ld + j]);
    red[q][c] = s;
    __syncthreads();
    if (q == 0 && j < N) part[(int64_t)blockIdx.y * N + j] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
}
__global__ void colsum_finish(int nchunk, int N, const float *part, float *db) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float s = 0.0f;
    for (int y = 0; y < nchunk; ++y) s += part[(int64_t)y * N + j];
    db[j] = s;
}

// Adam on the flat FP32 master parameters (c1 = 1/(1-b1^t), c2 = 1/(1-b2^t)),
// refreshing their BF16 copy p16 (same layout).
__global__ void adam_step(int64_t np, float *p, const float *g, float *m, float *v, float lr,
                          float b1, float b2, float eps, float c1, float c2, bf16_t *p16) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= np) return;
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.0f - b1) * gi;
    const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    const float pi = p[e] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    p[e] = pi;
    p16[e] = f2bf(pi);
}

__global__ void to_bf16(int64_t n, const float *x, bf16_t *y) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) y[e] = f2bf(x[e]);
}

__global__ void sum_partials(int nb, const float *part, float *out) {
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

}  // namespace ohcm

namespace ohcm {

// ---- fused inference + gate ------------------------------------------------

constexpr double kYFloor = 1e-20;  // mass-fraction floor of the log features
constexpr double kAgreeY = 1e-4;   // species the two-network agreement check covers

typedef __attribute__((__vector_size__(8 * sizeof(int)))) int i32x8;

// OCP FP8 E4M3 (bias 7, no infinities, 0x7f NaN, largest finite 448): round to
// nearest even, saturating.
__host__ __device__ inline uint8_t f2e4m3(float x) {
    const uint8_t sign = x < 0.0f ? 0x80 : 0x00;
    const float a = fabsf(x);
    if (!(a == a)) return 0x7f;
    if (a >= 448.0f) return sign | 0x7e;
    if (a < 0.015625f) {  // subnormal: m * 2^-9
        const int m = (int)rintf(a * 512.0f);
        return sign | (uint8_t)m;  // m == 8 is the smallest normal, 0x08
    }
    int ex;
    const float fr = frexpf(a, &ex);  // a = fr 2^ex, fr in [0.5, 1)
    int e = ex - 1;
    int m = (int)rintf((2.0f * fr - 1.0f) * 8.0f);
    if (m == 8) {
        m = 0;
        ++e;
    }
    const int be = e + 7;
    if (be > 15 || (be == 15 && m == 7)) return sign | 0x7e;
    return sign | (uint8_t)((be << 3) | m);
}
// The same rounding with the hardware conversion (v_cvt_pk_fp8_f32 is OCP
// E4M3 with round to nearest even on this target, NaN above 448, hence the
// clamp; a negative value that rounds to zero gives +0 instead of -0).
__device__ __forceinline__ uint8_t f2e4m3_hw(float x) {
    const float c = fminf(fmaxf(x, -448.0f), 448.0f);
    return (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(c, c, 0, false) & 0xff);
}
__host__ __device__ inline float e4m32f(uint8_t b) {
    const int e = (b >> 3) & 15, m = b & 7;
    const float v = e == 0 ? ldexpf((float)m, -9) : ldexpf(1.0f + 0.125f * (float)m, e - 7);
    return (b & 0x80) ? -v : v;
}

struct InferArgs {
    int64_t n;
    int ns, ne;                 // species, elements
    double *T, *Y;              // state, advanced in place where the gate passes
    const double *p;
    int64_t ldY;
    const double *dt;
    double dtFactor, dt_train;  // cells whose dt*dtFactor differs from dt_train fall back
    const bf16_t *W;            // packed BF16 weights, layer l at woff[l], [rows][K] padded
    const float *b;             // biases, layer l at boff[l] (padded rows zero)
    int woff[8], boff[8];
    int depth;                  // hidden layers
    const uint8_t *W8;          // FP8 copies of the hidden-to-hidden layers l = 1..depth-1
    const float *s8;            // their per-row scales
    int w8off[8], s8off[8];
    const float *in_mean, *in_rstd, *zlo, *zhi;  // input normalisation and OOD box
    const float *out_mean, *out_std;
    const float *delta;         // symmetric-log target scales (nullptr: plain increments)
    const float *peer;          // [n][nout] increments of a second network (nullptr: none)
    double peer_tol;            // largest accepted disagreement with the peer
    const double *elemW;        // [ns][ne]: atoms of element e in species k / W_k
    double gate_tol;            // max relative element-mass change
    double T_low, T_high;
    int *fail_list;             // indices of cells left to the integrator
    unsigned int *fail_count;
    int mode;                   // 0: advance and gate; 1: write increments to out32 only
    float *out32;               // mode 1: [n][nout] increments (dT, dln Y)
};

// network output k (denormalised) -> increment: identity, or the inverse of
// z = sign(d) ln(1 + |d| / delta_k)
__device__ __forceinline__ float out_increment(const InferArgs &a, const float *o, int k) {
    const float z = o[k] * a.out_std[k] + a.out_mean[k];
    if (!a.delta) return z;
    return copysignf(expm1f(fabsf(z)), z) * a.delta[k];
}

// Hidden layer: out[TM][W] = gelu(in[TM][K] W^T + b), 4 wavefronts each
// producing TM rows x W/4 columns.  IN8: FP8 activations and weights on the
// 32x32x64 f8f6f4 MFMA (unit block scales; per-row weight scale applied in
// the epilogue), else BF16 on 32x32x16.  OUT8: the epilogue writes FP8.
// Strides LDI / LDO count elements of the buffer's type.
template <int TM, int WD, int K, int LDI, int LDO, bool IN8, bool OUT8>
__device__ __forceinline__ void mlp_layer(const void *sin_, void *sout_, const void *Wl_,
                                          const float *ws, const float *bl) {
    constexpr int RT = TM / 32, CT = WD / 128;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    f32x16 acc[RT][CT];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[r][c] = f32x16{};
    const int i = lane & 31, h = lane >> 5;
    if constexpr (IN8) {
        const uint8_t *sin = (const uint8_t *)sin_, *Wl = (const uint8_t *)Wl_;
#pragma unroll
        for (int k0 = 0; k0 < K; k0 += 64) {
            i32x8 a[RT], b[CT];
#pragma unroll
            for (int r = 0; r < RT; ++r) a[r] = *(const i32x8 *)(sin + (r * 32 + i) * LDI + k0 + 32 * h);
#pragma unroll
            for (int c = 0; c < CT; ++c) {
                const int nn = wave * (CT * 32) + c * 32 + i;
                b[c] = *(const i32x8 *)(Wl + (int64_t)nn * K + k0 + 32 * h);
            }
#pragma unroll
            for (int r = 0; r < RT; ++r)
#pragma unroll
                for (int c = 0; c < CT; ++c)
                    acc[r][c] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                        a[r], b[c], acc[r][c], 0, 0, 0, 127, 0, 127);
        }
    } else {
        const bf16_t *sin = (const bf16_t *)sin_, *Wl = (const bf16_t *)Wl_;
        const int kh = 8 * h;
#pragma unroll
        for (int k0 = 0; k0 < K; k0 += 16) {
            bf16x8 a[RT], b[CT];
#pragma unroll
            for (int r = 0; r < RT; ++r) a[r] = *(const bf16x8 *)(sin + (r * 32 + i) * LDI + k0 + kh);
#pragma unroll
            for (int c = 0; c < CT; ++c) {
                const int nn = wave * (CT * 32) + c * 32 + i;
                b[c] = *(const bf16x8 *)(Wl + (int64_t)nn * K + k0 + kh);
            }
#pragma unroll
            for (int r = 0; r < RT; ++r)
#pragma unroll
                for (int c = 0; c < CT; ++c)
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[r], b[c], acc[r][c], 0, 0, 0);
        }
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        const int col = wave * (CT * 32) + c * 32 + i;
        const float bias = bl[col];
        const float scale = IN8 ? ws[col] : 1.0f;
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = r * 32 + 8 * (v / 4) + 4 * h + (v % 4);
                const float g = gelu(acc[r][c][v] * scale + bias);
                if constexpr (OUT8) ((uint8_t *)sout_)[row * LDO + col] = f2e4m3_hw(g);
                else ((bf16_t *)sout_)[row * LDO + col] = f2bf(g);
            }
    }
}

// One block of 256 threads advances TM cells: features (T, log p, log Y) ->
// normalised BF16 -> hidden layers -> increments (dT, dlog Y) -> new state,
// gated on the dt match, the input box (OOD), the element-mass change and,
// with a peer network, the two networks' agreement (the state then takes the
// mean of both increments).  FP8: the hidden-to-hidden layers run in FP8.
template <int TM, int WD, int KIN, int NOUTP, bool FP8>
__global__ __launch_bounds__(256) void mlp_infer(const InferArgs a) {
    constexpr int LD = WD + 8;   // LDS row stride (bf16) of the activation buffers
    constexpr int LD8 = WD + 16; // the same buffers holding FP8 rows (bytes)
    static_assert(TM * LD8 <= TM * LD * 2, "FP8 rows must fit the buffer");
    __shared__ __attribute__((aligned(16))) bf16_t buf[2][TM * LD];
    __shared__ int sfail[TM];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t c0 = (int64_t)blockIdx.x * TM;
    const int nin = a.ns + 2, nout = a.ns + 1;
    if (tid < TM) sfail[tid] = 0;
    __syncthreads();
    // ---- features, normalised, into buf[0] as [TM][KIN] (stride LD) ----
    for (int e = tid; e < TM * KIN; e += 256) {
        const int k = e / TM, r = e % TM;
        const int64_t cell = c0 + r;
        float z = 0.0f;
        if (k < nin && cell < a.n) {
            double x;
            if (k == 0) x = a.T[cell];
            else if (k == 1) x = log(a.p[cell]);
            else x = log(fmax(a.Y[(int64_t)(k - 2) * a.ldY + cell], kYFloor));
            z = ((float)x - a.in_mean[k]) * a.in_rstd[k];
            if (!(z >= a.zlo[k] && z <= a.zhi[k])) atomicOr(&sfail[r], 1);
        }
        buf[0][r * LD + k] = f2bf(z);
    }
    __syncthreads();
    // ---- hidden layers (ping-pong between the two buffers) ----
    const bool l1_fp8 = FP8 && a.depth > 1;
    if (l1_fp8) mlp_layer<TM, WD, KIN, LD, LD8, false, true>(buf[0], buf[1], a.W + a.woff[0], nullptr, a.b + a.boff[0]);
    else mlp_layer<TM, WD, KIN, LD, LD, false, false>(buf[0], buf[1], a.W + a.woff[0], nullptr, a.b + a.boff[0]);
    __syncthreads();
    int cur = 1;
    for (int l = 1; l < a.depth; ++l) {
        if constexpr (FP8) {
            if (l + 1 < a.depth)
                mlp_layer<TM, WD, WD, LD8, LD8, true, true>(buf[cur], buf[cur ^ 1], a.W8 + a.w8off[l],
                                                             a.s8 + a.s8off[l], a.b + a.boff[l]);
            else
                mlp_layer<TM, WD, WD, LD8, LD, true, false>(buf[cur], buf[cur ^ 1], a.W8 + a.w8off[l],
                                                             a.s8 + a.s8off[l], a.b + a.boff[l]);
        } else {
            mlp_layer<TM, WD, WD, LD, LD, false, false>(buf[cur], buf[cur ^ 1], a.W + a.woff[l],
                                                         nullptr, a.b + a.boff[l]);
        }
        __syncthreads();
        cur ^= 1;
    }
    // ---- output layer into FP32 (aliasing the free buffer) ----
    float *sout = (float *)buf[cur ^ 1];
    constexpr int LDF = NOUTP + 1;
    static_assert(TM * LDF * 4 <= TM * LD * 2, "output tile must fit the free buffer");
    {
        constexpr int RT = TM / 32, CTO = NOUTP / 32;
        const int i = lane & 31, kh = 8 * (lane >> 5);
        if (wave < RT * CTO) {
            const int r = wave / CTO, c = wave % CTO;
            const bf16_t *Wl = a.W + a.woff[a.depth];
            f32x16 acc = {};
#pragma unroll
            for (int k0 = 0; k0 < WD; k0 += 16) {
                const bf16x8 av = *(const bf16x8 *)(buf[cur] + (r * 32 + i) * LD + k0 + kh);
                const bf16x8 bv = *(const bf16x8 *)(Wl + (int64_t)(c * 32 + i) * WD + k0 + kh);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
            }
            const float bias = a.b[a.boff[a.depth] + c * 32 + i];
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = r * 32 + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
                sout[row * LDF + c * 32 + i] = acc[v] + bias;
            }
        }
    }
    __syncthreads();
    // ---- per cell: increments, new state, gate ----
    if (tid >= TM) return;
    const int64_t cell = c0 + tid;
    if (cell >= a.n) return;
    const float *o = sout + tid * LDF;
    if (a.mode == 1) {
        for (int k = 0; k < nout; ++k) a.out32[cell * nout + k] = out_increment(a, o, k);
        return;
    }
    bool ok = sfail[tid] == 0;
    ok = ok && fabs(a.dt[cell] * a.dtFactor - a.dt_train) <= 1e-9 * a.dt_train;
    const double T0 = a.T[cell];
    const float *q = a.peer ? a.peer + cell * nout : nullptr;
    double dT = out_increment(a, o, 0);
    if (q) {
        ok = ok && fabs(dT - (double)q[0]) <= a.peer_tol * T0;
        dT = 0.5 * (dT + (double)q[0]);
    }
    const double T1 = T0 + dT;
    ok = ok && T1 >= a.T_low && T1 <= a.T_high;
    // mass fractions from the log increments (kept in the free activation
    // buffer as [TM][64] doubles), then renormalised
    double *y1 = (double *)buf[cur] + tid * 64;
    static_assert(TM * 64 * 8 <= TM * LD * 2, "mass-fraction tile must fit the free buffer");
    double ys = 0.0;
    for (int k = 0; k < a.ns; ++k) {
        const double yk = a.Y[(int64_t)k * a.ldY + cell];
        const double y0 = fmax(yk, kYFloor);
        const double ds = out_increment(a, o, k + 1);
        const double dq = q ? (double)q[k + 1] : ds;
        const double d = q ? 0.5 * (ds + dq) : ds;
        const double v = y0 * exp(d);
        // the networks must agree on every species above kAgreeY before or
        // after the step (a decade below the 1e-3 "major" threshold, so
        // that a species predicted just under it is still checked)
        if (q && (yk >= kAgreeY || v >= kAgreeY)) ok = ok && fabs(ds - dq) <= a.peer_tol;
        y1[k] = v;
        ys += v;
    }
    ok = ok && ys > 0.0 && isfinite(ys);
    const double rys = 1.0 / ys;
    // element-mass change (chemistry conserves element moles per unit mass)
    double emax = 0.0;
    for (int e = 0; e < a.ne && ok; ++e) {
        double e0 = 0.0, e1 = 0.0;
        for (int k = 0; k < a.ns; ++k) {
            const double w = a.elemW[k * a.ne + e];
            e0 += w * a.Y[(int64_t)k * a.ldY + cell];
            e1 += w * y1[k] * rys;
        }
        if (e0 > 1e-12) emax = fmax(emax, fabs(e1 - e0) / e0);
    }
    ok = ok && emax <= a.gate_tol;
    if (ok) {
        a.T[cell] = T1;
        for (int k = 0; k < a.ns; ++k) a.Y[(int64_t)k * a.ldY + cell] = y1[k] * rys;
    } else {
        const unsigned int slot = atomicAdd(a.fail_count, 1u);
        a.fail_list[slot] = (int)cell;
    }
}

}  // namespace ohcm
