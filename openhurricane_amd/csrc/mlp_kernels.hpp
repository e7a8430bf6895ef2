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
    f32x16 acc = {};
    for (int k0 = 0; k0 < g.K; k0 += kTK) {
        // ---- stage A and B slices (256 threads x 8 elements each) ----
        {
            bf16_t va[8], vb[8];
            if (A_KC) {
                const int r = tid >> 2, kk = (tid & 3) * 8;
                const int m = m0 + r;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int k = k0 + kk + e;
                    va[e] = (m < g.M && k < g.K) ? g.A[(int64_t)m * g.lda + k] : (bf16_t)0;
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) sA[r][kk + e] = va[e];
            } else {
                const int kk = tid >> 3, r = (tid & 7) * 8;
                const int k = k0 + kk;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int m = m0 + r + e;
                    va[e] = (m < g.M && k < g.K) ? g.A[(int64_t)k * g.lda + m] : (bf16_t)0;
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
                    vb[e] = (n < g.N && k < g.K) ? g.B[(int64_t)n * g.ldb + k] : (bf16_t)0;
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) sB[r][kk + e] = vb[e];
            } else {
                const int kk = tid >> 3, r = (tid & 7) * 8;
                const int k = k0 + kk;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int n = n0 + r + e;
                    vb[e] = (n < g.N && k < g.K) ? g.B[(int64_t)k * g.ldb + n] : (bf16_t)0;
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
            g.C32[o] = x;
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
    const float *in_mean, *in_rstd, *zlo, *zhi;  // input normalisation and OOD box
    const float *out_mean, *out_std;
    const double *elemW;        // [ns][ne]: atoms of element e in species k / W_k
    double gate_tol;            // max relative element-mass change
    double T_low, T_high;
    int *fail_list;             // indices of cells left to the integrator
    unsigned int *fail_count;
    int mode;                   // 0: advance and gate; 1: write raw outputs to out32 only
    float *out32;               // mode 1: [n][nout] denormalised increments
};

// Hidden layer: out[TM][W] = gelu(in[TM][K] W^T + b), 4 wavefronts each
// producing TM rows x W/4 columns with 32x32x16 BF16 MFMA.
template <int TM, int WD, int K, int LDI, int LDO>
__device__ __forceinline__ void mlp_layer(const bf16_t *sin, bf16_t *sout, const bf16_t *Wl,
                                          const float *bl) {
    constexpr int RT = TM / 32, CT = WD / 128;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    f32x16 acc[RT][CT];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[r][c] = f32x16{};
    const int i = lane & 31, kh = 8 * (lane >> 5);
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
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        const int col = wave * (CT * 32) + c * 32 + i;
        const float bias = bl[col];
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = r * 32 + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
                sout[row * LDO + col] = f2bf(gelu(acc[r][c][v] + bias));
            }
    }
}

// One block of 256 threads advances TM cells: features (T, log p, log Y) ->
// normalised BF16 -> hidden layers -> increments (dT, dlog Y) -> new state,
// gated on the dt match, the input box (OOD) and the element-mass change.
template <int TM, int WD, int KIN, int NOUTP>
__global__ __launch_bounds__(256) void mlp_infer(const InferArgs a) {
    constexpr int LD = WD + 8;  // LDS row stride (bf16) of the activation buffers
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
    mlp_layer<TM, WD, KIN, LD, LD>(buf[0], buf[1], a.W + a.woff[0], a.b + a.boff[0]);
    __syncthreads();
    int cur = 1;
    for (int l = 1; l < a.depth; ++l) {
        mlp_layer<TM, WD, WD, LD, LD>(buf[cur], buf[cur ^ 1], a.W + a.woff[l], a.b + a.boff[l]);
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
        for (int k = 0; k < nout; ++k) a.out32[cell * nout + k] = o[k] * a.out_std[k] + a.out_mean[k];
        return;
    }
    bool ok = sfail[tid] == 0;
    ok = ok && fabs(a.dt[cell] * a.dtFactor - a.dt_train) <= 1e-9 * a.dt_train;
    const double T0 = a.T[cell];
    const double T1 = T0 + (double)(o[0] * a.out_std[0] + a.out_mean[0]);
    ok = ok && T1 >= a.T_low && T1 <= a.T_high;
    // mass fractions from the log increments (kept in the free activation
    // buffer as [TM][64] doubles), then renormalised
    double *y1 = (double *)buf[cur] + tid * 64;
    static_assert(TM * 64 * 8 <= TM * LD * 2, "mass-fraction tile must fit the free buffer");
    double ys = 0.0;
    for (int k = 0; k < a.ns; ++k) {
        const double y0 = fmax(a.Y[(int64_t)k * a.ldY + cell], kYFloor);
        const double v = y0 * exp((double)(o[k + 1] * a.out_std[k + 1] + a.out_mean[k + 1]));
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
