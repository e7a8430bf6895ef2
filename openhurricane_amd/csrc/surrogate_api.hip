// surrogate_api.hip -- C-ABI of the chemistry surrogate: an MLP trained on
// the integrator's own state pairs with the BF16 MFMA kernels of
// mlp_kernels.hpp, and the gated advance that falls back to the integrator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "ctx_internal.hpp"
#include "mlp_kernels.hpp"

using namespace ohci;
using ohcm::bf16_t;

namespace {

// target k: increment d (dT, dln Y) or, with delta, z = sign(d) ln(1 + |d| / delta_k)
__device__ __forceinline__ float target(double d, const float *delta, int k) {
    if (!delta) return (float)d;
    return (float)copysign(log1p(fabs(d) / (double)delta[k]), d);
}

__global__ void features_kernel(int64_t n, int ns, const double *T0, const double *p,
                                const double *Y0, int64_t ldY, const double *T1,
                                const double *Y1, const float *delta, float *X, float *Yt) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int nin = ns + 2, nout = ns + 1;
    X[c * nin + 0] = (float)T0[c];
    X[c * nin + 1] = (float)log(p[c]);
    for (int k = 0; k < ns; ++k) {
        const double l0 = log(fmax(Y0[(int64_t)k * ldY + c], ohcm::kYFloor));
        X[c * nin + 2 + k] = (float)l0;
        if (Yt)
            Yt[c * nout + 1 + k] =
                target(log(fmax(Y1[(int64_t)k * ldY + c], ohcm::kYFloor)) - l0, delta, 1 + k);
    }
    if (Yt) Yt[c * nout] = target(T1[c] - T0[c], delta, 0);
}

// packed inference copy: layer rows x K (zero padded), BF16
__global__ void pack_layer(int rows, int cols, int prow, int pcol, const float *W, bf16_t *P) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)prow * pcol) return;
    const int r = (int)(e / pcol), k = (int)(e % pcol);
    P[e] = (r < rows && k < cols) ? ohcm::f2bf(W[(int64_t)r * cols + k]) : (bf16_t)0;
}
__global__ void pack_bias(int rows, int prow, const float *b, float *P) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < prow) P[e] = e < rows ? b[e] : 0.0f;
}

// FP8 copy of a [rows][cols] layer: one block per row, scale = max|w| / 448
__global__ __launch_bounds__(256) void pack_layer_fp8(int cols, const float *W, uint8_t *P,
                                                      float *scale) {
    __shared__ float red[256];
    const int r = blockIdx.x, t = threadIdx.x;
    float m = 0.0f;
    for (int k = t; k < cols; k += 256) m = fmaxf(m, fabsf(W[(int64_t)r * cols + k]));
    red[t] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] = fmaxf(red[t], red[t + o]);
        __syncthreads();
    }
    const float sc = red[0] > 0.0f ? red[0] / 448.0f : 1.0f;
    for (int k = t; k < cols; k += 256) P[(int64_t)r * cols + k] = ohcm::f2e4m3(W[(int64_t)r * cols + k] / sc);
    if (t == 0) scale[r] = sc;
}

inline unsigned nblk(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

__global__ void zero_u32(unsigned int *p) { *p = 0u; }
inline int rup(int x, int m) { return (x + m - 1) / m * m; }

}  // namespace

struct ohc_surrogate {
    ohc_ctx *ctx = nullptr;
    int ns = 0, nin = 0, nout = 0, width = 0, depth = 0;
    int kin = 0, noutp = 0;  // padded input width and output rows of the inference copy
    int64_t np = 0;          // parameters
    std::vector<int64_t> w_off, b_off;  // flat FP32 layout: per layer W [out][in], b [out]
    std::vector<int> rows, cols;
    float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
    bf16_t *d_p16 = nullptr;
    // inference copy
    bf16_t *d_W = nullptr;
    float *d_b = nullptr;
    int woff[8] = {}, boff[8] = {};
    // FP8 copies of the hidden-to-hidden layers (inference precision 1)
    uint8_t *d_W8 = nullptr;
    float *d_s8 = nullptr;
    int w8off[8] = {}, s8off[8] = {};
    int precision = 0;
    bool dirty = true;
    // symmetric-log target scales (device, nout floats; unused when !has_delta)
    float *d_delta = nullptr;
    bool has_delta = false;
    // second network whose agreement the gate requires
    ohc_surrogate *peer = nullptr;
    double peer_tol = 0.0;
    float *d_peer_out = nullptr;
    int64_t peer_cap = 0;
    // normalisation, OOD box, loss weights (device, float)
    float *d_norm = nullptr;  // in_mean, in_rstd, zlo, zhi [nin]; out_mean, out_std, out_rstd, w [nout]
    double dt_train = 0.0;
    double *d_elemW = nullptr;
    int ne = 0;
    int step = 0;
    // batch workspace
    int64_t cap = 0;
    void *d_ws = nullptr;
    bf16_t *X16 = nullptr, *dOut = nullptr, *dZa = nullptr, *dZb = nullptr;
    std::vector<float *> Z;
    std::vector<bf16_t *> H;
    float *out = nullptr, *part = nullptr, *part2 = nullptr, *lsum = nullptr;
    float *wsplit = nullptr;  // split-K partial tiles of the weight gradients
    int *d_fail = nullptr;
    unsigned int *d_nfail = nullptr;
    int64_t fail_cap = 0;
};

namespace {

float *nm(ohc_surrogate *s, int which) {  // 0..3 in_*, 4..7 out_*
    const int a = s->nin, b = s->nout;
    const int64_t off[8] = {0, a, 2 * a, 3 * a, 4 * a, 4 * a + b, 4 * a + 2 * b, 4 * a + 3 * b};
    return s->d_norm + off[which];
}

constexpr int kMaxSplit = 16;  // split-K factor limit of the weight-gradient GEMMs

int ensure_ws(ohc_surrogate *s, int64_t B) {
    if (B <= s->cap) return OHC_OK;
    ohc_ctx *c = s->ctx;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (s->d_ws) (void)hipFree(s->d_ws);
    s->d_ws = nullptr;
    s->cap = 0;
    const int W = s->width;
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t bx = al(2 * B * s->nin), bz = al(4 * B * W), bh = al(2 * B * W),
                 bo = al(4 * B * s->nout), bdo = al(2 * B * s->nout),
                 bl = al(4 * (size_t)(nblk(B * s->nout) + 64)),
                 bp = bl + al(4 * (size_t)(B / ohcm::kRowChunk + 2) * std::max(W, s->nout));
    const size_t bw = al(4 * (size_t)kMaxSplit * W * std::max(std::max(W, s->nin), s->nout));
    const size_t total = bx + s->depth * (bz + bh) + bo + bdo + 2 * bh + bp + 256 + bw;
    HIPCHK(c, hipMalloc(&s->d_ws, total));
    char *q = (char *)s->d_ws;
    s->X16 = (bf16_t *)q;
    q += bx;
    s->Z.assign(s->depth, nullptr);
    s->H.assign(s->depth, nullptr);
    for (int l = 0; l < s->depth; ++l) {
        s->Z[l] = (float *)q;
        q += bz;
        s->H[l] = (bf16_t *)q;
        q += bh;
    }
    s->out = (float *)q;
    q += bo;
    s->dOut = (bf16_t *)q;
    q += bdo;
    s->dZa = (bf16_t *)q;
    q += bh;
    s->dZb = (bf16_t *)q;
    q += bh;
    s->lsum = (float *)q;
    q += 256;
    s->part = (float *)q;
    s->part2 = (float *)(q + bl);
    q += bp;
    s->wsplit = (float *)q;
    s->cap = B;
    return OHC_OK;
}

template <bool AK, bool BK, int EPI>
hipError_t gemm(hipStream_t st, int M, int N, int K, const bf16_t *A, int64_t lda, const bf16_t *B,
                int64_t ldb, float *C32, bf16_t *C16, int64_t ldc, const float *bias = nullptr,
                const float *Zs = nullptr) {
    ohcm::GemmArgs g{M, N, K, A, lda, B, ldb, C32, C16, ldc, bias, Zs, K, 0};
    const dim3 grid((unsigned)((N + ohcm::kTN - 1) / ohcm::kTN), (unsigned)((M + ohcm::kTM - 1) / ohcm::kTM));
    hipLaunchKernelGGL((ohcm::gemm_bf16<AK, BK, EPI>), grid, dim3(256), 0, st, g);
    return hipGetLastError();
}

// Weight gradients dW = dZ^T H reduce over the batch (K = batch rows) into a
// few 64 x 64 tiles; split K so that the launch fills the chip, each split
// writing its own partial tile into ws, then add the partials in a fixed
// order (deterministic).
hipError_t gemm_wgrad(hipStream_t st, int M, int N, int K, const bf16_t *A, int64_t lda,
                      const bf16_t *B, int64_t ldb, float *C32, int64_t ldc, float *ws) {
    const int gx = (N + ohcm::kTN - 1) / ohcm::kTN, gy = (M + ohcm::kTM - 1) / ohcm::kTM;
    int S = std::min(kMaxSplit, std::max(1, 256 / (gx * gy)));
    S = std::min(S, std::max(1, K / 256));  // at least 256 batch rows per split
    if (S <= 1) return gemm<false, false, ohcm::EPI_F32>(st, M, N, K, A, lda, B, ldb, C32, nullptr, ldc);
    const int kc = ((K + S - 1) / S + ohcm::kTK - 1) / ohcm::kTK * ohcm::kTK;
    S = (K + kc - 1) / kc;
    const int64_t zs = (int64_t)M * ldc;
    ohcm::GemmArgs g{M, N, K, A, lda, B, ldb, ws, nullptr, ldc, nullptr, nullptr, kc, zs};
    hipLaunchKernelGGL((ohcm::gemm_bf16<false, false, ohcm::EPI_F32>), dim3(gx, gy, S), dim3(256), 0, st,
                       g);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(ohcm::splitk_sum, dim3(nblk((int64_t)M * N)), dim3(256), 0, st, S, M, N, ldc,
                       zs, ws, C32);
    return hipGetLastError();
}

int colsum(ohc_surrogate *s, int64_t n, int N, const bf16_t *dZ, float *db) {
    ohc_ctx *c = s->ctx;
    const int nchunk = (int)((n + ohcm::kRowChunk - 1) / ohcm::kRowChunk);
    float *part = s->part2;
    hipLaunchKernelGGL(ohcm::colsum_partial, dim3(nblk(N, 64), nchunk), dim3(256), 0, c->stream, n,
                       N, (int64_t)N, dZ, part);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(ohcm::colsum_finish, dim3(nblk(N)), dim3(256), 0, c->stream, nchunk, N,
                       part, db);
    HIPCHK(c, hipGetLastError());
    return OHC_OK;
}

// Forward and backward over a batch of raw features / targets: gradients into
// d_g, loss (sum over the batch / n) into lsum.
int forward_backward(ohc_surrogate *s, int64_t n, const float *X, const float *Yt) {
    ohc_ctx *c = s->ctx;
    hipStream_t st = c->stream;
    int rc = ensure_ws(s, n);
    if (rc) return rc;
    const int W = s->width, L = s->depth, nin = s->nin, nout = s->nout;
    hipLaunchKernelGGL(ohcm::normalize_in, dim3(nblk(n * nin)), dim3(256), 0, st, n, nin, nin, X,
                       nm(s, 0), nm(s, 1), s->X16);
    HIPCHK(c, hipGetLastError());
    const bf16_t *P16 = s->d_p16;
    const float *P = s->d_p;
    // forward
    for (int l = 0; l < L; ++l) {
        const bf16_t *A = l == 0 ? s->X16 : s->H[l - 1];
        const int K = l == 0 ? nin : W;
        HIPCHK(c, (gemm<true, true, ohcm::EPI_BIAS_GELU>(st, (int)n, W, K, A, K, P16 + s->w_off[l], K,
                                                         s->Z[l], s->H[l], W, P + s->b_off[l])));
    }
    HIPCHK(c, (gemm<true, true, ohcm::EPI_BIAS_F32>(st, (int)n, nout, W, s->H[L - 1], W,
                                                    P16 + s->w_off[L], W, s->out, nullptr, nout,
                                                    P + s->b_off[L])));
    const unsigned nb = nblk(n * nout);
    hipLaunchKernelGGL(ohcm::loss_grad, dim3(nb), dim3(256), 0, st, n, nout, s->out, Yt, nm(s, 4),
                       nm(s, 6), nm(s, 7), s->dOut, s->part);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(ohcm::sum_partials, dim3(1), dim3(256), 0, st, (int)nb, s->part, s->lsum);
    HIPCHK(c, hipGetLastError());
    // backward
    float *G = s->d_g;
    HIPCHK(c, gemm_wgrad(st, nout, W, (int)n, s->dOut, nout, s->H[L - 1], W, G + s->w_off[L], W,
                         s->wsplit));
    if ((rc = colsum(s, n, nout, s->dOut, G + s->b_off[L]))) return rc;
    bf16_t *dz = s->dZa, *dz2 = s->dZb;
    HIPCHK(c, (gemm<true, false, ohcm::EPI_GELU_BWD>(st, (int)n, W, nout, s->dOut, nout,
                                                     P16 + s->w_off[L], W, nullptr, dz, W, nullptr,
                                                     s->Z[L - 1])));
    for (int l = L - 1; l >= 0; --l) {
        const bf16_t *Hin = l == 0 ? s->X16 : s->H[l - 1];
        const int K = l == 0 ? nin : W;
        HIPCHK(c, gemm_wgrad(st, W, K, (int)n, dz, W, Hin, K, G + s->w_off[l], K, s->wsplit));
        if ((rc = colsum(s, n, W, dz, G + s->b_off[l]))) return rc;
        if (l > 0) {
            HIPCHK(c, (gemm<true, false, ohcm::EPI_GELU_BWD>(st, (int)n, W, W, dz, W,
                                                             P16 + s->w_off[l], W, nullptr, dz2, W,
                                                             nullptr, s->Z[l - 1])));
            std::swap(dz, dz2);
        }
    }
    return OHC_OK;
}

int pack(ohc_surrogate *s) {
    if (!s->dirty) return OHC_OK;
    ohc_ctx *c = s->ctx;
    for (int l = 0; l <= s->depth; ++l) {
        const int prow = l < s->depth ? s->width : s->noutp;
        const int pcol = l == 0 ? s->kin : s->width;
        hipLaunchKernelGGL(pack_layer, dim3(nblk((int64_t)prow * pcol)), dim3(256), 0, c->stream,
                           s->rows[l], s->cols[l], prow, pcol, s->d_p + s->w_off[l],
                           s->d_W + s->woff[l]);
        HIPCHK(c, hipGetLastError());
        hipLaunchKernelGGL(pack_bias, dim3(nblk(prow)), dim3(256), 0, c->stream, s->rows[l], prow,
                           s->d_p + s->b_off[l], s->d_b + s->boff[l]);
        HIPCHK(c, hipGetLastError());
        if (l >= 1 && l < s->depth) {
            hipLaunchKernelGGL(pack_layer_fp8, dim3(s->width), dim3(256), 0, c->stream, s->width,
                               s->d_p + s->w_off[l], s->d_W8 + s->w8off[l], s->d_s8 + s->s8off[l]);
            HIPCHK(c, hipGetLastError());
        }
    }
    s->dirty = false;
    return OHC_OK;
}

template <int TM, int WD, int KIN, int NOUTP>
void launch_infer(bool fp8, const ohcm::InferArgs &a, hipStream_t st) {
    if (fp8)
        hipLaunchKernelGGL((ohcm::mlp_infer<TM, WD, KIN, NOUTP, true>), dim3(nblk(a.n, TM)), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((ohcm::mlp_infer<TM, WD, KIN, NOUTP, false>), dim3(nblk(a.n, TM)), dim3(256), 0, st, a);
}

template <int TM, int WD>
hipError_t infer_wd(int kin, int noutp, bool fp8, const ohcm::InferArgs &a, hipStream_t st) {
    if (noutp == 32) {
        if (kin == 16) launch_infer<TM, WD, 16, 32>(fp8, a, st);
        else if (kin == 32) launch_infer<TM, WD, 32, 32>(fp8, a, st);
        else if (kin == 48) launch_infer<TM, WD, 48, 32>(fp8, a, st);
        else launch_infer<TM, WD, 64, 32>(fp8, a, st);
    } else {
        if (kin == 16) launch_infer<TM, WD, 16, 64>(fp8, a, st);
        else if (kin == 32) launch_infer<TM, WD, 32, 64>(fp8, a, st);
        else if (kin == 48) launch_infer<TM, WD, 48, 64>(fp8, a, st);
        else launch_infer<TM, WD, 64, 64>(fp8, a, st);
    }
    return hipGetLastError();
}

int infer(ohc_surrogate *s, ohcm::InferArgs &a) {
    ohc_ctx *c = s->ctx;
    int rc = pack(s);
    if (rc) return rc;
    a.ns = s->ns;
    a.ne = s->ne;
    a.W = s->d_W;
    a.b = s->d_b;
    for (int l = 0; l <= s->depth; ++l) {
        a.woff[l] = s->woff[l];
        a.boff[l] = s->boff[l];
    }
    a.depth = s->depth;
    a.W8 = s->d_W8;
    a.s8 = s->d_s8;
    for (int l = 0; l <= s->depth; ++l) {
        a.w8off[l] = s->w8off[l];
        a.s8off[l] = s->s8off[l];
    }
    a.delta = s->has_delta ? s->d_delta : nullptr;
    a.in_mean = nm(s, 0);
    a.in_rstd = nm(s, 1);
    a.zlo = nm(s, 2);
    a.zhi = nm(s, 3);
    a.out_mean = nm(s, 4);
    a.out_std = nm(s, 5);
    a.elemW = s->d_elemW;
    a.dt_train = s->dt_train;
    a.T_low = c->opts.T_low;
    a.T_high = c->opts.T_high;
    if (a.n == 0) return OHC_OK;
    const bool fp8 = s->precision == 1;
    if (s->width == 256) HIPCHK(c, (infer_wd<64, 256>(s->kin, s->noutp, fp8, a, c->stream)));
    else HIPCHK(c, (infer_wd<32, 512>(s->kin, s->noutp, fp8, a, c->stream)));
    return OHC_OK;
}

}  // namespace

extern "C" {

int ohc_surrogate_create(ohc_ctx *c, int32_t width, int32_t depth, uint64_t seed,
                         ohc_surrogate **out) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!out) return fail(c, OHC_ERR_ARG, "null argument");
    *out = nullptr;
    if (width != 256 && width != 512) return fail(c, OHC_ERR_ARG, "width must be 256 or 512");
    if (depth < 1 || depth > 6) return fail(c, OHC_ERR_ARG, "depth must be in 1..6");
    const ohc::Mechanism &m = c->mech;
    if (m.ns + 2 > 64) return fail(c, OHC_ERR_UNSUPPORTED, "surrogate supports ns <= 62");
    if (!c->opts.device_pointers)
        return fail(c, OHC_ERR_ARG, "the surrogate needs a context with device_pointers = 1");
    auto s = std::make_unique<ohc_surrogate>();
    s->ctx = c;
    s->ns = m.ns;
    s->nin = m.ns + 2;
    s->nout = m.ns + 1;
    s->width = width;
    s->depth = depth;
    s->kin = rup(s->nin, 16);
    s->noutp = s->nout <= 32 ? 32 : 64;
    // flat parameter layout and Kaiming-uniform initialisation (seeded)
    std::mt19937_64 rng(seed);
    std::vector<float> p;
    int64_t pw = 0, pb = 0;
    for (int l = 0; l <= depth; ++l) {
        const int in = l == 0 ? s->nin : width, o = l < depth ? width : s->nout;
        s->rows.push_back(o);
        s->cols.push_back(in);
        s->w_off.push_back((int64_t)p.size());
        const double bound = std::sqrt(6.0 / in) * (l < depth ? 1.0 : 0.1);
        for (int64_t e = 0; e < (int64_t)o * in; ++e)
            p.push_back((float)((2.0 * ((rng() >> 11) * (1.0 / 9007199254740992.0)) - 1.0) * bound));
        s->b_off.push_back((int64_t)p.size());
        for (int e = 0; e < o; ++e) p.push_back(0.0f);
        // packed inference offsets (16-byte aligned)
        const int prow = l < depth ? width : s->noutp, pcol = l == 0 ? s->kin : width;
        s->woff[l] = (int)pw;
        s->boff[l] = (int)pb;
        pw += rup(prow * pcol, 64);
        pb += rup(prow, 64);
    }
    int64_t p8 = 0, q8 = 0;
    for (int l = 1; l < depth; ++l) {
        s->w8off[l] = (int)p8;
        s->s8off[l] = (int)q8;
        p8 += rup(width * width, 256);
        q8 += rup(width, 64);
    }
    s->np = (int64_t)p.size();
    ohc_surrogate *sp = s.get();
    HIPCHK(c, hipMalloc((void **)&sp->d_p, 4 * sp->np));
    HIPCHK(c, hipMalloc((void **)&sp->d_g, 4 * sp->np));
    HIPCHK(c, hipMalloc((void **)&sp->d_m, 4 * sp->np));
    HIPCHK(c, hipMalloc((void **)&sp->d_v, 4 * sp->np));
    HIPCHK(c, hipMalloc((void **)&sp->d_p16, 2 * sp->np));
    HIPCHK(c, hipMalloc((void **)&sp->d_W, 2 * pw));
    HIPCHK(c, hipMalloc((void **)&sp->d_b, 4 * pb));
    HIPCHK(c, hipMalloc((void **)&sp->d_norm, 4 * (4 * sp->nin + 4 * sp->nout)));
    HIPCHK(c, hipMalloc((void **)&sp->d_W8, std::max<int64_t>(p8, 256)));
    HIPCHK(c, hipMalloc((void **)&sp->d_s8, 4 * std::max<int64_t>(q8, 64)));
    HIPCHK(c, hipMalloc((void **)&sp->d_delta, 4 * sp->nout));
    HIPCHK(c, hipMemcpy(sp->d_p, p.data(), 4 * sp->np, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(sp->d_m, 0, 4 * sp->np));
    HIPCHK(c, hipMemset(sp->d_v, 0, 4 * sp->np));
    hipLaunchKernelGGL(ohcm::to_bf16, dim3(nblk(sp->np)), dim3(256), 0, c->stream, sp->np, sp->d_p,
                       sp->d_p16);
    HIPCHK(c, hipGetLastError());
    // default normalisation: identity, no OOD limit, unit loss weights
    std::vector<float> nmh(4 * sp->nin + 4 * sp->nout, 0.0f);
    for (int k = 0; k < sp->nin; ++k) {
        nmh[sp->nin + k] = 1.0f;
        nmh[2 * sp->nin + k] = -3.0e38f;
        nmh[3 * sp->nin + k] = 3.0e38f;
    }
    for (int k = 0; k < sp->nout; ++k) {
        nmh[4 * sp->nin + sp->nout + k] = 1.0f;
        nmh[4 * sp->nin + 2 * sp->nout + k] = 1.0f;
        nmh[4 * sp->nin + 3 * sp->nout + k] = 1.0f;
    }
    HIPCHK(c, hipMemcpy(sp->d_norm, nmh.data(), 4 * nmh.size(), hipMemcpyHostToDevice));
    // element moles per unit mass of each species
    sp->ne = std::max(m.ne, 1);
    std::vector<double> ew((size_t)sp->ns * sp->ne, 0.0);
    for (int k = 0; k < sp->ns; ++k)
        for (int e = 0; e < m.ne; ++e) ew[(size_t)k * sp->ne + e] = m.elem[(size_t)k * m.ne + e] / m.W[k];
    HIPCHK(c, hipMalloc((void **)&sp->d_elemW, 8 * ew.size()));
    HIPCHK(c, hipMemcpy(sp->d_elemW, ew.data(), 8 * ew.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = s.release();
    return OHC_OK;
}

int ohc_surrogate_free(ohc_surrogate *s) {
    if (!s) return OHC_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    for (void *q : {(void *)s->d_p, (void *)s->d_g, (void *)s->d_m, (void *)s->d_v, (void *)s->d_p16,
                    (void *)s->d_W, (void *)s->d_b, (void *)s->d_norm, (void *)s->d_elemW,
                    s->d_ws, (void *)s->d_fail, (void *)s->d_W8, (void *)s->d_s8,
                    (void *)s->d_delta, (void *)s->d_peer_out})
        if (q) (void)hipFree(q);
    delete s;
    return OHC_OK;
}

int ohc_surrogate_num_params(const ohc_surrogate *s, int64_t *n) {
    if (!s || !n) return fail(nullptr, OHC_ERR_ARG, "null argument");
    *n = s->np;
    return OHC_OK;
}

int ohc_surrogate_get_params(ohc_surrogate *s, float *host) {
    if (!s || !host) return fail(nullptr, OHC_ERR_ARG, "null argument");
    ohc_ctx *c = s->ctx;
    HIPCHK(c, hipMemcpyAsync(host, s->d_p, 4 * s->np, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OHC_OK;
}

int ohc_surrogate_set_params(ohc_surrogate *s, const float *host) {
    if (!s || !host) return fail(nullptr, OHC_ERR_ARG, "null argument");
    ohc_ctx *c = s->ctx;
    HIPCHK(c, hipMemcpyAsync(s->d_p, host, 4 * s->np, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(ohcm::to_bf16, dim3(nblk(s->np)), dim3(256), 0, c->stream, s->np, s->d_p, s->d_p16);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    s->dirty = true;
    return OHC_OK;
}

int ohc_surrogate_set_norm(ohc_surrogate *s, const float *in_mean, const float *in_std,
                           const float *in_lo, const float *in_hi, const float *out_mean,
                           const float *out_std, const float *out_weight, double dt_train) {
    if (!s || !in_mean || !in_std || !out_mean || !out_std)
        return fail(nullptr, OHC_ERR_ARG, "null argument");
    ohc_ctx *c = s->ctx;
    const int a = s->nin, b = s->nout;
    std::vector<float> h(4 * a + 4 * b);
    for (int k = 0; k < a; ++k) {
        if (!(in_std[k] > 0)) return fail(c, OHC_ERR_ARG, "input std must be > 0");
        h[k] = in_mean[k];
        h[a + k] = 1.0f / in_std[k];
        h[2 * a + k] = in_lo ? (in_lo[k] - in_mean[k]) / in_std[k] : -3.0e38f;
        h[3 * a + k] = in_hi ? (in_hi[k] - in_mean[k]) / in_std[k] : 3.0e38f;
    }
    for (int k = 0; k < b; ++k) {
        if (!(out_std[k] > 0)) return fail(c, OHC_ERR_ARG, "output std must be > 0");
        h[4 * a + k] = out_mean[k];
        h[4 * a + b + k] = out_std[k];
        h[4 * a + 2 * b + k] = 1.0f / out_std[k];
        h[4 * a + 3 * b + k] = out_weight ? out_weight[k] : 1.0f;
    }
    HIPCHK(c, hipMemcpyAsync(s->d_norm, h.data(), 4 * h.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    s->dt_train = dt_train;
    return OHC_OK;
}

int ohc_surrogate_features(ohc_surrogate *s, int64_t n, const double *T0, const double *p,
                           const double *Y0, int64_t ldY, const double *T1, const double *Y1,
                           float *X, float *Yt) {
    if (!s || !T0 || !p || !Y0 || !X || (Yt && (!T1 || !Y1)) || n < 0 || ldY < n)
        return fail(s ? s->ctx : nullptr, OHC_ERR_ARG, "bad arguments");
    ohc_ctx *c = s->ctx;
    if (n == 0) return OHC_OK;
    hipLaunchKernelGGL(features_kernel, dim3(nblk(n)), dim3(256), 0, c->stream, n, s->ns, T0, p, Y0,
                       ldY, T1, Y1, s->has_delta ? s->d_delta : nullptr, X, Yt);
    HIPCHK(c, hipGetLastError());
    return OHC_OK;
}

int ohc_surrogate_train_step(ohc_surrogate *s, int64_t n, const float *X, const float *Yt,
                             float lr, float *loss) {
    if (!s || !X || !Yt || n < 1 || n > (1 << 24)) return fail(s ? s->ctx : nullptr, OHC_ERR_ARG, "bad arguments");
    ohc_ctx *c = s->ctx;
    int rc = forward_backward(s, n, X, Yt);
    if (rc) return rc;
    s->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f / (1.0f - std::pow(b1, (float)s->step)),
                c2 = 1.0f / (1.0f - std::pow(b2, (float)s->step));
    hipLaunchKernelGGL(ohcm::adam_step, dim3(nblk(s->np)), dim3(256), 0, c->stream, s->np, s->d_p,
                       s->d_g, s->d_m, s->d_v, lr, b1, b2, 1e-8f, c1, c2, s->d_p16);
    HIPCHK(c, hipGetLastError());
    s->dirty = true;
    if (loss) {
        HIPCHK(c, hipMemcpyAsync(loss, s->lsum, 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return OHC_OK;
}

int ohc_surrogate_gradient(ohc_surrogate *s, int64_t n, const float *X, const float *Yt,
                           float *grad, float *loss) {
    if (!s || !X || !Yt || !grad || n < 1 || n > (1 << 24))
        return fail(s ? s->ctx : nullptr, OHC_ERR_ARG, "bad arguments");
    ohc_ctx *c = s->ctx;
    int rc = forward_backward(s, n, X, Yt);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(grad, s->d_g, 4 * s->np, hipMemcpyDeviceToHost, c->stream));
    if (loss) HIPCHK(c, hipMemcpyAsync(loss, s->lsum, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OHC_OK;
}

int ohc_surrogate_predict(ohc_surrogate *s, int64_t n, const double *T, const double *p,
                          const double *Y, int64_t ldY, float *dout) {
    if (!s || !T || !p || !Y || !dout || n < 0 || ldY < n)
        return fail(s ? s->ctx : nullptr, OHC_ERR_ARG, "bad arguments");
    ohcm::InferArgs a;
    std::memset(&a, 0, sizeof(a));
    a.n = n;
    a.T = const_cast<double *>(T);
    a.Y = const_cast<double *>(Y);
    a.p = p;
    a.ldY = ldY;
    a.mode = 1;
    a.out32 = dout;
    return infer(s, a);
}

int ohc_chem_solve_update_surrogate(ohc_ctx *c, ohc_surrogate *s, int64_t n, double *T,
                                    const double *p, const double *rho, double *Y, int64_t ldY,
                                    const double *dt, double dtFactor, double *subDt,
                                    double gate_tol, int64_t *n_fallback, ohc_stats *stats) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!s || s->ctx != c) return fail(c, OHC_ERR_ARG, "surrogate belongs to another context");
    if (!T || !p || !rho || !Y || !dt || !subDt || n < 0 || ldY < n || n > INT32_MAX)
        return fail(c, OHC_ERR_ARG, "bad arguments");
    if (!(s->dt_train > 0)) return fail(c, OHC_ERR_ARG, "surrogate has no training time step");
    if (n > s->fail_cap) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (s->d_fail) (void)hipFree(s->d_fail);
        s->d_fail = nullptr;
        HIPCHK(c, hipMalloc((void **)&s->d_fail, 4 * n + 256));
        s->d_nfail = (unsigned int *)((char *)s->d_fail + 4 * n + 128 - (4 * n) % 128);
        s->fail_cap = n;
    }
    hipLaunchKernelGGL(zero_u32, dim3(1), dim3(1), 0, c->stream, s->d_nfail);
    HIPCHK(c, hipGetLastError());
    if (s->peer) {
        // the peer's increments for every cell, from the same (unchanged) state
        ohc_surrogate *q = s->peer;
        if (n > q->peer_cap) {
            HIPCHK(c, hipStreamSynchronize(c->stream));
            if (q->d_peer_out) (void)hipFree(q->d_peer_out);
            q->d_peer_out = nullptr;
            q->peer_cap = 0;
            HIPCHK(c, hipMalloc((void **)&q->d_peer_out, 4 * (size_t)n * q->nout));
            q->peer_cap = n;
        }
        ohcm::InferArgs b;
        std::memset(&b, 0, sizeof(b));
        b.n = n;
        b.T = T;
        b.Y = Y;
        b.p = p;
        b.ldY = ldY;
        b.mode = 1;
        b.out32 = q->d_peer_out;
        if ((rc = infer(q, b))) return rc;
    }
    ohcm::InferArgs a;
    std::memset(&a, 0, sizeof(a));
    a.n = n;
    a.T = T;
    a.Y = Y;
    a.p = p;
    a.ldY = ldY;
    a.dt = dt;
    a.dtFactor = dtFactor;
    a.gate_tol = gate_tol;
    a.fail_list = s->d_fail;
    a.fail_count = s->d_nfail;
    a.mode = 0;
    a.peer = s->peer ? s->peer->d_peer_out : nullptr;
    a.peer_tol = s->peer_tol;
    if ((rc = infer(s, a))) return rc;
    unsigned int nf = 0;
    HIPCHK(c, hipMemcpyAsync(&nf, s->d_nfail, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (n_fallback) *n_fallback = nf;
    // the gated-out cells take the plain integrator (solveUpdate semantics)
    return chem_integrate(c, 1, n, nullptr, T, p, rho, nullptr, Y, ldY, dt, dtFactor, subDt,
                          nullptr, n, stats, s->d_fail, (int64_t)nf);
}

int ohc_surrogate_set_target_scale(ohc_surrogate *s, const float *delta) {
    if (!s) return fail(nullptr, OHC_ERR_ARG, "null argument");
    ohc_ctx *c = s->ctx;
    if (!delta) {
        s->has_delta = false;
        return OHC_OK;
    }
    for (int k = 0; k < s->nout; ++k)
        if (!(delta[k] > 0.0f) || !std::isfinite(delta[k]))
            return fail(c, OHC_ERR_ARG, "target scales must be finite and > 0");
    HIPCHK(c, hipMemcpyAsync(s->d_delta, delta, 4 * s->nout, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    s->has_delta = true;
    return OHC_OK;
}

int ohc_surrogate_set_precision(ohc_surrogate *s, int32_t precision) {
    if (!s) return fail(nullptr, OHC_ERR_ARG, "null argument");
    if (precision != 0 && precision != 1) return fail(s->ctx, OHC_ERR_ARG, "precision must be 0 or 1");
    s->precision = precision;
    return OHC_OK;
}

int ohc_surrogate_set_peer(ohc_surrogate *s, ohc_surrogate *peer, double tol) {
    if (!s) return fail(nullptr, OHC_ERR_ARG, "null argument");
    if (peer && (peer == s || peer->ctx != s->ctx || peer->ns != s->ns))
        return fail(s->ctx, OHC_ERR_ARG, "the peer must be another surrogate of the same context");
    if (peer && !(tol >= 0.0)) return fail(s->ctx, OHC_ERR_ARG, "peer tolerance must be >= 0");
    s->peer = peer;
    s->peer_tol = tol;
    return OHC_OK;
}

int ohc_surrogate_save(const ohc_surrogate *s, const char *path) {
    if (!s || !path) return fail(nullptr, OHC_ERR_ARG, "null argument");
    ohc_ctx *c = s->ctx;
    std::vector<float> p(s->np), nmh(4 * s->nin + 4 * s->nout);
    HIPCHK(c, hipMemcpy(p.data(), s->d_p, 4 * s->np, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(nmh.data(), s->d_norm, 4 * nmh.size(), hipMemcpyDeviceToHost));
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(c, OHC_ERR_ARG, std::string("cannot write ") + path);
    std::vector<float> dl(s->nout, 0.0f);
    if (s->has_delta) HIPCHK(c, hipMemcpy(dl.data(), s->d_delta, 4 * dl.size(), hipMemcpyDeviceToHost));
    // format 2: header, dt_train, normalisation, target scales (0: none), parameters
    const int32_t hdr[6] = {0x4f484353, 2, s->ns, s->width, s->depth, s->has_delta ? 1 : 0};
    bool ok = std::fwrite(hdr, sizeof hdr, 1, f) == 1 &&
              std::fwrite(&s->dt_train, 8, 1, f) == 1 &&
              std::fwrite(nmh.data(), 4, nmh.size(), f) == nmh.size() &&
              std::fwrite(dl.data(), 4, dl.size(), f) == dl.size() &&
              std::fwrite(p.data(), 4, p.size(), f) == p.size();
    ok = (std::fclose(f) == 0) && ok;
    return ok ? OHC_OK : fail(c, OHC_ERR_ARG, std::string("write failed: ") + path);
}

int ohc_surrogate_load(ohc_surrogate *s, const char *path) {
    if (!s || !path) return fail(nullptr, OHC_ERR_ARG, "null argument");
    ohc_ctx *c = s->ctx;
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(c, OHC_ERR_ARG, std::string("cannot read ") + path);
    int32_t hdr[6];
    double dtt = 0.0;
    std::vector<float> p(s->np), nmh(4 * s->nin + 4 * s->nout), dl(s->nout, 0.0f);
    bool ok = std::fread(hdr, sizeof hdr, 1, f) == 1 && hdr[0] == 0x4f484353 &&
              (hdr[1] == 1 || hdr[1] == 2) && hdr[2] == s->ns && hdr[3] == s->width &&
              hdr[4] == s->depth && std::fread(&dtt, 8, 1, f) == 1 &&
              std::fread(nmh.data(), 4, nmh.size(), f) == nmh.size() &&
              (hdr[1] == 1 || std::fread(dl.data(), 4, dl.size(), f) == dl.size()) &&
              std::fread(p.data(), 4, p.size(), f) == p.size();
    std::fclose(f);
    if (!ok) return fail(c, OHC_ERR_ARG, std::string("not a surrogate of this shape: ") + path);
    HIPCHK(c, hipMemcpy(s->d_norm, nmh.data(), 4 * nmh.size(), hipMemcpyHostToDevice));
    s->dt_train = dtt;
    s->has_delta = hdr[1] == 2 && hdr[5] == 1;
    if (s->has_delta) HIPCHK(c, hipMemcpy(s->d_delta, dl.data(), 4 * dl.size(), hipMemcpyHostToDevice));
    return ohc_surrogate_set_params(s, p.data());
}

}  // extern "C"
