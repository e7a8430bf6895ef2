// wave_kernels.hpp -- generic (run-time mechanism table) kernels: one
// wavefront integrates one cell at a time.
//
// The lane kernels keep a whole cell in one lane's registers, which works up
// to ~10 species.  Larger mechanisms (32 species / 150 reactions, 54 / 270)
// need a (ns+1)^2 Jacobian per cell, so here the 64 lanes of a wavefront share
// one cell:
//   * species k lives in lane k: thermo, net rate omega_k, Jacobian row k;
//   * reactions are evaluated 64 at a time, reaction r in lane r % 64, and
//     each reaction lane adds its contributions to omega and to the Jacobian
//     rows of its participants with LDS floating-point atomics (ds_add_f64),
//     straight from the registers that hold the reaction's data;
//   * the LU factorisation keeps row i of I - h J in lane i's registers and
//     eliminates with partial pivoting across lanes (pivot search by wave
//     reduction, pivot row broadcast lane to all), forward substitution
//     folded into the elimination;
//   * the ImplicitEuler control logic is uniform across the wavefront.
// The arithmetic follows the same reference functions as chem_device.hpp and
// lane_kernels.hpp (chemistrySource.cpp:863-1055, reactionTypes.cpp,
// reactionRateTypes, ImplicitEuler.cpp:40-342, SquareMatrix.hpp:185-238).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "chem_device.hpp"
#include "lane_kernels.hpp"

namespace ohcd {

constexpr int kWP = 3;  // distinct species per reaction side in the run-time tables

// How the LU's pivot row reaches the other lanes: 1 by v_readlane pairs (two
// vector ALU instructions per element, no LDS round trip), 2 by ds_bpermute
// pairs (no vector ALU work, but an LDS round trip on every step's critical
// path: 1.3x the LU time of 1 on the 32-species mechanism), 0 through an LDS
// row buffer written by the pivot lane (a barrier per elimination step).
#ifndef OHC_LU_BCAST
#define OHC_LU_BCAST 1
#endif
static_assert(kWP == 3, "the reaction kernels multiply three participant powers");

// species fields, [field][64]
enum {
    SP_W,
    SP_TLOW,
    SP_THIGH,
    SP_TCOM,
    SP_ALO,
    SP_AHI = SP_ALO + 7,
    SP_HLO = SP_AHI + 7,
    SP_HHI = SP_HLO + 4,
    SP_SLO = SP_HHI + 4,
    SP_SHI = SP_SLO + 3,
    SP_GLO = SP_SHI + 3,
    SP_GHI = SP_GLO + 4,
    SP_EKB = SP_GHI + 4,  // transport: eps/k [K]
    SP_SIG,               // transport: sigma [Angstrom]
    SP_NFIELD
};
// Reaction tables are chunk-major: reaction r = 64 c + l (the chunk a
// wavefront evaluates at once, l its lane) keeps field f at
// [(c * NFIELD + f) * 64 + l].
// Reaction 32-bit fields, read for every reaction:
//   RH_HDR  nf | nb << 2 | rev << 4 | kind << 6 | fot << 9 | GEN << 12
//   RH_IDX  fi0 | fi1 << 8 | fi2 << 16 | bi0 << 24;  RH_IDX2  bi1 | bi2 << 8
//   RH_FNU.. RH_SUMS  stoichiometric coefficients, orders and their sum as
//   float, exact for the small integers and halves of most mechanisms; a
//   reaction with any other value has GEN set and reads the FP64 copies in
//   the double table instead.
enum { RH_HDR, RH_IDX, RH_IDX2, RH_FNU, RH_BNU = RH_FNU + kWP, RH_FO = RH_BNU + kWP,
       RH_BO = RH_FO + kWP, RH_SUMS = RH_BO + kWP, RH_NFIELD };
constexpr unsigned kRhGen = 1u << 12;
// reaction double fields: the Arrhenius triple read for every reaction, then
// what only some reactions read (fall-off, explicit reverse, GEN copies)
enum {
    RD_A,
    RD_B,
    RD_TA,
    RD_FNU,
    RD_BNU = RD_FNU + kWP,
    RD_FO = RD_BNU + kWP,
    RD_BO = RD_FO + kWP,
    RD_A0,
    RD_B0,
    RD_TA0,
    RD_AR,
    RD_BR,
    RD_TAR,
    RD_FOP,
    RD_SUMS = RD_FOP + 5,
    RD_NFIELD
};

// Device copy of a mechanism (built by the host from the SoA tables).
struct WMech {
    int ns, nr, nrp;      // nrp = nr rounded up to a multiple of 64
    const double *sp;     // [SP_NFIELD][64]
    const unsigned *rh;   // [nrp / 64][RH_NFIELD][64]
    const double *rd;     // [nrp / 64][RD_NFIELD][64]
    const double *eff;    // [nr][64] third-body efficiencies
};

struct WaveParams {
    LaneParams L;
    WMech M;
    int exact_dkp;  // 1: exact d(Kp)/dT instead of the reference's sign convention
};

// Reads of the per-reaction tables through buffer resources: the lane is
// the one vector offset, the chunk and field a scalar offset, so a table read
// needs no vector address arithmetic.  Reads past a table return zero.
struct RxnTab {
    __amdgpu_buffer_rsrc_t rh, rd, ef;
};
__device__ __forceinline__ RxnTab rxn_tab(const WMech &M) {
    constexpr int kRsrcWord3 = 0x00020000;  // raw 32-bit data, CDNA
    RxnTab t;
    t.rh = __builtin_amdgcn_make_buffer_rsrc((void *)M.rh, 0, RH_NFIELD * M.nrp * 4, kRsrcWord3);
    t.rd = __builtin_amdgcn_make_buffer_rsrc((void *)M.rd, 0, RD_NFIELD * M.nrp * 8, kRsrcWord3);
    t.ef = __builtin_amdgcn_make_buffer_rsrc((void *)M.eff, 0, M.nr * 64 * 8, kRsrcWord3);
    return t;
}
// field f of reaction 64 c + lane (c wave-uniform)
__device__ __forceinline__ unsigned ldu(const RxnTab &t, int f, int c, int lane) {
    return __builtin_amdgcn_raw_buffer_load_b32(t.rh, lane * 4, (c * RH_NFIELD + f) * 256, 0);
}
__device__ __forceinline__ double ldf(const RxnTab &t, int f, int c, int lane) {
    return (double)__builtin_bit_cast(float, ldu(t, f, c, lane));
}
__device__ __forceinline__ double ldd(const RxnTab &t, int f, int c, int lane) {
    return __builtin_bit_cast(
        double, __builtin_amdgcn_raw_buffer_load_b64(t.rd, lane * 8, (c * RD_NFIELD + f) * 512, 0));
}
// efficiency of species k (wave-uniform) in reaction r
__device__ __forceinline__ double lde(const RxnTab &t, int r, int k) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(t.ef, r * 512, k * 8, 0));
}
// efficiency of species `lane` in reaction r (wave-uniform)
__device__ __forceinline__ double lde_col(const RxnTab &t, int r, int lane) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(t.ef, lane * 8, r * 512, 0));
}
__device__ __forceinline__ double rdlane(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
// lane (p4 / 4)'s v to every lane, through the LDS crossbar
__device__ __forceinline__ double bperm(double v, int p4) {
    const int lo = __builtin_amdgcn_ds_bpermute(p4, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(p4, __double2hiint(v));
    return __hiloint2double(hi, lo);
}
// Data-parallel-primitive move of a double within each row of 16 lanes
// (quad permutations and row mirrors; no LDS traffic).
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141;  // row_half_mirror
constexpr int kDppMirror = 0x140;      // row_mirror
// Wavefront reductions: pairwise within rows by DPP, then the four row values
// through scalar reads; every lane gets the same bits (each combine is
// symmetric in its two operands).
__device__ __forceinline__ double wsum(double v) {
    v += dpp<kDppXor1>(v);
    v += dpp<kDppXor2>(v);
    v += dpp<kDppHalfMirror>(v);
    v += dpp<kDppMirror>(v);
    return (rdlane(v, 0) + rdlane(v, 16)) + (rdlane(v, 32) + rdlane(v, 48));
}
__device__ __forceinline__ double wmax(double v) {
    v = fmax(v, dpp<kDppXor1>(v));
    v = fmax(v, dpp<kDppXor2>(v));
    v = fmax(v, dpp<kDppHalfMirror>(v));
    v = fmax(v, dpp<kDppMirror>(v));
    return fmax(fmax(rdlane(v, 0), rdlane(v, 16)), fmax(rdlane(v, 32), rdlane(v, 48)));
}
// Maximum over lanes 0 .. N-1 of v (v = 0 in lanes N .. 63): each row of 16
// lanes reduces by DPP, then the rows by the row broadcasts of lanes 15 and
// 31 (rows 1 and 3 take row 0 and 2's maximum, rows 2 and 3 then lane 31's),
// which leave the maximum of rows 0 .. r in row r.
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;
template <int N>
__device__ __forceinline__ unsigned wmax_u32_rows(unsigned v) {
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, kDppXor1, 0xf, 0xf, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, kDppXor2, 0xf, 0xf, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, kDppHalfMirror, 0xf, 0xf, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, kDppMirror, 0xf, 0xf, false));
    if (N > 16) v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, kDppRowBcast15, 0xa, 0xf, false));
    if (N > 32) v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, kDppRowBcast31, 0xc, 0xf, false));
    return __builtin_amdgcn_readlane(v, (N - 1) | 15);
}

// LDS floating-point add (ds_add_f64), result unused.
__device__ __forceinline__ void lds_add(double *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The general power, kept out of line: the kernels reach it only for
// non-integer reaction orders / exponents, and inlining its dozens of call
// sites would multiply the code the hot loop drags through the I-cache.
__device__ __noinline__ double pow_ool(double x, double p) { return pow(x, p); }

// x^p for a run-time exponent with the integer fast paths of cpow<P>.
__device__ __forceinline__ double rpow(double x, double p) {
    if (p == 1.0) return x;
    if (p == 2.0) return x * x;
    if (p == 0.0) return 1.0;
    if (p == 3.0) return x * x * x;
    return pow_ool(x, p);
}
__device__ __forceinline__ double xpow(const TFac &tf, double s) {
    if (s == 1.0) return tf.X;
    if (s == -1.0) return tf.Xinv;
    if (s == 2.0) return tf.X * tf.X;
    if (s == -2.0) return tf.Xinv * tf.Xinv;
    return pow_ool(tf.X, s);
}
// ArrheniusRR::k: A T^b exp(-Ta/T); with b = 0 or Ta = 0 the single exp gives
// the same value as the reference's skipped pow / exp.  F = float: the
// split-precision mode's FP32 rate constant.
template <class F = double>
__device__ __forceinline__ double rarrh(double A, double b, double Ta, const TFac &tf) {
    if constexpr (std::is_same_v<F, float>)
        return (double)((float)A * __expf((float)b * (float)tf.lnT - (float)Ta * (float)tf.invT));
    else
        return A * exp(b * tf.lnT - Ta * tf.invT);
}
__device__ __forceinline__ double rarrh_dkdT(double k, double b, double Ta, const TFac &tf) {
    return k * tf.invTm * (b + Ta * tf.invTm);
}

// Fall-off broadening (fallOffFunctions.cpp:54-168: F, DFDci, DFDT) for a
// run-time type and parameters, evaluated once per reaction together with
// the pieces both derivatives reuse:
//   dF/dc_i / F = gPr dPr/dc_i,   dF/dT / F = gT + gPr dPr/dT
// (the reference's DFDci and DFDT divided by F; both are linear in the
// derivative of Pr).  The published Troe and SRI forms as the reference
// writes them, with log10 x = ln x / ln 10, 10^x = exp(x ln 10) and
// a^X b^e = exp(X ln a + e ln b): one logarithm of Pr and of the centre /
// SRI base shared by F and its derivatives.  Both forms are a weighted sum of
// three exponentials in T, its logarithm and one final exponential:
//   Troe  Fcent = (1 - alpha) exp(-T/T3) + alpha exp(-T/T1) [+ exp(-T2/T)]
//   SRI   base  = exp(-T/c) + a exp(-b/T)
// so every lane of a wavefront evaluates the same five transcendentals with
// its own weights and exponents, and a chunk that mixes the forms pays for
// one set, not for each form's.
struct FoState {
    double F, gPr, gT;
};
template <bool WITH_J>
__device__ __forceinline__ FoState rfo_eval(int fot, const double *fp, const TFac &tf, double Pr) {
    constexpr double kLn10 = 2.302585092994045684;
    constexpr double kRln10 = 0.4342944819032518277;  // 1 / ln 10
    FoState o{1.0, 0.0, 0.0};
    if (fot == 0) return o;  // Lindemann
    const double T = tf.T;
    const double rPr = frcp(Pr);
    const double lnPr = log(fmax(Pr, kTiny));
    const double logPr = lnPr * kRln10;
    const bool sri = fot == 3, t4 = fot == 2;
    // exp(-T / p1) and exp(-T / p2) weighted w1, w2; exp(-p3 / T) weighted w3
    const double p1 = sri ? fp[2] : fp[1], p2 = fp[2];
    const double w1 = sri ? 1.0 : 1.0 - fp[0], w2 = sri ? 0.0 : fp[0];
    const double p3 = sri ? fp[1] : (t4 ? fp[3] : 0.0), w3 = sri ? fp[0] : (t4 ? 1.0 : 0.0);
    const double u1 = exp(-T / p1), u2 = exp(sri ? 0.0 : -T / p2), u3 = exp(-p3 * tf.invT);
    double base = w1 * u1 + w2 * u2;
    base += w3 * u3;
    if (!sri) base = fmax(base, kTiny);
    const double lnb = log(base);
    // Troe: F = 10^(log10 Fcent / (1 + inv^2)), inv = (log10 Pr + c) / (n - d (log10 Pr + c))
    const double logFc = lnb * kRln10;
    const double c = -0.4 - 0.67 * logFc, n = 0.75 - 1.27 * logFc, dd = 0.14;
    const double lpc = logPr + c;
    const double rnmd = frcp(n - dd * lpc);
    const double inv = lpc * rnmd;
    const double rq = frcp(1.0 + inv * inv);
    // SRI: F = d base^X T^e, X = 1 / (1 + log10(Pr)^2)
    const double X = frcp(1.0 + logPr * logPr);
    const double E = sri ? X * lnb + fp[4] * tf.lnT : lnb * rq;
    o.F = (sri ? fp[3] : 1.0) * exp(E);
    if (WITH_J) {
        // (d base / dT) / base
        double db = w1 * u1 / (-p1) + w2 * u2 / (-p2);
        db += w3 * u3 * (p3 * tf.invT * tf.invT);
        db *= frcp(base);
        // Troe: d(log10 F)/d(log10 Pr + c) and the T derivatives of log10
        // Fcent, c, n
        const double k = -logFc * rq * rq * 2.0 * lpc * rnmd * rnmd;
        const double dlF = db * kRln10;
        const double dc = -0.67 * dlF, dn = -1.27 * dlF;
        const double gPr_t = k * (1.0 + dd * inv) * rPr;
        const double gT_t = kLn10 * (dlF * rq + k * (dc * (1.0 + dd * inv) - inv * dn));
        const double gPr_s = -X * X * 2.0 * logPr * rPr * kRln10 * lnb;
        const double gT_s = fp[4] * tf.invT + X * db;
        o.gPr = sri ? gPr_s : gPr_t;
        o.gT = sri ? gT_s : gT_t;
    }
    return o;
}

// Molar G, dG/dT, Cp, H, dCp/dT of species k (JANAF.cpp:142-219), same
// operation order as species_G / species_CH of the lane kernels.  F = float:
// G and dG/dT in FP32 (split-precision mode; Cp, H, dCp/dT stay FP64).
template <class F = double>
__device__ __forceinline__ void wthermo(const WMech &M, int k, const TFac &tf, double &G,
                                        double &dG, double &Cp, double &H, double &dCp) {
    const double *S = M.sp;
    const double T = tf.T;
    const bool lo = T < S[SP_TCOM * 64 + k];
    const int oa = lo ? SP_ALO : SP_AHI, oh = lo ? SP_HLO : SP_HHI, os = lo ? SP_SLO : SP_SHI,
              og = lo ? SP_GLO : SP_GHI;
    double a[7];
#pragma unroll
    for (int c = 0; c < 7; ++c) a[c] = S[(oa + c) * 64 + k];
    const double h4 = S[(oh + 0) * 64 + k], h3 = S[(oh + 1) * 64 + k], h2 = S[(oh + 2) * 64 + k],
                 h1 = S[(oh + 3) * 64 + k];
    const double s4 = S[(os + 0) * 64 + k], s3 = S[(os + 1) * 64 + k], s2 = S[(os + 2) * 64 + k];
    const double g4 = S[(og + 0) * 64 + k], g3 = S[(og + 1) * 64 + k], g2 = S[(og + 2) * 64 + k],
                 g1 = S[(og + 3) * 64 + k];
    const double W = S[SP_W * 64 + k], Tl = S[SP_TLOW * 64 + k], Th = S[SP_THIGH * 64 + k];
    const bool out = (T < Tl) || (T > Th);
    const double T0 = out ? fmin(fmax(T, Tl), Th) : T;
    const double cp0 = ((((a[4] * T0 + a[3]) * T0 + a[2]) * T0 + a[1]) * T0 + a[0]);
    if constexpr (std::is_same_v<F, float>) {
        const float t = (float)T, t0 = (float)T0, lnT = (float)tf.lnT, iTm = (float)tf.invTm;
        float b[7];
#pragma unroll
        for (int c = 0; c < 7; ++c) b[c] = (float)a[c];
        const float c0 = ((((b[4] * t0 + b[3]) * t0 + b[2]) * t0 + b[1]) * t0 + b[0]);
        const float s0 = (((((float)s4 * t + (float)s3) * t + (float)s2) * t + b[1]) * t + b[0] * lnT + b[6]);
        const float g0 = (((((-(float)g4 * t - (float)g3) * t - (float)g2) * t - (float)g1) * t +
                           b[0] * (1.0f - lnT) - b[6]) * t + b[5]);
        const float ds0 = (b[0] * iTm + b[1] + (b[2] + (b[3] + b[4] * t) * t) * t);
        G = (double)(g0 * (float)W);
        dG = (double)((c0 - (s0 + t * ds0)) * (float)W);
    } else {
        const double s0 = ((((s4 * T + s3) * T + s2) * T + a[1]) * T + a[0] * tf.lnT + a[6]);
        const double g0 = (((((-g4 * T - g3) * T - g2) * T - g1) * T + a[0] * (1.0 - tf.lnT) - a[6]) * T + a[5]);
        const double ds0 = (a[0] * tf.invTm + a[1] + (a[2] + (a[3] + a[4] * T) * T) * T);
        G = g0 * W;
        dG = (cp0 - (s0 + T * ds0)) * W;
    }
    const double Tx = out ? (T < Tl ? Tl : Th) : T;
    double h = ((((h4 * Tx + h3) * Tx + h2) * Tx + h1) * Tx + a[0]) * Tx + a[5];
    if (out) h += cp0 * (T - Tx);
    const double dcp = (a[1] + T * (2.0 * a[2] + T * (3.0 * a[3] + 4.0 * a[4] * T)));
    Cp = cp0 * W;
    H = h * W;
    dCp = dcp * W;
}

// One reaction (lane-private): net rate q, and with J: dq/dT + dgamma/dT term
// s, the third-body coefficient gam (Jacobian term gam * eff_j), and the
// partial derivatives of q_f / q_r by each participant's concentration.
struct WRxn {
    double q, s, gam, dqf[kWP], dqr[kWP];
    int nf, nb, rev, fi[kWP], bi[kWP];
    double fnu[kWP], bnu[kWP];
};

// F = float: the split-precision mode -- Arrhenius rate constants and the
// equilibrium constant (Gibbs sums and exponential) in FP32; concentration
// products, rates, derivatives in FP64.
template <bool WITH_J, class FT = double>
__device__ __forceinline__ void wreaction(const WMech &M, int base, const TFac &tf,
                                          const double *sc, const double *sG, const double *sdG,
                                          int exact, WRxn &o) {
    const RxnTab tb = rxn_tab(M);
    const int lane = threadIdx.x, ch = base >> 6, r = base + lane;
    const unsigned hdr = ldu(tb, RH_HDR, ch, lane), ix = ldu(tb, RH_IDX, ch, lane),
                   ix2 = ldu(tb, RH_IDX2, ch, lane);
    const int nf = hdr & 3u, nb = (hdr >> 2) & 3u, rev = (hdr >> 4) & 3u, kind = (hdr >> 6) & 7u;
    const int fi[kWP] = {(int)(ix & 255u), (int)((ix >> 8) & 255u), (int)((ix >> 16) & 255u)};
    const int bi[kWP] = {(int)(ix >> 24), (int)(ix2 & 255u), (int)((ix2 >> 8) & 255u)};
    double fnu[kWP], bnu[kWP], fo[kWP], bo[kWP], cf[kWP], cb[kWP];
    if (hdr & kRhGen) {
#pragma unroll
        for (int i = 0; i < kWP; ++i) {
            fnu[i] = ldd(tb, RD_FNU + i, ch, lane);
            bnu[i] = ldd(tb, RD_BNU + i, ch, lane);
            fo[i] = ldd(tb, RD_FO + i, ch, lane);
            bo[i] = ldd(tb, RD_BO + i, ch, lane);
        }
    } else {
#pragma unroll
        for (int i = 0; i < kWP; ++i) {
            fnu[i] = ldf(tb, RH_FNU + i, ch, lane);
            bnu[i] = ldf(tb, RH_BNU + i, ch, lane);
            fo[i] = ldf(tb, RH_FO + i, ch, lane);
            bo[i] = ldf(tb, RH_BO + i, ch, lane);
        }
    }
#pragma unroll
    for (int i = 0; i < kWP; ++i) {
        cf[i] = i < nf ? sc[fi[i]] : 1.0;
        cb[i] = i < nb ? sc[bi[i]] : 1.0;
    }
    const double A = ldd(tb, RD_A, ch, lane), b = ldd(tb, RD_B, ch, lane),
                 Ta = ldd(tb, RD_TA, ch, lane);
    // third-body concentration, summed in species order (thirdBodyEfficiency::M)
    double Mtb = 0.0;
    if (kind != 0) {
#pragma unroll 8
        for (int k = 0; k < M.ns; ++k) Mtb += lde(tb, r, k) * sc[k];
    }
    double kf, k0 = 0.0, kinf = 0.0, Pr = 0.0, F = 1.0;
    FoState fs{1.0, 0.0, 0.0};
    double fp[5];
    const int fot = (hdr >> 9) & 7u;
    // one exponential for the three rate laws (divergent branches would
    // each evaluate their own)
    const double kA = rarrh<FT>(A, b, Ta, tf);
    if (kind == 0) {
        kf = kA;
    } else if (kind == 1) {
        kf = Mtb * kA;
    } else {
#pragma unroll
        for (int i = 0; i < 5; ++i) fp[i] = ldd(tb, RD_FOP + i, ch, lane);
        k0 = rarrh<FT>(ldd(tb, RD_A0, ch, lane), ldd(tb, RD_B0, ch, lane),
                       ldd(tb, RD_TA0, ch, lane), tf);
        kinf = kA;
        const double kinfz = kinf != 0 ? kinf : kVeryTiny;
        Pr = k0 / kinfz * Mtb;
        fs = rfo_eval<WITH_J>(fot, fp, tf, Pr);
        F = fs.F;
        kf = kind == 2 ? kinf * (Pr / (1.0 + Pr)) * F : k0 * (1.0 / (1.0 + Pr)) * F;
    }
    double kr = 0.0, rK = 0.0, dKc = 0.0;
    if (rev == 1) {
        // reactionTypes::Kp / Kc / DKcDT (reactionTypes.cpp:178-241)
        FT Kt = 0.0, dk = 0.0;
        const FT iTm = (FT)tf.invTm;
#pragma unroll
        for (int i = 0; i < kWP; ++i) {
            if (i < nf) {
                const FT g = (FT)sG[fi[i]];
                Kt -= g * (FT)fnu[i];
                if (WITH_J) dk -= ((FT)sdG[fi[i]] - g * iTm) * (FT)fnu[i];
            }
        }
#pragma unroll
        for (int i = 0; i < kWP; ++i) {
            if (i < nb) {
                const FT g = (FT)sG[bi[i]];
                Kt += g * (FT)bnu[i];
                if (WITH_J) dk += ((FT)sdG[bi[i]] - g * iTm) * (FT)bnu[i];
            }
        }
        Kt *= (FT)tf.invRuT;
        double Kp;
        if constexpr (std::is_same_v<FT, float>) Kp = (double)__expf(-Kt);
        else Kp = exp(-Kt);
        if (bad(Kp)) Kp = kRootVeryLarge;
        const double sS = (hdr & kRhGen) ? ldd(tb, RD_SUMS, ch, lane) : ldf(tb, RH_SUMS, ch, lane);
        const bool zero_s = fabs(sS) < kVeryTiny;
        double Xs = 1.0, Kc;
        if (zero_s) {
            Kc = Kp;
        } else {
            Xs = xpow(tf, sS);
            Kc = fmin(Kp * Xs, kVeryLarge);
        }
        rK = frcp(fmax(Kc, kVeryTiny));
        kr = Kc > kVeryTiny ? kf * rK : 0.0;
        if (WITH_J) {
            dk *= (FT)tf.invRuT;
            if (exact) dk = -dk;
            double dKp = Kp * (double)dk;
            if (bad(dKp)) dKp = kVeryLarge;
            if (zero_s) {
                dKc = dKp;
            } else {
                double v = (dKp + Kp * (-sS) * tf.invT) * Xs;
                if (bad(v)) v = kVeryLarge;
                dKc = v;
            }
        }
    } else if (rev == 2) {
        kr = rarrh<FT>(ldd(tb, RD_AR, ch, lane), ldd(tb, RD_BR, ch, lane),
                       ldd(tb, RD_TAR, ch, lane), tf);
        if (kind == 1) kr = Mtb * kr;
    }
    // concentration powers c^ord of every participant, computed once and
    // reused by q, dq/dc and dq/dT (absent participants contribute 1)
    double pf[kWP], pb[kWP];
#pragma unroll
    for (int i = 0; i < kWP; ++i) {
        pf[i] = i < nf ? rpow(cf[i], fo[i]) : 1.0;
        pb[i] = i < nb ? rpow(cb[i], bo[i]) : 1.0;
    }
    const double prf = pf[0] * pf[1] * pf[2], prb = pb[0] * pb[1] * pb[2];
    const double qf = kf * prf, qr = kr * prb;
    const double q = qf - qr;
    o.q = q;
    o.nf = nf;
    o.nb = nb;
    o.rev = rev;
#pragma unroll
    for (int i = 0; i < kWP; ++i) {
        o.fi[i] = fi[i];
        o.bi[i] = bi[i];
        o.fnu[i] = fnu[i];
        o.bnu[i] = bnu[i];
    }
    if (!WITH_J) return;
    // DqDci (reactionTypes.cpp:354-445): d(c_i^ord)/dc_i times the other
    // participants' powers
#pragma unroll
    for (int i = 0; i < kWP; ++i) {
        auto dpow = [](double c, double ord) {
            if (ord == 1.0) return 1.0;
            if (ord < 1.0) return c > kTiny ? ord * pow_ool(c, ord - 1.0) : 0.0;
            return ord * rpow(c, ord - 1.0);
        };
        const int i1 = (i + 1) % kWP, i2 = (i + 2) % kWP;
        o.dqf[i] = i < nf ? kf * dpow(cf[i], fo[i]) * (pf[i1] * pf[i2]) : 0.0;
        o.dqr[i] = i < nb ? kr * dpow(cb[i], bo[i]) * (pb[i1] * pb[i2]) : 0.0;
    }
    // third-body / fall-off terms q dln(gamma)/dc_j = gam * eff_j
    double gam = 0.0;
    if (kind == 1) {
        gam = q * (1.0 / fmax(Mtb, kTiny));
    } else if (kind >= 2 && Mtb > kTiny) {
        const double kinfz = kinf != 0 ? kinf : kVeryTiny;
        const double dPr = k0 / kinfz;
        const double dlF = fs.gPr * dPr;  // (dF/dc) / F
        const double g = kind == 2 ? dPr / (Pr * (1 + Pr)) + dlF : -dPr / (1 + Pr) + dlF;
        gam = bad(g) ? 0.0 : q * g;
    }
    o.gam = gam;
    // DqDT (reactionTypes.cpp:843-878)
    double dkf, dkr = 0.0;
    if (kind <= 1) dkf = rarrh_dkdT(kf, b, Ta, tf);
    else if (kind == 2) dkf = (Pr / (1.0 + Pr)) * F * rarrh_dkdT(kinf, b, Ta, tf);
    else
        dkf = (1.0 / (1.0 + Pr)) * F *
              rarrh_dkdT(k0, ldd(tb, RD_B0, ch, lane), ldd(tb, RD_TA0, ch, lane), tf);
    if (rev == 1) dkr = (dkf - kr * dKc) * rK;
    else if (rev == 2)
        dkr = rarrh_dkdT(kr, ldd(tb, RD_BR, ch, lane), ldd(tb, RD_TAR, ch, lane), tf);
    dkf *= prf;
    dkr *= prb;
    double dqdT = dkf - dkr, dgdT = 0.0;
    if (kind >= 2 && Mtb > kTiny) {
        const double k0z = k0 != 0 ? k0 : kVeryTiny, kinfz = kinf != 0 ? kinf : kVeryTiny;
        const double dPrdT =
            Pr * (rarrh_dkdT(k0, ldd(tb, RD_B0, ch, lane), ldd(tb, RD_TA0, ch, lane), tf) / k0z -
                  rarrh_dkdT(kinf, b, Ta, tf) / kinfz);
        const double dlFdT = fs.gT + fs.gPr * dPrdT;  // (dF/dT) / F
        const double gg = kind == 2 ? dPrdT / (Pr * (1 + Pr)) + dlFdT : -dPrdT / (1 + Pr) + dlFdT;
        dgdT = q * gg;
    }
    if (bad(dqdT)) dqdT = 0.0;
    if (bad(dgdT)) dgdT = 0.0;
    o.s = dqdT + dgdT;
}

// LDS work area of one wavefront.
template <int N>
struct WaveLDS {
    // row stride of the Jacobian: odd, so that a lane-per-row access pattern
    // (lane i reading row i) hits distinct LDS banks.  Every array is sized
    // for the N unknowns (not the 64 lanes) so that more wavefronts fit in
    // a CU's LDS.
    static constexpr int JS = N | 1;
    double J[N * JS];
    // species Cp and dCp/dT are read only by their own lane: above N = 40
    // they stay in registers, so that the 55-row kernel fits six wavefronts
    // in a CU's LDS
    static constexpr bool kCpLds = N <= 40;
    double c[N], G[N], dG[N], H[N];  // species k in [k]
    double Cp[kCpLds ? N : 1], dCp[kCpLds ? N : 1];
    double w[N];      // net production rates omega_k
#if OHC_LU_BCAST == 0
    // pivot row of the current elimination step (element j at [2 j], b at
    // [2 N]): the stride keeps the writes single 64-bit stores straight from
    // the row's registers instead of packed 128-bit stores that need copies
    double prow[2 * N + 1];
#endif
};

// dy/dt (and the Jacobian into L.J when WITH_J) at the wavefront-distributed
// state yv (lane i holds y_i; lane ns holds T).  Returns f_i in lane i.
template <int N, int FLOW, bool WITH_J, class F = double>
__device__ __forceinline__ double wave_eval(const WMech &M, WaveLDS<N> &L, double yv, int exact,
                                            double *cp_lane = nullptr) {
    const int lane = threadIdx.x, ns = M.ns;
    const double T = rdlane(yv, ns);
    const TFac tf = tfac(T);
    double cpr = 0.0, dcpr = 0.0;  // this lane's Cp and dCp/dT when not in LDS
    if (lane < ns) {
        L.c[lane] = fmax(yv, 0.0);
        double G, dG, Cp, H, dCp;
        wthermo<F>(M, lane, tf, G, dG, Cp, H, dCp);
        L.G[lane] = G;
        L.dG[lane] = dG;
        L.H[lane] = H;
        if constexpr (WaveLDS<N>::kCpLds) {
            L.Cp[lane] = Cp;
            L.dCp[lane] = dCp;
        } else {
            cpr = Cp;
            dcpr = dCp;
        }
    }
    if (WITH_J)
        for (int e = lane; e < N * WaveLDS<N>::JS; e += 64) L.J[e] = 0.0;
    if (lane < N) L.w[lane] = 0.0;
    __syncthreads();
    for (int base = 0; base < M.nr; base += 64) {
        const int r = base + lane;
        double gam = 0.0;
        WRxn o;
        o.nf = o.nb = 0;
        if (r < M.nr) {
            wreaction<WITH_J, F>(M, base, tf, L.c, L.G, L.dG, exact, o);
            // d(omega)/dt rows of the participants: sigma_a q, and with J the
            // row terms sigma_a dq/dc_j and sigma_a dq/dT
#pragma unroll
            for (int side = 0; side < 2; ++side) {
#pragma unroll
                for (int i = 0; i < kWP; ++i) {
                    const bool has = side == 0 ? i < o.nf : i < o.nb;
                    if (!has) continue;
                    const int a = side == 0 ? o.fi[i] : o.bi[i];
                    const double nu = side == 0 ? -o.fnu[i] : o.bnu[i];
                    lds_add(&L.w[a], nu * o.q);
                    if (WITH_J) {
                        double *row = L.J + a * WaveLDS<N>::JS;
#pragma unroll
                        for (int j = 0; j < kWP; ++j) {
                            if (j < o.nf) lds_add(&row[o.fi[j]], nu * o.dqf[j]);
                            // (dq/dc of the products is zero for an
                            // irreversible reaction: kr = 0)
                            if (j < o.nb && o.rev != 0) lds_add(&row[o.bi[j]], -nu * o.dqr[j]);
                        }
                        lds_add(&row[ns], nu * o.s);
                    }
                }
            }
            gam = o.gam;
        }
        if (WITH_J) {
            // third-body / fall-off terms sigma_a gam_r eff_rj, column j in lane
            // j, one reaction of the chunk after another.  The reaction lane
            // first folds gam into its participants' coefficients; the loop
            // then needs per participant only its row and coefficient from
            // that lane (v_readlane, wave-uniform) and one LDS atomic add per
            // column lane, with no read-modify-write round trip.  The
            // wavefront's LDS operations complete in program order, so every
            // element receives the same sums in the same order as the
            // participant rows above.
            unsigned long long tbm = __ballot(gam != 0.0);
            if (tbm) {
#pragma unroll
                for (int i = 0; i < kWP; ++i) {
                    o.fnu[i] = -o.fnu[i] * gam;
                    o.bnu[i] = o.bnu[i] * gam;
                }
            }
            const RxnTab tb = rxn_tab(M);
            while (tbm) {
                const int sl = __ffsll((long long)tbm) - 1;
                tbm &= tbm - 1ull;
                const double e = lde_col(tb, base + sl, lane);  // eff_rj, zero past ns
                const int nf = __builtin_amdgcn_readlane(o.nf, sl),
                          nb = __builtin_amdgcn_readlane(o.nb, sl);
#pragma unroll
                for (int i = 0; i < kWP; ++i) {
                    if (i < nf) {
                        const int k = __builtin_amdgcn_readlane(o.fi[i], sl);
                        const double v = rdlane(o.fnu[i], sl) * e;
                        if (lane < ns) lds_add(&L.J[k * WaveLDS<N>::JS + lane], v);
                    }
                    if (i < nb) {
                        const int k = __builtin_amdgcn_readlane(o.bi[i], sl);
                        const double v = rdlane(o.bnu[i], sl) * e;
                        if (lane < ns) lds_add(&L.J[k * WaveLDS<N>::JS + lane], v);
                    }
                }
            }
        }
    }
    __syncthreads();
    const double w = lane < ns ? L.w[lane] : 0.0;
    // temperature equation (chemistrySource.cpp:863-907, jacobian :993-1054)
    const bool sp = lane < ns;
    double cpv = cpr, dcpv = dcpr;
    if constexpr (WaveLDS<N>::kCpLds) {
        cpv = sp ? L.Cp[lane] : 0.0;
        dcpv = sp ? L.dCp[lane] : 0.0;
    }
    if (cp_lane) *cp_lane = cpv;
    double Hx = 0.0, Cx = 0.0;
    if (sp) {
        Hx = FLOW == 0 ? L.H[lane] : L.H[lane] - kRu * T;
        Cx = FLOW == 0 ? cpv : cpv - kRu;
    }
    const double cC = wsum(sp ? yv * Cx : 0.0);
    const double dTs = wsum(sp ? w * Hx : 0.0);
    const double rcC = frcp(cC);
    const double dT = dTs * rcC;
    if (WITH_J) {
        const double dcC = wsum(sp ? yv * dcpv : 0.0);
        double jt = 0.0;
        if (sp) {
            double v = 0.0;
#pragma unroll 8
            for (int j = 0; j < ns; ++j) {
                const double hx = FLOW == 0 ? L.H[j] : L.H[j] - kRu * T;
                v -= L.J[j * WaveLDS<N>::JS + lane] * hx;
            }
            v += Cx * dT;
            jt = v * rcC;
        }
        const double jT = sp ? L.J[lane * WaveLDS<N>::JS + ns] : 0.0;  // d(omega_k)/dT
        const double tt = wsum(sp ? (Cx * w + Hx * jT) : 0.0);
        __syncthreads();
        if (sp) L.J[ns * WaveLDS<N>::JS + lane] = jt;
        if (lane == 0) L.J[ns * WaveLDS<N>::JS + ns] = (dcC * dT - tt) * rcC;
        __syncthreads();
    }
    return sp ? w : (lane == ns ? -dT : 0.0);
}

// Solve (diag I - h J) x = b with the Jacobian rows in LDS: lane i owns row i;
// partial pivoting as SquareMatrix::LUDecomposeDPivoting (largest |pivot| to
// 2^-20 relative, lowest row among those), forward substitution folded into the
// elimination.  Returns x_i in lane i.  The elimination steps are unrolled at
// compile time, so step k touches only the N - k - 1 columns right of the
// pivot; FULL (n == N) drops the per-step bounds, so that the whole
// factorisation is one straight-line block the compiler can schedule across
// steps.  The pivot row reaches the other lanes as OHC_LU_BCAST selects; each
// pivot row keeps its step and reciprocal in two registers for the back
// substitution.
template <int N, bool FULL>
__device__ __forceinline__ double wave_lu_solve_n(WaveLDS<N> &L, int n, double h, double b,
                                                  double diag) {
    const int lane = threadIdx.x;
    if (FULL) n = N;
    const bool act = lane < n;
    // every lane loads a row that exists (lanes past the system read row 0),
    // so the N loads issue back to back without a branch and a wait each.
    // With n = N the lanes past the system keep row 0's values: they are
    // never pivots, so no other lane reads them, and their b and x are not
    // used.
    const double *Jr = L.J + (act ? lane : 0) * WaveLDS<N>::JS;
    double a[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        // -h J, then + diag on the diagonal: two roundings, as the reference
        // forms I - h J (ImplicitEuler.cpp:186-193), never one fused multiply-add
#pragma clang fp contract(off)
        const double x = Jr[j];
        const double v = FULL || (act && j < n) ? -h * x : 0.0;
        const bool dj = __builtin_amdgcn_inverse_ballot_w64(1ull << j);  // lane == j
        a[j] = ((FULL || act) && dj) ? v + diag : v;
    }
    // rows not yet used as a pivot, as a lane mask: the per-row predicates
    // of an elimination step are scalar mask operations
    unsigned long long avail = __builtin_amdgcn_ballot_w64(act);
    int stepv = N;       // the elimination step at which this row was the pivot
    double rpv = 0.0;    // and 1 / its pivot
    bool nan_col = false;  // a step found only NaN in its column
    sfor<0, N>([&](auto kc) {
        constexpr int k = kc;
        if (FULL || k < n) {
            // largest |a_ik| over the rows not yet used by one 32-bit
            // reduction of the high words of |a_ik| (exponent and the top 20
            // mantissa bits: the order of non-negative doubles to 2^-20
            // relative), the lowest row among those that reach it.  Rows
            // whose |a_ik| differ by less than that are equally good pivots;
            // the Jacobian's own rounding (its atomic sums) already decides
            // between them at that level.
            const double av = fabs(a[k]);
            const unsigned long long candm = avail & __builtin_amdgcn_ballot_w64(!isnan(av));
            const unsigned hi =
                __builtin_amdgcn_inverse_ballot_w64(candm) ? (unsigned)__double2hiint(av) : 0u;
            const unsigned mhi = wmax_u32_rows<N>(hi);
            const unsigned long long hit = candm & __builtin_amdgcn_ballot_w64(hi == mhi);
            const int p = hit ? (int)__ffsll((long long)hit) - 1 : 0;  // no hit only for NaN rows
            nan_col = nan_col || hit == 0;
            const unsigned long long pm = 1ull << p;
            const double pk = rdlane(a[k], p);
            const double rp = frcp(pk);
            // the pivot row keeps its step and reciprocal for the back
            // substitution (register selects: a masked LDS store would split
            // the straight-line block)
            const bool isp = __builtin_amdgcn_inverse_ballot_w64(pm);
            stepv = isp ? k : stepv;
            rpv = isp ? rp : rpv;
            avail &= ~pm;
            // (an eliminated row's a_ik is not needed again: the back
            // substitution reads column k only in the rows pivoted before
            // step k, and the multiplier goes straight into the update)
            const double l = __builtin_amdgcn_inverse_ballot_w64(avail) ? a[k] * rp : 0.0;
            // columns j >= n hold zeros in every row and stay zero, so the
            // update needs no per-column predicate
#if OHC_LU_BCAST == 2
            const int p4 = p * 4;
            sfor<k + 1, N>([&](auto j) { a[j] = fma(-l, bperm(a[j], p4), a[j]); });
            b = fma(-l, bperm(b, p4), b);
#elif OHC_LU_BCAST == 1
            sfor<k + 1, N>([&](auto j) { a[j] = fma(-l, rdlane(a[j], p), a[j]); });
            b = fma(-l, rdlane(b, p), b);
#else
            if (__builtin_amdgcn_inverse_ballot_w64(pm)) {
                sfor<k + 1, N>([&](auto j) { L.prow[2 * j] = a[j]; });
                L.prow[2 * N] = b;
            }
            // (the block is one wavefront, whose LDS accesses complete in
            // program order: the reads below see the row, and the next
            // step's row cannot overwrite them)
            __builtin_amdgcn_wave_barrier();
            sfor<k + 1, N>([&](auto j) { a[j] = fma(-l, L.prow[2 * j], a[j]); });
            b = fma(-l, L.prow[2 * N], b);
#endif
        }
    });
    double xo = 0.0;  // the unknown of this row's step, set at that step
    sfor<0, N>([&](auto kk) {
        constexpr int k = N - 1 - kk;
        if (FULL || k < n) {
            // x_k = b / pivot in the row pivoted at step k
            const bool own = stepv == k;
            const unsigned long long pkm = __builtin_amdgcn_ballot_w64(own);
            const int pk = pkm ? (int)__ffsll((long long)pkm) - 1 : 0;
            const double t = b * rpv;
            const double xk = rdlane(t, pk);
            // rows pivoted at step k or later have given their x already
            // (or give it now), so their b may change freely
            b = fma(-a[k], xk, b);
            xo = own ? t : xo;
        }
    });
    // x_k to lane k: every pivot row sends its unknown to the lane of its
    // step (one LDS permute); the lanes past the system, never pivots, keep
    // their own 0
    const int dst = (stepv < n ? stepv : lane) * 4;
    const int lo = __builtin_amdgcn_ds_permute(dst, __double2loint(xo));
    const int hi = __builtin_amdgcn_ds_permute(dst, __double2hiint(xo));
    // a column whose remaining rows are all NaN leaves steps without a
    // pivot row: the solution is NaN, as the reference's division by a NaN
    // pivot makes it (the caller then rejects the Newton step)
    return nan_col ? __builtin_nan("") : __hiloint2double(hi, lo);
}
template <int N>
__device__ __forceinline__ double wave_lu_solve(WaveLDS<N> &L, int n, double h, double b,
                                                double diag = 1.0) {
    return n == N ? wave_lu_solve_n<N, true>(L, n, h, b, diag)
                  : wave_lu_solve_n<N, false>(L, n, h, b, diag);
}

// Persistent integrator: each wavefront takes one cell at a time from the
// queue and runs chemistrySource::solve / solveUpdate on it (same control
// flow as ie_lane_kernel, uniform across the wavefront).
// METH 0: ImplicitEuler (ImplicitEuler.cpp:40-342); METH 2: ImplicitTrapezoidal
// (ImplicitTrapezoidal.cpp:99-335) -- Newton on y - y0 - h/2 (f0 + f(y)) with
// I - h/2 J, error from the second difference of f times h^3/12, and f at
// the step start (reused from the accepted attempt's error estimate, the same
// state, instead of a fresh evaluation); METH 3: BDF23 at its default order 2
// (BDF23.cpp:54-509) -- variable-step BDF2 (order 1 on a solve's first step)
// on a0 y + a1 y_n + a2 y_n-1 = h f(y), first step min(0.1 tn, 1e-8), doubled
// on steps 4 and 5, error y - y0 - h f(y) on the first three steps and the
// third difference of the step-start states after, and an attempt whose error
// grew from the previous attempt's but stays below 5 accepted as it is.
// F = float: the split-precision mode (FP32 Gibbs energies, equilibrium and
// Arrhenius rate constants; FP64 concentration products, net rates,
// Jacobian, LU, error norms and sums), ImplicitEuler only.
template <int N, int FLOW, int METH = 0, class F = double>
__global__ __launch_bounds__(64, (N <= 40 ? 3 : 2)) void ie_wave_kernel(const WaveParams WP) {
    const LaneParams &P = WP.L;
    const WMech &M = WP.M;
    const int lane = threadIdx.x, ns = M.ns, n = ns + 1;
    __shared__ WaveLDS<N> L;
    __shared__ unsigned long long s_cell;
    const bool sp = lane < ns;
    const double W = sp ? M.sp[SP_W * 64 + lane] : 1.0;
    const double errg = pow(P.maxsc / P.safe, -1.0 / P.alpha);
    long long c_cells = 0, c_steps = 0, c_att = 0, c_newt = 0, c_rhs = 0, c_maxs = 0, c_fail = 0;
    double dtmin = __builtin_huge_val();
    for (;;) {
        if (lane == 0) s_cell = atomicAdd(P.counter, 1ull);
        __syncthreads();
        const unsigned long long qi = s_cell;
        __syncthreads();
        if (qi >= (unsigned long long)P.n) break;
        const long long cell = P.order ? (long long)P.order[qi] : (long long)qi;
        const double T0 = P.T[cell];
        const double dtc = P.dt[cell] * P.dtFactor;
        const double sub_in = P.subDt[cell];
        if (T0 < P.Tthr) {
            if (lane == 0) {
                P.subDt[cell] = dtc;
                if (P.cost) P.cost[cell] = 0;
            }
            if (P.mode == 0 && sp) P.R[lane * P.ldR + cell] = 0.0;
            continue;
        }
        const double rho = P.rho[cell];
        const double pr = FLOW == 0 ? P.p[cell] : rho;
        const double c0 = sp ? rho * P.Y[lane * P.ldY + cell] / W : 0.0;
        double y = sp ? c0 : (lane == ns ? T0 : 0.0);  // chemistry state, lane i = y_i
        double subDt = fmin(dtc, fmin(sub_in > 0.0 ? sub_in : dtc, P.dt_max));
        double left = dtc;
        long long csteps = 0, citer = 0;
        bool failed = false;
        // ImplicitTrapezoidal: f at the step start (valid when fnext_ok), f at
        // the start of the previous step
        double fnext = 0.0, flast = 0.0;
        bool fnext_ok = false;
        // chemistrySource::solve: while (leftTime > 0) ODE solves of the remaining interval
        while (left > 0.0 && !failed) {
            double tn = left;
            // ImplicitEuler::solve(t0, tn, dt0, y)
            double t = 0.0, dt = METH == 3 ? subDt : 0.2 * subDt;
            bool last = false, done = false;
            double yl1 = y, lts0 = 0.0, lts1 = 0.0, lts2 = 0.0;
            double h0 = y, h1 = y, h2 = y;  // BDF23 step-start history (yklast_)
            for (int nstep = 0; nstep < P.max_steps; ++nstep) {
                const double y0 = y;
                int order = 0;
                if constexpr (METH == 3) {
                    if (nstep == 0) {
                        dt = fmin(0.1 * tn, 1e-8);
                        order = 1;
                        h0 = y;
                    } else if (nstep == 1) {
                        order = 2;
                        h1 = h0;
                        h0 = y;
                    } else {
                        order = 2;
                        if (nstep > 3 && nstep <= 5) dt *= 2.0;
                        h2 = h1;
                        h1 = h0;
                        h0 = y;
                    }
                }
                const double dtTry = dt;
                if ((t + dt - tn) * (t + dt) > 0) {
                    dt = tn - t;
                    last = true;
                } else if (t + dt == tn) {
                    last = true;
                }
                // adaptiveSolve
                double dtl = dt;
                lts2 = lts1;
                lts1 = lts0;
                double lastError = kLarge;  // BDF23
                int chk = 0;
                double yk, error;
                double f0 = 0.0, fk = 0.0;  // ImplicitTrapezoidal: f(y0), f(yk)
                bool last_conv = false;
                if constexpr (METH == 2) {
                    if (fnext_ok) {
                        f0 = fnext;
                    } else {
                        ++c_rhs;
                        f0 = wave_eval<N, FLOW, false>(M, L, y0, WP.exact_dkp);
                    }
                    fnext_ok = false;
                }
                for (;;) {
                    lts0 = dtl;
                    ++c_att;
                    // BDF23::setAlpha (orders 1 and 2) and the history term
                    double ba0 = 1.0, yh = 0.0;
                    if constexpr (METH == 3) {
                        if (order == 1) {
                            yh = -1.0 * h0;
                        } else {
                            const double D = lts0, D1 = lts1;
                            double a1 = -(D + D1) / (D * D1), a2 = D / (D1 * (D + D1));
                            double a0 = -(a1 + a2);
                            a0 *= D;
                            a1 *= D;
                            a2 *= D;
                            ba0 = a0;
                            yh = a2 * h1;
                            yh += a1 * h0;
                        }
                    }
                    // ImplicitEuler::solve(t, dt, y0, y): Newton iterations
                    yk = y0;
                    int count = 0;
                    double lastErr = 0.0, maxErr = 0.0;
                    bool conv = true;
                    for (;;) {
                        const double f = wave_eval<N, FLOW, true, F>(M, L, yk, WP.exact_dkp);
                        ++c_newt;
                        ++citer;
                        const double h = METH == 2 ? 0.5 * dtl : dtl;
                        double rhs, dg = 1.0;
                        if constexpr (METH == 2) {
                            rhs = -yk + y0 + 0.5 * dtl * (f0 + f);
                        } else if constexpr (METH == 3) {
                            rhs = -(ba0 * yk + yh - dtl * f);
                            dg = ba0;
                        } else {
                            rhs = -yk + y0 + dtl * f;
                        }
                        const double d = wave_lu_solve<N>(L, n, h, lane < n ? rhs : 0.0, dg);
                        if (__any(lane < n && bad(d))) {
                            conv = false;
                            maxErr = 1000.0;
                            break;
                        }
                        yk += d;
                        const double tol = P.atol + P.rtol * fabs(yk);
                        maxErr = wmax(lane < n ? fabs(d) * frcp(tol) : 0.0);
                        if (count > 0 && (maxErr / fmax(lastErr, kTiny)) > 2 && count > 5) {
                            conv = false;
                            maxErr = 1000.0;
                            break;
                        }
                        lastErr = maxErr;
                        ++count;
                        if (isnan(maxErr)) {
                            conv = false;
                            maxErr = 1000.0;
                            break;
                        }
                        if (count > 1000) {
                            conv = false;
                            break;
                        }
                        if (maxErr <= 1.0) break;
                        if (citer > P.max_cell_iters) {
                            conv = false;
                            failed = true;
                            break;
                        }
                    }
                    if (failed) break;
                    last_conv = conv;
                    if (!conv) {
                        error = maxErr;
                    } else {
                        double e;
                        if constexpr (METH == 3) {
                            if (nstep <= 2) {
                                ++c_rhs;
                                const double f = wave_eval<N, FLOW, false>(M, L, yk, WP.exact_dkp);
                                e = yk - y0 - dtl * f;
                            } else {
                                const double D = lts0, D1 = lts1, D2 = lts2;
                                const double c1 = 1.0 / D, c2 = (1.0 / D1) * (1.0 + D / D1),
                                             c3 = D / (D1 * D2), ltef = (D + D1) / 6.0;
                                e = (c1 * (yk - h0) - c2 * (h0 - h1) + c3 * (h1 - h2)) * ltef;
                            }
                        } else if constexpr (METH == 2) {
                            ++c_rhs;
                            fk = wave_eval<N, FLOW, false>(M, L, yk, WP.exact_dkp);
                            if (nstep == 0) {
                                e = yk - y0 - dtl * fk;
                            } else {
                                const double dtn = lts0, dtn1 = lts1;
                                const double c1 = 2.0 / (dtn * (dtn + dtn1));
                                const double c2 = 2.0 / (dtn1 * (dtn + dtn1));
                                const double LTEF = dtn * dtn * dtn / 12.0;
                                e = (c1 * (fk - f0) - c2 * (f0 - flast)) * LTEF;
                            }
                        } else if (nstep == 0) {
                            ++c_rhs;
                            const double f = wave_eval<N, FLOW, false, F>(M, L, yk, WP.exact_dkp);
                            e = yk - y0 - dtl * f;
                        } else {
                            const double dtn = lts0, dtn1 = lts1;
                            const double f1 = 2.0 / (dtn * (dtn + dtn1));
                            const double f2 = 2.0 / (dtn1 * (dtn + dtn1));
                            const double LTEF = dtn * dtn * 0.5;
                            e = (f1 * (yk - y0) - f2 * (y0 - yl1)) * LTEF;
                        }
                        const bool nan_y = __any(lane < n && isnan(yk));
                        const double tol = P.atol_inc + P.rtol_inc * fmax(fabs(y0), fabs(yk));
                        const double m = wmax(lane < n ? fabs(e) * frcp(tol) : 0.0);
                        error = nan_y ? 1.5 : m;
                    }
                    // checkNewY, then the step-size control of adaptiveSolve
                    const double y0T = rdlane(y0, ns), ykT = rdlane(yk, ns);
                    const bool ok = !((ykT - y0T) > P.Talpha * y0T);
                    if (!ok) dtl *= P.Tbeta;
                    if (!ok && chk < 100) {
                        ++chk;
                        if (dtl <= kVeryTiny) {
                            failed = true;
                            break;
                        }
                        continue;
                    }
                    if constexpr (METH == 3) {
                        if (lastError < error && error < 5.0) break;
                    }
                    if (error > 1.0) {
                        dtl *= fmax(P.safe * pow(error, -P.alpha), P.minsc);
                        if (dtl <= kVeryTiny) {
                            failed = true;
                            break;
                        }
                        lastError = error;
                        continue;
                    }
                    break;
                }
                if (failed) break;
                // accepted step
                t += dtl;
                yl1 = y0;
                y = yk;
                if constexpr (METH == 2) {
                    flast = f0;
                    // f at the new state, unless the attempt was accepted
                    // without an error estimate (Newton stopped at its cap)
                    fnext = fk;
                    fnext_ok = last_conv;
                }
                if (error > errg)
                    dt = fmin(fmax(P.safe * pow(error, -P.alpha), P.minsc), P.maxsc) * dtl;
                else
                    dt = fmax(P.safe * P.maxsc * dtl, dtl);
                ++c_steps;
                ++csteps;
                if ((t - tn) * tn >= 0) {
                    if (nstep > 0 && last) subDt = fmin(dtTry, tn);
                    done = true;
                    break;
                }
            }
            if (failed) break;
            if (!done) tn = t;  // step budget exhausted: time reached
            left -= tn;
        }
        if (failed) ++c_fail;
        // outputs (chemistrySource.cpp:1494-1609)
        if (lane == 0) {
            P.subDt[cell] = subDt;
            if (P.cost) P.cost[cell] = (int)(citer < 0x7fffffffll ? citer : 0x7fffffffll);
        }
        if (P.mode == 0) {
            if (sp) P.R[lane * P.ldR + cell] = (y - c0) * (1.0 / dtc) * W;
        } else {
            // ciToYi (chemistrySource.cpp:1222-1240)
            const double yy = sp ? fmax(0.0, y) : 0.0;
            const double rs = wsum(yy * W);
            const double rr = 1.0 / fmax(rs, kTiny);
            const double yv = yy * W * rr;
            const double ys = wsum(yv);
            const double ri = 1.0 / ys;
            if (sp) P.Yout[lane * P.ldY + cell] = ys != 1.0 ? yv * ri : yv;
            const double Tn = rdlane(y, ns);
            if (lane == 0) P.Tout[cell] = Tn;
        }
        ++c_cells;
        dtmin = fmin(dtmin, subDt);
        c_maxs = csteps > c_maxs ? csteps : c_maxs;
    }
    if (lane == 0) {
        atomicAdd(&P.stats[0], (unsigned long long)c_cells);
        atomicAdd(&P.stats[1], (unsigned long long)c_steps);
        atomicAdd(&P.stats[2], (unsigned long long)c_att);
        atomicAdd(&P.stats[3], (unsigned long long)c_newt);
        atomicAdd(&P.stats[4], (unsigned long long)c_rhs);
        atomicMax(&P.stats[5], (unsigned long long)c_maxs);
        if (dtmin < __builtin_huge_val())
            atomicMin(&P.stats[6], (unsigned long long)__double_as_longlong(dtmin));
        atomicAdd(&P.stats[7], (unsigned long long)c_fail);
    }
}

// ---- MTS explicit Euler (the reference GPU path's MTSEuler solver type:
// solveChemODEsMTS, CUDAChemistrySourceODEs.cu:684-778, with timescale /
// MTS_whichGroup / MTS_GroupDT :199-309).  The species are grouped by the
// decade of their time scale tci_k = min(200, 1 / (|d omega_k/d c_k| + 1e-20))
// relative to the flow step; group by group, fastest first, explicit Euler
// steps of 0.8 times the group's smallest time scale advance the whole state
// with the faster groups' rates frozen, until the group's increments are
// within atol + rtol |y| (the last group runs to the end of the step).  A
// group without species is skipped, the time-scale sums start at zero and a
// cell fails after max_steps steps (where the reference's device code leaves
// these undefined).

// d(omega_k)/d(c_k) of the mass-action terms, kf and kr held fixed
// (reactionTable::dWdciAllSpecies, reactionTableCUDA.cu:331-385), lane k.
template <int N>
__device__ __forceinline__ double wave_mts_diag(const WMech &M, WaveLDS<N> &L, double yv,
                                                int exact) {
    const int lane = threadIdx.x, ns = M.ns;
    const double T = rdlane(yv, ns);
    const TFac tf = tfac(T);
    if (lane < ns) {
        L.c[lane] = fmax(yv, 0.0);
        double G, dG, Cp, H, dCp;
        wthermo(M, lane, tf, G, dG, Cp, H, dCp);
        L.G[lane] = G;
        L.dG[lane] = dG;
    }
    if (lane < N) L.w[lane] = 0.0;
    __syncthreads();
    for (int base = 0; base < M.nr; base += 64) {
        const int r = base + lane;
        if (r < M.nr) {
            WRxn o;
            wreaction<true>(M, base, tf, L.c, L.G, L.dG, exact, o);
#pragma unroll
            for (int i = 0; i < kWP; ++i) {
                if (i < o.nf) {
                    const int a = o.fi[i];
                    double d = -o.fnu[i] * o.dqf[i];
#pragma unroll
                    for (int j = 0; j < kWP; ++j)
                        if (j < o.nb && o.bi[j] == a) d += o.bnu[j] * o.dqf[i];
                    lds_add(&L.w[a], d);
                }
                if (i < o.nb) {
                    const int a = o.bi[i];
                    double d = -o.bnu[i] * o.dqr[i];
#pragma unroll
                    for (int j = 0; j < kWP; ++j)
                        if (j < o.nf && o.fi[j] == a) d += o.fnu[j] * o.dqr[i];
                    lds_add(&L.w[a], d);
                }
            }
        }
    }
    __syncthreads();
    const double d = lane < ns ? L.w[lane] : 0.0;
    __syncthreads();
    return d;
}

__device__ __forceinline__ double wmin(double v) { return -wmax(-v); }

template <int N, int FLOW>
__global__ __launch_bounds__(64, (N <= 40 ? 3 : 2)) void mts_wave_kernel(const WaveParams WP) {
    const LaneParams &P = WP.L;
    const WMech &M = WP.M;
    const int lane = threadIdx.x, ns = M.ns, n = ns + 1;
    __shared__ WaveLDS<N> L;
    __shared__ unsigned long long s_cell;
    const bool sp = lane < ns;
    const double W = sp ? M.sp[SP_W * 64 + lane] : 1.0;
    constexpr double kNone = 1.0e300;
    long long c_cells = 0, c_steps = 0, c_maxs = 0, c_fail = 0;
    double dtmin = __builtin_huge_val();
    for (;;) {
        if (lane == 0) s_cell = atomicAdd(P.counter, 1ull);
        __syncthreads();
        const unsigned long long qi = s_cell;
        __syncthreads();
        if (qi >= (unsigned long long)P.n) break;
        const long long cell = P.order ? (long long)P.order[qi] : (long long)qi;
        const double T0 = P.T[cell];
        const double tn = P.dt[cell] * P.dtFactor;
        const double rho = P.rho[cell];
        const double c0 = sp ? rho * P.Y[lane * P.ldY + cell] / W : 0.0;
        double y = sp ? c0 : (lane == ns ? T0 : 0.0);
        // timescale / MTS_whichGroup / MTS_GroupDT
        const double dd = wave_mts_diag<N>(M, L, y, WP.exact_dkp);
        const double tci = sp ? fmin(200.0, 1.0 / (fabs(dd) + 1e-20)) : kNone;
        const double tmin = wmin(tci);
        const int Nm = tn > tmin ? (int)fmax(0.0, floor(log10(tn / tmin))) + 1 : 1;
        const int ig = sp ? Nm - mts_group(tn, tci) : Nm;  // group index, fastest = 0
        double dtDid = 0.0;
        long long nst = 0;
        bool failed = false;
        int im = (int)wmin(sp ? (double)ig : kNone);
        while (im < Nm && !failed) {
            const double dtG = wmin(sp && ig == im ? tci : kNone);
            while (dtDid < tn) {
                if (nst >= P.max_steps) {
                    failed = true;
                    break;
                }
                double dt = 0.8 * dtG;
                if (dtDid + dt > tn) dt = tn - dtDid;
                double f = wave_eval<N, FLOW, false>(M, L, y, WP.exact_dkp);
                if (sp && ig < im) f = 0.0;  // faster groups are frozen
                dtDid += dt;
                const double err = dt * f, ay0 = fabs(y);
                if (lane < n) y += err;
                const double tol = P.atol + P.rtol * fmax(fabs(y), ay0);
                const double e = wmax(sp && ig == im ? fabs(err) / tol : 0.0);
                ++nst;
                if (im != Nm - 1 && e < 1.0) break;
            }
            im = (int)fmin((double)Nm, wmin(sp && ig > im ? (double)ig : kNone));
        }
        if (lane == 0 && P.cost) P.cost[cell] = (int)(nst < 0x7fffffffll ? nst : 0x7fffffffll);
        if (P.mode == 0) {
            if (sp) P.R[lane * P.ldR + cell] = (y - c0) * (1.0 / tn) * W;
        } else {
            // device ciToYi (CUDAChemistrySourceODEs.cu:545-586): Y = c W / sum c W
            const double r = wsum(sp ? y * W : 0.0);
            if (sp) P.Yout[lane * P.ldY + cell] = y * W * (1.0 / r);
            const double Tn = rdlane(y, ns);
            if (lane == 0) P.Tout[cell] = Tn;
        }
        ++c_cells;
        c_steps += nst;
        if (failed) ++c_fail;
        c_maxs = nst > c_maxs ? nst : c_maxs;
        dtmin = fmin(dtmin, P.subDt[cell]);
    }
    if (lane == 0) {
        atomicAdd(&P.stats[0], (unsigned long long)c_cells);
        atomicAdd(&P.stats[1], (unsigned long long)c_steps);
        atomicAdd(&P.stats[2], (unsigned long long)c_steps);
        atomicAdd(&P.stats[4], (unsigned long long)c_steps);
        atomicMax(&P.stats[5], (unsigned long long)c_maxs);
        if (dtmin < __builtin_huge_val() && dtmin >= 0.0)
            atomicMin(&P.stats[6], (unsigned long long)__double_as_longlong(dtmin));
        atomicAdd(&P.stats[7], (unsigned long long)c_fail);
    }
}

// Source terms and their (coupled) diagonal, one wavefront per cell; same
// definitions as source_lane_kernel.
template <int N>
__global__ __launch_bounds__(64) void source_wave_kernel(const SourceParams P, const WMech M) {
    __shared__ WaveLDS<N> L;
    const int lane = threadIdx.x, ns = M.ns;
    const int64_t cell = blockIdx.x;
    if (cell >= P.n) return;
    const bool sp = lane < ns;
    const double T = P.T[cell], rho = P.rho[cell];
    const double W = sp ? M.sp[SP_W * 64 + lane] : 1.0;
    const double Yk = sp ? P.Y[lane * P.ldY + cell] : 0.0;
    const double y = sp ? rho * Yk / W : (lane == ns ? T : 0.0);
    double cpl;
    const double w = wave_eval<N, 0, true>(M, L, y, P.exact, &cpl);
    if (sp) P.R[lane * P.ldR + cell] = w * W;
    if (!P.D) return;
    const double Ri = kRu / W;
    const double ea = sp ? L.H[lane] / W - Ri * T : 0.0;
    const double cv = wsum(sp ? (cpl / W - Ri) * Yk : 0.0);
    const double b = 1.0 / (rho * cv);
    const double *row = L.J + lane * WaveLDS<N>::JS;
    if (!P.without_last) {
        if (sp) P.D[lane * P.ldD + cell] = row[lane] + row[ns] * W * (-ea) * b;
    } else {
        const int l = ns - 1;
        const double Wl = M.sp[SP_W * 64 + l];
        const double el = rdlane(ea, l);
        if (lane < l)
            P.D[lane * P.ldD + cell] = row[lane] - row[l] * (W / Wl) + row[ns] * W * (el - ea) * b;
    }
}

// The GPU path's calculateSourceTermsImp[Async][Hybrid] (calcChemSource-
// TermsKimDiagonal, CUDAChemSourceKimDiagonalKernel.cu:39-105, with
// reactionTable::qfr<dOmegaType>, reactionTableCUDA.inl:142-273), one
// wavefront per cell: R_k = omega_k W_k (FP64) and the mass-action diagonal
// d(omega_k)/d(c_k) of species k < ns-1 with the last species substituted
// (d c_ns / d c_k = -W_k / W_ns), accumulated and stored as DT (double, or
// float for the Hybrid variant: each term formed in FP64, rounded to float,
// summed in float).
struct KimParams {
    int64_t n;
    const double *T, *rho, *Y;
    int64_t ldY;
    double *R;
    int64_t ldR;
    void *D;  // DT [ns-1][ldD], nullable
    int64_t ldD;
    int exact;
};

// Batched dense solves A x = b (SquareMatrix::LUDecomposeDPivoting +
// LUBacksubstituteDPivoting): entry (i, j) of system m at A[(i n + j) ldA + m]
// (the layout of the Jacobian output), b_i / x_i at b[i ldb + m] / x[i ldx + m].
struct LuParams {
    int64_t nmat;
    int n;
    const double *A;
    int64_t ldA;
    const double *b;
    int64_t ldb;
    double *x;
    int64_t ldx;
};

// One wavefront per system: rows into the LDS matrix, then the integrators'
// LU with h = -1 and no diagonal shift, so that it factors A itself.
template <int N>
__global__ __launch_bounds__(64) void lu_wave_kernel(const LuParams P) {
    __shared__ WaveLDS<N> L;
    const int lane = threadIdx.x, n = P.n;
    const int64_t m = blockIdx.x;
    for (int e = lane; e < n * n; e += 64)
        L.J[(e / n) * WaveLDS<N>::JS + e % n] = P.A[(int64_t)e * P.ldA + m];
    __syncthreads();
    const double b = lane < n ? P.b[(int64_t)lane * P.ldb + m] : 0.0;
    const double x = wave_lu_solve<N>(L, n, -1.0, b, 0.0);
    if (lane < n) P.x[(int64_t)lane * P.ldx + m] = x;
}

template <class DT>
__device__ __forceinline__ void lds_add_t(DT *p, DT v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <int N, class DT>
__global__ __launch_bounds__(64) void kim_wave_kernel(const KimParams P, const WMech M) {
    __shared__ WaveLDS<N> L;
    __shared__ DT s_d[64];
    const int lane = threadIdx.x, ns = M.ns;
    const int64_t cell = blockIdx.x;
    if (cell >= P.n) return;
    const bool sp = lane < ns;
    const double T = P.T[cell], rho = P.rho[cell];
    const double W = sp ? M.sp[SP_W * 64 + lane] : 1.0;
    const TFac tf = tfac(T);
    if (sp) {
        L.c[lane] = fmax(rho * P.Y[lane * P.ldY + cell] / W, 0.0);
        double G, dG, Cp, H, dCp;
        wthermo(M, lane, tf, G, dG, Cp, H, dCp);
        L.G[lane] = G;
        L.dG[lane] = dG;
    }
    if (lane < N) L.w[lane] = 0.0;
    s_d[lane] = DT(0);
    __syncthreads();
    const int last = ns - 1;
    const double Wns = M.sp[SP_W * 64 + last];
    for (int base = 0; base < M.nr; base += 64) {
        const int r = base + lane;
        if (r < M.nr) {
            WRxn o;
            wreaction<true>(M, base, tf, L.c, L.G, L.dG, P.exact, o);
#pragma unroll
            for (int i = 0; i < kWP; ++i) {
                if (i < o.nf) lds_add(&L.w[o.fi[i]], -o.fnu[i] * o.q);
                if (i < o.nb) lds_add(&L.w[o.bi[i]], o.bnu[i] * o.q);
            }
            // qfr<dOmegaType>: forward participants (side 0, derivative dqf)
            // then reverse participants (side 1, dqr)
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const int np = side == 0 ? o.nf : o.nb, no = side == 0 ? o.nb : o.nf;
                const int *id = side == 0 ? o.fi : o.bi;
                const int *oid = side == 0 ? o.bi : o.fi;
                const double *nu = side == 0 ? o.fnu : o.bnu;
                const double *onu = side == 0 ? o.bnu : o.fnu;
                const double *dq = side == 0 ? o.dqf : o.dqr;
#pragma unroll
                for (int i = 0; i < kWP; ++i) {
                    if (i >= np) continue;
                    const int spi = id[i];
                    if (spi != last) {
                        double d = -nu[i] * dq[i];
#pragma unroll
                        for (int j = 0; j < kWP; ++j)
                            if (j < no && oid[j] == spi) d += onu[j] * dq[i];
                        lds_add_t<DT>(&s_d[spi], (DT)d);
                    } else {
#pragma unroll
                        for (int j = 0; j < kWP; ++j) {
                            if (j < np && id[j] != spi)
                                lds_add_t<DT>(&s_d[id[j]],
                                              (DT)(M.sp[SP_W * 64 + id[j]] / Wns * nu[j] * dq[i]));
                            if (j < no && oid[j] != spi)
                                lds_add_t<DT>(&s_d[oid[j]],
                                              (DT)(-M.sp[SP_W * 64 + oid[j]] / Wns * onu[j] * dq[i]));
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    if (sp) P.R[lane * P.ldR + cell] = L.w[lane] * W;
    if (P.D && lane < last) ((DT *)P.D)[lane * P.ldD + cell] = s_d[lane];
}

// dy/dt and J at given states, one wavefront per cell.
template <int N, int FLOW, class F = double>
__global__ __launch_bounds__(64) void eval_wave_kernel(const EvalParams P, const WMech M, int exact) {
    __shared__ WaveLDS<N> L;
    const int lane = threadIdx.x, ns = M.ns, n = ns + 1;
    const int64_t cell = blockIdx.x;
    if (cell >= P.n) return;
    const double y = lane < n ? P.y[lane * P.ldy + cell] : 0.0;
    const double pr = P.pr[cell];
    (void)pr;
    double f;
    if (P.J) f = wave_eval<N, FLOW, true, F>(M, L, y, exact);
    else f = wave_eval<N, FLOW, false, F>(M, L, y, exact);
    if (lane < n) P.f[lane * P.ldf + cell] = f;
    if (P.J && lane < n)
        for (int j = 0; j < n; ++j) P.J[(lane * n + j) * P.ldJ + cell] = L.J[lane * WaveLDS<N>::JS + j];
}

}  // namespace ohcd
