// chem_device.hpp -- register-resident FP64 chemistry for one cell per lane.
//
// Everything here is specialised at compile time on a mechanism descriptor M
// (generated from the mechanism tables, see codegen.cpp): species and
// reaction indices are template constants, so the state vector, the RHS and
// the dense (ns+1)^2 Jacobian live in VGPRs with no dynamic indexing and the
// reaction loop is straight-line code.  The arithmetic restates
//   JANAF thermo        src/speciesTransport/thermo/src/JANAF.cpp:142-219
//   rate laws           src/speciesTransport/reactions/reactionRateTypes/src/*.cpp
//   Kp / Kc / q / DqDci / DqDT
//                       src/speciesTransport/reactions/reactionTypes/src/reactionTypes.cpp:178-445, :843-878
//   DyDt / jacobian     src/speciesTransport/chemistrySource/src/chemistrySource.cpp:863-1055
//   Doolittle LU        src/Hurricane/dataStructure/containers/Matrices/include/SquareMatrix.hpp:185-238
//   back substitution   src/Hurricane/mathematicalTools/LinearEquations/include/LUDecompose.hpp:68-96
// with FP64 rewrites that change results only at the rounding level:
// A*T^b*exp(-Ta/T) is evaluated as A*exp(b*lnT - Ta/T), integer powers as
// products, and repeated divisions by one value as a reciprocal multiply.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace ohcd {

constexpr double kRu = 1.380649e-23 * 6.02214076e23 * 1e3;  // J/kmol/K
constexpr double kPatm = 1.01325e5;
constexpr double kTiny = 1.0e-15;
constexpr double kVeryTiny = 1.0e-300;
constexpr double kLarge = 1.0e15;
constexpr double kVeryLarge = 1.0e300;
constexpr double kRootVeryLarge = 1.0e150;

constexpr int kMaxP = 6;

// One reaction of a compile-time mechanism (all units SI kmol-m^3-s-K).
struct RxnD {
    int nf, nb;
    int fi[kMaxP], bi[kMaxP];
    double fnu[kMaxP], bnu[kMaxP], fo[kMaxP], bo[kMaxP];
    int rev;   // 0 irreversible, 1 Kc, 2 REV parameters
    int kind;  // 0 Arrhenius, 1 third body, 2 fall-off, 3 chemically activated
    int fot;   // 0 Lindemann, 1 Troe3, 2 Troe4, 3 SRI
    double A, b, Ta;     // kf, or k_inf for fall-off kinds
    double A0, b0, Ta0;  // k0 for fall-off kinds
    double Ar, br, Tar;  // REV parameters
    double fop[5];
    double sumStoi;
    int tb;  // row of M::EFF, -1 if none
};

constexpr double cabs(double x) { return x < 0 ? -x : x; }
// compile-time square root (Newton iterations to convergence) for constants
constexpr double csqrt(double x) {
    if (!(x > 0)) return 0.0;
    double r = x > 1 ? x : 1.0;
    for (int i = 0; i < 200; ++i) {
        const double n = 0.5 * (r + x / r);
        if (n == r) break;
        r = n;
    }
    return r;
}

template <int B, int E, class F>
__device__ __forceinline__ void sfor(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        sfor<B + 1, E>(f);
    }
}

__device__ __forceinline__ bool bad(double x) { return isnan(x) || isinf(x); }

// 1/x from the hardware reciprocal estimate refined by two Newton steps
// (within 1 ulp of the correctly rounded quotient for normal x; 0 -> inf,
// inf -> 0, NaN -> NaN).  Used where a quotient feeds a tolerance ratio, a
// pivot or a Kc reciprocal, never for a value that is compared bit-exactly.
__device__ __forceinline__ double frcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    if (isinf(x) || x == 0.0) return r;
    double e = fma(-x, r, 1.0);
    r = fma(e, r, r);
    e = fma(-x, r, 1.0);
    return fma(e, r, r);
}

// x^p for a compile-time exponent p >= 0 (integer exponents as products).
template <class P>
__device__ __forceinline__ double cpow(double x) {
    constexpr double p = P::value;
    if constexpr (p == 0.0) return 1.0;
    else if constexpr (p == 1.0) return x;
    else if constexpr (p == 2.0) return x * x;
    else if constexpr (p == 3.0) return x * x * x;
    else return pow(x, p);
}

// Lane-private thermo at one temperature: molar Cp, dCp/dT, H, G, dG/dT.
template <class M>
struct Thermo {
    double Cp[M::NS], dCp[M::NS], H[M::NS], G[M::NS], dG[M::NS];
};

// Coefficient set of species k at T (JANAF.inl:34-40) -- a[0..6] and the
// host-precomputed quotients a_i/5.. (h), a_i/4.. (s), a_i/20.. (g).
// R = -1 picks the set per lane (T < Tcommon: low); R = 0 / 1 fixes the low /
// high set, for code where every lane of the wavefront is known to be on that
// side, so the coefficients stay compile-time constants instead of per-lane
// selects (two moves and two selects per FP64 value).
template <class M, int k, int R = -1>
struct Coef {
    double a[7], h4, h3, h2, h1, s4, s3, s2, g4, g3, g2, g1;
    __device__ __forceinline__ explicit Coef(double T) {
        const bool lo = R < 0 ? T < M::Tcommon[k] : R == 0;
#pragma unroll
        for (int c = 0; c < 7; ++c) a[c] = lo ? M::alo[k][c] : M::ahi[k][c];
        h4 = lo ? M::hlo[k][0] : M::hhi[k][0];
        h3 = lo ? M::hlo[k][1] : M::hhi[k][1];
        h2 = lo ? M::hlo[k][2] : M::hhi[k][2];
        h1 = lo ? M::hlo[k][3] : M::hhi[k][3];
        s4 = lo ? M::slo[k][0] : M::shi[k][0];
        s3 = lo ? M::slo[k][1] : M::shi[k][1];
        s2 = lo ? M::slo[k][2] : M::shi[k][2];
        g4 = lo ? M::glo[k][0] : M::ghi[k][0];
        g3 = lo ? M::glo[k][1] : M::ghi[k][1];
        g2 = lo ? M::glo[k][2] : M::ghi[k][2];
        g1 = lo ? M::glo[k][3] : M::ghi[k][3];
    }
};

// Highest Tcommon of the mechanism.
template <class M>
constexpr double tcommon_max() {
    double v = M::Tcommon[0];
    for (int k = 1; k < M::NS; ++k) v = M::Tcommon[k] > v ? M::Tcommon[k] : v;
    return v;
}

// Whether every active lane of the wavefront is at or above every Tcommon
// (the high coefficient sets apply throughout).
template <class M>
__device__ __forceinline__ bool wave_all_high(double T) {
    constexpr double Th = tcommon_max<M>();
    return __all(T >= Th);
}

// cp0 with the JANAF temperature clamp (JANAF.cpp:155-159), per unit mass.
template <class M, int k, int R>
__device__ __forceinline__ double janaf_cp0(const Coef<M, k, R> &c, double T) {
    const double Tl = M::Tlow[k], Th = M::Thigh[k];
    const double T0 = (T < Tl || T > Th) ? fmin(fmax(T, Tl), Th) : T;
    const double *a = c.a;
    return ((((a[4] * T0 + a[3]) * T0 + a[2]) * T0 + a[1]) * T0 + a[0]);
}

// Molar G and dG/dT of species k (JANAF.cpp:199-212, thermo.inl:195-205).
template <class M, int k, int R>
__device__ __forceinline__ void species_G(const Coef<M, k, R> &c, double T, double lnT,
                                          double invTm, double &G, double &dG) {
    const double *a = c.a;
    const double cp0 = janaf_cp0(c, T);
    const double s0 = ((((c.s4 * T + c.s3) * T + c.s2) * T + a[1]) * T + a[0] * lnT + a[6]);
    const double g0 =
        (((((-c.g4 * T - c.g3) * T - c.g2) * T - c.g1) * T + a[0] * (1.0 - lnT) - a[6]) * T + a[5]);
    const double ds0 = (a[0] * invTm + a[1] + (a[2] + (a[3] + a[4] * T) * T) * T);
    constexpr double W = M::W[k];
    G = g0 * W;
    dG = (cp0 - (s0 + T * ds0)) * W;
}

template <class M, int k>
__device__ __forceinline__ void species_G(double T, double lnT, double invTm, double &G,
                                          double &dG) {
    species_G<M, k>(Coef<M, k>(T), T, lnT, invTm, G, dG);
}

// Molar Cp, H and dCp/dT of species k (JANAF.cpp:155-189, :214-217).
template <class M, int k, int R>
__device__ __forceinline__ void species_CH(const Coef<M, k, R> &c, double T, double &Cp, double &H,
                                           double &dCp) {
    const double *a = c.a;
    const double cp0 = janaf_cp0(c, T);
    const double Tl = M::Tlow[k], Th = M::Thigh[k];
    const bool out = (T < Tl) || (T > Th);
    const double Tx = out ? (T < Tl ? Tl : Th) : T;
    double h = ((((c.h4 * Tx + c.h3) * Tx + c.h2) * Tx + c.h1) * Tx + a[0]) * Tx + a[5];
    if (out) h += cp0 * (T - Tx);
    const double dcp = (a[1] + T * (2.0 * a[2] + T * (3.0 * a[3] + 4.0 * a[4] * T)));
    constexpr double W = M::W[k];
    Cp = cp0 * W;
    H = h * W;
    dCp = dcp * W;
}

template <class M, int k>
__device__ __forceinline__ void species_CH(double T, double &Cp, double &H, double &dCp) {
    species_CH<M, k>(Coef<M, k>(T), T, Cp, H, dCp);
}

// species_G evaluated in FP32 (the split-precision mode's Gibbs terms).
template <class M, int k, int R = -1>
__device__ __forceinline__ void species_G32(double Td, float lnT, float invTm, double &G,
                                            double &dG) {
    const bool lo = R < 0 ? Td < M::Tcommon[k] : R == 0;
    const float T = (float)Td;
    float a[7];
#pragma unroll
    for (int c = 0; c < 7; ++c) a[c] = lo ? (float)M::alo[k][c] : (float)M::ahi[k][c];
    const float s4 = lo ? (float)M::slo[k][0] : (float)M::shi[k][0];
    const float s3 = lo ? (float)M::slo[k][1] : (float)M::shi[k][1];
    const float s2 = lo ? (float)M::slo[k][2] : (float)M::shi[k][2];
    const float g4 = lo ? (float)M::glo[k][0] : (float)M::ghi[k][0];
    const float g3 = lo ? (float)M::glo[k][1] : (float)M::ghi[k][1];
    const float g2 = lo ? (float)M::glo[k][2] : (float)M::ghi[k][2];
    const float g1 = lo ? (float)M::glo[k][3] : (float)M::ghi[k][3];
    const float Tl = (float)M::Tlow[k], Th = (float)M::Thigh[k];
    const float T0 = (T < Tl || T > Th) ? fminf(fmaxf(T, Tl), Th) : T;
    const float cp0 = ((((a[4] * T0 + a[3]) * T0 + a[2]) * T0 + a[1]) * T0 + a[0]);
    const float s0 = ((((s4 * T + s3) * T + s2) * T + a[1]) * T + a[0] * lnT + a[6]);
    const float g0 = (((((-g4 * T - g3) * T - g2) * T - g1) * T + a[0] * (1.0f - lnT) - a[6]) * T + a[5]);
    const float ds0 = (a[0] * invTm + a[1] + (a[2] + (a[3] + a[4] * T) * T) * T);
    constexpr float W = (float)M::W[k];
    G = (double)(g0 * W);
    dG = (double)((cp0 - (s0 + T * ds0)) * W);
}

template <class M, int k>
__device__ __forceinline__ void species_thermo(double T, double lnT, double invTm, Thermo<M> &th) {
    species_G<M, k>(T, lnT, invTm, th.G[k], th.dG[k]);
    species_CH<M, k>(T, th.Cp[k], th.H[k], th.dCp[k]);
}

// ArrheniusRR::k with the reference's skip of pow/exp when b or Ta vanish;
// F = float evaluates the rate constant in FP32 (split-precision mode).
template <int kindsel, class R, class F = double>
__device__ __forceinline__ double arrh(double lnT, double invT) {
    constexpr double A = kindsel == 0 ? R::A : (kindsel == 1 ? R::A0 : R::Ar);
    constexpr double b = kindsel == 0 ? R::b : (kindsel == 1 ? R::b0 : R::br);
    constexpr double Ta = kindsel == 0 ? R::Ta : (kindsel == 1 ? R::Ta0 : R::Tar);
    constexpr bool hb = cabs(b) > kVeryTiny, ht = cabs(Ta) > kVeryTiny;
    if constexpr (std::is_same_v<F, float>) {
        const float l = (float)lnT, it = (float)invT;
        if constexpr (hb && ht) return (double)((float)A * __expf((float)b * l - (float)Ta * it));
        else if constexpr (hb) return (double)((float)A * __expf((float)b * l));
        else if constexpr (ht) return (double)((float)A * __expf(-(float)Ta * it));
        else return A;
    } else {
        if constexpr (hb && ht) return A * exp(b * lnT - Ta * invT);
        else if constexpr (hb) return A * exp(b * lnT);
        else if constexpr (ht) return A * exp(-Ta * invT);
        else return A;
    }
}
template <int kindsel, class R>
__device__ __forceinline__ double arrh_dkdT(double kj, double invTm) {
    constexpr double b = kindsel == 0 ? R::b : (kindsel == 1 ? R::b0 : R::br);
    constexpr double Ta = kindsel == 0 ? R::Ta : (kindsel == 1 ? R::Ta0 : R::Tar);
    return kj * invTm * (b + Ta * invTm);
}

// Reaction r exposed as a type with static constexpr members.
template <class M, int r>
struct RT {
    static constexpr RxnD d = M::R[r];
    static constexpr double A = d.A, b = d.b, Ta = d.Ta, A0 = d.A0, b0 = d.b0, Ta0 = d.Ta0,
                            Ar = d.Ar, br = d.br, Tar = d.Tar;
};

template <class M, int r, int side, int i>
struct Ord {
    static constexpr double value = side == 0 ? M::R[r].fo[i] : M::R[r].bo[i];
};
template <class M, int r, int side, int i>
struct Nu {
    static constexpr double value = side == 0 ? M::R[r].fnu[i] : M::R[r].bnu[i];
};
template <class M, int r, int side, int i>
struct OrdM1 {
    static constexpr double value = (side == 0 ? M::R[r].fo[i] : M::R[r].bo[i]) - 1.0;
};

// fallOffFunctions.cpp:54-168 (Lindemann F = 1).
template <class M, int r>
__device__ __forceinline__ double fo_F(double T, double Pr) {
    constexpr RxnD d = M::R[r];
    if constexpr (d.fot == 3) {
        const double logPr = log10(fmax(Pr, kTiny));
        const double X = 1.0 / (1.0 + logPr * logPr);
        return d.fop[3] * pow(d.fop[0] * exp(-d.fop[1] / T) + exp(-T / d.fop[2]), X) *
               pow(T, d.fop[4]);
    } else if constexpr (d.fot == 1 || d.fot == 2) {
        double fc = (1.0 - d.fop[0]) * exp(-T / d.fop[1]) + d.fop[0] * exp(-T / d.fop[2]);
        if constexpr (d.fot == 2) fc += exp(-d.fop[3] / T);
        const double logFcent = log10(fmax(fc, kTiny));
        const double c = -0.4 - 0.67 * logFcent, n = 0.75 - 1.27 * logFcent, dd = 0.14;
        const double logPr = log10(fmax(Pr, kTiny));
        const double q = (logPr + c) / (n - dd * (logPr + c));
        return pow(10.0, logFcent / (1.0 + q * q));
    } else {
        return 1.0;
    }
}
template <class M, int r>
__device__ __forceinline__ double fo_DFDci(double T, double Pr, double F, double dPr) {
    constexpr RxnD d = M::R[r];
    if constexpr (d.fot == 3) {
        const double lp = log10(fmax(Pr, kTiny));
        const double X = 1.0 / (1.0 + lp * lp);
        const double dXdc = -X * X * 2.0 * log10(Pr) / Pr / log(10.0) * dPr;
        return F * dXdc * log(d.fop[0] * exp(-d.fop[1] / T) + exp(-T / d.fop[2]));
    } else if constexpr (d.fot == 1 || d.fot == 2) {
        double fc = (1.0 - d.fop[0]) * exp(-T / d.fop[1]) + d.fop[0] * exp(-T / d.fop[2]);
        if constexpr (d.fot == 2) fc += exp(-d.fop[3] / T);
        const double logFcent = log10(fmax(fc, kTiny));
        const double c = -0.4 - 0.67 * logFcent, n = 0.75 - 1.27 * logFcent, dd = 0.14;
        const double logPrpc = log10(fmax(Pr, kTiny)) + c;
        const double nmd = n - dd * logPrpc;
        const double inv = logPrpc / nmd;
        const double q = 1.0 + inv * inv;
        return F * ((-logFcent) / (q * q)) * (2.0 * logPrpc / (nmd * nmd)) * (1.0 + dd * inv) *
               (1.0 / Pr * dPr);
    } else {
        return 0.0;
    }
}
template <class M, int r>
__device__ __forceinline__ double fo_DFDT(double T, double Pr, double F, double dPrdT) {
    constexpr RxnD d = M::R[r];
    if constexpr (d.fot == 3) {
        const double a = d.fop[0], b = d.fop[1], c = d.fop[2], e = d.fop[4];
        const double lp = log10(fmax(Pr, kTiny));
        const double X = 1.0 / (1.0 + lp * lp);
        const double dXdPr = -X * X * 2.0 * log10(Pr) / Pr / log(10.0);
        const double ea = exp(-b / T), ec = exp(-T / c);
        return F * (e / T + dXdPr * dPrdT * log(a * ea + ec) +
                    X * (a * b * ea / (T * T) - ec / c) / (a * ea + ec));
    } else if constexpr (d.fot == 1 || d.fot == 2) {
        const double ln10 = log(10.0);
        const double logPr = log10(fmax(Pr, kTiny));
        const double e1 = exp(-T / d.fop[1]), e2 = exp(-T / d.fop[2]);
        double Fc = (1.0 - d.fop[0]) * e1 + d.fop[0] * e2;
        double dFc = (1.0 - d.fop[0]) * e1 / (-d.fop[1]) + d.fop[0] * e2 / (-d.fop[2]);
        if constexpr (d.fot == 2) {
            const double e3 = exp(-d.fop[3] / T);
            Fc += e3;
            dFc += e3 * (d.fop[3] / (T * T));
        }
        Fc = fmax(Fc, kTiny);
        const double logFcent = log10(Fc);
        const double dlF = 1.0 / (Fc * ln10) * dFc;
        const double c = -0.4 - 0.67 * logFcent, dc = -0.67 * dlF;
        const double n = 0.75 - 1.27 * logFcent, dn = -1.27 * dlF;
        const double dd = 0.14;
        const double dlogPr = dPrdT / Pr / ln10;
        const double lpc = logPr + c;
        const double nmd = n - dd * lpc;
        const double inv = lpc / nmd;
        const double q = 1.0 + inv * inv;
        const double dPP = 2.0 * lpc / (nmd * nmd) *
                           ((dlogPr + dc) - lpc / nmd * (dn - dd * (dlogPr + dc)));
        return F * ln10 * (dlF / q - logFcent / (q * q) * dPP);
    } else {
        return 0.0;
    }
}

// Per-evaluation temperature factors shared by all reactions.
struct TFac {
    double T, lnT, invT, invTm, invRuT, X, Xinv;
};

__device__ __forceinline__ TFac tfac(double T) {
    TFac f;
    f.T = T;
    const double Tm = fmax(T, kTiny);
    f.lnT = log(Tm);
    f.invT = frcp(T);
    f.invTm = frcp(Tm);
    f.invRuT = frcp(Tm * kRu);
    f.X = kPatm / (kRu * T);  // Patm / (Ru T), the Kc conversion base
    f.Xinv = (kRu * T) / kPatm;
    return f;
}

template <class S>
__device__ __forceinline__ double Xpow(const TFac &tf) {
    constexpr double s = S::value;
    if constexpr (s == 1.0) return tf.X;
    else if constexpr (s == -1.0) return tf.Xinv;
    else if constexpr (s == 2.0) return tf.X * tf.X;
    else if constexpr (s == -2.0) return tf.Xinv * tf.Xinv;
    else return pow(tf.X, s);
}
template <class M, int r>
struct SumS {
    static constexpr double value = M::R[r].sumStoi;
};

// Equilibrium terms of reversible reaction r: Kc, 1/max(Kc, veryTiny) and
// (optionally) dKc/dT -- reactionTypes.cpp:178-241 (dKp/dT with the
// reference's sign convention).
template <class M, int r>
struct MSumS {
    static constexpr double value = -M::R[r].sumStoi;
};

// Equilibrium terms of reversible reaction r: Kc, rK = 1/max(Kc, veryTiny)
// and (optionally) dKc/dT -- reactionTypes.cpp:178-241 (dKp/dT with the
// reference's sign convention).
template <class M, int r, bool WITH_D, class F = double>
__device__ __forceinline__ void equilibrium(const TFac &tf, const Thermo<M> &th, double &Kc,
                                            double &rK, double &dKc, bool exact) {
    constexpr RxnD d = M::R[r];
    constexpr bool zero_s = cabs(d.sumStoi) < kVeryTiny;
    constexpr bool kc_species = M::KC_SPECIES && std::is_same_v<F, double>;
    double Kp, rKp = 0.0, dk = 0.0;
    if constexpr (std::is_same_v<F, float>) {
        // FP32 Gibbs sums and exponential (split-precision mode)
        float Kt = 0.0f, dkf = 0.0f;
        const float iRuT = (float)tf.invRuT, iTm = (float)tf.invTm;
        sfor<0, d.nf>([&](auto i) {
            constexpr int k = d.fi[i];
            Kt -= (float)th.G[k] * (float)d.fnu[i];
            if constexpr (WITH_D) dkf -= ((float)th.dG[k] - (float)th.G[k] * iTm) * (float)d.fnu[i];
        });
        sfor<0, d.nb>([&](auto i) {
            constexpr int k = d.bi[i];
            Kt += (float)th.G[k] * (float)d.bnu[i];
            if constexpr (WITH_D) dkf += ((float)th.dG[k] - (float)th.G[k] * iTm) * (float)d.bnu[i];
        });
        Kp = (double)__expf(-Kt * iRuT);
        dk = (double)(dkf * iRuT);
        if (bad(Kp)) Kp = kRootVeryLarge;
    } else if constexpr (kc_species) {
        // th.G = E_k = exp(-G_k/(Ru T)), th.H = 1/E_k, th.dG = (dG_k - G_k/T)/(Ru T):
        // Kp = prod_products E^nu prod_reactants (1/E)^nu == exp(-sum nu G/(Ru T)).
        double Pn = 1.0, Qd = 1.0, Pi = 1.0, Qi = 1.0;
        sfor<0, d.nb>([&](auto i) {
            Pn *= cpow<Nu<M, r, 1, i>>(th.G[d.bi[i]]);
            Pi *= cpow<Nu<M, r, 1, i>>(th.H[d.bi[i]]);
        });
        sfor<0, d.nf>([&](auto i) {
            Qd *= cpow<Nu<M, r, 0, i>>(th.H[d.fi[i]]);
            Qi *= cpow<Nu<M, r, 0, i>>(th.G[d.fi[i]]);
        });
        Kp = Pn * Qd;
        rKp = Pi * Qi;
        if constexpr (WITH_D) {
            sfor<0, d.nf>([&](auto i) { dk -= th.dG[d.fi[i]] * d.fnu[i]; });
            sfor<0, d.nb>([&](auto i) { dk += th.dG[d.bi[i]] * d.bnu[i]; });
        }
        if (bad(Kp)) {
            Kp = kRootVeryLarge;
            rKp = 1.0 / kRootVeryLarge;
        }
    } else {
        double Kt = 0.0;
        sfor<0, d.nf>([&](auto i) { Kt -= th.G[d.fi[i]] * d.fnu[i]; });
        sfor<0, d.nb>([&](auto i) { Kt += th.G[d.bi[i]] * d.bnu[i]; });
        Kt *= tf.invRuT;
        Kp = exp(-Kt);
        if constexpr (WITH_D) {
            sfor<0, d.nf>([&](auto i) {
                constexpr int k = d.fi[i];
                dk -= (th.dG[k] - th.G[k] * tf.invTm) * d.fnu[i];
            });
            sfor<0, d.nb>([&](auto i) {
                constexpr int k = d.bi[i];
                dk += (th.dG[k] - th.G[k] * tf.invTm) * d.bnu[i];
            });
            dk *= tf.invRuT;
        }
        if (bad(Kp)) Kp = kRootVeryLarge;
    }
    double Xs = 1.0;
    if constexpr (zero_s) {
        Kc = Kp;
    } else {
        Xs = Xpow<SumS<M, r>>(tf);
        Kc = fmin(Kp * Xs, kVeryLarge);
    }
    if constexpr (kc_species) {
        double rKc = rKp;
        if constexpr (!zero_s) rKc = Kc == kVeryLarge ? 1.0 / kVeryLarge : rKp * Xpow<MSumS<M, r>>(tf);
        rK = Kc > kVeryTiny ? rKc : 1.0 / kVeryTiny;
    } else {
        rK = frcp(fmax(Kc, kVeryTiny));
    }
    if constexpr (WITH_D) {
        if (exact) dk = -dk;  // exact d(Kp)/dT = -Kp d(sum nu G/(Ru T))/dT
        double dKp = Kp * dk;
        if (bad(dKp)) dKp = kVeryLarge;
        if constexpr (zero_s) dKc = dKp;
        else {
            double v = (dKp + Kp * (-d.sumStoi) * tf.invT) * Xs;
            if (bad(v)) v = kVeryLarge;
            dKc = v;
        }
    } else {
        dKc = 0.0;
    }
}

// Third-body concentration M = sum eff_k c_k of efficiency group g
// (thirdBodyEfficiency::M, thirdBodyEfficiency.hpp:58-77).
template <class M, int g>
__device__ __forceinline__ double third_M(const double *ci) {
    double m = 0.0;
    sfor<0, M::NS>([&](auto k) {
        constexpr double e = M::EFF[g][k];
        if constexpr (e != 0.0) {
            if constexpr (e == 1.0) m += ci[k];
            else m += e * ci[k];
        }
    });
    return m;
}

// Forward or reverse rate-constant law k(p, T, c) (kindsel 0: forward table,
// 2: REV table) and the fall-off helpers.
template <class M, int r>
struct FallOff {
    double k0, kinf, Pr, F;
};

template <class M, int r>
__device__ __forceinline__ FallOff<M, r> falloff(const TFac &tf, double Mtb) {
    using R = RT<M, r>;
    FallOff<M, r> f;
    f.k0 = arrh<1, R>(tf.lnT, tf.invT);
    f.kinf = arrh<0, R>(tf.lnT, tf.invT);
    const double kinfz = f.kinf != 0 ? f.kinf : kVeryTiny;
    f.Pr = f.k0 / kinfz * Mtb;
    f.F = fo_F<M, r>(tf.T, f.Pr);
    return f;
}

// Evaluate dy/dt (and the Jacobian when WITH_J) at y = [c_1..c_ns, T] for the
// constant-pressure reactor (FLOW == 0) or constant volume (FLOW == 1, pr =
// rho0).  f has NS+1 entries, J is (NS+1)x(NS+1) row-major in registers;
// exact selects the exact d(Kp)/dT over the reference's sign convention.
// F = float is the split-precision mode: Gibbs energies, equilibrium and
// Arrhenius rate constants in FP32, everything else (concentration products,
// net rates, Jacobian, temperature equation) in FP64.
// UNI: when every lane of the wavefront is above every Tcommon, evaluate the
// thermo with the high coefficient sets as constants (else per-lane selects).
// MA: mass-action Jacobian only -- the third-body and fall-off q dln(gamma)/dc
// terms are left out, so that J[k][k] is reactionTable::dWdciAllSpecies's
// diagonal (the MTS time scales).
template <class M, bool WITH_J, int FLOW, class F = double, bool UNI = false, bool MA = false>
__device__ __forceinline__ void eval_fJ(const double (&y)[M::NS + 1], double pr,
                                        double (&f)[M::NS + 1],
                                        double (&J)[M::NS + 1][M::NS + 1], bool exact = false) {
    constexpr int NS = M::NS, N = NS + 1;
    constexpr bool fp32 = std::is_same_v<F, float>;
    const double T = y[NS];
    const TFac tf = tfac(T);
    const bool hi = UNI && wave_all_high<M>(T);
    Thermo<M> th;  // G / dG now; Cp / H / dCp after the reaction loop
    auto gibbs = [&](auto Rc) {
        constexpr int R = decltype(Rc)::value;
        sfor<0, NS>([&](auto k) {
            if constexpr (fp32)
                species_G32<M, k, R>(T, (float)tf.lnT, (float)tf.invTm, th.G[k], th.dG[k]);
            else
                species_G<M, k>(Coef<M, k, R>(T), T, tf.lnT, tf.invTm, th.G[k], th.dG[k]);
            if constexpr (M::KC_SPECIES && !fp32) {
                const double gRT = th.G[k] * tf.invRuT;
                th.dG[k] = (th.dG[k] - th.G[k] * tf.invTm) * tf.invRuT;
                th.G[k] = exp(-gRT);
                th.H[k] = frcp(th.G[k]);  // exp(+G/(Ru T)), slot reused until the T row
            }
        });
    };
    if (hi) gibbs(std::integral_constant<int, 1>{});
    else gibbs(std::integral_constant<int, -1>{});
    double ci[NS];
    sfor<0, NS>([&](auto k) { ci[k] = fmax(y[k], 0.0); });
    // third-body concentrations and the per-group row weights of dln(M)/dc
    double Mg[M::NTB > 0 ? M::NTB : 1], rMg[M::NTB > 0 ? M::NTB : 1];
    double u[M::NTB > 0 ? M::NTB : 1][NS];
    sfor<0, M::NTB>([&](auto g) {
        Mg[g] = third_M<M, g>(ci);
        rMg[g] = 1.0 / fmax(Mg[g], kTiny);
        sfor<0, NS>([&](auto a) { u[g][a] = 0.0; });
    });
    sfor<0, N>([&](auto i) {
        f[i] = 0.0;
        if constexpr (WITH_J) sfor<0, N>([&](auto j) { J[i][j] = 0.0; });
    });
    sfor<0, M::NR>([&](auto rr) {
        constexpr int r = rr;
        constexpr RxnD d = M::R[r];
        using R = RT<M, r>;
        // ---- rate constants
        double Mtb = 0.0;
        if constexpr (d.kind != 0) Mtb = Mg[d.tb];
        double kf, kr = 0.0, Kc = 0.0, dKc = 0.0, rKc = 0.0;
        FallOff<M, r> fo{};
        if constexpr (d.kind == 0) kf = arrh<0, R, F>(tf.lnT, tf.invT);
        else if constexpr (d.kind == 1) kf = Mtb * arrh<0, R, F>(tf.lnT, tf.invT);
        else {
            fo = falloff<M, r>(tf, Mtb);
            if constexpr (d.kind == 2) kf = fo.kinf * (fo.Pr / (1.0 + fo.Pr)) * fo.F;
            else kf = fo.k0 * (1.0 / (1.0 + fo.Pr)) * fo.F;
        }
        if constexpr (d.rev == 1) {
            equilibrium<M, r, WITH_J, F>(tf, th, Kc, rKc, dKc, exact);
            kr = Kc > kVeryTiny ? kf * rKc : 0.0;
        } else if constexpr (d.rev == 2) {
            kr = arrh<2, R, F>(tf.lnT, tf.invT);
            if constexpr (d.kind == 1) kr = Mtb * kr;
        }
        // ---- concentration products
        double qf = kf, qr = kr;
        sfor<0, d.nf>([&](auto i) { qf *= cpow<Ord<M, r, 0, i>>(ci[d.fi[i]]); });
        sfor<0, d.nb>([&](auto i) { qr *= cpow<Ord<M, r, 1, i>>(ci[d.bi[i]]); });
        const double q = qf - qr;
        sfor<0, d.nf>([&](auto i) { f[d.fi[i]] -= d.fnu[i] * q; });
        sfor<0, d.nb>([&](auto i) { f[d.bi[i]] += d.bnu[i] * q; });
        if constexpr (WITH_J) {
            // ---- DqDci (reactionTypes.cpp:354-445)
            auto dside = [&](auto side, double k0v, auto si) {
                constexpr int sd = decltype(side)::value;
                constexpr int nn = sd == 0 ? d.nf : d.nb;
                double dq = k0v;
                sfor<0, nn>([&](auto j) {
                    constexpr int sj = sd == 0 ? d.fi[j] : d.bi[j];
                    constexpr double ord = sd == 0 ? d.fo[j] : d.bo[j];
                    if constexpr (j == decltype(si)::value) {
                        if constexpr (ord == 1.0) {
                        } else if constexpr (ord < 1.0) {
                            dq *= ci[sj] > kTiny ? ord * pow(ci[sj], ord - 1.0) : 0.0;
                        } else {
                            dq *= ord * cpow<OrdM1<M, r, sd, j>>(ci[sj]);
                        }
                    } else {
                        dq *= cpow<Ord<M, r, sd, j>>(ci[sj]);
                    }
                });
                return dq;
            };
            sfor<0, d.nf>([&](auto i) {
                constexpr int si = d.fi[i];
                const double dq = dside(std::integral_constant<int, 0>{}, kf, i);
                sfor<0, d.nf>([&](auto j) { J[d.fi[j]][si] -= d.fnu[j] * dq; });
                sfor<0, d.nb>([&](auto j) { J[d.bi[j]][si] += d.bnu[j] * dq; });
            });
            sfor<0, d.nb>([&](auto i) {
                constexpr int si = d.bi[i];
                const double dq = dside(std::integral_constant<int, 1>{}, kr, i);
                sfor<0, d.nf>([&](auto j) { J[d.fi[j]][si] += d.fnu[j] * dq; });
                sfor<0, d.nb>([&](auto j) { J[d.bi[j]][si] -= d.bnu[j] * dq; });
            });
            // third-body terms q * dln(gamma)/dc_i
            if constexpr (MA) {
            } else if constexpr (d.kind == 1) {
                // J(a, i) += sigma_a q eff_i / max(M, tiny): accumulate sigma_a q / M per
                // row, the eff_i outer product is applied once per group below.
                const double w = q * rMg[d.tb];
                sfor<0, d.nf>([&](auto j) { u[d.tb][d.fi[j]] -= d.fnu[j] * w; });
                sfor<0, d.nb>([&](auto j) { u[d.tb][d.bi[j]] += d.bnu[j] * w; });
            } else if constexpr (d.kind >= 2) {
                if (Mtb > kTiny) {
                    const double kinfz = fo.kinf != 0 ? fo.kinf : kVeryTiny;
                    sfor<0, NS>([&](auto i) {
                        constexpr double e = M::EFF[d.tb][i];
                        const double dPr = e * fo.k0 / kinfz;
                        const double dF = fo_DFDci<M, r>(T, fo.Pr, fo.F, dPr);
                        double g;
                        if constexpr (d.kind == 2) g = dPr / (fo.Pr * (1 + fo.Pr)) + dF / fo.F;
                        else g = -dPr / (1 + fo.Pr) + dF / fo.F;
                        if (!bad(g)) {
                            sfor<0, d.nf>([&](auto j) { J[d.fi[j]][i] -= d.fnu[j] * g * q; });
                            sfor<0, d.nb>([&](auto j) { J[d.bi[j]][i] += d.bnu[j] * g * q; });
                        }
                    });
                }
            }
            // ---- DqDT (reactionTypes.cpp:843-878)
            double dkf, dkr = 0.0;
            if constexpr (d.kind <= 1) dkf = arrh_dkdT<0, R>(kf, tf.invTm);
            else if constexpr (d.kind == 2) dkf = (fo.Pr / (1.0 + fo.Pr)) * fo.F * arrh_dkdT<0, R>(fo.kinf, tf.invTm);
            else dkf = (1.0 / (1.0 + fo.Pr)) * fo.F * arrh_dkdT<1, R>(fo.k0, tf.invTm);
            if constexpr (d.rev == 1) dkr = (dkf - kr * dKc) * rKc;
            else if constexpr (d.rev == 2) dkr = arrh_dkdT<2, R>(kr, tf.invTm);
            sfor<0, d.nf>([&](auto i) { dkf *= cpow<Ord<M, r, 0, i>>(ci[d.fi[i]]); });
            sfor<0, d.nb>([&](auto i) { dkr *= cpow<Ord<M, r, 1, i>>(ci[d.bi[i]]); });
            double dqdT = dkf - dkr;
            double dgdT = 0.0;
            if constexpr (d.kind >= 2) {
                if (Mtb > kTiny) {
                    const double k0z = fo.k0 != 0 ? fo.k0 : kVeryTiny;
                    const double kinfz = fo.kinf != 0 ? fo.kinf : kVeryTiny;
                    const double dPrdT = fo.Pr * (arrh_dkdT<1, R>(fo.k0, tf.invTm) / k0z -
                                                  arrh_dkdT<0, R>(fo.kinf, tf.invTm) / kinfz);
                    const double dFdT = fo_DFDT<M, r>(T, fo.Pr, fo.F, dPrdT);
                    double gg;
                    if constexpr (d.kind == 2) gg = dPrdT / (fo.Pr * (1 + fo.Pr)) + dFdT / fo.F;
                    else gg = -dPrdT / (1 + fo.Pr) + dFdT / fo.F;
                    dgdT = q * gg;
                }
            }
            if (bad(dqdT)) dqdT = 0.0;
            if (bad(dgdT)) dgdT = 0.0;
            const double s = dqdT + dgdT;
            sfor<0, d.nf>([&](auto i) { J[d.fi[i]][NS] -= d.fnu[i] * s; });
            sfor<0, d.nb>([&](auto i) { J[d.bi[i]][NS] += d.bnu[i] * s; });
        }
    });
    if constexpr (WITH_J && !MA) {
        sfor<0, M::NTB>([&](auto g) {
            sfor<0, NS>([&](auto a) {
                if constexpr (M::TBROW[g][a]) {
                    sfor<0, NS>([&](auto i) {
                        constexpr double e = M::EFF[g][i];
                        if constexpr (e == 1.0) J[a][i] += u[g][a];
                        else if constexpr (e != 0.0) J[a][i] += u[g][a] * e;
                    });
                }
            });
        });
    }
    // ---- temperature row (chemistrySource.cpp:993-1020 / :1021-1054)
    auto cph = [&](auto Rc) {
        constexpr int R = decltype(Rc)::value;
        sfor<0, NS>([&](auto k) {
            species_CH<M, k>(Coef<M, k, R>(T), T, th.Cp[k], th.H[k], th.dCp[k]);
        });
    };
    if (hi) cph(std::integral_constant<int, 1>{});
    else cph(std::integral_constant<int, -1>{});
    double Hx[NS], Cx[NS];
    if constexpr (FLOW == 0) {
        sfor<0, NS>([&](auto k) {
            Hx[k] = th.H[k];
            Cx[k] = th.Cp[k];
        });
    } else {
        sfor<0, NS>([&](auto k) {
            Hx[k] = th.H[k] - kRu * T;  // ea W = (ha - Ri T) W
            Cx[k] = th.Cp[k] - kRu;     // Cv W = (cp - Ri) W
        });
    }
    double cC = 0.0, dT = 0.0, dcC = 0.0;
    sfor<0, NS>([&](auto k) {
        cC += y[k] * Cx[k];
        if constexpr (WITH_J) dcC += y[k] * th.dCp[k];
        dT += f[k] * Hx[k];
    });
    const double rcC = frcp(cC);
    dT *= rcC;
    f[NS] = -dT;
    if constexpr (WITH_J) {
        sfor<0, NS>([&](auto i) {
            double v = 0.0;
            sfor<0, NS>([&](auto j) { v -= J[j][i] * Hx[j]; });
            v += Cx[i] * dT;
            J[NS][i] = v * rcC;
        });
        double v = 0.0;
        sfor<0, NS>([&](auto i) { v -= (Cx[i] * f[i] + Hx[i] * J[i][NS]); });
        v += dcC * dT;
        J[NS][NS] = v * rcC;
    }
}

// In-place Doolittle LU with partial pivoting of A (registers) applied to b
// at the same time (the reference's recorded row swaps replayed on b give the
// same sequence), then forward/back substitution: b <- A^-1 b.
// Row swaps are done with selects and skipped when no lane of the wavefront
// needs them.
template <int N>
__device__ __forceinline__ void lu_solve(double (&A)[N][N], double (&b)[N]) {
    double dinv[N];
    sfor<0, N>([&](auto kk) {
        constexpr int k = kk;
        double s[N];
        sfor<k, N>([&](auto i) {
            double sum = 0.0;
            sfor<0, k>([&](auto t) { sum += A[i][t] * A[t][k]; });
            s[i] = A[i][k] - sum;
        });
        // a row swap is needed where some |s_i| (i > k) exceeds |s_k|; only
        // then is the pivot row located (the first such maximum, as in the
        // sequential search)
        double mo = 0.0;
        sfor<k + 1, N>([&](auto i) { mo = fmax(mo, fabs(s[i])); });
        if (__any(mo > fabs(s[k]))) {
            int ik = k;
            double maxs = s[k];
            sfor<k + 1, N>([&](auto i) {
                const bool g = fabs(s[i]) > fabs(maxs);
                ik = g ? (int)i : ik;
                maxs = g ? s[i] : maxs;
            });
            sfor<k + 1, N>([&](auto i) {
                const bool sw = (ik == (int)i);
                if (__any(sw)) {
                    sfor<0, N>([&](auto t) {
                        const double a = A[k][t], c = A[i][t];
                        A[k][t] = sw ? c : a;
                        A[i][t] = sw ? a : c;
                    });
                    const double a = s[k], c = s[i];
                    s[k] = sw ? c : a;
                    s[i] = sw ? a : c;
                    const double u = b[k], v = b[i];
                    b[k] = sw ? v : u;
                    b[i] = sw ? u : v;
                }
            });
        }
        A[k][k] = s[k];
        sfor<k + 1, N>([&](auto j) {
            double sum = 0.0;
            sfor<0, k>([&](auto t) { sum += A[k][t] * A[t][j]; });
            A[k][j] -= sum;
        });
        dinv[k] = frcp(A[k][k]);
        sfor<k + 1, N>([&](auto i) { A[i][k] = s[i] * dinv[k]; });
    });
    sfor<0, N>([&](auto i) {
        double sum = 0.0;
        sfor<0, i>([&](auto t) { sum += A[i][t] * b[t]; });
        b[i] -= sum;
    });
    sfor<0, N>([&](auto ii) {
        constexpr int i = N - 1 - ii;
        double sum = 0.0;
        sfor<i + 1, N>([&](auto t) { sum += A[i][t] * b[t]; });
        b[i] = (b[i] - sum) * dinv[i];
    });
}

}  // namespace ohcd
