/*
 * ohc_oracle.c -- FP64 CPU oracle of OpenHurricane's per-cell chemistry path.
 *
 * TEST INFRASTRUCTURE ONLY (see ohc_oracle.h).  Plain C99 + pthreads.
 * Reference file:line citations are relative to the OpenHurricane tree;
 * abbreviations: CS = src/speciesTransport/chemistrySource,
 * RX = src/speciesTransport/reactions, TH = src/speciesTransport/thermo,
 * TR = src/speciesTransport/transport, ODE = src/Hurricane/mathematicalTools/ODEs.
 */
#include "ohc_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* src/Hurricane/common/constants/physicalConstant.hpp:117-189 */
#define ORC_R (1.380649e-23 * 6.02214076e23)
#define ORC_RU (ORC_R * 1e3)
#define ORC_PATM 1.01325e5
/* src/Hurricane/dataStructure/baseDataType/include/doubleReal.hpp:40-45 */
#define TINY 1.0e-15
#define VERYTINY 1.0e-300
#define LARGE 1.0e15
#define VERYLARGE 1.0e300
#define ROOTVERYLARGE 1.0e150

#define NMAX (OHC_MAXNS + 1)

struct orc_mech {
    int ns, nr;
    double *W, *Ri, *Tlow, *Thigh, *Tcommon;
    double *alo, *ahi; /* NASA coefficients scaled by Ri (JANAF.cpp:42-57) */
    double *ekb, *sigma;
    int *nf, *nb, *fi, *bi, *rev, *kind, *fot;
    double *fnu, *bnu, *ford, *bord, *arr, *arrlo, *arrrev, *fo, *eff, *sumStoi;
};

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }
static double sqr(double a) { return a * a; }
static double notZeroD(double s) { return s != 0 ? s : VERYTINY; }

static double *dup_d(const double *s, size_t n) {
    double *d = (double *)malloc(n * sizeof(double) + 8);
    if (s) memcpy(d, s, n * sizeof(double));
    else memset(d, 0, n * sizeof(double));
    return d;
}
static int *dup_i(const int32_t *s, size_t n) {
    int *d = (int *)malloc(n * sizeof(int) + 8);
    for (size_t i = 0; i < n; ++i) d[i] = s ? s[i] : 0;
    return d;
}

orc_mech *orc_mech_new(const ohc_mech_tables *t) {
    /* The restated rate laws are those of the reference's GPU tables
     * (reversibleReactions.cpp:64-187): PLOG and CHEB are not among them. */
    for (int r = 0; r < t->nr; ++r)
        if (t->kind[r] < OHC_RATE_ARRHENIUS || t->kind[r] > OHC_RATE_CHEMACT) return NULL;
    orc_mech *m = (orc_mech *)calloc(1, sizeof(orc_mech));
    const int ns = t->ns, nr = t->nr, P = OHC_MAXP;
    m->ns = ns;
    m->nr = nr;
    m->W = dup_d(t->W, ns);
    m->Ri = dup_d(NULL, ns);
    m->Tlow = dup_d(t->Tlow, ns);
    m->Thigh = dup_d(t->Thigh, ns);
    m->Tcommon = dup_d(t->Tcommon, ns);
    m->alo = dup_d(NULL, 7 * ns);
    m->ahi = dup_d(NULL, 7 * ns);
    for (int k = 0; k < ns; ++k) {
        /* species.cpp:47 Ri = Ru / W; JANAF.cpp:42-49 coefficients * Ri */
        m->Ri[k] = ORC_RU / m->W[k];
        for (int c = 0; c < 7; ++c) {
            m->alo[7 * k + c] = t->nasa_lo[7 * k + c] * m->Ri[k];
            m->ahi[7 * k + c] = t->nasa_hi[7 * k + c] * m->Ri[k];
        }
    }
    m->ekb = dup_d(t->ekb, ns);
    m->sigma = dup_d(t->sigma, ns);
    m->nf = dup_i(t->nf, nr);
    m->nb = dup_i(t->nb, nr);
    m->fi = dup_i(t->f_idx, (size_t)nr * P);
    m->bi = dup_i(t->b_idx, (size_t)nr * P);
    m->fnu = dup_d(t->f_nu, (size_t)nr * P);
    m->bnu = dup_d(t->b_nu, (size_t)nr * P);
    m->ford = dup_d(t->f_ord, (size_t)nr * P);
    m->bord = dup_d(t->b_ord, (size_t)nr * P);
    m->rev = dup_i(t->rev, nr);
    m->kind = dup_i(t->kind, nr);
    m->fot = dup_i(t->fo_type, nr);
    m->arr = dup_d(t->arr, (size_t)nr * 3);
    m->arrlo = dup_d(t->arr_lo, (size_t)nr * 3);
    m->arrrev = dup_d(t->arr_rev, (size_t)nr * 3);
    m->fo = dup_d(t->fo, (size_t)nr * 5);
    m->eff = dup_d(t->eff, (size_t)nr * ns);
    m->sumStoi = dup_d(NULL, nr);
    for (int r = 0; r < nr; ++r) {
        /* reactionTypes.cpp:153-160 sumStoi_ = sum(nu_b) - sum(nu_f) */
        double s = 0;
        for (int i = 0; i < m->nf[r]; ++i) s -= m->fnu[r * P + i];
        for (int i = 0; i < m->nb[r]; ++i) s += m->bnu[r * P + i];
        m->sumStoi[r] = s;
    }
    return m;
}

void orc_mech_free(orc_mech *m) {
    if (!m) return;
    double *dp[] = {m->W, m->Ri, m->Tlow, m->Thigh, m->Tcommon, m->alo, m->ahi, m->ekb,
                    m->sigma, m->fnu, m->bnu, m->ford, m->bord, m->arr, m->arrlo, m->arrrev,
                    m->fo, m->eff, m->sumStoi};
    for (size_t i = 0; i < sizeof(dp) / sizeof(dp[0]); ++i) free(dp[i]);
    int *ip[] = {m->nf, m->nb, m->fi, m->bi, m->rev, m->kind, m->fot};
    for (size_t i = 0; i < sizeof(ip) / sizeof(ip[0]); ++i) free(ip[i]);
    free(m);
}

/* ------------------------------------------------------------------ */
/* JANAF species thermo, per unit mass: TH/src/JANAF.cpp:142-219,      */
/* TH/include/JANAF.inl:34-84                                          */
/* ------------------------------------------------------------------ */
static const double *coeffs(const orc_mech *m, int k, double T) {
    return T < m->Tcommon[k] ? m->alo + 7 * k : m->ahi + 7 * k; /* JANAF.inl:34-40 */
}
static double limitT(const orc_mech *m, int k, double T) { /* JANAF.inl:43-52 */
    if (T < m->Tlow[k] || T > m->Thigh[k]) return dmin(dmax(T, m->Tlow[k]), m->Thigh[k]);
    return T;
}
static double cp0(const orc_mech *m, int k, double T) { /* JANAF.cpp:155-159 */
    const double *a = coeffs(m, k, T);
    const double T0 = limitT(m, k, T);
    return ((((a[4] * T0 + a[3]) * T0 + a[2]) * T0 + a[1]) * T0 + a[0]);
}
static double ha0(const orc_mech *m, int k, double T) { /* JANAF.cpp:161-189 */
    const double *a = coeffs(m, k, T);
    if (T < m->Tlow[k] || T > m->Thigh[k]) {
        const double T0 = T < m->Tlow[k] ? m->Tlow[k] : m->Thigh[k];
        double h0 =
            ((((a[4] / 5.0 * T0 + a[3] / 4.0) * T0 + a[2] / 3.0) * T0 + a[1] / 2.0) * T0 + a[0]) *
                T0 +
            a[5];
        double c = ((((a[4] * T0 + a[3]) * T0 + a[2]) * T0 + a[1]) * T0 + a[0]);
        h0 += c * (T - T0);
        return h0;
    }
    return (((((a[4] / 5.0 * T + a[3] / 4.0) * T + a[2] / 3.0) * T + a[1] / 2.0) * T + a[0]) * T +
            a[5]);
}
static double s0(const orc_mech *m, int k, double T) { /* JANAF.cpp:199-203 */
    const double *a = coeffs(m, k, T);
    return ((((a[4] / 4.0 * T + a[3] / 3.0) * T + a[2] / 2.0) * T + a[1]) * T +
            a[0] * log(dmax(T, TINY)) + a[6]);
}
static double g0(const orc_mech *m, int k, double T) { /* JANAF.cpp:205-212 */
    const double *a = coeffs(m, k, T);
    return (((((-a[4] / 20.0 * T - a[3] / 12.0) * T - a[2] / 6.0) * T - a[1] / 2.0) * T +
             a[0] * (1 - log(dmax(T, TINY))) - a[6]) *
                T +
            a[5]);
}
static double Ds0DT(const orc_mech *m, int k, double T) { /* JANAF.inl:80-83 */
    const double *a = coeffs(m, k, T);
    return (a[0] / dmax(T, TINY) + a[1] + (a[2] + (a[3] + a[4] * T) * T) * T);
}
static double Dcp_pDT(const orc_mech *m, int k, double T) { /* JANAF.cpp:214-217 */
    const double *a = coeffs(m, k, T);
    return (a[1] + T * (2.0 * a[2] + T * (3.0 * a[3] + 4.0 * a[4] * T)));
}
/* thermo.inl molar wrappers: Cp0 = cp0 W (:48), Ha0 = ha0 W (:158),
 * G0 = g0 W (:195), DG0DT = (Dha0DT - (s0 + T Ds0DT)) W (:199-205),
 * DCp_pDT = Dcp_pDT W (:62). */
typedef struct {
    double Cp[NMAX], dCp[NMAX], H[NMAX], G[NMAX], dG[NMAX];
} thermo_ws;

static void species_thermo(const orc_mech *m, double T, thermo_ws *w) {
    for (int k = 0; k < m->ns; ++k) {
        const double W = m->W[k];
        w->Cp[k] = cp0(m, k, T) * W;
        w->dCp[k] = Dcp_pDT(m, k, T) * W;
        w->H[k] = ha0(m, k, T) * W;
        w->G[k] = g0(m, k, T) * W;
        w->dG[k] = (cp0(m, k, T) - (s0(m, k, T) + T * Ds0DT(m, k, T))) * W;
    }
}

void orc_species_thermo(const orc_mech *m, double T, double *Cp, double *dCpdT, double *H,
                        double *G, double *dGdT, double *S) {
    thermo_ws w;
    species_thermo(m, T, &w);
    for (int k = 0; k < m->ns; ++k) {
        if (Cp) Cp[k] = w.Cp[k];
        if (dCpdT) dCpdT[k] = w.dCp[k];
        if (H) H[k] = w.H[k];
        if (G) G[k] = w.G[k];
        if (dGdT) dGdT[k] = w.dG[k];
        if (S) S[k] = s0(m, k, T) * m->W[k];
    }
}

/* ------------------------------------------------------------------ */
/* Rate constants: RX/reactionRateTypes                                */
/* ------------------------------------------------------------------ */
static double arrh_k(const double *a, double T) { /* ArrheniusRR.cpp:37-48 */
    double k = a[0];
    if (fabs(a[1]) > VERYTINY) k *= pow(T, a[1]);
    if (fabs(a[2]) > VERYTINY) k *= exp(-a[2] / T);
    return k;
}
static double arrh_dkdT(double kj, const double *a, double T) { /* ArrheniusRR.hpp:96-100 */
    double invT = 1.0 / dmax(T, TINY);
    return kj * invT * (a[1] + a[2] * invT);
}
static double third_M(const orc_mech *m, int r, const double *c) { /* thirdBodyEfficiency.hpp:58-77 */
    const double *e = m->eff + (size_t)r * m->ns;
    double M = 0.0;
    for (int i = 0; i < m->ns; ++i) M += e[i] * c[i];
    return M;
}

/* fallOffFunctions.cpp:54-168 (Lindemann: F = 1, derivatives 0) */
static double fo_F(const orc_mech *m, int r, double T, double Pr) {
    const double *f = m->fo + 5 * r;
    switch (m->fot[r]) {
    case OHC_FO_SRI: {
        double logPr = log10(dmax(Pr, TINY));
        double X = 1.0 / (1.0 + sqr(logPr));
        return f[3] * pow(f[0] * exp(-f[1] / T) + exp(-T / f[2]), X) * pow(T, f[4]);
    }
    case OHC_FO_TROE3:
    case OHC_FO_TROE4: {
        const double alpha = f[0], Tsss = f[1], Ts = f[2], Tss = f[3];
        double logFcent;
        if (m->fot[r] == OHC_FO_TROE3)
            logFcent = log10(dmax((1.0 - alpha) * exp(-T / Tsss) + alpha * exp(-T / Ts), TINY));
        else
            logFcent = log10(dmax(
                (1.0 - alpha) * exp(-T / Tsss) + alpha * exp(-T / Ts) + exp(-Tss / T), TINY));
        double c = -0.4 - 0.67 * logFcent;
        double n = 0.75 - 1.27 * logFcent;
        const double d = 0.14;
        double logPr = log10(dmax(Pr, TINY));
        return pow(10.0, logFcent / (1.0 + sqr((logPr + c) / (n - d * (logPr + c)))));
    }
    default:
        return 1.0;
    }
}
static double fo_DFDci(const orc_mech *m, int r, double T, double Pr, double F, double dPr) {
    const double *f = m->fo + 5 * r;
    switch (m->fot[r]) {
    case OHC_FO_SRI: {
        const double X = 1.0 / (1.0 + sqr(log10(dmax(Pr, TINY))));
        const double dXdc = -X * X * 2.0 * log10(Pr) / Pr / log(10.0) * dPr;
        return (F * dXdc * log(f[0] * exp(-f[1] / T) + exp(-T / f[2])));
    }
    case OHC_FO_TROE3:
    case OHC_FO_TROE4: {
        const double alpha = f[0], Tsss = f[1], Ts = f[2], Tss = f[3];
        const double logPr = log10(dmax(Pr, TINY));
        double logFcent;
        if (m->fot[r] == OHC_FO_TROE3)
            logFcent = log10(dmax((1.0 - alpha) * exp(-T / Tsss) + alpha * exp(-T / Ts), TINY));
        else
            logFcent = log10(dmax(
                (1.0 - alpha) * exp(-T / Tsss) + alpha * exp(-T / Ts) + exp(-Tss / T), TINY));
        double c = -0.4 - 0.67 * logFcent;
        double n = 0.75 - 1.27 * logFcent;
        const double d = 0.14;
        const double logPrpc = logPr + c;
        const double dlogPrpc = d * logPrpc;
        const double nmd = n - dlogPrpc;
        const double inv = logPrpc / nmd;
        return (F * ((-logFcent) / sqr(1.0 + sqr(inv))) * (2.0 * logPrpc / sqr(nmd)) *
                (1.0 + d * inv) * (1.0 / Pr * dPr));
    }
    default:
        return 0.0;
    }
}
static double fo_DFDT(const orc_mech *m, int r, double T, double Pr, double F, double dPrdT) {
    const double *f = m->fo + 5 * r;
    switch (m->fot[r]) {
    case OHC_FO_SRI: {
        const double a = f[0], b = f[1], c = f[2], e = f[4];
        const double X = 1.0 / (1.0 + sqr(log10(dmax(Pr, TINY))));
        const double dXdPr = -X * X * 2.0 * log10(Pr) / Pr / log(10.0);
        return (F * (e / T + dXdPr * dPrdT * log(a * exp(-b / T) + exp(-T / c)) +
                     X * (a * b * exp(-b / T) / sqr(T) - exp(-T / c) / c) /
                         (a * exp(-b / T) + exp(-T / c))));
    }
    case OHC_FO_TROE3:
    case OHC_FO_TROE4: {
        const double alpha = f[0], Tsss = f[1], Ts = f[2], Tss = f[3];
        const double ln10 = log(10.0);
        const double logPr = log10(dmax(Pr, TINY));
        double Fcent, DFcentDT;
        if (m->fot[r] == OHC_FO_TROE3) {
            Fcent = dmax((1.0 - alpha) * exp(-T / Tsss) + alpha * exp(-T / Ts), TINY);
            DFcentDT = (1.0 - alpha) * exp(-T / Tsss) / (-Tsss) + alpha * exp(-T / Ts) / (-Ts);
        } else {
            Fcent = dmax((1.0 - alpha) * exp(-T / Tsss) + alpha * exp(-T / Ts) + exp(-Tss / T),
                         TINY);
            DFcentDT = (1.0 - alpha) * exp(-T / Tsss) / (-Tsss) + alpha * exp(-T / Ts) / (-Ts) +
                       exp(-Tss / T) * (Tss / sqr(T));
        }
        const double logFcent = log10(Fcent);
        const double DlogFcentDT = 1.0 / (Fcent * ln10) * DFcentDT;
        const double c = -0.4 - 0.67 * logFcent;
        const double DcDT = -0.67 * DlogFcentDT;
        const double n = 0.75 - 1.27 * logFcent;
        const double DnDT = -1.27 * DlogFcentDT;
        const double d = 0.14;
        const double dlogPrDT = dPrdT / Pr / ln10;
        const double logPrpc = logPr + c;
        const double nmd = n - d * logPrpc;
        const double inv = logPrpc / nmd;
        const double dPPDT = 2.0 * logPrpc / sqr(nmd) *
                             ((dlogPrDT + DcDT) - logPrpc / nmd * (DnDT - d * (dlogPrDT + DcDT)));
        return (F * ln10 *
                (DlogFcentDT / (1.0 + sqr(inv)) - logFcent / sqr(1.0 + sqr(inv)) * dPPDT));
    }
    default:
        return 0.0;
    }
}

/* kf: reversibleReactions.hpp:58-60 -> rate law k(p, T, c) */
static double rate_k(const orc_mech *m, int r, const double *arr, double T, const double *c) {
    switch (m->kind[r]) {
    case OHC_RATE_THIRDBODY: /* ArrheniusThirdBodyRR.cpp:39-43 */
        return third_M(m, r, c) * arrh_k(arr, T);
    case OHC_RATE_FALLOFF: { /* unimolecularFallOffRR.cpp:37-44 */
        double k0 = arrh_k(m->arrlo + 3 * r, T);
        double kinf = arrh_k(m->arr + 3 * r, T);
        double Pr = k0 / notZeroD(kinf) * third_M(m, r, c);
        return kinf * (Pr / (1.0 + Pr)) * fo_F(m, r, T, Pr);
    }
    case OHC_RATE_CHEMACT: { /* chemicallyActicatedBimolecularRR.cpp:36-44 */
        double k0 = arrh_k(m->arrlo + 3 * r, T);
        double kinf = arrh_k(m->arr + 3 * r, T);
        double Pr = k0 / notZeroD(kinf) * third_M(m, r, c);
        return k0 * (1.0 / (1.0 + Pr)) * fo_F(m, r, T, Pr);
    }
    default:
        return arrh_k(arr, T);
    }
}
static double rate_dkdT(const orc_mech *m, int r, const double *arr, double kj, double T,
                        const double *c) {
    switch (m->kind[r]) {
    case OHC_RATE_FALLOFF: { /* unimolecularFallOffRR.cpp:46-54 */
        const double k0 = arrh_k(m->arrlo + 3 * r, T);
        const double kinf = arrh_k(m->arr + 3 * r, T);
        const double Pr = k0 / notZeroD(kinf) * third_M(m, r, c);
        return (Pr / (1.0 + Pr)) * fo_F(m, r, T, Pr) * arrh_dkdT(kinf, m->arr + 3 * r, T);
    }
    case OHC_RATE_CHEMACT: { /* chemicallyActicatedBimolecularRR.cpp:46-54 */
        const double k0 = arrh_k(m->arrlo + 3 * r, T);
        const double kinf = arrh_k(m->arr + 3 * r, T);
        const double Pr = k0 / notZeroD(kinf) * third_M(m, r, c);
        return (1.0 / (1.0 + Pr)) * fo_F(m, r, T, Pr) * arrh_dkdT(k0, m->arrlo + 3 * r, T);
    }
    default: /* Arrhenius and third body (ArrheniusThirdBodyRR.hpp DkDT) */
        return arrh_dkdT(kj, arr, T);
    }
}
/* gamDGamDT: 0 for Arrhenius / third body; unimolecularFallOffRR.cpp:80-99,
 * chemicallyActicatedBimolecularRR.cpp:77-95 */
static double DGGDT(const orc_mech *m, int r, double p, double T, const double *c) {
    if (m->kind[r] != OHC_RATE_FALLOFF && m->kind[r] != OHC_RATE_CHEMACT) return 0.0;
    const double M = third_M(m, r, c);
    if (!(M > TINY)) return 0.0;
    const double k0 = arrh_k(m->arrlo + 3 * r, T);
    const double kinf = arrh_k(m->arr + 3 * r, T);
    const double Pr = k0 * M / notZeroD(kinf);
    const double F = fo_F(m, r, T, Pr);
    const double dPrdT = Pr * (arrh_dkdT(k0, m->arrlo + 3 * r, T) / notZeroD(k0) -
                               arrh_dkdT(kinf, m->arr + 3 * r, T) / notZeroD(kinf));
    const double dFdT = fo_DFDT(m, r, T, Pr, F, dPrdT);
    if (m->kind[r] == OHC_RATE_FALLOFF) return (dPrdT / (Pr * (1 + Pr)) + dFdT / F);
    return (-dPrdT / (1 + Pr) + dFdT / F);
}
/* gamDGamDCi: ArrheniusThirdBodyRR.cpp:45-59, unimolecularFallOffRR.cpp:56-78,
 * chemicallyActicatedBimolecularRR.cpp:56-75 */
static void DGGDc(const orc_mech *m, int r, double p, double T, const double *c, double *g) {
    const int ns = m->ns;
    const double *e = m->eff + (size_t)r * ns;
    const double M = third_M(m, r, c);
    if (m->kind[r] == OHC_RATE_THIRDBODY) {
        for (int i = 0; i < ns; ++i) g[i] = e[i] / dmax(M, TINY);
        return;
    }
    if (!(M > TINY)) {
        for (int i = 0; i < ns; ++i) g[i] = 0;
        return;
    }
    const double k0 = arrh_k(m->arrlo + 3 * r, T);
    const double kinf = arrh_k(m->arr + 3 * r, T);
    const double Pr = k0 * M / notZeroD(kinf);
    const double F = fo_F(m, r, T, Pr);
    for (int i = 0; i < ns; ++i) {
        const double dPrdci = e[i] * k0 / notZeroD(kinf);
        const double dFdci = fo_DFDci(m, r, T, Pr, F, dPrdci);
        if (m->kind[r] == OHC_RATE_FALLOFF) g[i] = (dPrdci / (Pr * (1 + Pr)) + dFdci / F);
        else g[i] = (-dPrdci / (1 + Pr) + dFdci / F);
    }
}

/* ------------------------------------------------------------------ */
/* Equilibrium constants: RX/reactionTypes/src/reactionTypes.cpp:178-241 */
/* ------------------------------------------------------------------ */
static double Kp(const orc_mech *m, int r, double T, const thermo_ws *w) {
    const int P = OHC_MAXP;
    double Kt = 0.0;
    for (int i = 0; i < m->nf[r]; ++i) Kt -= w->G[m->fi[r * P + i]] * m->fnu[r * P + i];
    for (int i = 0; i < m->nb[r]; ++i) Kt += w->G[m->bi[r * P + i]] * m->bnu[r * P + i];
    Kt /= (dmax(T, TINY) * ORC_RU);
    const double e = exp(-Kt);
    if (isnan(e) || isinf(e)) return ROOTVERYLARGE;
    return e;
}
/* reactionTypes.cpp:198-215; exact = 1 uses dKp/dT = -Kp d(Kt)/dT (test aid). */
static double DKpDT(const orc_mech *m, int r, double T, const thermo_ws *w, int exact) {
    const int P = OHC_MAXP;
    const double kp = Kp(m, r, T, w);
    double d = 0.0;
    for (int i = 0; i < m->nf[r]; ++i) {
        int k = m->fi[r * P + i];
        d -= (w->dG[k] - w->G[k] / dmax(T, TINY)) * m->fnu[r * P + i];
    }
    for (int i = 0; i < m->nb[r]; ++i) {
        int k = m->bi[r * P + i];
        d += (w->dG[k] - w->G[k] / dmax(T, TINY)) * m->bnu[r * P + i];
    }
    d /= (dmax(T, TINY) * ORC_RU);
    const double dkpdt = exact ? -kp * d : kp * d;
    if (isnan(dkpdt) || isinf(dkpdt)) return VERYLARGE;
    return dkpdt;
}
static double Kc(const orc_mech *m, int r, double T, const thermo_ws *w) { /* :217-226 */
    const double s = m->sumStoi[r];
    if (fabs(s) < VERYTINY) return Kp(m, r, T, w);
    return dmin(Kp(m, r, T, w) * pow(ORC_PATM / (ORC_RU * T), s), VERYLARGE);
}
static double DKcDT(const orc_mech *m, int r, double T, const thermo_ws *w, int exact) { /* :228-241 */
    const double s = m->sumStoi[r];
    if (fabs(s) < VERYTINY) return DKpDT(m, r, T, w, exact);
    const double X = pow(ORC_PATM / (ORC_RU * T), s);
    const double kp = Kp(m, r, T, w);
    const double dkcdt = (DKpDT(m, r, T, w, exact) + kp * (-s) / T) * X;
    if (isnan(dkcdt) || isinf(dkcdt)) return VERYLARGE;
    return dkcdt;
}
/* kr(kf): reversibleReactions.hpp:63-70 / REV params / irreversible 0 */
static double rxn_kr(const orc_mech *m, int r, double kf, double T, const double *c,
                     const thermo_ws *w) {
    if (m->rev[r] == OHC_RXN_REVERSIBLE) {
        const double kc = Kc(m, r, T, w);
        return fabs(kc) > VERYTINY ? kf / kc : 0.0;
    }
    if (m->rev[r] == OHC_RXN_REV_PARAMS) return rate_k(m, r, m->arrrev + 3 * r, T, c);
    return 0.0;
}
static double rxn_dkrdT(const orc_mech *m, int r, double kr, double T, const double *c,
                        double dkfdT, const thermo_ws *w, int exact) {
    if (m->rev[r] == OHC_RXN_REVERSIBLE) { /* reversibleReactions.hpp:120-124 */
        const double kc = dmax(Kc(m, r, T, w), VERYTINY);
        return (dkfdT - kr * DKcDT(m, r, T, w, exact)) / kc;
    }
    if (m->rev[r] == OHC_RXN_REV_PARAMS) return arrh_dkdT(kr, m->arrrev + 3 * r, T);
    return 0.0;
}

/* q(p, T, c, qf, qr, kf, kr): reactionTypes.cpp:260-276 */
static double rxn_q(const orc_mech *m, int r, double p, double T, const double *c,
                    const thermo_ws *w, double *qf, double *qr, double *kf, double *kr) {
    const int P = OHC_MAXP;
    double f = rate_k(m, r, m->arr + 3 * r, T, c);
    double b = rxn_kr(m, r, f, T, c, w);
    *kf = f;
    *kr = b;
    for (int i = 0; i < m->nf[r]; ++i) f *= pow(dmax(c[m->fi[r * P + i]], 0.0), m->ford[r * P + i]);
    for (int i = 0; i < m->nb[r]; ++i) b *= pow(dmax(c[m->bi[r * P + i]], 0.0), m->bord[r * P + i]);
    *qf = f;
    *qr = b;
    return f - b;
}

void orc_rates(const orc_mech *m, double p, double T, const double *c, double *kf, double *kr,
               double *qf, double *qr) {
    thermo_ws w;
    species_thermo(m, T, &w);
    double cc[NMAX];
    for (int k = 0; k < m->ns; ++k) cc[k] = dmax(c[k], 0.0);
    for (int r = 0; r < m->nr; ++r) {
        double a, b, e, g;
        rxn_q(m, r, p, T, cc, &w, &a, &b, &e, &g);
        if (kf) kf[r] = e;
        if (kr) kr[r] = g;
        if (qf) qf[r] = a;
        if (qr) qr[r] = b;
    }
}

void orc_equilibrium(const orc_mech *m, double T, double *kc, double *dkc) {
    thermo_ws w;
    species_thermo(m, T, &w);
    for (int r = 0; r < m->nr; ++r) {
        if (kc) kc[r] = Kc(m, r, T, &w);
        if (dkc) dkc[r] = DKcDT(m, r, T, &w, 0);
    }
}

/* omega: chemistrySource.cpp:280-291 with reactionTypes.cpp:278-293 */
static void omega_ws(const orc_mech *m, double p, double T, const double *c, const thermo_ws *w,
                     double *dcdt) {
    const int P = OHC_MAXP;
    for (int k = 0; k < m->ns; ++k) dcdt[k] = 0.0;
    for (int r = 0; r < m->nr; ++r) {
        double qf, qr, kf, kr;
        const double q = rxn_q(m, r, p, T, c, w, &qf, &qr, &kf, &kr);
        for (int i = 0; i < m->nf[r]; ++i) dcdt[m->fi[r * P + i]] -= m->fnu[r * P + i] * q;
        for (int i = 0; i < m->nb[r]; ++i) dcdt[m->bi[r * P + i]] += m->bnu[r * P + i] * q;
    }
}

void orc_omega(const orc_mech *m, double p, double T, const double *c, double *w) {
    thermo_ws tw;
    species_thermo(m, T, &tw);
    double cc[NMAX];
    for (int k = 0; k < m->ns; ++k) cc[k] = dmax(c[k], 0.0);
    omega_ws(m, p, T, cc, &tw, w);
}

/* Mixture gas constant Rm = Ru * sum(Y/W) (speciesList::Rm) */
static double mix_Rm(const orc_mech *m, const double *Y) {
    double s = 0.0;
    for (int k = 0; k < m->ns; ++k) s += Y[k] / m->W[k];
    return ORC_RU * s;
}

/* DyDt: chemistrySource.cpp:863-907 */
void orc_dydt(const orc_mech *m, int flow, double pr, const double *c, double *dcdt) {
    const int ns = m->ns;
    const double T = c[ns];
    thermo_ws w;
    species_thermo(m, T, &w);
    double ci[NMAX] = {0};
    if (flow == 0) {
        for (int i = 0; i < ns; ++i) ci[i] = dmax(c[i], 0.0);
        omega_ws(m, pr, T, ci, &w, dcdt);
        double cCp = 0.0, DTdt = 0.0;
        for (int i = 0; i < ns; ++i) {
            cCp += c[i] * w.Cp[i];
            DTdt += dcdt[i] * w.H[i];
        }
        DTdt /= cCp;
        dcdt[ns] = -DTdt;
    } else {
        const double rho = pr;
        double yi[NMAX];
        for (int i = 0; i < ns; ++i) {
            yi[i] = c[i] * m->W[i] / rho;
            ci[i] = c[i];
        }
        const double p = rho * mix_Rm(m, yi) * T;
        omega_ws(m, p, T, ci, &w, dcdt);
        double cvmix = 0.0, DTdt = 0.0;
        for (int i = 0; i < ns; ++i) {
            /* cv_p = cp0 - Ri; ea = ha0 - Ri T (thermo.inl:68-84, :176) */
            cvmix += yi[i] * (w.Cp[i] / m->W[i] - m->Ri[i]);
            DTdt += dcdt[i] * (w.H[i] / m->W[i] - m->Ri[i] * T) * m->W[i];
        }
        DTdt /= (cvmix * rho);
        dcdt[ns] = -DTdt;
    }
}

/* jacobian: chemistrySource.cpp:967-1055; DqDci reactionTypes.cpp:354-445;
 * DqDT reactionTypes.cpp:843-878 */
void orc_jacobian(const orc_mech *m, int flow, double pr, const double *c, double *dfdt,
                  double *dfdy, int exact) {
    const int ns = m->ns, n = ns + 1, P = OHC_MAXP;
    const double T = c[ns];
    thermo_ws w;
    species_thermo(m, T, &w);
    double ci[NMAX], dgg[NMAX];
    for (int i = 0; i < ns; ++i) ci[i] = dmax(c[i], 0.0);
    for (int i = 0; i < n * n; ++i) dfdy[i] = 0.0;
    double p = pr, rho = 0.0;
    if (flow == 1) {
        rho = pr;
        double yi[NMAX];
        for (int i = 0; i < ns; ++i) yi[i] = ci[i] * m->W[i] / rho;
        p = rho * mix_Rm(m, yi) * T;
    }
    for (int i = 0; i < n; ++i) dfdt[i] = 0.0;
#define JAC(i, j) dfdy[(i) * n + (j)]
    for (int r = 0; r < m->nr; ++r) {
        double qf, qr, kf, kr;
        const double qfr = rxn_q(m, r, p, T, ci, &w, &qf, &qr, &kf, &kr);
        const int nf = m->nf[r], nb = m->nb[r];
        const int *fi = m->fi + r * P, *bi = m->bi + r * P;
        const double *fnu = m->fnu + r * P, *bnu = m->bnu + r * P;
        const double *fo = m->ford + r * P, *bo = m->bord + r * P;
        for (int i = 0; i < nf; ++i) dfdt[fi[i]] -= fnu[i] * qfr;
        for (int i = 0; i < nb; ++i) dfdt[bi[i]] += bnu[i] * qfr;
        /* DqDci */
        for (int i = 0; i < nf; ++i) {
            const int si = fi[i];
            double d = kf;
            for (int j = 0; j < nf; ++j) {
                const int sj = fi[j];
                const double ord = fo[j];
                if (i == j) {
                    if (ord == 1.0) continue;
                    else if (ord < 1.0) d *= ci[sj] > TINY ? ord * pow(ci[sj], ord - 1.0) : 0.0;
                    else d *= ord * pow(dmax(ci[sj], 0.0), ord - 1.0);
                } else {
                    d *= pow(dmax(ci[sj], 0.0), ord);
                }
            }
            for (int j = 0; j < nf; ++j) JAC(fi[j], si) -= fnu[j] * d;
            for (int j = 0; j < nb; ++j) JAC(bi[j], si) += bnu[j] * d;
        }
        for (int i = 0; i < nb; ++i) {
            const int si = bi[i];
            double d = kr;
            for (int j = 0; j < nb; ++j) {
                const int sj = bi[j];
                const double ord = bo[j];
                if (i == j) {
                    if (ord == 1.0) continue;
                    else if (ord < 1.0) d *= ci[sj] > TINY ? ord * pow(ci[sj], ord - 1.0) : 0.0;
                    else d *= ord * pow(dmax(ci[sj], 0.0), ord - 1.0);
                } else {
                    d *= pow(dmax(ci[sj], 0.0), ord);
                }
            }
            for (int j = 0; j < nf; ++j) JAC(fi[j], si) += fnu[j] * d;
            for (int j = 0; j < nb; ++j) JAC(bi[j], si) -= bnu[j] * d;
        }
        if (m->kind[r] != OHC_RATE_ARRHENIUS) {
            DGGDc(m, r, p, T, ci, dgg);
            for (int i = 0; i < ns; ++i) {
                if (isnan(dgg[i]) || isinf(dgg[i])) continue;
                for (int j = 0; j < nf; ++j) JAC(fi[j], i) -= fnu[j] * dgg[i] * qfr;
                for (int j = 0; j < nb; ++j) JAC(bi[j], i) += bnu[j] * dgg[i] * qfr;
            }
        }
        /* DqDT */
        double dkfdT = rate_dkdT(m, r, m->arr + 3 * r, kf, T, ci);
        double dkrdT = rxn_dkrdT(m, r, kr, T, ci, dkfdT, &w, exact);
        for (int i = 0; i < nf; ++i) dkfdT *= pow(dmax(ci[fi[i]], 0.0), fo[i]);
        for (int i = 0; i < nb; ++i) dkrdT *= pow(dmax(ci[bi[i]], 0.0), bo[i]);
        double dqjdT = dkfdT - dkrdT;
        double dgdT = qfr * DGGDT(m, r, p, T, ci);
        if (isnan(dqjdT) || isinf(dqjdT)) dqjdT = 0;
        if (isnan(dgdT) || isinf(dgdT)) dgdT = 0;
        for (int i = 0; i < nf; ++i) JAC(fi[i], ns) -= fnu[i] * (dqjdT + dgdT);
        for (int i = 0; i < nb; ++i) JAC(bi[i], ns) += bnu[i] * (dqjdT + dgdT);
    }
    if (flow == 0) {
        double cCp = 0.0, DTdt = 0.0, dcCpdT = 0.0;
        for (int i = 0; i < ns; ++i) {
            cCp += c[i] * w.Cp[i];
            dcCpdT += c[i] * w.dCp[i];
            DTdt += dfdt[i] * w.H[i];
        }
        DTdt /= cCp;
        dfdt[ns] = -DTdt;
        for (int i = 0; i < ns; ++i) {
            double d = JAC(ns, i);
            for (int j = 0; j < ns; ++j) d -= JAC(j, i) * w.H[j];
            d += w.Cp[i] * DTdt;
            d /= cCp;
            JAC(ns, i) = d;
        }
        double dd = 0.0;
        for (int i = 0; i < ns; ++i) dd -= (w.Cp[i] * dfdt[i] + w.H[i] * JAC(i, ns));
        dd += dcCpdT * DTdt;
        dd /= cCp;
        JAC(ns, ns) = dd;
    } else {
        double ea[NMAX], Cv[NMAX];
        for (int i = 0; i < ns; ++i) {
            ea[i] = (w.H[i] / m->W[i] - m->Ri[i] * T) * m->W[i];
            Cv[i] = (w.Cp[i] / m->W[i] - m->Ri[i]) * m->W[i];
        }
        double cCv = 0.0, DTdt = 0.0, dcCvdT = 0.0;
        for (int i = 0; i < ns; ++i) {
            cCv += c[i] * Cv[i];
            dcCvdT += c[i] * w.dCp[i];
            DTdt += dfdt[i] * ea[i];
        }
        DTdt /= cCv;
        dfdt[ns] = -DTdt;
        for (int i = 0; i < ns; ++i) {
            double d = JAC(ns, i);
            for (int j = 0; j < ns; ++j) d -= JAC(j, i) * ea[j];
            d += Cv[i] * DTdt;
            d /= cCv;
            JAC(ns, i) = d;
        }
        double dd = 0.0;
        for (int i = 0; i < ns; ++i) dd -= (Cv[i] * dfdt[i] + ea[i] * JAC(i, ns));
        dd += dcCvdT * DTdt;
        dd /= cCv;
        JAC(ns, ns) = dd;
    }
#undef JAC
}

/* ------------------------------------------------------------------ */
/* Linear algebra: SquareMatrix.hpp:185-238, LUDecompose.hpp:68-96      */
/* ------------------------------------------------------------------ */
static int lu_decompose(int n, double *A, int *M) {
    int sign = 1;
    double s[NMAX];
#define A_(i, j) A[(i) * n + (j)]
    for (int k = 0; k < n; ++k) {
        double maxs = A_(k, k);
        int ik = k;
        for (int i = k; i < n; ++i) {
            double sumlu = 0.0;
            for (int t = 0; t < k; ++t) sumlu += A_(i, t) * A_(t, k);
            s[i] = A_(i, k) - sumlu;
            if (i == k) maxs = s[i];
            if (fabs(s[i]) > fabs(maxs)) {
                ik = i;
                maxs = s[i];
            }
        }
        M[k] = ik;
        if (ik != k) {
            for (int t = 0; t < n; ++t) {
                double tmp = A_(k, t);
                A_(k, t) = A_(ik, t);
                A_(ik, t) = tmp;
            }
            double tmp = s[k];
            s[k] = s[ik];
            s[ik] = tmp;
            sign = -sign;
        }
        A_(k, k) = s[k];
        for (int j = k + 1; j < n; ++j) {
            double sumlu = 0.0;
            for (int t = 0; t < k; ++t) sumlu += A_(k, t) * A_(t, j);
            A_(k, j) -= sumlu;
        }
        for (int i = k + 1; i < n; ++i) {
            if (k < n - 1) A_(i, k) = s[i] / A_(k, k);
        }
    }
    return sign;
}
static void lu_backsub(int n, const double *A, double *b, const int *M) {
    for (int k = 0; k < n - 1; ++k) {
        int t = M[k];
        if (t != k) {
            double tmp = b[k];
            b[k] = b[t];
            b[t] = tmp;
        }
    }
    for (int i = 0; i < n; ++i) {
        double sumly = 0.0;
        for (int t = 0; t < i; ++t) sumly += A_(i, t) * b[t];
        b[i] -= sumly;
    }
    for (int i = n - 1; i >= 0; --i) {
        double sumux = 0.0;
        for (int t = i + 1; t < n; ++t) sumux += A_(i, t) * b[t];
        b[i] = (b[i] - sumux) / A_(i, i);
    }
#undef A_
}

int orc_lu_solve(int n, double *A, double *b) {
    int M[NMAX];
    int sign = lu_decompose(n, A, M);
    lu_backsub(n, A, b, M);
    return sign;
}

/* ------------------------------------------------------------------ */
/* ImplicitEuler: ODE/src/ImplicitEuler.cpp:40-342, ODEsSolver.cpp      */
/* ------------------------------------------------------------------ */
typedef struct {
    const orc_mech *m;
    const ohc_opts *o;
    int n;
    double pr; /* p0 (ConstantPressure) or rho0 (ConstantVolume) */
    double J[NMAX * NMAX], f[NMAX], yk[NMAX], yTemp[NMAX], ylast0[NMAX], ylast1[NMAX];
    double dydtLast[NMAX]; /* ImplicitTrapezoidal: f at the start of the previous step */
    double ykl[4][NMAX];   /* BDF23: step-start states, newest first */
    double lts4[4];        /* BDF23: lastTimeStep_[0..3] */
    double alpha[4];       /* BDF23: coefficients of the current order */
    int nalpha;
    double lastTimeStep[2];
    int countStep;
    int64_t steps, attempts, newton, rhs, cell_steps;
} ie_ws;

/* ImplicitEuler::maxIncrement, ImplicitEuler.cpp:40-52 */
static double maxIncrement(const ie_ws *s, const double *y0, const double *y, const double *e) {
    double maxE = 0.0;
    for (int i = 0; i < s->n; ++i) {
        if (isnan(y[i])) return 1.5;
        double tol = s->o->atol_inc + s->o->rtol_inc * dmax(fabs(y0[i]), fabs(y[i]));
        maxE = dmax(maxE, fabs(e[i]) / tol);
    }
    return maxE;
}
/* ODEsSolver::maxError, ODEsSolver.cpp:37-46 */
static double maxError(const ie_ws *s, const double *y0, const double *y, const double *e) {
    double maxE = 0.0;
    for (int i = 0; i < s->n; ++i) {
        double tol = s->o->atol + s->o->rtol * dmax(fabs(y0[i]), fabs(y[i]));
        maxE = dmax(maxE, fabs(e[i]) / tol);
    }
    return maxE;
}

/* NewtonIteration, ImplicitEuler.cpp:161-251 (autoUpdateJacobian off) */
static double newton(ie_ws *s, const double *y0, double dt, double *ykp1, int *conv) {
    const int n = s->n;
    int count = 0;
    double maxErr = 0.0, lastErr = 0.0;
    int M[NMAX];
    double err[NMAX];
    do {
        orc_jacobian(s->m, s->o->flow_type, s->pr, ykp1, s->f, s->J, s->o->jacobian);
        s->newton++;
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j < n; ++j) s->J[i * n + j] = -dt * s->J[i * n + j];
            s->J[i * n + i] += 1.0;
        }
        lu_decompose(n, s->J, M);
        for (int i = 0; i < n; ++i) err[i] = -ykp1[i] + y0[i] + dt * s->f[i];
        lu_backsub(n, s->J, err, M);
        for (int i = 0; i < n; ++i) {
            if (isnan(err[i]) || isinf(err[i])) {
                *conv = 0;
                return 1000.0;
            }
            ykp1[i] += err[i];
        }
        maxErr = maxError(s, ykp1, ykp1, err);
        if (count > 0) {
            double diver = maxErr / dmax(lastErr, TINY);
            if (diver > 2 && count > 5) {
                *conv = 0;
                return 1000.0;
            }
        }
        lastErr = maxErr;
        count++;
        if (isnan(maxErr)) {
            *conv = 0;
            return 1000.0;
        }
        if (count > 1000) {
            *conv = 0;
            return maxErr;
        }
    } while (maxErr > 1.0);
    *conv = 1;
    return maxErr;
}

/* ImplicitEuler::solve(t0, y0, dydt0, dt, y), ImplicitEuler.cpp:101-142 */
static double ie_step(ie_ws *s, const double *y0, double dt, double *y) {
    const int n = s->n;
    double e[NMAX];
    memcpy(s->yk, y0, sizeof(double) * n);
    int conv = 1;
    double errN = newton(s, y0, dt, s->yk, &conv);
    memcpy(y, s->yk, sizeof(double) * n);
    if (!conv) return errN;
    if (s->countStep == 0) {
        orc_dydt(s->m, s->o->flow_type, s->pr, s->yk, s->f);
        s->rhs++;
        for (int i = 0; i < n; ++i) e[i] = s->yk[i] - y0[i] - dt * s->f[i];
        return maxIncrement(s, y0, s->yk, e);
    }
    const double dtn = s->lastTimeStep[0], dtn1 = s->lastTimeStep[1];
    const double f1 = 2.0 / (dtn * (dtn + dtn1));
    const double f2 = 2.0 / (dtn1 * (dtn + dtn1));
    const double LTEF = sqr(dtn) * 0.5;
    for (int i = 0; i < n; ++i) {
        e[i] = f1 * (s->yk[i] - y0[i]) - f2 * (y0[i] - s->ylast1[i]);
        e[i] *= LTEF;
    }
    return maxIncrement(s, y0, s->yk, e);
}

/* chemistrySource::checkNewY, chemistrySource.cpp:1152-1164 */
static int checkNewY(ie_ws *s, double *dt, const double *yOld, const double *yNew) {
    const int ns = s->n - 1;
    if ((yNew[ns] - yOld[ns]) > s->o->T_alpha * yOld[ns]) {
        *dt *= s->o->T_beta;
        return 0;
    }
    return 1;
}

/* ImplicitEuler::adaptiveSolve, ImplicitEuler.cpp:253-292 */
static int adaptive(ie_ws *s, double *t, double *dt0, double *y) {
    const ohc_opts *o = s->o;
    double dt = *dt0, error = 0.0;
    memcpy(s->yTemp, y, sizeof(double) * s->n);
    s->lastTimeStep[1] = s->lastTimeStep[0];
    int count = 0;
    do {
        s->lastTimeStep[0] = dt;
        s->attempts++;
        error = ie_step(s, y, dt, s->yTemp);
        if (!checkNewY(s, &dt, y, s->yTemp) && count < 100) {
            error = 2;
            count++;
            if (dt <= VERYTINY) return OHC_ERR_NUMERIC;
            continue;
        }
        if (error > 1.0) {
            double scale = dmax(o->safe_scale * pow(error, -o->alpha_inc), o->min_scale);
            dt *= scale;
            if (dt <= VERYTINY) return OHC_ERR_NUMERIC;
        }
    } while (error > 1.0);
    *t += dt;
    memcpy(y, s->yTemp, sizeof(double) * s->n);
    if (error > pow(o->max_scale / o->safe_scale, -1.0 / o->alpha_inc))
        *dt0 = dmin(dmax(o->safe_scale * pow(error, -o->alpha_inc), o->min_scale), o->max_scale) *
               dt;
    else
        *dt0 = dmax(o->safe_scale * o->max_scale * dt, dt);
    s->steps++;
    s->cell_steps++;
    return 0;
}

/* ImplicitEuler::solve(t0, tn, dt0, y), ImplicitEuler.cpp:294-342 */
static int ie_solve(ie_ws *s, double *tn, double *dt0, double *y) {
    const double t0 = 0.0;
    double t = t0, dt = 0.2 * (*dt0), dtReached = t0;
    int last = 0;
    for (int n = 0; n < s->o->max_steps; ++n) {
        s->countStep = n;
        if (n == 0) {
            memcpy(s->ylast0, y, sizeof(double) * s->n);
        } else {
            memcpy(s->ylast1, s->ylast0, sizeof(double) * s->n);
            memcpy(s->ylast0, y, sizeof(double) * s->n);
        }
        double dtTry0 = dt;
        if ((t + dt - *tn) * (t + dt - t0) > 0) {
            dt = *tn - t;
            last = 1;
        } else if (t + dt == *tn) {
            last = 1;
        }
        int rc = adaptive(s, &t, &dt, y);
        if (rc) return rc;
        dtReached = t;
        if ((t - *tn) * (*tn - t0) >= 0) {
            if (n > 0 && last) *dt0 = dmin(dtTry0, *tn);
            return 0;
        }
    }
    *tn = dtReached;
    return 0;
}

/* ------------------------------------------------------------------ */
/* ImplicitTrapezoidal: ODE/src/ImplicitTrapezoidal.cpp:99-335          */
/* ------------------------------------------------------------------ */
/* NewtonIteration, ImplicitTrapezoidal.cpp:170-250 (autoUpdateJacobian off):
 * y - y0 - dt/2 (f0 + f(y)) = 0 with the matrix I - dt/2 J(y) */
static double tr_newton(ie_ws *s, const double *y0, const double *f0, double dt, double *ykp1,
                        int *conv) {
    const int n = s->n;
    int count = 0;
    double maxErr = 0.0, lastErr = 0.0;
    int M[NMAX];
    double err[NMAX];
    do {
        orc_jacobian(s->m, s->o->flow_type, s->pr, ykp1, s->f, s->J, s->o->jacobian);
        s->newton++;
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j < n; ++j) s->J[i * n + j] = -0.5 * dt * s->J[i * n + j];
            s->J[i * n + i] += 1.0;
        }
        lu_decompose(n, s->J, M);
        for (int i = 0; i < n; ++i) err[i] = -ykp1[i] + y0[i] + 0.5 * dt * (f0[i] + s->f[i]);
        lu_backsub(n, s->J, err, M);
        for (int i = 0; i < n; ++i) {
            if (isnan(err[i]) || isinf(err[i])) {
                *conv = 0;
                return 1000.0;
            }
            ykp1[i] += err[i];
        }
        maxErr = maxError(s, ykp1, ykp1, err);
        if (count > 0) {
            double diver = maxErr / dmax(lastErr, TINY);
            if (diver > 2 && count > 5) {
                *conv = 0;
                return 1000.0;
            }
        }
        lastErr = maxErr;
        count++;
        if (isnan(maxErr)) {
            *conv = 0;
            return 1000.0;
        }
        if (count > 1000) {
            *conv = 0;
            return maxErr;
        }
    } while (maxErr > 1.0);
    *conv = 1;
    return maxErr;
}

/* ImplicitTrapezoidal::solve(t0, y0, dydt0, dt, y), ImplicitTrapezoidal.cpp:99-142:
 * first step y - y0 - dt f(y); later the second difference of f, dt^3/12 */
static double tr_step(ie_ws *s, const double *y0, const double *f0, double dt, double *y) {
    const int n = s->n;
    double e[NMAX], f1v[NMAX];
    memcpy(s->yk, y0, sizeof(double) * n);
    int conv = 1;
    double errN = tr_newton(s, y0, f0, dt, s->yk, &conv);
    memcpy(y, s->yk, sizeof(double) * n);
    if (!conv) return errN;
    orc_dydt(s->m, s->o->flow_type, s->pr, s->yk, f1v);
    s->rhs++;
    if (s->countStep == 0) {
        for (int i = 0; i < n; ++i) e[i] = s->yk[i] - y0[i] - dt * f1v[i];
        return maxIncrement(s, y0, s->yk, e);
    }
    const double dtn = s->lastTimeStep[0], dtn1 = s->lastTimeStep[1];
    const double c1 = 2.0 / (dtn * (dtn + dtn1));
    const double c2 = 2.0 / (dtn1 * (dtn + dtn1));
    const double LTEF = dtn * dtn * dtn / 12.0;
    for (int i = 0; i < n; ++i) {
        e[i] = c1 * (f1v[i] - f0[i]) - c2 * (f0[i] - s->dydtLast[i]);
        e[i] *= LTEF;
    }
    return maxIncrement(s, y0, s->yk, e);
}

/* ImplicitTrapezoidal::adaptiveSolve, ImplicitTrapezoidal.cpp:253-292 */
static int tr_adaptive(ie_ws *s, double *t, double *dt0, double *y) {
    const ohc_opts *o = s->o;
    double dt = *dt0, error = 0.0, f0[NMAX];
    orc_dydt(s->m, o->flow_type, s->pr, y, f0);
    s->rhs++;
    memcpy(s->yTemp, y, sizeof(double) * s->n);
    s->lastTimeStep[1] = s->lastTimeStep[0];
    int count = 0;
    do {
        s->lastTimeStep[0] = dt;
        s->attempts++;
        error = tr_step(s, y, f0, dt, s->yTemp);
        if (!checkNewY(s, &dt, y, s->yTemp) && count < 100) {
            error = 2;
            count++;
            if (dt <= VERYTINY) return OHC_ERR_NUMERIC;
            continue;
        }
        if (error > 1.0) {
            double scale = dmax(o->safe_scale * pow(error, -o->alpha_inc), o->min_scale);
            dt *= scale;
            if (dt <= VERYTINY) return OHC_ERR_NUMERIC;
        }
    } while (error > 1.0);
    memcpy(s->dydtLast, f0, sizeof(double) * s->n);
    *t += dt;
    memcpy(y, s->yTemp, sizeof(double) * s->n);
    if (error > pow(o->max_scale / o->safe_scale, -1.0 / o->alpha_inc))
        *dt0 = dmin(dmax(o->safe_scale * pow(error, -o->alpha_inc), o->min_scale), o->max_scale) *
               dt;
    else
        *dt0 = dmax(o->safe_scale * o->max_scale * dt, dt);
    s->steps++;
    s->cell_steps++;
    return 0;
}

/* ImplicitTrapezoidal::solve(t0, tn, dt0, y), ImplicitTrapezoidal.cpp:294-335 */
static int tr_solve(ie_ws *s, double *tn, double *dt0, double *y) {
    const double t0 = 0.0;
    double t = t0, dt = 0.2 * (*dt0), dtReached = t0;
    int last = 0;
    for (int n = 0; n < s->o->max_steps; ++n) {
        s->countStep = n;
        double dtTry0 = dt;
        if ((t + dt - *tn) * (t + dt - t0) > 0) {
            dt = *tn - t;
            last = 1;
        } else if (t + dt == *tn) {
            last = 1;
        }
        int rc = tr_adaptive(s, &t, &dt, y);
        if (rc) return rc;
        dtReached = t;
        if ((t - *tn) * (*tn - t0) >= 0) {
            if (n > 0 && last) *dt0 = dmin(dtTry0, *tn);
            return 0;
        }
    }
    *tn = dtReached;
    return 0;
}

/* ------------------------------------------------------------------ */
/* BDF23 (order 2, its default): ODE/src/BDF23.cpp:54-509               */
/* ------------------------------------------------------------------ */
/* BDF23::setAlpha, BDF23.cpp:54-112 (orders 1 and 2) */
static void bdf_set_alpha(ie_ws *s, int iord) {
    if (iord == 1) {
        s->nalpha = 2;
        s->alpha[0] = 1.0;
        s->alpha[1] = -1.0;
        return;
    }
    const double Dtn = s->lts4[0], Dtn1 = s->lts4[1];
    s->nalpha = 3;
    s->alpha[1] = -(Dtn + Dtn1) / (Dtn * Dtn1);
    s->alpha[2] = Dtn / (Dtn1 * (Dtn + Dtn1));
    s->alpha[0] = -(s->alpha[1] + s->alpha[2]);
    for (int k = 0; k < 3; ++k) s->alpha[k] *= Dtn;
}

/* BDF23::NewtonIteration, BDF23.cpp:283-389 (autoUpdateJacobian off):
 * alpha0 y + yhist - dt f(y) = 0 with the matrix alpha0 I - dt J */
static double bdf_newton(ie_ws *s, const double *yh, double dt, double *ykp1, int *conv) {
    const int n = s->n;
    int count = 0;
    double maxErr = 0.0, lastErr = 0.0;
    int M[NMAX];
    double err[NMAX];
    const double a0 = s->alpha[0];
    do {
        orc_jacobian(s->m, s->o->flow_type, s->pr, ykp1, s->f, s->J, s->o->jacobian);
        s->newton++;
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j < n; ++j) s->J[i * n + j] = -dt * s->J[i * n + j];
            s->J[i * n + i] += a0;
        }
        lu_decompose(n, s->J, M);
        for (int i = 0; i < n; ++i) err[i] = -(a0 * ykp1[i] + yh[i] - dt * s->f[i]);
        lu_backsub(n, s->J, err, M);
        for (int i = 0; i < n; ++i) {
            if (isnan(err[i]) || isinf(err[i])) {
                *conv = 0;
                return 1000.0;
            }
            ykp1[i] += err[i];
        }
        maxErr = maxError(s, ykp1, ykp1, err);
        if (count > 0) {
            double diver = maxErr / dmax(lastErr, TINY);
            if (diver > 2 && count > 5) {
                *conv = 0;
                return 1000.0;
            }
        }
        lastErr = maxErr;
        count++;
        if (isnan(maxErr)) {
            *conv = 0;
            return 1000.0;
        }
        if (count > 1000) {
            *conv = 0;
            return maxErr;
        }
    } while (maxErr > 1.0);
    *conv = 1;
    return maxErr;
}

/* BDF23::solve(t0, y0, dydt0, dt, y), BDF23.cpp:176-268 (order 2): the first
 * iOrder + 1 steps of a solve estimate y - y0 - dt f(y); later ones the
 * third difference of the step-start states, (dt_n + dt_n-1)/6 */
static double bdf_step(ie_ws *s, const double *y0, double dt, double *y) {
    const int n = s->n;
    double yh[NMAX], e[NMAX];
    for (int i = 0; i < n; ++i) {
        double v = 0.0;
        for (int k = s->nalpha - 1; k > 0; --k) v += s->alpha[k] * s->ykl[k - 1][i];
        yh[i] = v;
    }
    memcpy(s->yk, y0, sizeof(double) * n);
    int conv = 1;
    double errN = bdf_newton(s, yh, dt, s->yk, &conv);
    memcpy(y, s->yk, sizeof(double) * n);
    if (!conv) return errN;
    if (s->countStep <= 2) {
        orc_dydt(s->m, s->o->flow_type, s->pr, s->yk, s->f);
        s->rhs++;
        for (int i = 0; i < n; ++i) e[i] = s->yk[i] - y0[i] - dt * s->f[i];
        return maxIncrement(s, y0, s->yk, e);
    }
    const double dtn = s->lts4[0], dtn1 = s->lts4[1], dtn2 = s->lts4[2];
    const double f1 = 1.0 / dtn;
    const double f2 = (1.0 / dtn1) * (1.0 + dtn / dtn1);
    const double f3 = dtn / (dtn1 * dtn2);
    const double ltef = (dtn + dtn1) / 6.0;
    for (int i = 0; i < n; ++i) {
        e[i] = f1 * (s->yk[i] - s->ykl[0][i]) - f2 * (s->ykl[0][i] - s->ykl[1][i]) +
               f3 * (s->ykl[1][i] - s->ykl[2][i]);
        e[i] *= ltef;
    }
    return maxIncrement(s, y0, s->yk, e);
}

/* BDF23::adaptiveSolve, BDF23.cpp:391-437: an attempt whose error grew from
 * the previous attempt's but stays below 5 is accepted as it is */
static int bdf_adaptive(ie_ws *s, int iCorder, double *t, double *dt0, double *y) {
    const ohc_opts *o = s->o;
    double dt = *dt0, error = 0.0, lastError = LARGE;
    s->lts4[3] = s->lts4[2];
    s->lts4[2] = s->lts4[1];
    s->lts4[1] = s->lts4[0];
    int count = 0;
    do {
        s->lts4[0] = dt;
        bdf_set_alpha(s, iCorder);
        s->attempts++;
        error = bdf_step(s, y, dt, s->yTemp);
        if (!checkNewY(s, &dt, y, s->yTemp) && count < 100) {
            count++;
            error = 2;
            if (dt <= VERYTINY) return OHC_ERR_NUMERIC;
            continue;
        }
        if (lastError < error && error < 5) break;
        if (error > 1.0) {
            double scale = dmax(o->safe_scale * pow(error, -o->alpha_inc), o->min_scale);
            dt *= scale;
            if (dt <= VERYTINY) return OHC_ERR_NUMERIC;
        }
        lastError = error;
    } while (error > 1.0);
    *t += dt;
    memcpy(y, s->yTemp, sizeof(double) * s->n);
    if (error > pow(o->max_scale / o->safe_scale, -1.0 / o->alpha_inc))
        *dt0 = dmin(dmax(o->safe_scale * pow(error, -o->alpha_inc), o->min_scale), o->max_scale) *
               dt;
    else
        *dt0 = dmax(o->safe_scale * o->max_scale * dt, dt);
    s->steps++;
    s->cell_steps++;
    return 0;
}

/* BDF23::solve(t0, tn, dt0, y), BDF23.cpp:439-509: first step
 * min(0.1 tn, 1e-8) at order 1, then order 2; the step doubles on steps 4
 * and 5; the history shifts before each step */
static int bdf_solve(ie_ws *s, double *tn, double *dt0, double *y) {
    const int n_ = s->n;
    const double t0 = 0.0;
    double t = t0, dt = *dt0, dtReached = t0;
    int last = 0;
    for (int n = 0; n < s->o->max_steps; ++n) {
        s->countStep = n;
        int iCorder;
        if (n == 0) {
            dt = dmin(0.1 * (*tn), 1e-8);
            iCorder = 1;
            memcpy(s->ykl[0], y, sizeof(double) * n_);
        } else if (n == 1) {
            iCorder = 2;
            memcpy(s->ykl[1], s->ykl[0], sizeof(double) * n_);
            memcpy(s->ykl[0], y, sizeof(double) * n_);
        } else {
            iCorder = 2;
            if (n <= 3) {
                memcpy(s->ykl[2], s->ykl[1], sizeof(double) * n_);
            } else {
                if (n <= 5) dt *= 2.0;
                memcpy(s->ykl[3], s->ykl[2], sizeof(double) * n_);
                memcpy(s->ykl[2], s->ykl[1], sizeof(double) * n_);
            }
            memcpy(s->ykl[1], s->ykl[0], sizeof(double) * n_);
            memcpy(s->ykl[0], y, sizeof(double) * n_);
        }
        double dtTry0 = dt;
        if ((t + dt - *tn) * (t + dt - t0) > 0) {
            dt = *tn - t;
            last = 1;
        } else if (t + dt == *tn) {
            last = 1;
        }
        int rc = bdf_adaptive(s, iCorder, &t, &dt, y);
        if (rc) return rc;
        dtReached = t;
        if ((t - *tn) * (*tn - t0) >= 0) {
            if (n > 0 && last) *dt0 = dmin(dtTry0, *tn);
            return 0;
        }
    }
    *tn = dtReached;
    return 0;
}

/* Per-cell driver: chemistrySource::solve(dt, f) body (:1513-1554) and
 * solve(t, dT, p, T, rho, yi) (:1300-1325).  mode 0: R output, 1: update. */
static int cell_integrate(ie_ws *s, double T, double p, double rho, double leftTime,
                          double *subDt, double *c) {
    const int ns = s->m->ns;
    double y[NMAX];
    s->pr = s->o->flow_type == 0 ? p : rho;
    double Tc = T;
    while (leftTime > 0.0) {
        double dtt = leftTime;
        /* ODESolver::solve, ODESolver.inl:74-90 */
        for (int i = 0; i < ns; ++i) y[i] = c[i];
        y[ns] = Tc;
        int rc = s->o->integrator == 2   ? tr_solve(s, &dtt, subDt, y)
                 : s->o->integrator == 3 ? bdf_solve(s, &dtt, subDt, y)
                                         : ie_solve(s, &dtt, subDt, y);
        if (rc) return rc;
        Tc = y[ns];
        for (int i = 0; i < ns; ++i) c[i] = y[i];
        leftTime -= dtt;
    }
    c[ns] = Tc;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Reference GPU integrator (comparison mode): explicit Euler with a      */
/* step-doubling error estimate -- EulerCUDA::{solve, adaptiveSolve,     */
/* solve(t0,tn,dt0,...)} (CS/src/EulerCUDA.cu:34-62, :130-224),          */
/* ODEsSolverCUDA::maxError (CS/src/ODEsSolverCUDA.cu:34-48) and the      */
/* per-cell loop of solveChemODEs (CS/src/CUDAChemistrySourceODEs.cu:     */
/* 588-640).  The state is y = [c_1..c_ns, T] as in the implicit path.    */
/* ------------------------------------------------------------------ */
#define EE_MAX_TINY_REJECTS 1000 /* rejected attempts at dt = tiny before a cell fails */

/* one attempt of size dt from y0 with f0 = f(y0): y <- two half steps,
 * returns max |2 (y_full - y_half)| / (atol + rtol max(|y_full|, |y_half|)) */
static double ee_attempt(ie_ws *s, const double *y0, const double *f0, double dt, double *y) {
    const int n = s->n;
    double ym[NMAX], fm[NMAX];
    for (int j = 0; j < n; ++j) ym[j] = y0[j] + 0.5 * dt * f0[j];
    orc_dydt(s->m, s->o->flow_type, s->pr, ym, fm);
    s->rhs++;
    double e = 0.0;
    for (int j = 0; j < n; ++j) {
        const double yh = ym[j] + 0.5 * dt * fm[j];
        const double yf = y0[j] + dt * f0[j];
        const double err = 2.0 / (1.0 - 2.0) * (yh - yf);
        const double tol = s->o->atol + s->o->rtol * dmax(fabs(yf), fabs(yh));
        e = dmax(e, fabs(err) / tol);
        y[j] = yh;
    }
    return e;
}

/* EulerCUDA::adaptiveSolve: y advanced in place, *dt0 the next step */
static int ee_adaptive(ie_ws *s, double *t, double *dt0, double *y) {
    const int n = s->n;
    double f0[NMAX], y1[NMAX];
    orc_dydt(s->m, s->o->flow_type, s->pr, y, f0);
    s->rhs++;
    double dt = *dt0, err;
    int tiny_rejects = 0;
    for (;;) {
        s->attempts++;
        err = ee_attempt(s, y, f0, dt, y1);
        if (!(err > 1.0)) break;
        dt *= dmax(0.8 * pow(err, -0.25), 0.25);
        if (dt <= TINY) {
            dt = TINY;
            if (++tiny_rejects > EE_MAX_TINY_REJECTS) return OHC_ERR_NUMERIC;
        }
    }
    *t += dt;
    if (err > pow(10.0 / 0.8, -1.0 / 0.25))
        *dt0 = dmin(dmax(0.8 * pow(err, -0.25), 0.25), 10.0) * dt;
    else
        *dt0 = 0.8 * 10.0 * dt;
    memcpy(y, y1, sizeof(double) * n);
    s->steps++;
    s->cell_steps++;
    return 0;
}

/* EulerCUDA::solve(odes, t0, tn, dt0, ...) */
static int ee_solve(ie_ws *s, double *tn, double *dt0, double *y) {
    const double t0 = 0.0;
    double t = t0, dt = *dt0, dtReached = t0;
    int last = 0;
    for (int n = 0; n < s->o->max_steps; ++n) {
        const double dtTry0 = dt;
        if ((t + dt - *tn) * (t + dt - t0) > 0) {
            dt = *tn - t;
            last = 1;
        } else if (t + dt == *tn) {
            last = 1;
        }
        int rc = ee_adaptive(s, &t, &dt, y);
        if (rc) return rc;
        dtReached = t;
        if ((t - *tn) * (*tn - t0) >= 0) {
            if (n > 0 && last) *dt0 = dmin(dtTry0, *tn);
            return 0;
        }
    }
    *tn = dtReached;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Reference GPU integrator, multi-time-scale variant (MTSEuler solver    */
/* type, CS/src/CUDAChemistrySource.cpp:104-105, :389-392):               */
/* solveChemODEsMTS (CS/src/CUDAChemistrySourceODEs.cu:684-778) with      */
/* timescale / MTS_whichGroup / MTS_GroupDT (:199-309) and                */
/* reactionTable::dWdciAllSpecies (RX/cuda/src/reactionTableCUDA.cu:      */
/* 331-385).  Where the reference's device code is undefined, the evident */
/* intent is restated: the time-scale sums start at zero (the reference   */
/* accumulates into uninitialised shared memory), a group that holds no   */
/* species is skipped (the reference steps with dt = 0.8 * -1), and a     */
/* cell fails after max_steps explicit steps (the reference has no cap).  */
/* Concentrations enter the rates clamped at 0, as in every other path.   */
/* ------------------------------------------------------------------ */

/* d(omega_k)/d(c_k) of the mass-action terms with kf, kr (third-body and
 * fall-off factors included) held fixed: dWdciAllSpecies */
static void mts_diag(const orc_mech *m, double T, const double *c, double *d) {
    const int P = OHC_MAXP, ns = m->ns;
    thermo_ws w;
    species_thermo(m, T, &w);
    for (int k = 0; k < ns; ++k) d[k] = 0.0;
    for (int r = 0; r < m->nr; ++r) {
        const double kf = rate_k(m, r, m->arr + 3 * r, T, c);
        const double kr = rxn_kr(m, r, kf, T, c, &w);
        for (int side = 0; side < 2; ++side) {
            const int np = side == 0 ? m->nf[r] : m->nb[r];
            const int *id = side == 0 ? m->fi + r * P : m->bi + r * P;
            const double *nu = side == 0 ? m->fnu + r * P : m->bnu + r * P;
            const double *od = side == 0 ? m->ford + r * P : m->bord + r * P;
            const int no = side == 0 ? m->nb[r] : m->nf[r];
            const int *oid = side == 0 ? m->bi + r * P : m->fi + r * P;
            const double *onu = side == 0 ? m->bnu + r * P : m->fnu + r * P;
            for (int i = 0; i < np; ++i) {
                double pf = side == 0 ? kf : kr;
                for (int k = 0; k < np; ++k) {
                    const double ck = c[id[k]], ok = od[k];
                    if (i == k) {
                        if (ok < 1.0) pf = ck > TINY ? pf * ok * pow(ck + VERYTINY, ok - 1.0) : 0.0;
                        else if (ok != 1.0) pf *= ok * pow(ck, ok - 1.0);
                    } else {
                        pf *= pow(ck, ok);
                    }
                }
                d[id[i]] -= nu[i] * pf;
                for (int j = 0; j < no; ++j)
                    if (oid[j] == id[i]) d[id[i]] += onu[j] * pf;
            }
        }
    }
}

/* timescale: tci_k = min(200, 1 / (|d omega_k / d c_k| + 1e-20)) */
void orc_mts_timescales(const orc_mech *m, double T, const double *c, double *tci) {
    double cc[NMAX], d[NMAX];
    for (int k = 0; k < m->ns; ++k) cc[k] = dmax(c[k], 0.0);
    mts_diag(m, T, cc, d);
    for (int k = 0; k < m->ns; ++k) tci[k] = dmin(200.0, 1.0 / (fabs(d[k]) + 1e-20));
}

/* MTS_whichGroup: 1 + decades of dtBase / tci (1 for tci >= dtBase) */
static int mts_group(double dtBase, double tci) {
    if (!(dtBase > tci)) return 1;
    const int g = (int)log10(dtBase / tci);
    return (g > 0 ? g : 0) + 1;
}

static int cmp_d(const void *a, const void *b) {
    const double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* One cell over tn: y = [c_1..c_ns, T] advanced in place; *Nm_out receives
 * the number of groups; returns 0 or OHC_ERR_NUMERIC (step cap). */
static int mts_cell(ie_ws *s, double tn, double *y, int *Nm_out) {
    const orc_mech *m = s->m;
    const int ns = m->ns, n = ns + 1;
    double tci[NMAX], srt[NMAX], dtG[NMAX], f[NMAX];
    int grp[NMAX];
    orc_mts_timescales(m, y[ns], y, tci);
    for (int k = 0; k < ns; ++k) grp[k] = mts_group(tn, tci[k]);
    /* MTS_GroupDT: the smallest time scale of each group, fastest group first */
    memcpy(srt, tci, sizeof(double) * ns);
    qsort(srt, (size_t)ns, sizeof(double), cmp_d);
    const int Nm = tn > srt[0] ? (int)dmax(0.0, floor(log10(tn / srt[0]))) + 1 : 1;
    for (int i = 0; i < Nm; ++i) {
        dtG[i] = -1.0;
        for (int j = 0; j < ns; ++j)
            if (Nm - mts_group(tn, srt[j]) == i) {
                dtG[i] = srt[j];
                break;
            }
    }
    if (Nm_out) *Nm_out = Nm;
    double dtDid = 0.0;
    int64_t nsteps = 0;
    for (int im = 0; im < Nm; ++im) {
        if (!(dtG[im] > 0.0)) continue;
        while (dtDid < tn) {
            if (nsteps >= s->o->max_steps) return OHC_ERR_NUMERIC;
            double dt = 0.8 * dtG[im];
            if (dtDid + dt > tn) dt = tn - dtDid;
            orc_dydt(m, s->o->flow_type, s->pr, y, f);
            s->rhs++;
            for (int k = 0; k < ns; ++k)
                if (Nm - grp[k] < im) f[k] = 0.0; /* faster groups are frozen */
            dtDid += dt;
            double emax = 0.0;
            for (int i = 0; i < n; ++i) {
                const double err = dt * f[i], ay0 = fabs(y[i]);
                y[i] += err;
                const double tol = s->o->atol + s->o->rtol * dmax(fabs(y[i]), ay0);
                /* fmax: a NaN increment (a cell the explicit steps blew up) does not
                 * decide the group's end, as in the device reductions */
                if (i < ns && Nm - grp[i] == im) emax = fmax(emax, fabs(err) / tol);
            }
            ++nsteps;
            s->steps++;
            s->attempts++;
            s->cell_steps++;
            if (im != Nm - 1 && emax < 1.0) break;
        }
    }
    return 0;
}

int orc_mts_cell(const orc_mech *m, const ohc_opts *o, double pr, double tn, double *y,
                 int *Nm, int64_t *steps) {
    ie_ws *s = (ie_ws *)calloc(1, sizeof(ie_ws));
    s->m = m;
    s->o = o;
    s->n = m->ns + 1;
    s->pr = pr;
    const int rc = mts_cell(s, tn, y, Nm);
    if (steps) *steps = s->steps;
    free(s);
    return rc;
}

typedef struct {
    const orc_mech *m;
    const ohc_opts *o;
    long lo, hi;
    int update;
    double *T;
    const double *Tin, *p, *rho, *dt;
    double *Y;
    const double *Yin;
    long ldY, ldR;
    double dtFactor, *subDt, *R;
    ohc_stats st;
    int *cell_steps;
    int rc;
} solve_job;

static void *solve_range(void *arg) {
    solve_job *j = (solve_job *)arg;
    const orc_mech *m = j->m;
    const int ns = m->ns;
    ie_ws *s = (ie_ws *)calloc(1, sizeof(ie_ws));
    s->m = m;
    s->o = j->o;
    s->n = ns + 1;
    j->st.dt_min = LARGE;
    double c[NMAX], c0[NMAX];
    for (long cell = j->lo; cell < j->hi; ++cell) {
        const double Ti = j->update ? j->T[cell] : j->Tin[cell];
        const double dtc = j->dt[cell] * j->dtFactor;
        double dtI = j->subDt[cell] > 0 ? j->subDt[cell] : dtc; /* chemistrySource.inl:86-95 */
        if (j->cell_steps) j->cell_steps[cell] = 0;
        if (j->o->integrator == 1 || j->o->integrator == 4) {
            /* CUDAChemistrySource::solve / solveUpdate: every cell, dt0 = subDt
             * (explicit Euler) or the MTS variant, which leaves subDt as it is */
            const double pi = j->p[cell], rhoi = j->rho[cell];
            const double *Ysrc = j->update ? j->Y : j->Yin;
            double y[NMAX];
            for (int k = 0; k < ns; ++k) {
                c0[k] = rhoi * Ysrc[k * j->ldY + cell] / m->W[k];
                y[k] = c0[k];
            }
            y[ns] = Ti;
            s->pr = j->o->flow_type == 0 ? pi : rhoi;
            s->cell_steps = 0;
            double left = dtc, sub = dtI;
            int rc = 0;
            if (j->o->integrator == 4) {
                sub = j->subDt[cell];
                rc = mts_cell(s, dtc, y, NULL);
            } else {
                while (left > 0.0 && !rc) {
                    double dtn = left;
                    rc = ee_solve(s, &dtn, &sub, y);
                    left -= dtn;
                }
            }
            if (rc) {
                j->rc = rc;
                break;
            }
            j->st.cells++;
            j->st.dt_min = dmin(j->st.dt_min, sub);
            if (s->cell_steps > j->st.max_cell_steps) j->st.max_cell_steps = s->cell_steps;
            if (j->cell_steps) j->cell_steps[cell] = (int)s->cell_steps;
            j->subDt[cell] = sub;
            if (!j->update) {
                for (int k = 0; k < ns; ++k)
                    j->R[k * j->ldR + cell] = (y[k] - c0[k]) / dtc * m->W[k];
            } else {
                /* device ciToYi (CUDAChemistrySourceODEs.cu:545-586): Y = c W / sum c W */
                double r = 0.0;
                for (int k = 0; k < ns; ++k) r += y[k] * m->W[k];
                for (int k = 0; k < ns; ++k) j->Y[k * j->ldY + cell] = y[k] * m->W[k] / r;
                j->T[cell] = y[ns];
            }
            continue;
        }
        if (Ti < j->o->T_threshold) {
            j->subDt[cell] = dtc;
            if (!j->update)
                for (int k = 0; k < ns; ++k) j->R[k * j->ldR + cell] = 0.0;
            continue;
        }
        const double pi = j->p[cell], rhoi = j->rho[cell];
        const double *Ysrc = j->update ? j->Y : j->Yin;
        for (int k = 0; k < ns; ++k) {
            c[k] = rhoi * Ysrc[k * j->ldY + cell] / m->W[k];
            c0[k] = c[k];
        }
        double subDt = dmin(dtc, dmin(dtI, j->o->dt_max));
        s->cell_steps = 0;
        int rc = cell_integrate(s, Ti, pi, rhoi, dtc, &subDt, c);
        if (rc) {
            j->rc = rc;
            break;
        }
        j->st.cells++;
        j->st.dt_min = dmin(j->st.dt_min, subDt);
        if (s->cell_steps > j->st.max_cell_steps) j->st.max_cell_steps = s->cell_steps;
        if (j->cell_steps) j->cell_steps[cell] = (int)s->cell_steps;
        j->subDt[cell] = subDt;
        if (!j->update) {
            for (int k = 0; k < ns; ++k)
                j->R[k * j->ldR + cell] = (c[k] - c0[k]) / dtc * m->W[k];
        } else {
            /* ciToYi(ci, yi), chemistrySource.cpp:1222-1240 */
            double yk[NMAX], r = 0.0, ys = 0.0;
            for (int k = 0; k < ns; ++k) {
                yk[k] = dmax(0.0, c[k]);
                r += yk[k] * m->W[k];
            }
            for (int k = 0; k < ns; ++k) {
                yk[k] = yk[k] * m->W[k] / dmax(r, TINY);
                ys += yk[k];
            }
            if (ys != 1.0)
                for (int k = 0; k < ns; ++k) yk[k] /= ys;
            for (int k = 0; k < ns; ++k) j->Y[k * j->ldY + cell] = yk[k];
            j->T[cell] = c[ns];
        }
    }
    j->st.steps = s->steps;
    j->st.attempts = s->attempts;
    j->st.newton_iters = s->newton;
    j->st.rhs_evals = s->rhs;
    free(s);
    return NULL;
}

static int run_jobs(solve_job *proto, long n, int nthreads, ohc_stats *stats) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > n && n > 0) nthreads = (int)n;
    solve_job *jobs = (solve_job *)calloc((size_t)nthreads, sizeof(solve_job));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    for (int t = 0; t < nthreads; ++t) {
        jobs[t] = *proto;
        jobs[t].lo = n * t / nthreads;
        jobs[t].hi = n * (t + 1) / nthreads;
        if (nthreads > 1) pthread_create(&th[t], NULL, solve_range, &jobs[t]);
    }
    if (nthreads == 1) solve_range(&jobs[0]);
    int rc = 0;
    ohc_stats s;
    memset(&s, 0, sizeof(s));
    s.dt_min = LARGE;
    for (int t = 0; t < nthreads; ++t) {
        if (nthreads > 1) pthread_join(th[t], NULL);
        if (jobs[t].rc) rc = jobs[t].rc;
        s.cells += jobs[t].st.cells;
        s.steps += jobs[t].st.steps;
        s.attempts += jobs[t].st.attempts;
        s.newton_iters += jobs[t].st.newton_iters;
        s.rhs_evals += jobs[t].st.rhs_evals;
        if (jobs[t].st.max_cell_steps > s.max_cell_steps)
            s.max_cell_steps = jobs[t].st.max_cell_steps;
        s.dt_min = dmin(s.dt_min, jobs[t].st.dt_min);
    }
    if (stats) *stats = s;
    free(jobs);
    free(th);
    return rc;
}

int orc_chem_solve(const orc_mech *m, const ohc_opts *o, long n, const double *T,
                   const double *p, const double *rho, const double *Y, long ldY,
                   const double *dt, double dtFactor, double *subDt, double *R, long ldR,
                   int nthreads, ohc_stats *stats, int *cell_steps) {
    solve_job j;
    memset(&j, 0, sizeof(j));
    j.m = m;
    j.o = o;
    j.update = 0;
    j.Tin = T;
    j.p = p;
    j.rho = rho;
    j.Yin = Y;
    j.ldY = ldY;
    j.dt = dt;
    j.dtFactor = dtFactor;
    j.subDt = subDt;
    j.R = R;
    j.ldR = ldR;
    j.cell_steps = cell_steps;
    return run_jobs(&j, n, nthreads, stats);
}

int orc_chem_solve_update(const orc_mech *m, const ohc_opts *o, long n, double *T,
                          const double *p, const double *rho, double *Y, long ldY,
                          const double *dt, double dtFactor, double *subDt, int nthreads,
                          ohc_stats *stats, int *cell_steps) {
    solve_job j;
    memset(&j, 0, sizeof(j));
    j.m = m;
    j.o = o;
    j.update = 1;
    j.T = T;
    j.p = p;
    j.rho = rho;
    j.Y = Y;
    j.ldY = ldY;
    j.dt = dt;
    j.dtFactor = dtFactor;
    j.subDt = subDt;
    j.cell_steps = cell_steps;
    return run_jobs(&j, n, nthreads, stats);
}

/* ------------------------------------------------------------------ */
/* Source terms: chemistrySource.cpp:1371-1450 with omega(rho,p,T,...)  */
/* (:319-389) and omegaCoupled (:422-480); diagonal DqDci :447-584.     */
/* ------------------------------------------------------------------ */
static void diag_dqdci(const orc_mech *m, int r, double kf, double kr, double qfr, double p,
                       double T, const double *c, double *diag, int calcLast) {
    const int ns = m->ns, P = OHC_MAXP;
    const int nf = m->nf[r], nb = m->nb[r];
    const int *fi = m->fi + r * P, *bi = m->bi + r * P;
    const double *fnu = m->fnu + r * P, *bnu = m->bnu + r * P;
    const double *fo = m->ford + r * P, *bo = m->bord + r * P;
    for (int i = 0; i < nf; ++i) {
        const int si = fi[i];
        double d = kf;
        for (int j = 0; j < nf; ++j) {
            const int sj = fi[j];
            const double ord = fo[j];
            if (i == j) {
                if (ord == 1.0) continue;
                else if (ord < 1.0) d *= c[sj] > TINY ? ord * pow(c[sj], ord - 1.0) : 0.0;
                else d *= ord * pow(dmax(c[sj], 0.0), ord - 1.0);
            } else {
                d *= pow(dmax(c[sj], 0.0), ord);
            }
        }
        if (calcLast || si != ns - 1) {
            diag[si] -= fnu[i] * d;
            for (int j = 0; j < nb; ++j)
                if (si == bi[j]) diag[si] += bnu[j] * d;
        } else {
            for (int j = 0; j < nf; ++j)
                if (fi[j] != si) diag[fi[j]] += fnu[j] * d * m->W[fi[j]] / m->W[si];
            for (int j = 0; j < nb; ++j)
                if (bi[j] != si) diag[bi[j]] -= bnu[j] * d * m->W[bi[j]] / m->W[si];
        }
    }
    for (int i = 0; i < nb; ++i) {
        const int si = bi[i];
        double d = kr;
        for (int j = 0; j < nb; ++j) {
            const int sj = bi[j];
            const double ord = bo[j];
            if (i == j) {
                if (ord == 1.0) continue;
                else if (ord < 1.0) d *= c[sj] > TINY ? ord * pow(c[sj], ord - 1.0) : 0.0;
                else d *= ord * pow(dmax(c[sj], 0.0), ord - 1.0);
            } else {
                d *= pow(dmax(c[sj], 0.0), ord);
            }
        }
        if (calcLast || si != ns - 1) {
            for (int j = 0; j < nf; ++j)
                if (si == fi[j]) diag[si] += fnu[j] * d;
            diag[si] -= bnu[i] * d;
        } else {
            for (int j = 0; j < nf; ++j)
                if (fi[j] != si) diag[fi[j]] -= fnu[j] * d * m->W[fi[j]] / m->W[si];
            for (int j = 0; j < nb; ++j)
                if (bi[j] != si) diag[bi[j]] += bnu[j] * d * m->W[bi[j]] / m->W[si];
        }
    }
    if (m->kind[r] != OHC_RATE_ARRHENIUS) {
        double g[NMAX];
        DGGDc(m, r, p, T, c, g);
        const int nsm1 = ns - 1;
        for (int j = 0; j < nf; ++j) {
            const int sj = fi[j];
            if (calcLast) diag[sj] -= fnu[j] * g[sj] * qfr;
            else if (sj != nsm1)
                diag[sj] -= fnu[j] * (g[sj] - g[nsm1] * m->W[sj] / m->W[nsm1]) * qfr;
        }
        for (int j = 0; j < nb; ++j) {
            const int sj = bi[j];
            if (calcLast) diag[sj] += bnu[j] * g[sj] * qfr;
            else if (sj != nsm1)
                diag[sj] += bnu[j] * (g[sj] - g[nsm1] * m->W[sj] / m->W[nsm1]) * qfr;
        }
    }
}

void orc_chem_source(const orc_mech *m, const ohc_opts *o, long n, const double *T,
                     const double *p, const double *rho, const double *Y, long ldY, double *R,
                     long ldR, double *diagJ, long ldD, int withoutLastSpec) {
    const int ns = m->ns, P = OHC_MAXP;
    (void)o;
    for (long cell = 0; cell < n; ++cell) {
        const double Ti = T[cell], pi = p[cell], rhoi = rho[cell];
        double c[NMAX], w[NMAX], dg[NMAX], ds[NMAX], yi[NMAX];
        for (int k = 0; k < ns; ++k) {
            c[k] = rhoi * Y[k * ldY + cell] / m->W[k];
            yi[k] = Y[k * ldY + cell];
        }
        thermo_ws tw;
        species_thermo(m, Ti, &tw);
        if (!diagJ) {
            omega_ws(m, pi, Ti, c, &tw, w);
        } else {
            /* omega(rho,p,T,c,yi,w,diagJac) / omegaCoupled non-simplified branch */
            for (int k = 0; k < ns; ++k) w[k] = dg[k] = ds[k] = 0.0;
            const int calcLast = !withoutLastSpec;
            for (int r = 0; r < m->nr; ++r) {
                double qf, qr, kf, kr;
                const double q = rxn_q(m, r, pi, Ti, c, &tw, &qf, &qr, &kf, &kr);
                for (int i = 0; i < m->nf[r]; ++i) w[m->fi[r * P + i]] -= m->fnu[r * P + i] * q;
                for (int i = 0; i < m->nb[r]; ++i) w[m->bi[r * P + i]] += m->bnu[r * P + i] * q;
                diag_dqdci(m, r, kf, kr, q, pi, Ti, c, dg, calcLast);
                /* DqDT into dwdT, reactionTypes.cpp:880-909 */
                double dkfdT = rate_dkdT(m, r, m->arr + 3 * r, kf, Ti, c);
                double dkrdT = rxn_dkrdT(m, r, kr, Ti, c, dkfdT, &tw, o->jacobian);
                for (int i = 0; i < m->nf[r]; ++i)
                    dkfdT *= pow(dmax(c[m->fi[r * P + i]], 0.0), m->ford[r * P + i]);
                for (int i = 0; i < m->nb[r]; ++i)
                    dkrdT *= pow(dmax(c[m->bi[r * P + i]], 0.0), m->bord[r * P + i]);
                const double dqjdT = dkfdT - dkrdT;
                const double dgdT = q * DGGDT(m, r, pi, Ti, c);
                for (int i = 0; i < m->nf[r]; ++i)
                    ds[m->fi[r * P + i]] -= m->fnu[r * P + i] * (dqjdT + dgdT);
                for (int i = 0; i < m->nb[r]; ++i)
                    ds[m->bi[r * P + i]] += m->bnu[r * P + i] * (dqjdT + dgdT);
            }
            /* thermoList::cv_p(p, T, yi) = sum yi (cp0_i - Ri) */
            double cv = 0.0;
            for (int k = 0; k < ns; ++k) cv += (tw.Cp[k] / m->W[k] - m->Ri[k]) * yi[k];
            if (!withoutLastSpec) {
                for (int i = 0; i < ns; ++i) {
                    const double ea = tw.H[i] / m->W[i] - m->Ri[i] * Ti;
                    dg[i] += ds[i] * m->W[i] * (-ea) / (rhoi * cv);
                }
            } else {
                const double ens = tw.H[ns - 1] / m->W[ns - 1] - m->Ri[ns - 1] * Ti;
                const double betae = 1.0 / (rhoi * cv);
                for (int i = 0; i < ns - 1; ++i) {
                    const double ea = tw.H[i] / m->W[i] - m->Ri[i] * Ti;
                    dg[i] += ds[i] * m->W[i] * (ens - ea) * betae;
                }
            }
            const int nd = withoutLastSpec ? ns - 1 : ns;
            for (int k = 0; k < nd; ++k) diagJ[k * ldD + cell] = dg[k];
        }
        for (int k = 0; k < ns; ++k) R[k * ldR + cell] = w[k] * m->W[k];
    }
}

/* ------------------------------------------------------------------ */
/* The GPU path's calculateSourceTermsImp[Async][Hybrid]                 */
/* (CS/src/CUDAChemistrySource.cpp:152-250): calcChemSourceTermsKim-     */
/* Diagonal (CS/src/CUDAChemSourceKimDiagonalKernel.cu:39-105) with      */
/* reactionTable::omegaCoupled3 / qfr<dOmegaType> (RX/cuda/include/      */
/* reactionTableCUDA.inl:142-300).  R_k = omega_k W_k in FP64; the        */
/* diagonal d(omega_k)/d(c_k) of the mass-action terms (kf, kr with their */
/* third-body and fall-off factors held fixed) for k < ns-1, with the last */
/* species eliminated by mass conservation (d c_ns / d c_k = -W_k/W_ns),  */
/* accumulated in dOmegaType: double, or float for the Hybrid variant    */
/* (every product of the diagonal in float, as the reference casts).     */
/* ------------------------------------------------------------------ */
#define KIM_DIAG(TYPE, POW)                                                                   \
    do {                                                                                      \
        for (int side = 0; side < 2; ++side) {                                                \
            const int np = side == 0 ? m->nf[r] : m->nb[r];                                   \
            const int *id = side == 0 ? m->fi + r * P : m->bi + r * P;                        \
            const double *nu = side == 0 ? m->fnu + r * P : m->bnu + r * P;                   \
            const double *od = side == 0 ? m->ford + r * P : m->bord + r * P;                 \
            const int no = side == 0 ? m->nb[r] : m->nf[r];                                   \
            const int *oid = side == 0 ? m->bi + r * P : m->fi + r * P;                       \
            const double *onu = side == 0 ? m->bnu + r * P : m->fnu + r * P;                  \
            for (int i = 0; i < np; ++i) {                                                    \
                TYPE pf = (TYPE)(side == 0 ? kf : kr);                                        \
                for (int k = 0; k < np; ++k) {                                                \
                    const TYPE ck = (TYPE)c[id[k]], ok = (TYPE)od[k];                         \
                    if (i == k) {                                                             \
                        if (ok < (TYPE)1) pf = ck > TINY ? pf * ok * POW(ck, ok - (TYPE)1) : (TYPE)0; \
                        else if (ok != (TYPE)1) pf *= ok * POW(ck, ok - (TYPE)1);             \
                    } else {                                                                  \
                        pf *= POW(ck, ok);                                                    \
                    }                                                                         \
                }                                                                             \
                const int spi = id[i];                                                        \
                if (spi != ns - 1) {                                                          \
                    acc[spi] += -(TYPE)nu[i] * pf;                                            \
                    for (int j = 0; j < no; ++j)                                              \
                        if (oid[j] == spi) acc[spi] += (TYPE)onu[j] * pf;                     \
                } else {                                                                      \
                    const TYPE Wns = (TYPE)m->W[spi];                                         \
                    for (int j = 0; j < np; ++j)                                              \
                        if (id[j] != spi) acc[id[j]] += (TYPE)m->W[id[j]] / Wns * (TYPE)nu[j] * pf; \
                    for (int j = 0; j < no; ++j)                                              \
                        if (oid[j] != spi) acc[oid[j]] += -(TYPE)m->W[oid[j]] / Wns * (TYPE)onu[j] * pf; \
                }                                                                             \
            }                                                                                 \
        }                                                                                     \
    } while (0)

void orc_chem_source_kim(const orc_mech *m, long n, const double *T, const double *p,
                         const double *rho, const double *Y, long ldY, double *R, long ldR,
                         void *diag, long ldD, int fp32) {
    const int ns = m->ns, P = OHC_MAXP;
    for (long cell = 0; cell < n; ++cell) {
        const double Ti = T[cell], pi = p[cell], rhoi = rho[cell];
        double c[NMAX], w[NMAX], dd[NMAX];
        float df[NMAX];
        for (int k = 0; k < ns; ++k) {
            c[k] = dmax(rhoi * Y[k * ldY + cell] / m->W[k], 0.0);
            w[k] = dd[k] = 0.0;
            df[k] = 0.0f;
        }
        thermo_ws tw;
        species_thermo(m, Ti, &tw);
        for (int r = 0; r < m->nr; ++r) {
            double qf, qr, kf, kr;
            const double q = rxn_q(m, r, pi, Ti, c, &tw, &qf, &qr, &kf, &kr);
            for (int i = 0; i < m->nf[r]; ++i) w[m->fi[r * P + i]] -= m->fnu[r * P + i] * q;
            for (int i = 0; i < m->nb[r]; ++i) w[m->bi[r * P + i]] += m->bnu[r * P + i] * q;
            if (fp32) {
                float *acc = df;
                KIM_DIAG(float, powf);
            } else {
                double *acc = dd;
                KIM_DIAG(double, pow);
            }
        }
        for (int k = 0; k < ns; ++k) R[k * ldR + cell] = w[k] * m->W[k];
        if (diag) {
            for (int k = 0; k < ns - 1; ++k) {
                if (fp32) ((float *)diag)[k * ldD + cell] = df[k];
                else ((double *)diag)[k * ldD + cell] = dd[k];
            }
        }
    }
}
#undef KIM_DIAG

/* ------------------------------------------------------------------ */
/* T from h: thermoList.cpp:504-533, limit :152-163                     */
/* ------------------------------------------------------------------ */
/* speciesList::MWbyYi (speciesList.cpp:106-116) -> Rm = Ru / W (speciesList.cpp:83-86) */
static double mix_Rm_yi(const orc_mech *m, const double *yi) {
    double Wm;
    if (m->ns == 1) {
        Wm = m->W[0];
    } else {
        double s = 0.0;
        for (int k = 0; k < m->ns; ++k) s += yi[k] / m->W[k];
        Wm = 1.0 / dmax(s, VERYTINY);
    }
    return ORC_RU / Wm;
}

/* thermoList::T (thermoList.cpp:504-533) with F = ha_p, dF/dT = cp_p (THa_p,
 * thermoList.inl:328-340) or, energy = 1, F = ea_p = ha_p - Rm T and dF/dT =
 * cv_p = cp_p - Rm (TEa_p / TEa_rho, thermoList.inl:343-370; perfect gas:
 * p / rho = Rm T, cpMcvm = Rm, perfectGas.hpp:524-540); limit() clamps to
 * [T_low, T_high] and sets iFlag 0 / 1 / 2 (thermoList.cpp:152-163). */
static void T_from_f(const orc_mech *m, const ohc_opts *o, long n, const double *f_in,
                     const double *Y, long ldY, double *T, int *iflag, int energy) {
    const int ns = m->ns;
    for (long cell = 0; cell < n; ++cell) {
        double yi[NMAX];
        for (int k = 0; k < ns; ++k) yi[k] = Y[k * ldY + cell];
        const double f = f_in[cell];
        const double Rm = energy ? mix_Rm_yi(m, yi) : 0.0;
        double T0 = T[cell], Test = T0, Tnew = T0, Ttol = T0 * 1.0e-4;
        int iter = 0, flag = 1;
        do {
            Test = Tnew;
            double hm = 0.0, cpm = 0.0;
            if (ns == 1) {
                hm = ha0(m, 0, Test);
                cpm = cp0(m, 0, Test);
            } else {
                for (int k = 0; k < ns; ++k) hm += ha0(m, k, Test) * yi[k];
                for (int k = 0; k < ns; ++k) cpm += cp0(m, k, Test) * yi[k];
            }
            if (energy) {
                hm -= Rm * Test;
                cpm -= Rm;
            }
            double Tn = Test - (hm - f) / cpm;
            if (Tn < o->T_low) {
                flag = 0;
                Tn = o->T_low;
            } else if (Tn > o->T_high) {
                flag = 2;
                Tn = o->T_high;
            } else {
                flag = 1;
            }
            Tnew = Tn;
            if (iter++ > 100) break;
        } while (fabs(Tnew - Test) > Ttol);
        T[cell] = Tnew;
        if (iflag) iflag[cell] = flag;
    }
}

void orc_T_from_h(const orc_mech *m, const ohc_opts *o, long n, const double *h,
                  const double *p, const double *Y, long ldY, double *T, int *iflag) {
    (void)p;
    T_from_f(m, o, n, h, Y, ldY, T, iflag, 0);
}

void orc_T_from_e(const orc_mech *m, const ohc_opts *o, long n, const double *e,
                  const double *p_or_rho, const double *Y, long ldY, double *T, int *iflag) {
    (void)p_or_rho;
    T_from_f(m, o, n, e, Y, ldY, T, iflag, 1);
}

/* Mixture cp_p, cv_p, ha_p, ea_p (thermoList.cpp:165-290, thermoList.inl:145-156,
 * :289-298): sums in species order, then cv = cp - Rm and ea = ha - Rm T. */
void orc_mix_thermo(const orc_mech *m, long n, const double *T, const double *p,
                    const double *Y, long ldY, double *cp, double *cv, double *ha, double *ea) {
    const int ns = m->ns;
    (void)p;
    for (long cell = 0; cell < n; ++cell) {
        double yi[NMAX];
        for (int k = 0; k < ns; ++k) yi[k] = Y[k * ldY + cell];
        const double Tc = T[cell];
        double cpm = 0.0, hm = 0.0;
        for (int k = 0; k < ns; ++k) {
            cpm += cp0(m, k, Tc) * yi[k];
            hm += ha0(m, k, Tc) * yi[k];
        }
        const double Rm = mix_Rm_yi(m, yi);
        if (cp) cp[cell] = cpm;
        if (cv) cv[cell] = cpm - Rm;
        if (ha) ha[cell] = hm;
        if (ea) ea[cell] = hm - Rm * Tc;
    }
}

/* ------------------------------------------------------------------ */
/* Transport: kineticTheory.cpp:37-168, kineticTheory.inl:32-34,        */
/* transportList.cpp:31-56 (phiPre, WiWj), :286-334 (muKappa),          */
/* speciesList.cpp:106-116 (MWbyYi), :208-222 (Yi2Xi)                   */
/* ------------------------------------------------------------------ */
static double kt_mu(const orc_mech *m, int k, double T) {
    const double Tstar = T * (1.0 / m->ekb[k]);
    const double om = 1.147 * exp(-0.145 * log(Tstar)) + 1.0 / sqr(Tstar + 0.5);
    return 2.6693e-6 * sqrt(m->W[k] * T) / (sqr(m->sigma[k]) * om);
}

void orc_mix_props(const orc_mech *m, long n, const double *T, const double *p,
                   const double *Y, long ldY, double *mu, double *kappa, double *Dk, long ldD,
                   double *cp, double *ha, double *hk) {
    const int ns = m->ns;
    for (long cell = 0; cell < n; ++cell) {
        const double Tc = T[cell], pc = p[cell];
        double yi[NMAX], xi[NMAX], mui[NMAX], cpi[NMAX], kap[NMAX];
        for (int k = 0; k < ns; ++k) yi[k] = Y[k * ldY + cell];
        double cpm = 0.0, ham = 0.0;
        for (int k = 0; k < ns; ++k) {
            cpi[k] = cp0(m, k, Tc);
            cpm += cpi[k] * yi[k];
            const double h = ha0(m, k, Tc);
            ham += h * yi[k];
            if (hk) hk[k * ldD + cell] = h;
        }
        if (cp) cp[cell] = cpm;
        if (ha) ha[cell] = ham;
        /* Yi2Xi */
        if (ns == 1) {
            xi[0] = 1.0;
        } else {
            double wm = 0.0, xt = 0.0;
            for (int k = 0; k < ns; ++k) wm += yi[k] / m->W[k];
            wm = 1.0 / dmax(wm, VERYTINY);
            for (int k = 0; k < ns; ++k) {
                xi[k] = yi[k] * wm / m->W[k];
                xi[k] = dmax(0.0, dmin(1.0, xi[k]));
                xt += xi[k];
            }
            if (xt != 1.0)
                for (int k = 0; k < ns; ++k) xi[k] /= xt;
        }
        for (int k = 0; k < ns; ++k) {
            mui[k] = kt_mu(m, k, Tc);
            kap[k] = mui[k] * m->Ri[k] * (cpi[k] / m->Ri[k] + 1.25);
        }
        if (mu || kappa) {
            double mum = 0.0, km = 0.0;
            if (ns == 1) {
                mum = mui[0];
                km = kap[0];
            } else {
                const double c8 = 1.0 / sqrt(8.0);
                for (int i = 0; i < ns; ++i) {
                    double phi = 0.0;
                    for (int j = 0; j < ns; ++j) {
                        const double phiP = c8 / sqrt(1.0 + m->W[i] / m->W[j]);
                        const double wiwj = sqrt(m->W[j] / m->W[i]);
                        const double t = 1.0 + sqrt(mui[i] / mui[j] * wiwj);
                        phi += phiP * xi[j] * sqr(t);
                    }
                    mum += mui[i] * xi[i] / phi;
                    km += kap[i] * xi[i] / phi;
                }
            }
            if (mu) mu[cell] = mum;
            if (kappa) kappa[cell] = km;
        }
        if (Dk) {
            const double Tv1p5 = Tc * sqrt(Tc);
            for (int i = 0; i < ns; ++i) {
                const double invW = 1.0 / m->W[i];
                double sum = 0.0;
                for (int j = 0; j < ns; ++j) {
                    if (j == i) continue;
                    const double sij = 0.5 * (m->sigma[i] + m->sigma[j]);
                    const double cst = 1.858e-7 * sqrt(invW + 1.0 / m->W[j]) / sqr(sij) * ORC_PATM;
                    const double Tstar = Tc * (1.0 / sqrt(m->ekb[i] * m->ekb[j]));
                    const double Tss = 1.0 / sqr(Tstar + 0.5);
                    const double OmegaDij = exp(-0.145 * log(Tstar)) + Tss;
                    const double Dij = Tv1p5 * cst / (pc * OmegaDij);
                    sum += xi[j] / Dij;
                }
                double Dim = 0.0;
                if (xi[i] < 1.0 - VERYTINY) Dim = (1.0 - xi[i]) / sum;
                if (isinf(Dim)) Dim = 0.0;
                Dk[i * ldD + cell] = Dim;
            }
        }
    }
}
