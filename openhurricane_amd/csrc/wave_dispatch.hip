// wave_dispatch.hip -- dispatch of the generic (run-time table) kernels over
// the padded ODE sizes N.  Each size is instantiated in its own translation
// unit (wave_size.hip compiled with -DOHC_WAVE_N=N), so they build in parallel.
#include "wave_dispatch.hpp"

#include "wave_kernels.hpp"

namespace ohcd {

#define OHC_WAVE_DECL(n)                                                           \
    hipError_t wave_integrate_##n(const WaveParams &, int, hipStream_t);             \
    hipError_t wave_eval_##n(const EvalParams &, const WMech &, int, hipStream_t);  \
    hipError_t wave_source_##n(const SourceParams &, const WMech &, hipStream_t);    \
    hipError_t wave_kim_##n(const KimParams &, const WMech &, int, hipStream_t);     \
    hipError_t wave_lu_##n(const LuParams &, hipStream_t);                           \
    const void *wave_fn_##n();
OHC_WAVE_DECL(12)
OHC_WAVE_DECL(16)
OHC_WAVE_DECL(24)
OHC_WAVE_DECL(33)
OHC_WAVE_DECL(40)
OHC_WAVE_DECL(48)
OHC_WAVE_DECL(55)
OHC_WAVE_DECL(64)
#undef OHC_WAVE_DECL

namespace {
struct Entry {
    int N;
    hipError_t (*integrate)(const WaveParams &, int, hipStream_t);
    hipError_t (*eval)(const EvalParams &, const WMech &, int, hipStream_t);
    hipError_t (*source)(const SourceParams &, const WMech &, hipStream_t);
    hipError_t (*kim)(const KimParams &, const WMech &, int, hipStream_t);
    hipError_t (*lu)(const LuParams &, hipStream_t);
    const void *(*fn)();
};

#define OHC_WAVE(n) \
    Entry{n,           &wave_integrate_##n, &wave_eval_##n, &wave_source_##n, &wave_kim_##n, \
          &wave_lu_##n, &wave_fn_##n}
const Entry kWave[] = {OHC_WAVE(12), OHC_WAVE(16), OHC_WAVE(24), OHC_WAVE(33),
                       OHC_WAVE(40), OHC_WAVE(48), OHC_WAVE(55), OHC_WAVE(64)};
#undef OHC_WAVE

const Entry *pick(int n) {
    for (const Entry &e : kWave)
        if (e.N >= n) return &e;
    return nullptr;
}
}  // namespace

int wave_size(int n) {
    const Entry *e = pick(n);
    return e ? e->N : -1;
}

hipError_t wave_integrate(const WaveParams &p, int grid, hipStream_t s) {
    const Entry *e = pick(p.M.ns + 1);
    return e ? e->integrate(p, grid, s) : hipErrorInvalidValue;
}

hipError_t wave_eval(const EvalParams &p, const WMech &m, int exact, hipStream_t s) {
    const Entry *e = pick(m.ns + 1);
    return e ? e->eval(p, m, exact, s) : hipErrorInvalidValue;
}

hipError_t wave_source(const SourceParams &p, const WMech &m, hipStream_t s) {
    const Entry *e = pick(m.ns + 1);
    return e ? e->source(p, m, s) : hipErrorInvalidValue;
}

hipError_t wave_kim(const KimParams &p, const WMech &m, int fp32, hipStream_t s) {
    const Entry *e = pick(m.ns + 1);
    return e ? e->kim(p, m, fp32, s) : hipErrorInvalidValue;
}

hipError_t wave_lu(const LuParams &p, hipStream_t s) {
    const Entry *e = pick(p.n);
    return e ? e->lu(p, s) : hipErrorInvalidValue;
}

// ---- mixture thermo + transport for run-time tables ----------------------
// One wavefront per cell, species k in lane k; the mixture rules are wavefront
// reductions (wsum) and the pair sums of Wilke's rule and of the
// mixture-averaged diffusivities run in lane i over j with the other
// species' values broadcast from LDS.  Same formulas as props_lane_kernel
// (JANAF.cpp:155-189, speciesList.cpp:208-222, kineticTheory.cpp:37-168,
// kineticTheory.inl:32-34, transportList.cpp:309-333).  F = double: FP64;
// F = float (split precision): the kinetic-theory species properties,
// Wilke's rule and the binary / mixture-averaged diffusivities in FP32, the
// thermo, mole fractions and wavefront sums in FP64.
constexpr int kPropWaves = 4;  // cells (wavefronts) per 256-thread block

template <class F>
__global__ __launch_bounds__(256) void props_wave_kernel(const PropParams P, const WMech M) {
    __shared__ double s_mu[kPropWaves][64], s_x[kPropWaves][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t cell = (int64_t)blockIdx.x * kPropWaves + wv;
    if (cell >= P.n) return;  // uniform per wavefront; no block barrier below
    const int ns = M.ns;
    const bool sp = lane < ns;
    const double *S = M.sp;
    const double T = P.T[cell], p = P.p[cell];
    const TFac tf = tfac(T);
    const double W = S[SP_W * 64 + lane];
    const double y = sp ? P.Y[lane * P.ldY + cell] : 0.0;
    double cpk = 0.0, hk = 0.0;
    if (sp) {
        double G, dG, Cp, H, dCp;
        wthermo(M, lane, tf, G, dG, Cp, H, dCp);
        cpk = Cp / W;
        hk = H / W;
        if (P.hk) P.hk[lane * P.ldD + cell] = hk;
    }
    const double cpm = wsum(cpk * y), ham = wsum(hk * y);
    if (lane == 0) {
        if (P.cp) P.cp[cell] = cpm;
        if (P.ha) P.ha[cell] = ham;
    }
    if (P.cv || P.ea) {
        // perfect gas: cv = cp - Rm, ea = ha - Rm T (thermoList.inl:145-156, :289-298)
        const double Wm = ns == 1 ? S[SP_W * 64] : 1.0 / fmax(wsum(sp ? y / W : 0.0), kVeryTiny);
        const double Rm = kRu / Wm;
        if (lane == 0) {
            if (P.cv) P.cv[cell] = cpm - Rm;
            if (P.ea) P.ea[cell] = ham - Rm * T;
        }
    }
    if (!(P.mu || P.kappa || P.Dk)) return;
    // Yi2Xi
    double x;
    if (ns == 1) {
        x = sp ? 1.0 : 0.0;
    } else {
        const double wm = 1.0 / fmax(wsum(sp ? y * (1.0 / W) : 0.0), kVeryTiny);
        x = sp ? fmax(0.0, fmin(1.0, y * wm / W)) : 0.0;
        const double xt = wsum(x);
        if (xt != 1.0) x *= 1.0 / xt;
    }
    // species viscosity and conductivity
    const F ekb = (F)S[SP_EKB * 64 + lane], sig = (F)S[SP_SIG * 64 + lane];
    const F Tf = (F)T, Wf = (F)W;
    F mu = 0, ka = 0;
    if (sp) {
        const F Ts = Tf * (F(1) / ekb);
        const F om = F(1.147) * exp(F(-0.145) * log(Ts)) + F(1) / ((Ts + F(0.5)) * (Ts + F(0.5)));
        mu = F(2.6693e-6) * sqrt(Wf * Tf) / ((sig * sig) * om);
        const F Ri = F(kRu) / Wf;
        ka = mu * Ri * ((F)cpk / Ri + F(1.25));
    }
    s_mu[wv][lane] = (double)mu;
    s_x[wv][lane] = x;
    __builtin_amdgcn_wave_barrier();
    if (P.mu || P.kappa) {
        double num = 0.0, nk = 0.0;
        if (ns == 1) {
            num = mu;
            nk = ka;
        } else if (sp) {
            F phi = 0;
            for (int j = 0; j < ns; ++j) {
                const F Wj = (F)S[SP_W * 64 + j];
                const F phiP = (F(1) / sqrt(F(8))) / sqrt(F(1) + Wf / Wj);
                const F tt = F(1) + sqrt(mu / (F)s_mu[wv][j] * sqrt(Wj / Wf));
                phi += phiP * (F)s_x[wv][j] * (tt * tt);
            }
            const double rphi = 1.0 / (double)phi;
            num = (double)mu * x * rphi;
            nk = (double)ka * x * rphi;
        }
        const double mum = wsum(num), km = wsum(nk);
        if (lane == 0) {
            if (P.mu) P.mu[cell] = mum;
            if (P.kappa) P.kappa[cell] = km;
        }
    }
    if (P.Dk && sp) {
        const F Tv1p5 = Tf * sqrt(Tf), rp = F(1) / (F)p;
        F sum = 0;
        for (int j = 0; j < ns; ++j) {
            if (j == lane) continue;
            const F Wj = (F)S[SP_W * 64 + j], sj = (F)S[SP_SIG * 64 + j],
                    ej = (F)S[SP_EKB * 64 + j];
            const F sij = F(0.5) * (sig + sj);
            const F cst = F(1.858e-7) * sqrt(F(1) / Wf + F(1) / Wj) / (sij * sij) * F(kPatm);
            const F Ts = Tf * (F(1) / sqrt(ekb * ej));
            const F Od = exp(F(-0.145) * log(Ts)) + F(1) / ((Ts + F(0.5)) * (Ts + F(0.5)));
            const F Dij = Tv1p5 * cst * rp / Od;
            sum += (F)s_x[wv][j] / Dij;
        }
        double Dim = x < 1.0 - kVeryTiny ? (1.0 - x) / (double)sum : 0.0;
        if (isinf(Dim)) Dim = 0.0;
        P.Dk[lane * P.ldD + cell] = Dim;
    }
}

// thermoList::T (thermoList.cpp:504-533) for run-time tables: one lane per
// cell, species loop over the table (uniform addresses across the lanes).
__global__ __launch_bounds__(256) void tfromh_table_kernel(const ThParams P, const WMech M) {
    const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= P.n) return;
    const double hstar = P.h[cell];
    const double T0 = P.T[cell], Ttol = T0 * 1.0e-4;
    double Rm = 0.0;
    if (P.energy) {  // TEa: F = ha - Rm T, dF/dT = cp - Rm (thermoList.inl:343-370)
        double s = 0.0;
        for (int k = 0; k < M.ns; ++k) s += P.Y[k * P.ldY + cell] / M.sp[SP_W * 64 + k];
        Rm = kRu / (M.ns == 1 ? M.sp[SP_W * 64] : 1.0 / fmax(s, kVeryTiny));
    }
    double Test = T0, Tnew = T0;
    int iter = 0, flag = 1;
    do {
        Test = Tnew;
        const TFac tf = tfac(Test);
        double hm = 0.0, cpm = 0.0;
        for (int k = 0; k < M.ns; ++k) {
            double G, dG, Cp, H, dCp;
            wthermo(M, k, tf, G, dG, Cp, H, dCp);
            const double W = M.sp[SP_W * 64 + k], yk = P.Y[k * P.ldY + cell];
            hm += H / W * yk;
            cpm += Cp / W * yk;
        }
        if (P.energy) {
            hm -= Rm * Test;
            cpm -= Rm;
        }
        double Tn = Test - (hm - hstar) / cpm;
        if (Tn < P.Tlow) {
            flag = 0;
            Tn = P.Tlow;
        } else if (Tn > P.Thigh) {
            flag = 2;
            Tn = P.Thigh;
        } else {
            flag = 1;
        }
        Tnew = Tn;
        if (iter++ > 100) break;
    } while (fabs(Tnew - Test) > Ttol);
    P.T[cell] = Tnew;
    if (P.iflag) P.iflag[cell] = flag;
}

hipError_t wave_props(const PropParams &p, const WMech &m, hipStream_t s) {
    const unsigned grid = (unsigned)((p.n + kPropWaves - 1) / kPropWaves);
    if (p.prec == 1)
        hipLaunchKernelGGL(props_wave_kernel<float>, dim3(grid), dim3(64 * kPropWaves), 0, s, p, m);
    else
        hipLaunchKernelGGL(props_wave_kernel<double>, dim3(grid), dim3(64 * kPropWaves), 0, s, p, m);
    return hipGetLastError();
}

hipError_t wave_tfromh(const ThParams &p, const WMech &m, hipStream_t s) {
    hipLaunchKernelGGL(tfromh_table_kernel, dim3((unsigned)((p.n + 255) / 256)), dim3(256), 0, s, p, m);
    return hipGetLastError();
}

const void *wave_integrate_fn(int n) {
    const Entry *e = pick(n);
    return e ? e->fn() : nullptr;
}

}  // namespace ohcd
