// ohc_api.cpp -- C-ABI of libohc.so (declared in include/ohc.h).
//
// Host side of the drop-in boundary: mechanism objects, per-device contexts
// (one HIP stream each, no globals shared between contexts), argument checks,
// host<->device staging for host-pointer calls and kernel dispatch.  Every
// entry point returns a status; nothing aborts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ohc.h"
#include "lane_kernels.hpp"
#include "mech.hpp"
#include "sched.hpp"
#include "wave_dispatch.hpp"
#include "wave_kernels.hpp"

#include "ctx_internal.hpp"

namespace ohci {

thread_local std::string g_err;

int fail(ohc_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    g_err = msg;
    return code;
}

__global__ void reset_ctl(unsigned long long *ctl) {
    const int i = threadIdx.x;
    if (i < 9) ctl[i] = (i == 7) ? 0x7ff0000000000000ull : 0ull;  // stats[6] = +inf bits
}

bool read_file(const char *path, std::string &out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream s;
    s << f.rdbuf();
    out = s.str();
    return true;
}

int check_ctx(ohc_ctx *c) {
    if (!c) return fail(nullptr, OHC_ERR_ARG, "null context");
    HIPCHK(c, hipSetDevice(c->device));
    return OHC_OK;
}

bool host_slices(ohc_ctx *c, int64_t n, size_t bytes_per_cell, int64_t *m) {
    if (c->opts.device_pointers || c->in_slices || n <= 0) return false;
    int64_t cells = c->opts.max_slice_cells;
    if (cells <= 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess || fr == 0) return false;
        cells = (int64_t)(fr / 2 / std::max<size_t>(bytes_per_cell, 1));
    }
    *m = std::max<int64_t>(cells, 1);
    return n > *m;
}

// Runs f(s, k) over the slices [s, s + k) of n cells with c->in_slices set.
template <class F>
int run_slices(ohc_ctx *c, int64_t n, int64_t m, F &&f) {
    c->in_slices = true;
    int rc = OHC_OK;
    for (int64_t s = 0; s < n && rc == OHC_OK; s += m) rc = f(s, std::min(m, n - s));
    c->in_slices = false;
    return rc;
}

void free_sched(ohc_ctx *c) {
    for (void *p : {(void *)c->d_cost, (void *)c->d_order, (void *)c->d_bins})
        if (p) (void)hipFree(p);
    c->d_cost = c->d_order = nullptr;
    c->d_bins = nullptr;
    c->sched_n = 0;
    c->sched_ready = false;
}

// Buffers of the longest-first schedule for n cells (reallocated when n changes).
int ensure_sched(ohc_ctx *c, int64_t n) {
    if (c->sched_n == n && c->d_cost) return OHC_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_sched(c);
    if (n > (int64_t)INT32_MAX) return fail(c, OHC_ERR_ARG, "schedule supports n < 2^31");
    const size_t b = sizeof(int) * (size_t)n;
    HIPCHK(c, hipMalloc((void **)&c->d_cost, b));
    HIPCHK(c, hipMalloc((void **)&c->d_order, b));
    HIPCHK(c, hipMalloc((void **)&c->d_bins, sizeof(unsigned) * 2 * ohcs::kBins));
    c->sched_n = n;
    c->sched_ready = false;
    return OHC_OK;
}

// Device tables of the generic kernels: species fields [field][64], reaction
// fields chunk-major [nrp / 64][field][64] (the 64 reactions a wavefront
// evaluates at once, field by field) -- a packed 32-bit table read for every
// reaction and an FP64 table -- and efficiencies [nr][64].
int build_wmech(ohc_ctx *c) {
    using namespace ohcd;
    const ohc::Mechanism &m = c->mech;
    const int ns = m.ns, nr = m.nr, P = OHC_MAXP;
    if (ns + 1 > 64 || wave_size(ns + 1) < 0)
        return fail(c, OHC_ERR_UNSUPPORTED, "generic kernels support ns <= 63");
    for (int r = 0; r < nr; ++r)
        if (m.nf[r] > kWP || m.nb[r] > kWP)
            return fail(c, OHC_ERR_UNSUPPORTED,
                        "generic kernels support at most 3 distinct species per reaction side");
    const int nrp = std::max(64, (nr + 63) / 64 * 64);
    std::vector<double> sp((size_t)SP_NFIELD * 64, 0.0);
    for (int k = 0; k < 64; ++k) {
        sp[SP_W * 64 + k] = 1.0;
        sp[SP_EKB * 64 + k] = 1.0;
        sp[SP_SIG * 64 + k] = 1.0;
    }
    if (m.has_transport)
        for (int k = 0; k < ns; ++k) {
            sp[SP_EKB * 64 + k] = m.ekb[k];
            sp[SP_SIG * 64 + k] = m.sigma[k];
        }
    for (int k = 0; k < ns; ++k) {
        const double Ri = ohc::kRu / m.W[k];
        sp[SP_W * 64 + k] = m.W[k];
        sp[SP_TLOW * 64 + k] = m.Tlow[k];
        sp[SP_THIGH * 64 + k] = m.Thigh[k];
        sp[SP_TCOM * 64 + k] = m.Tcommon[k];
        for (int side = 0; side < 2; ++side) {
            const double *raw = side == 0 ? &m.nasa_lo[7 * k] : &m.nasa_hi[7 * k];
            double a[7];
            for (int i = 0; i < 7; ++i) a[i] = raw[i] * Ri;
            const int oa = side == 0 ? SP_ALO : SP_AHI, oh = side == 0 ? SP_HLO : SP_HHI,
                      os = side == 0 ? SP_SLO : SP_SHI, og = side == 0 ? SP_GLO : SP_GHI;
            for (int i = 0; i < 7; ++i) sp[(oa + i) * 64 + k] = a[i];
            const double hq[4] = {a[4] / 5.0, a[3] / 4.0, a[2] / 3.0, a[1] / 2.0};
            const double sq[3] = {a[4] / 4.0, a[3] / 3.0, a[2] / 2.0};
            const double gq[4] = {a[4] / 20.0, a[3] / 12.0, a[2] / 6.0, a[1] / 2.0};
            for (int i = 0; i < 4; ++i) sp[(oh + i) * 64 + k] = hq[i];
            for (int i = 0; i < 3; ++i) sp[(os + i) * 64 + k] = sq[i];
            for (int i = 0; i < 4; ++i) sp[(og + i) * 64 + k] = gq[i];
        }
    }
    // Table slot q holds reaction perm[q]: reactions grouped by rate type,
    // reversibility and fall-off form (stable, so file order within a group),
    // so that a 64-reaction chunk of the wavefront mostly takes one code path.
    // Only the order of floating-point sums into omega and J changes.
    std::vector<int> perm(nr);
    for (int r = 0; r < nr; ++r) perm[r] = r;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) {
        auto key = [&](int r) { return m.kind[r] * 100 + m.rev[r] * 10 + m.fo_type[r]; };
        return key(a) < key(b);
    });
    auto xh = [](int f, int q) { return ((size_t)(q / 64) * RH_NFIELD + f) * 64 + q % 64; };
    auto xd = [](int f, int q) { return ((size_t)(q / 64) * RD_NFIELD + f) * 64 + q % 64; };
    auto fbits = [](double v) {
        const float f = (float)v;
        uint32_t u;
        std::memcpy(&u, &f, 4);
        return u;
    };
    std::vector<uint32_t> rh((size_t)RH_NFIELD * nrp, 0u);
    std::vector<double> rd((size_t)RD_NFIELD * nrp, 0.0), eff((size_t)std::max(nr, 1) * 64, 0.0);
    for (int q = 0; q < nr; ++q) {
        const int r = perm[q];
        double s = 0.0;
        bool gen = false;  // a coefficient, order or their sum float does not hold exactly
        auto put_f = [&](int f, double v) {
            rh[xh(f, q)] = fbits(v);
            gen = gen || (double)(float)v != v;
        };
        uint32_t fidx[kWP] = {0, 0, 0}, bidx[kWP] = {0, 0, 0};
        for (int i = 0; i < m.nf[r]; ++i) {
            fidx[i] = (uint32_t)m.f_idx[r * P + i];
            rd[xd(RD_FNU + i, q)] = m.f_nu[r * P + i];
            rd[xd(RD_FO + i, q)] = m.f_ord[r * P + i];
            put_f(RH_FNU + i, m.f_nu[r * P + i]);
            put_f(RH_FO + i, m.f_ord[r * P + i]);
            s -= m.f_nu[r * P + i];
        }
        for (int i = 0; i < m.nb[r]; ++i) {
            bidx[i] = (uint32_t)m.b_idx[r * P + i];
            rd[xd(RD_BNU + i, q)] = m.b_nu[r * P + i];
            rd[xd(RD_BO + i, q)] = m.b_ord[r * P + i];
            put_f(RH_BNU + i, m.b_nu[r * P + i]);
            put_f(RH_BO + i, m.b_ord[r * P + i]);
            s += m.b_nu[r * P + i];
        }
        rd[xd(RD_SUMS, q)] = s;
        put_f(RH_SUMS, s);
        rh[xh(RH_HDR, q)] = (uint32_t)m.nf[r] | (uint32_t)m.nb[r] << 2 | (uint32_t)m.rev[r] << 4 |
                            (uint32_t)m.kind[r] << 6 | (uint32_t)m.fo_type[r] << 9 |
                            (gen ? kRhGen : 0u);
        rh[xh(RH_IDX, q)] = fidx[0] | fidx[1] << 8 | fidx[2] << 16 | bidx[0] << 24;
        rh[xh(RH_IDX2, q)] = bidx[1] | bidx[2] << 8;
        for (int i = 0; i < 3; ++i) {
            rd[xd(RD_A + i, q)] = m.arr[3 * r + i];
            rd[xd(RD_A0 + i, q)] = m.arr_lo[3 * r + i];
            rd[xd(RD_AR + i, q)] = m.arr_rev[3 * r + i];
        }
        for (int i = 0; i < 5; ++i) rd[xd(RD_FOP + i, q)] = m.fo[5 * r + i];
        if (m.kind[r] != OHC_RATE_ARRHENIUS)
            for (int k = 0; k < ns; ++k) eff[(size_t)q * 64 + k] = m.eff[(size_t)r * ns + k];
    }
    // one device allocation, 256-byte aligned pieces
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t b_sp = al(sp.size() * 8), b_rh = al(rh.size() * 4), b_rd = al(rd.size() * 8),
                 b_ef = al(eff.size() * 8);
    const size_t total = b_sp + b_rh + b_rd + b_ef;
    std::vector<unsigned char> h(total, 0);
    size_t off = 0;
    auto put = [&](const void *src, size_t bytes, size_t span) {
        std::memcpy(h.data() + off, src, bytes);
        const size_t o = off;
        off += span;
        return o;
    };
    const size_t o_sp = put(sp.data(), sp.size() * 8, b_sp), o_rh = put(rh.data(), rh.size() * 4, b_rh),
                 o_rd = put(rd.data(), rd.size() * 8, b_rd), o_ef = put(eff.data(), eff.size() * 8, b_ef);
    HIPCHK(c, hipMalloc(&c->d_wm, total));
    HIPCHK(c, hipMemcpy(c->d_wm, h.data(), total, hipMemcpyHostToDevice));
    unsigned char *d = (unsigned char *)c->d_wm;
    c->wm.ns = ns;
    c->wm.nr = nr;
    c->wm.nrp = nrp;
    c->wm.sp = (const double *)(d + o_sp);
    c->wm.rh = (const unsigned *)(d + o_rh);
    c->wm.rd = (const double *)(d + o_rd);
    c->wm.eff = (const double *)(d + o_ef);
    return OHC_OK;
}

int read_stats(ohc_ctx *c, ohc_stats *st) {
    unsigned long long h[9];
    HIPCHK(c, hipMemcpyAsync(h, c->d_ctl, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return stats_from_words(c, h + 1, st);
}

void base_params(const ohc_ctx *c, ohcd::LaneParams &P) {
    std::memset(&P, 0, sizeof(P));
    const ohc_opts &o = c->opts;
    P.flow = o.flow_type;
    P.max_steps = o.max_steps;
    P.atol = o.atol;
    P.rtol = o.rtol;
    P.atol_inc = o.atol_inc;
    P.rtol_inc = o.rtol_inc;
    P.safe = o.safe_scale;
    P.minsc = o.min_scale;
    P.maxsc = o.max_scale;
    P.alpha = o.alpha_inc;
    P.Tthr = o.T_threshold;
    P.Talpha = o.T_alpha;
    P.Tbeta = o.T_beta;
    P.dt_max = o.dt_max;
    P.max_cell_iters = 50000000;
    P.exact_dkp = o.jacobian;
    P.prec = o.precision;
    P.integrator = o.integrator;
}

hipError_t launch_integrator(const ohc_ctx *c, const ohcd::LaneParams &P) {
    if (P.n <= 0) return hipSuccess;
    if (c->lane) {
        const int grid = (int)std::min<int64_t>(c->grid, (P.n + 63) / 64);
        return c->lane->integrate(P, grid, c->stream);
    }
    ohcd::WaveParams W;
    W.L = P;
    W.M = c->wm;
    W.exact_dkp = c->opts.jacobian;
    const int grid = (int)std::min<int64_t>(c->grid, P.n);
    return ohcd::wave_integrate(W, grid, c->stream);
}

int stats_from_words(ohc_ctx *c, const unsigned long long *h, ohc_stats *st) {
    if (st) {
        st->cells = (int64_t)h[0];
        st->steps = (int64_t)h[1];
        st->attempts = (int64_t)h[2];
        st->newton_iters = (int64_t)h[3];
        st->rhs_evals = (int64_t)h[4];
        st->max_cell_steps = (int64_t)h[5];
        double d;
        std::memcpy(&d, &h[6], sizeof(d));
        st->dt_min = h[0] ? d : 1.0e15;
    }
    if (h[7]) {
        return fail(c, OHC_ERR_NUMERIC,
                    std::to_string(h[7]) +
                        " cell(s) failed: step size underflow (dt <= 1e-300) or iteration cap");
    }
    return OHC_OK;
}

int chem_integrate(ohc_ctx *c, int mode, int64_t n, const double *T, double *Tio,
                   const double *p, const double *rho, const double *Y, double *Yio,
                   int64_t ldY, const double *dt, double dtFactor, double *subDt, double *R,
                   int64_t ldR, ohc_stats *stats, const int *list, int64_t nlist) {
    int rc = check_ctx(c);
    if (rc) return rc;
    const int ns = c->mech.ns;
    if (n < 0) return fail(c, OHC_ERR_ARG, "n < 0");
    if (n == 0) {
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->dt_min = 1.0e15;
        }
        return OHC_OK;
    }
    if (ldY < n || (mode == 0 && ldR < n))
        return fail(c, OHC_ERR_ARG, "leading dimension smaller than n");
    if (!(mode == 0 ? T : Tio) || !p || !rho || !(mode == 0 ? Y : Yio) || !dt || !subDt ||
        (mode == 0 && !R))
        return fail(c, OHC_ERR_ARG, "null array argument");
    if (!(dtFactor > 0)) return fail(c, OHC_ERR_ARG, "dtFactor must be > 0");
    if (!c->lane && !c->wave)
        return fail(c, OHC_ERR_UNSUPPORTED, "no device integrator for this mechanism");
    int64_t m = 0;
    if (!list && host_slices(c, n, sizeof(double) * (5 + 2 * (size_t)ns), &m)) {
        ohc_stats acc;
        std::memset(&acc, 0, sizeof(acc));
        acc.dt_min = 1.0e15;
        rc = run_slices(c, n, m, [&](int64_t s, int64_t k) {
            ohc_stats one;
            const int r = chem_integrate(c, mode, k, off(T, s), off(Tio, s), off(p, s),
                                         off(rho, s), off(Y, s), off(Yio, s), ldY, off(dt, s),
                                         dtFactor, off(subDt, s), off(R, s), ldR, &one, nullptr, 0);
            if (r) return r;
            acc.cells += one.cells;
            acc.steps += one.steps;
            acc.attempts += one.attempts;
            acc.newton_iters += one.newton_iters;
            acc.rhs_evals += one.rhs_evals;
            acc.max_cell_steps = std::max(acc.max_cell_steps, one.max_cell_steps);
            acc.dt_min = std::min(acc.dt_min, one.dt_min);
            return OHC_OK;
        });
        if (rc) return rc;
        if (stats) *stats = acc;
        return OHC_OK;
    }
    // the longest-first schedule needs the same cells call after call: not
    // within slices
    const bool sched = c->opts.schedule && !list && !c->in_slices;
    const size_t sv = sizeof(double) * (size_t)n;
    Staging st(c);
    void *dT, *dp, *drho, *dY, *ddt, *dsub, *dR = nullptr;
    if ((rc = st.map(mode == 0 ? (const void *)T : Tio, sv, true, mode == 1, &dT))) return rc;
    if ((rc = st.map(p, sv, true, false, &dp))) return rc;
    if ((rc = st.map(rho, sv, true, false, &drho))) return rc;
    if ((rc = st.map_rows(mode == 0 ? (const void *)Y : Yio, ns, n, ldY, sizeof(double), true,
                          mode == 1, &dY)))
        return rc;
    if ((rc = st.map(dt, sv, true, false, &ddt))) return rc;
    if ((rc = st.map(subDt, sv, true, true, &dsub))) return rc;
    if (mode == 0) {
        if ((rc = st.map_rows(R, ns, n, ldR, sizeof(double), false, true, &dR))) return rc;
    }
    ohcd::LaneParams P;
    base_params(c, P);
    const ohc_opts &o = c->opts;
    P.n = n;
    P.T = (const double *)dT;
    P.p = (const double *)dp;
    P.rho = (const double *)drho;
    P.Y = (const double *)dY;
    P.dt = (const double *)ddt;
    P.ldY = ldY;
    P.dtFactor = dtFactor;
    P.subDt = (double *)dsub;
    P.R = (double *)dR;
    P.ldR = ldR;
    P.Tout = (double *)dT;
    P.Yout = (double *)dY;
    P.mode = mode;
    if (list) {
        // integrate only the listed cells (device array of cell indices)
        P.n = nlist;
        P.order = list;
    } else if (sched) {
        if ((rc = ensure_sched(c, n))) return rc;
        P.order = c->sched_ready ? c->d_order : nullptr;
        P.cost = c->d_cost;
    }
    P.counter = c->d_ctl;
    P.stats = c->d_ctl + 1;
    hipLaunchKernelGGL(reset_ctl, dim3(1), dim3(64), 0, c->stream, c->d_ctl);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, launch_integrator(c, P));
    if (sched) {
        // next call on these cells: most expensive first
        HIPCHK(c, ohcs::order_by_cost(c->d_cost, c->d_order, c->d_bins, n, c->stream));
        c->sched_ready = true;
    }
    if ((rc = st.finish())) return rc;
    if (stats || !o.device_pointers) return read_stats(c, stats);
    return OHC_OK;
}

}  // namespace ohci

using namespace ohci;

extern "C" {

int ohc_opts_default(ohc_opts *o) {
    if (!o) return fail(nullptr, OHC_ERR_ARG, "null opts");
    std::memset(o, 0, sizeof(*o));
    o->flow_type = 0;
    o->max_steps = 10000;
    o->atol = 1.0e-15;
    o->rtol = 1.0e-4;
    o->atol_inc = 1.0e-7;
    o->rtol_inc = 1.0e-3;
    o->safe_scale = 0.8;
    o->min_scale = 0.25;
    o->max_scale = 2.0;
    o->alpha_inc = 0.25;
    o->T_threshold = 200.0;
    o->T_alpha = 0.2;
    o->T_beta = 0.25;
    o->dt_max = 1.0e15;
    o->T_low = 50.0;
    o->T_high = 5000.0;
    o->max_slice_cells = 0;
    return OHC_OK;
}

int ohc_mech_load(const char *chemkin, const char *thermo, const char *transport,
                  ohc_mech **out) {
    if (!chemkin || !out) return fail(nullptr, OHC_ERR_ARG, "null argument");
    *out = nullptr;
    std::string ck, th, tr;
    if (!read_file(chemkin, ck)) return fail(nullptr, OHC_ERR_ARG, std::string("cannot read ") + chemkin);
    if (thermo && !read_file(thermo, th)) return fail(nullptr, OHC_ERR_ARG, std::string("cannot read ") + thermo);
    if (transport && !read_file(transport, tr))
        return fail(nullptr, OHC_ERR_ARG, std::string("cannot read ") + transport);
    auto m = std::make_unique<ohc_mech>();
    try {
        ohc::parse_chemkin(ck, thermo ? &th : nullptr, transport ? &tr : nullptr, m->m);
    } catch (const std::exception &e) {
        return fail(nullptr, OHC_ERR_PARSE, e.what());
    }
    *out = m.release();
    return OHC_OK;
}

int ohc_mech_from_tables(const ohc_mech_tables *t, ohc_mech **out) {
    if (!t || !out) return fail(nullptr, OHC_ERR_ARG, "null argument");
    *out = nullptr;
    auto m = std::make_unique<ohc_mech>();
    try {
        ohc::load_from_tables(*t, m->m);
    } catch (const std::exception &e) {
        return fail(nullptr, OHC_ERR_ARG, e.what());
    }
    *out = m.release();
    return OHC_OK;
}

int ohc_mech_get_tables(const ohc_mech *m, ohc_mech_tables *out) {
    if (!m || !out) return fail(nullptr, OHC_ERR_ARG, "null argument");
    *out = m->m.tables();
    if (!m->m.has_transport) {
        out->ekb = nullptr;
        out->sigma = nullptr;
    }
    return OHC_OK;
}

int ohc_mech_species_name(const ohc_mech *m, int32_t k, const char **name) {
    if (!m || !name || k < 0 || k >= m->m.ns) return fail(nullptr, OHC_ERR_ARG, "bad species index");
    *name = m->m.species[k].c_str();
    return OHC_OK;
}

int ohc_mech_synthetic(int32_t ns, int32_t nr, uint64_t seed, double frac_rev, double frac_tb,
                       ohc_mech **out) {
    if (!out) return fail(nullptr, OHC_ERR_ARG, "null argument");
    *out = nullptr;
    auto m = std::make_unique<ohc_mech>();
    try {
        ohc::synthetic_mechanism(ns, nr, seed, frac_rev, frac_tb, m->m);
    } catch (const std::exception &e) {
        return fail(nullptr, OHC_ERR_ARG, e.what());
    }
    *out = m.release();
    return OHC_OK;
}

int ohc_mech_hash(const ohc_mech *m, uint64_t *hash) {
    if (!m || !hash) return fail(nullptr, OHC_ERR_ARG, "null argument");
    *hash = m->m.hash();
    return OHC_OK;
}

int ohc_mech_free(ohc_mech *m) {
    delete m;
    return OHC_OK;
}

int ohc_ctx_create(const ohc_mech *m, const ohc_opts *opts, int32_t device, ohc_ctx **out) {
    if (!m || !out) return fail(nullptr, OHC_ERR_ARG, "null argument");
    *out = nullptr;
    auto c = std::make_unique<ohc_ctx>();
    c->mech = m->m;
    if (opts) c->opts = *opts;
    else ohc_opts_default(&c->opts);
    const ohc_opts &o = c->opts;
    if (o.flow_type != 0 && o.flow_type != 1) return fail(nullptr, OHC_ERR_ARG, "flow_type must be 0 or 1");
    if (o.max_steps < 1) return fail(nullptr, OHC_ERR_ARG, "max_steps must be >= 1");
    if (!(o.rtol > 0) || !(o.atol >= 0) || !(o.rtol_inc > 0) || !(o.atol_inc >= 0))
        return fail(nullptr, OHC_ERR_ARG, "tolerances must be positive");
    if (!(o.safe_scale > 0) || !(o.min_scale > 0) || !(o.max_scale > 0) || !(o.alpha_inc > 0) ||
        !(o.T_beta > 0) || !(o.T_beta < 1))
        return fail(nullptr, OHC_ERR_ARG, "step-control factors out of range");
    if (o.precision != 0 && o.precision != 1) return fail(nullptr, OHC_ERR_ARG, "precision must be 0 or 1");
    if (o.jacobian != 0 && o.jacobian != 1) return fail(nullptr, OHC_ERR_ARG, "jacobian must be 0 or 1");
    if (o.kernel < 0 || o.kernel > 2) return fail(nullptr, OHC_ERR_ARG, "kernel must be 0, 1 or 2");
    // The GPU rate laws are those of the reference's GPU tables: a PLOG or
    // CHEB reaction stops it with "Unsupported reaction rate type"
    // (reversibleReactions.cpp:186, irreversibleReactions.cpp:188).
    for (int r = 0; r < c->mech.nr; ++r)
        if (c->mech.kind[r] == OHC_RATE_PLOG || c->mech.kind[r] == OHC_RATE_CHEB)
            return fail(nullptr, OHC_ERR_UNSUPPORTED,
                        "unsupported reaction rate type (" +
                            std::string(c->mech.kind[r] == OHC_RATE_PLOG ? "PLOG" : "CHEB") +
                            ") in reaction " + c->mech.rxn_names[r] +
                            ": the GPU path evaluates Arrhenius, third-body, fall-off and "
                            "chemically-activated rates");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, OHC_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(nullptr, OHC_ERR_ARG, "device index out of range");
    c->device = device;
    ohc_ctx *cp = c.get();
    HIPCHK(cp, hipSetDevice(device));
    HIPCHK(cp, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = true;
    HIPCHK(cp, hipMalloc((void **)&c->d_ctl, 16 * sizeof(unsigned long long)));
    // select the specialised kernel set
    int cnt = 0;
    const ohcd::LaneEntry *ent = ohc_lane_entries(&cnt);
    const unsigned long long h = c->mech.hash(),
                             th = c->mech.has_transport ? c->mech.hash_transport() : 0ull;
    std::string names;
    for (int i = 0; i < cnt; ++i) {
        names += (i ? ", " : "") + std::string(ent[i].name);
        if (ent[i].hash == h && (o.kernel == 0 || o.kernel == 1) && o.integrator != 3) {
            c->lane = &ent[i];
            c->lane_transport = th != 0 && ent[i].thash == th;
        }
    }
    if (o.integrator < 0 || o.integrator > 4)
        return fail(nullptr, OHC_ERR_ARG, "integrator must be 0, 1, 2, 3 or 4");
    if (o.integrator >= 2 && o.precision != 0)
        return fail(nullptr, OHC_ERR_UNSUPPORTED,
                    "the ImplicitTrapezoidal, BDF23 and MTS integrators run in FP64 only");
    if (o.integrator == 3 && o.kernel == 1)
        return fail(nullptr, OHC_ERR_UNSUPPORTED, "the BDF23 integrator runs on the generic kernels");
    if (o.integrator == 1 && (!c->lane || o.precision != 0))
        return fail(nullptr, OHC_ERR_UNSUPPORTED,
                    "the explicit Euler comparison integrator runs on specialised FP64 kernels only");
    if (!c->lane && o.kernel == 1)
        return fail(nullptr, OHC_ERR_UNSUPPORTED,
                    "no specialised kernel compiled for this mechanism; compiled: " +
                        (names.empty() ? std::string("none") : names));
    if (!c->lane) {
        int rc = build_wmech(cp);
        if (rc) return fail(nullptr, rc, cp->err);
        c->wave = true;
    } else if (build_wmech(cp) != OHC_OK) {
        // a compiled mechanism the run-time tables cannot hold: only the
        // entry points that need them (ohc_chem_source_kim) are unavailable
        if (c->d_wm) (void)hipFree(c->d_wm);
        c->d_wm = nullptr;
        cp->err.clear();
    }
    {
        hipDeviceProp_t prop;
        HIPCHK(cp, hipGetDeviceProperties(&prop, device));
        int per_cu = 0;
        const void *fn = c->lane ? c->lane->integrate_fn : ohcd::wave_integrate_fn(c->mech.ns + 1);
        HIPCHK(cp, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64, 0));
        c->cus = prop.multiProcessorCount;
        c->max_per_cu = std::max(per_cu, 1);
        if (o.waves_per_cu > 0) per_cu = std::min(per_cu, o.waves_per_cu);
        if (per_cu < 1) per_cu = 1;
        c->grid = prop.multiProcessorCount * per_cu;
        c->kname = c->lane ? std::string("lane:") + c->lane->name
                           : "wave:" + std::to_string(ohcd::wave_size(c->mech.ns + 1));
    }
    *out = c.release();
    return OHC_OK;
}

int ohc_ctx_free(ohc_ctx *c) {
    if (!c) return OHC_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void *p : c->stage)
        if (p) (void)hipFree(p);
    if (c->d_ctl) (void)hipFree(c->d_ctl);
    if (c->d_wm) (void)hipFree(c->d_wm);
    free_sched(c);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return OHC_OK;
}

int ohc_ctx_set_stream(ohc_ctx *c, void *stream) {
    int rc = check_ctx(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->own_stream) HIPCHK(c, hipStreamDestroy(c->stream));
    c->stream = (hipStream_t)stream;
    c->own_stream = false;
    return OHC_OK;
}

int ohc_ctx_autotune(ohc_ctx *c, int64_t n, const double *T, const double *p, const double *rho,
                     const double *Y, int64_t ldY, const double *dt, double dtFactor,
                     int32_t *waves_per_cu, double *best_ms) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (n < 1 || ldY < n || !T || !p || !rho || !Y || !dt || !(dtFactor > 0))
        return fail(c, OHC_ERR_ARG, "bad arguments");
    const int ns = c->mech.ns;
    const size_t sv = sizeof(double) * (size_t)n;
    const size_t sY = sizeof(double) * (size_t)(ldY * (ns - 1) + n);
    const bool dev = c->opts.device_pointers != 0;
    // private copies of the advanced state: the caller's arrays stay untouched
    std::vector<double> hT, hY, hsub;
    double *dT = nullptr, *dY = nullptr, *dsub = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // every HIP failure below records the error and falls through to the
    // common clean-up, which frees what was allocated and restores the grid
    auto ok = [&](hipError_t e, const char *what) {
        if (e == hipSuccess) return true;
        if (!rc) rc = fail(c, OHC_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
        return false;
    };
    if (dev) {
        if (ok(hipMalloc((void **)&dT, sv), "hipMalloc") && ok(hipMalloc((void **)&dY, sY), "hipMalloc"))
            ok(hipMalloc((void **)&dsub, sv), "hipMalloc");
    } else {
        hsub.assign((size_t)n, 0.0);
    }
    if (!rc && ok(hipEventCreate(&e0), "hipEventCreate")) ok(hipEventCreate(&e1), "hipEventCreate");
    auto restore = [&]() -> int {
        if (dev) {
            HIPCHK(c, hipMemcpyAsync(dT, T, sv, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(dY, Y, sY, hipMemcpyDeviceToDevice, c->stream));
        } else {
            hT.assign(T, T + n);
            hY.assign(Y, Y + (size_t)(ldY * (ns - 1) + n));
        }
        return OHC_OK;
    };
    auto advance = [&]() -> int {
        return dev ? chem_integrate(c, 1, n, nullptr, dT, p, rho, nullptr, dY, ldY, dt, dtFactor,
                                    dsub, nullptr, n, nullptr, nullptr, 0)
                   : chem_integrate(c, 1, n, nullptr, hT.data(), p, rho, nullptr, hY.data(), ldY,
                                    dt, dtFactor, hsub.data(), nullptr, n, nullptr, nullptr, 0);
    };
    const int keep = c->grid;
    int best = 0;
    float best_t = 0.0f;
    // from the most resident blocks down; stop once a size is clearly slower
    // than the best so far (the step time is unimodal in the grid size)
    for (int w = c->max_per_cu; w >= 1 && !rc; --w) {
        c->grid = c->cus * w;
        if (dev) {
            if (!ok(hipMemsetAsync(dsub, 0, sv, c->stream), "hipMemsetAsync")) break;
        } else {
            std::fill(hsub.begin(), hsub.end(), 0.0);
        }
        // first flow step sets the sub-step memory and the cell order; the
        // second, from the same state, is timed
        if ((rc = restore()) || (rc = advance()) || (rc = restore())) break;
        if (!ok(hipEventRecord(e0, c->stream), "hipEventRecord")) break;
        if ((rc = advance())) break;
        float ms = 0.0f;
        if (!ok(hipEventRecord(e1, c->stream), "hipEventRecord") ||
            !ok(hipEventSynchronize(e1), "hipEventSynchronize") ||
            !ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime"))
            break;
        if (best == 0 || ms < best_t) {
            best = w;
            best_t = ms;
        } else if (ms > 1.03f * best_t) {
            break;
        }
    }
    (void)hipStreamSynchronize(c->stream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (dT) (void)hipFree(dT);
    if (dY) (void)hipFree(dY);
    if (dsub) (void)hipFree(dsub);
    if (rc || best == 0) {
        c->grid = keep;
        return rc ? rc : fail(c, OHC_ERR_DEVICE, "autotune found no working configuration");
    }
    c->grid = c->cus * best;
    c->opts.waves_per_cu = best;
    if (waves_per_cu) *waves_per_cu = best;
    if (best_ms) *best_ms = best_t;
    return OHC_OK;
}

int ohc_sync(ohc_ctx *c) {
    int rc = check_ctx(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OHC_OK;
}

const char *ohc_last_error(const ohc_ctx *c) { return c ? c->err.c_str() : g_err.c_str(); }

const char *ohc_ctx_kernel_name(const ohc_ctx *c) { return c ? c->kname.c_str() : ""; }

int ohc_chem_solve(ohc_ctx *c, int64_t n, const double *T, const double *p, const double *rho,
                   const double *Y, int64_t ldY, const double *dt, double dtFactor,
                   double *subDt, double *R, int64_t ldR, ohc_stats *stats) {
    return chem_integrate(c, 0, n, T, nullptr, p, rho, Y, nullptr, ldY, dt, dtFactor, subDt, R,
                          ldR, stats, nullptr, 0);
}

int ohc_chem_solve_update(ohc_ctx *c, int64_t n, double *T, const double *p, const double *rho,
                          double *Y, int64_t ldY, const double *dt, double dtFactor,
                          double *subDt, ohc_stats *stats) {
    return chem_integrate(c, 1, n, nullptr, T, p, rho, nullptr, Y, ldY, dt, dtFactor, subDt,
                          nullptr, n, stats, nullptr, 0);
}

int ohc_chem_source(ohc_ctx *c, int64_t n, const double *T, const double *p, const double *rho,
                    const double *Y, int64_t ldY, double *R, int64_t ldR, double *diagJ,
                    int64_t ldD, int32_t withoutLastSpec) {
    int rc = check_ctx(c);
    if (rc) return rc;
    const int ns = c->mech.ns;
    if (n < 0 || ldY < n || ldR < n || (diagJ && ldD < n)) return fail(c, OHC_ERR_ARG, "bad sizes");
    if (!T || !p || !rho || !Y || !R) return fail(c, OHC_ERR_ARG, "null array argument");
    if (withoutLastSpec != 0 && withoutLastSpec != 1)
        return fail(c, OHC_ERR_ARG, "withoutLastSpec must be 0 or 1");
    if (n == 0) return OHC_OK;
    if (!c->lane && !c->wave) return fail(c, OHC_ERR_UNSUPPORTED, "no device kernels for this mechanism");
    if (!c->lane && n > 0x7fffffffll) return fail(c, OHC_ERR_ARG, "n too large");
    int64_t m = 0;
    if (host_slices(c, n, sizeof(double) * (3 + 3 * (size_t)ns), &m))
        return run_slices(c, n, m, [&](int64_t s, int64_t k) {
            return ohc_chem_source(c, k, T + s, p + s, rho + s, Y + s, ldY, R + s, ldR,
                                   off(diagJ, s), ldD, withoutLastSpec);
        });
    Staging st(c);
    const size_t sv = sizeof(double) * (size_t)n;
    void *dT, *dp, *drho, *dY, *dR, *dD = nullptr;
    if ((rc = st.map(T, sv, true, false, &dT))) return rc;
    if ((rc = st.map(p, sv, true, false, &dp))) return rc;
    if ((rc = st.map(rho, sv, true, false, &drho))) return rc;
    if ((rc = st.map_rows(Y, ns, n, ldY, sizeof(double), true, false, &dY))) return rc;
    if ((rc = st.map_rows(R, ns, n, ldR, sizeof(double), false, true, &dR))) return rc;
    if (diagJ && (rc = st.map_rows(diagJ, ns - withoutLastSpec, n, ldD, sizeof(double), false,
                                   true, &dD)))
        return rc;
    ohcd::SourceParams P{n,   (const double *)dT, (const double *)dp, (const double *)drho,
                         (const double *)dY, ldY, (double *)dR, ldR, (double *)dD, ldD,
                         withoutLastSpec, c->opts.jacobian};
    if (c->lane) HIPCHK(c, c->lane->source(P, c->stream));
    else HIPCHK(c, ohcd::wave_source(P, c->wm, c->stream));
    return st.finish();
}

int ohc_chem_source_kim(ohc_ctx *c, int64_t n, const double *T, const double *p,
                        const double *rho, const double *Y, int64_t ldY, double *R, int64_t ldR,
                        void *diag, int64_t ldD, int32_t diag_fp32) {
    int rc = check_ctx(c);
    if (rc) return rc;
    const int ns = c->mech.ns;
    if (n < 0 || ldY < n || ldR < n || (diag && ldD < n)) return fail(c, OHC_ERR_ARG, "bad sizes");
    if (!T || !p || !rho || !Y || !R) return fail(c, OHC_ERR_ARG, "null array argument");
    if (diag_fp32 != 0 && diag_fp32 != 1) return fail(c, OHC_ERR_ARG, "diag_fp32 must be 0 or 1");
    if (ns < 2) return fail(c, OHC_ERR_UNSUPPORTED, "the substituted diagonal needs ns >= 2");
    if (n == 0) return OHC_OK;
    if (!c->d_wm)
        return fail(c, OHC_ERR_UNSUPPORTED,
                    "no run-time device tables for this mechanism (ns <= 63 and at most 3 "
                    "distinct species per reaction side)");
    if (n > 0x7fffffffll) return fail(c, OHC_ERR_ARG, "n too large");
    const size_t esz = diag_fp32 ? sizeof(float) : sizeof(double);
    int64_t m = 0;
    if (host_slices(c, n, sizeof(double) * (3 + 2 * (size_t)ns) + esz * (size_t)ns, &m))
        return run_slices(c, n, m, [&](int64_t s, int64_t k) {
            return ohc_chem_source_kim(c, k, T + s, p + s, rho + s, Y + s, ldY, R + s, ldR,
                                       diag ? (void *)((char *)diag + esz * (size_t)s) : nullptr,
                                       ldD, diag_fp32);
        });
    Staging st(c);
    const size_t sv = sizeof(double) * (size_t)n;
    void *dT, *dp, *drho, *dY, *dR, *dD = nullptr;
    if ((rc = st.map(T, sv, true, false, &dT))) return rc;
    if ((rc = st.map(p, sv, true, false, &dp))) return rc;
    if ((rc = st.map(rho, sv, true, false, &drho))) return rc;
    if ((rc = st.map_rows(Y, ns, n, ldY, sizeof(double), true, false, &dY))) return rc;
    if ((rc = st.map_rows(R, ns, n, ldR, sizeof(double), false, true, &dR))) return rc;
    if (diag && (rc = st.map_rows(diag, ns - 1, n, ldD, esz, false, true, &dD))) return rc;
    ohcd::KimParams P{n,   (const double *)dT, (const double *)drho, (const double *)dY, ldY,
                      (double *)dR, ldR, dD, ldD, c->opts.jacobian};
    HIPCHK(c, ohcd::wave_kim(P, c->wm, diag_fp32, c->stream));
    return st.finish();
}

static int eval_common(ohc_ctx *c, int64_t n, const double *y, int64_t ldy, const double *p,
                       double *f, int64_t ldf, double *J, int64_t ldJ) {
    int rc = check_ctx(c);
    if (rc) return rc;
    const int N = c->mech.ns + 1;
    if (n < 0 || ldy < n || ldf < n || (J && ldJ < n)) return fail(c, OHC_ERR_ARG, "bad sizes");
    if (!y || !p || !f) return fail(c, OHC_ERR_ARG, "null array argument");
    if (n == 0) return OHC_OK;
    if (!c->lane && !c->wave) return fail(c, OHC_ERR_UNSUPPORTED, "no device kernels for this mechanism");
    if (!c->lane && n > 0x7fffffffll) return fail(c, OHC_ERR_ARG, "n too large");
    int64_t m = 0;
    if (host_slices(c, n, sizeof(double) * (1 + 2 * (size_t)N + (J ? (size_t)N * N : 0)), &m))
        return run_slices(c, n, m, [&](int64_t s, int64_t k) {
            return eval_common(c, k, y + s, ldy, p + s, f + s, ldf, off(J, s), ldJ);
        });
    Staging st(c);
    void *dy, *dp, *df, *dJ = nullptr;
    if ((rc = st.map_rows(y, N, n, ldy, sizeof(double), true, false, &dy))) return rc;
    if ((rc = st.map(p, sizeof(double) * (size_t)n, true, false, &dp))) return rc;
    if ((rc = st.map_rows(f, N, n, ldf, sizeof(double), false, true, &df))) return rc;
    if (J && (rc = st.map_rows(J, (int64_t)N * N, n, ldJ, sizeof(double), false, true, &dJ)))
        return rc;
    ohcd::EvalParams P{n,     (const double *)dy, ldy, (const double *)dp, (double *)df,
                       ldf,   (double *)dJ,       ldJ, c->opts.flow_type,  c->opts.jacobian,
                       c->opts.precision};
    if (c->lane) HIPCHK(c, c->lane->eval(P, c->stream));
    else HIPCHK(c, ohcd::wave_eval(P, c->wm, c->opts.jacobian, c->stream));
    return st.finish();
}

int ohc_eval_rhs(ohc_ctx *c, int64_t n, const double *y, int64_t ldy, const double *p,
                 double *f, int64_t ldf) {
    return eval_common(c, n, y, ldy, p, f, ldf, nullptr, n);
}

int ohc_eval_jacobian(ohc_ctx *c, int64_t n, const double *y, int64_t ldy, const double *p,
                      double *f, int64_t ldf, double *J, int64_t ldJ) {
    if (!J) return fail(c, OHC_ERR_ARG, "null J");
    return eval_common(c, n, y, ldy, p, f, ldf, J, ldJ);
}

int ohc_lu_solve(ohc_ctx *c, int64_t nmat, int32_t n, const double *A, int64_t ldA,
                 const double *b, int64_t ldb, double *x, int64_t ldx) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (n < 1 || n > 64) return fail(c, OHC_ERR_ARG, "n must be in 1..64");
    if (nmat < 0 || ldA < nmat || ldb < nmat || ldx < nmat) return fail(c, OHC_ERR_ARG, "bad sizes");
    if (!A || !b || !x) return fail(c, OHC_ERR_ARG, "null array argument");
    if (nmat == 0) return OHC_OK;
    if (nmat > 0x7fffffffll) return fail(c, OHC_ERR_ARG, "nmat too large");
    int64_t m = 0;
    if (host_slices(c, nmat, sizeof(double) * ((size_t)n * n + 2 * (size_t)n), &m))
        return run_slices(c, nmat, m, [&](int64_t s, int64_t k) {
            return ohc_lu_solve(c, k, n, A + s, ldA, b + s, ldb, x + s, ldx);
        });
    Staging st(c);
    void *dA, *db, *dx;
    if ((rc = st.map_rows(A, (int64_t)n * n, nmat, ldA, sizeof(double), true, false, &dA))) return rc;
    if ((rc = st.map_rows(b, n, nmat, ldb, sizeof(double), true, false, &db))) return rc;
    if ((rc = st.map_rows(x, n, nmat, ldx, sizeof(double), false, true, &dx))) return rc;
    ohcd::LuParams P{nmat, n, (const double *)dA, ldA, (const double *)db, ldb, (double *)dx, ldx};
    HIPCHK(c, ohcd::wave_lu(P, c->stream));
    return st.finish();
}

static int t_from_f(ohc_ctx *c, int64_t n, const double *h, const double *Y, int64_t ldY,
                    double *T, int32_t *iflag, int energy) {
    int rc = check_ctx(c);
    if (rc) return rc;
    const int ns = c->mech.ns;
    if (n < 0 || ldY < n) return fail(c, OHC_ERR_ARG, "bad sizes");
    if (!h || !Y || !T) return fail(c, OHC_ERR_ARG, "null array argument");
    if (n == 0) return OHC_OK;
    if (!c->lane && !c->wave) return fail(c, OHC_ERR_UNSUPPORTED, "no device kernels for this mechanism");
    int64_t m = 0;
    if (host_slices(c, n, sizeof(double) * (2 + (size_t)ns) + sizeof(int32_t), &m))
        return run_slices(c, n, m, [&](int64_t s, int64_t k) {
            return t_from_f(c, k, h + s, Y + s, ldY, T + s, off(iflag, s), energy);
        });
    Staging st(c);
    void *dh, *dY, *dT, *dI = nullptr;
    const size_t sv = sizeof(double) * (size_t)n;
    if ((rc = st.map(h, sv, true, false, &dh))) return rc;
    if ((rc = st.map_rows(Y, ns, n, ldY, sizeof(double), true, false, &dY))) return rc;
    if ((rc = st.map(T, sv, true, true, &dT))) return rc;
    if (iflag && (rc = st.map(iflag, sizeof(int32_t) * (size_t)n, false, true, &dI))) return rc;
    ohcd::ThParams P{n, (const double *)dh, (const double *)dY, ldY, (double *)dT, (int *)dI,
                     c->opts.T_low, c->opts.T_high, energy};
    if (c->lane) HIPCHK(c, c->lane->tfromh(P, c->stream));
    else HIPCHK(c, ohcd::wave_tfromh(P, c->wm, c->stream));
    return st.finish();
}

int ohc_thermo_T_from_h(ohc_ctx *c, int64_t n, const double *h, const double *p,
                        const double *Y, int64_t ldY, double *T, int32_t *iflag) {
    (void)p;
    return t_from_f(c, n, h, Y, ldY, T, iflag, 0);
}

int ohc_thermo_T_from_e(ohc_ctx *c, int64_t n, const double *e, const double *p_or_rho,
                        const double *Y, int64_t ldY, double *T, int32_t *iflag) {
    (void)p_or_rho;
    return t_from_f(c, n, e, Y, ldY, T, iflag, 1);
}

static int mix_common(ohc_ctx *c, int64_t n, const double *T, const double *p, const double *Y,
                      int64_t ldY, double *mu, double *kappa, double *Dk, int64_t ldD,
                      double *cp, double *ha, double *hk, double *cv, double *ea) {
    int rc = check_ctx(c);
    if (rc) return rc;
    const int ns = c->mech.ns;
    if (n < 0 || ldY < n || ((Dk || hk) && ldD < n)) return fail(c, OHC_ERR_ARG, "bad sizes");
    if (!T || !p || !Y) return fail(c, OHC_ERR_ARG, "null array argument");
    if ((mu || kappa || Dk) && !c->mech.has_transport)
        return fail(c, OHC_ERR_ARG, "mechanism has no transport data");
    if (n == 0) return OHC_OK;
    if (c->lane && (mu || kappa || Dk) && !c->lane_transport)
        return fail(c, OHC_ERR_UNSUPPORTED, "no device property kernels for this mechanism");
    if (!c->lane && !c->wave)
        return fail(c, OHC_ERR_UNSUPPORTED, "no device property kernels for this mechanism");
    int64_t m = 0;
    if (host_slices(c, n, sizeof(double) * (8 + 3 * (size_t)ns), &m))
        return run_slices(c, n, m, [&](int64_t s, int64_t k) {
            return mix_common(c, k, T + s, p + s, Y + s, ldY, off(mu, s), off(kappa, s),
                              off(Dk, s), ldD, off(cp, s), off(ha, s), off(hk, s), off(cv, s),
                              off(ea, s));
        });
    Staging st(c);
    const size_t sv = sizeof(double) * (size_t)n;
    void *dT, *dp, *dY, *dmu, *dk, *dD, *dcp, *dha, *dhk, *dcv, *dea;
    if ((rc = st.map(T, sv, true, false, &dT))) return rc;
    if ((rc = st.map(p, sv, true, false, &dp))) return rc;
    if ((rc = st.map_rows(Y, ns, n, ldY, sizeof(double), true, false, &dY))) return rc;
    if ((rc = st.map(mu, sv, false, true, &dmu))) return rc;
    if ((rc = st.map(kappa, sv, false, true, &dk))) return rc;
    if ((rc = st.map_rows(Dk, ns, n, ldD, sizeof(double), false, true, &dD))) return rc;
    if ((rc = st.map(cp, sv, false, true, &dcp))) return rc;
    if ((rc = st.map(ha, sv, false, true, &dha))) return rc;
    if ((rc = st.map_rows(hk, ns, n, ldD, sizeof(double), false, true, &dhk))) return rc;
    if ((rc = st.map(cv, sv, false, true, &dcv))) return rc;
    if ((rc = st.map(ea, sv, false, true, &dea))) return rc;
    ohcd::PropParams P{n,          (const double *)dT, (const double *)dp, (const double *)dY,
                       ldY,        (double *)dmu,      (double *)dk,       (double *)dD,
                       ldD,        (double *)dcp,      (double *)dha,      (double *)dhk,
                       c->opts.precision, (double *)dcv, (double *)dea};
    if (c->lane) HIPCHK(c, c->lane->props(P, c->stream));
    else HIPCHK(c, ohcd::wave_props(P, c->wm, c->stream));
    return st.finish();
}

int ohc_mix_thermo(ohc_ctx *c, int64_t n, const double *T, const double *p, const double *Y,
                   int64_t ldY, double *cp, double *cv, double *ha, double *ea) {
    return mix_common(c, n, T, p, Y, ldY, nullptr, nullptr, nullptr, n, cp, ha, nullptr, cv, ea);
}

int ohc_mix_props(ohc_ctx *c, int64_t n, const double *T, const double *p, const double *Y,
                  int64_t ldY, double *mu, double *kappa, double *Dk, int64_t ldD, double *cp,
                  double *ha, double *hk) {
    return mix_common(c, n, T, p, Y, ldY, mu, kappa, Dk, ldD, cp, ha, hk, nullptr, nullptr);
}

}  // extern "C"
