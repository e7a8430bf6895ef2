// codegen.cpp -- emit a compile-time mechanism descriptor for the lane kernels.
//
// The descriptor is a struct of static constexpr tables (RxnD per reaction,
// scaled NASA coefficients, efficiencies, transport constants).  The kernel
// templates in chem_device.hpp / lane_kernels.hip consume it with template
// constant indices, so the compiler keeps every per-cell vector and the
// Jacobian in registers.  All derived constants (R_i, NASA coefficients x R_i,
// the coefficient quotients a_k/5 ... a_k/20) are computed here in double
// exactly as the reference computes them at run time and printed with 17
// significant digits, which round-trips every double.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <sstream>
#include <stdexcept>
#include <string>

#include "codegen.hpp"

namespace ohc {

namespace {
std::string d17(double v) {
    char b[64];
    std::snprintf(b, sizeof(b), "%.17g", v);
    std::string s(b);
    if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
    if (s == "inf") return "__builtin_huge_val()";
    return s;
}
}  // namespace

std::string codegen_mechanism(const Mechanism &m, const std::string &name) {
    const int P = OHC_MAXP, ns = m.ns, nr = m.nr;
    for (int r = 0; r < nr; ++r)
        if (m.kind[r] > OHC_RATE_CHEMACT)
            throw std::runtime_error("reaction " + m.rxn_names[r] +
                                     ": PLOG/CHEB rates have no GPU kernel");
    std::ostringstream o;
    o << "// Generated from the mechanism tables by codegen.cpp -- do not edit.\n";
    o << "#pragma once\n#include \"chem_device.hpp\"\n\n";
    o << "struct Mech_" << name << " {\n";
    o << "  static constexpr unsigned long long HASH = " << m.hash() << "ull;\n";
    o << "  static constexpr unsigned long long THASH = " << (m.has_transport ? m.hash_transport() : 0ull)
      << "ull;\n";
    o << "  static constexpr int NS = " << ns << ", NR = " << nr << ";\n";
    auto arr1 = [&](const char *nm, const std::vector<double> &v) {
        o << "  static constexpr double " << nm << "[" << v.size() << "] = {";
        for (size_t i = 0; i < v.size(); ++i) o << (i ? ", " : "") << d17(v[i]);
        o << "};\n";
    };
    std::vector<double> Ri(ns);
    for (int k = 0; k < ns; ++k) Ri[k] = kRu / m.W[k];
    arr1("W", m.W);
    arr1("Ri", Ri);
    arr1("Tlow", m.Tlow);
    arr1("Thigh", m.Thigh);
    arr1("Tcommon", m.Tcommon);
    std::vector<double> ekb = m.ekb, sig = m.sigma;
    if (!m.has_transport) {
        ekb.assign(ns, 1.0);
        sig.assign(ns, 1.0);
    }
    arr1("EKB", ekb);
    arr1("SIGMA", sig);
    auto coef = [&](const char *nm, const std::vector<double> &raw, int which) {
        // which: 0 = scaled a0..a6, 1 = h quotients, 2 = s quotients, 3 = g quotients
        int w = which == 0 ? 7 : (which == 2 ? 3 : 4);
        o << "  static constexpr double " << nm << "[" << ns << "][" << w << "] = {";
        for (int k = 0; k < ns; ++k) {
            double a[7];
            for (int c = 0; c < 7; ++c) a[c] = raw[7 * k + c] * Ri[k];
            o << (k ? ", " : "") << "{";
            if (which == 0) {
                for (int c = 0; c < 7; ++c) o << (c ? ", " : "") << d17(a[c]);
            } else if (which == 1) {
                o << d17(a[4] / 5.0) << ", " << d17(a[3] / 4.0) << ", " << d17(a[2] / 3.0) << ", "
                  << d17(a[1] / 2.0);
            } else if (which == 2) {
                o << d17(a[4] / 4.0) << ", " << d17(a[3] / 3.0) << ", " << d17(a[2] / 2.0);
            } else {
                o << d17(a[4] / 20.0) << ", " << d17(a[3] / 12.0) << ", "
                  << d17(a[2] / 6.0) << ", " << d17(a[1] / 2.0);
            }
            o << "}";
        }
        o << "};\n";
    };
    coef("alo", m.nasa_lo, 0);
    coef("ahi", m.nasa_hi, 0);
    coef("hlo", m.nasa_lo, 1);
    coef("hhi", m.nasa_hi, 1);
    coef("slo", m.nasa_lo, 2);
    coef("shi", m.nasa_hi, 2);
    coef("glo", m.nasa_lo, 3);
    coef("ghi", m.nasa_hi, 3);
    // Third-body efficiency rows, one per distinct efficiency vector: the
    // reactions of a group share M = sum eff_k c_k and their dln(M)/dc_i
    // Jacobian terms are applied as one outer product per group.
    std::vector<int> tbrow(nr, -1);
    std::vector<std::vector<double>> groups;
    for (int r = 0; r < nr; ++r) {
        if (m.kind[r] == OHC_RATE_ARRHENIUS) continue;
        std::vector<double> e(m.eff.begin() + (size_t)r * ns, m.eff.begin() + (size_t)(r + 1) * ns);
        int g = -1;
        for (size_t q = 0; q < groups.size(); ++q)
            if (groups[q] == e) g = (int)q;
        if (g < 0) {
            g = (int)groups.size();
            groups.push_back(e);
        }
        tbrow[r] = g;
    }
    const int ntb = (int)groups.size();
    o << "  static constexpr int NTB = " << ntb << ";\n";
    o << "  static constexpr double EFF[" << (ntb ? ntb : 1) << "][" << ns << "] = {";
    if (!ntb) {
        o << "{";
        for (int k = 0; k < ns; ++k) o << (k ? ", " : "") << "0.0";
        o << "}";
    }
    for (int g = 0; g < ntb; ++g) {
        o << (g ? ", " : "") << "{";
        for (int k = 0; k < ns; ++k) o << (k ? ", " : "") << d17(groups[g][k]);
        o << "}";
    }
    o << "};\n";
    // rows touched by the plain third-body reactions of each group
    o << "  static constexpr bool TBROW[" << (ntb ? ntb : 1) << "][" << ns << "] = {";
    for (int g = 0; g < std::max(ntb, 1); ++g) {
        o << (g ? ", " : "") << "{";
        for (int k = 0; k < ns; ++k) {
            bool hit = false;
            for (int r = 0; r < nr; ++r) {
                if (tbrow[r] != g || m.kind[r] != OHC_RATE_THIRDBODY) continue;
                for (int i = 0; i < m.nf[r]; ++i) hit |= m.f_idx[r * P + i] == k;
                for (int i = 0; i < m.nb[r]; ++i) hit |= m.b_idx[r * P + i] == k;
            }
            o << (k ? ", " : "") << (hit ? "true" : "false");
        }
        o << "}";
    }
    o << "};\n";
    // Equilibrium constants may be assembled from per-species Boltzmann factors
    // exp(-G_k/(Ru T)) when those stay inside the double range for every
    // temperature the integrator can see (threshold 200 K .. 6000 K).
    bool kc_species = true;
    for (int k = 0; k < ns && kc_species; ++k) {
        for (double T = 200.0; T <= 6000.0; T += 10.0) {
            const double *a = (T < m.Tcommon[k] ? &m.nasa_lo[7 * k] : &m.nasa_hi[7 * k]);
            const double g = a[0] * (1 - std::log(T)) - a[1] * T / 2 - a[2] * T * T / 6 -
                             a[3] * T * T * T / 12 - a[4] * T * T * T * T / 20 + a[5] / T - a[6];
            if (!(std::fabs(g) < 600.0)) {
                kc_species = false;
                break;
            }
        }
    }
    o << "  static constexpr bool KC_SPECIES = " << (kc_species ? "true" : "false") << ";\n";
    o << "  static constexpr ohcd::RxnD R[" << (nr ? nr : 1) << "] = {\n";
    for (int r = 0; r < nr; ++r) {
        double s = 0.0;
        for (int i = 0; i < m.nf[r]; ++i) s -= m.f_nu[r * P + i];
        for (int i = 0; i < m.nb[r]; ++i) s += m.b_nu[r * P + i];
        auto ilist = [&](const std::vector<int32_t> &v) {
            std::string t = "{";
            for (int i = 0; i < P; ++i) t += (i ? ", " : "") + std::to_string(v[r * P + i]);
            return t + "}";
        };
        auto dlist = [&](const std::vector<double> &v, int stride, int n) {
            std::string t = "{";
            for (int i = 0; i < n; ++i) t += (i ? ", " : "") + d17(v[(size_t)r * stride + i]);
            return t + "}";
        };
        o << "    {" << m.nf[r] << ", " << m.nb[r] << ", " << ilist(m.f_idx) << ", "
          << ilist(m.b_idx) << ", " << dlist(m.f_nu, P, P) << ", " << dlist(m.b_nu, P, P) << ", "
          << dlist(m.f_ord, P, P) << ", " << dlist(m.b_ord, P, P) << ", " << m.rev[r] << ", "
          << m.kind[r] << ", " << m.fo_type[r] << ", " << d17(m.arr[3 * r]) << ", "
          << d17(m.arr[3 * r + 1]) << ", " << d17(m.arr[3 * r + 2]) << ", "
          << d17(m.arr_lo[3 * r]) << ", " << d17(m.arr_lo[3 * r + 1]) << ", "
          << d17(m.arr_lo[3 * r + 2]) << ", " << d17(m.arr_rev[3 * r]) << ", "
          << d17(m.arr_rev[3 * r + 1]) << ", " << d17(m.arr_rev[3 * r + 2]) << ", "
          << dlist(m.fo, 5, 5) << ", " << d17(s) << ", " << tbrow[r] << "},  // "
          << m.rxn_names[r] << "\n";
    }
    if (!nr) o << "    {}\n";
    o << "  };\n};\n";
    return o.str();
}

}  // namespace ohc
