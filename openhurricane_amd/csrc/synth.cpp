// synth.cpp -- seeded synthetic element-balanced C/H/O/N mechanisms.
//
// The generator writes ordinary CHEMKIN-II reaction, NASA-7 thermo and
// transport texts and loads them through the same parser as real files, so a
// synthetic mechanism is exactly what a user would get from those files (and
// the texts can be written out and read by any CHEMKIN-compatible code).
//
// Species come from a fixed list of real C/H/O/N formulas (first ns entries).
// Thermo: cp/R(T) rises smoothly from a low-T value set by the molecule's
// shape (2.5 atoms, 3.5 diatomics, 4 larger) towards its classical high-T
// limit; each NASA-7 range is a least-squares quartic fit of that curve, h and
// s are anchored at 298.15 K to a random heat of formation and entropy and made
// continuous at 1000 K.  Reactions: a random reactant set (A + B, or A, or
// A + B + M) and an element-balanced product set found among the species
// (single species or pairs); duplicates are skipped; each reaction is
// written in the direction in which its Gibbs energy at 1500 K decreases, so
// irreversible steps relax a mixture towards equilibrium.  Rate constants:
// A log-uniform in [1e6, 1e14] (mol-cm-s) taken as the pre-exponential
// factor at 1000 K (k = A (T/1000)^beta exp(-Ea/RT)), beta ~
// U(-1.5, 1.5), Ea ~ U(0, 45) kcal/mol.  frac_rev of the reactions are
// reversible (kr from Kc); frac_tb carry a third body, one third of those as
// Lindemann / Troe / SRI fall-off.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "mech.hpp"

namespace ohc {
namespace {

struct Rng {  // SplitMix64: a platform-independent stream
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double u() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
    double u(double a, double b) { return a + (b - a) * u(); }
    int i(int n) { return (int)(u() * n); }
};

struct Formula {
    const char *name;
    int c, h, o, n;
};

// Real C/H/O/N formulas; the first entries are the ones a fuel/air case needs.
const Formula kSpecies[] = {
    {"H2", 0, 2, 0, 0},     {"O2", 0, 0, 2, 0},     {"N2", 0, 0, 0, 2},
    {"H2O", 0, 2, 1, 0},    {"H", 0, 1, 0, 0},      {"O", 0, 0, 1, 0},
    {"OH", 0, 1, 1, 0},     {"HO2", 0, 1, 2, 0},    {"H2O2", 0, 2, 2, 0},
    {"CH4", 1, 4, 0, 0},    {"CO", 1, 0, 1, 0},     {"CO2", 1, 0, 2, 0},
    {"CH3", 1, 3, 0, 0},    {"CH2", 1, 2, 0, 0},    {"CH", 1, 1, 0, 0},
    {"HCO", 1, 1, 1, 0},    {"CH2O", 1, 2, 1, 0},   {"CH3O", 1, 3, 1, 0},
    {"CH3OH", 1, 4, 1, 0},  {"C2H2", 2, 2, 0, 0},   {"C2H3", 2, 3, 0, 0},
    {"C2H4", 2, 4, 0, 0},   {"C2H5", 2, 5, 0, 0},   {"C2H6", 2, 6, 0, 0},
    {"HCCO", 2, 1, 1, 0},   {"CH2CO", 2, 2, 1, 0},  {"N", 0, 0, 0, 1},
    {"NO", 0, 0, 1, 1},     {"NO2", 0, 0, 2, 1},    {"N2O", 0, 0, 1, 2},
    {"NH", 0, 1, 0, 1},     {"NH2", 0, 2, 0, 1},    {"NH3", 0, 3, 0, 1},
    {"NNH", 0, 1, 0, 2},    {"HNO", 0, 1, 1, 1},    {"HCN", 1, 1, 0, 1},
    {"CN", 1, 0, 0, 1},     {"NCO", 1, 0, 1, 1},    {"HNCO", 1, 1, 1, 1},
    {"C", 1, 0, 0, 0},      {"C2H", 2, 1, 0, 0},    {"CH2OH", 1, 3, 1, 0},
    {"CH3CHO", 2, 4, 1, 0}, {"CH2CHO", 2, 3, 1, 0}, {"C3H8", 3, 8, 0, 0},
    {"C3H7", 3, 7, 0, 0},   {"C3H6", 3, 6, 0, 0},   {"C3H5", 3, 5, 0, 0},
    {"C3H4", 3, 4, 0, 0},   {"C3H3", 3, 3, 0, 0},   {"HOCO", 1, 1, 2, 0},
    {"N2H2", 0, 2, 0, 2},   {"N2H3", 0, 3, 0, 2},   {"N2H4", 0, 4, 0, 2},
    {"HONO", 0, 1, 2, 1},   {"CH3O2", 1, 3, 2, 0},  {"C2H5OH", 2, 6, 1, 0},
    {"H2CN", 1, 2, 0, 1},   {"HCNO", 1, 1, 1, 1},   {"HOCN", 1, 1, 1, 1},
    {"HCNN", 1, 1, 0, 2},   {"CH3CO", 2, 3, 1, 0},  {"C2H4O", 2, 4, 1, 0},
    {"CH2CHOH", 2, 4, 1, 0},
};
constexpr int kNumSpecies = sizeof(kSpecies) / sizeof(kSpecies[0]);

using Comp = std::array<int, 4>;
Comp comp_of(const Formula &f) { return {f.c, f.h, f.o, f.n}; }
Comp add(const Comp &a, const Comp &b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2], a[3] + b[3]}; }
int atoms(const Formula &f) { return f.c + f.h + f.o + f.n; }

// Least-squares quartic cp/R on [a, b] (in t = T/1000), 64 samples.
void fit_quartic(double (*cp)(double, const double *), const double *par, double a, double b,
                 double out[5]) {
    double M[5][6] = {};
    for (int s = 0; s < 64; ++s) {
        const double T = a + (b - a) * s / 63.0, t = T / 1000.0;
        double p[5] = {1, t, t * t, t * t * t, t * t * t * t};
        const double y = cp(T, par);
        for (int i = 0; i < 5; ++i) {
            for (int j = 0; j < 5; ++j) M[i][j] += p[i] * p[j];
            M[i][5] += p[i] * y;
        }
    }
    for (int k = 0; k < 5; ++k) {  // Gauss-Jordan with partial pivoting
        int piv = k;
        for (int i = k + 1; i < 5; ++i)
            if (std::fabs(M[i][k]) > std::fabs(M[piv][k])) piv = i;
        for (int j = 0; j < 6; ++j) std::swap(M[k][j], M[piv][j]);
        for (int i = 0; i < 5; ++i) {
            if (i == k) continue;
            const double f = M[i][k] / M[k][k];
            for (int j = k; j < 6; ++j) M[i][j] -= f * M[k][j];
        }
    }
    for (int i = 0; i < 5; ++i) out[i] = M[i][5] / M[i][i] / std::pow(1000.0, i);
}

double cp_shape(double T, const double *p) {  // p = {cp0, cpinf, theta}
    return p[0] + (p[1] - p[0]) * (1.0 - std::exp(-(T - 200.0) / p[2]));
}

double h_over_R(const double a[7], double T) {
    return a[0] * T + a[1] * T * T / 2 + a[2] * T * T * T / 3 + a[3] * std::pow(T, 4) / 4 +
           a[4] * std::pow(T, 5) / 5 + a[5];
}
double s_over_R(const double a[7], double T) {
    return a[0] * std::log(T) + a[1] * T + a[2] * T * T / 2 + a[3] * T * T * T / 3 +
           a[4] * std::pow(T, 4) / 4 + a[6];
}

std::string fmt(const char *f, double v) {
    char b[64];
    std::snprintf(b, sizeof b, f, v);
    return b;
}

}  // namespace

void synthetic_chemkin(int ns, int nr, uint64_t seed, double frac_rev, double frac_tb,
                       std::string &chem, std::string &thermo, std::string &transport) {
    if (ns < 4 || ns > std::min(OHC_MAXNS, kNumSpecies))
        throw std::runtime_error("synthetic mechanism: ns must be in 4.." +
                                 std::to_string(std::min(OHC_MAXNS, kNumSpecies)));
    if (nr < 1 || nr > 100000) throw std::runtime_error("synthetic mechanism: bad nr");
    if (!(frac_rev >= 0 && frac_rev <= 1 && frac_tb >= 0 && frac_tb <= 1))
        throw std::runtime_error("synthetic mechanism: fractions must be in [0, 1]");
    Rng rng{seed * 0x2545F4914F6CDD1Dull + 0x1234567ull};

    // ---- species, thermo, transport
    thermo = "THERMO ALL\n   300.000  1000.000  5000.000\n";
    transport.clear();
    std::string species_line;
    std::vector<Comp> comp(ns);
    std::vector<double> g1500(ns);  // G/(R T) at 1500 K, orients irreversible reactions
    for (int k = 0; k < ns; ++k) {
        const Formula &f = kSpecies[k];
        comp[k] = comp_of(f);
        const int na = atoms(f);
        double par[3];
        par[0] = na == 1 ? 2.5 : (na == 2 ? 3.5 : 4.0);
        par[1] = na == 1 ? 2.5 : (na == 2 ? 4.5 : 3.0 * na - 2.0);
        par[2] = rng.u(600.0, 1500.0);
        const bool element_ref = (!std::string(f.name).compare("H2") ||
                                  !std::string(f.name).compare("O2") ||
                                  !std::string(f.name).compare("N2"));
        const double dHf = element_ref ? 0.0 : rng.u(-120e3, 200e3);            // J/mol
        const double S298 = 150.0 + 20.0 * na + rng.u(-15.0, 15.0);             // J/mol/K
        double lo[7] = {}, hi[7] = {};
        fit_quartic(cp_shape, par, 300.0, 1000.0, lo);
        fit_quartic(cp_shape, par, 1000.0, 5000.0, hi);
        // cp continuous at 1000 K: shift the high-range constant
        auto cpv = [](const double *a, double T) {
            return a[0] + a[1] * T + a[2] * T * T + a[3] * T * T * T + a[4] * T * T * T * T;
        };
        hi[0] += cpv(lo, 1000.0) - cpv(hi, 1000.0);
        const double R = 8.31446261815324;
        lo[5] = dHf / R - h_over_R(lo, 298.15);
        lo[6] = S298 / R - s_over_R(lo, 298.15);
        hi[5] = h_over_R(lo, 1000.0) - h_over_R(hi, 1000.0);
        hi[6] = s_over_R(lo, 1000.0) - s_over_R(hi, 1000.0);
        g1500[k] = h_over_R(hi, 1500.0) / 1500.0 - s_over_R(hi, 1500.0);
        // record line 1: name (cols 1-18), composition (25-44), phase, T range, '1'
        char l1[128];
        const char sym[4] = {'C', 'H', 'O', 'N'};
        const int cnt[4] = {f.c, f.h, f.o, f.n};
        std::string el;
        for (int e = 0; e < 4; ++e) {
            char b[16];
            if (cnt[e] > 0) std::snprintf(b, sizeof b, "%-2c%3d", sym[e], cnt[e] % 1000);
            else std::snprintf(b, sizeof b, "     ");
            el += b;
        }
        std::snprintf(l1, sizeof l1, "%-18s      %-20sG%10.3f%10.3f%8.2f      1\n", f.name,
                      el.c_str(), 300.0, 5000.0, 1000.0);
        thermo += l1;
        auto line = [&](const double *v, int nv, int no) {
            std::string s;
            for (int i = 0; i < nv; ++i) s += fmt("%15.8E", v[i]);
            for (int i = nv; i < 5; ++i) s += std::string(15, ' ');
            s += "    " + std::to_string(no) + "\n";
            thermo += s;
        };
        double c15[15];
        for (int i = 0; i < 7; ++i) c15[i] = hi[i], c15[7 + i] = lo[i];
        line(c15, 5, 2);
        line(c15 + 5, 5, 3);
        line(c15 + 10, 4, 4);
        // transport: geometry, eps/k, sigma, dipole, polarizability, Zrot
        const int geom = na == 1 ? 0 : (na == 2 ? 1 : 2);
        char tl[160];
        std::snprintf(tl, sizeof tl, "%-16s %d %9.3f %9.3f %9.3f %9.3f %9.3f\n", f.name, geom,
                      80.0 + 40.0 * na + rng.u(-30.0, 30.0),
                      2.5 + 0.25 * std::pow((double)na, 0.8) + rng.u(-0.2, 0.2), 0.0, 0.0,
                      na == 1 ? 0.0 : 1.0);
        transport += tl;
        species_line += std::string(f.name) + " ";
    }
    thermo += "END\n";

    // ---- element-balanced reactions
    std::map<Comp, std::vector<int>> singles;
    std::map<Comp, std::vector<std::pair<int, int>>> pairs;
    for (int a = 0; a < ns; ++a) {
        singles[comp[a]].push_back(a);
        for (int b = a; b < ns; ++b) pairs[add(comp[a], comp[b])].push_back({a, b});
    }
    std::set<std::pair<std::vector<int>, std::vector<int>>> seen;
    chem = "ELEMENTS\nC H O N\nEND\nSPECIES\n" + species_line + "\nEND\nREACTIONS\n";
    auto side = [&](const std::vector<int> &v) {
        std::string s;
        for (size_t i = 0; i < v.size(); ++i) s += (i ? "+" : "") + std::string(kSpecies[v[i]].name);
        return s;
    };
    auto eff_line = [&]() {
        std::string s;
        const char *enh[] = {"H2", "H2O", "CO2", "CO", "CH4"};
        const double lo[] = {2.0, 6.0, 2.0, 1.5, 2.0}, hi[] = {2.5, 12.0, 3.8, 1.9, 3.0};
        for (int e = 0; e < 5; ++e)
            for (int k = 0; k < ns; ++k)
                if (!std::string(kSpecies[k].name).compare(enh[e]))
                    s += std::string(enh[e]) + "/" + fmt("%.3f", rng.u(lo[e], hi[e])) + "/ ";
        return s.empty() ? s : s + "\n";
    };
    int made = 0, tries = 0;
    const int n_tb = (int)std::lround(frac_tb * nr);
    const int n_rev = (int)std::lround(frac_rev * nr);
    while (made < nr) {
        if (++tries > 1000000) throw std::runtime_error("synthetic mechanism: too few reactions");
        const bool tb = made < n_tb;
        const bool fall = tb && (made % 3 == 0);
        const bool rev = (made * 7919) % nr < n_rev;  // spread reversible ones over kinds
        std::vector<int> R;
        const double form = rng.u();
        if (!tb && form < 0.15) R = {rng.i(ns)};
        else R = {rng.i(ns), rng.i(ns)};
        std::sort(R.begin(), R.end());
        Comp S = R.size() == 1 ? comp[R[0]] : add(comp[R[0]], comp[R[1]]);
        std::vector<std::vector<int>> options;
        if (singles.count(S) && R.size() == 2)
            for (int a : singles[S]) options.push_back({a});
        if (pairs.count(S))
            for (auto &pr : pairs[S]) options.push_back({pr.first, pr.second});
        if (fall) {  // prefer associations A + B (+M) = AB (+M)
            std::vector<std::vector<int>> assoc;
            for (auto &o : options)
                if (o.size() == 1) assoc.push_back(o);
            if (!assoc.empty()) options = assoc;
        }
        std::vector<std::vector<int>> ok;
        for (auto &o : options) {
            std::vector<int> P = o;
            std::sort(P.begin(), P.end());
            if (P == R) continue;
            if (seen.count({R, P}) || seen.count({P, R})) continue;
            ok.push_back(P);
        }
        if (ok.empty()) continue;
        std::vector<int> P = ok[rng.i((int)ok.size())];
        seen.insert({R, P});
        // write the reaction in its downhill direction (Delta G(1500 K) < 0), so
        // that irreversible steps relax a mixture towards equilibrium
        auto gsum = [&](const std::vector<int> &v) {
            double g = 0.0;
            for (int k : v) g += g1500[k];
            return g;
        };
        if (gsum(P) > gsum(R)) std::swap(R, P);
        // rate constant: A is k(1000 K) in mol-cm-s units -> CHEMKIN A = k / 1000^beta
        const double k1000 = std::pow(10.0, rng.u(6.0, 14.0));
        const double beta = rng.u(-1.5, 1.5);
        const double Ea = rng.u(0.0, 45000.0);  // cal/mol
        const double A = k1000 / std::pow(1000.0, beta);
        const char *arrow = rev ? "=" : "=>";
        std::string eq;
        if (!tb) eq = side(R) + arrow + side(P);
        else if (!fall) eq = side(R) + "+M" + arrow + side(P) + "+M";
        else eq = side(R) + "(+M)" + arrow + side(P) + "(+M)";
        chem += eq + "  " + fmt("%.6E", A) + " " + fmt("%.4f", beta) + " " + fmt("%.2f", Ea) + "\n";
        if (fall) {
            // k0 [cm^6/mol^2/s] = 1e4..1e6 cm^3/mol times k_inf: Pr ~ 0.1..10 near 1 atm
            const double A0 = A * std::pow(10.0, rng.u(4.0, 6.0));
            chem += "LOW/ " + fmt("%.6E", A0) + " " + fmt("%.4f", beta - rng.u(0.5, 1.5)) + " " +
                    fmt("%.2f", Ea * rng.u(0.3, 0.9)) + " /\n";
            const int f = (made / 3) % 3;
            if (f == 1)
                chem += "TROE/ " + fmt("%.4f", rng.u(0.3, 0.8)) + " " + fmt("%.1f", rng.u(50, 300)) +
                        " " + fmt("%.1f", rng.u(1000, 3000)) + " " + fmt("%.1f", rng.u(3000, 8000)) +
                        " /\n";
            else if (f == 2)
                chem += "SRI/ " + fmt("%.4f", rng.u(0.5, 1.5)) + " " + fmt("%.1f", rng.u(100, 500)) +
                        " " + fmt("%.1f", rng.u(800, 2000)) + " /\n";
        }
        if (tb) chem += eff_line();
        ++made;
    }
    chem += "END\n";
}

void synthetic_mechanism(int ns, int nr, uint64_t seed, double frac_rev, double frac_tb,
                         Mechanism &m) {
    std::string chem, thermo, transport;
    synthetic_chemkin(ns, nr, seed, frac_rev, frac_tb, chem, thermo, transport);
    parse_chemkin(chem, &thermo, &transport, m);
}

}  // namespace ohc
