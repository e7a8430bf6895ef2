// mech.cpp -- CHEMKIN-II / NASA-7 / transport loaders and table packing.
//
// Unit rules follow chemkinFileRead::changeUnit and addReactionToList
// (src/speciesTransport/chemistryRead/src/chemkinFileRead.cpp:1349-1560,
// :1654-1699): A is scaled by 1e-3^(sum of forward orders - 1) for
// mole-cm-s input (MOLECULES adds N_A powers), third-body and fall-off k0
// gain one more 1e-3, chemically-activated k_inf one less; activation
// energies become Ta = E / (R in the E unit).  Participants of each side are
// merged by name and kept in name order, with order = stoichiometric
// coefficient (chemkinFileRead.cpp:148-175, :902-927).  Third-body
// efficiencies default to 1 (chemkinFileRead.cpp:1050-1060).
#include "mech.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <map>
#include <sstream>
#include <stdexcept>

namespace ohc {

namespace {

std::string upper(std::string s) {
    for (auto &c : s) c = (char)std::toupper((unsigned char)c);
    return s;
}
std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}
std::string strip_comment(const std::string &line) {
    size_t p = line.find('!');
    return p == std::string::npos ? line : line.substr(0, p);
}
std::vector<std::string> split_ws(const std::string &s) {
    std::vector<std::string> out;
    std::istringstream is(s);
    std::string t;
    while (is >> t) out.push_back(t);
    return out;
}
std::vector<std::string> lines_of(const std::string &text) {
    std::vector<std::string> out;
    std::string cur;
    for (char c : text) {
        if (c == '\n') {
            out.push_back(cur);
            cur.clear();
        } else if (c != '\r') {
            cur.push_back(c == '\t' ? ' ' : c);
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}
double to_double(const std::string &s, const char *what) {
    std::string t = trim(s);
    for (auto &c : t)
        if (c == 'd' || c == 'D') c = 'e';
    char *end = nullptr;
    double v = std::strtod(t.c_str(), &end);
    if (t.empty() || end == t.c_str() || *end != '\0')
        throw std::runtime_error(std::string("cannot read number '") + s + "' (" + what + ")");
    return v;
}

struct Section {
    std::vector<std::string> lines;
    std::string header;  // the keyword line remainder (REACTIONS units)
    bool found = false;
};

// Extract the lines between a section keyword (any of the abbreviations) and
// the next END line.
Section section(const std::vector<std::string> &lines, const std::vector<const char *> &kws) {
    Section s;
    for (size_t i = 0; i < lines.size(); ++i) {
        std::string u = upper(trim(strip_comment(lines[i])));
        if (u.empty()) continue;
        auto tok = split_ws(u);
        bool hit = false;
        for (auto kw : kws)
            if (tok[0] == kw) hit = true;
        if (!hit) continue;
        s.found = true;
        s.header = trim(u.substr(tok[0].size()));
        // Items may follow the keyword on the same line ("ELEMENTS O N H END").
        std::string rest = trim(strip_comment(lines[i])).substr(tok[0].size());
        size_t j = i + 1;
        std::vector<std::string> body;
        // "THERMO ALL" is an option of the keyword, not an item
        if (!trim(rest).empty() && upper(trim(rest)) != "ALL") body.push_back(rest);
        for (; j < lines.size(); ++j) body.push_back(lines[j]);
        for (auto &b : body) {
            std::string bu = upper(trim(strip_comment(b)));
            auto bt = split_ws(bu);
            auto endpos = std::find(bt.begin(), bt.end(), "END");
            if (endpos != bt.end()) {
                // keep tokens before END on this line
                std::string keep;
                for (auto it = bt.begin(); it != endpos; ++it) keep += *it + " ";
                if (!trim(keep).empty()) s.lines.push_back(keep);
                return s;
            }
            s.lines.push_back(b);
        }
        return s;
    }
    return s;
}

struct NasaEntry {
    double Tlow, Thigh, Tcommon;
    double hi[7], lo[7];
    std::map<std::string, double> atoms;
};

// NASA-7 four-line records: name in cols 1-18, element/count pairs in
// cols 25-44 (+ 74-78), Tlow/Thigh/Tcommon in cols 46-73, then 15-column
// coefficients a1..a7 (high range) and a1..a7 (low range).
std::map<std::string, NasaEntry> parse_nasa(const std::vector<std::string> &lines) {
    std::map<std::string, NasaEntry> out;
    double gT[3] = {300.0, 1000.0, 5000.0};
    size_t i = 0;
    // optional global temperature line
    while (i < lines.size() && trim(strip_comment(lines[i])).empty()) ++i;
    if (i < lines.size()) {
        auto t = split_ws(strip_comment(lines[i]));
        bool nums = t.size() >= 3;
        for (size_t k = 0; k < 3 && k < t.size() && nums; ++k) {
            char *e = nullptr;
            std::strtod(t[k].c_str(), &e);
            if (*e != '\0') nums = false;
        }
        if (nums) {
            for (int k = 0; k < 3; ++k) gT[k] = to_double(t[k], "global T");
            // file order is Tlow Tcommon Thigh
            ++i;
        }
    }
    auto field = [](const std::string &l, size_t a, size_t n) {
        return a < l.size() ? l.substr(a, std::min(n, l.size() - a)) : std::string();
    };
    while (i < lines.size()) {
        std::string l1 = lines[i];
        if (trim(l1).empty() || trim(l1)[0] == '!') {
            ++i;
            continue;
        }
        if (i + 3 >= lines.size()) break;
        std::string name = upper(trim(field(l1, 0, 18)));
        {
            auto nt = split_ws(name);
            name = nt.empty() ? "" : nt[0];
        }
        NasaEntry e;
        e.Tlow = gT[0];
        e.Tcommon = gT[1];
        e.Thigh = gT[2];
        auto parse_atoms = [&](const std::string &s) {
            for (size_t k = 0; k < s.size(); k += 5) {
                std::string sym = trim(field(s, k, 2));
                std::string cnt = trim(field(s, k + 2, 3));
                if (sym.empty() || cnt.empty()) continue;
                double c = std::strtod(cnt.c_str(), nullptr);
                if (c != 0.0) e.atoms[upper(sym)] += c;
            }
        };
        parse_atoms(field(l1, 24, 20));
        std::string t1 = trim(field(l1, 45, 10)), t2 = trim(field(l1, 55, 10)),
                    t3 = trim(field(l1, 65, 8));
        if (!t1.empty()) e.Tlow = to_double(t1, "Tlow");
        if (!t2.empty()) e.Thigh = to_double(t2, "Thigh");
        if (!t3.empty()) e.Tcommon = to_double(t3, "Tcommon");
        parse_atoms(field(l1, 73, 5));
        double c[15];
        int nc = 0;
        for (int r = 1; r <= 3; ++r) {
            const std::string &l = lines[i + r];
            int per = r < 3 ? 5 : 4;
            for (int k = 0; k < per; ++k) c[nc++] = to_double(field(l, 15 * k, 15), "NASA coeff");
        }
        for (int k = 0; k < 7; ++k) e.hi[k] = c[k];
        for (int k = 0; k < 7; ++k) e.lo[k] = c[7 + k];
        out[name] = e;
        i += 4;
    }
    return out;
}

struct SideTerm {
    std::string name;
    double nu;
};

// "2OH + H2 + M" -> merged map; returns third-body marker status.
std::map<std::string, double> parse_side(const std::string &side, bool &has_m, bool &has_pm,
                                         std::string &pm_species) {
    std::map<std::string, double> out;
    std::string s = side;
    // fall-off "(+M)" or "(+SPECIES)"
    size_t lp = s.find("(+");
    if (lp != std::string::npos) {
        size_t rp = s.find(')', lp);
        if (rp == std::string::npos) throw std::runtime_error("unbalanced (+M) in '" + side + "'");
        pm_species = upper(trim(s.substr(lp + 2, rp - lp - 2)));
        has_pm = true;
        s = s.substr(0, lp) + s.substr(rp + 1);
    }
    // Split on '+'; a '+' that ends a name and is followed by another '+' or
    // the end of the side is an ion charge and stays in the name.
    std::vector<std::string> terms;
    std::string cur;
    for (size_t i = 0; i < s.size(); ++i) {
        char ch = s[i];
        bool ion = ch == '+' && !trim(cur).empty() && (i + 1 == s.size() || s[i + 1] == '+');
        if (ch == '+' && !ion) {
            terms.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(ch);
        }
    }
    terms.push_back(cur);
    for (auto &t : terms) {
        std::string u = upper(trim(t));
        if (u.empty()) continue;
        if (u == "M") {
            has_m = true;
            continue;
        }
        size_t k = 0;
        while (k < u.size() && (std::isdigit((unsigned char)u[k]) || u[k] == '.')) ++k;
        double nu = 1.0;
        std::string name = u;
        if (k > 0 && k < u.size()) {
            nu = to_double(u.substr(0, k), "stoichiometric coefficient");
            name = trim(u.substr(k));
        }
        out[name] += nu;
    }
    return out;
}

int unit_code_E(const std::string &u) {
    if (u == "CAL/MOLE") return 0;
    if (u == "KCAL/MOLE") return 1;
    if (u == "JOULES/MOLE") return 2;
    if (u == "KJOULES/MOLE" || u == "KJOULE/MOLE") return 3;
    if (u == "KELVINS") return 4;
    if (u == "EVOLTS") return 5;
    return -1;
}

// chemkinFileRead::changeUnit
void change_unit(double order, bool molecules, int uE, bool third_body, double &Af,
                 double &Ef) {
    const double RRjoule = kR, RRcal = kR / kCal;
    const double conc = 0.001;
    if (molecules) {
        Af = third_body ? std::pow(kNA, order) : std::pow(kNA, order - 1);
        Af *= std::pow(conc, order - 1);
    } else {
        Af = std::pow(conc, order - 1);
    }
    switch (uE) {
    case 0: Ef = RRcal; break;
    case 1: Ef = RRcal / 1000.0; break;
    case 2: Ef = RRjoule; break;
    case 3: Ef = RRjoule / 1000.0; break;
    case 4: Ef = 1.0; break;
    case 5: Ef = RRjoule / kEv; break;
    default: Ef = RRcal;
    }
}

// One group of a reaction's auxiliary line: a name followed by an optional
// "/values/" field ("LOW / 1e14 0 0 /", "H2O/12/", or a bare "DUP").
struct AuxGroup {
    std::string name, vals;
    bool has_vals = false;
};
std::vector<AuxGroup> aux_groups(const std::string &u) {
    std::vector<AuxGroup> out;
    const size_t n = u.size();
    size_t i = 0;
    auto ws = [&](size_t j) { return std::isspace((unsigned char)u[j]) != 0; };
    while (i < n) {
        while (i < n && ws(i)) ++i;
        if (i >= n) break;
        AuxGroup g;
        const size_t a = i;
        while (i < n && !ws(i) && u[i] != '/') ++i;
        g.name = u.substr(a, i - a);
        size_t j = i;
        while (j < n && ws(j)) ++j;
        if (j < n && u[j] == '/') {
            const size_t b = u.find('/', j + 1);
            if (b == std::string::npos) throw std::runtime_error("unterminated /.../ in '" + u + "'");
            g.vals = u.substr(j + 1, b - j - 1);
            g.has_vals = true;
            i = b + 1;
        }
        if (g.name.empty()) throw std::runtime_error("/.../ values without a name in '" + u + "'");
        out.push_back(g);
    }
    return out;
}

}  // namespace

double atomic_weight(const std::string &e) {
    // CHEMKIN-II atomic weight table (g/mol), the values OpenHurricane tabulates.
    static const std::map<std::string, double> tab = {
        {"E", 5.45e-4}, {"H", 1.00797}, {"D", 2.01410}, {"T", 3.01604}, {"HE", 4.00260},
        {"C", 12.01115}, {"N", 14.00670}, {"O", 15.99940}, {"F", 18.99840}, {"NE", 20.18300},
        {"AR", 39.94800}, {"S", 32.06400}, {"CL", 35.45300}, {"SI", 28.08600}, {"KR", 83.80000}};
    auto it = tab.find(upper(e));
    if (it == tab.end()) throw std::runtime_error("unknown element '" + e + "'");
    return it->second;
}

void Mechanism::resize_species(int n) {
    ns = n;
    species.resize(n);
    W.assign(n, 0.0);
    Tlow.assign(n, 300.0);
    Thigh.assign(n, 5000.0);
    Tcommon.assign(n, 1000.0);
    nasa_lo.assign(7 * n, 0.0);
    nasa_hi.assign(7 * n, 0.0);
    elem.assign((size_t)n * std::max(ne, 1), 0.0);
    ekb.assign(n, 0.0);
    sigma.assign(n, 0.0);
    geom.assign(n, 0.0);
    dipole.assign(n, 0.0);
    polar.assign(n, 0.0);
    zrot.assign(n, 0.0);
}

int Mechanism::add_reaction() {
    const int P = OHC_MAXP;
    int r = nr++;
    nf.push_back(0);
    nb.push_back(0);
    f_idx.resize((size_t)nr * P, 0);
    b_idx.resize((size_t)nr * P, 0);
    f_nu.resize((size_t)nr * P, 0.0);
    b_nu.resize((size_t)nr * P, 0.0);
    f_ord.resize((size_t)nr * P, 0.0);
    b_ord.resize((size_t)nr * P, 0.0);
    rev.push_back(OHC_RXN_IRREVERSIBLE);
    kind.push_back(OHC_RATE_ARRHENIUS);
    fo_type.push_back(OHC_FO_LINDEMANN);
    arr.resize((size_t)nr * 3, 0.0);
    arr_lo.resize((size_t)nr * 3, 0.0);
    arr_rev.resize((size_t)nr * 3, 0.0);
    fo.resize((size_t)nr * 5, 0.0);
    eff.resize((size_t)nr * ns, 0.0);
    aux_off.push_back((int32_t)aux.size());
    rxn_names.emplace_back();
    return r;
}

void Mechanism::set_aux(int r, const std::vector<double> &v) {
    if (r != nr - 1 || aux_off[r + 1] != aux_off[r])
        throw std::runtime_error("rate data must be attached to the last reaction once");
    aux.insert(aux.end(), v.begin(), v.end());
    aux_off[r + 1] = (int32_t)aux.size();
}

int Mechanism::species_index(const std::string &name) const {
    for (int k = 0; k < ns; ++k)
        if (species[k] == name) return k;
    return -1;
}

ohc_mech_tables Mechanism::tables() const {
    ohc_mech_tables t;
    std::memset(&t, 0, sizeof(t));
    t.ns = ns;
    t.nr = nr;
    t.ne = ne;
    t.W = W.data();
    t.Tlow = Tlow.data();
    t.Thigh = Thigh.data();
    t.Tcommon = Tcommon.data();
    t.nasa_lo = nasa_lo.data();
    t.nasa_hi = nasa_hi.data();
    t.elem = elem.data();
    t.ekb = ekb.data();
    t.sigma = sigma.data();
    t.nf = nf.data();
    t.nb = nb.data();
    t.f_idx = f_idx.data();
    t.b_idx = b_idx.data();
    t.f_nu = f_nu.data();
    t.b_nu = b_nu.data();
    t.f_ord = f_ord.data();
    t.b_ord = b_ord.data();
    t.rev = rev.data();
    t.kind = kind.data();
    t.arr = arr.data();
    t.arr_lo = arr_lo.data();
    t.arr_rev = arr_rev.data();
    t.fo_type = fo_type.data();
    t.fo = fo.data();
    t.eff = eff.data();
    t.aux_off = aux_off.data();
    t.aux = aux.data();
    return t;
}

uint64_t Mechanism::hash() const {
    // FNV-1a over the numeric tables that the kernels consume.
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void *p, size_t n) {
        const unsigned char *c = (const unsigned char *)p;
        for (size_t i = 0; i < n; ++i) {
            h ^= c[i];
            h *= 1099511628211ull;
        }
    };
    auto vd = [&](const std::vector<double> &v) { mix(v.data(), v.size() * sizeof(double)); };
    auto vi = [&](const std::vector<int32_t> &v) { mix(v.data(), v.size() * sizeof(int32_t)); };
    mix(&ns, sizeof(ns));
    mix(&nr, sizeof(nr));
    vd(W), vd(Tlow), vd(Thigh), vd(Tcommon), vd(nasa_lo), vd(nasa_hi);
    vi(nf), vi(nb), vi(f_idx), vi(b_idx), vd(f_nu), vd(b_nu), vd(f_ord), vd(b_ord);
    vi(rev), vi(kind), vi(fo_type), vd(arr), vd(arr_lo), vd(arr_rev), vd(fo), vd(eff);
    if (!aux.empty()) vi(aux_off), vd(aux);  // leaves the hash of other mechanisms as it was
    return h;
}

uint64_t Mechanism::hash_transport() const {
    uint64_t h = hash() ^ 0x9e3779b97f4a7c15ull;
    auto mix = [&](const std::vector<double> &v) {
        const unsigned char *c = (const unsigned char *)v.data();
        for (size_t i = 0; i < v.size() * sizeof(double); ++i) {
            h ^= c[i];
            h *= 1099511628211ull;
        }
    };
    mix(ekb);
    mix(sigma);
    return h;
}

// Layout of the PLOG / CHEB data of reaction r (ohc_mech_tables.aux).
static void check_aux(const Mechanism &m, int r) {
    const double *a = m.aux.data() + m.aux_off[r];
    const int n = m.aux_off[r + 1] - m.aux_off[r];
    const std::string where = " in reaction " + m.rxn_names[r];
    if (m.kind[r] == OHC_RATE_PLOG) {
        if (n < 4 || n % 4) throw std::runtime_error("PLOG needs 4 values per pressure point" + where);
        for (int i = 0; i < n; i += 4) {
            if (!(a[i] > 0)) throw std::runtime_error("PLOG pressure must be > 0" + where);
            if (i && a[i] < a[i - 4]) throw std::runtime_error("PLOG pressures not ascending" + where);
        }
    } else if (m.kind[r] == OHC_RATE_CHEB) {
        const bool dims = n >= 6 && a[0] >= 1 && a[1] >= 1 && a[0] <= 64 && a[1] <= 64 &&
                          a[0] == std::floor(a[0]) && a[1] == std::floor(a[1]) &&
                          n == 6 + (int)a[0] * (int)a[1];
        if (!dims) throw std::runtime_error("CHEB needs N, M (1..64) and N*M coefficients" + where);
        if (!(a[2] > 0 && a[3] > a[2])) throw std::runtime_error("TCHEB needs 0 < Tmin < Tmax" + where);
        if (!(a[4] > 0 && a[5] > a[4])) throw std::runtime_error("PCHEB needs 0 < Pmin < Pmax" + where);
    } else if (n != 0) {
        throw std::runtime_error("PLOG/CHEB data on a reaction of another rate law" + where);
    }
}

void load_from_tables(const ohc_mech_tables &t, Mechanism &m) {
    if (t.ns <= 0 || t.ns > OHC_MAXNS) throw std::runtime_error("ns out of range");
    if (t.nr < 0) throw std::runtime_error("nr < 0");
    const int P = OHC_MAXP;
    m = Mechanism();
    m.ne = t.ne;
    m.resize_species(t.ns);
    for (int k = 0; k < t.ns; ++k) m.species[k] = "S" + std::to_string(k);
    auto cp = [](std::vector<double> &d, const double *s, size_t n) {
        if (s) d.assign(s, s + n);
    };
    cp(m.W, t.W, t.ns);
    cp(m.Tlow, t.Tlow, t.ns);
    cp(m.Thigh, t.Thigh, t.ns);
    cp(m.Tcommon, t.Tcommon, t.ns);
    cp(m.nasa_lo, t.nasa_lo, 7 * (size_t)t.ns);
    cp(m.nasa_hi, t.nasa_hi, 7 * (size_t)t.ns);
    if (t.ne > 0) cp(m.elem, t.elem, (size_t)t.ns * t.ne);
    if (t.ekb && t.sigma) {
        cp(m.ekb, t.ekb, t.ns);
        cp(m.sigma, t.sigma, t.ns);
        m.has_transport = true;
    }
    for (int r = 0; r < t.nr; ++r) {
        m.add_reaction();
        m.rxn_names[r] = "R" + std::to_string(r);
    }
    auto cpi = [](std::vector<int32_t> &d, const int32_t *s, size_t n) {
        if (s) d.assign(s, s + n);
    };
    const size_t nr = t.nr;
    cpi(m.nf, t.nf, nr);
    cpi(m.nb, t.nb, nr);
    cpi(m.f_idx, t.f_idx, nr * P);
    cpi(m.b_idx, t.b_idx, nr * P);
    cp(m.f_nu, t.f_nu, nr * P);
    cp(m.b_nu, t.b_nu, nr * P);
    cp(m.f_ord, t.f_ord, nr * P);
    cp(m.b_ord, t.b_ord, nr * P);
    cpi(m.rev, t.rev, nr);
    cpi(m.kind, t.kind, nr);
    cp(m.arr, t.arr, nr * 3);
    cp(m.arr_lo, t.arr_lo, nr * 3);
    cp(m.arr_rev, t.arr_rev, nr * 3);
    cpi(m.fo_type, t.fo_type, nr);
    cp(m.fo, t.fo, nr * 5);
    cp(m.eff, t.eff, nr * (size_t)t.ns);
    for (int r = 0; r < t.nr; ++r) {
        if (m.nf[r] < 0 || m.nf[r] > P || m.nb[r] < 0 || m.nb[r] > P)
            throw std::runtime_error("participant count out of range in reaction " +
                                     std::to_string(r));
        for (int i = 0; i < m.nf[r]; ++i)
            if (m.f_idx[r * P + i] < 0 || m.f_idx[r * P + i] >= t.ns)
                throw std::runtime_error("species index out of range");
        for (int i = 0; i < m.nb[r]; ++i)
            if (m.b_idx[r * P + i] < 0 || m.b_idx[r * P + i] >= t.ns)
                throw std::runtime_error("species index out of range");
        if (m.kind[r] < 0 || m.kind[r] > OHC_RATE_CHEB || m.rev[r] < 0 ||
            m.rev[r] > OHC_RXN_REV_PARAMS)
            throw std::runtime_error("bad reaction type code");
        if (m.kind[r] == OHC_RATE_PLOG || m.kind[r] == OHC_RATE_CHEB) {
            if (!t.aux_off || !t.aux)
                throw std::runtime_error("PLOG/CHEB reaction " + std::to_string(r) +
                                         " without aux_off/aux tables");
            const int32_t a = t.aux_off[r], b = t.aux_off[r + 1];
            if (a < 0 || b < a) throw std::runtime_error("aux_off not ascending");
            m.aux.insert(m.aux.end(), t.aux + a, t.aux + b);
        }
        m.aux_off[r + 1] = (int32_t)m.aux.size();
        check_aux(m, r);
    }
    for (int k = 0; k < t.ns; ++k)
        if (!(m.W[k] > 0)) throw std::runtime_error("molecular weight must be > 0");
}

void parse_chemkin(const std::string &chem, const std::string *thermo,
                   const std::string *transport, Mechanism &m) {
    const int P = OHC_MAXP;
    m = Mechanism();
    auto L = lines_of(chem);
    // ---- elements
    Section es = section(L, {"ELEMENTS", "ELEMENT", "ELEM"});
    if (!es.found) throw std::runtime_error("no ELEMENTS section");
    for (auto &l : es.lines)
        for (auto &t : split_ws(upper(strip_comment(l)))) {
            std::string e = t;
            size_t sl = e.find('/');
            if (sl != std::string::npos) e = e.substr(0, sl);
            if (!e.empty()) m.elements.push_back(e);
        }
    m.ne = (int)m.elements.size();
    // ---- species
    Section ss = section(L, {"SPECIES", "SPECIE", "SPEC"});
    if (!ss.found) throw std::runtime_error("no SPECIES section");
    std::vector<std::string> sp;
    for (auto &l : ss.lines)
        for (auto &t : split_ws(upper(strip_comment(l)))) sp.push_back(t);
    if (sp.empty() || (int)sp.size() > OHC_MAXNS)
        throw std::runtime_error("species count must be in 1.." + std::to_string(OHC_MAXNS));
    m.resize_species((int)sp.size());
    m.species = sp;
    // ---- thermo
    std::map<std::string, NasaEntry> nasa;
    Section ts = section(L, {"THERMO", "THER"});
    if (ts.found) {
        nasa = parse_nasa(ts.lines);
    }
    if (thermo) {
        auto TL = lines_of(*thermo);
        Section t2 = section(TL, {"THERMO", "THER"});
        auto more = parse_nasa(t2.found ? t2.lines : TL);
        for (auto &kv : more)
            if (!nasa.count(kv.first)) nasa[kv.first] = kv.second;
    }
    for (int k = 0; k < m.ns; ++k) {
        auto it = nasa.find(m.species[k]);
        if (it == nasa.end()) throw std::runtime_error("no thermo data for " + m.species[k]);
        const NasaEntry &e = it->second;
        m.Tlow[k] = e.Tlow;
        m.Thigh[k] = e.Thigh;
        m.Tcommon[k] = e.Tcommon;
        for (int c = 0; c < 7; ++c) {
            m.nasa_lo[7 * k + c] = e.lo[c];
            m.nasa_hi[7 * k + c] = e.hi[c];
        }
        double w = 0.0;
        for (auto &a : e.atoms) {
            w += a.second * atomic_weight(a.first);
            auto ei = std::find(m.elements.begin(), m.elements.end(), a.first);
            if (ei == m.elements.end())
                throw std::runtime_error("element " + a.first + " of " + m.species[k] +
                                         " not declared");
            m.elem[(size_t)k * m.ne + (ei - m.elements.begin())] = a.second;
        }
        if (!(w > 0)) throw std::runtime_error("zero molecular weight for " + m.species[k]);
        m.W[k] = w;
    }
    // ---- reactions
    Section rs = section(L, {"REACTIONS", "REACTION", "REAC"});
    if (!rs.found) throw std::runtime_error("no REACTIONS section");
    bool molecules = false;
    int uE = 0;
    for (auto &t : split_ws(rs.header)) {
        if (t == "MOLECULES") molecules = true;
        else if (t == "MOLES") molecules = false;
        else if (unit_code_E(t) >= 0) uE = unit_code_E(t);
    }
    struct Pending {
        std::map<std::string, double> f, b;
        bool reversible = true, has_m = false, has_pm = false;
        std::string pm;
        double A = 0, beta = 0, E = 0;
        std::vector<double> low, high, troe, sri, revp;
        std::vector<double> plog, cheb, tcheb, pcheb;
        std::map<std::string, double> effs;
        std::map<std::string, double> ford, rord;
        std::string name;
    };
    std::vector<Pending> pend;
    for (auto &raw : rs.lines) {
        std::string line = trim(strip_comment(raw));
        if (line.empty()) continue;
        std::string u = upper(line);
        bool is_rxn = (u.find('=') != std::string::npos);
        if (is_rxn) {
            auto tok = split_ws(line);
            if (tok.size() < 4) throw std::runtime_error("malformed reaction line: " + line);
            Pending p;
            p.A = to_double(tok[tok.size() - 3], "A");
            p.beta = to_double(tok[tok.size() - 2], "beta");
            p.E = to_double(tok[tok.size() - 1], "E");
            // reaction text = everything before the three numbers
            std::string eq;
            for (size_t i = 0; i + 3 < tok.size(); ++i) eq += tok[i] + " ";
            eq = trim(eq);
            p.name = eq;
            std::string ueq = upper(eq);
            size_t pos;
            std::string lhs, rhs;
            if ((pos = ueq.find("<=>")) != std::string::npos) {
                lhs = ueq.substr(0, pos);
                rhs = ueq.substr(pos + 3);
            } else if ((pos = ueq.find("=>")) != std::string::npos) {
                lhs = ueq.substr(0, pos);
                rhs = ueq.substr(pos + 2);
                p.reversible = false;
            } else {
                pos = ueq.find('=');
                lhs = ueq.substr(0, pos);
                rhs = ueq.substr(pos + 1);
            }
            bool m1 = false, m2 = false, pm1 = false, pm2 = false;
            std::string s1, s2;
            p.f = parse_side(lhs, m1, pm1, s1);
            p.b = parse_side(rhs, m2, pm2, s2);
            if (m1 != m2 || pm1 != pm2) throw std::runtime_error("unbalanced third body: " + eq);
            p.has_m = m1;
            p.has_pm = pm1;
            p.pm = s1;
            pend.push_back(p);
            continue;
        }
        if (pend.empty()) throw std::runtime_error("auxiliary line before any reaction: " + line);
        Pending &p = pend.back();
        // Every "NAME /values/" group of the line in turn, and bare keywords
        // (DUP): a line may carry LOW together with TROE, or LOW followed by
        // efficiencies (chemkinFileRead.cpp:256-263 handles each keyword found
        // on the line).  Keywords match whole names only.
        for (const AuxGroup &g : aux_groups(u)) {
            const std::string &k = g.name;
            auto vals = [&]() {
                if (!g.has_vals) throw std::runtime_error("missing /.../ values after " + k + " in: " + line);
                std::vector<double> v;
                for (auto &t : split_ws(g.vals)) v.push_back(to_double(t, k.c_str()));
                return v;
            };
            if (k == "DUP" || k == "DUPLICATE") {
                if (g.has_vals) throw std::runtime_error("DUP takes no values: " + line);
            } else if (k == "LOW") {
                p.low = vals();
            } else if (k == "HIGH") {
                p.high = vals();
            } else if (k == "TROE") {
                p.troe = vals();
            } else if (k == "SRI") {
                p.sri = vals();
            } else if (k == "REV") {
                p.revp = vals();
            } else if (k == "FORD" || k == "RORD") {
                if (!g.has_vals) throw std::runtime_error("missing /.../ values after " + k + " in: " + line);
                auto nm = split_ws(g.vals);
                if (nm.size() != 2) throw std::runtime_error("bad FORD/RORD: " + line);
                (k == "FORD" ? p.ford : p.rord)[nm[0]] = to_double(nm[1], "order");
            } else if (k == "PLOG") {  // chemkinFileRead.cpp:606-644: one point per group
                auto v = vals();
                if (v.size() != 4) throw std::runtime_error("PLOG needs /p A beta E/ in: " + line);
                p.plog.insert(p.plog.end(), v.begin(), v.end());
            } else if (k == "CHEB") {  // :645-685: every CHEB group appends its numbers
                auto v = vals();
                p.cheb.insert(p.cheb.end(), v.begin(), v.end());
            } else if (k == "TCHEB" || k == "PCHEB") {  // :686-755
                auto v = vals();
                if (v.size() != 2) throw std::runtime_error(k + " needs /min max/ in: " + line);
                (k == "TCHEB" ? p.tcheb : p.pcheb) = v;
            } else if (k == "LT" ||
                       k == "RLT" || k == "JAN" || k == "FIT1" || k == "HV" || k == "TDEP" ||
                       k == "EXCI" || k == "MOME" || k == "XSMI" || k == "UNITS" ||
                       k == "COLLEFF") {
                throw std::runtime_error("unsupported CHEMKIN keyword " + k + " in: " + line);
            } else if (g.has_vals) {
                // third-body efficiency "SP/val/"
                p.effs[k] = to_double(g.vals, "efficiency");
            } else {
                throw std::runtime_error("unrecognised auxiliary keyword " + k + " in: " + line);
            }
        }
    }
    for (auto &p : pend) {
        int r = m.add_reaction();
        m.rxn_names[r] = p.name;
        auto fill = [&](const std::map<std::string, double> &side,
                        const std::map<std::string, double> &ord, int32_t *idx, double *nu,
                        double *o, int32_t &cnt) {
            if ((int)side.size() > P) throw std::runtime_error("too many participants: " + p.name);
            cnt = 0;
            for (auto &kv : side) {  // std::map: name order
                int k = m.species_index(kv.first);
                if (k < 0) throw std::runtime_error("unknown species " + kv.first + " in " + p.name);
                idx[cnt] = k;
                nu[cnt] = kv.second;
                auto oi = ord.find(kv.first);
                o[cnt] = oi == ord.end() ? kv.second : oi->second;
                ++cnt;
            }
        };
        fill(p.f, p.ford, &m.f_idx[r * P], &m.f_nu[r * P], &m.f_ord[r * P], m.nf[r]);
        fill(p.b, p.rord, &m.b_idx[r * P], &m.b_nu[r * P], &m.b_ord[r * P], m.nb[r]);
        // element balance (chemkinFileRead::checkElementBalance)
        for (int e = 0; e < m.ne; ++e) {
            double bal = 0.0;
            for (int i = 0; i < m.nf[r]; ++i) bal -= m.f_nu[r * P + i] * m.elem[m.f_idx[r * P + i] * m.ne + e];
            for (int i = 0; i < m.nb[r]; ++i) bal += m.b_nu[r * P + i] * m.elem[m.b_idx[r * P + i] * m.ne + e];
            if (std::fabs(bal) > 1e-6) throw std::runtime_error("reaction not element-balanced: " + p.name);
        }
        double ordf = 0.0, ordb = 0.0;
        for (int i = 0; i < m.nf[r]; ++i) ordf += m.f_ord[r * P + i];
        for (int i = 0; i < m.nb[r]; ++i) ordb += m.b_ord[r * P + i];
        // efficiencies
        const bool tb = p.has_m || p.has_pm;
        if (tb) {
            double *e = &m.eff[(size_t)r * m.ns];
            if (p.has_pm && p.pm != "M") {
                int k = m.species_index(p.pm);
                if (k < 0) throw std::runtime_error("unknown enhanced species " + p.pm);
                for (int i = 0; i < m.ns; ++i) e[i] = 0.0;
                e[k] = 1.0;
            } else {
                for (int i = 0; i < m.ns; ++i) e[i] = 1.0;
            }
            for (auto &kv : p.effs) {
                int k = m.species_index(kv.first);
                if (k < 0) throw std::runtime_error("unknown species in efficiencies: " + kv.first);
                e[k] = kv.second;
            }
        }
        double Af, Ef, Afr = 0, Efr = 0;
        change_unit(ordf, molecules, uE, tb, Af, Ef);
        if (!p.revp.empty()) change_unit(ordb, molecules, uE, tb, Afr, Efr);
        const double base = 1e-3;
        m.rev[r] = !p.reversible ? OHC_RXN_IRREVERSIBLE
                                 : (p.revp.empty() ? OHC_RXN_REVERSIBLE : OHC_RXN_REV_PARAMS);
        const bool plog = !p.plog.empty(),
                   cheb = !p.cheb.empty() || !p.tcheb.empty() || !p.pcheb.empty();
        if (plog || cheb) {
            // chemkinFileRead.cpp:606-755 (reading) and :1595-1650 (units):
            // PLOG pressures atm -> Pa, A and E as the main line; CHEB
            // pressure limits atm -> Pa, coefficients as written.
            if (plog && cheb) throw std::runtime_error("PLOG and CHEB in one reaction: " + p.name);
            if (!p.low.empty() || !p.high.empty() || !p.troe.empty() || !p.sri.empty())
                throw std::runtime_error("LOW/HIGH/TROE/SRI with a PLOG or CHEB rate: " + p.name);
            if (!p.revp.empty())
                throw std::runtime_error("REV cannot be used with a PLOG or CHEB rate: " + p.name);
            m.arr[3 * r + 0] = p.A * Af;  // main line, kept as read; the law does not use it
            m.arr[3 * r + 1] = p.beta;
            m.arr[3 * r + 2] = p.E / Ef;
            std::vector<double> v;
            if (plog) {
                if (p.has_pm)
                    throw std::runtime_error("PLOG cannot be used in a (+M) reaction: " + p.name);
                m.kind[r] = OHC_RATE_PLOG;
                // ascending pressure, equal pressures in reading order
                // (logInterpolationRR::checkSizeAndPresureRank's bubble sort)
                std::vector<int> ord(p.plog.size() / 4);
                for (size_t i = 0; i < ord.size(); ++i) ord[i] = (int)i;
                std::stable_sort(ord.begin(), ord.end(),
                                 [&](int a, int b) { return p.plog[4 * a] < p.plog[4 * b]; });
                for (int i : ord) {
                    const double *q = &p.plog[4 * i];
                    v.insert(v.end(), {q[0] * kPatm, q[1] * Af, q[2], q[3] / Ef});
                }
            } else {
                if (!p.has_pm)
                    throw std::runtime_error("CHEB needs a (+M) pressure-dependent reaction: " + p.name);
                if (p.cheb.size() < 2)
                    throw std::runtime_error("TCHEB/PCHEB without CHEB coefficients: " + p.name);
                m.kind[r] = OHC_RATE_CHEB;
                const double Tmin = p.tcheb.empty() ? 300.0 : p.tcheb[0],
                             Tmax = p.tcheb.empty() ? 2500.0 : p.tcheb[1],
                             Pmin = p.pcheb.empty() ? 0.001 : p.pcheb[0],
                             Pmax = p.pcheb.empty() ? 100.0 : p.pcheb[1];
                v = {p.cheb[0], p.cheb[1], Tmin, Tmax, Pmin * kPatm, Pmax * kPatm};
                v.insert(v.end(), p.cheb.begin() + 2, p.cheb.end());
            }
            m.set_aux(r, v);
            check_aux(m, r);
            continue;
        }
        if (p.has_pm) {
            if (!p.revp.empty()) throw std::runtime_error("REV with fall-off not supported: " + p.name);
            const bool chemact = !p.high.empty();
            if (!chemact && p.low.size() != 3) throw std::runtime_error("fall-off needs LOW: " + p.name);
            m.kind[r] = chemact ? OHC_RATE_CHEMACT : OHC_RATE_FALLOFF;
            const std::vector<double> &aux = chemact ? p.high : p.low;
            if (aux.size() != 3) throw std::runtime_error("bad LOW/HIGH: " + p.name);
            double *ki = &m.arr[3 * r], *k0 = &m.arr_lo[3 * r];
            if (!chemact) {  // main line = k_inf, LOW = k0
                ki[0] = p.A * Af;
                ki[1] = p.beta;
                ki[2] = p.E / Ef;
                k0[0] = aux[0] * base * Af;
                k0[1] = aux[1];
                k0[2] = aux[2] / Ef;
            } else {  // main line = k0, HIGH = k_inf
                k0[0] = p.A * Af;
                k0[1] = p.beta;
                k0[2] = p.E / Ef;
                ki[0] = aux[0] * Af / base;
                ki[1] = aux[1];
                ki[2] = aux[2] / Ef;
            }
            if (!p.troe.empty()) {
                if (p.troe.size() != 3 && p.troe.size() != 4) throw std::runtime_error("bad TROE: " + p.name);
                m.fo_type[r] = p.troe.size() == 3 ? OHC_FO_TROE3 : OHC_FO_TROE4;
                for (size_t i = 0; i < p.troe.size(); ++i) m.fo[5 * r + i] = p.troe[i];
            } else if (!p.sri.empty()) {
                if (p.sri.size() != 3 && p.sri.size() != 5) throw std::runtime_error("bad SRI: " + p.name);
                m.fo_type[r] = OHC_FO_SRI;
                m.fo[5 * r + 3] = 1.0;
                m.fo[5 * r + 4] = 0.0;
                for (size_t i = 0; i < p.sri.size(); ++i) m.fo[5 * r + i] = p.sri[i];
            } else {
                m.fo_type[r] = OHC_FO_LINDEMANN;
            }
        } else {
            m.kind[r] = p.has_m ? OHC_RATE_THIRDBODY : OHC_RATE_ARRHENIUS;
            const double b = p.has_m ? base : 1.0;
            m.arr[3 * r + 0] = b * p.A * Af;
            m.arr[3 * r + 1] = p.beta;
            m.arr[3 * r + 2] = p.E / Ef;
            if (!p.revp.empty()) {
                if (p.revp.size() != 3) throw std::runtime_error("bad REV: " + p.name);
                m.arr_rev[3 * r + 0] = b * p.revp[0] * Afr;
                m.arr_rev[3 * r + 1] = p.revp[1];
                m.arr_rev[3 * r + 2] = p.revp[2] / Efr;
            }
        }
    }
    // ---- transport
    if (transport) {
        std::map<std::string, std::vector<double>> tr;
        for (auto &l : lines_of(*transport)) {
            auto t = split_ws(strip_comment(l));
            if (t.size() < 7) continue;
            std::vector<double> v;
            for (int i = 1; i < 7; ++i) v.push_back(to_double(t[i], "transport"));
            tr[upper(t[0])] = v;
        }
        for (int k = 0; k < m.ns; ++k) {
            auto it = tr.find(m.species[k]);
            if (it == tr.end()) throw std::runtime_error("no transport data for " + m.species[k]);
            m.geom[k] = it->second[0];
            m.ekb[k] = it->second[1];
            m.sigma[k] = it->second[2];
            m.dipole[k] = it->second[3];
            m.polar[k] = it->second[4];
            m.zrot[k] = it->second[5];
        }
        m.has_transport = true;
    }
}

}  // namespace ohc
