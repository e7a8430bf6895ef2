// mech.hpp -- host-side mechanism model for the MI355X chemistry engine.
//
// A Mechanism holds species data (molecular weight, NASA-7 polynomials,
// element composition, Lennard-Jones transport data) and reactions as
// structure-of-arrays by reaction, already converted to the SI kmol-m^3-s-K
// unit system the reference works in (chemkinFileRead::changeUnit,
// src/speciesTransport/chemistryRead/src/chemkinFileRead.cpp:1654-1699).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ohc.h"

namespace ohc {

constexpr double kR = 1.380649e-23 * 6.02214076e23;  // J/mol/K (k * N_A)
constexpr double kRu = kR * 1e3;                     // J/kmol/K
constexpr double kPatm = 1.01325e5;                  // Pa
constexpr double kCal = 4.184;                       // thermochemical calorie
constexpr double kNA = 6.02214076e23;
constexpr double kEv = 1.602176634e-19;

struct Mechanism {
    std::vector<std::string> elements, species, rxn_names;
    int32_t ns = 0, nr = 0, ne = 0;
    bool has_transport = false;
    // species
    std::vector<double> W, Tlow, Thigh, Tcommon, nasa_lo, nasa_hi, elem;
    std::vector<double> ekb, sigma, geom, dipole, polar, zrot;
    // reactions (OHC_MAXP slots per side)
    std::vector<int32_t> nf, nb, f_idx, b_idx, rev, kind, fo_type;
    std::vector<double> f_nu, b_nu, f_ord, b_ord, arr, arr_lo, arr_rev, fo, eff;
    // PLOG / CHEB data: reaction r owns aux[aux_off[r], aux_off[r+1])
    std::vector<int32_t> aux_off{0};
    std::vector<double> aux;

    void resize_species(int n);
    int add_reaction();  // appends zeroed reaction slots, returns its index
    void set_aux(int r, const std::vector<double> &v);  // r must be the last reaction
    ohc_mech_tables tables() const;
    uint64_t hash() const;            // kinetics + thermo tables
    uint64_t hash_transport() const;  // hash() combined with transport data
    int species_index(const std::string &name) const;
};

// Parse CHEMKIN-II reaction file text (+ optional NASA-7 thermo text when the
// reaction file has no THERMO section, + optional transport text).
// Throws std::runtime_error with a message on malformed input.
void parse_chemkin(const std::string &chem, const std::string *thermo,
                   const std::string *transport, Mechanism &m);
void load_from_tables(const ohc_mech_tables &t, Mechanism &m);

// Seeded synthetic element-balanced mechanism: the CHEMKIN, NASA-7 and
// transport texts, and the mechanism parsed from them.
void synthetic_chemkin(int ns, int nr, uint64_t seed, double frac_rev, double frac_tb,
                       std::string &chem, std::string &thermo, std::string &transport);
void synthetic_mechanism(int ns, int nr, uint64_t seed, double frac_rev, double frac_tb,
                         Mechanism &m);

double atomic_weight(const std::string &elem);  // CHEMKIN-II atomic weights

}  // namespace ohc
