// codegen_main.cpp -- build-time tool: mechanism files -> descriptor header.
//   ohc_codegen NAME CHEMKIN THERMO|- TRANSPORT|- OUT.h
//   ohc_codegen --synthetic NS NR SEED FRAC_REV FRAC_TB OUTDIR
//       writes OUTDIR/{mech.che, thermo.dat, transport.dat} of the seeded
//       synthetic mechanism (the same one ohc_mech_synthetic builds)
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>

#include "codegen.hpp"

static std::string slurp(const char *p) {
    std::ifstream f(p, std::ios::binary);
    if (!f) throw std::runtime_error(std::string("cannot read ") + p);
    std::ostringstream s;
    s << f.rdbuf();
    return s.str();
}

static int write_synthetic(char **a) {
    std::string ck, th, tr;
    ohc::synthetic_chemkin(std::atoi(a[0]), std::atoi(a[1]), std::strtoull(a[2], nullptr, 10),
                           std::atof(a[3]), std::atof(a[4]), ck, th, tr);
    const std::string dir = a[5];
    const std::pair<const char *, std::string *> files[] = {
        {"/mech.che", &ck}, {"/thermo.dat", &th}, {"/transport.dat", &tr}};
    for (auto &f : files) {
        std::ofstream o(dir + f.first);
        o << *f.second;
        if (!o) throw std::runtime_error("cannot write " + dir + f.first);
    }
    return 0;
}

int main(int argc, char **argv) {
    if (argc == 8 && std::string(argv[1]) == "--synthetic") {
        try {
            return write_synthetic(argv + 2);
        } catch (const std::exception &e) {
            std::cerr << "ohc_codegen: " << e.what() << "\n";
            return 1;
        }
    }
    if (argc != 6) {
        std::cerr << "usage: ohc_codegen NAME CHEMKIN THERMO|- TRANSPORT|- OUT.h\n";
        return 2;
    }
    try {
        std::string ck = slurp(argv[2]), th, tr;
        const bool hth = std::string(argv[3]) != "-", htr = std::string(argv[4]) != "-";
        if (hth) th = slurp(argv[3]);
        if (htr) tr = slurp(argv[4]);
        ohc::Mechanism m;
        ohc::parse_chemkin(ck, hth ? &th : nullptr, htr ? &tr : nullptr, m);
        std::ofstream o(argv[5]);
        o << ohc::codegen_mechanism(m, argv[1]);
        if (!o) throw std::runtime_error(std::string("cannot write ") + argv[5]);
    } catch (const std::exception &e) {
        std::cerr << "ohc_codegen: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
