// codegen_main.cpp -- build-time tool: mechanism files -> descriptor header.
//   ohc_codegen NAME CHEMKIN THERMO|- TRANSPORT|- OUT.h
#include <cstdio>
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

int main(int argc, char **argv) {
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
