// synth.cpp -- seeded synthetic element-balanced mechanisms.
#include <stdexcept>

#include "mech.hpp"

namespace ohc {

void synthetic_mechanism(int ns, int nr, uint64_t seed, double frac_rev, double frac_tb,
                         Mechanism &m) {
    (void)ns, (void)nr, (void)seed, (void)frac_rev, (void)frac_tb, (void)m;
    throw std::runtime_error("synthetic mechanism generator not built yet");
}

}  // namespace ohc
