// codegen.hpp -- compile-time mechanism descriptor emission.
#pragma once

#include <string>

#include "mech.hpp"

namespace ohc {
std::string codegen_mechanism(const Mechanism &m, const std::string &name);
}
