// Jump-ahead polynomials for Xoshiro256++ (see rng_jump.cpp).
#pragma once

#include <cstdint>

#include "core.hpp"

namespace cpm {

struct JumpPoly {
    uint64_t w[4];  // coefficient of x^i is bit i
};

// x^distance mod p(x), p the generator's characteristic polynomial.
JumpPoly xo_jump_poly(uint64_t distance);
// s <- the state `distance` draws ahead (q = xo_jump_poly(distance)).
void xo_apply_jump(Xoshiro& s, const JumpPoly& q);
// Throws if jumping disagrees with stepping.
void xo_jump_selfcheck();

}  // namespace cpm
