// Reverse-mode adjoints of recorded chunks: the host side that turns a chunk into an adjoint
// program (adjoint.hpp) and the specialised source for it.
#pragma once

#include "scihip.h"

#include <cstdint>
#include <string>
#include <vector>

namespace scihip {

    // An adjoint program as the host builds it, before it is copied into an Adjoint<T>.
    struct AdjointPlan {
        uint8_t precision = SCIHIP_SINGLE;
        bool seeded = false;    // the result's adjoint is an upstream gradient array (the last input)
        bool loss = false;      // the sum of the result is folded
        bool contract = false;  // ... as the sum of left * right (a contracted chunk)
        uint32_t arguments = 0; // inputs, the upstream gradient included
        uint32_t gradients = 0; // the chunk's inputs
        uint32_t registers = 0; // inputs + constants + values
        uint32_t left = 0, right = 0;
        std::vector<uint32_t> word;  // code | target << 8 | first << 16 | second << 24
        std::vector<uint8_t> third;
        std::vector<uint8_t> slot;   // register of each input's gradient
        std::vector<double> constants;
    };

    // Emits the adjoint program of a well-formed chunk. Returns SCIHIP_OK, or SCIHIP_EINVAL with a
    // message when the program would exceed the kernel's capacity.
    int plan_adjoint(const scihip_chunk &chunk, bool seeded, bool loss, AdjointPlan &plan, std::string &error);

    // Straight-line HIP source of the plan's specialised kernel, `unroll` vectors per thread per sweep.
    std::string generate_adjoint(const AdjointPlan &plan, int unroll);

    // The compiled-kernel cache key of a plan.
    std::string adjoint_key(const AdjointPlan &plan);

} // namespace scihip
