// Launchers for the kernels compiled into the library, one translation unit per element type
// (kernels_f32.hip, kernels_f64.hip) so that they build in parallel.
#pragma once

#include "interpreter.hpp"
#include "scihip.h"

#include <hip/hip_runtime.h>

namespace scihip {

    // Register-file sizes of the interpreter variants: small files stay in VGPRs with W lanes per
    // thread; the widest (every input, constant and temporary a chunk can hold, plus the output)
    // keeps its file in scratch memory and runs one lane per thread.
    constexpr int kSmall = 8;
    constexpr int kMedium = 16;
    constexpr int kLarge = 196;

    template <typename T> struct Traits;
    template <> struct Traits<float> {
        static constexpr int width = 4; // lanes per thread: one 16-byte access per array
    };
    template <> struct Traits<double> {
        static constexpr int width = 2;
    };

    // Launches the interpreter for `registers` (kSmall, kMedium or kLarge); mode 0 evaluates,
    // 1 reduces. Returns the launch error.
    hipError_t interpret(const Launch<float> &launch, int registers, int mode, unsigned blocks, hipStream_t stream);
    hipError_t interpret(const Launch<double> &launch, int registers, int mode, unsigned blocks, hipStream_t stream);

    // Folds the `count` partials a reduction kernel left, in workgroup order, into *result (one
    // workgroup; NaN when a minimum / maximum saw one).
    hipError_t fold_partials(uint32_t fold, const Partial *partials, unsigned count, double *result, hipStream_t stream);

    // The function of each interpreter variant, for occupancy queries.
    const void *interpreter_function(int precision, int registers, int mode);

    // Seeded uniform data, bit-identical to the oracle's generator.
    hipError_t fill_uniform(void *output, int precision, size_t count, uint64_t seed, double lower, double upper,
                            uint64_t nan_seed, double nan_fraction, hipStream_t stream);

} // namespace scihip
