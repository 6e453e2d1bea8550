// Device-resident transpose and per-axis reductions (axis.hip).
#pragma once

#include "interpreter.hpp"

#include <hip/hip_runtime.h>

namespace scihip {

    // A chunk folded along one axis of a rows x columns row-major matrix (every input of the chunk is
    // such a matrix): output[r] folds row r (axis 1), output[c] folds column c (axis 0).
    template <typename T> struct AxisLaunch {
        Program<T> p;         // p.count = rows * columns; p.aligned: every row starts 16-byte aligned
        T *output;
        Partial *partials;    // axis 0: slices x columns
        uint64_t rows, columns;
    };

    // out (columns x rows) = in^T (rows x columns), both row-major.
    template <typename T> hipError_t transpose(const T *in, T *out, int64_t rows, int64_t columns, hipStream_t stream);

    // Launches the fold for the interpreter variant `registers`; axis 0 runs `slices` row slices.
    template <typename T>
    hipError_t fold_axis(const AxisLaunch<T> &launch, int registers, int axis, uint32_t slices, hipStream_t stream);

    // out[0, n) = value.
    template <typename T> hipError_t fill_value(T *out, uint64_t n, T value, hipStream_t stream);

} // namespace scihip
