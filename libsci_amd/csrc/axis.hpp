// Device-resident transpose and per-axis reductions (axis.hip).
#pragma once

#include "interpreter.hpp"

#include <hip/hip_runtime.h>

namespace scihip {

    // out (columns x rows) = in^T (rows x columns), both row-major.
    template <typename T> hipError_t transpose(const T *in, T *out, int64_t rows, int64_t columns, hipStream_t stream);

    // Launches the fold for the interpreter variant `registers`; axis 0 runs `slices` row slices.
    template <typename T>
    hipError_t fold_axis(const AxisLaunch<T> &launch, int registers, int axis, uint32_t slices, hipStream_t stream);

    // Folds the slices' partials of a column fold (axis 0) into the output, in slice order.
    template <typename T> hipError_t finish_axis(const AxisLaunch<T> &launch, uint32_t slices, hipStream_t stream);

    // out[0, n) = value.
    template <typename T> hipError_t fill_value(T *out, uint64_t n, T value, hipStream_t stream);

} // namespace scihip
