// Small kernels of the solvers (linear.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace scihip {

    hipError_t convert(const double *source, float *target, int64_t count, hipStream_t stream);
    hipError_t convert(const float *source, double *target, int64_t count, hipStream_t stream);
    // target += delta, in double.
    hipError_t accumulate(double *target, const float *delta, int64_t count, hipStream_t stream);
    hipError_t accumulate(double *target, const double *delta, int64_t count, hipStream_t stream);
    // target = a (row-major, converted to the target's type; only columns copy_from onwards) and
    // sums[i] = sum_j |a_ij| over the whole row (the infinity norm is their maximum), reading a once
    // for both. copy_from is even.
    hipError_t copy_with_row_sums(const double *a, int64_t rows, int64_t columns, double *target, double *sums,
                                  hipStream_t stream, int64_t copy_from = 0);
    hipError_t copy_with_row_sums(const double *a, int64_t rows, int64_t columns, float *target, double *sums,
                                  hipStream_t stream, int64_t copy_from = 0);
    // Columns [0, count) of target = a's (converted), the rest left alone.
    hipError_t copy_columns(const double *a, int64_t rows, int64_t columns, int64_t count, double *target,
                            hipStream_t stream);
    hipError_t copy_columns(const double *a, int64_t rows, int64_t columns, int64_t count, float *target,
                            hipStream_t stream);
    // Folds column_max partials into host memory the device can write (pinned): `sets` sets of `width`
    // columns to host_out[s * width + c]; with extra > 0, one more single column of `extra` chunks to
    // host_out[sets * width]; then *error (0 if null) to host_out[sets * width + 1].
    hipError_t fold_norms(const double *partial, int64_t chunks, int64_t width, int sets, int64_t extra,
                          const unsigned *error, double *host_out, hipStream_t stream);
    // Partial column maxima of |x|: column_max_chunks(rows) rows of `width` values.
    int64_t column_max_chunks(int64_t rows);
    hipError_t column_max(const double *x, int64_t rows, int64_t width, double *partial, hipStream_t stream);

} // namespace scihip
