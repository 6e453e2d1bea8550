// Small elementwise and reduction kernels of the solvers: precision changes, x += d, norms.
#include "linear.hpp"

#include <algorithm>

namespace scihip {

    namespace {

        template <typename D, typename S>
        __global__ __launch_bounds__(256) void convert_kernel(const S *source, D *target, int64_t count) {
            for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
                 i += (int64_t)gridDim.x * blockDim.x)
                target[i] = (D)source[i];
        }

        template <typename S>
        __global__ __launch_bounds__(256) void accumulate_kernel(double *target, const S *delta, int64_t count) {
            for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
                 i += (int64_t)gridDim.x * blockDim.x)
                target[i] = target[i] + (double)delta[i];
        }

        // One workgroup per row: the row's columns from copy_from on copied to target (converted to D)
        // and the sum of |a_ij| over the whole row, two elements per thread per 16-byte load where the
        // rows are 16-byte aligned (copy_from even then).
        template <typename D>
        __global__ __launch_bounds__(256) void copy_row_sums_kernel(const double *a, int64_t columns, D *target,
                                                                    double *sums, int64_t copy_from) {
            const int64_t row = blockIdx.x;
            const double *source = a + row * columns;
            D *out = target + row * columns;
            double s = 0.0;
            if (columns % 2 == 0 && reinterpret_cast<uintptr_t>(a) % 16 == 0) {
                typedef double Two __attribute__((ext_vector_type(2)));
                for (int64_t j = 2 * (int64_t)threadIdx.x; j < columns; j += 2 * blockDim.x) {
                    const Two v = *reinterpret_cast<const Two *>(source + j);
                    if (j >= copy_from) {
                        out[j] = (D)v[0];
                        out[j + 1] = (D)v[1];
                    }
                    s += __builtin_fabs(v[0]);
                    s += __builtin_fabs(v[1]);
                }
            } else {
                for (int64_t j = threadIdx.x; j < columns; j += blockDim.x) {
                    if (j >= copy_from) out[j] = (D)source[j];
                    s += __builtin_fabs(source[j]);
                }
            }
#pragma unroll
            for (int offset = 32; offset > 0; offset >>= 1) s += __shfl_xor(s, offset, 64);
            __shared__ double waves[4];
            if ((threadIdx.x & 63) == 0) waves[threadIdx.x >> 6] = s;
            __syncthreads();
            if (threadIdx.x == 0) sums[row] = (waves[0] + waves[1]) + (waves[2] + waves[3]);
        }

        // Columns [0, count) of every row of a (row length columns) copied to target (converted to D).
        template <typename D>
        __global__ __launch_bounds__(256) void copy_columns_kernel(const double *a, int64_t rows, int64_t columns,
                                                                   int64_t count, D *target) {
            const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
            if (idx >= rows * count) return;
            const int64_t r = idx / count, c = idx % count;
            target[r * columns + c] = (D)a[r * columns + c];
        }

        // Partial column maxima of |x| over row chunks: partial[chunk * width + c].
        __global__ __launch_bounds__(256) void column_max_kernel(const double *x, int64_t rows, int64_t width,
                                                                 int64_t chunk, double *partial) {
            const int64_t c = threadIdx.x % 64 + (int64_t)blockIdx.y * 64;
            const int64_t lane_rows = threadIdx.x / 64;
            const int64_t first = (int64_t)blockIdx.x * chunk;
            const int64_t last = std::min(rows, first + chunk);
            double m = 0.0;
            bool nan = false;
            if (c < width)
                for (int64_t r = first + lane_rows; r < last; r += 4) {
                    const double v = __builtin_fabs(x[r * width + c]);
                    nan = nan || __builtin_isnan(v);
                    m = v > m ? v : m;
                }
            __shared__ double best[4][64];
            best[lane_rows][threadIdx.x % 64] = nan ? __builtin_nan("") : m;
            __syncthreads();
            if (lane_rows == 0 && c < width) {
                double v = best[0][threadIdx.x];
                for (int k = 1; k < 4; ++k) {
                    const double o = best[k][threadIdx.x];
                    v = (__builtin_isnan(v) || __builtin_isnan(o)) ? v + o : (o > v ? o : v);
                }
                partial[(int64_t)blockIdx.x * width + c] = v;
            }
        }

        // One workgroup folds column_max partials over their chunks into pinned host memory: sets of
        // `width` columns (chunks rows each, set s at partial + s * chunks * width) to out[s * width +
        // c], then (extra > 0) one more single column of `extra` chunk maxima at partial + sets *
        // chunks * width to out[sets * width], then *error to the next word. Each fold keeps the
        // larger of the running value and the next partial, starting from 0 (a NaN partial is passed
        // over, as a host loop of std::max from 0 would do).
        __global__ __launch_bounds__(256) void fold_norms_kernel(const double *partial, int64_t chunks, int64_t width,
                                                                 int sets, int64_t extra, const unsigned *error,
                                                                 double *out) {
            for (int64_t idx = threadIdx.x; idx < sets * width; idx += blockDim.x) {
                const int64_t s = idx / width, c = idx % width;
                double m = 0.0;
                for (int64_t k = 0; k < chunks; ++k) {
                    const double v = partial[(s * chunks + k) * width + c];
                    m = m < v ? v : m;
                }
                out[idx] = m;
            }
            if (threadIdx.x == 0) {
                if (extra > 0) {
                    double m = 0.0;
                    for (int64_t k = 0; k < extra; ++k) {
                        const double v = partial[sets * chunks * width + k];
                        m = m < v ? v : m;
                    }
                    out[sets * width] = m;
                }
                out[sets * width + 1] = error ? (double)__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
            }
        }

        unsigned grid_for(int64_t count) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((count + 255) / 256, 8192)); }

    } // namespace

    hipError_t convert(const double *source, float *target, int64_t count, hipStream_t stream) {
        if (count == 0) return hipSuccess;
        hipLaunchKernelGGL((convert_kernel<float, double>), dim3(grid_for(count)), dim3(256), 0, stream, source, target, count);
        return hipGetLastError();
    }

    hipError_t convert(const float *source, double *target, int64_t count, hipStream_t stream) {
        if (count == 0) return hipSuccess;
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid_for(count)), dim3(256), 0, stream, source, target, count);
        return hipGetLastError();
    }

    hipError_t accumulate(double *target, const float *delta, int64_t count, hipStream_t stream) {
        if (count == 0) return hipSuccess;
        hipLaunchKernelGGL(accumulate_kernel<float>, dim3(grid_for(count)), dim3(256), 0, stream, target, delta, count);
        return hipGetLastError();
    }

    hipError_t accumulate(double *target, const double *delta, int64_t count, hipStream_t stream) {
        if (count == 0) return hipSuccess;
        hipLaunchKernelGGL(accumulate_kernel<double>, dim3(grid_for(count)), dim3(256), 0, stream, target, delta, count);
        return hipGetLastError();
    }

    hipError_t copy_with_row_sums(const double *a, int64_t rows, int64_t columns, double *target, double *sums,
                                  hipStream_t stream, int64_t copy_from) {
        if (rows == 0) return hipSuccess;
        hipLaunchKernelGGL(copy_row_sums_kernel<double>, dim3((unsigned)rows), dim3(256), 0, stream, a, columns, target,
                           sums, copy_from);
        return hipGetLastError();
    }

    hipError_t copy_with_row_sums(const double *a, int64_t rows, int64_t columns, float *target, double *sums,
                                  hipStream_t stream, int64_t copy_from) {
        if (rows == 0) return hipSuccess;
        hipLaunchKernelGGL(copy_row_sums_kernel<float>, dim3((unsigned)rows), dim3(256), 0, stream, a, columns, target,
                           sums, copy_from);
        return hipGetLastError();
    }

    hipError_t copy_columns(const double *a, int64_t rows, int64_t columns, int64_t count, double *target,
                            hipStream_t stream) {
        if (rows == 0 || count == 0) return hipSuccess;
        hipLaunchKernelGGL(copy_columns_kernel<double>, dim3((unsigned)((rows * count + 255) / 256)), dim3(256), 0,
                           stream, a, rows, columns, count, target);
        return hipGetLastError();
    }

    hipError_t copy_columns(const double *a, int64_t rows, int64_t columns, int64_t count, float *target,
                            hipStream_t stream) {
        if (rows == 0 || count == 0) return hipSuccess;
        hipLaunchKernelGGL(copy_columns_kernel<float>, dim3((unsigned)((rows * count + 255) / 256)), dim3(256), 0,
                           stream, a, rows, columns, count, target);
        return hipGetLastError();
    }

    // Chunks of about 64 rows (16 per thread row), at most 512: enough workgroups that the loads of a
    // 16384-row column set are a few latencies deep (1024-row chunks took 96 us for 4 MB).
    int64_t column_max_chunks(int64_t rows) { return std::max<int64_t>(1, std::min<int64_t>(512, (rows + 63) / 64)); }

    hipError_t fold_norms(const double *partial, int64_t chunks, int64_t width, int sets, int64_t extra,
                          const unsigned *error, double *host_out, hipStream_t stream) {
        hipLaunchKernelGGL(fold_norms_kernel, dim3(1), dim3(256), 0, stream, partial, chunks, width, sets, extra, error,
                           host_out);
        return hipGetLastError();
    }

    hipError_t column_max(const double *x, int64_t rows, int64_t width, double *partial, hipStream_t stream) {
        if (rows == 0 || width == 0) return hipSuccess;
        const int64_t chunks = column_max_chunks(rows);
        const int64_t chunk = (rows + chunks - 1) / chunks;
        hipLaunchKernelGGL(column_max_kernel, dim3((unsigned)chunks, (unsigned)((width + 63) / 64)), dim3(256), 0,
                           stream, x, rows, width, chunk, partial);
        return hipGetLastError();
    }

} // namespace scihip
