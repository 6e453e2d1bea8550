// Device-resident transpose and per-axis reductions of row-major matrices. The reference transposes
// on the host into a host array (include/operations/manipulations.hpp:36-62) and has no axis
// reductions; each row's fold here means what scihip_reduce means for that row alone (the sum in
// double, NaN propagating; minimum / maximum skipping NaN but returning NaN if one was seen; a
// contracted sum folding exact products), cast to the element type as reduce's result is
// (include/array/expression.hpp:198).
//
//  - transpose: square tiles through LDS (one padding column against bank conflicts); full tiles
//    move 16 bytes per lane on both sides, edge tiles element by element. 8 B/float of traffic.
//  - fold_rows: one workgroup per row; threads stride along the row W lanes at a time (16-byte loads
//    when every row starts 16-byte aligned), fold in double, then across the workgroup.
//  - fold_columns: each thread owns W adjacent columns and walks one slice of the rows, so a wave
//    reads 64 * W consecutive elements of a row per step; the slices' partials are folded in slice
//    order by finish_columns, so results do not depend on timing.
#include "axis.hpp"
#include "kernels.hpp"

namespace scihip {

    namespace {

        template <typename T> struct Wide;
        template <> struct Wide<float> {
            typedef float type __attribute__((ext_vector_type(4)));
            static constexpr int lanes = 4;
        };
        template <> struct Wide<double> {
            typedef double type __attribute__((ext_vector_type(2)));
            static constexpr int lanes = 2;
        };

        // Tile edge per element type: 128 x 128 floats (66 KB of LDS), 64 x 64 doubles (33 KB). The
        // larger float tile writes 512-byte runs of each output row, which measured 12% faster than
        // 64 x 64 tiles' 256-byte runs.
        template <typename T> constexpr int kTileOf = sizeof(T) == 4 ? 128 : 64;

        template <typename T>
        __global__ __launch_bounds__(256) void transpose_tiles(const T *__restrict__ in, T *__restrict__ out,
                                                               int64_t rows, int64_t columns, bool vectors) {
            constexpr int kTile = kTileOf<T>;
            __shared__ T tile[kTile][kTile + 1];
            typedef typename Wide<T>::type V;
            constexpr int L = Wide<T>::lanes;
            constexpr int per_row = kTile / L;         // threads across one tile row
            constexpr int rows_per_pass = 256 / per_row;
            constexpr int passes = kTile / rows_per_pass;
            const int64_t r0 = int64_t(blockIdx.y) * kTile, c0 = int64_t(blockIdx.x) * kTile;
            const int t = threadIdx.x;
            const int i0 = t / per_row, j = (t % per_row) * L;
            const bool full = vectors && r0 + kTile <= rows && c0 + kTile <= columns;
            if (full) {
                V v[passes]; // every load of the tile in flight before the first LDS write
#pragma unroll
                for (int q = 0; q < passes; ++q)
                    v[q] = *reinterpret_cast<const V *>(in + (r0 + i0 + q * rows_per_pass) * columns + c0 + j);
#pragma unroll
                for (int q = 0; q < passes; ++q)
#pragma unroll
                    for (int k = 0; k < L; ++k) tile[i0 + q * rows_per_pass][j + k] = v[q][k];
            } else {
                for (int e = t; e < kTile * kTile; e += 256) {
                    const int i = e / kTile, jj = e % kTile;
                    if (r0 + i < rows && c0 + jj < columns) tile[i][jj] = in[(r0 + i) * columns + c0 + jj];
                }
            }
            __syncthreads();
            // out[c][r] = in[r][c]: output row c0 + i holds input column c0 + i.
            if (full) {
#pragma unroll
                for (int q = 0; q < passes; ++q) {
                    const int i = i0 + q * rows_per_pass;
                    V v;
#pragma unroll
                    for (int k = 0; k < L; ++k) v[k] = tile[j + k][i];
                    *reinterpret_cast<V *>(out + (c0 + i) * rows + r0 + j) = v;
                }
            } else {
                for (int e = t; e < kTile * kTile; e += 256) {
                    const int i = e / kTile, jj = e % kTile;
                    if (c0 + i < columns && r0 + jj < rows) out[(c0 + i) * rows + r0 + jj] = tile[jj][i];
                }
            }
        }

        template <typename T> __device__ __forceinline__ T result_of(double value, uint32_t poisoned) {
            return poisoned ? T(__builtin_nan("")) : T(value);
        }

        // One workgroup per row: the fold of the chunk's result over that row.
        template <typename T, int W, int R> __global__ __launch_bounds__(256) void fold_rows(const AxisLaunch<T> a) {
            const Program<T> &p = a.p;
            SCIHIP_FILE(T, W, R);
            constants<T, W, R>(p, regs);
            const uint64_t row = blockIdx.x;
            const uint64_t first = row * a.columns, end = first + a.columns;
            const uint32_t length = p.contract ? p.length - 1 : p.length;
            double total = identity(p.fold);
            uint32_t poisoned = 0;
            for (uint64_t at = first + uint64_t(threadIdx.x) * W; at < end; at += 256 * W) {
                const bool full = at + W <= end;
                load<T, W, R>(p, regs, at, full);
                run<T, W, R>(p, regs, length);
                if (p.contract) contract<T, W>(total, regs.get(p.left), regs.get(p.right), at, full, end);
                else consume<T, W>(p.fold, total, poisoned, regs.get(p.result), at, full, end);
            }
            const Partial folded = fold_block(p.fold, total, poisoned);
            if (threadIdx.x == 0) a.output[row] = result_of<T>(folded.value, folded.poisoned);
        }

        // Each thread: W adjacent columns over the rows of slice blockIdx.y; one Partial per column.
        template <typename T, int W, int R>
        __global__ __launch_bounds__(256) void fold_columns(const AxisLaunch<T> a) {
            const Program<T> &p = a.p;
            SCIHIP_FILE(T, W, R);
            constants<T, W, R>(p, regs);
            const uint64_t column = (uint64_t(blockIdx.x) * 256 + threadIdx.x) * W;
            if (column >= a.columns) return;
            const uint64_t per_slice = (a.rows + gridDim.y - 1) / gridDim.y;
            const uint64_t begin = uint64_t(blockIdx.y) * per_slice;
            const uint64_t end = begin + per_slice < a.rows ? begin + per_slice : a.rows;
            const uint32_t length = p.contract ? p.length - 1 : p.length;
            const uint32_t fold = p.fold;
            const bool full = column + W <= a.columns;
            double total[W];
            uint32_t poisoned[W];
#pragma unroll
            for (int k = 0; k < W; ++k) {
                total[k] = identity(fold);
                poisoned[k] = 0;
            }
            for (uint64_t r = begin; r < end; ++r) {
                const uint64_t at = r * a.columns + column;
                load<T, W, R>(p, regs, at, full);
                run<T, W, R>(p, regs, length);
                if (p.contract) {
                    const Lanes<T, W> l = regs.get(p.left), m = regs.get(p.right);
#pragma unroll
                    for (int k = 0; k < W; ++k) total[k] = fused(double(l.v[k]), double(m.v[k]), total[k]);
                } else {
                    const Lanes<T, W> v = regs.get(p.result);
#pragma unroll
                    for (int k = 0; k < W; ++k) fold_in(fold, total[k], poisoned[k], double(v.v[k]));
                }
            }
            Partial *slice = a.partials + uint64_t(blockIdx.y) * a.columns;
#pragma unroll
            for (int k = 0; k < W; ++k)
                if (full || column + k < a.columns) slice[column + k] = Partial{total[k], poisoned[k], 0u};
        }

        // out[c] = the fold of the slices' partials of column c, in slice order; 16 partials' loads are
        // in flight at a time.
        template <typename T>
        __global__ __launch_bounds__(256) void finish_columns(const Partial *partials, T *out, uint64_t columns,
                                                              uint32_t slices, uint32_t fold) {
            constexpr uint32_t kBatch = 16;
            const uint64_t c = uint64_t(blockIdx.x) * 256 + threadIdx.x;
            if (c >= columns) return;
            double value = identity(fold);
            uint32_t flag = 0;
            for (uint32_t s0 = 0; s0 < slices; s0 += kBatch) {
                Partial part[kBatch];
#pragma unroll
                for (uint32_t i = 0; i < kBatch; ++i)
                    if (s0 + i < slices) part[i] = partials[uint64_t(s0 + i) * columns + c];
#pragma unroll
                for (uint32_t i = 0; i < kBatch; ++i)
                    if (s0 + i < slices) {
                        value = combine(fold, value, part[i].value);
                        flag |= part[i].poisoned;
                    }
            }
            out[c] = result_of<T>(value, flag);
        }

        template <typename T> __global__ __launch_bounds__(256) void fill(T *out, uint64_t n, T value) {
            const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
            if (i < n) out[i] = value;
        }

    } // namespace

    template <typename T> hipError_t fill_value(T *out, uint64_t n, T value, hipStream_t stream) {
        if (n == 0) return hipSuccess;
        hipLaunchKernelGGL((fill<T>), dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, out, n, value);
        return hipGetLastError();
    }

    template <typename T>
    hipError_t transpose(const T *in, T *out, int64_t rows, int64_t columns, hipStream_t stream) {
        if (rows == 0 || columns == 0) return hipSuccess;
        constexpr int L = Wide<T>::lanes;
        const bool vectors = rows % L == 0 && columns % L == 0 && (reinterpret_cast<uintptr_t>(in) & 15u) == 0 &&
                             (reinterpret_cast<uintptr_t>(out) & 15u) == 0;
        constexpr int kTile = kTileOf<T>;
        const dim3 grid(unsigned((columns + kTile - 1) / kTile), unsigned((rows + kTile - 1) / kTile));
        hipLaunchKernelGGL((transpose_tiles<T>), grid, dim3(256), 0, stream, in, out, rows, columns, vectors);
        return hipGetLastError();
    }

    template <typename T>
    hipError_t fold_axis(const AxisLaunch<T> &launch, int registers, int axis, uint32_t slices, hipStream_t stream) {
        constexpr int W = Traits<T>::width;
        if (axis == 1) {
            const dim3 grid(unsigned(launch.rows));
            if (registers == kSmall) hipLaunchKernelGGL((fold_rows<T, W, kSmall>), grid, dim3(256), 0, stream, launch);
            else if (registers == kMedium) hipLaunchKernelGGL((fold_rows<T, W, kMedium>), grid, dim3(256), 0, stream, launch);
            else hipLaunchKernelGGL((fold_rows<T, 1, kLarge>), grid, dim3(256), 0, stream, launch);
        } else {
            const int w = registers == kLarge ? 1 : W;
            const dim3 grid(unsigned((launch.columns + 256 * w - 1) / (256 * w)), slices);
            if (registers == kSmall) hipLaunchKernelGGL((fold_columns<T, W, kSmall>), grid, dim3(256), 0, stream, launch);
            else if (registers == kMedium) hipLaunchKernelGGL((fold_columns<T, W, kMedium>), grid, dim3(256), 0, stream, launch);
            else hipLaunchKernelGGL((fold_columns<T, 1, kLarge>), grid, dim3(256), 0, stream, launch);
            if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
            return finish_axis<T>(launch, slices, stream);
        }
        return hipGetLastError();
    }

    template <typename T> hipError_t finish_axis(const AxisLaunch<T> &launch, uint32_t slices, hipStream_t stream) {
        hipLaunchKernelGGL((finish_columns<T>), dim3(unsigned((launch.columns + 255) / 256)), dim3(256), 0, stream,
                           launch.partials, launch.output, launch.columns, slices, launch.p.fold);
        return hipGetLastError();
    }

    template hipError_t transpose<float>(const float *, float *, int64_t, int64_t, hipStream_t);
    template hipError_t transpose<double>(const double *, double *, int64_t, int64_t, hipStream_t);
    template hipError_t fold_axis<float>(const AxisLaunch<float> &, int, int, uint32_t, hipStream_t);
    template hipError_t fold_axis<double>(const AxisLaunch<double> &, int, int, uint32_t, hipStream_t);
    template hipError_t finish_axis<float>(const AxisLaunch<float> &, uint32_t, hipStream_t);
    template hipError_t finish_axis<double>(const AxisLaunch<double> &, uint32_t, hipStream_t);
    template hipError_t fill_value<float>(float *, uint64_t, float, hipStream_t);
    template hipError_t fill_value<double>(double *, uint64_t, double, hipStream_t);

} // namespace scihip
