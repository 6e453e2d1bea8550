// LU with partial pivoting, triangular solves (see lu.hpp).
#include "lu.hpp"
#include "product.hpp"

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace scihip {

    namespace {

        constexpr unsigned kSpinLimit = 1u << 21; // a bounded wait: about two seconds

        __device__ __forceinline__ unsigned relaxed(unsigned *p) {
            return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }

        // Values other workgroups read inside the same kernel travel through relaxed agent-scope
        // atomic stores and loads: those go to the device's point of coherence, so no release
        // (L2 write-back) or acquire (L2 invalidate) fence is needed; a workgroup barrier, which
        // waits for every wave's stores to complete, orders the data before the flag that
        // announces it.
        template <typename T> __device__ __forceinline__ void put(T *p, T v) {
            if constexpr (sizeof(T) == 8)
                __hip_atomic_store(reinterpret_cast<unsigned long long *>(p), __builtin_bit_cast(unsigned long long, v),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else
                __hip_atomic_store(reinterpret_cast<unsigned *>(p), __builtin_bit_cast(unsigned, v), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        template <typename T> __device__ __forceinline__ T get(const T *p) {
            if constexpr (sizeof(T) == 8)
                return __builtin_bit_cast(T, __hip_atomic_load(reinterpret_cast<const unsigned long long *>(p),
                                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            else
                return __builtin_bit_cast(T, __hip_atomic_load(reinterpret_cast<const unsigned *>(p), __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_AGENT));
        }

        // Workgroup waits until *counter >= target (thread 0 polls with relaxed atomic loads; what the
        // flag announces is read with get()). Gives up, and raises *error, past kSpinLimit polls or as
        // soon as another workgroup has given up.
        __device__ __forceinline__ void wait_until(unsigned *counter, unsigned target, unsigned *error) {
            if (threadIdx.x == 0) {
                unsigned spins = 0;
                while (relaxed(counter) < target) {
                    if (++spins > kSpinLimit) {
                        __hip_atomic_fetch_or(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    if ((spins & 63u) == 0 && relaxed(error)) break;
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __syncthreads();
        }

        // The better of two pivot candidates: the larger |a|, and on a tie the lower row (the first
        // maximum, include/linear/decompose.hpp:35-38).
        template <typename T> __device__ __forceinline__ void better(T &v, int64_t &r, T ov, int64_t orow) {
            if (ov > v || (ov == v && orow < r)) {
                v = ov;
                r = orow;
            }
        }

        // One DPP move of a 32-bit value within rows of 16 lanes (every lane reads a lane of its row).
        template <int control> __device__ __forceinline__ unsigned dpp(unsigned x) {
            return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, control, 0xF, 0xF, false);
        }

        template <int control, typename T> __device__ __forceinline__ void dpp_step(T &v, int64_t &r) {
            T ov;
            if constexpr (sizeof(T) == 8) {
                const unsigned long long bits = __builtin_bit_cast(unsigned long long, v);
                const unsigned long long moved = (unsigned long long)dpp<control>((unsigned)(bits >> 32)) << 32 |
                                                 dpp<control>((unsigned)bits);
                ov = __builtin_bit_cast(T, moved);
            } else {
                ov = __builtin_bit_cast(T, dpp<control>(__builtin_bit_cast(unsigned, v)));
            }
            const unsigned long long rb = (unsigned long long)r;
            const int64_t orow = (int64_t)((unsigned long long)dpp<control>((unsigned)(rb >> 32)) << 32 |
                                           dpp<control>((unsigned)rb));
            better(v, r, ov, orow);
        }

        // The best candidate across the wave; every lane ends with it. Within each row of 16 lanes by
        // DPP (lane pairs at distance 1 and 2, then mirrored halves and mirrored rows: each step pairs
        // every lane with one holding a disjoint set of candidates), then across rows by shuffles.
        template <typename T> __device__ __forceinline__ void wave_best(T &v, int64_t &r) {
            dpp_step<0xB1>(v, r);  // quad_perm [1, 0, 3, 2]
            dpp_step<0x4E>(v, r);  // quad_perm [2, 3, 0, 1]
            dpp_step<0x141>(v, r); // row_half_mirror
            dpp_step<0x140>(v, r); // row_mirror
#pragma unroll
            for (int offset = 16; offset <= 32; offset <<= 1) {
                const T ov = __shfl_xor(v, offset, 64);
                const int64_t orow = __shfl_xor(r, offset, 64);
                better(v, r, ov, orow);
            }
        }

        // The largest 64-bit value across the wave; every lane ends with it.
        __device__ __forceinline__ unsigned long long wave_max(unsigned long long x) {
            auto step = [&](unsigned long long o) { x = o > x ? o : x; };
            step((unsigned long long)dpp<0xB1>((unsigned)(x >> 32)) << 32 | dpp<0xB1>((unsigned)x));
            step((unsigned long long)dpp<0x4E>((unsigned)(x >> 32)) << 32 | dpp<0x4E>((unsigned)x));
            step((unsigned long long)dpp<0x141>((unsigned)(x >> 32)) << 32 | dpp<0x141>((unsigned)x));
            step((unsigned long long)dpp<0x140>((unsigned)(x >> 32)) << 32 | dpp<0x140>((unsigned)x));
            step(__shfl_xor(x, 16, 64));
            step(__shfl_xor(x, 32, 64));
            return x;
        }

        // A workgroup's candidate for one column as one 64-bit word, which is also its arrival flag:
        // bits 63-32 the float bits of |a| (monotonic in |a| for a non-negative value; exact for float
        // matrices, rounded for double ones), bits 31-8 the complement of the row's offset from the
        // panel's first row (so that the larger word has the lower row on equal keys), bits 7-0 the
        // column number + 1. A workgroup without a candidate publishes key 0 and the largest offset.
        template <typename T>
        __device__ __forceinline__ unsigned long long candidate_word(T best, int64_t row, int64_t j0, unsigned epoch) {
            unsigned long long word = epoch & 0xFFu;
            if (row != LLONG_MAX) {
                const float key = (float)best;
                const unsigned bits = key == key ? __builtin_bit_cast(unsigned, key) : 0u; // NaN never wins
                word |= (unsigned long long)bits << 32 | (unsigned long long)(0xFFFFFFu - (unsigned)(row - j0)) << 8;
            }
            return word;
        }
        __device__ __forceinline__ int64_t row_of(unsigned long long word, int64_t j0) {
            return j0 + (int64_t)(0xFFFFFFu - (unsigned)((word >> 8) & 0xFFFFFFu));
        }

        constexpr int kLine = 16;          // 64-bit words per 128-byte line: one line per candidate word
        constexpr int kPanelThreads = 64;  // one wave per panel workgroup: its barriers cost nothing
        constexpr int kOwned = (kPanelRows + kPanelThreads - 1) / kPanelThreads; // panel rows per lane, at most

        // Factors columns [j0, j0 + w) of rows [j0, n) (see lu.hpp, step 1). Workgroup b is one wave
        // and owns rows [j0 + b * kPanelRows, ...); it keeps their panel part in LDS for the whole
        // launch, lane l searching rows l, l + 64, ... Per column, every workgroup publishes its candidate
        // (first maximum of |a| among its rows at or below the diagonal) with that row's panel values,
        // then its candidate word (candidate_word), which is its arrival flag and carries the candidate,
        // so one round trip gathers every candidate; every workgroup takes the largest word, swaps and
        // eliminates; the next column's candidate is published before the bulk of the elimination,
        // which then overlaps the words' round trip. Publications have two parities (column c uses
        // c & 1): a workgroup overwrites parity p only after every word of the column in between has
        // arrived, that is after every workgroup has finished reading parity p. value[2][G] (exact
        // |a|, for ties of the float key), rows[2][G][kPanel], diagonal[2][kPanel], words[2][G] (one
        // 128-byte line each).
        template <typename T>
        __global__ __launch_bounds__(kPanelThreads) void panel(T *a, int64_t n, int64_t j0, int w, int64_t *pivots,
                                                                T *pub_value, T *pub_rows, T *diagonal,
                                                                unsigned long long *words, unsigned *error) {
            constexpr int stride = kPanel + 1;
            const int G = gridDim.x, b = blockIdx.x, t = threadIdx.x;
            const int64_t first = j0 + (int64_t)b * kPanelRows;
            const int mine = (int)min((int64_t)kPanelRows, n - first); // rows this workgroup owns
            __shared__ T rows[kPanelRows * stride];
            __shared__ T pivot_row[kPanel];
            __shared__ T diagonal_row[kPanel];
            __shared__ T multipliers[kPanelRows];

            for (int i = 0; i < mine; ++i)
                for (int k = t; k < w; k += kPanelThreads) rows[i * stride + k] = a[(first + i) * n + j0 + k];
            __syncthreads();

            // Publishes this workgroup's candidate for column `col` (the first maximum of |a| among its
            // rows at or below the diagonal) with the candidate's panel row, and the diagonal row if it
            // owns it; then raises its arrival flag for `col`. Column `col` of the rows is up to date;
            // with `pending`, their columns beyond it still lack the previous column's update, which
            // is applied to the published values (x - l * u, as the elimination will store it).
            auto publish = [&](int col, bool pending) {
                const int64_t dd = j0 + col;
                const int par = col & 1;
                auto value = [&](int i, int k) -> T {
                    const T x = rows[i * stride + k];
                    return (pending && k > col) ? x - multipliers[i] * pivot_row[k] : x;
                };
                T best = T(-1);
                int64_t row = LLONG_MAX;
#pragma unroll
                for (int o = 0; o < kOwned; ++o) {
                    const int i = t + o * kPanelThreads;
                    if (i < mine && first + i >= dd) better(best, row, T(__builtin_fabs(rows[i * stride + col])), first + i);
                }
                wave_best(best, row);
                if (t == 0) put(&pub_value[par * G + b], best); // the exact value, for ties of the float key
                if (row != LLONG_MAX)
                    for (int k = t; k < w; k += kPanelThreads)
                        put(&pub_rows[((int64_t)par * G + b) * kPanel + k], value((int)(row - first), k));
                if (dd >= first && dd < first + mine)
                    for (int k = t; k < w; k += kPanelThreads) put(&diagonal[par * kPanel + k], value((int)(dd - first), k));
                __syncthreads(); // every publication has completed before the word that announces it
                if (t == 0)
                    __hip_atomic_store(&words[((int64_t)par * G + b) * kLine],
                                       candidate_word<T>(best, row, j0, (unsigned)col + 1), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            };

            publish(0, false);
            for (int c = 0; c < w; ++c) {
                const int64_t d = j0 + c;
                const int parity = c & 1;
                const unsigned epoch = (unsigned)c + 1;
                // A: every workgroup gathers: lane g polls workgroup g's candidate word until it carries
                // this column, and the wave takes the largest word, so all workgroups pick the same pivot.
                // Equal float keys of different doubles are settled with the exact values.
                unsigned long long mine_best = 0; // this lane's largest word, and how many share its key
                int sharing = 0;
                for (int g = t; g < G; g += kPanelThreads) {
                    unsigned long long *word = &words[((int64_t)parity * G + g) * kLine];
                    unsigned spins = 0;
                    unsigned long long x;
                    while (((x = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) & 0xFFu) !=
                           (epoch & 0xFFu)) {
                        if (++spins > kSpinLimit) {
                            __hip_atomic_fetch_or(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            break;
                        }
                        if ((spins & 63u) == 0 && relaxed(error)) break;
                        __builtin_amdgcn_s_sleep(1);
                    }
                    if (sharing == 0 || (x >> 32) > (mine_best >> 32)) {
                        mine_best = x;
                        sharing = 1;
                    } else if ((x >> 32) == (mine_best >> 32)) {
                        if ((x >> 8) > (mine_best >> 8)) mine_best = x;
                        ++sharing;
                    }
                }
                const unsigned long long top = wave_max(sharing ? mine_best >> 8 : 0); // key << 24 | ~offset
                int64_t p = row_of(top << 8, j0);
                if constexpr (sizeof(T) == 8) {
                    const bool holds = sharing > 0 && (mine_best >> 32) == (top >> 24);
                    const unsigned long long holders = __ballot(holds);
                    if (__ballot(holds && sharing > 1) || (holders & (holders - 1))) {
                        // Two or more candidates share the float key: compare the exact values. The words
                        // of this parity are not rewritten before this workgroup publishes again.
                        T v = T(-1);
                        p = LLONG_MAX;
                        for (int g = t; g < G; g += kPanelThreads) {
                            const unsigned long long x = __hip_atomic_load(&words[((int64_t)parity * G + g) * kLine],
                                                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            if ((x >> 32) == (top >> 24) && (x & 0xFFu) == (epoch & 0xFFu))
                                better(v, p, get(&pub_value[parity * G + g]), row_of(x, j0));
                        }
                        wave_best(v, p);
                    }
                }
                if (p < d || p >= n) p = d; // only after a failed wait: keep every index in range

                // B: swap.
                const int owner = (int)((p - j0) / kPanelRows); // the workgroup that published row p
                for (int k = t; k < w; k += kPanelThreads) {
                    pivot_row[k] = get(&pub_rows[((int64_t)parity * G + owner) * kPanel + k]);
                    diagonal_row[k] = get(&diagonal[parity * kPanel + k]);
                }
                __syncthreads();
                if (p != d) {
                    if (p >= first && p < first + mine)
                        for (int k = t; k < w; k += kPanelThreads) rows[(p - first) * stride + k] = diagonal_row[k];
                    if (d >= first && d < first + mine)
                        for (int k = t; k < w; k += kPanelThreads) rows[(d - first) * stride + k] = pivot_row[k];
                }
                if (b == 0 && t == 0) pivots[d] = p;
                __syncthreads();

                // C: multipliers; column c + 1 and the next candidate first (look-ahead), then the rest of
                // the rank-one update while the other workgroups' candidates travel.
                const T pivot = pivot_row[c];
                const bool eliminate = pivot != T(0); // a zero pivot skips the column (decompose.hpp:44-45)
                if (eliminate) {
#pragma unroll
                    for (int o = 0; o < kOwned; ++o) {
                        const int i = t + o * kPanelThreads;
                        T l = T(0);
                        if (i < mine && first + i > d) {
                            l = rows[i * stride + c] / pivot;
                            rows[i * stride + c] = l;
                            if (c + 1 < w) rows[i * stride + c + 1] = rows[i * stride + c + 1] - l * pivot_row[c + 1];
                        }
                        if (i < kPanelRows) multipliers[i] = l;
                    }
                    __syncthreads();
                }
                if (c + 1 < w) publish(c + 1, eliminate);
                if (eliminate) {
                    // Lane t updates columns c + 2 + t, c + 66 + t of every row below the diagonal; the
                    // iterations are independent (eight LDS loads in flight before any store).
                    const int lo = (int)max((int64_t)0, d + 1 - first);
                    for (int k = c + 2 + t; k < w; k += kPanelThreads) {
                        const T u = pivot_row[k];
                        T *column = rows + k;
                        int i = lo;
                        for (; i + 8 <= mine; i += 8) {
                            T x[8], m[8];
#pragma unroll
                            for (int j = 0; j < 8; ++j) {
                                x[j] = column[(i + j) * stride];
                                m[j] = multipliers[i + j];
                            }
#pragma unroll
                            for (int j = 0; j < 8; ++j) column[(i + j) * stride] = x[j] - m[j] * u;
                        }
                        for (; i < mine; ++i) column[i * stride] = column[i * stride] - multipliers[i] * u;
                    }
                }
                __syncthreads();
            }

            for (int i = 0; i < mine; ++i)
                for (int k = t; k < w; k += kPanelThreads) a[(first + i) * n + j0 + k] = rows[i * stride + k];
        }

        // Applies the panel's swaps to every column outside it, in step order.
        template <typename T>
        __global__ __launch_bounds__(256) void swap_rows(T *a, int64_t n, int64_t j0, int w, const int64_t *pivots) {
            const int64_t raw = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
            if (raw >= n - w) return;
            const int64_t col = raw < j0 ? raw : raw + w;
            for (int k = 0; k < w; ++k) {
                const int64_t d = j0 + k, p = pivots[d];
                if (p != d) {
                    const T x = a[d * n + col];
                    a[d * n + col] = a[p * n + col];
                    a[p * n + col] = x;
                }
            }
        }

        // Sync-free blocked substitution over columns [col0, col0 + kSolveColumns) of x (order x width):
        // workgroup q solves block q (lower, unit diagonal) or block nb - 1 - q (upper), after folding in
        // every block it depends on as soon as that block's ready flag reaches `epoch`.
        template <typename T, bool upper>
        __global__ __launch_bounds__(256) void solve_blocks(const T *f, int64_t n, T *x, int64_t width, int64_t col0,
                                                            unsigned *flags, unsigned epoch, unsigned *error) {
            const int nb = gridDim.x, q = blockIdx.x, t = threadIdx.x;
            const int blk = upper ? nb - 1 - q : q;
            const int64_t r0 = (int64_t)blk * kSolveRows;
            const int rows = (int)min((int64_t)kSolveRows, n - r0);
            const int cols = (int)min((int64_t)kSolveColumns, width - col0);
            __shared__ T rhs[kSolveRows][kSolveColumns + 1];
            __shared__ T tile[kSolveRows][kSolveRows + 1];
            __shared__ T known[kSolveRows][kSolveColumns + 1];

            for (int idx = t; idx < kSolveRows * kSolveColumns; idx += blockDim.x) {
                const int i = idx / kSolveColumns, c = idx % kSolveColumns;
                rhs[i][c] = (i < rows && c < cols) ? x[(r0 + i) * width + col0 + c] : T(0);
            }
            for (int qq = 0; qq < q; ++qq) {
                const int other = upper ? nb - 1 - qq : qq;
                wait_until(&flags[other], epoch, error);
                const int64_t o0 = (int64_t)other * kSolveRows;
                const int orows = (int)min((int64_t)kSolveRows, n - o0);
                for (int idx = t; idx < kSolveRows * kSolveRows; idx += blockDim.x) {
                    const int i = idx / kSolveRows, k = idx % kSolveRows;
                    tile[i][k] = (i < rows && k < orows) ? f[(r0 + i) * n + o0 + k] : T(0);
                }
                for (int idx = t; idx < kSolveRows * kSolveColumns; idx += blockDim.x) {
                    const int k = idx / kSolveColumns, c = idx % kSolveColumns;
                    known[k][c] = (k < orows && c < cols) ? get(&x[(o0 + k) * width + col0 + c]) : T(0);
                }
                __syncthreads();
                const int c = t % kSolveColumns;
                for (int i = t / kSolveColumns; i < kSolveRows; i += blockDim.x / kSolveColumns) {
                    T s = rhs[i][c];
                    for (int k = 0; k < kSolveRows; ++k) s = s - tile[i][k] * known[k][c];
                    rhs[i][c] = s;
                }
                __syncthreads();
            }
            // The diagonal block: kSolveColumns columns, eight threads per column splitting the sum.
            for (int idx = t; idx < kSolveRows * kSolveRows; idx += blockDim.x) {
                const int i = idx / kSolveRows, k = idx % kSolveRows;
                tile[i][k] = (i < rows && k < rows) ? f[(r0 + i) * n + r0 + k] : T(0);
            }
            __syncthreads();
            const int c = t >> 3, s = t & 7;
            for (int step = 0; step < rows; ++step) {
                const int i = upper ? rows - 1 - step : step;
                T partial = T(0);
                if (upper) {
                    for (int k = i + 1 + s; k < rows; k += 8) partial = partial + tile[i][k] * rhs[k][c];
                } else {
                    for (int k = s; k < i; k += 8) partial = partial + tile[i][k] * rhs[k][c];
                }
                partial = partial + __shfl_xor(partial, 1, 64);
                partial = partial + __shfl_xor(partial, 2, 64);
                partial = partial + __shfl_xor(partial, 4, 64);
                if (s == 0) {
                    const T y = rhs[i][c] - partial;
                    rhs[i][c] = upper ? y / tile[i][i] : y;
                }
                __syncthreads();
            }
            for (int idx = t; idx < rows * cols; idx += blockDim.x) {
                const int i = idx / cols, cc = idx % cols;
                put(&x[(r0 + i) * width + col0 + cc], rhs[i][cc]);
            }
            __syncthreads();
            if (t == 0) __hip_atomic_store(&flags[blk], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }

        // Inverts each kBlock x kBlock diagonal block of the unit lower L (upper false) or of U (upper
        // true) in LDS, one workgroup per block, and stores it transposed: out[b][k][r] = inv(block
        // b)[r][k]. Lower: M L = I gives, for i > j, M[i][j] = -(L[i][j] + sum_{j<k<i} M[i][k] L[k][j]),
        // columns from the last; upper: N U = I gives N[j][j] = 1 / U[j][j] and, for i < j,
        // N[i][j] = -(sum_{i<=k<j} N[i][k] U[k][j]) / U[j][j], columns from the first. Column j is
        // overwritten only after every thread has read it.
        template <typename T, bool upper>
        __global__ __launch_bounds__(256) void invert_diagonal(const T *f, int64_t n, int size, T *out) {
            constexpr int stride = kBlock + 1;
            __shared__ T m[kBlock * stride];
            const int64_t r0 = (int64_t)blockIdx.x * kBlock;
            const int t = threadIdx.x;
            // A block of `size` < kBlock rows is completed by the identity, whose inverse it keeps.
            for (int idx = t; idx < kBlock * kBlock; idx += blockDim.x) {
                const int i = idx / kBlock, k = idx % kBlock;
                m[i * stride + k] = (i < size && k < size) ? f[(r0 + i) * n + r0 + k] : (i == k ? T(1) : T(0));
            }
            __syncthreads();
            const int i = t; // threads below kBlock each carry one row of the inverse
            for (int step = 0; step < kBlock; ++step) {
                const int j = upper ? step : kBlock - 1 - step;
                T v = T(0);
                const bool active = i < kBlock && (upper ? i <= j : i > j);
                if (active) {
                    if (upper) {
                        if (i == j) {
                            v = T(1) / m[j * stride + j];
                        } else {
                            T sum = T(0);
                            for (int k = i; k < j; ++k) sum = __builtin_fma(m[i * stride + k], m[k * stride + j], sum);
                            v = -sum / m[j * stride + j];
                        }
                    } else {
                        T sum = m[i * stride + j];
                        for (int k = j + 1; k < i; ++k) sum = __builtin_fma(m[i * stride + k], m[k * stride + j], sum);
                        v = -sum;
                    }
                }
                __syncthreads();
                if (active) m[i * stride + j] = v;
                __syncthreads();
            }
            T *block = out + (int64_t)blockIdx.x * kBlock * kBlock;
            for (int idx = t; idx < kBlock * kBlock; idx += blockDim.x) {
                const int r = idx % kBlock, k = idx / kBlock; // consecutive threads write consecutive r
                const T value = upper ? (k >= r ? m[r * stride + k] : T(0))
                                      : (k < r ? m[r * stride + k] : (k == r ? T(1) : T(0)));
                block[(int64_t)k * kBlock + r] = value;
            }
        }

        // x[r0 + r][c] = sum_k inv[r][k] x[r0 + k][c] over one diagonal block (inverse given transposed,
        // inverse_t[k][r]); one workgroup of kBlock threads per column c. Also writes the packed B
        // operand of the update that follows, packed[c][r] (width x kBlock).
        template <typename T>
        __global__ __launch_bounds__(kBlock) void apply_inverse(const T *inverse_t, T *x, int64_t r0, int64_t width,
                                                                T *packed) {
            __shared__ T column[kBlock];
            const int64_t c = blockIdx.x;
            const int r = threadIdx.x;
            column[r] = x[(r0 + r) * width + c];
            __syncthreads();
            T sum = T(0);
#pragma unroll 16
            for (int k = 0; k < kBlock; ++k) sum = __builtin_fma(inverse_t[k * kBlock + r], column[k], sum);
            x[(r0 + r) * width + c] = sum;
            packed[c * kBlock + r] = sum;
        }

        template <typename T, typename S>
        __global__ __launch_bounds__(256) void permute_rows(const S *source, T *target, const int64_t *rows,
                                                            int64_t order, int64_t width) {
            const int64_t total = order * width;
            for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
                 idx += (int64_t)gridDim.x * blockDim.x) {
                const int64_t k = idx / width, c = idx % width;
                target[idx] = (T)source[rows[k] * width + c];
            }
        }

        const char *kSpinFailure = "a bounded wait in the LU kernels gave up";

        template <typename T> struct PackMode;
        template <> struct PackMode<double> {
            static bool wide() { return true; }
        };
        template <> struct PackMode<float> {
            static bool wide() { return false; }
        };

    } // namespace

    size_t factor_scratch(int64_t order, int mode) {
        (void)mode;
        const size_t kp = (size_t)kPanel;
        // L21 and U12 packed (up to 8 bytes, split pairs), L11^-1 and its packed copy, the panel's
        // candidate words (two parities, one 128-byte line each).
        return 4 * (size_t)order * kp * 8 + 2 * kp * kp * 8 + 2 * ((size_t)order / kPanelRows + 2) * 128 + 1024;
    }

    namespace {

        // SCIHIP_TRACE=lu: HIP events around every stage of the factorisation, totals on stderr.
        struct StageTimer {
            bool on = false;
            hipStream_t stream;
            hipEvent_t marks[2] = {nullptr, nullptr};
            double totals[6] = {0, 0, 0, 0, 0, 0};
            explicit StageTimer(hipStream_t s) : stream(s) {
                const char *text = std::getenv("SCIHIP_TRACE");
                on = text && std::strstr(text, "lu");
                if (on) {
                    (void)hipEventCreate(&marks[0]);
                    (void)hipEventCreate(&marks[1]);
                    (void)hipEventRecord(marks[0], stream);
                }
            }
            void stage(int index) {
                if (!on) return;
                (void)hipEventRecord(marks[1], stream);
                (void)hipEventSynchronize(marks[1]);
                float ms = 0;
                (void)hipEventElapsedTime(&ms, marks[0], marks[1]);
                totals[index] += ms;
                std::swap(marks[0], marks[1]);
            }
            ~StageTimer() {
                if (!on) return;
                std::fprintf(stderr,
                             "[scihip] lu stages ms: panel (not hidden) %.1f swap %.1f u12 %.1f pack %.1f update %.1f\n",
                             totals[0], totals[1], totals[2], totals[3], totals[4]);
                (void)hipEventDestroy(marks[0]);
                (void)hipEventDestroy(marks[1]);
            }
        };

    } // namespace

    template <typename T>
    int factor(T *a, int64_t n, int64_t *pivots, int mode, Workspace &ws, void *scratch, size_t scratch_bytes,
               hipStream_t stream, const char **error) {
        *error = nullptr;
        if (scratch_bytes < factor_scratch(n, mode)) {
            *error = "LU scratch too small";
            return -1;
        }
        const bool wide = PackMode<T>::wide();
        const int own = wide ? kFloat64 : kFloat32; // U12 = L11^-1 A12 runs in the matrix's precision
        const size_t kp = kPanel;
        unsigned char *base = static_cast<unsigned char *>(scratch);
        const size_t operand = (size_t)n * kp * 8; // bytes of one packed n x kPanel operand, at most
        void *la_hi = base, *la_lo = base + operand;
        void *ub_hi = base + 2 * operand, *ub_lo = base + 3 * operand;
        T *inverse = reinterpret_cast<T *>(base + 4 * operand);
        void *inverse_packed = base + 4 * operand + kp * kp * 8;
        unsigned *flags = reinterpret_cast<unsigned *>(base + 4 * operand + 2 * kp * kp * 8);
        const int max_groups = (int)((n + kPanelRows - 1) / kPanelRows);
        T *pub_value = static_cast<T *>(ws.publish);
        T *pub_rows = pub_value + 2 * max_groups + 2;
        T *diagonal = pub_rows + (size_t)2 * max_groups * kPanel;
        unsigned long long *words = reinterpret_cast<unsigned long long *>(flags);

        StageTimer timer(stream);
        // Factors the panel at column j0 on stream `s`.
        auto launch_panel = [&](int64_t j0, hipStream_t s) -> hipError_t {
            int w = (int)std::min<int64_t>(kPanel, n - j0);
            const int groups = (int)((n - j0 + kPanelRows - 1) / kPanelRows);
            hipError_t e = hipMemsetAsync(words, 0, (size_t)2 * (groups + 1) * kLine * sizeof(unsigned long long), s);
            if (e != hipSuccess) return e;
            void *args[] = {&a, &n, (void *)&j0, (void *)&w, &pivots, &pub_value, &pub_rows, &diagonal, &words,
                            &ws.error};
            e = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&panel<T>), dim3(groups),
                                           dim3(kPanelThreads), args, 0, s);
            if (e != hipSuccess) return e;
            // L11^-1, transposed (inverse[k][r] = L11^-1[r][k]), for U12 = L11^-1 A12.
            hipLaunchKernelGGL((invert_diagonal<T, false>), dim3(1), dim3(256), 0, s, a + j0 * n + j0, n, w, inverse);
            return hipGetLastError();
        };
        const size_t eb = (size_t)element_bytes(mode);
        hipError_t e = launch_panel(0, stream);
        if (e != hipSuccess) return (int)e;
        for (int64_t j0 = 0; j0 < n; j0 += kPanel) {
            const int w = (int)std::min<int64_t>(kPanel, n - j0);
            if (j0 > 0) { // this panel ran on the side stream, during the previous update
                e = hipStreamWaitEvent(stream, ws.panel_done, 0);
                if (e != hipSuccess) return (int)e;
            }
            timer.stage(0);
            if (n > w)
                hipLaunchKernelGGL(swap_rows<T>, dim3((unsigned)((n - w + 255) / 256)), dim3(256), 0, stream, a, n, j0,
                                   w, (const int64_t *)pivots);
            timer.stage(1);
            const int64_t rest = n - j0 - w;
            if (rest > 0) {
                // U12 = L11^-1 A12: the packed copy of A12 is the product's operand, so the result can
                // overwrite A12 in place.
                e = pack(own, a + j0 * n + j0 + w, wide, rest, w, n, true, ub_hi, nullptr, kp, stream);
                if (e == hipSuccess) e = pack(own, inverse, wide, w, w, kPanel, true, inverse_packed, nullptr, kp, stream);
                if (e != hipSuccess) return (int)e;
                ProductArgs solve{inverse_packed, nullptr, ub_hi, nullptr, a + j0 * n + j0 + w, w, rest, (int64_t)kp,
                                  (int64_t)kp, (int64_t)kp, n, 1.0, 0.0};
                e = product(own, solve, stream);
                if (e != hipSuccess) return (int)e;
                timer.stage(2);
                // A22 -= L21 U12 on the matrix cores, in the requested mode: the next panel's columns
                // first, then (while that panel is factored on the side stream) the rest.
                e = pack(mode, a + (j0 + w) * n + j0, wide, rest, w, n, false, la_hi, la_lo, kp, stream);
                if (e == hipSuccess) e = pack(mode, a + j0 * n + j0 + w, wide, rest, w, n, true, ub_hi, ub_lo, kp, stream);
                if (e != hipSuccess) return (int)e;
                timer.stage(3);
                const int64_t next = std::min<int64_t>(kPanel, rest);
                T *c = a + (j0 + w) * n + (j0 + w);
                ProductArgs narrow{la_hi, la_lo, ub_hi, ub_lo, c, rest, next, (int64_t)kp, (int64_t)kp, (int64_t)kp, n,
                                   -1.0, 1.0};
                e = product(mode, narrow, stream);
                if (e == hipSuccess) e = hipEventRecord(ws.columns_ready, stream);
                if (e == hipSuccess) e = hipStreamWaitEvent(ws.side, ws.columns_ready, 0);
                if (e == hipSuccess) e = launch_panel(j0 + w, ws.side);
                if (e == hipSuccess) e = hipEventRecord(ws.panel_done, ws.side);
                if (e != hipSuccess) return (int)e;
                if (rest > next) {
                    const unsigned char *b_hi = static_cast<const unsigned char *>(ub_hi) + (size_t)next * kp * eb;
                    const unsigned char *b_lo =
                            ub_lo ? static_cast<const unsigned char *>(ub_lo) + (size_t)next * kp * eb : nullptr;
                    ProductArgs wide_update{la_hi, la_lo, b_hi, mode == kSplit ? b_lo : nullptr, c + next, rest,
                                            rest - next, (int64_t)kp, (int64_t)kp, (int64_t)kp, n, -1.0, 1.0};
                    e = product(mode, wide_update, stream);
                    if (e != hipSuccess) return (int)e;
                }
                timer.stage(4);
            }
        }
        return hipGetLastError();
    }

    template <typename T> hipError_t invert_blocks(const T *f, int64_t n, T *inverses, hipStream_t stream) {
        if (!blocked_substitution(n)) return hipErrorInvalidValue;
        const unsigned blocks = (unsigned)(n / kBlock);
        hipLaunchKernelGGL((invert_diagonal<T, false>), dim3(blocks), dim3(256), 0, stream, f, n, kBlock, inverses);
        hipLaunchKernelGGL((invert_diagonal<T, true>), dim3(blocks), dim3(256), 0, stream, f, n, kBlock,
                           inverses + n * kBlock);
        return hipGetLastError();
    }

    template <typename T>
    int substitute(const T *f, int64_t n, T *x, int64_t width, Workspace &ws, unsigned &epoch, hipStream_t stream,
                   const char **error, const T *inverses, T *packed) {
        *error = nullptr;
        if (inverses && packed && blocked_substitution(n)) {
            // Block by block: x_j = inv(L_jj) x_j, then the rows below lose L[below, j] x_j (forward);
            // x_j = inv(U_jj) x_j, then the rows above lose U[above, j] x_j (backward). The updates
            // read L and U in place (row stride n) on the matrix cores.
            const int mode = sizeof(T) == 8 ? kFloat64 : kFloat32;
            const int64_t nb = n / kBlock;
            const T *inverse_lower = inverses, *inverse_upper = inverses + n * kBlock;
            for (int pass = 0; pass < 2; ++pass) {
                const bool upper = pass == 1;
                for (int64_t step = 0; step < nb; ++step) {
                    const int64_t j = upper ? nb - 1 - step : step;
                    const int64_t r0 = j * kBlock;
                    hipLaunchKernelGGL(apply_inverse<T>, dim3((unsigned)width), dim3(kBlock), 0, stream,
                                       (upper ? inverse_upper : inverse_lower) + j * kBlock * kBlock, x, r0, width, packed);
                    const int64_t rows = upper ? r0 : n - r0 - kBlock;
                    if (rows == 0) continue;
                    const T *a_block = upper ? f + r0 : f + (r0 + kBlock) * n + r0;
                    T *c_block = upper ? x : x + (r0 + kBlock) * width;
                    ProductArgs update{a_block, nullptr, packed, nullptr, c_block, rows, width, (int64_t)kBlock,
                                       n, (int64_t)kBlock, width, -1.0, 1.0};
                    const hipError_t e = product(mode, update, stream);
                    if (e != hipSuccess) return (int)e;
                }
            }
            return hipGetLastError();
        }
        const int blocks = (int)((n + kSolveRows - 1) / kSolveRows);
        if ((size_t)blocks > ws.flag_count) {
            *error = "solve flags too few";
            return -1;
        }
        for (int64_t col0 = 0; col0 < width; col0 += kSolveColumns) {
            for (int pass = 0; pass < 2; ++pass) {
                const unsigned e = ++epoch; // every launch waits on values no earlier launch has written
                void *args[] = {(void *)&f, &n, &x, &width, &col0, &ws.flags, (void *)&e, &ws.error};
                const void *kernel = pass == 0 ? reinterpret_cast<const void *>(&solve_blocks<T, false>)
                                               : reinterpret_cast<const void *>(&solve_blocks<T, true>);
                const hipError_t err = hipLaunchCooperativeKernel(kernel, dim3(blocks), dim3(256), args, 0, stream);
                if (err != hipSuccess) return (int)err;
            }
        }
        return hipGetLastError();
    }

    template <typename T, typename S>
    hipError_t permute(const S *source, T *target, const int64_t *rows, int64_t order, int64_t width,
                       hipStream_t stream) {
        const int64_t total = order * width;
        const unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 4096);
        if (total == 0) return hipSuccess;
        hipLaunchKernelGGL((permute_rows<T, S>), dim3(blocks), dim3(256), 0, stream, source, target, rows, order, width);
        return hipGetLastError();
    }

    template int factor<double>(double *, int64_t, int64_t *, int, Workspace &, void *, size_t, hipStream_t,
                                const char **);
    template int factor<float>(float *, int64_t, int64_t *, int, Workspace &, void *, size_t, hipStream_t,
                               const char **);
    template int substitute<double>(const double *, int64_t, double *, int64_t, Workspace &, unsigned &, hipStream_t,
                                    const char **, const double *, double *);
    template int substitute<float>(const float *, int64_t, float *, int64_t, Workspace &, unsigned &, hipStream_t,
                                   const char **, const float *, float *);
    template hipError_t invert_blocks<double>(const double *, int64_t, double *, hipStream_t);
    template hipError_t invert_blocks<float>(const float *, int64_t, float *, hipStream_t);
    template hipError_t permute<double, double>(const double *, double *, const int64_t *, int64_t, int64_t,
                                                hipStream_t);
    template hipError_t permute<float, double>(const double *, float *, const int64_t *, int64_t, int64_t, hipStream_t);
    template hipError_t permute<float, float>(const float *, float *, const int64_t *, int64_t, int64_t, hipStream_t);

    const char *spin_failure() { return kSpinFailure; }

} // namespace scihip
