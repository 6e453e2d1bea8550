// LU with partial pivoting, triangular solves (see lu.hpp).
#include "lu.hpp"
#include "product.hpp"

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>

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

        // The largest 64-bit value across the wave, in every lane (the wave's lanes all active): within
        // each row of 16 lanes by DPP, then the four rows' results read with v_readlane (cross-row
        // shuffles go through the LDS unit, each a round trip).
        __device__ __forceinline__ unsigned long long wave_max_rows(unsigned long long x) {
            auto step = [&](unsigned long long o) { x = o > x ? o : x; };
            step((unsigned long long)dpp<0xB1>((unsigned)(x >> 32)) << 32 | dpp<0xB1>((unsigned)x));
            step((unsigned long long)dpp<0x4E>((unsigned)(x >> 32)) << 32 | dpp<0x4E>((unsigned)x));
            step((unsigned long long)dpp<0x141>((unsigned)(x >> 32)) << 32 | dpp<0x141>((unsigned)x));
            step((unsigned long long)dpp<0x140>((unsigned)(x >> 32)) << 32 | dpp<0x140>((unsigned)x));
            unsigned long long best = 0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const unsigned long long v =
                        (unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(x >> 32), 16 * r) << 32 |
                        (unsigned)__builtin_amdgcn_readlane((int)x, 16 * r);
                best = v > best ? v : best;
            }
            return best;
        }
        // The sum across the wave, in every lane (the wave's lanes all active), the same way.
        __device__ __forceinline__ int wave_sum_rows(int x) {
            x += (int)dpp<0xB1>((unsigned)x);
            x += (int)dpp<0x4E>((unsigned)x);
            x += (int)dpp<0x141>((unsigned)x);
            x += (int)dpp<0x140>((unsigned)x);
            return __builtin_amdgcn_readlane(x, 0) + __builtin_amdgcn_readlane(x, 16) + __builtin_amdgcn_readlane(x, 32) +
                   __builtin_amdgcn_readlane(x, 48);
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

        __device__ __forceinline__ bool has_candidate(unsigned long long word) { return ((word >> 8) & 0xFFFFFFu) != 0; }

        constexpr int kLine = 16;          // 64-bit slots per 128-byte line: one line per workgroup and column
        constexpr int kParities = 4;       // publication buffers, used in turn by consecutive columns
        constexpr int kPanelThreads = 256; // four waves per panel workgroup
        constexpr int kWaves = kPanelThreads / 64;
        constexpr int kRowWaves = kPanelRows / 64;          // the waves whose threads own a row each
        constexpr int kParts = kPanelThreads / kPanelRows;  // threads sharing a row in the bulk update
        constexpr int kPoll = 2;           // lines a thread polls at once
        static_assert(kPanelRows % 64 == 0 && kPanelThreads % kPanelRows == 0 && kPanelThreads == 2 * kPanel,
                      "panel layout: a thread per row, two threads per panel column");

        // Values that panel workgroups read from each other travel in 64-bit slots: the upper half
        // holds 32 bits of the value, the lower half a tag (launch << 8 | column + 1). Each slot is
        // stored with one 64-bit atomic store and read whole (a 64-bit atomic load, or one half of a
        // 16-byte load), so a reader that finds the expected tag in every slot of a value holds the
        // current value, in whatever order the writer's stores land.
        template <typename T> constexpr int kSlots = (int)(sizeof(T) / 4);

        template <typename T> __device__ __forceinline__ void put_tagged(unsigned long long *slots, T v, unsigned tag) {
            auto half = [&](unsigned long long *slot, unsigned bits) {
                __hip_atomic_store(slot, (unsigned long long)bits << 32 | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            };
            if constexpr (sizeof(T) == 8) {
                const unsigned long long bits = __builtin_bit_cast(unsigned long long, v);
                half(slots, (unsigned)(bits >> 32));
                half(slots + 1, (unsigned)bits);
            } else {
                half(slots, __builtin_bit_cast(unsigned, v));
            }
        }
        template <typename T> __device__ __forceinline__ bool current(const unsigned long long *slot, unsigned tag) {
            bool ok = true;
#pragma unroll
            for (int i = 0; i < kSlots<T>; ++i) ok = ok && (unsigned)slot[i] == tag;
            return ok;
        }
        template <typename T> __device__ __forceinline__ T value_of(const unsigned long long *slot) {
            if constexpr (sizeof(T) == 8)
                return __builtin_bit_cast(T, (slot[0] & 0xFFFFFFFF00000000ull) | (slot[1] >> 32));
            else
                return __builtin_bit_cast(T, (unsigned)(slot[0] >> 32));
        }

        // 16-byte loads of two slots that bypass the caches that are not coherent across the device
        // (the agent-scope cache policy), through a buffer resource over a publication buffer.
        constexpr int kBufferFlags = 0x00020000; // raw buffer, 32-bit elements
        constexpr int kCoherent = 16;            // cache policy: device (agent) scope
        __device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_over(const void *base) {
            return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, 0x7FFFFFFF, kBufferFlags);
        }
        __device__ __forceinline__ void load_pair(__amdgpu_buffer_rsrc_t buffer, unsigned byte, unsigned long long *slot) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(buffer, byte, 0, kCoherent);
            slot[0] = (unsigned long long)v[1] << 32 | v[0];
            slot[1] = (unsigned long long)v[3] << 32 | v[2];
        }

        // A bounded wait for a tagged value (after a failed wait the value is garbage and *error is set).
        template <typename T> __device__ T take(const unsigned long long *slots, unsigned tag, unsigned *error) {
            unsigned long long x[kSlots<T>];
            unsigned spins = 0;
            for (;;) {
#pragma unroll
                for (int i = 0; i < kSlots<T>; ++i)
                    x[i] = __hip_atomic_load(slots + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (current<T>(x, tag)) break;
                if (++spins > kSpinLimit) {
                    __hip_atomic_fetch_or(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                if ((spins & 63u) == 0 && relaxed(error)) break;
                __builtin_amdgcn_s_sleep(1);
            }
            return value_of<T>(x);
        }

        // x - m u with one rounding: every place of the panel that applies an elimination computes it
        // this way, so a value rebuilt from a publication equals the one kept in place.
        template <typename T> __device__ __forceinline__ T eliminated(T x, T m, T u) {
            if constexpr (sizeof(T) == 8) return __builtin_fma(-m, u, x);
            else return __builtin_fmaf(-m, u, x);
        }

        // Fields of a candidate line after slot 0 (the candidate word), kSlots<T> slots each: the
        // candidate's value in the column, its value in the next column as stored when published,
        // and the multiplier of the update the next column still lacks then (0 if none); the same
        // three for the diagonal row, in the line of the workgroup that owns it. A double's fields
        // start at slot 2, so that the diagonal's three begin on a 16-byte pair of their own.
        enum : int { kValue, kNext, kFactor, kDiagonal, kDiagonalNext, kDiagonalFactor };
        template <typename T> __device__ __forceinline__ constexpr int field(int f) {
            return kSlots<T> == 2 ? 2 + 2 * f : 1 + f;
        }
        template <typename T> constexpr int kMainPairs = (field<T>(kDiagonal) + 1) / 2; // word + candidate fields
        template <typename T> constexpr int kAllPairs = (field<T>(kDiagonalFactor + 1) + 1) / 2;

        // What slot s of a line holds: 0 nothing, 1 the candidate word, 2 a tagged half of a field, 3 a
        // tagged half of a next-column field (written only when the column has a next one).
        template <typename T> __device__ __forceinline__ constexpr int slot_kind(int s) {
            if (s == 0) return 1;
            for (int f = kValue; f <= kDiagonalFactor; ++f)
                if (s >= field<T>(f) && s < field<T>(f) + kSlots<T>) return f == kNext || f == kDiagonalNext ? 3 : 2;
            return 0;
        }
        // Whether the 16-byte pair x (slots s0 and s0 + 1 of a line) belongs to this column: the word's
        // epoch, the fields' tags.
        template <typename T>
        __device__ __forceinline__ bool pair_current(const unsigned long long *x, int s0, unsigned tag, unsigned epoch,
                                                     bool more) {
            bool ok = true;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int kind = slot_kind<T>(s0 + h);
                if (kind == 1) ok = ok && (x[h] & 0xFFu) == epoch;
                else if (kind == 2 || (kind == 3 && more)) ok = ok && (unsigned)x[h] == tag;
            }
            return ok;
        }

        // Workgroup barrier for LDS traffic only: the publications (global stores) in flight are not
        // waited for, nothing in the panel needs them ordered.
        __device__ __forceinline__ void settle() {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        }

        // A pivot candidate's rank: the bits of |x| plus one, which order non-negative values as the
        // values do; NaN ranks 0 with an absent candidate, below every number (it never wins,
        // decompose.hpp:35-38 compares with >).
        template <typename T> __device__ __forceinline__ unsigned long long rank_of(T x) {
            if (x != x) return 0;
            if constexpr (sizeof(T) == 8) return __builtin_bit_cast(unsigned long long, T(__builtin_fabs(x))) + 1;
            else return (unsigned long long)__builtin_bit_cast(unsigned, T(__builtin_fabsf(x))) + 1;
        }

        __device__ __forceinline__ int wave_sum(int x) {
#pragma unroll
            for (int offset = 1; offset < 64; offset <<= 1) x += __shfl_xor(x, offset, 64);
            return x;
        }

        // Factors columns [j0, j0 + w) of rows [j0, n) (see lu.hpp, step 1). Workgroup b (four waves)
        // owns rows [j0 + b * kPanelRows, ...) and keeps their panel part in LDS for the whole launch;
        // thread t < kPanelRows owns row t of them.
        //
        // Per column c, every workgroup publishes its candidate (first maximum of |a| among its rows at
        // or below the diagonal): the candidate's panel row, the diagonal row if it owns it, and a line
        // holding the candidate word (candidate_word, also the arrival flag) and the fields the next
        // step needs at once. Every workgroup polls every line together with its fields, so one round
        // trip gives the pivot p, the pivot value and the pivot row's next entry, and the diagonal
        // row's two entries: it swaps and eliminates column c, updates column c + 1 and publishes its
        // candidate for c + 1 before it reads the rest of the pivot row and of the diagonal row, whose
        // round trip overlaps that of the next column's lines. The rest of column c's rank-one update
        // is done by the waves that do not poll (part 1) while the others gather column c + 1.
        //
        // Publications are made as the rows stand in LDS, whose columns beyond c + 1 still lack column
        // c's update then; the reader applies it (eliminated() with the published multiplier and the
        // previous pivot row, which every workgroup keeps). The workgroup that receives the old
        // diagonal row in place of the pivot row has only its columns c and c + 1 at that point; if
        // that row is its next candidate or the next diagonal row, it reads the rest first.
        //
        // Buffers (kParities of each, column c uses c % kParities; a buffer is rewritten three
        // columns later, after every workgroup has read it): lines[4][G][kLine], rows[4][G][kPanel *
        // kSlots] and diagonal[4][kPanel * kSlots], in tagged slots. Lines are zeroed before each
        // launch; rows and diagonal are told apart by the launch number in their tags.
        template <typename T>
        __global__ __launch_bounds__(kPanelThreads) void panel(T *a, int64_t n, int64_t j0, int w, int64_t *pivots,
                                                                unsigned long long *lines,
                                                                unsigned long long *published,
                                                                unsigned long long *diagonal, unsigned launch,
                                                                unsigned *error, int spread, __bf16 *pack_hi,
                                                                __bf16 *pack_lo) {
            constexpr int stride = kPanel + 1;
            constexpr int S = kSlots<T>;
            // With spread > 1 the grid is spread times the panel's workgroups and only every spread-th
            // block works: blocks are dealt to the XCDs round-robin, so with spread = 8 the panel's
            // workgroups all run on one XCD and exchange their candidates through one L2.
            if (spread > 1 && blockIdx.x % spread != 0) return;
            const int G = gridDim.x / spread, b = blockIdx.x / spread, t = threadIdx.x;
            const int lane = t & 63, wave = t >> 6;
            const int k_own = t & (kPanel - 1), part = t / kPanel; // a panel column and which half of the threads
            const int64_t first = j0 + (int64_t)b * kPanelRows;
            const int mine = (int)min((int64_t)kPanelRows, n - first); // rows this workgroup owns
            __shared__ T rows[kPanelRows * stride];
            __shared__ T pivot_rows[2][kPanel]; // the pivot row of column c (c & 1) and of column c - 1
            __shared__ T multipliers[kPanelRows];
            __shared__ unsigned long long wave_word[kWaves]; // gather: each wave's largest word
            __shared__ int wave_sharing[kWaves];             // and how many lines carry its key
            __shared__ T wave_fields[kWaves][3];             // and that word's fields
            __shared__ T diagonal_fields[3];
            __shared__ T search_value[kWaves];
            __shared__ unsigned long long search_rank[kWaves];
            __shared__ int64_t search_row[kWaves];

            const __amdgpu_buffer_rsrc_t line_buffer = buffer_over(lines);
            const __amdgpu_buffer_rsrc_t row_buffer = buffer_over(published);
            const __amdgpu_buffer_rsrc_t diagonal_buffer = buffer_over(diagonal);

            for (int i = part; i < mine; i += kParts)
                if (k_own < w) rows[i * stride + k_own] = a[(first + i) * n + j0 + k_own];
            if (t < kPanelRows) multipliers[t] = T(0);
            settle();

            auto tag_of = [&](int col) { return launch << 8 | (unsigned)(col + 1); };
            auto line_index = [&](int col, int g) { return ((int64_t)(col & (kParities - 1)) * G + g) * kLine; };
            auto row_index = [&](int col, int g) { return ((int64_t)(col & (kParities - 1)) * G + g) * kPanel * S; };
            auto diagonal_index = [&](int col) { return (int64_t)(col & (kParities - 1)) * kPanel * S; };

            // This workgroup's candidate row for column `col` (the same in every thread; LLONG_MAX if none):
            // the largest rank, on equal ranks the lowest row (rows follow the threads).
            auto search = [&](int col) -> int64_t {
                const int64_t dd = j0 + col;
                const unsigned long long rank =
                        t < mine && first + t >= dd ? rank_of(rows[t * stride + col]) : 0ull;
                const unsigned long long top = wave_max_rows(rank);
                const unsigned long long hits = __ballot(rank == top && top != 0);
                if (lane == 0) {
                    search_rank[wave] = top;
                    search_row[wave] = hits ? first + wave * 64 + (__ffsll((long long)hits) - 1) : LLONG_MAX;
                }
                settle();
                unsigned long long best = search_rank[0];
                int64_t row = search_row[0];
#pragma unroll
                for (int q = 1; q < kRowWaves; ++q)
                    if (search_rank[q] > best) {
                        best = search_rank[q];
                        row = search_row[q];
                    }
                return row;
            };
            // Publishes the candidate `row` for column `col`, and the diagonal row if this workgroup owns
            // it; `pending`: the rows' columns beyond col still lack the update of column col - 1, whose
            // multipliers are in `multipliers`. Every publication store comes from the waves of part 1,
            // and every load of another workgroup's publication from those of part 0: a wave's wait for
            // a load also waits for the stores it issued before (one counter counts both), and an
            // agent-scope store is acknowledged only once it reaches memory, so the waves that poll
            // never store.
            auto publish = [&](int col, int64_t row, bool pending) {
                const int64_t dd = j0 + col;
                const unsigned tag = tag_of(col);
                unsigned long long *line = lines + line_index(col, b);
                auto send = [&](int i, unsigned long long *target, int fields) {
                    if (k_own < w) put_tagged(target + k_own * S, rows[i * stride + k_own], tag);
                    if (k_own == 0) {
                        put_tagged(line + field<T>(fields), rows[i * stride + col], tag);
                        if (col + 1 < w) put_tagged(line + field<T>(fields + 1), rows[i * stride + col + 1], tag);
                        put_tagged(line + field<T>(fields + 2), pending ? multipliers[i] : T(0), tag);
                    }
                };
                if (part != 1) return;
                if (row != LLONG_MAX) send((int)(row - first), published + row_index(col, b), kValue);
                else if (t == kPanelRows) { // no candidate: the fields still go out (zeros), every slot tagged
                    put_tagged(line + field<T>(kValue), T(0), tag);
                    if (col + 1 < w) put_tagged(line + field<T>(kNext), T(0), tag);
                    put_tagged(line + field<T>(kFactor), T(0), tag);
                }
                if (dd >= first && dd < first + mine) send((int)(dd - first), diagonal + diagonal_index(col), kDiagonal);
                if (t == kPanelRows) {
                    const T best = row != LLONG_MAX ? T(__builtin_fabs(rows[(row - first) * stride + col])) : T(-1);
                    __hip_atomic_store(line, candidate_word<T>(best, row, j0, (unsigned)col + 1), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                }
            };

            // The rest of column col's rank-one update (columns col + 2 onwards; the diagonal row dd, the
            // multipliers in `multipliers`, the pivot row in pivot_rows[col & 1]), a thread of part 1 per
            // row. A batch loads all its entries before it stores any, so it costs one LDS round trip;
            // the last batch of a row is cut to a quarter, half or all of kUpdateBatch entries and its
            // indices clamped to the row's end, where the repeated entries store the same value again.
            // The batches depend only on col, so every branch here is uniform across the wave.
            constexpr int kUpdateBatch = sizeof(T) == 8 ? 16 : 32;
            auto update_rest = [&](int col, int64_t dd) {
                const int i = t % kPanelRows;
                if (i >= mine || first + i <= dd) return;
                const T l = multipliers[i];
                const T *const v = pivot_rows[col & 1];
                T *const row = rows + i * stride;
                const int end = w;
                int k = col + 2;
                auto batch = [&](auto size) {
                    constexpr int B = decltype(size)::value;
                    T x[B], y[B];
#pragma unroll
                    for (int j = 0; j < B; ++j) {
                        const int at = min(k + j, end - 1);
                        x[j] = row[at];
                        y[j] = v[at];
                    }
#pragma unroll
                    for (int j = 0; j < B; ++j) row[min(k + j, end - 1)] = eliminated(x[j], l, y[j]);
                    k += B;
                };
                while (end - k > kUpdateBatch) batch(std::integral_constant<int, kUpdateBatch>());
                if (end - k > kUpdateBatch / 2) batch(std::integral_constant<int, kUpdateBatch>());
                else if (end - k > kUpdateBatch / 4) batch(std::integral_constant<int, kUpdateBatch / 2>());
                else if (end - k > 0) batch(std::integral_constant<int, kUpdateBatch / 4>());
            };

            publish(0, search(0), false);
            bool previous = false; // whether column c - 1 eliminated (its update is pending in publications)
            for (int c = 0; c < w; ++c) {
                const int64_t d = j0 + c;
                const unsigned tag = tag_of(c);
                const unsigned epoch = ((unsigned)c + 1) & 0xFFu;
                const int owner_d = (int)((d - j0) / kPanelRows);
                const bool more = c + 1 < w;
                T *const u = pivot_rows[c & 1];
                const T *const u_before = pivot_rows[(c + 1) & 1];

                // F of column c - 1: part 1 applies the rest of its rank-one update while part 0 polls
                // for this column (nothing before this column's multipliers, after the barrier that ends
                // the gather, reads those entries).
                if (part == 1 && previous) update_rest(c - 1, d - 1);

                // A: gather. Thread g < kPanelRows takes the lines of workgroups g, g + 128, ... (kPoll per
                // pass) with their fields once each carries this column, and keeps the largest word with its
                // fields; the workgroup takes the largest word, so all workgroups pick the same pivot. Equal
                // float keys of different doubles are settled with the exact values.
                //
                // One lane's loads of one line are served one after another, so the P pairs of a line are
                // loaded by P lanes: lane t polls pair p of line t ^ p (p < P, a group of P adjacent lanes
                // covering P lines), until each of its pairs carries this column, then takes pair p of its
                // own line t from lane t ^ p (DPP within the group). Lanes 64 .. 64 + D - 1 also poll the D
                // pairs of the diagonal row's fields in workgroup owner_d's line. A poll is then one round
                // trip, with the same loads as one lane per line.
                unsigned long long mine_best = 0; // this thread's largest word, and how many share its key
                int sharing = 0;
                T value = T(0), next = T(0), factor = T(0); // the fields of that word's line
                for (int g0 = 0; part == 0 && g0 < G; g0 += kPoll * kPanelRows) {
                    constexpr int P = kMainPairs<T>, D = kAllPairs<T> - kMainPairs<T>;
                    static_assert(P <= 4 && (P & (P - 1)) == 0 && kPoll * P + 1 <= 32, "gather layout");
                    const int diagonal_pair = g0 == 0 ? t - 64 : -1; // this lane's pair of the diagonal fields
                    unsigned long long got[kPoll][P][2] = {}, got_diagonal[2] = {};
                    unsigned waiting = 0; // bit q * P + p: pair p of line g0 + (t ^ p) + q * kPanelRows; bit kPoll * P: the diagonal pair
#pragma unroll
                    for (int q = 0; q < kPoll; ++q)
#pragma unroll
                        for (int p = 0; p < P; ++p)
                            if (g0 + (t ^ p) + q * kPanelRows < G) waiting |= 1u << (q * P + p);
                    if (diagonal_pair >= 0 && diagonal_pair < D) waiting |= 1u << (kPoll * P);
                    unsigned spins = 0;
                    while (waiting) {
#pragma unroll
                        for (int q = 0; q < kPoll; ++q)
#pragma unroll
                            for (int p = 0; p < P; ++p)
                                if (waiting >> (q * P + p) & 1u)
                                    load_pair(line_buffer, (unsigned)(line_index(c, g0 + (t ^ p) + q * kPanelRows) * 8) + 16 * p,
                                              got[q][p]);
                        if (waiting >> (kPoll * P) & 1u)
                            load_pair(line_buffer, (unsigned)(line_index(c, owner_d) * 8) + 16 * (P + diagonal_pair),
                                      got_diagonal);
#pragma unroll
                        for (int q = 0; q < kPoll; ++q)
#pragma unroll
                            for (int p = 0; p < P; ++p)
                                if ((waiting >> (q * P + p) & 1u) && pair_current<T>(got[q][p], 2 * p, tag, epoch, more))
                                    waiting &= ~(1u << (q * P + p));
                        if (waiting >> (kPoll * P) & 1u) {
                            bool ok = false;
#pragma unroll
                            for (int e = 0; e < D; ++e)
                                if (diagonal_pair == e) ok = pair_current<T>(got_diagonal, 2 * (P + e), tag, epoch, more);
                            if (ok) waiting &= ~(1u << (kPoll * P));
                        }
                        if (!waiting) break;
                        if (++spins > kSpinLimit) {
                            __hip_atomic_fetch_or(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            break;
                        }
                        if ((spins & 63u) == 0 && relaxed(error)) break;
                        __builtin_amdgcn_s_sleep(1);
                    }
                    // The diagonal row's fields, each from the pair that holds it.
                    if (diagonal_pair >= 0 && diagonal_pair < D) {
#pragma unroll
                        for (int e = 0; e < D; ++e)
#pragma unroll
                            for (int f = 0; f < 3; ++f) {
                                constexpr int kFirst = 2 * kMainPairs<T>;
                                const int at = field<T>(kDiagonal + f) - kFirst - 2 * e; // slot within pair e
                                if (diagonal_pair == e && at >= 0 && at < 2) diagonal_fields[f] = value_of<T>(got_diagonal + at);
                            }
                    }
                    // Every lane of the wave is here: pair p of line g0 + t + q * kPanelRows from lane t ^ p.
#pragma unroll
                    for (int q = 0; q < kPoll; ++q) {
                        unsigned long long s[2 * P];
#pragma unroll
                        for (int p = 0; p < P; ++p)
#pragma unroll
                            for (int h = 0; h < 2; ++h) {
                                const unsigned long long x = got[q][p][h];
                                unsigned lo = (unsigned)x, hi = (unsigned)(x >> 32);
                                if (p == 1) lo = dpp<0xB1>(lo), hi = dpp<0xB1>(hi); // quad_perm [1, 0, 3, 2]
                                if (p == 2) lo = dpp<0x4E>(lo), hi = dpp<0x4E>(hi); // quad_perm [2, 3, 0, 1]
                                if (p == 3) lo = dpp<0x1B>(lo), hi = dpp<0x1B>(hi); // quad_perm [3, 2, 1, 0]
                                s[2 * p + h] = (unsigned long long)hi << 32 | lo;
                            }
                        if (g0 + t + q * kPanelRows >= G) continue;
                        const unsigned long long x = s[0];
                        bool take_it = false;
                        if (sharing == 0 || (x >> 32) > (mine_best >> 32)) {
                            take_it = true;
                            sharing = 1;
                        } else if ((x >> 32) == (mine_best >> 32)) {
                            take_it = (x >> 8) > (mine_best >> 8);
                            ++sharing;
                        }
                        if (take_it) {
                            mine_best = x;
                            value = value_of<T>(s + field<T>(kValue));
                            next = value_of<T>(s + field<T>(kNext));
                            factor = value_of<T>(s + field<T>(kFactor));
                        }
                    }
                }
                // The workgroup's largest word: each wave's, with its fields and how many lines of the
                // wave carry its key, then the largest of the waves'.
                {
                    const unsigned long long top = wave_max_rows(sharing ? mine_best >> 8 : 0); // key << 24 | ~offset
                    const bool holder = sharing > 0 && (mine_best >> 8) == top;
                    // lines sharing the float key: only double keys can tie (they are rounded)
                    const int carrying = sizeof(T) == 8 ? wave_sum_rows(sharing > 0 && (mine_best >> 32) == (top >> 24)
                                                                                ? sharing : 0)
                                                        : 0;
                    if (holder) {
                        wave_fields[wave][0] = value;
                        wave_fields[wave][1] = next;
                        wave_fields[wave][2] = factor;
                    }
                    if (lane == 0) {
                        wave_word[wave] = top;
                        wave_sharing[wave] = carrying;
                    }
                }
                settle();
                unsigned long long top = wave_word[0];
                int from = 0;
#pragma unroll
                for (int q = 1; q < kWaves; ++q)
                    if (wave_word[q] > top) {
                        top = wave_word[q];
                        from = q;
                    }
                int64_t p = row_of(top << 8, j0);
                T pivot = wave_fields[from][0], pivot_next = wave_fields[from][1], pivot_factor = wave_fields[from][2];
                if constexpr (sizeof(T) == 8) {
                    int carrying = 0;
#pragma unroll
                    for (int q = 0; q < kWaves; ++q)
                        if ((wave_word[q] >> 24) == (top >> 24)) carrying += wave_sharing[q];
                    if (carrying > 1) {
                        // Two or more candidates share the float key: compare the exact values. The lines
                        // of this column are not rewritten before this workgroup publishes again.
                        T v = T(-1);
                        p = LLONG_MAX;
                        for (int g = t; g < G; g += kPanelThreads) {
                            const unsigned long long *line = lines + line_index(c, g);
                            const unsigned long long x = __hip_atomic_load(line, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            if ((x >> 32) == (top >> 24) && (x & 0xFFu) == epoch && has_candidate(x))
                                better(v, p, T(__builtin_fabs(take<T>(line + field<T>(kValue), tag, error))), row_of(x, j0));
                        }
                        wave_best(v, p);
                        if (lane == 0) {
                            search_value[wave] = v;
                            search_row[wave] = p;
                        }
                        settle();
                        v = search_value[0];
                        p = search_row[0];
#pragma unroll
                        for (int q = 1; q < kWaves; ++q) better(v, p, search_value[q], search_row[q]);
                        if (p >= j0 && p < n) {
                            const unsigned long long *line = lines + line_index(c, (int)((p - j0) / kPanelRows));
                            pivot = take<T>(line + field<T>(kValue), tag, error);
                            pivot_next = more ? take<T>(line + field<T>(kNext), tag, error) : T(0);
                            pivot_factor = take<T>(line + field<T>(kFactor), tag, error);
                        }
                    }
                }
                const T diagonal_value = diagonal_fields[0], diagonal_next = diagonal_fields[1],
                        diagonal_factor = diagonal_fields[2];
                // A NaN diagonal stays the pivot: the reference starts from the diagonal and takes a row
                // only if |a| is larger (decompose.hpp:35-37), which nothing is than NaN.
                if (diagonal_value != diagonal_value) p = d;
                if (p < d || p >= n) p = d; // only after a failed wait: keep every index in range
                // With p = d the pivot row is the diagonal row, which is always published.
                const bool at_diagonal = p == d;
                if (at_diagonal) {
                    pivot = diagonal_value;
                    pivot_next = diagonal_next;
                    pivot_factor = diagonal_factor;
                }
                const int owner_p = (int)((p - j0) / kPanelRows);
                // entry c + 1 of the pivot row and of the diagonal row, with column c - 1's update
                const T u1 = !more ? T(0) : previous ? eliminated(pivot_next, pivot_factor, u_before[c + 1]) : pivot_next;
                const T e1 = !more ? T(0) : previous ? eliminated(diagonal_next, diagonal_factor, u_before[c + 1]) : diagonal_next;

                // B: swap, columns c and c + 1 for now (by the thread that owns row p, which eliminates it
                // next). Workgroup owner_p puts the diagonal row in place of row p; workgroup owner_d
                // receives the pivot row in row d after publication.
                const bool swapped = p != d;
                const bool receiver = swapped && b == owner_p;
                if (receiver && t == (int)(p - first)) {
                    rows[t * stride + c] = diagonal_value;
                    if (more) rows[t * stride + c + 1] = e1;
                }
                if (b == 0 && t == kPanelRows) pivots[d] = p;

                // C: multipliers and column c + 1 (look-ahead), each thread on its row.
                const bool eliminate = pivot != T(0); // a zero pivot skips the column (decompose.hpp:44-45)
                if (eliminate && t < kPanelRows) {
                    T l = T(0);
                    if (t < mine && first + t > d) {
                        l = rows[t * stride + c] / pivot;
                        rows[t * stride + c] = l;
                        if (more) rows[t * stride + c + 1] = eliminated(rows[t * stride + c + 1], l, u1);
                    }
                    multipliers[t] = l;
                }

                // Entry k_own of a published row (not columns c and c + 1; `slots` is the row's first slot
                // in `buffer`'s buffer `base`); entries beyond c + 1 get column c - 1's update (multiplier m).
                auto receive = [&](__amdgpu_buffer_rsrc_t buffer, const unsigned long long *base, int64_t slots, T m,
                                   T *out) {
                    const int k = k_own;
                    if (k >= w || k == c || k == c + 1) return;
                    const int64_t at = slots + (int64_t)k * S;
                    unsigned long long x[2];
                    unsigned spins = 0;
                    for (;;) {
                        if constexpr (S == 2)
                            load_pair(buffer, (unsigned)(at * 8), x);
                        else
                            x[0] = __hip_atomic_load(base + at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (current<T>(x, tag)) break;
                        if (++spins > kSpinLimit) {
                            __hip_atomic_fetch_or(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            break;
                        }
                        if ((spins & 63u) == 0 && relaxed(error)) break;
                        __builtin_amdgcn_s_sleep(1);
                    }
                    T v = value_of<T>(x);
                    if (k > c + 1 && previous) v = eliminated(v, m, u_before[k]);
                    out[k] = v;
                };
                bool received = false;

                // D: the candidate for column c + 1 goes out before the rest of this column's work.
                if (more) {
                    const int64_t row = search(c + 1); // its barrier also orders B and C before publication
                    if (receiver && (row == p || d + 1 == p)) {
                        // the swapped-in row is needed whole now
                        if (part == 0) receive(diagonal_buffer, diagonal, diagonal_index(c), diagonal_factor, rows + (p - first) * stride);
                        received = true;
                        settle();
                    }
                    publish(c + 1, row, eliminate);
                } else {
                    settle();
                }

                // E: the rest of the pivot row and, at workgroup owner_p, of the diagonal row (both by part
                // 0); then row d at workgroup owner_d.
                if (part == 0) {
                    if (at_diagonal) receive(diagonal_buffer, diagonal, diagonal_index(c), pivot_factor, u);
                    else receive(row_buffer, published, row_index(c, owner_p), pivot_factor, u);
                    if (t == 0) {
                        u[c] = pivot;
                        if (more) u[c + 1] = u1;
                    }
                    if (receiver && !received)
                        receive(diagonal_buffer, diagonal, diagonal_index(c), diagonal_factor, rows + (p - first) * stride);
                }
                settle();
                if (swapped && b == owner_d && part == 0 && k_own < w) rows[(d - first) * stride + k_own] = u[k_own];
                // F (the rest of the rank-one update) follows at the top of the next column.
                previous = eliminate;
            }
            settle(); // the last column's row d

            for (int i = part; i < mine; i += kParts)
                if (k_own < w) a[(first + i) * n + j0 + k_own] = rows[i * stride + k_own];
            // pack_hi (float factors with a bf16 or bf16x3 update): L21 (the rows below the panel's first
            // kPanel) also in the product's packed operand layout, as pack() would make it after the row
            // swaps, which leave the panel's columns alone: row r - j0 - kPanel, kPanel entries apart.
            if constexpr (sizeof(T) == 4) {
                if (pack_hi && w == kPanel)
                    for (int i = part; i < mine; i += kParts) {
                        const int64_t r = first + i - (j0 + kPanel);
                        if (r < 0) continue;
                        if (pack_lo) pack_element<__bf16, true>(rows[i * stride + k_own], pack_hi, pack_lo, r * kPanel + k_own);
                        else pack_element<__bf16, false>(rows[i * stride + k_own], pack_hi, pack_lo, r * kPanel + k_own);
                    }
            }
        }

        // The panel in strips: columns are factored kStrip at a time. Per column only the strip's
        // entries of each row change; the candidate line carries the candidate's whole strip row, so
        // one exchange between workgroups per column gives every workgroup the pivot row it needs. The
        // rows' entries outside the strip stay as they were at the strip's start (not updated, not
        // swapped) until the strip ends; then the rows that moved are exchanged once and every row
        // takes the strip's updates. Every entry still gets the reference's eliminations one at a
        // time, in column order, each as eliminated(), so the result equals panel()'s bit for bit.
        template <typename T> constexpr int kStrip = sizeof(T) == 8 ? 8 : 16;       // columns per strip
        template <typename T> constexpr int kStripRows = 128; // rows per workgroup (a thread of waves 2-3 each)
        constexpr int kStripThreads = 256;
        constexpr int kRestSlots = 2;     // rest buffers, used in turn by consecutive strips
        constexpr int kStripLineMax = 32; // 64-bit slots per line (256 bytes)
        constexpr int kLineLanes = 16;    // lanes that load one line, a 16-byte pair each, in one instruction
        // Lines a polling lane loads at once (G + 1 lines take (G + 1) / (8 kPollLoads<T>) passes, one round
        // trip each): one pass up to an order of 18176 (float) or 17280 (double). The unrolled poll is
        // kept this short so that the per-column code stays in the instruction cache.
        template <typename T> constexpr int kPollLoads = sizeof(T) == 8 ? 17 : 9;
        // From this many workgroups on, the gather takes two steps: every workgroup polls only the words of
        // the lines (16 bytes of each), then loads the winner's line and the diagonal line. Every workgroup
        // reading every whole line made the polls' traffic the bound (a first poll of 4.7 K cycles with 128
        // workgroups of double rows against 1.2 K with one).
        constexpr int kTwoStepGroups = 48;

        // Slots of a line for column c of a strip [s0, s1). Each workgroup's line holds the candidate word
        // (candidate_word; also the arrival flag), the candidate's strip row (entries s0..s1-1 as they
        // stand: multipliers left of c, then the values; zeros past a short strip or without a candidate)
        // and the multiplier of the update its entries right of c still lack (0 if none). One more line,
        // after the workgroups', holds the same for the diagonal row; its word carries only the column.
        // Values take kSlots<T> tagged slots each, a double's two from a 16-byte boundary.
        template <typename T> struct StripLine {
            static constexpr int S = kSlots<T>, B = kStrip<T>;
            static constexpr int values = S == 2 ? 2 : 1;
            static constexpr int factor = values + B * S;
            static constexpr int end = factor + S;
            static constexpr int pairs = (end + 1) / 2; // 16-byte pairs a reader loads
            static_assert(pairs <= kLineLanes && end <= kStripLineMax, "strip line");
        };

        // A 32-bit OR over each row of 16 lanes (DPP rotations within the row).
        __device__ __forceinline__ unsigned row_or(unsigned x) {
            x |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x121, 0xF, 0xF, false); // row_ror:1
            x |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x122, 0xF, 0xF, false); // row_ror:2
            x |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x124, 0xF, 0xF, false); // row_ror:4
            x |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x128, 0xF, 0xF, false); // row_ror:8
            return x;
        }

        // Factors columns [j0, j0 + w) of rows [j0, n), as panel() does. Workgroup b owns rows
        // [j0 + b R, ...) and keeps their panel part in LDS. Waves 0-1 make every load of another
        // workgroup's publication; thread 128 + r of waves 2-3 owns row r and holds the row's entries in the
        // current strip in registers (a wave's wait for a load also waits for the stores it issued before,
        // so the waves that poll never store).
        //
        // Per column c of strip [s0, s1), as in panel(): gather every line, pick the pivot p (the first
        // maximum; NaN never wins; a NaN diagonal stays), swap the strip entries of rows p and d = j0 + c
        // (the new ones come in the lines), divide the multipliers and update column c + 1, publish the
        // candidate for c + 1 (its entries right of c + 1 still lack column c's update, which the reader
        // applies), then update the strip's other columns. Each row's thread does all of this on its own
        // registers, publishing its row before updating it, so no barrier separates the publication from
        // the update.
        //
        // At the strip's end every workgroup knows the strip's swaps. The rows that moved (the pivot rows,
        // now U rows at d_q, and the rows that arrive at the pivots' places) are published by their owners
        // into the rest buffer with their entries outside the strip (left of it: multipliers of the
        // panel's earlier strips; right of it: values), one slot per row: slot q < B for the U row of
        // strip column q, slot B + q for the row arriving at p_q. Every workgroup reads the U rows and
        // applies the strip's part of L11 to their entries right of the strip (U = L11^-1 rows, in the
        // reference's order); each reads the rows arriving at its own places; then every row below the
        // strip takes the strip's eliminations, column by column.
        //
        // The lines are polled kLineLanes lanes to a line, each lane loading one 16-byte pair, so that one
        // load instruction covers whole lines: a lane issuing several loads of one line waits for them one
        // after another (a round trip of 4.4 us for a 19-slot line polled that way, 1.0 us with one load
        // per lane, measured on MI355X). For the same reason a line is stored one pair per lane.
        //
        // Buffers: lines[kParities][G + 1][kStripLineMax] (cleared before each launch), rests[kRestSlots]
        // [2 B][kPanel] values in tagged slots, told apart by the launch and strip in their tags.
        template <typename T>
        __global__ __launch_bounds__(kStripThreads) void panel_strips(T *a, int64_t n, int64_t j0, int w, int64_t *pivots,
                                                                      unsigned long long *lines,
                                                                      unsigned long long *rests, unsigned launch,
                                                                      unsigned *error) {
            using Line = StripLine<T>;
            constexpr int S = kSlots<T>, B = kStrip<T>, R = kStripRows<T>;
            constexpr int stride = kPanel + 1;
            constexpr int pollers = 128;              // waves 0-1
            constexpr int shares = kStripThreads / R; // threads per row in the strip-end update
            static_assert(R == kStripThreads - pollers, "a row thread per row");
            const int G = gridDim.x, b = blockIdx.x, t = threadIdx.x;
            const int lane = t & 63, wave = t >> 6;
            const int tp = t - pollers; // >= 0: a thread of waves 2-3, which own the rows
            const int64_t first = j0 + (int64_t)b * R;
            const int mine = (int)min((int64_t)R, n - first);
            const bool owner = tp >= 0 && tp < mine; // this thread holds row first + tp
            const int64_t my_row = first + tp;
            __shared__ T rows[R * stride];
            __shared__ T strip_u[B][B];          // the pivot row of strip column q: its strip entries, current
            __shared__ T strip_ut[B][B];         // the same transposed (L11 by columns)
            __shared__ T u_rest[kPanel][B];      // strip end: the U rows' entries outside the strip (column-major)
            __shared__ T incoming[B][kPanel];    // strip end: rows arriving at this workgroup's places
            __shared__ int64_t strip_pivot[B];
            __shared__ int slot_source[2 * B];   // strip end: the row (offset from j0) slot j carries, or -1
            __shared__ int slot_target[2 * B];   // and the place it goes to
            __shared__ unsigned long long wave_word[kWaves];
            __shared__ int wave_sharing[kWaves];
            __shared__ T wave_fields[kWaves][B + 1];
            __shared__ T diagonal_fields[B + 1];
            __shared__ T tie_fields[B + 1];
            __shared__ T search_value[kWaves];
            __shared__ unsigned long long search_rank[kWaves];
            __shared__ int64_t search_row[kWaves];
            __shared__ unsigned long long staged[2][2 * kLineLanes]; // a line's slots on their way out (per wave)

            const __amdgpu_buffer_rsrc_t line_buffer = buffer_over(lines);
            const __amdgpu_buffer_rsrc_t rest_buffer = buffer_over(rests);

            for (int e = t; e < mine * kPanel; e += kStripThreads) {
                const int r = e / kPanel, k = e % kPanel;
                if (k < w) rows[r * stride + k] = a[(first + r) * n + j0 + k];
            }
            settle();

            auto tag_of = [&](int col) { return launch << 8 | (unsigned)(col + 1); };
            // slot index of line g (g = G: the diagonal line) for column col
            auto line_at = [&](int col, int g) {
                return ((int64_t)(col & (kParities - 1)) * (G + 1) + g) * kStripLineMax;
            };

            T sv[B];    // this thread's row, entries s0 .. s0 + B - 1 of the current strip
            T l = T(0); // and its multiplier of the current column
            auto load_strip = [&](int s0, int width) {
#pragma unroll
                for (int k = 0; k < B; ++k) sv[k] = owner && k < width ? rows[tp * stride + s0 + k] : T(0);
            };
            auto store_strip = [&](int s0, int width) {
#pragma unroll
                for (int k = 0; k < B; ++k)
                    if (owner && k < width) rows[tp * stride + s0 + k] = sv[k];
            };

            // This workgroup's candidate row for column `col`, strip entry `k` (the same in every thread;
            // LLONG_MAX if none): the largest rank, on equal ranks the lowest row.
            auto search = [&](int col, int k) -> int64_t {
                const int64_t dd = j0 + col;
                const unsigned long long rank = owner && my_row >= dd ? rank_of(sv[k]) : 0ull;
                const unsigned long long top = wave_max_rows(rank);
                const unsigned long long hits = __ballot(rank == top && top != 0);
                if (lane == 0) {
                    search_rank[wave] = top;
                    search_row[wave] = hits ? first + (wave - 2) * 64 + (__ffsll((long long)hits) - 1) : LLONG_MAX;
                }
                settle();
                unsigned long long best = search_rank[2];
                int64_t row = search_row[2];
                if (search_rank[3] > best) row = search_row[3];
                return row;
            };
            // Publishes the line of column `col` (strip [s0, s0 + width)): this workgroup's candidate `row` and
            // the diagonal row if this workgroup owns it (into line G), each from the registers of the thread
            // that holds it (without a candidate, row thread 0 writes an empty line). `pending`: the rows'
            // entries right of col still lack the previous column's update, with multiplier `l`. The thread
            // stages the line's slots in LDS and the lanes of its wave store one 16-byte pair each, the whole
            // 256-byte line. (Reading the holder's registers with v_readlane instead of staging measured
            // slower: bf16 solve at n = 16384 108.7 against 103.1 ms.)
            auto publish = [&](int col, int s0, int width, int64_t row, bool pending) {
                if (tp < 0) return;
                const int64_t dd = j0 + col;
                const unsigned tag = tag_of(col);
                const int which_wave = wave - 2;
                const int line_wave = row == LLONG_MAX ? 0 : (int)((row - first) / 64);
                const bool diagonal_here = dd >= first && dd < first + mine;
                const int diagonal_wave = diagonal_here ? (int)((dd - first) / 64) : -1;
                auto stage = [&](bool empty, bool diagonal) {
                    unsigned long long slots[2 * kLineLanes]; // the whole line, padding zero
#pragma unroll
                    for (int e = 0; e < 2 * kLineLanes; ++e) slots[e] = 0;
                    const int q = col - s0;
                    T current = T(0);
#pragma unroll
                    for (int k = 0; k < B; ++k) {
                        const T v = !empty && k < width ? sv[k] : T(0);
                        if (k == q) current = v;
                        if constexpr (S == 2) {
                            const unsigned long long bits = __builtin_bit_cast(unsigned long long, v);
                            slots[Line::values + 2 * k] = (bits & 0xFFFFFFFF00000000ull) | tag;
                            slots[Line::values + 2 * k + 1] = bits << 32 | tag;
                        } else {
                            slots[Line::values + k] = (unsigned long long)__builtin_bit_cast(unsigned, v) << 32 | tag;
                        }
                    }
                    const T f = !empty && pending ? l : T(0);
                    if constexpr (S == 2) {
                        const unsigned long long bits = __builtin_bit_cast(unsigned long long, f);
                        slots[Line::factor] = (bits & 0xFFFFFFFF00000000ull) | tag;
                        slots[Line::factor + 1] = bits << 32 | tag;
                    } else {
                        slots[Line::factor] = (unsigned long long)__builtin_bit_cast(unsigned, f) << 32 | tag;
                    }
                    slots[0] = diagonal ? (unsigned long long)(((unsigned)col + 1) & 0xFFu)
                                        : candidate_word<T>(empty ? T(-1) : T(__builtin_fabs(current)),
                                                            empty ? LLONG_MAX : my_row, j0, (unsigned)col + 1);
#pragma unroll
                    for (int e = 0; e < 2 * kLineLanes; ++e) staged[which_wave][e] = slots[e];
                };
                // one wave stages and stores a line at a time (both lines may come from one wave)
#pragma unroll
                for (int kind = 0; kind < 2; ++kind) {
                    const bool diagonal = kind == 1;
                    if (diagonal ? diagonal_wave != which_wave : line_wave != which_wave) continue;
                    const bool writer = diagonal ? owner && my_row == dd
                                                 : (row == LLONG_MAX ? tp == 0 : owner && my_row == row);
                    if (writer) stage(!diagonal && row == LLONG_MAX, diagonal);
                    // the staging thread's LDS writes before the wave's reads (one wave: in order)
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
                    if (lane < kLineLanes) {
                        const unsigned long long x0 = staged[which_wave][2 * lane], x1 = staged[which_wave][2 * lane + 1];
                        typedef unsigned v4u __attribute__((ext_vector_type(4)));
                        const v4u v = {(unsigned)x0, (unsigned)(x0 >> 32), (unsigned)x1, (unsigned)(x1 >> 32)};
                        __builtin_amdgcn_raw_buffer_store_b128(v, line_buffer,
                                                               (unsigned)((line_at(col, diagonal ? G : b) + 2 * lane) * 8), 0,
                                                               kCoherent);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
                }
            };

            // Which rows arrive at place x in the last strip's swaps (the slot, or -1).
            auto arriving_at = [&](int64_t x) -> int {
                int arriving = -1;
#pragma unroll
                for (int j = B; j < 2 * B; ++j)
                    if (slot_source[j] >= 0 && j0 + slot_target[j] == x) arriving = j - B;
                return arriving;
            };
            // Strip [ps0, ps1)'s end, right of the strip, columns [k0, k1), in row r of this workgroup: a U row
            // takes its solved entries, a row below the strip (an arriving row's entries) the strip's
            // eliminations (mask: the strip's columns whose pivot was not zero; the strip is full, B columns).
            auto update_right = [&](int r, int ps0, int ps1, unsigned mask, int k0, int k1) {
                if (k0 >= k1) return;
                const int64_t x = first + r;
                T *row = rows + r * stride;
                if (x < j0 + ps1) {
                    const int qq = (int)(x - (j0 + ps0));
                    for (int k = k0; k < k1; ++k) row[k] = u_rest[k][qq];
                    return;
                }
                const int arriving = arriving_at(x);
                const T *from = arriving >= 0 ? incoming[arriving] : row;
                T lr[B];
#pragma unroll
                for (int h = 0; h < B; ++h) lr[h] = row[ps0 + h];
                if (mask == (B == 32 ? ~0u : (1u << B) - 1u)) {
                    // four columns at a time: four independent chains of subtractions; u_rest's entries read
                    // four at a time
                    int k = k0;
                    for (; k + 4 <= k1; k += 4) {
                        T v[4];
#pragma unroll
                        for (int j = 0; j < 4; ++j) v[j] = from[k + j];
#pragma unroll
                        for (int r0 = 0; r0 < B; r0 += 4) {
                            T u[4][4];
#pragma unroll
                            for (int j = 0; j < 4; ++j)
#pragma unroll
                                for (int h = 0; h < 4; ++h) u[j][h] = u_rest[k + j][r0 + h];
#pragma unroll
                            for (int h = 0; h < 4; ++h)
#pragma unroll
                                for (int j = 0; j < 4; ++j) v[j] = eliminated(v[j], lr[r0 + h], u[j][h]);
                        }
#pragma unroll
                        for (int j = 0; j < 4; ++j) row[k + j] = v[j];
                    }
                    for (; k < k1; ++k) {
                        T v = from[k];
#pragma unroll
                        for (int h = 0; h < B; ++h) v = eliminated(v, lr[h], u_rest[k][h]);
                        row[k] = v;
                    }
                } else {
                    for (int k = k0; k < k1; ++k) {
                        T v = from[k];
#pragma unroll
                        for (int h = 0; h < B; ++h)
                            if (mask >> h & 1u) v = eliminated(v, lr[h], u_rest[k][h]);
                        row[k] = v;
                    }
                }
            };

            load_strip(0, min(B, w));
            publish(0, 0, min(B, w), search(0, 0), false);
            for (int s0 = 0; s0 < w; s0 += B) {
                const int s1 = min(s0 + B, w), width = s1 - s0;
                bool previous = false;           // whether the previous column of this strip eliminated
                unsigned eliminated_columns = 0; // the strip's columns whose pivot was not zero
                for (int c = s0; c < s1; ++c) {
                    const int q = c - s0;
                    const int64_t d = j0 + c;
                    const unsigned tag = tag_of(c);
                    const unsigned epoch = ((unsigned)c + 1) & 0xFFu;
                    const int owner_d = (int)((d - j0) / R);
                    const bool more = c + 1 < s1;

                    // A: gather. Below kTwoStepGroups workgroups, lane t of waves 0-1 loads pair t % 16 of lines
                    // t / 16 + 8 i (the diagonal line, index G, among them) until each carries this column; each
                    // group of 16 lanes keeps the largest word of its lines with its pair of that line, and the
                    // waves then take the largest. From kTwoStepGroups on, the lanes poll the words only, and the
                    // winner's and the diagonal line follow (step 2, after the pivot is known).
                    unsigned long long best_word = 0;
                    unsigned best_pair[4] = {0u, 0u, 0u, 0u};
                    unsigned diagonal_pair[4] = {0u, 0u, 0u, 0u}; // this lane's pair of the diagonal line
                    int sharing = 0; // lines of the group whose float key equals its best one's
                    const int pair = t % kLineLanes, group = t / kLineLanes;
                    constexpr int groups = pollers / kLineLanes;
                    // A pair's values into an array of B + 1 (the strip entries, then the factor).
                    auto unpack = [&](const unsigned *x, T *fields) {
                        if constexpr (S == 2) {
                            const int k = pair - Line::values / 2; // the value this pair holds
                            const T v = __builtin_bit_cast(T, (unsigned long long)x[1] << 32 | x[3]);
                            if (k >= 0 && k < B) fields[k] = v;
                            else if (2 * pair == Line::factor) fields[B] = v;
                        } else {
#pragma unroll
                            for (int h = 0; h < 2; ++h) {
                                const int slot = 2 * pair + h;
                                const T v = __builtin_bit_cast(T, x[2 * h + 1]);
                                if (slot >= Line::values && slot < Line::factor) fields[slot - Line::values] = v;
                                else if (slot == Line::factor) fields[B] = v;
                            }
                        }
                    };
                    const bool two_step = G >= kTwoStepGroups;
                    if (two_step) {
                        // Step 1: lane t loads the first pair (the word) of lines t, t + 128, ...
                        for (int base = 0; t < pollers && base < G; base += pollers) {
                            const int g = min(base + t, G - 1);
                            bool waiting = base + t < G;
                            unsigned long long x = 0;
                            // Three polls in flight, issued one after another: the wait ends about a third
                            // of a round trip after the word lands instead of up to a whole one.
                            const unsigned byte = (unsigned)(line_at(c, g) * 8);
                            auto issue = [&]() {
                                return __builtin_amdgcn_raw_buffer_load_b128(line_buffer, byte, 0, kCoherent);
                            };
                            auto done = [&](decltype(issue()) v) { // true once every lane of the wave has its word
                                const unsigned long long y = (unsigned long long)v[1] << 32 | v[0];
                                if (waiting && (y & 0xFFu) == epoch) {
                                    waiting = false;
                                    x = y;
                                }
                                return __ballot(waiting) == 0;
                            };
                            auto v0 = issue();
                            __builtin_amdgcn_s_sleep(2);
                            auto v1 = issue();
                            __builtin_amdgcn_s_sleep(2);
                            auto v2 = issue();
                            unsigned spins = 0;
                            for (;;) {
                                if (done(v0)) break;
                                v0 = issue();
                                if (done(v1)) break;
                                v1 = issue();
                                if (done(v2)) break;
                                v2 = issue();
                                spins += 3;
                                if (spins > kSpinLimit) {
                                    __hip_atomic_fetch_or(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                    break;
                                }
                                if ((spins & 63u) < 3 && relaxed(error)) break;
                            }
                            if (base + t >= G) continue;
                            if (sharing == 0 || (x >> 32) > (best_word >> 32)) {
                                best_word = x;
                                sharing = 1;
                            } else if ((x >> 32) == (best_word >> 32)) {
                                if ((x >> 8) > (best_word >> 8)) best_word = x;
                                ++sharing;
                            }
                        }
                    } else {
                        // Every branch around a load is uniform across the wave (a load under a branch taken by some
                        // lanes only was compiled to wait for the loads before it): a wave issues `count` loads per
                        // poll, lanes past the last line or the last pair loading the diagonal line or their line's
                        // last pair again.
                        const int wave_group0 = (t / 64) * (64 / kLineLanes); // the wave's first group
                        for (int pass0 = 0; t < pollers && wave_group0 + pass0 <= G; pass0 += groups * kPollLoads<T>) {
                            const int count = min(kPollLoads<T>, (G - wave_group0 - pass0) / groups + 1);
                            unsigned pending = 0;
    #pragma unroll
                            for (int i = 0; i < kPollLoads<T>; ++i)
                                if (i < count && group + pass0 + groups * i <= G) pending |= 1u << i;
                            const int loaded_pair = min(pair, Line::pairs - 1);
                            unsigned got[kPollLoads<T>][4];
                            unsigned spins = 0;
                            if (count == 1) {
                                // one line per lane: three polls in flight, as in the two-step gather
                                const int g = min(group + pass0, G);
                                const unsigned byte = (unsigned)((line_at(c, g) + 2 * loaded_pair) * 8);
                                auto issue = [&]() {
                                    return __builtin_amdgcn_raw_buffer_load_b128(line_buffer, byte, 0, kCoherent);
                                };
                                auto done = [&](decltype(issue()) v) {
                                    bool ok;
                                    if (pair == 0) ok = (v[0] & 0xFFu) == epoch && (S == 2 || v[2] == tag);
                                    else ok = v[0] == tag && (2 * pair + 1 >= Line::end || v[2] == tag);
                                    if (pending & 1u) {
                                        got[0][0] = v[0];
                                        got[0][1] = v[1];
                                        got[0][2] = v[2];
                                        got[0][3] = v[3];
                                    }
                                    pending &= row_or(!ok && pair < Line::pairs ? 1u : 0u);
                                    return __ballot(pending != 0) == 0;
                                };
                                auto v0 = issue();
                                __builtin_amdgcn_s_sleep(2);
                                auto v1 = issue();
                                __builtin_amdgcn_s_sleep(2);
                                auto v2 = issue();
                                for (;;) {
                                    if (done(v0)) break;
                                    v0 = issue();
                                    if (done(v1)) break;
                                    v1 = issue();
                                    if (done(v2)) break;
                                    v2 = issue();
                                    spins += 3;
                                    if (spins > kSpinLimit) {
                                        __hip_atomic_fetch_or(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                        break;
                                    }
                                    if ((spins & 63u) < 3 && relaxed(error)) break;
                                }
                            } else
                            for (;;) {
    #pragma unroll
                                for (int i = 0; i < kPollLoads<T>; ++i) {
                                    if (i >= count) continue;
                                    const int g = min(group + pass0 + groups * i, G);
                                    const auto v = __builtin_amdgcn_raw_buffer_load_b128(
                                            line_buffer, (unsigned)((line_at(c, g) + 2 * loaded_pair) * 8), 0, kCoherent);
                                    got[i][0] = v[0];
                                    got[i][1] = v[1];
                                    got[i][2] = v[2];
                                    got[i][3] = v[3];
                                }
                                unsigned stale = 0; // lines whose pair in this lane is not current yet
    #pragma unroll
                                for (int i = 0; i < kPollLoads<T>; ++i) {
                                    if (i >= count) continue;
                                    bool ok;
                                    if (pair == 0) ok = (got[i][0] & 0xFFu) == epoch && (S == 2 || got[i][2] == tag);
                                    else ok = got[i][0] == tag && (2 * pair + 1 >= Line::end || got[i][2] == tag);
                                    if (!ok && pair < Line::pairs) stale |= 1u << i;
                                }
                                pending &= row_or(stale); // a line is done when every lane of its group is
                                if (__ballot(pending != 0) == 0) break;
                                if (++spins > kSpinLimit) {
                                    __hip_atomic_fetch_or(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                    break;
                                }
                                if ((spins & 63u) == 0 && relaxed(error)) break;
                                __builtin_amdgcn_s_sleep(1);
                            }
    #pragma unroll
                            for (int i = 0; i < kPollLoads<T>; ++i) {
                                if (i >= count) continue;
                                const int g = group + pass0 + groups * i;
                                if (g == G) {
    #pragma unroll
                                    for (int h = 0; h < 4; ++h) diagonal_pair[h] = got[i][h];
                                }
                                // the line's word, from the group's first lane
                                const unsigned lo = (unsigned)__shfl((int)got[i][0], t & ~(kLineLanes - 1), 64);
                                const unsigned hi = (unsigned)__shfl((int)got[i][1], t & ~(kLineLanes - 1), 64);
                                const unsigned long long x = (unsigned long long)hi << 32 | lo;
                                if (g >= G) continue;
                                bool take_it = false;
                                if (sharing == 0 || (x >> 32) > (best_word >> 32)) {
                                    take_it = true;
                                    sharing = 1;
                                } else if ((x >> 32) == (best_word >> 32)) {
                                    take_it = (x >> 8) > (best_word >> 8);
                                    ++sharing;
                                }
                                if (take_it) {
                                    best_word = x;
    #pragma unroll
                                    for (int h = 0; h < 4; ++h) best_pair[h] = got[i][h];
                                }
                            }
                        }
                    // the group that polled the diagonal line (line G of every pass is polled by group G % 8)
                    if (t < pollers && group == G % groups && pair < Line::pairs) unpack(diagonal_pair, diagonal_fields);
                    }
                    {
                        const bool leader = t < pollers && (two_step || pair == 0);
                        const unsigned long long top = wave_max_rows(sharing ? best_word >> 8 : 0); // key << 24 | ~offset
                        const bool holder = t < pollers && sharing > 0 && (best_word >> 8) == top && (top & 0xFFFFFFu) != 0;
                        // lines sharing the float key: only double keys can tie (they are rounded)
                        const int carrying = sizeof(T) == 8 ? wave_sum_rows(leader && sharing > 0 && (best_word >> 32) == (top >> 24)
                                                                                    ? sharing : 0)
                                                            : 0;
                        if (!two_step && holder && pair < Line::pairs) unpack(best_pair, wave_fields[wave]);
                        if (lane == 0) {
                            wave_word[wave] = top;
                            wave_sharing[wave] = carrying;
                        }
                    }
                    settle();
                    unsigned long long top = wave_word[0];
                    int from = 0;
#pragma unroll
                    for (int v = 1; v < kWaves; ++v)
                        if (wave_word[v] > top) {
                            top = wave_word[v];
                            from = v;
                        }
                    int64_t p = row_of(top << 8, j0);
                    const T *pf = wave_fields[from];
                    if constexpr (sizeof(T) == 8) {
                        int carrying = 0;
#pragma unroll
                        for (int v = 0; v < kWaves; ++v)
                            if ((wave_word[v] >> 24) == (top >> 24)) carrying += wave_sharing[v];
                        if (carrying > 1) {
                            // Two or more candidates share the float key: compare the exact values. The lines
                            // of this column are not rewritten before this workgroup publishes again.
                            T v = T(-1);
                            p = LLONG_MAX;
                            for (int g = t; g < G; g += kStripThreads) {
                                const unsigned long long *line = lines + line_at(c, g);
                                const unsigned long long x =
                                        __hip_atomic_load(line, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                if ((x >> 32) == (top >> 24) && (x & 0xFFu) == epoch && has_candidate(x))
                                    better(v, p, T(__builtin_fabs(take<T>(line + Line::values + q * S, tag, error))),
                                           row_of(x, j0));
                            }
                            wave_best(v, p);
                            if (lane == 0) {
                                search_value[wave] = v;
                                search_row[wave] = p;
                            }
                            settle();
                            v = search_value[0];
                            p = search_row[0];
#pragma unroll
                            for (int u = 1; u < kWaves; ++u) better(v, p, search_value[u], search_row[u]);
                            if (!two_step && p >= j0 && p < n) {
                                const unsigned long long *line = lines + line_at(c, (int)((p - j0) / R));
                                if (t < width) tie_fields[t] = take<T>(line + Line::values + t * S, tag, error);
                                else if (t == B) tie_fields[B] = take<T>(line + Line::factor, tag, error);
                                settle();
                                pf = tie_fields;
                            }
                        }
                    }
                    if (two_step) {
                        // Step 2: the winner's line (lanes 0-15) and the diagonal line (lanes 16-31).
                        const bool found = p >= j0 && p < n;
                        if (t < 2 * kLineLanes) {
                            const bool winner = t < kLineLanes;
                            const int g = winner && found ? (int)((p - j0) / R) : G;
                            const int loaded_pair = min(pair, Line::pairs - 1);
                            bool waiting = pair < Line::pairs;
                            unsigned got[4] = {0u, 0u, 0u, 0u};
                            unsigned spins = 0;
                            for (;;) {
                                const auto v = __builtin_amdgcn_raw_buffer_load_b128(
                                        line_buffer, (unsigned)((line_at(c, g) + 2 * loaded_pair) * 8), 0, kCoherent);
                                got[0] = v[0];
                                got[1] = v[1];
                                got[2] = v[2];
                                got[3] = v[3];
                                bool ok;
                                if (pair == 0) ok = (got[0] & 0xFFu) == epoch && (S == 2 || got[2] == tag);
                                else ok = got[0] == tag && (2 * pair + 1 >= Line::end || got[2] == tag);
                                if (ok) waiting = false;
                                if (__ballot(waiting) == 0) break;
                                if (++spins > kSpinLimit) {
                                    __hip_atomic_fetch_or(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                    break;
                                }
                                if ((spins & 63u) == 0 && relaxed(error)) break;
                                __builtin_amdgcn_s_sleep(1);
                            }
                            if (pair < Line::pairs) unpack(got, winner ? wave_fields[0] : diagonal_fields);
                        }
                        settle();
                        pf = wave_fields[0];
                    }
                    // A NaN diagonal stays the pivot: the reference starts from the diagonal and takes a row
                    // only if |a| is larger (decompose.hpp:35-37), which nothing is than NaN.
                    if (diagonal_fields[q] != diagonal_fields[q]) p = d;
                    if (p < d || p >= n) p = d; // only after a failed wait: keep every index in range
                    const bool swapped = p != d;
                    if (!swapped) pf = diagonal_fields;
                    // Strip entry k of a published row as it stands now: entries right of c still lack the
                    // previous column's update (its multiplier is the row's field B).
                    auto now = [&](const T *f, int k) -> T {
                        return previous && k > q ? eliminated(f[k], f[B], strip_u[q - 1][k]) : f[k];
                    };
                    const T pivot = pf[q];
                    const bool eliminate = pivot != T(0); // a zero pivot skips the column (decompose.hpp:44-45)
                    const int owner_p = (int)((p - j0) / R);

                    // B: swap the strip entries of rows p and d; record the pivot row and the step.
                    if (swapped && b == owner_p && my_row == p)
#pragma unroll
                        for (int k = 0; k < B; ++k)
                            if (k < width) sv[k] = now(diagonal_fields, k);
                    if (swapped && b == owner_d && my_row == d)
#pragma unroll
                        for (int k = 0; k < B; ++k)
                            if (k < width) sv[k] = now(pf, k);
                    if (t < width) {
                        const T v = now(pf, t);
                        strip_u[q][t] = v;
                        strip_ut[t][q] = v;
                    }
                    if (t == 0) {
                        strip_pivot[q] = p;
                        if (b == 0) pivots[d] = p;
                    }
                    if (eliminate) eliminated_columns |= 1u << q;

                    // C: multipliers and column c + 1 (look-ahead), each row's thread on its row.
                    const bool below = owner && my_row > d;
                    l = T(0);
                    if (eliminate && below) {
                        const T u1 = more ? now(pf, q + 1) : T(0);
                        l = sv[q] / pivot;
                        sv[q] = l;
                        if (more) sv[q + 1] = eliminated(sv[q + 1], l, u1);
                    }

                    // D: the candidate for column c + 1 goes out, then F: the strip's other columns. The next
                    // column's first barrier, after its gather, orders F before anything reads the rows again.
                    if (more) {
                        const int64_t candidate = search(c + 1, q + 1); // its barrier also orders B
                        publish(c + 1, s0, width, candidate, eliminate);
                        if (eliminate && below) {
#pragma unroll
                            for (int k = 0; k < B; ++k)
                                if (k >= q + 2 && k < width) sv[k] = eliminated(sv[k], l, strip_u[q][k]);
                        }
                    }
                    previous = eliminate;
                }
                store_strip(s0, width);
                settle();
                // A panel of one strip is done; after the last of several, the rows that moved still take
                // their entries left of the strip.
                if (s0 == 0 && s1 == w) break;

                // The strip's end. Slots: which row each carries and where it goes (wave 0).
                const int nq = width;
                if (t < 2 * B) {
                    int64_t sp[B];
#pragma unroll
                    for (int r = 0; r < B; ++r) sp[r] = strip_pivot[r]; // entries r >= nq are not used
                    const int qq = t % B;
                    const bool u_slot = t < B;
                    int source = -1, target = -1;
                    if (qq < nq) {
                        const int64_t place = u_slot ? j0 + s0 + qq : strip_pivot[qq];
                        bool wanted = u_slot || place >= j0 + s1;
#pragma unroll
                        for (int r = 0; r < B; ++r)
                            if (!u_slot && r < qq) wanted = wanted && sp[r] != place; // the first of equal places
                        int64_t x = place; // the row that ends at `place`: the strip's swaps traced backwards
#pragma unroll
                        for (int r = B - 1; r >= 0; --r) {
                            if (r >= nq) continue;
                            const int64_t dr = j0 + s0 + r;
                            if (x == dr) x = sp[r];
                            else if (x == sp[r]) x = dr;
                        }
                        if (wanted && (u_slot || x != place)) {
                            source = (int)(x - j0);
                            target = (int)(place - j0);
                        }
                    }
                    slot_source[t] = source;
                    slot_target[t] = target;
                }
                settle();
                const int span = w - s1;
                const unsigned rest_tag = tag_of(s1 - 1);
                unsigned long long *rest = rests + (size_t)((s0 / B) % kRestSlots) * 2 * B * kPanel * S;
                // The rows in the slots move whole: their entries left of the strip (the multipliers of the
                // panel's earlier strips) and right of it. Entry index e < outside counts those columns.
                const int outside = w - width;
                auto column_of = [&](int e) { return e < s0 ? e : e - s0 + s1; };
                // Owners publish them (waves 2-3).
                if (tp >= 0 && tp < outside) {
                    const int k = column_of(tp);
#pragma unroll
                    for (int j = 0; j < 2 * B; ++j) {
                        const int64_t source = j0 + slot_source[j];
                        if (slot_source[j] >= 0 && source >= first && source < first + mine)
                            put_tagged(rest + ((size_t)j * kPanel + k) * S, rows[(source - first) * stride + k], rest_tag);
                    }
                }
                // Every workgroup reads the U rows, and the rows arriving at its places (waves 0-1). A 16-byte
                // load takes two float entries (k even: s0 and s1 are multiples of the strip width) or one
                // double.
                if (t < pollers) {
                    constexpr int per_load = 2 / S;
                    const int left_loads = s0 / per_load;
                    const int loads = left_loads + (span + per_load - 1) / per_load;
                    // lanes past the last entry load the first one again (every load is unconditional per lane)
                    const int lane_load = t < loads ? t : 0;
                    const int k = lane_load < left_loads ? lane_load * per_load : s1 + (lane_load - left_loads) * per_load;
                    const int k_end = lane_load < left_loads ? s0 : w; // entries k + 1 at or past this are not published
                    const unsigned rest_byte = (unsigned)((rest - rests) * 8);
                    // Two batches of B slots: the U rows (every workgroup), then the arriving rows (the workgroups
                    // they arrive at). Which slots a workgroup reads is the same in all its lanes.
#pragma unroll 1
                    for (int batch = 0; batch < 2; ++batch) {
                        unsigned needed = 0;
#pragma unroll
                        for (int jj = 0; jj < B; ++jj) {
                            const int j = batch * B + jj;
                            const int64_t target = j0 + slot_target[j];
                            if (slot_source[j] >= 0 && (batch == 0 || (target >= first && target < first + mine)))
                                needed |= 1u << jj;
                        }
                        if (!needed) continue;
                        unsigned waiting = t < loads ? needed : 0u;
                        unsigned long long got[B][2];
                        unsigned spins = 0;
                        for (;;) {
#pragma unroll
                            for (int jj = 0; jj < B; ++jj)
                                if (needed >> jj & 1u)
                                    load_pair(rest_buffer,
                                              rest_byte + (unsigned)(((size_t)(batch * B + jj) * kPanel + k) * S * 8), got[jj]);
#pragma unroll
                            for (int jj = 0; jj < B; ++jj) {
                                if (!(waiting >> jj & 1u)) continue;
                                bool ok;
                                if constexpr (S == 2) ok = current<T>(got[jj], rest_tag);
                                else ok = (unsigned)got[jj][0] == rest_tag && (k + 1 >= k_end || (unsigned)got[jj][1] == rest_tag);
                                if (!ok) continue;
                                waiting &= ~(1u << jj);
                                T *target = batch == 0 ? nullptr : incoming[jj];
                                if (batch == 0) u_rest[k][jj] = value_of<T>(got[jj]);
                                else target[k] = value_of<T>(got[jj]);
                                if constexpr (S == 1) {
                                    if (k + 1 < k_end) {
                                        if (batch == 0) u_rest[k + 1][jj] = value_of<T>(got[jj] + 1);
                                        else target[k + 1] = value_of<T>(got[jj] + 1);
                                    }
                                }
                            }
                            if (__ballot(waiting != 0) == 0) break;
                            if (++spins > kSpinLimit) {
                                __hip_atomic_fetch_or(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                break;
                            }
                            if ((spins & 63u) == 0 && relaxed(error)) break;
                            __builtin_amdgcn_s_sleep(1);
                        }
                    }
                }
                settle();
                // U rows right of the strip: row q of the strip subtracts l[q][r] u[r] for r = 0, 1, ..., q - 1
                // in that order, skipping columns whose pivot was zero; a thread per column. The rows are
                // taken together, r by r, so each step's subtractions are independent. (Entries right of the
                // strip exist only after a strip of full width: nq = B here.)
                const bool all_eliminated = eliminated_columns == (B == 32 ? ~0u : (1u << B) - 1u);
                if (t < span) {
                    const int k = s1 + t;
                    T u[B];
#pragma unroll
                    for (int qq = 0; qq < B; ++qq) u[qq] = u_rest[k][qq];
                    if (all_eliminated) {
#pragma unroll
                        for (int r = 0; r < B - 1; ++r)
#pragma unroll
                            for (int qq = r + 1; qq < B; ++qq) u[qq] = eliminated(u[qq], strip_ut[r][qq], u[r]);
                    } else {
#pragma unroll
                        for (int r = 0; r < B - 1; ++r) {
                            if (!(eliminated_columns >> r & 1u)) continue;
#pragma unroll
                            for (int qq = r + 1; qq < B; ++qq) u[qq] = eliminated(u[qq], strip_ut[r][qq], u[r]);
                        }
                    }
#pragma unroll
                    for (int qq = 0; qq < B; ++qq) u_rest[k][qq] = u[qq];
                }
                settle();
                // Every row from the strip's first place down: a U row takes its entries outside the strip
                // (left: as they were; right: solved), an arriving row its entries left of the strip, and
                // every row below the strip the strip's eliminations right of it (an arriving row's entries).
                {
                    const int i = t % R, part = t / R;
                    const int64_t x = first + i;
                    if (i < mine && x >= j0 + s0) {
                        T *row = rows + i * stride;
                        const int end = s1 + span * (part + 1) / shares;
                        const int begin = s1 + span * part / shares;
                        const int left_end = s0 * (part + 1) / shares, left_begin = s0 * part / shares;
                        if (x < j0 + s1) {
                            const int qq = (int)(x - (j0 + s0));
                            for (int k = left_begin; k < left_end; ++k) row[k] = u_rest[k][qq];
                        } else {
                            const int arriving = arriving_at(x);
                            if (arriving >= 0)
                                for (int k = left_begin; k < left_end; ++k) row[k] = incoming[arriving][k];
                        }
                        update_right(i, s0, s1, eliminated_columns, begin, end);
                    }
                }
                settle();
                if (s1 < w) {
                    load_strip(s1, min(B, w - s1));
                    publish(s1, s1, min(B, w - s1), search(s1, 0), false);
                }
            }

            for (int e = t; e < mine * kPanel; e += kStripThreads) {
                const int r = e / kPanel, k = e % kPanel;
                if (k < w) a[(first + r) * n + j0 + k] = rows[r * stride + k];
            }
        }

        // U12 = L11^-1 A12 in place (rows [j0, j0 + kPanel) of columns [j0 + kPanel, n)) by forward
        // substitution with the unit lower L11. Row r subtracts l[r][s] times row s for s = 0, 1, ...,
        // r - 1 in that order, which is the order in which the reference's elimination reaches A12
        // (decompose.hpp:47-53), so NaN and infinity reach exactly the entries it reaches. Each
        // workgroup copies L11's strict lower part to LDS, where every lane reads the same entry at
        // once. Thread k owns column k of A12 and takes its rows a quarter (32 rows) at a
        // time, in registers: the quarter receives the updates of each solved quarter before it (whose
        // rows it reads back), then is solved itself and stored. The code is a 32 x 32 update and a
        // 32-row triangle, small enough for the instruction cache. `l` is L11 and `u` A12 (row stride n).
        // Threads per workgroup: one wave for fp64, four for fp32 (whose entries take half the LDS
        // bandwidth), as measured.
        template <typename T> constexpr int kSolveU12Threads = sizeof(T) == 8 ? 64 : 256;
        template <typename T>
        __global__ __launch_bounds__(kSolveU12Threads<T>) void solve_u12(const T *__restrict__ l, T *__restrict__ u,
                                                                         int64_t n, int64_t rest) {
            constexpr int Q = 32;
            constexpr int kThreads = kSolveU12Threads<T>;
            constexpr int kRows = kThreads / 64; // rows of L11 copied per pass, one per wave
            __shared__ T lower[kPanel * kPanel];
            // L11's strict lower part, 32 passes' loads in flight at a time (each lane two entries).
            const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
            for (int r0 = 0; r0 < kPanel; r0 += 32 * kRows) {
                T near[32], far[32];
#pragma unroll
                for (int i = 0; i < 32; ++i) {
                    const int r = r0 + i * kRows + wave;
                    near[i] = r < kPanel && lane < r ? l[(int64_t)r * n + lane] : T(0);
                    far[i] = r < kPanel && lane + 64 < r ? l[(int64_t)r * n + lane + 64] : T(0);
                }
#pragma unroll
                for (int i = 0; i < 32; ++i) {
                    const int r = r0 + i * kRows + wave;
                    if (r < kPanel) {
                        lower[r * kPanel + lane] = near[i];
                        lower[r * kPanel + lane + 64] = far[i];
                    }
                }
            }
            __syncthreads();
            const int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x;
            if (k >= rest) return;
            T *__restrict__ column = u + k;
            for (int q = 0; q < kPanel; q += Q) {
                T y[Q];
#pragma unroll
                for (int i = 0; i < Q; ++i) y[i] = column[(int64_t)(q + i) * n];
                for (int p0 = 0; p0 < q; p0 += Q) {
                    T x[Q];
#pragma unroll
                    for (int j = 0; j < Q; ++j) x[j] = column[(int64_t)(p0 + j) * n];
                    const T *block = lower + q * kPanel + p0;
#pragma unroll
                    for (int j = 0; j < Q; ++j)
#pragma unroll
                        for (int i = 0; i < Q; ++i) y[i] = eliminated(y[i], block[i * kPanel + j], x[j]);
                }
                const T *block = lower + q * kPanel + q;
#pragma unroll
                for (int i = 1; i < Q; ++i)
#pragma unroll
                    for (int j = 0; j < i; ++j) y[i] = eliminated(y[i], block[i * kPanel + j], y[j]);
#pragma unroll
                for (int i = 0; i < Q; ++i) column[(int64_t)(q + i) * n] = y[i];
            }
        }

        // The same U12 = L11^-1 A12, kPanel columns per workgroup, with the rows split over kU12Blocks
        // blocks of kU12Rows so that eight waves per column pair share the work: thread (q, k) holds rows
        // [q R, q R + R) of column k. With sixteen waves per compute unit instead of one, memory and LDS
        // latencies overlap (fp64 at n = 16384, all of a step's columns: 65 -> 47-53 us). Column c is
        // also stored as row c of the product's packed B operand at hi / lo (rows `stride` elements
        // apart, its kPanel entries first; hi == nullptr stores none), which saves a packing launch on
        // the chain of panels. Block s is solved by its own threads, which leave
        // the solved values in LDS; then every later block subtracts their contribution. Each entry
        // receives exactly solve_u12's operations in solve_u12's order (row r subtracts l[r][s] times
        // row s for s = 0, 1, ..., r - 1), so the results are the same bits; the longest chain of
        // dependent operations is a third of a thread's whole column in solve_u12.
        constexpr int kU12Rows = 16;
        constexpr int kU12Blocks = kPanel / kU12Rows;
        constexpr int kU12NarrowThreads = kPanel * kU12Blocks; // 1024
        template <typename T, typename D, bool split>
        __global__ __launch_bounds__(kU12NarrowThreads) void solve_u12_narrow(const T *__restrict__ l, T *__restrict__ u,
                                                                               int64_t n, int64_t columns, D *hi, D *lo,
                                                                               int64_t stride) {
            constexpr int R = kU12Rows;
            // float: L11 transposed (entry (r, c) at c * kLowerStride + r, rows 16-byte aligned), so that one
            // 16-byte LDS read gives four entries of a column; double: row-major.
            constexpr int kLowerStride = sizeof(T) == 4 ? kPanel + 4 : kPanel;
            __shared__ __attribute__((aligned(16))) T lower[kPanel * kLowerStride];
            __shared__ T solved[R][kPanel];
            const int t = (int)threadIdx.x;
            for (int e = t; e < kPanel * kPanel; e += kU12NarrowThreads) {
                const int r = e / kPanel, c = e % kPanel;
                const T v = c < r ? l[(int64_t)r * n + c] : T(0);
                if constexpr (sizeof(T) == 4) lower[c * kLowerStride + r] = v;
                else lower[e] = v;
            }
            const int k = t % kPanel, q = t / kPanel; // q is the same in a whole wave
            const int64_t index = (int64_t)blockIdx.x * kPanel + k; // the column
            const bool live = index < columns;
            T *__restrict__ column = u + index;
            T y[R];
#pragma unroll
            for (int i = 0; i < R; ++i) y[i] = live ? column[(int64_t)(q * R + i) * n] : T(0);
            __syncthreads();
            for (int s = 0; s < kU12Blocks; ++s) {
                if (q == s) {
                    if constexpr (sizeof(T) == 4) {
                        // by columns: y[i] still takes j = 0, 1, ..., i - 1 in order
                        typedef T Four __attribute__((ext_vector_type(4)));
#pragma unroll
                        for (int j = 0; j + 1 < R; ++j) {
                            const Four *column_j = reinterpret_cast<const Four *>(lower + (s * R + j) * kLowerStride + s * R);
#pragma unroll
                            for (int h = (j + 1) / 4; h < R / 4; ++h) {
                                const Four m = column_j[h];
#pragma unroll
                                for (int e = 0; e < 4; ++e)
                                    if (4 * h + e > j) y[4 * h + e] = eliminated(y[4 * h + e], m[e], y[j]);
                            }
                        }
                    } else {
                        const T *block = lower + (s * R) * kPanel + s * R;
#pragma unroll
                        for (int i = 1; i < R; ++i)
#pragma unroll
                            for (int j = 0; j < i; ++j) y[i] = eliminated(y[i], block[i * kPanel + j], y[j]);
                    }
#pragma unroll
                    for (int i = 0; i < R; ++i) solved[i][k] = y[i];
                    if (live) {
#pragma unroll
                        for (int i = 0; i < R; ++i) column[(int64_t)(s * R + i) * n] = y[i];
                        if (hi)
#pragma unroll
                            for (int i = 0; i < R; ++i) pack_element<D, split>(y[i], hi, lo, index * stride + s * R + i);
                    }
                }
                __syncthreads();
                if (q > s) {
                    // fp64 reads four entries of a row of L at once (at n = 16384, with the packed copy
                    // stored too: 53 -> 38 us per step); float four entries of a column of the transposed
                    // copy (four of a row measured 24 -> 33 us per step). Each y[i] takes j = 0, 1, ... in
                    // order either way.
                    const T *block = lower + (q * R) * kPanel + s * R;
                    if constexpr (sizeof(T) == 8) {
                        typedef T Four __attribute__((ext_vector_type(4)));
#pragma unroll
                        for (int j = 0; j < R; j += 4) {
                            T x[4];
#pragma unroll
                            for (int h = 0; h < 4; ++h) x[h] = solved[j + h][k];
#pragma unroll
                            for (int i = 0; i < R; ++i) {
                                const Four m = *reinterpret_cast<const Four *>(block + i * kPanel + j);
#pragma unroll
                                for (int h = 0; h < 4; ++h) y[i] = eliminated(y[i], m[h], x[h]);
                            }
                        }
                    } else {
                        typedef T Four __attribute__((ext_vector_type(4)));
#pragma unroll
                        for (int j = 0; j < R; ++j) {
                            const T x = solved[j][k];
                            const Four *column_j = reinterpret_cast<const Four *>(lower + (s * R + j) * kLowerStride + q * R);
#pragma unroll
                            for (int h = 0; h < R / 4; ++h) {
                                const Four m = column_j[h];
#pragma unroll
                                for (int e = 0; e < 4; ++e) y[4 * h + e] = eliminated(y[4 * h + e], m[e], x);
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }

        // U12 = L11^-1 A12 for `columns` columns by solve_u12_narrow, packed in mode `mode`'s input type
        // at hi / lo (see there; hi == nullptr: not packed).
        template <typename T>
        hipError_t solve_and_pack_u12(const T *l11, T *a12, int64_t n, int64_t columns, int mode, void *hi, void *lo,
                                      int64_t stride, hipStream_t s) {
            const dim3 grid((unsigned)((columns + kPanel - 1) / kPanel)), block(kU12NarrowThreads);
            if (sizeof(T) == 8 || mode == kFloat32 || !hi) {
                hipLaunchKernelGGL((solve_u12_narrow<T, T, false>), grid, block, 0, s, l11, a12, n, columns,
                                   static_cast<T *>(hi), nullptr, stride);
            } else if constexpr (sizeof(T) == 8) {
                return hipErrorInvalidValue; // fp64 factors pack as fp64 only
            } else if (mode == kBfloat16)
                hipLaunchKernelGGL((solve_u12_narrow<T, __bf16, false>), grid, block, 0, s, l11, a12, n, columns,
                                   static_cast<__bf16 *>(hi), nullptr, stride);
            else
                hipLaunchKernelGGL((solve_u12_narrow<T, __bf16, true>), grid, block, 0, s, l11, a12, n, columns,
                                   static_cast<__bf16 *>(hi), static_cast<__bf16 *>(lo), stride);
            return hipGetLastError();
        }

        // Applies the panel's swaps, in step order, to the columns outside it numbered [from, to) (column
        // c < j0 is number c, column c >= j0 + w number c - w).
        template <typename T>
        __global__ __launch_bounds__(256) void swap_rows(T *a, int64_t n, int64_t j0, int w, const int64_t *pivots,
                                                         int64_t from, int64_t to) {
            const int64_t raw = from + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
            if (raw >= to) return;
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
        // true), one workgroup per block, and stores it transposed: out[b][k][r] = inv(block b)[r][k].
        // Recursive 2 x 2 blocking in place in LDS: the sixteen-row diagonal blocks are inverted first,
        // a thread per row in registers (lower: M L = I row by row, M[i][j] = -(L[i][j] + sum_{j<k<i}
        // M[i][k] L[k][j]); upper: N U = I, N[i][i] = 1 / U[i][i], N[i][j] = -(sum_{i<=k<j} N[i][k]
        // U[k][j]) / U[j][j]); then, for diagonal blocks of 32, 64 and 128 rows, the off-diagonal half
        // becomes -inv(B) C inv(A) (lower, C below the diagonal halves A and B already inverted) or
        // -inv(A) B inv(C) (upper, B right of them), as two products whose sums run over the nonzero
        // triangle only, all threads sharing each level's outputs. (A row's sequential chain of 8,000
        // multiply-adds took about 290 us per block; this takes a few.) L's diagonal, which holds U's in
        // the packed factors, is taken as 1.
        template <typename T, bool upper>
        __global__ __launch_bounds__(256) void invert_diagonal(const T *f, int64_t n, int size, T *out, int first) {
            constexpr int S = kBlock;
            constexpr int stride = sizeof(T) == 4 ? S + 1 : S; // doubles: 128 KB, the padding does not fit
            constexpr int kBase = 16;
            __shared__ T m[S * stride];
            const int64_t r0 = (int64_t)(first + (int)blockIdx.x) * S;
            const int t = threadIdx.x;
            // A block of `size` < kBlock rows is completed by the identity, whose inverse it keeps.
            for (int i0 = 0; i0 < S; i0 += 32) {
                T v[16];
#pragma unroll
                for (int h = 0; h < 16; ++h) {
                    const int i = i0 + 2 * h + (t >> 7), k = t & (S - 1);
                    v[h] = (i < size && k < size) ? f[(r0 + i) * n + r0 + k] : (i == k ? T(1) : T(0));
                }
#pragma unroll
                for (int h = 0; h < 16; ++h) m[(i0 + 2 * h + (t >> 7)) * stride + (t & (S - 1))] = v[h];
            }
            __syncthreads();
            auto at = [&](int i, int k) -> T & { return m[i * stride + k]; };
            // The sixteen-row diagonal blocks, a thread per row: the row of the inverse in registers, the
            // block's other rows read from LDS before anyone writes.
            {
                T row[kBase];
                const int i = t % kBase, base = (t / kBase) * kBase;
                if (t < S) {
#pragma unroll
                    for (int j = 0; j < kBase; ++j) row[j] = T(0);
                    if (!upper) {
                        row[i] = T(1);
#pragma unroll
                        for (int j = kBase - 1; j >= 0; --j) {
                            if (j >= i) continue;
                            T s = at(base + i, base + j);
#pragma unroll
                            for (int k = j + 1; k < kBase; ++k)
                                if (k < i) s = s + row[k] * at(base + k, base + j);
                            row[j] = -s;
                        }
                    } else {
#pragma unroll
                        for (int j = 0; j < kBase; ++j) {
                            if (j < i) continue;
                            if (j == i) {
                                row[j] = T(1) / at(base + j, base + j);
                                continue;
                            }
                            T s = T(0);
#pragma unroll
                            for (int k = 0; k < kBase; ++k)
                                if (k >= i && k < j) s = s + row[k] * at(base + k, base + j);
                            row[j] = -s / at(base + j, base + j);
                        }
                    }
                }
                __syncthreads();
                if (t < S)
#pragma unroll
                    for (int j = 0; j < kBase; ++j)
                        if (upper ? j >= i : j <= i) at(base + i, base + j) = row[j];
                __syncthreads();
            }
            // Levels of 32, 64 and 128 rows: each block's off-diagonal half from its inverted diagonal
            // halves, h x h outputs per block, spread over the threads (consecutive threads consecutive
            // columns), in two products with a barrier after each.
            for (int s = 2 * kBase; s <= S; s *= 2) {
                const int h = s / 2, outputs = (S / s) * h * h; // at most 4096: 16 per thread
                T part[16];
                // first product: lower T = C inv(A) (sum over k >= c, inv(A)[c][c] = 1); upper T = B inv(C)
                // (sum over k <= c)
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int idx = t + 256 * q;
                    if (idx >= outputs) continue;
                    const int o = (idx / (h * h)) * s, r = (idx % (h * h)) / h, c = idx % h;
                    T acc = T(0);
                    if (!upper) {
                        acc = at(o + h + r, o + c);
                        for (int k = c + 1; k < h; ++k) acc = acc + at(o + h + r, o + k) * at(o + k, o + c);
                    } else {
                        for (int k = 0; k <= c; ++k) acc = acc + at(o + r, o + h + k) * at(o + h + k, o + h + c);
                    }
                    part[q] = acc;
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int idx = t + 256 * q;
                    if (idx >= outputs) continue;
                    const int o = (idx / (h * h)) * s, r = (idx % (h * h)) / h, c = idx % h;
                    if (!upper) at(o + h + r, o + c) = part[q];
                    else at(o + r, o + h + c) = part[q];
                }
                __syncthreads();
                // second product: lower X = -inv(B) T (sum over k <= r, inv(B)[r][r] = 1); upper
                // X = -inv(A) T (sum over k >= r)
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int idx = t + 256 * q;
                    if (idx >= outputs) continue;
                    const int o = (idx / (h * h)) * s, r = (idx % (h * h)) / h, c = idx % h;
                    T acc = T(0);
                    if (!upper) {
                        acc = at(o + h + r, o + c);
                        for (int k = 0; k < r; ++k) acc = acc + at(o + h + r, o + h + k) * at(o + h + k, o + c);
                    } else {
                        for (int k = r; k < h; ++k) acc = acc + at(o + r, o + k) * at(o + k, o + h + c);
                    }
                    part[q] = -acc;
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int idx = t + 256 * q;
                    if (idx >= outputs) continue;
                    const int o = (idx / (h * h)) * s, r = (idx % (h * h)) / h, c = idx % h;
                    if (!upper) at(o + h + r, o + c) = part[q];
                    else at(o + r, o + h + c) = part[q];
                }
                __syncthreads();
            }
            T *block = out + (int64_t)(first + (int)blockIdx.x) * S * S;
            for (int idx = t; idx < S * S; idx += 256) {
                const int r = idx % S, k = idx / S; // consecutive threads write consecutive r
                const T value = upper ? (k >= r ? at(r, k) : T(0)) : (k < r ? at(r, k) : (k == r ? T(1) : T(0)));
                block[(int64_t)k * S + r] = value;
            }
        }

        // One triangle of the factors for up to kTriangleColumns right-hand sides, all diagonal blocks in
        // one launch. Four workgroups share each diagonal block, one per quarter (32 rows) of it, so that
        // the product on the chain from block to block is spread over four compute units. Workgroups 4 q
        // .. 4 q + 3 take block j (q for the lower triangle, nb - 1 - q for the upper):
        // - for every block i the block depends on (i < j lower, i > j upper, in the order they complete)
        //   except the last two, the workgroup waits for x_i's publication, then subtracts F_ji x_i over
        //   its quarter's rows;
        // - it publishes its quarter of the result and gathers the other three (published at about the
        //   same time by its siblings), and applies its rows of inv(D_jj);
        // - only then it waits for the last two blocks, whose contributions it subtracts through its rows
        //   of H_j = inv(D_jj) F_j,j-2 and G_j = inv(D_jj) F_j,j-1 (from invert_blocks): one 32-row
        //   product after each wait.
        // Its rows of x_j are stored into x and published in tagged slots (tag: this launch's number) at
        // published[((j * kBlock + r) * kTriangleColumns + c) * kSlots]; the quarters before inv(D_jj)
        // go to the same place in `partial`. A workgroup waits on workgroups of lower index, which are
        // dispatched before it, and on its three siblings, which are dispatched with it. Each output
        // element is the same sequence of matrix-core steps as when one workgroup took the whole block.
        constexpr int kTriangleColumns = 32;
        constexpr int kQuarter = kBlock / 4; // rows of a diagonal block per workgroup
        template <typename T> struct Quad;
        template <> struct Quad<float> {
            typedef float type __attribute__((ext_vector_type(4)));
        };
        template <> struct Quad<double> {
            typedef double type __attribute__((ext_vector_type(4)));
        };
        // One 16 x 16 x 4 matrix-core step, acc += a b (lane l: a = A[l % 16][l / 16], b = B[l / 16][l % 16]).
        template <typename T>
        __device__ __forceinline__ void mfma(typename Quad<T>::type &acc, T a, T b) {
            if constexpr (sizeof(T) == 8) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
            else acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
        }

        // Two workgroups per compute unit for float, so that the 4 x 128 workgroups of order 16384 are
        // resident at once; double keeps one (its registers are full), and its later workgroups start as
        // earlier ones finish.
        template <typename T, bool upper>
        __global__ __launch_bounds__(256, sizeof(T) == 8 ? 1 : 2) void solve_triangle(
            const T *f, const T *inverses_t, const T *links, const T *links2, T *x, int64_t n, int64_t width,
            int64_t col0, unsigned long long *published, unsigned long long *partial, unsigned tag, unsigned *error) {
            typedef typename Quad<T>::type V;
            constexpr int S = kSlots<T>;
            constexpr int C = kTriangleColumns;
            constexpr int kStride = kBlock + 4; // LDS row of fl: 16 lanes of consecutive rows on distinct banks
            __shared__ T xl[kBlock][C + 1];      // x_i, later the block's reduced right-hand side
            __shared__ T fl[kQuarter * kStride]; // this quarter's rows of the 128 x 128 matrix being applied
            const int nb = (int)(n / kBlock), q = (int)(blockIdx.x >> 2), s = (int)(blockIdx.x & 3);
            const int j = upper ? nb - 1 - q : q;
            const int t = threadIdx.x, lane = t & 63, wave = t >> 6, g = lane >> 4, c16 = lane & 15;
            const int cols = (int)min((int64_t)C, width - col0);
            const int64_t r0 = (int64_t)j * kBlock + kQuarter * s; // first row of this quarter
            // Wave w holds one 16 x 16 tile of the quarter's 32 x 32 result (a matrix-core accumulator):
            // rows 16 (w / 2) .., columns 16 (w % 2) ..; element e is row 16 (w / 2) + (fp32: 4 g + e;
            // fp64: g + 4 e), column 16 (w % 2) + c16.
            const int tile_row = 16 * (wave >> 1), col = 16 * (wave & 1) + c16;
            auto row_of = [&](int e) { return tile_row + (sizeof(T) == 8 ? g + 4 * e : 4 * g + e); };
            V acc;
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[e] = col < cols ? x[(r0 + row_of(e)) * width + col0 + col] : T(0);

            // The quarter's rows of a 128 x 128 matrix M whose entry (r, k) is at m[r * row + k * column]
            // (m pointing at the quarter's first row), 16 entries per thread, all loads issued together
            // so that a dependency costs one memory latency. Consecutive threads take consecutive
            // addresses: along k when `row` is the long stride, along r otherwise.
            typedef T Block[16];
            auto fetch = [&](const T *m, int64_t row, int64_t column, Block &block) {
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int idx = t + 256 * e;
                    const int r = column == 1 ? idx >> 7 : idx & (kQuarter - 1);
                    const int k = column == 1 ? idx & (kBlock - 1) : idx >> 5;
                    block[e] = m[r * row + (int64_t)k * column];
                }
            };
            // M from fetch() (the same strides) into fl, then into += sign * M xl over k = 0 .. kBlock - 1.
            auto stage = [&](const Block &block, int64_t column) {
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int idx = t + 256 * e;
                    const int r = column == 1 ? idx >> 7 : idx & (kQuarter - 1);
                    const int k = column == 1 ? idx & (kBlock - 1) : idx >> 5;
                    fl[r * kStride + k] = block[e];
                }
                __syncthreads();
            };
            auto multiply = [&](V &into, T sign) {
#pragma unroll 8
                for (int kk = 0; kk < kBlock; kk += 4)
                    mfma<T>(into, sign * fl[(tile_row + c16) * kStride + kk + g], xl[kk + g][col]);
                __syncthreads();
            };
            auto apply = [&](const Block &block, int64_t column, V &into, T sign) {
                stage(block, column);
                multiply(into, sign);
            };

            // The 128 x 32 block i of `from` into xl once published, 16 entries per thread with their
            // loads in flight together; entries of rows in quarter `own` (0-3) are left to the caller.
            auto gather = [&](const unsigned long long *from, int i, int own) {
                unsigned long long slot[16][2];
                unsigned waiting = 0;
#pragma unroll
                for (int e = 0; e < 16; ++e) // entry e is row (t + 256 e) / C: quarter e / 4
                    if ((t + 256 * e) % C < cols && (e >> 2) != own) waiting |= 1u << e;
                unsigned spins = 0;
                while (waiting) {
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        if (!(waiting >> e & 1u)) continue;
                        const int idx = t + 256 * e, r = idx / C, c = idx % C;
                        const unsigned long long *at = from + (((int64_t)i * kBlock + r) * C + c) * S;
#pragma unroll
                        for (int h = 0; h < S; ++h)
                            slot[e][h] = __hip_atomic_load(at + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        if (!(waiting >> e & 1u) || !current<T>(slot[e], tag)) continue;
                        waiting &= ~(1u << e);
                        const int idx = t + 256 * e;
                        xl[idx / C][idx % C] = value_of<T>(slot[e]);
                    }
                    if (!waiting) break;
                    if (++spins > kSpinLimit) {
                        __hip_atomic_fetch_or(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    if ((spins & 63u) == 0 && relaxed(error)) break;
                    __builtin_amdgcn_s_sleep(1);
                }
                if (cols < C)
                    for (int idx = t; idx < kBlock * C; idx += 256)
                        if (idx % C >= cols) xl[idx / C][idx % C] = T(0);
                __syncthreads();
            };
            auto depend = [&](int i) {
                Block block;
                fetch(f + r0 * n + (int64_t)i * kBlock, n, 1, block); // F_ji, rows n apart; in flight while x_i is awaited
                gather(published, i, -1);
                apply(block, 1, acc, T(-1));
            };

            // Every dependency but the two that complete last (blocks j - 2 and j - 1, or j + 2 and j + 1
            // for the upper triangle), whose contributions go through links: H_j = inv(D_jj) F_j,j-2 and
            // G_j = inv(D_jj) F_j,j-1. Then the inverse and the first link are applied while block j - 1
            // is still being solved, and only G_j's product follows its publication.
            if (!upper)
                for (int i = 0; i + 2 < j; ++i) depend(i);
            else
                for (int i = nb - 1; i > j + 2; --i) depend(i);
            const int last = upper ? j + 1 : j - 1, second = upper ? j + 2 : j - 2;
            const bool linked = upper ? j + 1 < nb : j > 0, linked2 = upper ? j + 2 < nb : j > 1;

            // The block's reduced right-hand side: this quarter from acc (published for the siblings),
            // the others gathered from theirs.
            Block block;
            // this quarter's rows of inv(D_jj) (given transposed, inv[r][k] at inverses_t[k * kBlock + r])
            fetch(inverses_t + (int64_t)j * kBlock * kBlock + kQuarter * s, 1, kBlock, block);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = row_of(e);
                xl[kQuarter * s + r][col] = acc[e];
                if (col < cols) put_tagged(partial + ((r0 + r) * C + col) * S, acc[e], tag);
            }
            gather(partial, j, s);
            // out = inv(D_jj) acc, then out -= H_j x_second, then out -= G_j x_last (row-major links)
            V out = V{T(0), T(0), T(0), T(0)};
            apply(block, kBlock, out, T(1));
            // The links are staged in LDS before their blocks are awaited, so that only the product
            // follows the last wait.
            if (linked2) {
                fetch(links2 + ((int64_t)j * kBlock + kQuarter * s) * kBlock, kBlock, 1, block);
                stage(block, 1);
                gather(published, second, -1);
                multiply(out, T(-1));
            }
            if (linked) {
                fetch(links + ((int64_t)j * kBlock + kQuarter * s) * kBlock, kBlock, 1, block);
                stage(block, 1);
                gather(published, last, -1);
                multiply(out, T(-1));
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = row_of(e);
                if (col >= cols) continue;
                x[(r0 + r) * width + col0 + col] = out[e];
                put_tagged(published + ((r0 + r) * C + col) * S, out[e], tag);
            }
        }

        // inv(D_jj) F_j,j-d (lower) or inv(D_jj) F_j,j+d (upper) for d = `distance` (1: G_j, 2: H_j),
        // row-major, for solve_triangle; workgroup j computes block j's 128 x 128 entries, thread t
        // column t % kBlock of rows [64 (t / kBlock), + 64), over 32-row slices of both operands staged
        // in LDS (the inverse read where all lanes share an entry). Each entry is the fma chain over
        // m = 0, 1, ..., kBlock - 1.
        template <typename T, bool upper, int distance>
        __global__ __launch_bounds__(256) void link_blocks(const T *f, int64_t n, const T *inverses_t, T *links,
                                                           int first) {
            constexpr int M = 32;
            __shared__ T inverse_s[M][kBlock];
            __shared__ T block_s[M][kBlock];
            const int nb = (int)(n / kBlock), j = first + (int)blockIdx.x;
            const int neighbour = upper ? j + distance : j - distance;
            if (neighbour < 0 || neighbour >= nb) return;
            const int t = (int)threadIdx.x, k = t % kBlock, r0 = 64 * (t / kBlock);
            const T *inverse = inverses_t + (int64_t)j * kBlock * kBlock; // inverse[m * kBlock + r] = inv[r][m]
            const T *block = f + (int64_t)j * kBlock * n + (int64_t)neighbour * kBlock;
            T acc[64];
#pragma unroll
            for (int i = 0; i < 64; ++i) acc[i] = T(0);
            for (int m0 = 0; m0 < kBlock; m0 += M) {
                // all of a slice's loads in flight before any is stored
                T iv[M * kBlock / 256], bv[M * kBlock / 256];
#pragma unroll
                for (int h = 0; h < M * kBlock / 256; ++h) {
                    const int e = t + 256 * h, mm = e / kBlock, c = e % kBlock;
                    iv[h] = inverse[(int64_t)(m0 + mm) * kBlock + c];
                    bv[h] = block[(int64_t)(m0 + mm) * n + c];
                }
#pragma unroll
                for (int h = 0; h < M * kBlock / 256; ++h) {
                    const int e = t + 256 * h, mm = e / kBlock, c = e % kBlock;
                    inverse_s[mm][c] = iv[h];
                    block_s[mm][c] = bv[h];
                }
                __syncthreads();
                for (int mm = 0; mm < M; ++mm) {
                    const T b = block_s[mm][k];
#pragma unroll
                    for (int i = 0; i < 64; ++i) {
                        if constexpr (sizeof(T) == 8) acc[i] = __builtin_fma(inverse_s[mm][r0 + i], b, acc[i]);
                        else acc[i] = __builtin_fmaf(inverse_s[mm][r0 + i], b, acc[i]);
                    }
                }
                __syncthreads();
            }
#pragma unroll
            for (int i = 0; i < 64; ++i) links[(int64_t)j * kBlock * kBlock + (int64_t)(r0 + i) * kBlock + k] = acc[i];
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

        // Slice blockIdx.y of R = A X for a narrow X (n x w, w <= 32, row-major) and an order that is a
        // multiple of kBlock: workgroup (b, s) multiplies rows [128 b, 128 b + 128) of A by X over the
        // s-th share of the kBlock-wide column blocks of A, on the fp64 matrix cores (each wave 32 rows
        // x 32 columns, as in solve_triangle), and writes its partial product to partials[s] (n x w).
        // A is read in place, row-major, once: the residual of a refinement is bound by that read,
        // where a general product's 128-column tile would spend four times the matrix-core work on
        // the padding of X's 32 columns. Each 32-column chunk of the next A block is loaded as soon as
        // the same chunk of the current one is in LDS, and X's next rows while the current are used.
        __global__ __launch_bounds__(256) void residual_slices(const double *a, int64_t n, const double *x, int64_t w,
                                                               double *partials) {
            typedef typename Quad<double>::type V;
            __shared__ double xl[kBlock][33];
            __shared__ double fl[kBlock][33];
            const int t = (int)threadIdx.x, lane = t & 63, wave = t >> 6, g = lane >> 4, c16 = lane & 15;
            const int64_t r0 = (int64_t)blockIdx.x * kBlock;
            const int64_t blocks = n / kBlock, first = blocks * blockIdx.y / gridDim.y,
                          end = blocks * (blockIdx.y + 1) / gridDim.y;
            const int cols = (int)min((int64_t)32, w);
            auto row_of = [&](int i, int e) { return 32 * wave + 16 * i + g + 4 * e; };
            V acc[2][2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int h = 0; h < 2; ++h) acc[i][h] = V{0.0, 0.0, 0.0, 0.0};
            double block[4][16], xs[16];
            auto fetch_chunk = [&](int64_t kb, int c) {
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int idx = t + 256 * e, r = idx >> 5, k = idx & 31;
                    block[c][e] = a[(r0 + r) * n + kb * kBlock + 32 * c + k];
                }
            };
            auto fetch_x = [&](int64_t kb) {
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int idx = t + 256 * e, k = idx >> 5, c = idx & 31;
                    xs[e] = c < cols ? x[(kb * kBlock + k) * w + c] : 0.0;
                }
            };
            if (first < end) {
#pragma unroll
                for (int c = 0; c < 4; ++c) fetch_chunk(first, c);
                fetch_x(first);
            }
            for (int64_t kb = first; kb < end; ++kb) {
                const bool more = kb + 1 < end;
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int idx = t + 256 * e;
                    xl[idx >> 5][idx & 31] = xs[e];
                }
                if (more) fetch_x(kb + 1);
#pragma unroll
                for (int c = 0; c < 4; ++c) {
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        const int idx = t + 256 * e;
                        fl[idx >> 5][idx & 31] = block[c][e];
                    }
                    __syncthreads();
                    if (more) fetch_chunk(kb + 1, c);
#pragma unroll
                    for (int kk = 0; kk < 32; kk += 4) {
                        double am[2], bm[2];
#pragma unroll
                        for (int i = 0; i < 2; ++i) am[i] = fl[32 * wave + 16 * i + c16][kk + g];
#pragma unroll
                        for (int h = 0; h < 2; ++h) bm[h] = xl[32 * c + kk + g][16 * h + c16];
#pragma unroll
                        for (int i = 0; i < 2; ++i)
#pragma unroll
                            for (int h = 0; h < 2; ++h) mfma<double>(acc[i][h], am[i], bm[h]);
                    }
                    __syncthreads();
                }
            }
            double *out = partials + (int64_t)blockIdx.y * n * w;
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int c = 16 * h + c16;
                        if (c < cols) out[(r0 + row_of(i, e)) * w + c] = acc[i][h][e];
                    }
        }

        // r -= the sum of the slices' partial products, in slice order.
        __global__ __launch_bounds__(256) void subtract_slices(const double *partials, double *r, int64_t count,
                                                               int slices) {
            const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
            if (idx >= count) return;
            double sum = partials[idx];
            for (int s = 1; s < slices; ++s) sum += partials[(int64_t)s * count + idx];
            r[idx] = -sum + r[idx];
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
        // L21 and U12 packed (up to 8 bytes, split pairs), then two sets of the panel's candidate lines
        // (kParities buffers, one line of up to kStripLineMax slots per workgroup of at least 128 rows).
        // A second set of L21 buffers after them (alternate steps use alternate sets).
        return 6 * (size_t)order * kPanel * 8 + 2 * kParities * ((size_t)order / kPanelRows + 2) * kStripLineMax * 8 +
               1024;
    }

    size_t panel_publish_bytes(int64_t order, size_t element) {
        const size_t groups = (size_t)((order + kPanelRows - 1) / kPanelRows);
        const size_t slots = element / 4; // tagged 64-bit slots per value
        // panel(): candidate rows and the diagonal row; panel_strips(): the rest buffers after them
        return kParities * (groups + 1) * kPanel * slots * 8 + (size_t)kRestSlots * 2 * 16 * kPanel * slots * 8;
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

    namespace {
        // invert_blocks for diagonal blocks [first, first + count): L's inverses and links (lower), U's
        // (upper).
        template <typename T>
        hipError_t invert_range(const T *f, int64_t n, T *inverses, int first, int count, hipStream_t stream,
                                bool lower = true, bool upper = true) {
            if (count <= 0) return hipSuccess;
            const dim3 grid((unsigned)count);
            if (lower) {
                hipLaunchKernelGGL((invert_diagonal<T, false>), grid, dim3(256), 0, stream, f, n, kBlock, inverses,
                                   first);
                hipLaunchKernelGGL((link_blocks<T, false, 1>), grid, dim3(256), 0, stream, f, n, inverses,
                                   inverses + 2 * n * kBlock, first);
                hipLaunchKernelGGL((link_blocks<T, false, 2>), grid, dim3(256), 0, stream, f, n, inverses,
                                   inverses + 4 * n * kBlock, first);
            }
            if (upper) {
                hipLaunchKernelGGL((invert_diagonal<T, true>), grid, dim3(256), 0, stream, f, n, kBlock,
                                   inverses + n * kBlock, first);
                hipLaunchKernelGGL((link_blocks<T, true, 1>), grid, dim3(256), 0, stream, f, n, inverses + n * kBlock,
                                   inverses + 3 * n * kBlock, first);
                hipLaunchKernelGGL((link_blocks<T, true, 2>), grid, dim3(256), 0, stream, f, n, inverses + n * kBlock,
                                   inverses + 5 * n * kBlock, first);
            }
            return hipGetLastError();
        }
    } // namespace

    // The share of each fp64 update that runs beside the next panel: kPanelMsFp64 milliseconds of work
    // at kAlongsideRateFp64, the rate measured for an fp64 update beside a panel on MI355X. An fp64
    // panel takes about 0.6 ms at n = 16384, but the fp64 solve there measured best with 0.4 ms of
    // work beside it (0.1 to 1.1 ms swept: 158, 153, 148, 146, 150, 155, 161 ms at 0.1, 0.2, 0.3,
    // 0.4, 0.6, 0.8, 1.1): the update beside the panel slows it down.
    constexpr double kPanelMsFp64 = 0.4;
    constexpr double kAlongsideRateFp64 = 25e12;
    // Float steps pair while this many columns or more remain: in later steps the update hides
    // beside the panel anyway, and a pair only adds its repacking and 128-row product to the chain of
    // panels. At n = 16384 (threshold 0 / 4096 / 8192 / 12288) the fp32 solve measured 104.2 / 103.8 /
    // 103.3 / 104.4 ms and the bf16 one 94.0 / 92.6 / 92.6 / 92.8 ms (mean of two runs on one box).
    constexpr int64_t kFloatPairColumns = 8192;
    // A float pair's update splits while this many columns or more remain beyond the next panel's.
    constexpr int64_t kFloatSplitColumns = 10240;

    // Panels before the end from which the diagonal blocks done so far are inverted on the third stream
    // (the last panels leave most compute units idle; the blocks after are inverted after the last one).
    constexpr int kInvertTail = 24;

    template <typename T>
    int factor(T *a, int64_t n, int64_t *pivots, int mode, Workspace &ws, void *scratch, size_t scratch_bytes,
               hipStream_t stream, const char **error, T *inverses) {
        *error = nullptr;
        if (inverses && (!blocked_substitution(n) || !ws.third || !ws.inverted || !ws.upper_inverted)) {
            *error = "LU inverses need an order that is a multiple of 128 (at least 256) and the third stream";
            return -1;
        }
        const int blocks_total = (int)(n / kBlock);
        const int early_blocks = inverses ? std::max(0, blocks_total - kInvertTail) : 0; // inverted on the third stream
        if (scratch_bytes < factor_scratch(n, mode)) {
            *error = "LU scratch too small";
            return -1;
        }
        const bool wide = PackMode<T>::wide();
        const size_t kp = kPanel;
        unsigned char *base = static_cast<unsigned char *>(scratch);
        const size_t operand = (size_t)n * kp * 8; // bytes of one packed n x kPanel operand, at most
        // Packed L21, two sets used by alternate steps: a panel packs its own L21 (float factors with a
        // bf16 or bf16x3 update) while the update beside it still reads the previous step's.
        const size_t line_region = 2 * kParities * ((size_t)n / kPanelRows + 2) * kStripLineMax * 8;
        void *const la_his[2] = {base, base + 4 * operand + line_region};
        void *const la_los[2] = {base + operand, base + 5 * operand + line_region};
        void *ub_hi = base + 2 * operand, *ub_lo = base + 3 * operand;
        unsigned long long *lines = reinterpret_cast<unsigned long long *>(base + 4 * operand);
        const int max_groups = (int)((n + kPanelRows - 1) / kPanelRows);
        if (ws.publish_bytes < panel_publish_bytes(n, sizeof(T))) {
            *error = "LU publication buffer too small";
            return -1;
        }
        unsigned long long *published = static_cast<unsigned long long *>(ws.publish);
        unsigned long long *diagonal = published + (size_t)kParities * max_groups * kPanel * kSlots<T>;
        unsigned long long *rests = published + (size_t)kParities * (max_groups + 1) * kPanel * kSlots<T>;
        // SCIHIP_LU_PANEL=strips: the strip panel kernel (panel_strips) instead of the row-wise one (panel).
        // Both give the same bits; the row-wise one is faster at n = 16384 (fp32 panels 455 against 709 us,
        // fp64 555 against 735 us on MI355X), so it stays the default.
        const char *panel_text = std::getenv("SCIHIP_LU_PANEL");
        const bool strips = panel_text && std::strcmp(panel_text, "strips") == 0;

        StageTimer timer(stream);
        // Panel workgroups the device holds at once, and how many panel workgroups are confined to one
        // XCD (SCIHIP_LU_CONFINE=g: panels of at most g workgroups, capped at what one XCD holds).
        int resident = 0, confine = 0, confine_limit = 0;
        if (const char *text = std::getenv("SCIHIP_LU_CONFINE")) confine_limit = std::max(0, std::atoi(text));
        {
            int device = 0, units = 0, per_unit = 0;
            hipError_t e = hipGetDevice(&device);
            if (e == hipSuccess) e = hipDeviceGetAttribute(&units, hipDeviceAttributeMultiprocessorCount, device);
            if (e == hipSuccess)
                e = strips ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                     &per_unit, reinterpret_cast<const void *>(&panel_strips<T>), kStripThreads, 0)
                           : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                     &per_unit, reinterpret_cast<const void *>(&panel<T>), kPanelThreads, 0);
            if (e != hipSuccess) return (int)e;
            resident = units * per_unit;
            confine = std::min(units / 8 * per_unit, confine_limit);
        }
        // Clears candidate line set `set` on stream `s` (each launch expects its lines zero). Panels
        // run on the main stream and alternate between the sets; the set of the panel after next is
        // cleared on the side stream meanwhile.
        const size_t line_bytes = (size_t)kParities * (max_groups + 1) * kStripLineMax * sizeof(unsigned long long);
        unsigned long long *line_sets[2] = {lines, lines + line_bytes / sizeof(unsigned long long)};
        auto clear_lines = [&](hipStream_t s, int set) { return hipMemsetAsync(line_sets[set], 0, line_bytes, s); };
        // The row swaps of the panel at j0 (width w) applied to the columns outside it numbered [from, to)
        // (see swap_rows), on stream `s`.
        auto swap_columns = [&](int64_t j0, int w, int64_t from, int64_t to, hipStream_t s) -> hipError_t {
            if (to <= from) return hipSuccess;
            hipLaunchKernelGGL(swap_rows<T>, dim3((unsigned)((to - from + 255) / 256)), dim3(256), 0, s, a, n, j0, w,
                               (const int64_t *)pivots, from, to);
            return hipGetLastError();
        };
        // Whether the panel at column j0 also stores its L21 packed (into L21 set (j0 / kPanel) & 1): for
        // a float factorisation with a bf16 or bf16x3 update, unless its step pairs with the previous one
        // (a pair packs both steps' L21 after its row swaps) or nothing follows the panel.
        auto packs_l21 = [&](int64_t j0, bool pairs) {
            return sizeof(T) == 4 && (mode == kBfloat16 || mode == kSplit) && !strips && !pairs && n - j0 > kPanel;
        };
        // Factors the panel at column j0 on stream `s` with candidate line set `set` (already clear);
        // `pack`: also store its L21 packed (packs_l21).
        auto launch_panel = [&](int64_t j0, hipStream_t s, int set, bool pack) -> hipError_t {
            unsigned long long *lines = line_sets[set];
            __bf16 *pack_hi = pack ? static_cast<__bf16 *>(la_his[(j0 / kPanel) & 1]) : nullptr;
            __bf16 *pack_lo = pack && mode == kSplit ? static_cast<__bf16 *>(la_los[(j0 / kPanel) & 1]) : nullptr;
            int w = (int)std::min<int64_t>(kPanel, n - j0);
            const int groups = (int)((n - j0 + kPanelRows - 1) / kPanelRows);
            hipError_t e = hipSuccess;
            // The launch number goes into the publications' tags (24 bits); when it wraps, the buffers
            // are cleared so that no slot of an earlier launch carries a tag that is current again.
            if (++ws.launches > 0xFFFFFFu) {
                e = hipMemsetAsync(ws.publish, 0, ws.publish_bytes, s);
                if (e != hipSuccess) return e;
                ws.launches = 1;
            }
            unsigned launch = ws.launches;
            if (strips) {
                const int strip_groups = (int)((n - j0 + kStripRows<T> - 1) / kStripRows<T>);
                if (strip_groups <= resident) {
                    hipLaunchKernelGGL((panel_strips<T>), dim3(strip_groups), dim3(kStripThreads), 0, s, a, n, j0, w,
                                       pivots, lines, rests, launch, ws.error);
                } else {
                    void *strip_args[] = {&a, &n, (void *)&j0, (void *)&w, &pivots, &lines, &rests, &launch, &ws.error};
                    e = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&panel_strips<T>), dim3(strip_groups),
                                                   dim3(kStripThreads), strip_args, 0, s);
                }
                if (e == hipSuccess) e = hipGetLastError();
                return e;
            }
            int spread = 1;
            void *args[] = {&a, &n, (void *)&j0, (void *)&w, &pivots, &lines, &published, &diagonal, &launch,
                            &ws.error, &spread, &pack_hi, &pack_lo};
            // The workgroups wait for each other, so all of them must be resident at once. When the grid
            // fits the device, a plain launch is enough: workgroups queued behind the concurrent update
            // start as soon as its workgroups finish, which they do without waiting on the panel (and
            // every wait is bounded). Larger grids take the cooperative launch, which checks residency.
            if (groups <= confine) // one XCD holds them all (every XCD has units / 8 CUs)
                hipLaunchKernelGGL((panel<T>), dim3(groups * 8), dim3(kPanelThreads), 0, s, a, n, j0, w, pivots,
                                   lines, published, diagonal, launch, ws.error, 8, pack_hi, pack_lo);
            else if (groups <= resident)
                hipLaunchKernelGGL((panel<T>), dim3(groups), dim3(kPanelThreads), 0, s, a, n, j0, w, pivots, lines,
                                   published, diagonal, launch, ws.error, 1, pack_hi, pack_lo);
            else
                e = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&panel<T>), dim3(groups),
                                               dim3(kPanelThreads), args, 0, s);
            if (e == hipSuccess) e = hipGetLastError();
            return e;
        };
        const size_t eb = (size_t)element_bytes(mode);

        // SCIHIP_LU_PAIR=c (for tests at small orders): every step pairs while c or more columns remain,
        // and an fp64 step leaves at least half its columns to the next one.
        int64_t pair_columns = kFloatPairColumns;
        // SCIHIP_LU_SPLIT=c: a float pair's update splits (step_float) while c or more columns remain beyond
        // the next panel's (0: never).
        int64_t split_columns = kFloatSplitColumns;
        if (const char *text = std::getenv("SCIHIP_LU_SPLIT")) split_columns = std::strtoll(text, nullptr, 10);
        const char *pair_text = std::getenv("SCIHIP_LU_PAIR");
        const bool forced = pair_text != nullptr;
        if (forced) pair_columns = std::strtoll(pair_text, nullptr, 10);
        // fp64 steps in pairs. The trailing update, a product of depth 128, reads and writes all of C
        // for 2 x 128 flops per element; each product workgroup loads and stores its C tile around a
        // loop of only 32 matrix-core steps, so the C traffic adds to the math rather than hiding
        // under it, and a product of depth 256 does twice the work per C pass. So a
        // step leaves most of its update undone ("carried"): it updates the next panel's columns, and,
        // beside that panel, a panel's time worth of the last columns; the columns between wait. The
        // next step brings its own panel rows up to date with the carried update, solves its U12, and
        // updates the waiting columns with both steps' updates as one product of depth 256 - [L21 of
        // both] (adjacent columns of the matrix, read in place) times [U12 of both] (the two halves of
        // 256-element packed rows, the two steps' columns aligned) - for half the C traffic; the
        // columns already done get its own update alone. The row swaps in between act on the carried
        // L21's rows and on C alike, so the pair stays consistent.
        bool carry = false;
        int64_t carried_end = 0; // carried: columns [next panel's end, carried_end) wait for the update
        auto step_pair = [&](auto j0, int w, int64_t rest) -> hipError_t { // generic: compiled for double only
            constexpr int64_t kPair = 2 * kPanel; // elements per packed U12 row
            T *ub = static_cast<T *>(ub_hi); // rest x kPair; spans the ub_hi and ub_lo regions (fp64 only)
            const int64_t next = std::min<int64_t>(kPanel, rest);
            const T *l11 = a + j0 * n + j0;
            T *a12 = a + j0 * n + j0 + w;
            T *c = a + (j0 + w) * n + (j0 + w);
            const T *l21 = a + (j0 + w) * n + j0;          // this step's L21, in place
            const int64_t waiting = carry ? carried_end - (j0 + w) : 0; // carried columns [0, waiting) of this step
            hipError_t e = hipSuccess;
            if (carry) { // this panel's rows, carried columns: the previous step's update
                ProductArgs slab{a + j0 * n + (j0 - kPanel), nullptr, ub + kPanel * kPair, nullptr, a12, w, waiting,
                                 kPanel, n, kPair, n, -1.0, 1.0};
                e = product(kFloat64, slab, stream);
                if (e != hipSuccess) return e;
            }
            // U12, solved in place and packed: this step's column t goes to packed row t (first half),
            // or, after a carry, to row kPanel + t (second half), beside the previous step's U12 of the
            // same matrix column.
            T *u = carry ? ub + kPanel * kPair : ub; // packed row of this step's first column
            e = solve_and_pack_u12(l11, a12, n, rest, kFloat64, carry ? u + kPanel : u, nullptr, kPair, stream);
            if (e != hipSuccess) return e;
            timer.stage(2);
            // A product over this step's columns [from, from + count): depth 256 (both steps) or 128
            // (this step's update alone).
            // Beside the panel (`alongside`, on the side stream) it is sized for one workgroup per compute
            // unit; after the panel, on the main stream, for two.
            auto update = [&](int64_t from, int64_t count, bool both, bool alongside) -> hipError_t {
                if (count <= 0) return hipSuccess;
                const T *left = both ? l21 - kPanel : l21;
                const T *right = u + from * kPair + (carry && !both ? kPanel : 0);
                ProductArgs part{left, nullptr, right, nullptr, c + from, rest, count, both ? kPair : kPanel, n,
                                 kPair, n, -1.0, 1.0};
                return alongside ? product_alongside(kFloat64, part, ws.side) : product(kFloat64, part, stream);
            };
            e = update(0, next, carry, false); // the next panel's columns (carried ones, after a carry)
            // The next panel follows on the main stream at once; the update beside it runs on the side
            // stream, whose end the next step waits for (panel_done marks it).
            const int set = (int)((j0 + w) / kPanel) & 1;
            if (e == hipSuccess) e = hipEventRecord(ws.columns_ready, stream);
            if (e == hipSuccess) e = launch_panel(j0 + w, stream, set, false);
            if (e == hipSuccess) e = hipStreamWaitEvent(ws.side, ws.columns_ready, 0);
            if (e == hipSuccess) e = clear_lines(ws.side, set ^ 1); // for the panel after next
            if (e != hipSuccess) return e;
            // Beside the panel, at one workgroup per compute unit, about a panel's time of depth-128
            // work (a share of the columns from the measured rates).
            auto panel_share = [&](int64_t columns) -> int64_t {
                const double flops = 2.0 * (double)rest * (double)columns * (double)kPanel;
                const double share = kPanelMsFp64 * 1e-3 * kAlongsideRateFp64 / flops;
                if (share >= 1.0) return columns;
                return std::min(columns, std::max<int64_t>(kPanel, (int64_t)(share * (double)columns) / kPanel * kPanel));
            };
            if (carry) {
                // Columns already carrying the previous update: this step's alone, beside the panel; then
                // the waiting ones with both steps', after the panel at two workgroups per compute unit.
                e = update(waiting, rest - waiting, false, true);
                if (e == hipSuccess && waiting > next) e = update(next, waiting - next, true, false);
                carry = false;
            } else if (rest - next >= kPanel &&
                       (forced ? rest >= pair_columns : panel_share(rest - next) < rest - next)) {
                // Carry: the last columns beside the panel, leaving at least the following panel's
                // columns (and all between) to the next step.
                int64_t beside = std::min(panel_share(rest - next), rest - next - kPanel);
                if (forced) beside = std::min(beside, (rest - next - kPanel) / 2 / kPanel * kPanel);
                e = update(rest - beside, beside, false, true);
                carry = true;
                carried_end = j0 + w + rest - beside;
            } else if (rest > next) {
                // All of it fits beside the panel (late steps), or too few columns are left to carry.
                e = update(next, rest - next, false, true);
            }
            if (e == hipSuccess) e = hipEventRecord(ws.panel_done, ws.side);
            return e;
        };

        // Float steps in pairs too. A float factorisation waits on the chain of panels, except in its
        // first steps, where each trailing update (all of C read and written) outlasts the panel beside
        // it. So a fresh step updates only the next panel's columns, which lets that panel run alone,
        // and carries the rest; the next step brings its panel rows up to date (a 128-row product),
        // then updates all its columns with both steps' updates as products of depth 256 beside its
        // panel: [L21 of both] (in place for fp32; packed after that step's row swaps otherwise, with
        // the panel rows first for the 128-row product) times [U12 of both] (packed as for fp64). In
        // either step the next panel's U12 columns are solved first (the row-parallel kernel), and the
        // rest while that panel is factored. A fresh step with fewer than two panels of columns left
        // updates them all itself.
        auto step_float = [&](auto j0, int w, int64_t rest) -> hipError_t { // generic: compiled for float only
            constexpr int64_t kPair = 2 * kPanel; // elements per packed row of a pair
            const int64_t next = std::min<int64_t>(kPanel, rest);
            const T *l11 = a + j0 * n + j0;
            T *a12 = a + j0 * n + j0 + w;
            T *c = a + (j0 + w) * n + (j0 + w);
            const bool in_place = mode == kFloat32;
            const bool split = mode == kSplit;
            void *const la_hi = la_his[(j0 / kPanel) & 1], *const la_lo = la_los[(j0 / kPanel) & 1];
            // Element `index` of a packed buffer (nullptr stays nullptr).
            auto at = [&](void *base, int64_t index) -> void * {
                return base ? static_cast<void *>(static_cast<unsigned char *>(base) + (size_t)index * eb) : nullptr;
            };
            const bool pair = carry;
            hipError_t e = hipSuccess;
            if (pair) {
                // L21 of both steps, rows from this panel's first (A's columns j0 - kPanel .. j0 + kPanel).
                if (!in_place)
                    e = pack(mode, a + j0 * n + (j0 - kPanel), wide, rest + w, kPair, n, false, la_hi, la_lo, kPair,
                             stream);
                // This panel's rows, all this step's columns: the previous step's update.
                if (e == hipSuccess) {
                    ProductArgs slab{in_place ? static_cast<const void *>(a + j0 * n + (j0 - kPanel)) : la_hi,
                                     split ? la_lo : nullptr, at(ub_hi, kPanel * kPair), split ? at(ub_lo, kPanel * kPair) : nullptr,
                                     a12, w, rest, kPanel, in_place ? n : kPair, kPair, n, -1.0, 1.0};
                    e = product(mode, slab, stream);
                }
                if (e != hipSuccess) return e;
            } else if (!in_place && !packs_l21(j0, false)) { // (else the panel packed its L21)
                e = pack(mode, a + (j0 + w) * n + j0, wide, rest, w, n, false, la_hi, la_lo, kp, stream);
                if (e != hipSuccess) return e;
            }
            // The left operand over C's rows: both steps' L21 (depth 256) or this step's (depth 128).
            const void *left = in_place ? static_cast<const void *>(a + (j0 + w) * n + (pair ? j0 - kPanel : j0))
                                        : pair ? at(la_hi, w * kPair) : la_hi;
            const void *left_lo = split ? (pair ? at(la_lo, w * kPair) : la_lo) : nullptr;
            const int64_t left_ld = in_place ? n : pair ? kPair : (int64_t)kp;
            const int64_t depth = pair ? kPair : kPanel;
            // This step's U12 column t goes to packed row t (first half), or, in a pair, to row kPanel + t
            // (second half), beside the previous step's U12 of the same matrix column.
            const int64_t row0 = pair ? kPanel : 0, half = pair ? kPanel : 0;
            auto pack_u12 = [&](int64_t from, int64_t count, hipStream_t s) -> hipError_t {
                const int64_t index = (row0 + from) * kPair + half;
                return pack(mode, a12 + from, wide, count, w, n, true, at(ub_hi, index), split ? at(ub_lo, index) : nullptr,
                            kPanel, s, kPair);
            };
            // C's columns [from, from + count) -= left times the packed U12 rows of those columns.
            auto update = [&](int64_t from, int64_t count, hipStream_t s) -> hipError_t {
                const int64_t index = (row0 + from) * kPair;
                ProductArgs part{left, left_lo, at(ub_hi, index), split ? at(ub_lo, index) : nullptr, c + from, rest,
                                 count, depth, left_ld, kPair, n, -1.0, 1.0};
                return from == 0 ? product(mode, part, s) : product_alongside(mode, part, s);
            };
            // The next panel's U12 columns, solved in place and packed.
            e = solve_and_pack_u12(l11, a12, n, next, mode, at(ub_hi, row0 * kPair + half),
                                   split ? at(ub_lo, row0 * kPair + half) : nullptr, kPair, stream);
            timer.stage(2);
            if (e == hipSuccess) e = update(0, next, stream); // the next panel's columns
            // The next panel follows on the main stream at once; the rest of the step runs beside it on
            // the side stream, whose end the next step waits for (panel_done marks it).
            const int set = (int)((j0 + w) / kPanel) & 1;
            const bool next_pairs = !pair && rest - next >= kPanel && rest >= pair_columns; // the next step pairs
            if (e == hipSuccess) e = hipEventRecord(ws.columns_ready, stream);
            if (e == hipSuccess) e = launch_panel(j0 + w, stream, set, packs_l21(j0 + w, next_pairs));
            if (e == hipSuccess) e = hipStreamWaitEvent(ws.side, ws.columns_ready, 0);
            if (e == hipSuccess) e = clear_lines(ws.side, set ^ 1); // for the panel after next
            if (!pair) { // the columns the main stream did not swap (the loop below swapped the next panel's)
                if (e == hipSuccess) e = swap_columns(j0, w, 0, j0, ws.side);
                if (e == hipSuccess) e = swap_columns(j0, w, j0 + next, n - w, ws.side);
            }
            if (e != hipSuccess) return e;
            if (rest > next) { // U12's other columns, and their packed copy
                const int64_t others = rest - next;
                hipLaunchKernelGGL(solve_u12<T>, dim3((unsigned)((others + kSolveU12Threads<T> - 1) / kSolveU12Threads<T>)),
                                   dim3(kSolveU12Threads<T>), 0, ws.side, l11, a12 + next, n, others);
                e = hipGetLastError();
                if (e == hipSuccess) e = pack_u12(next, others, ws.side);
                if (e != hipSuccess) return e;
            }
            carry = next_pairs;
            // A pair with at least split_columns columns beyond the next panel's leaves the update of the
            // last half of them out of what the next step waits for: that step pairs with nothing, swaps
            // only its next panel's columns on the chain and does the rest of its work on the side stream,
            // behind this update, which then runs beside its panel too.
            int64_t later = 0;
            if (pair && split_columns > 0 && rest - next >= split_columns)
                later = std::max<int64_t>(0, std::min((rest - next) / 2 / kPanel * kPanel, rest - next - kPanel));
            if (rest > next && !carry) e = update(next, rest - next - later, ws.side); // beside the panel
            if (e == hipSuccess) e = hipEventRecord(ws.panel_done, ws.side);
            if (e == hipSuccess && later > 0) e = update(rest - later, later, ws.side);
            return e;
        };

        hipError_t e = hipMemsetAsync(lines, 0, 2 * line_bytes, stream); // both sets
        if (e == hipSuccess) e = launch_panel(0, stream, 0, packs_l21(0, false));
        if (e != hipSuccess) return (int)e;
        for (int64_t j0 = 0; j0 < n; j0 += kPanel) {
            const int w = (int)std::min<int64_t>(kPanel, n - j0);
            if (j0 > 0 || ws.head_pending) { // the previous step's work on the side stream (or the caller's)
                e = hipStreamWaitEvent(stream, ws.panel_done, 0);
                if (e != hipSuccess) return (int)e;
                ws.head_pending = false;
            }
            timer.stage(0);
            const int64_t rest = n - j0 - w;
            // A float step that does not pair with the previous one needs only the next panel's columns
            // swapped on the chain of panels; its side stream swaps the others (step_float).
            const bool narrow_swap = sizeof(T) == 4 && !carry && rest > 0;
            e = narrow_swap ? swap_columns(j0, w, j0, j0 + std::min<int64_t>(kPanel, rest), stream)
                            : swap_columns(j0, w, 0, n - w, stream);
            if (e != hipSuccess) return (int)e;
            timer.stage(1);
            if (rest > 0) {
                if constexpr (sizeof(T) == 8) e = step_pair(j0, w, rest);
                else e = step_float(j0, w, rest);
                if (e != hipSuccess) return (int)e;
                timer.stage(4);
            }
            // Blocks [0, early_blocks) are final once this step's work is (later steps touch only rows
            // and columns beyond them): their inverses and links go on the third stream, which nothing
            // waits for until the end.
            if (early_blocks > 0 && j0 / kPanel == early_blocks) {
                e = hipStreamWaitEvent(ws.third, ws.panel_done, 0);
                if (e == hipSuccess) e = invert_range(static_cast<const T *>(a), n, inverses, 0, early_blocks, ws.third);
                if (e == hipSuccess) e = hipEventRecord(ws.inverted, ws.third);
                if (e != hipSuccess) return (int)e;
            }
        }
        if (inverses) {
            // The last blocks: L's part on this stream (the forward substitution comes next), U's on the
            // third stream, which only the backward substitution waits for (substitute, upper_pending).
            const int late = blocks_total - early_blocks;
            e = invert_range(static_cast<const T *>(a), n, inverses, early_blocks, late, stream, true, false);
            if (e == hipSuccess && early_blocks > 0) e = hipStreamWaitEvent(stream, ws.inverted, 0);
            if (e == hipSuccess) e = hipEventRecord(ws.columns_ready, stream);
            if (e == hipSuccess) e = hipStreamWaitEvent(ws.third, ws.columns_ready, 0);
            if (e == hipSuccess)
                e = invert_range(static_cast<const T *>(a), n, inverses, early_blocks, late, ws.third, false, true);
            if (e == hipSuccess) e = hipEventRecord(ws.upper_inverted, ws.third);
            if (e != hipSuccess) return (int)e;
            ws.upper_pending = true;
        }
        return hipGetLastError();
    }

    bool narrow_residual_fits(int64_t order, int64_t width) { return blocked_substitution(order) && width <= 32; }

    hipError_t narrow_residual(const double *a, int64_t n, const double *x, int64_t w, double *r, double *partials,
                               hipStream_t stream) {
        if (!narrow_residual_fits(n, w)) return hipErrorInvalidValue;
        hipLaunchKernelGGL(residual_slices, dim3((unsigned)(n / kBlock), kResidualSlices), dim3(256), 0, stream, a, n, x,
                           w, partials);
        hipLaunchKernelGGL(subtract_slices, dim3((unsigned)((n * w + 255) / 256)), dim3(256), 0, stream,
                           (const double *)partials, r, n * w, kResidualSlices);
        return hipGetLastError();
    }


    template <typename T> hipError_t invert_blocks(const T *f, int64_t n, T *inverses, hipStream_t stream) {
        if (!blocked_substitution(n)) return hipErrorInvalidValue;
        return invert_range(f, n, inverses, 0, (int)(n / kBlock), stream);
    }

    template <typename T>
    int substitute(const T *f, int64_t n, T *x, int64_t width, Workspace &ws, unsigned &epoch, hipStream_t stream,
                   const char **error, const T *inverses, void *scratch) {
        *error = nullptr;
        if (inverses && scratch && blocked_substitution(n)) {
            // Each triangle in one launch per 32 right-hand sides (solve_triangle), the forward pass with
            // the inverses of L's diagonal blocks, the backward pass with U's.
            const int64_t nb = n / kBlock;
            const T *inverse_lower = inverses, *inverse_upper = inverses + n * kBlock;
            unsigned long long *published = static_cast<unsigned long long *>(scratch);
            unsigned long long *partial = published + n * kTriangleColumns * 2; // the quarters before inv(D_jj)
            for (int64_t col0 = 0; col0 < width; col0 += kTriangleColumns)
                for (int pass = 0; pass < 2; ++pass) {
                    const unsigned tag = ++epoch;
                    if (pass == 1 && ws.upper_pending) { // U's last inverses, made beside the forward pass
                        if (hipError_t e = hipStreamWaitEvent(stream, ws.upper_inverted, 0)) return (int)e;
                        ws.upper_pending = false;
                    }
                    if (pass == 0)
                        hipLaunchKernelGGL((solve_triangle<T, false>), dim3((unsigned)(4 * nb)), dim3(256), 0, stream,
                                           f, inverse_lower, inverses + 2 * n * kBlock, inverses + 4 * n * kBlock, x,
                                           n, width, col0, published, partial, tag, ws.error);
                    else
                        hipLaunchKernelGGL((solve_triangle<T, true>), dim3((unsigned)(4 * nb)), dim3(256), 0, stream,
                                           f, inverse_upper, inverses + 3 * n * kBlock, inverses + 5 * n * kBlock, x,
                                           n, width, col0, published, partial, tag, ws.error);
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
                                const char **, double *);
    template int factor<float>(float *, int64_t, int64_t *, int, Workspace &, void *, size_t, hipStream_t,
                               const char **, float *);
    template int substitute<double>(const double *, int64_t, double *, int64_t, Workspace &, unsigned &, hipStream_t,
                                    const char **, const double *, void *);
    template int substitute<float>(const float *, int64_t, float *, int64_t, Workspace &, unsigned &, hipStream_t,
                                   const char **, const float *, void *);

    size_t substitute_scratch(int64_t order) { return 2 * (size_t)order * kTriangleColumns * 2 * 8; }
    template hipError_t invert_blocks<double>(const double *, int64_t, double *, hipStream_t);
    template hipError_t invert_blocks<float>(const float *, int64_t, float *, hipStream_t);
    template hipError_t permute<double, double>(const double *, double *, const int64_t *, int64_t, int64_t,
                                                hipStream_t);
    template hipError_t permute<float, double>(const double *, float *, const int64_t *, int64_t, int64_t, hipStream_t);
    template hipError_t permute<float, float>(const float *, float *, const int64_t *, int64_t, int64_t, hipStream_t);

    const char *spin_failure() { return kSpinFailure; }

} // namespace scihip
