// Matrix products on the gfx950 matrix cores (see product.hpp).
//
// Tiling: a 256-thread workgroup (four wave64s, 2 x 2) computes a 128 x 128 tile of C; each wave a
// 64 x 64 sub-tile as 4 x 4 MFMA 16 x 16 blocks held in accumulator registers. A and B (packed, the
// shared dimension contiguous) pass through LDS in k-tiles of 128 bytes per row (64 for kSplit),
// double-buffered: the next k-tile's global loads are in flight while the current one is
// multiplied. LDS rows are padded by 16 bytes so that the 16 lanes reading 16 consecutive rows hit
// distinct banks. Workgroups are numbered so that consecutive tiles run on the same XCD (blocks are
// dealt round-robin over the 8 XCDs) and neighbouring tiles share rows of A, for L2 reuse.
//
// MFMA 16 x 16 register layout (lane l, group g = l / 16, r = l % 16): the A operand is row r of the
// block at k-slot g, the B operand column r at k-slot g; accumulator element i is C[4g + i][r] for the
// fp32-accumulating instructions and C[g + 4i][r] for v_mfma_f64_16x16x4_f64 (measured on gfx950).
// Each lane reads 16 bytes of its row per step, so one step covers 64 bytes of k per row; for fp32
// (fp64) those 16 bytes feed four (two) MFMAs, each over the k values at the same position of every
// lane group's 16 bytes: a fixed permutation of k, the same for A and B, so the dot products are
// unchanged.
#include "product.hpp"

#include <algorithm>
#include <type_traits>

namespace scihip {

    namespace {

        typedef float v4f __attribute__((ext_vector_type(4)));
        typedef double v4d __attribute__((ext_vector_type(4)));
        typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
        typedef unsigned int v4u __attribute__((ext_vector_type(4)));

        constexpr int kTile = 128;      // rows of C (and columns) per workgroup
        constexpr int kThreads = 256;
        constexpr int kPad = 16;        // bytes of padding per LDS row
        constexpr int kXcds = 8;
        constexpr int kGroup = 8;       // tile rows per L2 group

        template <int M> struct Shape {
            static constexpr int row = M == kSplit ? 64 : 128;       // bytes of a k-tile row
            static constexpr int stride = row + kPad;                // LDS bytes per row
            static constexpr int operands = M == kSplit ? 4 : 2;     // A_hi, B_hi (, A_lo, B_lo)
            static constexpr int chunks = kTile * row / 16 / kThreads; // 16-byte loads per thread per operand
            static constexpr int steps = row / 64;                   // 64-byte k steps per k-tile
            static constexpr int k = row / element_bytes(M);         // k values per k-tile
        };

        template <int M> struct Accumulator {
            typedef v4f type;
        };
        template <> struct Accumulator<kFloat64> {
            typedef v4d type;
        };

        union Fragment {
            v4u raw;
            v4f f;
            v8bf b;
            double d[2];
        };

        template <int M>
        __device__ __forceinline__ void multiply(typename Accumulator<M>::type &acc, const Fragment &a,
                                                 const Fragment &b, const Fragment &a_lo, const Fragment &b_lo) {
            if constexpr (M == kFloat32) {
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.f[0], b.f[0], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.f[1], b.f[1], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.f[2], b.f[2], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.f[3], b.f[3], acc, 0, 0, 0);
            } else if constexpr (M == kBfloat16) {
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.b, b.b, acc, 0, 0, 0);
            } else if constexpr (M == kSplit) {
                // the small cross terms first, then the leading term
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_lo.b, b.b, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.b, b_lo.b, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.b, b.b, acc, 0, 0, 0);
            } else {
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a.d[0], b.d[0], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a.d[1], b.d[1], acc, 0, 0, 0);
            }
        }

        // Which 128 x 128 tile this workgroup computes: XCD-contiguous numbering, then groups of
        // kGroup tile rows walked column by column.
        __device__ __forceinline__ void tile_of(int tiles_m, int tiles_n, int &tm, int &tn) {
            const int total = tiles_m * tiles_n;
            int id = blockIdx.x;
            if (total % kXcds == 0) id = (id % kXcds) * (total / kXcds) + id / kXcds;
            const int width = kGroup * tiles_n;
            const int group = id / width;
            const int first = group * kGroup;
            const int rows = min(tiles_m - first, kGroup);
            const int within = id % width;
            tm = first + within % rows;
            tn = within / rows;
        }

        // Waves: waves per SIMD the registers are sized for (2: two workgroups per compute unit).
        template <int M, int Waves = 2> __global__ __launch_bounds__(kThreads, Waves) void gemm(const ProductArgs p) {
            typedef Shape<M> S;
            typedef typename Accumulator<M>::type Acc;
            extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
            constexpr int eb = element_bytes(M);
            constexpr int stage_bytes = S::operands * kTile * S::stride;

            const int tiles_m = (int)((p.m + kTile - 1) / kTile);
            const int tiles_n = (int)((p.n + kTile - 1) / kTile);
            int tm, tn;
            tile_of(tiles_m, tiles_n, tm, tn);
            const int64_t row0 = (int64_t)tm * kTile, col0 = (int64_t)tn * kTile;

            const int t = threadIdx.x;
            const int lane = t & 63, wave = t >> 6;
            const int g = lane >> 4, r = lane & 15;
            const int wm = wave >> 1, wn = wave & 1;

            // Global sources of this thread's chunks: operand o (0 A_hi, 1 B_hi, 2 A_lo, 3 B_lo).
            const unsigned char *source[S::operands];
            int64_t limit[S::operands], ld[S::operands], first[S::operands];
            source[0] = static_cast<const unsigned char *>(p.a_hi);
            source[1] = static_cast<const unsigned char *>(p.b_hi);
            if constexpr (S::operands == 4) {
                source[2] = static_cast<const unsigned char *>(p.a_lo);
                source[3] = static_cast<const unsigned char *>(p.b_lo);
            }
#pragma unroll
            for (int o = 0; o < S::operands; ++o) {
                const bool is_a = (o & 1) == 0;
                limit[o] = is_a ? p.m : p.n;
                ld[o] = (is_a ? p.lda : p.ldb) * eb;
                first[o] = is_a ? row0 : col0;
            }

            v4u staged[S::operands][S::chunks];
            auto fetch = [&](int64_t k_byte) {
#pragma unroll
                for (int o = 0; o < S::operands; ++o)
#pragma unroll
                    for (int j = 0; j < S::chunks; ++j) {
                        const int c = t + j * kThreads;
                        const int row = c / (S::row / 16), piece = c % (S::row / 16);
                        const int64_t global_row = first[o] + row;
                        v4u value = {0u, 0u, 0u, 0u};
                        if (global_row < limit[o])
                            value = *reinterpret_cast<const v4u *>(source[o] + global_row * ld[o] + k_byte + piece * 16);
                        staged[o][j] = value;
                    }
            };
            auto deposit = [&](int stage) {
#pragma unroll
                for (int o = 0; o < S::operands; ++o)
#pragma unroll
                    for (int j = 0; j < S::chunks; ++j) {
                        const int c = t + j * kThreads;
                        const int row = c / (S::row / 16), piece = c % (S::row / 16);
                        *reinterpret_cast<v4u *>(lds + stage * stage_bytes + (o * kTile + row) * S::stride + piece * 16) =
                                staged[o][j];
                    }
            };

            Acc acc[4][4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = Acc{0, 0, 0, 0};

            // k-tiles [k_begin, k_begin + tiles_k): all of them, or this workgroup's slice of them.
            const int64_t tiles_all = p.kp / S::k;
            const int64_t k_begin = tiles_all * blockIdx.y / gridDim.y;
            const int64_t tiles_k = tiles_all * (blockIdx.y + 1) / gridDim.y - k_begin;
            if (tiles_k > 0) {
                fetch(k_begin * S::row);
                deposit(0);
                __syncthreads();
            }
            for (int64_t kt = 0; kt < tiles_k; ++kt) {
                const int stage = (int)(kt & 1);
                if (kt + 1 < tiles_k) fetch((k_begin + kt + 1) * S::row);
                const unsigned char *base = lds + stage * stage_bytes;
#pragma unroll
                for (int s = 0; s < S::steps; ++s) {
                    Fragment a[4], b[4], a_lo[4], b_lo[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int arow = wm * 64 + i * 16 + r, brow = wn * 64 + i * 16 + r;
                        const int offset = s * 64 + g * 16;
                        a[i].raw = *reinterpret_cast<const v4u *>(base + arow * S::stride + offset);
                        b[i].raw = *reinterpret_cast<const v4u *>(base + (kTile + brow) * S::stride + offset);
                        if constexpr (M == kSplit) {
                            a_lo[i].raw = *reinterpret_cast<const v4u *>(base + (2 * kTile + arow) * S::stride + offset);
                            b_lo[i].raw = *reinterpret_cast<const v4u *>(base + (3 * kTile + brow) * S::stride + offset);
                        }
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < 4; ++j) multiply<M>(acc[i][j], a[i], b[j], a_lo[i], b_lo[j]);
                }
                if (kt + 1 < tiles_k) deposit(stage ^ 1);
                __syncthreads();
            }

            // Epilogue: accumulator element e of block (i, j) is C[row0 + 64 wm + 16 i + 4 g + e] for fp32
            // accumulators and C[row0 + 64 wm + 16 i + g + 4 e] for fp64 ones, column col0 + 64 wn + 16 j + r.
            // With beta != 0, the 16 C values of each block row are loaded before any of them is stored,
            // so their latencies overlap (stores to the same array would otherwise order every load
            // after the store before it).
            typedef typename std::conditional<M == kFloat64, double, float>::type CT;
            CT *const c_base = static_cast<CT *>(p.c) + (int64_t)blockIdx.y * p.slice_stride;
            const CT alpha = (CT)p.alpha, beta = (CT)p.beta;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                CT old[4][4];
                if (beta != CT(0)) {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const int64_t row = row0 + wm * 64 + i * 16 + (M == kFloat64 ? g + 4 * e : g * 4 + e);
                            const int64_t col = col0 + wn * 64 + j * 16 + r;
                            old[j][e] = (row < p.m && col < p.n) ? c_base[row * p.ldc + col] : CT(0);
                        }
                }
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int64_t row = row0 + wm * 64 + i * 16 + (M == kFloat64 ? g + 4 * e : g * 4 + e);
                        const int64_t col = col0 + wn * 64 + j * 16 + r;
                        if (row >= p.m || col >= p.n) continue;
                        const CT v = (CT)acc[i][j][e];
                        c_base[row * p.ldc + col] = beta == CT(0) ? (alpha == CT(1) ? v : alpha * v)
                                                                  : alpha * v + beta * old[j][e];
                    }
            }
        }

        // The large-tile bf16 product for long shared dimensions. A 256-thread workgroup computes a
        // 256 x 256 tile of C, each wave 128 x 128 as 4 x 4 blocks of v_mfma_f32_32x32x16_bf16
        // (256 accumulators per lane, one wave per SIMD). Operands stream global -> LDS with
        // global_load_lds_dwordx4 (each lane's 16 bytes land at the wave's LDS base + 16 * lane), in
        // k-tiles of 64 values (128 bytes per row), double-buffered: the next k-tile's loads are in
        // flight while the current one is multiplied, and one barrier per k-tile both publishes the
        // next tile and retires the current one. LDS rows are stored with their eight 16-byte chunks
        // permuted, chunk c of row r at position c ^ ((r / 2) % 8): the 16 lanes that read one chunk
        // of 16 consecutive rows in one LDS pass then cover all 64 banks once. The permutation costs
        // nothing: each lane simply fetches the chunk that belongs at its LDS position.
        // 32 x 32 x 16 layout (measured on gfx950): lane l supplies row (column) l % 32 of A (B) at k
        // 8 (l / 32) ... + 7; accumulator v of lane l is C[8 (v / 4) + 4 (l / 32) + v % 4][l % 32].
        constexpr int kWide = 256;          // rows and columns of C per workgroup
        constexpr int kWideRow = 128;       // bytes of a k-tile row (64 bf16)
        constexpr int kWideStage = 2 * kWide * kWideRow; // A and B of one k-tile: 64 KiB

        typedef float v16f __attribute__((ext_vector_type(16)));
        typedef __attribute__((address_space(3))) void *lds_pointer;

        __device__ __forceinline__ int wide_position(int row, int chunk) { return chunk ^ ((row >> 1) & 7); }

        template <int Waves = 1>
        __global__ __launch_bounds__(kThreads, Waves) void gemm_wide(const ProductArgs p) {
            extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
            const int tiles_m = (int)((p.m + kWide - 1) / kWide);
            const int tiles_n = (int)((p.n + kWide - 1) / kWide);
            int tm, tn;
            tile_of(tiles_m, tiles_n, tm, tn);
            const int64_t row0 = (int64_t)tm * kWide, col0 = (int64_t)tn * kWide;
            // the wave number is the same in every lane: say so, so that LDS bases (M0) are scalar
            const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
            const int wm = wave >> 1, wn = wave & 1;

            // This lane's part of each direct load: wave `wave` fills the LDS blocks 8 wave ... 8 wave + 7
            // of each operand (1 KiB each: 8 rows of 128 bytes); lane l of block b writes row 8 b + l / 8
            // at position l % 8 and so fetches chunk (l % 8) ^ ((row / 2) % 8) of that row, which is
            // (l % 8) ^ (4 (b % 2) + l / 16). The loads go through buffer descriptors based at the
            // tile's first row and sized to the rows that exist: the per-lane offset (row and chunk) is
            // range-checked, so rows past the matrix read zeros; the k offset rides in soffset.
            const int64_t lda = p.lda * 2, ldb = p.ldb * 2;
            // (every field is the same in all lanes; readfirstlane puts the descriptors in SGPRs, where
            // the buffer instructions need them)
            auto scalar = [](uint64_t x) {
                return (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(x >> 32)) << 32 |
                       (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)x);
            };
            const int64_t a_bytes = (p.m - row0) * lda, b_bytes = (p.n - col0) * ldb;
            const uint32_t a_records =
                    __builtin_amdgcn_readfirstlane(a_bytes > 0xFFFFFFFFll ? 0xFFFFFFFFu : (uint32_t)a_bytes);
            const uint32_t b_records =
                    __builtin_amdgcn_readfirstlane(b_bytes > 0xFFFFFFFFll ? 0xFFFFFFFFu : (uint32_t)b_bytes);
            const __amdgpu_buffer_rsrc_t a_rsrc = __builtin_amdgcn_make_buffer_rsrc(
                    (void *)scalar((uint64_t)(static_cast<const unsigned char *>(p.a_hi) + row0 * lda)), 0, a_records,
                    0x00020000);
            const __amdgpu_buffer_rsrc_t b_rsrc = __builtin_amdgcn_make_buffer_rsrc(
                    (void *)scalar((uint64_t)(static_cast<const unsigned char *>(p.b_hi) + col0 * ldb)), 0, b_records,
                    0x00020000);
            uint32_t a_offset[8], b_offset[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int row = (wave * 8 + i) * 8 + (lane >> 3);
                const int chunk = wide_position(row, lane & 7);
                a_offset[i] = (uint32_t)(row * lda + chunk * 16);
                b_offset[i] = (uint32_t)(row * ldb + chunk * 16);
            }
            auto issue = [&](int64_t k_byte, int stage) {
                unsigned char *base = lds + stage * kWideStage;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(a_rsrc, (lds_pointer)(base + (wave * 8 + i) * 1024), 16,
                                                              a_offset[i], (int)k_byte, 0, 0);
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(
                            b_rsrc, (lds_pointer)(base + kWide * kWideRow + (wave * 8 + i) * 1024), 16, b_offset[i],
                            (int)k_byte, 0, 0);
                }
            };

            v16f acc[4][4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;

            // Fragment addresses within a stage: block i of A is rows 128 wm + 32 i + l % 32, block j of
            // B rows 128 wn + 32 j + l % 32; k-step s (16 values) reads chunk 2 s + l / 32.
            int a_row[4], b_row[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a_row[i] = wm * 128 + i * 32 + (lane & 31);
                b_row[i] = wn * 128 + i * 32 + (lane & 31);
            }
            const int half = lane >> 5;

            const int64_t tiles_all = p.kp / 64;
            const int64_t k_begin = tiles_all * blockIdx.y / gridDim.y;
            const int64_t tiles_k = tiles_all * (blockIdx.y + 1) / gridDim.y - k_begin;
            if (tiles_k > 0) {
                issue(k_begin * kWideRow, 0);
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
            }
            for (int64_t kt = 0; kt < tiles_k; ++kt) {
                const int stage = (int)(kt & 1);
                if (kt + 1 < tiles_k) issue((k_begin + kt + 1) * kWideRow, stage ^ 1);
                const unsigned char *a_base = lds + stage * kWideStage, *b_base = a_base + kWide * kWideRow;
                // Fragments in two register sets: step s + 1's LDS reads are issued before step s's
                // sixteen MFMAs, so their latency hides under the matrix work.
                v8bf a[2][4], b[2][4];
                auto read = [&](int set, int s) {
                    const int chunk = 2 * s + half;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        a[set][i] = *reinterpret_cast<const v8bf *>(a_base + a_row[i] * kWideRow +
                                                                    wide_position(a_row[i], chunk) * 16);
                        b[set][i] = *reinterpret_cast<const v8bf *>(b_base + b_row[i] * kWideRow +
                                                                    wide_position(b_row[i], chunk) * 16);
                    }
                };
                read(0, 0);
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    if (s < 3) read((s + 1) & 1, s + 1);
                    // keep the compiler from sinking those reads next to their MFMAs
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[s & 1][i], b[s & 1][j], acc[i][j],
                                                                                 0, 0, 0);
                    __builtin_amdgcn_sched_barrier(0);
                }
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
            }

            float *const c_base = static_cast<float *>(p.c) + (int64_t)blockIdx.y * p.slice_stride;
            const float alpha = (float)p.alpha, beta = (float)p.beta;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int64_t col = col0 + wn * 128 + j * 32 + (lane & 31);
                    float old[16];
                    if (beta != 0.0f) {
#pragma unroll
                        for (int v = 0; v < 16; ++v) {
                            const int64_t row = row0 + wm * 128 + i * 32 + (v >> 2) * 8 + half * 4 + (v & 3);
                            old[v] = (row < p.m && col < p.n) ? c_base[row * p.ldc + col] : 0.0f;
                        }
                    }
#pragma unroll
                    for (int v = 0; v < 16; ++v) {
                        const int64_t row = row0 + wm * 128 + i * 32 + (v >> 2) * 8 + half * 4 + (v & 3);
                        if (row >= p.m || col >= p.n) continue;
                        const float x = acc[i][j][v];
                        c_base[row * p.ldc + col] = beta == 0.0f ? (alpha == 1.0f ? x : alpha * x)
                                                                  : alpha * x + beta * old[v];
                    }
                }
        }

        // 64 x 64 tiles through LDS, so that both the plain and the transposed copy read and write
        // global memory along rows.
        template <typename S, typename D, bool split>
        __global__ __launch_bounds__(256) void pack_tile(const S *source, int64_t rows, int64_t depth, int64_t ld,
                                                         bool transpose, D *hi, D *lo, int64_t kp, int64_t stride) {
            __shared__ S tile[64][65];
            const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
            const int64_t k0 = (int64_t)blockIdx.x * 64, i0 = (int64_t)blockIdx.y * 64;
            for (int y = ty; y < 64; y += 4) {
                if (transpose) { // source is depth x rows: element (i, k) = source[k * ld + i]
                    const int64_t k = k0 + y, i = i0 + tx;
                    tile[tx][y] = (k < depth && i < rows) ? source[k * ld + i] : S(0);
                } else {
                    const int64_t i = i0 + y, k = k0 + tx;
                    tile[y][tx] = (k < depth && i < rows) ? source[i * ld + k] : S(0);
                }
            }
            __syncthreads();
            for (int y = ty; y < 64; y += 4) {
                const int64_t i = i0 + y, k = k0 + tx;
                if (i >= rows || k >= kp) continue;
                pack_element<D, split>(tile[y][tx], hi, lo, i * stride + k);
            }
        }

        template <int M> constexpr int lds_bytes() { return 2 * Shape<M>::operands * kTile * Shape<M>::stride; }

        // C[r][c] = alpha * sum over slices s (in order) of partials[s][r][c] + beta * C[r][c].
        template <typename CT>
        __global__ __launch_bounds__(256) void fold_slices(const CT *partials, CT *c, int64_t m, int64_t n,
                                                           int64_t ldc, int slices, CT alpha, CT beta) {
            const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
            if (idx >= m * n) return;
            const int64_t r = idx / n, col = idx % n;
            CT sum = partials[idx];
            for (int s = 1; s < slices; ++s) sum += partials[(int64_t)s * m * n + idx];
            CT *target = c + r * ldc + col;
            *target = beta == CT(0) ? alpha * sum : alpha * sum + beta * *target;
        }

    } // namespace

    namespace {
        // The large-tile kernel takes bf16 products whose output spans at least a 256 x 256 tile and
        // whose shared dimension has at least 8 k-tiles of 64; shorter ones (such as the LU's rank-128
        // updates, bound by reading and writing C) keep the 128 x 128 kernel.
        bool wide_fits(int mode, const ProductArgs &args) {
            // enough 256 x 256 tiles to give every compute unit one (fewer run faster as 128 x 128 tiles),
            // and 32-bit buffer offsets: a tile's 256 rows of the operands span less than 2 GiB
            const int64_t tiles = ((args.m + kWide - 1) / kWide) * ((args.n + kWide - 1) / kWide);
            return mode == kBfloat16 && tiles >= 256 && args.kp % 64 == 0 && args.kp >= 512 &&
                   int64_t(kWide) * std::max(args.lda, args.ldb) * 2 < (int64_t(1) << 31);
        }

        hipError_t launch_gemm(int mode, const ProductArgs &args, int slices, hipStream_t stream, bool alongside = false) {
            if (wide_fits(mode, args)) {
                const int64_t wide_tiles = ((args.m + kWide - 1) / kWide) * ((args.n + kWide - 1) / kWide);
                hipLaunchKernelGGL(gemm_wide<1>, dim3((unsigned)wide_tiles, (unsigned)slices), dim3(kThreads),
                                   2 * kWideStage, stream, args);
                return hipGetLastError();
            }
            const int64_t tiles = ((args.m + kTile - 1) / kTile) * ((args.n + kTile - 1) / kTile);
            if (tiles == 0) return hipSuccess;
            const dim3 grid((unsigned)tiles, (unsigned)slices), block(kThreads);
            switch (mode) {
            case kFloat32:
                hipLaunchKernelGGL(gemm<kFloat32>, grid, block, lds_bytes<kFloat32>(), stream, args);
                break;
            case kBfloat16:
                hipLaunchKernelGGL(gemm<kBfloat16>, grid, block, lds_bytes<kBfloat16>(), stream, args);
                break;
            case kSplit:
                hipLaunchKernelGGL(gemm<kSplit>, grid, block, lds_bytes<kSplit>(), stream, args);
                break;
            default:
                if (alongside) hipLaunchKernelGGL((gemm<kFloat64, 1>), grid, block, lds_bytes<kFloat64>(), stream, args);
                else hipLaunchKernelGGL((gemm<kFloat64, 2>), grid, block, lds_bytes<kFloat64>(), stream, args);
                break;
            }
            return hipGetLastError();
        }
    } // namespace

    hipError_t product_split(int mode, const ProductArgs &args, int slices, void *partials, hipStream_t stream) {
        if (slices <= 1) return product(mode, args, stream);
        if (args.m == 0 || args.n == 0) return hipSuccess;
        ProductArgs sliced = args;
        sliced.c = partials;
        sliced.ldc = args.n;
        sliced.alpha = 1.0;
        sliced.beta = 0.0;
        sliced.slice_stride = args.m * args.n;
        hipError_t e = launch_gemm(mode, sliced, slices, stream);
        if (e != hipSuccess) return e;
        const unsigned blocks = (unsigned)((args.m * args.n + 255) / 256);
        if (mode == kFloat64)
            hipLaunchKernelGGL(fold_slices<double>, dim3(blocks), dim3(256), 0, stream,
                               static_cast<const double *>(partials), static_cast<double *>(args.c), args.m, args.n,
                               args.ldc, slices, args.alpha, args.beta);
        else
            hipLaunchKernelGGL(fold_slices<float>, dim3(blocks), dim3(256), 0, stream,
                               static_cast<const float *>(partials), static_cast<float *>(args.c), args.m, args.n,
                               args.ldc, slices, (float)args.alpha, (float)args.beta);
        return hipGetLastError();
    }

    int split_slices(int mode, int64_t m, int64_t n, int64_t kp) {
        const int64_t tiles = ((m + kTile - 1) / kTile) * ((n + kTile - 1) / kTile);
        if (tiles == 0 || tiles >= 256) return 1;
        const int64_t by_depth = kp / k_tile(mode) / 8, by_chip = (1024 + tiles - 1) / tiles;
        return (int)std::max<int64_t>(1, std::min<int64_t>({16, by_depth, by_chip}));
    }

    hipError_t product(int mode, const ProductArgs &args, hipStream_t stream) {
        ProductArgs whole = args;
        whole.slice_stride = 0;
        return launch_gemm(mode, whole, 1, stream);
    }

    hipError_t product_alongside(int mode, const ProductArgs &args, hipStream_t stream) {
        ProductArgs whole = args;
        whole.slice_stride = 0;
        return launch_gemm(mode, whole, 1, stream, true);
    }

    hipError_t pack(int mode, const void *source, bool wide, int64_t rows, int64_t depth, int64_t ld, bool transpose,
                    void *hi, void *lo, int64_t kp, hipStream_t stream, int64_t stride) {
        if (rows == 0 || kp == 0) return hipSuccess;
        if (stride == 0) stride = kp;
        const dim3 grid((unsigned)((kp + 63) / 64), (unsigned)((rows + 63) / 64)), block(256);
        if (mode == kFloat64) {
            if (!wide) return hipErrorInvalidValue;
            hipLaunchKernelGGL((pack_tile<double, double, false>), grid, block, 0, stream,
                               static_cast<const double *>(source), rows, depth, ld, transpose,
                               static_cast<double *>(hi), nullptr, kp, stride);
        } else if (mode == kFloat32) {
            if (wide)
                hipLaunchKernelGGL((pack_tile<double, float, false>), grid, block, 0, stream,
                                   static_cast<const double *>(source), rows, depth, ld, transpose,
                                   static_cast<float *>(hi), nullptr, kp, stride);
            else
                hipLaunchKernelGGL((pack_tile<float, float, false>), grid, block, 0, stream,
                                   static_cast<const float *>(source), rows, depth, ld, transpose,
                                   static_cast<float *>(hi), nullptr, kp, stride);
        } else if (mode == kBfloat16) {
            if (wide)
                hipLaunchKernelGGL((pack_tile<double, __bf16, false>), grid, block, 0, stream,
                                   static_cast<const double *>(source), rows, depth, ld, transpose,
                                   static_cast<__bf16 *>(hi), nullptr, kp, stride);
            else
                hipLaunchKernelGGL((pack_tile<float, __bf16, false>), grid, block, 0, stream,
                                   static_cast<const float *>(source), rows, depth, ld, transpose,
                                   static_cast<__bf16 *>(hi), nullptr, kp, stride);
        } else {
            if (wide)
                hipLaunchKernelGGL((pack_tile<double, __bf16, true>), grid, block, 0, stream,
                                   static_cast<const double *>(source), rows, depth, ld, transpose,
                                   static_cast<__bf16 *>(hi), static_cast<__bf16 *>(lo), kp, stride);
            else
                hipLaunchKernelGGL((pack_tile<float, __bf16, true>), grid, block, 0, stream,
                                   static_cast<const float *>(source), rows, depth, ld, transpose,
                                   static_cast<__bf16 *>(hi), static_cast<__bf16 *>(lo), kp, stride);
        }
        return hipGetLastError();
    }

} // namespace scihip
