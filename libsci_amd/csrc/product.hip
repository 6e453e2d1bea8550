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

        // The large-tile bf16 product for long shared dimensions. Its operands are packed in slabs
        // (product.hpp): each 32-value k-stage of 256 rows is one contiguous 16 KiB block, rows 64 bytes
        // apart with their four 16-byte chunks permuted (chunk c of row r at position c ^ ((r / 4) % 4)).
        // A persistent grid, one 256-thread workgroup per compute unit, walks the 256 x 256 tiles of C;
        // each wave computes 128 x 128 of a tile as 4 x 4 blocks of v_mfma_f32_32x32x16_bf16 (256
        // accumulators per lane, one wave per SIMD).
        // - Stages stream global -> LDS with buffer_load_dwordx4 ... lds, lane l's 16 bytes landing at
        //   the wave's LDS base + 16 l, so each load copies 1 KiB of a slab verbatim. A ring of five
        //   slots (160 KiB, all of the LDS) keeps the loads four stages ahead.
        // - Each stage is two k-steps of 16. A wave's fragment reads for one k-step go out while the
        //   other k-step's sixteen MFMAs run, and the one barrier per stage sits between the two
        //   halves, so neither the LDS reads nor the barrier leave the matrix cores idle for long.
        // - The permutation makes the 16 lanes of one LDS pass (one chunk of 16 consecutive rows) cover
        //   all 64 banks once.
        // - The MFMA takes the B fragment as its first operand, so it computes blocks of C transposed:
        //   accumulator v of lane l is C[32 i + l % 32][32 j + 8 (v / 4) + 4 (l / 32) + v % 4], four
        //   consecutive columns per lane and group of four, stored as one 16-byte vector.
        // 32 x 32 x 16 layout (measured on gfx950): lane l supplies row l % 32 of its first operand and
        // column l % 32 of its second, at k 8 (l / 32) ... + 7; accumulator v of lane l is row
        // 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32 of the product.
        constexpr int kWide = kSlabRows;   // rows and columns of C per tile
        constexpr int kStageBytes = kSlabRows * kSlabK * 2;    // one operand's stage: 16 KiB
        constexpr int kStageLoads = kStageBytes / 1024 / 4;      // 1 KiB loads per wave per operand
        // Per mode: operands per stage (A, B; bf16x3 adds A_lo, B_lo) and ring slots in 160 KiB of LDS.
        template <int Mode> struct SlabRing {
            static constexpr int operands = Mode == kSplit ? 4 : 2;
            static constexpr int slot = operands * kStageBytes;
            static constexpr int slots = Mode == kSplit ? 2 : 5;
            static constexpr int lds = slot * slots;
        };

        typedef float v16f __attribute__((ext_vector_type(16)));
        typedef __attribute__((address_space(3))) void *lds_pointer;

        __device__ __forceinline__ int slab_position(int row, int chunk) { return chunk ^ ((row >> 2) & 3); }

        // Waits until at most `pending` bf16 stages of this wave's loads are in flight (the newest ones:
        // loads complete in order for the count).
        __device__ __forceinline__ void wait_stages(int pending) {
            static_assert(2 * kStageLoads == 8, "the counts below assume eight loads per stage");
            switch (pending) {
            case 0: asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); break;
            case 1: asm volatile("s_waitcnt vmcnt(8)" ::: "memory"); break;
            case 2: asm volatile("s_waitcnt vmcnt(16)" ::: "memory"); break;
            default: asm volatile("s_waitcnt vmcnt(24)" ::: "memory"); break;
            }
        }

        // A workgroup barrier that waits for nothing in flight: the compiler's own barrier would first
        // wait for every global load outstanding, the very loads the ring keeps running ahead. Each
        // wave has waited (wait_stages) for its share of the stage the others will read, and the LDS
        // reads it issued before the barrier have returned their data to the MFMAs that consumed them;
        // the clobber keeps the compiler from moving LDS accesses across it.
        __device__ __forceinline__ void lds_barrier() { asm volatile("s_barrier" ::: "memory"); }

        template <int Mode>
        __global__ __launch_bounds__(kThreads, 1) void gemm_slabs(const ProductArgs p, int slices) {
            typedef SlabRing<Mode> Ring;
            constexpr int Stages = Ring::slots, kSlotBytes = Ring::slot;
            // stages a tile's prologue issues: all but one slot for bf16 (each stage refills the slot the
            // one before it used), both slots for bf16x3 (each stage refills its own slot)
            constexpr int Ahead = Mode == kSplit ? Stages : Stages - 1;
            static_assert(Mode == kSplit || (Stages >= 3 && Stages <= 5), "wait_stages counts up to three pending stages");
            extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
            const int tiles_m = (int)((p.m + kWide - 1) / kWide);
            const int tiles_n = (int)((p.n + kWide - 1) / kWide);
            const int tiles = tiles_m * tiles_n;
            // the wave number is the same in every lane: say so, so that LDS bases (M0) are scalar
            const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
            const int wm = wave >> 1, wn = wave & 1, half = lane >> 5;
            const int64_t stages_all = p.kp / kSlabK;
            const uint32_t slab = (uint32_t)(stages_all * kStageBytes); // bytes of one 256-row slab
            // work item w: tile w % tiles over the k-stages of slice w / tiles, [first, first + count)
            const int items = tiles * slices;
            auto range = [&](int w, int64_t &first, int64_t &count) {
                const int slice = w / tiles;
                first = stages_all * slice / slices;
                count = stages_all * (slice + 1) / slices - first;
            };

            // This lane's part of each load: wave w copies the 1 KiB pieces 4 w ... 4 w + 3 of each
            // operand's stage, lane l the 16 bytes at 16 l of a piece. The stage offset rides in soffset.
            uint32_t lane_offset[kStageLoads];
#pragma unroll
            for (int i = 0; i < kStageLoads; ++i) lane_offset[i] = (uint32_t)((wave * kStageLoads + i) * 1024 + lane * 16);
            // Fragment addresses within a slot: block i of A is rows 128 wm + 32 i + l % 32, block j of
            // B rows 128 wn + 32 j + l % 32; k-step s reads chunk 2 s + l / 32 of its row.
            int a_at[2][4], b_at[2][4];
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int a_row = wm * 128 + i * 32 + (lane & 31), b_row = wn * 128 + i * 32 + (lane & 31);
                    a_at[s][i] = a_row * 64 + slab_position(a_row, 2 * s + half) * 16;
                    b_at[s][i] = kStageBytes + b_row * 64 + slab_position(b_row, 2 * s + half) * 16;
                }
            // (every field is the same in all lanes; readfirstlane puts the descriptors in SGPRs, where
            // the buffer instructions need them)
            auto scalar = [](uint64_t x) {
                return (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(x >> 32)) << 32 |
                       (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)x);
            };
            auto descriptor = [&](const void *base, int index) {
                return __builtin_amdgcn_make_buffer_rsrc(
                        (void *)scalar((uint64_t)(static_cast<const unsigned char *>(base) + (uint64_t)index * slab)),
                        0, slab, 0x00020000);
            };
            // The operands' descriptors in slot order: A, B, then (bf16x3) A_lo, B_lo. (Descriptors are
            // sizeless values: no arrays or structs of them.)
            auto issue = [&](__amdgpu_buffer_rsrc_t a, __amdgpu_buffer_rsrc_t b, __amdgpu_buffer_rsrc_t a_lo,
                             __amdgpu_buffer_rsrc_t b_lo, int64_t stage, int slot) {
                unsigned char *base = lds + slot * kSlotBytes;
                const int k_byte = (int)(stage * kStageBytes);
#pragma unroll
                for (int i = 0; i < kStageLoads; ++i) {
                    const int piece = (wave * kStageLoads + i) * 1024;
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(a, (lds_pointer)(base + piece), 16, lane_offset[i], k_byte, 0, 0);
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(b, (lds_pointer)(base + kStageBytes + piece), 16, lane_offset[i],
                                                              k_byte, 0, 0);
                    if constexpr (Ring::operands == 4) {
                        __builtin_amdgcn_raw_ptr_buffer_load_lds(a_lo, (lds_pointer)(base + 2 * kStageBytes + piece), 16,
                                                                  lane_offset[i], k_byte, 0, 0);
                        __builtin_amdgcn_raw_ptr_buffer_load_lds(b_lo, (lds_pointer)(base + 3 * kStageBytes + piece), 16,
                                                                  lane_offset[i], k_byte, 0, 0);
                    }
                }
            };
            auto place = [&](int t, int &tm, int &tn) {
                // XCD-contiguous numbering, then groups of kGroup tile rows walked column by column
                int id = t;
                if (tiles % kXcds == 0) id = (id % kXcds) * (tiles / kXcds) + id / kXcds;
                const int width = kGroup * tiles_n;
                const int first = (id / width) * kGroup;
                const int rows = min(tiles_m - first, kGroup);
                const int within = id % width;
                tm = first + within % rows;
                tn = within / rows;
            };
            auto prologue = [&](int w) {
                int tm, tn;
                int64_t first, count;
                place(w % tiles, tm, tn);
                range(w, first, count);
                const __amdgpu_buffer_rsrc_t a = descriptor(p.a_hi, tm), b = descriptor(p.b_hi, tn);
                const __amdgpu_buffer_rsrc_t a_lo = Ring::operands == 4 ? descriptor(p.a_lo, tm) : a;
                const __amdgpu_buffer_rsrc_t b_lo = Ring::operands == 4 ? descriptor(p.b_lo, tn) : b;
#pragma unroll
                for (int s = 0; s < Ahead; ++s)
                    if (s < count) issue(a, b, a_lo, b_lo, first + s, s);
            };

            const bool vector_store = p.beta == 0.0 && (p.ldc & 3) == 0 && ((uintptr_t)p.c & 15) == 0;
            const float alpha = (float)p.alpha, beta = (float)p.beta;
            if ((int)blockIdx.x < items) prologue(blockIdx.x);
            for (int w = blockIdx.x; w < items; w += gridDim.x) {
                int tm, tn;
                int64_t first, stages;
                place(w % tiles, tm, tn);
                range(w, first, stages);
                const __amdgpu_buffer_rsrc_t a = descriptor(p.a_hi, tm), b = descriptor(p.b_hi, tn);
                const __amdgpu_buffer_rsrc_t a_lo = Ring::operands == 4 ? descriptor(p.a_lo, tm) : a;
                const __amdgpu_buffer_rsrc_t b_lo = Ring::operands == 4 ? descriptor(p.b_lo, tn) : b;
                v16f acc[4][4];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
#pragma unroll
                        for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;

                if constexpr (Mode == kBfloat16) {
                    v8bf x[2][4], y[2][4]; // fragments of k-step 0 (x) and 1 (y) of the current stage: [A, B][block]
                    // Fragment q of a k-step in the order the MFMAs first need them: A0, B0, B1, B2, B3,
                    // A1, A2, A3 (MFMA q runs block (q / 4, q % 4)).
                    auto read_one = [&](v8bf (&f)[2][4], int slot, int step, int q) {
                        const unsigned char *base = lds + slot * kSlotBytes;
                        const int operand = (q == 0 || q >= 5) ? 0 : 1;
                        const int block = q == 0 ? 0 : (q >= 5 ? q - 4 : q - 1);
                        f[operand][block] = *reinterpret_cast<const v8bf *>(
                                base + (operand == 0 ? a_at[step][block] : b_at[step][block]));
                    };
                    // Load q of a stage: operand q % 2 (A, B), this wave's 1 KiB piece q / 2.
                    auto load_one = [&](int64_t stage, int slot, int q) {
                        unsigned char *base = lds + slot * kSlotBytes + (q & 1) * kStageBytes;
                        const int piece = (wave * kStageLoads + (q >> 1)) * 1024;
                        __builtin_amdgcn_raw_ptr_buffer_load_lds(q & 1 ? b : a, (lds_pointer)(base + piece), 16,
                                                                  lane_offset[q >> 1], (int)(stage * kStageBytes), 0, 0);
                    };
                    auto mfma_one = [&](const v8bf (&f)[2][4], int q) {
                        acc[q >> 2][q & 3] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[1][q & 3], f[0][q >> 2],
                                                                                     acc[q >> 2][q & 3], 0, 0, 0);
                    };

                    // this tile's first stages were issued by the prologue (with the previous tile's stores
                    // after them in the count, so wait for everything)
                    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                    lds_barrier();
#pragma unroll
                    for (int q = 0; q < 8; ++q) read_one(x, 0, 0, q);
                    int slot = 0;
                    // One stage, two halves of sixteen MFMAs. With one wave per SIMD every other instruction
                    // takes an issue slot of its own, so the LDS reads and the loads are spread between the
                    // MFMAs (one after every other MFMA), where each issues while an MFMA runs:
                    // - first half: k-step 0's MFMAs, with k-step 1's eight fragment reads;
                    // - then (unless this is the last stage) the wait for stage kt + 1 and the barrier;
                    // - second half: k-step 1's MFMAs, with the eight loads refilling the ring slot of stage
                    //   kt - 1 (which every wave is past) with stage kt + Stages - 1, and the next stage's
                    //   k-step 0 reads.
                    // `more` and `load` are constants at every call, so nothing in a half is conditional.
                    auto advance = [&](int64_t kt, bool more, bool load, int pending) {
                        const int next = slot + 1 == Stages ? 0 : slot + 1;
                        const int refill = slot == 0 ? Stages - 1 : slot - 1;
#pragma unroll
                        for (int q = 0; q < 16; ++q) {
                            mfma_one(x, q);
                            if ((q & 1) == 0) read_one(y, slot, 1, q >> 1);
                            __builtin_amdgcn_sched_barrier(0);
                        }
                        if (more) {
                            wait_stages(pending);
                            lds_barrier();
                        }
#pragma unroll
                        for (int q = 0; q < 16; ++q) {
                            mfma_one(y, q);
                            if (load && (q & 1) == 0) load_one(first + kt + Stages - 1, refill, q >> 1);
                            if (more && (q & 1) == 1) read_one(x, next, 0, q >> 1);
                            __builtin_amdgcn_sched_barrier(0);
                        }
                        slot = next;
                    };
                    // stage kt + 1 has landed once only the stages after it, up to kt + Stages - 2, are pending
                    int64_t kt = 0;
                    for (; kt < stages - Stages + 1; ++kt) advance(kt, true, true, Stages - 3);
                    for (; kt < stages - 1 && kt < stages - Stages + 2; ++kt) advance(kt, true, false, Stages - 3);
                    for (; kt < stages - 1; ++kt) advance(kt, true, false, (int)(stages - 2 - kt));
                    advance(stages - 1, false, false, 0);
                } else {
                    // bf16x3: two slots; a stage is 96 MFMAs per wave (three per block and k-step), laid
                    // out as the bf16 stage is: k-step 0's MFMAs with k-step 1's reads, the barrier, then
                    // k-step 1's MFMAs with the loads of stage kt + 2 into this stage's slot and the next
                    // stage's k-step 0 reads (one read or load after every third MFMA). The loads have a
                    // stage of 96 MFMAs to land.
                    v8bf x[4][4], y[4][4]; // k-steps 0 and 1: [A, B, A_lo, B_lo][block]
                    // fragment q in the order the MFMAs first need them: block (0, 0)'s B, A_lo, B_lo, A;
                    // then B and B_lo of blocks (0, 1) ... (0, 3); then A_lo and A of rows 1 ... 3
                    auto read_one = [&](v8bf (&f)[4][4], int slot, int step, int q) {
                        constexpr int kOperand[16] = {1, 2, 3, 0, 1, 3, 1, 3, 1, 3, 2, 0, 2, 0, 2, 0};
                        constexpr int kBlock[16] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 1, 1, 2, 2, 3, 3};
                        const int operand = kOperand[q], block = kBlock[q];
                        const unsigned char *base = lds + slot * kSlotBytes + (operand >= 2 ? 2 * kStageBytes : 0);
                        f[operand][block] = *reinterpret_cast<const v8bf *>(
                                base + ((operand & 1) == 0 ? a_at[step][block] : b_at[step][block]));
                    };
                    // load q: operand q % 4 (A, B, A_lo, B_lo), this wave's 1 KiB piece q / 4
                    auto load_one = [&](int64_t stage, int slot, int q) {
                        const int operand = q & 3;
                        unsigned char *base = lds + slot * kSlotBytes + operand * kStageBytes;
                        const int piece = (wave * kStageLoads + (q >> 2)) * 1024;
                        __builtin_amdgcn_raw_ptr_buffer_load_lds(operand == 0 ? a : operand == 1 ? b : operand == 2 ? a_lo : b_lo,
                                                                  (lds_pointer)(base + piece), 16, lane_offset[q >> 2],
                                                                  (int)(stage * kStageBytes), 0, 0);
                    };
                    // MFMA q: block (q / 12, (q / 3) % 4), term q % 3: the small cross terms first, then
                    // the leading term
                    auto mfma_one = [&](const v8bf (&f)[4][4], int q) {
                        const int i = q / 12, j = (q / 3) % 4, term = q % 3;
                        const v8bf &left = term == 1 ? f[3][j] : f[1][j];
                        const v8bf &right = term == 0 ? f[2][i] : f[0][i];
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(left, right, acc[i][j], 0, 0, 0);
                    };
                    // this tile's first two stages were issued by the prologue (with the previous tile's
                    // stores after them in the count, so wait for everything)
                    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                    lds_barrier();
#pragma unroll
                    for (int q = 0; q < 16; ++q) read_one(x, 0, 0, q);
                    auto advance = [&](int64_t kt, bool more, bool load) {
                        const int slot = (int)(kt & 1);
#pragma unroll
                        for (int q = 0; q < 48; ++q) {
                            mfma_one(x, q);
                            if (q % 3 == 0) read_one(y, slot, 1, q / 3);
                            __builtin_amdgcn_sched_barrier(0);
                        }
                        if (more) {
                            // stage kt + 1 has landed for every wave, and every wave's reads of this slot
                            // have returned (its loads come next)
                            asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
                            lds_barrier();
                        }
#pragma unroll
                        for (int q = 0; q < 48; ++q) {
                            mfma_one(y, q);
                            if (load && q % 3 == 0) load_one(first + kt + 2, slot, q / 3);
                            if (more && q % 3 == 1) read_one(x, slot ^ 1, 0, q / 3);
                            __builtin_amdgcn_sched_barrier(0);
                        }
                    };
                    int64_t kt = 0;
                    for (; kt < stages - 2; ++kt) advance(kt, true, true);
                    for (; kt < stages - 1; ++kt) advance(kt, true, false);
                    advance(stages - 1, false, false);
                }
                // every wave has read its last fragments: the next tile's first stages may land
                lds_barrier();
                if (w + (int)gridDim.x < items) prologue(w + gridDim.x);

                // Epilogue: block (i, j) of this wave covers rows row0 + 32 i + l % 32 and columns
                // col0 + 32 j + 8 g + 4 (l / 32) + e for accumulator v = 4 g + e.
                const int64_t row0 = (int64_t)tm * kWide + wm * 128, col0 = (int64_t)tn * kWide + wn * 128;
                float *const c_base = static_cast<float *>(p.c) + (int64_t)(w / tiles) * p.slice_stride;
                const bool interior = row0 + 128 <= p.m && col0 + 128 <= p.n;
                if (vector_store && interior) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        float *row = c_base + (row0 + i * 32 + (lane & 31)) * p.ldc + col0 + 4 * half;
#pragma unroll
                        for (int j = 0; j < 4; ++j)
#pragma unroll
                            for (int g = 0; g < 4; ++g) {
                                const v4f value = {acc[i][j][4 * g], acc[i][j][4 * g + 1], acc[i][j][4 * g + 2],
                                                   acc[i][j][4 * g + 3]};
                                *reinterpret_cast<v4f *>(row + j * 32 + g * 8) = alpha == 1.0f ? value : alpha * value;
                            }
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int64_t row = row0 + i * 32 + (lane & 31);
                        if (row >= p.m) continue;
#pragma unroll
                        for (int j = 0; j < 4; ++j)
#pragma unroll
                            for (int v = 0; v < 16; ++v) {
                                const int64_t col = col0 + j * 32 + (v >> 2) * 8 + 4 * half + (v & 3);
                                if (col >= p.n) continue;
                                float *target = c_base + row * p.ldc + col;
                                const float value = acc[i][j][v];
                                *target = beta == 0.0f ? (alpha == 1.0f ? value : alpha * value)
                                                       : alpha * value + beta * *target;
                            }
                    }
                }
            }
        }

        // Packs one 32-value k-stage of one 256-row slab per workgroup. Thread t writes the 16-byte
        // positions t, t + 256, t + 512 and t + 768 of the stage (row q / 4, position q % 4, holding
        // chunk (q % 4) ^ ((row / 4) % 4), k 8 chunk ... + 7), so each wave writes 1 KiB in order.
        // Plain sources: each thread reads its chunk's eight values in place (two 16-byte loads when
        // the rows allow). Transposed sources: the workgroup first copies the 32 x 256 block, read
        // along the source's rows, into LDS.
        template <typename S> constexpr int transpose_rows() { return kSlabK; }
        template <typename S> struct Quad;
        template <> struct Quad<float> { typedef float4 type; };
        template <> struct Quad<double> { typedef double4 type; };

        template <typename S, bool split>
        __global__ __launch_bounds__(256) void pack_slab(const S *source, int64_t rows, int64_t depth, int64_t ld,
                                                         bool transpose, uint4 *out, uint4 *out_lo, int64_t kp) {
            typedef typename Quad<S>::type S4;
            constexpr int kStride = kSlabRows + 4; // LDS row stride, 16-byte aligned rows
            __shared__ __attribute__((aligned(16))) S tile[transpose_rows<S>()][kStride];
            const int64_t stage = blockIdx.x, row0 = (int64_t)blockIdx.y * kSlabRows, k0 = stage * kSlabK;
            const int t = threadIdx.x;
            // four consecutive elements of a source row at once when rows keep 16-byte alignment
            const bool quads = (ld & 3) == 0 && ((uintptr_t)source & (sizeof(S4) - 1)) == 0;
            if (transpose) { // element (i, k) = source[k * ld + i]
#pragma unroll
                for (int pass = 0; pass < kSlabK / 4; ++pass) {
                    const int kk = pass * 4 + (t >> 6), c = (t & 63) * 4;
                    const int64_t k = k0 + kk, i = row0 + c;
                    S4 value;
                    if (quads && k < depth && i + 3 < rows) {
                        value = *reinterpret_cast<const S4 *>(source + k * ld + i);
                    } else {
                        S v[4];
#pragma unroll
                        for (int e = 0; e < 4; ++e) v[e] = (k < depth && i + e < rows) ? source[k * ld + i + e] : S(0);
                        value = S4{v[0], v[1], v[2], v[3]};
                    }
                    *reinterpret_cast<S4 *>(&tile[kk][c]) = value;
                }
                __syncthreads();
            }
            uint4 *const stage_out = out + ((int64_t)blockIdx.y * (kp / kSlabK) + stage) * (kSlabRows * kSlabK / 8);
#pragma unroll
            for (int q4 = 0; q4 < 4; ++q4) {
                const int q = t + q4 * 256, r = q >> 2, chunk = slab_position(r, q & 3);
                const int64_t i = row0 + r, k = k0 + chunk * 8;
                S v[8];
                if (transpose) {
#pragma unroll
                    for (int e = 0; e < 8; ++e) v[e] = tile[chunk * 8 + e][r];
                } else if (quads && i < rows && k + 8 <= depth) {
                    const S4 lo = *reinterpret_cast<const S4 *>(source + i * ld + k);
                    const S4 hi = *reinterpret_cast<const S4 *>(source + i * ld + k + 4);
                    v[0] = lo.x; v[1] = lo.y; v[2] = lo.z; v[3] = lo.w;
                    v[4] = hi.x; v[5] = hi.y; v[6] = hi.z; v[7] = hi.w;
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) v[e] = (i < rows && k + e < depth) ? source[i * ld + k + e] : S(0);
                }
                // as pack_element: the value rounded to float, then to bf16; the low part the rounded
                // remainder, zero where the value is not finite
                __bf16 h[8], l[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float f = (float)v[e];
                    h[e] = (__bf16)f;
                    if constexpr (split) l[e] = (__bf16)(__builtin_isfinite(f) ? f - (float)h[e] : 0.0f);
                }
                stage_out[q] = __builtin_bit_cast(uint4, h);
                if constexpr (split) (out_lo + (stage_out - out))[q] = __builtin_bit_cast(uint4, l);
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
        // Compute units of the current device (the persistent kernel's grid), cached per device.
        int compute_units() {
            static int cached[64] = {};
            int device = 0;
            if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= 64) return 256;
            if (cached[device] == 0) {
                int count = 0;
                if (hipDeviceGetAttribute(&count, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || count <= 0)
                    count = 256;
                cached[device] = count;
            }
            return cached[device];
        }

        hipError_t launch_gemm(int mode, const ProductArgs &args, int slices, hipStream_t stream, bool alongside = false) {
            if (args.slabs) {
                if ((mode != kBfloat16 && mode != kSplit) || slices < 1 || args.kp % kSlabK != 0 ||
                    args.kp / kSlabK < slices)
                    return hipErrorInvalidValue;
                const int64_t items = ((args.m + kWide - 1) / kWide) * ((args.n + kWide - 1) / kWide) * slices;
                if (items == 0) return hipSuccess;
                const int grid = (int)std::min<int64_t>(items, compute_units());
                if (mode == kSplit)
                    hipLaunchKernelGGL(gemm_slabs<kSplit>, dim3(grid), dim3(kThreads), SlabRing<kSplit>::lds, stream,
                                       args, slices);
                else
                    hipLaunchKernelGGL(gemm_slabs<kBfloat16>, dim3(grid), dim3(kThreads), SlabRing<kBfloat16>::lds,
                                       stream, args, slices);
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

    int slab_slices(int mode, int64_t m, int64_t n, int64_t kp) {
        const int64_t tiles = ((m + kWide - 1) / kWide) * ((n + kWide - 1) / kWide), stages = kp / kSlabK;
        // a slab's stage offsets within 32 bits
        if ((mode != kBfloat16 && mode != kSplit) || tiles == 0 || kp % kSlabK != 0 ||
            kp * int64_t(kSlabRows) * 2 >= (int64_t(1) << 31))
            return 0;
        // enough 256 x 256 tiles to give every compute unit one, over at least 16 stages
        if (tiles >= 256) return stages >= 16 ? 1 : 0;
        // fewer tiles: split the stages so that tiles x slices fills the chip, each slice at least 32
        // stages (k >= 1024), at most 16 slices; below half the chip the 128 x 128 kernel does better
        const int64_t slices = std::min<int64_t>({(256 + tiles - 1) / tiles, stages / 32, 16});
        return slices >= 2 && tiles * slices >= 128 ? (int)slices : 0;
    }

    hipError_t pack_slabs(const void *source, bool wide, int64_t rows, int64_t depth, int64_t ld, bool transpose,
                          void *out, void *out_lo, int64_t kp, hipStream_t stream) {
        if (rows == 0 || kp == 0) return hipSuccess;
        if (kp % kSlabK != 0) return hipErrorInvalidValue;
        const dim3 grid((unsigned)(kp / kSlabK), (unsigned)((rows + kSlabRows - 1) / kSlabRows)), block(256);
        uint4 *hi = static_cast<uint4 *>(out), *lo = static_cast<uint4 *>(out_lo);
        if (wide) {
            const double *x = static_cast<const double *>(source);
            if (lo) hipLaunchKernelGGL((pack_slab<double, true>), grid, block, 0, stream, x, rows, depth, ld, transpose, hi, lo, kp);
            else hipLaunchKernelGGL((pack_slab<double, false>), grid, block, 0, stream, x, rows, depth, ld, transpose, hi, lo, kp);
        } else {
            const float *x = static_cast<const float *>(source);
            if (lo) hipLaunchKernelGGL((pack_slab<float, true>), grid, block, 0, stream, x, rows, depth, ld, transpose, hi, lo, kp);
            else hipLaunchKernelGGL((pack_slab<float, false>), grid, block, 0, stream, x, rows, depth, ld, transpose, hi, lo, kp);
        }
        return hipGetLastError();
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
