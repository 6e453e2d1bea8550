// Matrix products on the gfx950 matrix cores: C = alpha * A * B + beta * C, row-major, in four
// modes, replacing the reference's FMA-only product shader (shaders/product.slang:35-131):
//   kFloat32  fp32 operands on v_mfma_f32_16x16x4_f32 (fp32 accumulation)
//   kBfloat16 operands rounded to bf16, v_mfma_f32_16x16x32_bf16 (fp32 accumulation)
//   kSplit    bf16x3: a = a_hi + a_lo with both parts bf16, and a*b ~ a_hi*b_hi + a_hi*b_lo + a_lo*b_hi
//             (three bf16 MFMAs per step, fp32 accumulation)
//   kFloat64  fp64 operands on v_mfma_f64_16x16x4_f64
// Operands are first packed (packing kernels below): A as M x Kp and B transposed as N x Kp, both
// with the shared dimension contiguous and zero-padded to a whole number of k-tiles, in the MFMA
// input type. The product kernel then streams both through LDS the same way ("TN" layout).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace scihip {

    enum ProductMode : int { kFloat32 = 0, kBfloat16 = 1, kSplit = 2, kFloat64 = 3 };

    struct ProductArgs {
        const void *a_hi; // M x Kp, row stride lda (elements)
        const void *a_lo; // kSplit only
        const void *b_hi; // N x Kp (B transposed), row stride ldb
        const void *b_lo; // kSplit only
        void *c;          // M x N, row stride ldc; float (fp32 / bf16 modes) or double (fp64)
        int64_t m, n, kp;
        int64_t lda, ldb, ldc;
        double alpha, beta;
        int64_t slice_stride; // split-K launches: elements between consecutive slices' outputs
        int slabs = 0;       // kBfloat16 and kSplit: the operands are packed in slabs (pack_slabs), lda/ldb unused
    };

    // Slab layout of the large-tile bf16 product: an operand of R rows over Kp is ceil(R / 256)
    // slabs of 256 rows (rows past R zero), each Kp / 32 stages of 16 KiB, one stage holding the 32
    // k values of all 256 rows: row r at byte 64 r, its 16-byte chunk c (k 8 c ... 8 c + 7) at
    // position c ^ ((r / 4) % 4). One stage is one contiguous block, copied to LDS as it lies.
    constexpr int kSlabRows = 256;
    constexpr int kSlabK = 32;
    inline size_t slab_bytes(int64_t rows, int64_t kp) {
        return size_t((rows + kSlabRows - 1) / kSlabRows) * kSlabRows * size_t(kp) * 2;
    }

    // Whether a product of an m x n output over kp (a multiple of k_tile(mode)) takes the large-tile
    // kernel, and over how many slices of its shared dimension (0: it does not): bf16 and bf16x3; 1 with at
    // least one 256 x 256 tile per compute unit; fewer tiles split the stages (product_split) so that
    // tiles x slices fills the chip.
    int slab_slices(int mode, int64_t m, int64_t n, int64_t kp);

    // Packs a row-major float (wide: double) source into slabs, as pack() does into rows: element
    // (i, k) is src[i][k] (transpose false) or src[k][i] (transpose true), rounded to float and then
    // to bf16, zero for k >= depth or i >= rows; with out_lo (bf16x3) also the low parts, as
    // pack_element makes them.
    hipError_t pack_slabs(const void *source, bool wide, int64_t rows, int64_t depth, int64_t ld, bool transpose,
                          void *out, void *out_lo, int64_t kp, hipStream_t stream);

    // Bytes of one k-tile row in LDS and the k extent it holds, per mode.
    constexpr int k_tile(int mode) { return mode == kFloat32 ? 32 : mode == kFloat64 ? 16 : mode == kSplit ? 32 : 64; }
    constexpr int element_bytes(int mode) { return mode == kFloat32 ? 4 : mode == kFloat64 ? 8 : 2; }

    // Launches the product of packed operands. Kp must be a multiple of k_tile(mode).
    hipError_t product(int mode, const ProductArgs &args, hipStream_t stream);

    // The same, with the fp64 kernel sized for one workgroup per compute unit instead of two, for a
    // product that runs beside a kernel whose workgroups each need most of a compute unit's LDS (the
    // fp64 LU panel): as the product's workgroups finish, whole compute units come free for it.
    hipError_t product_alongside(int mode, const ProductArgs &args, hipStream_t stream);

    // The same product with the k-tiles split into `slices` ranges run by separate workgroups (for
    // products with few output tiles and a long shared dimension): each slice writes its partial
    // product to `partials` (slices x m x n, in the output type), then a second kernel folds the
    // slices in order into C with alpha and beta, so the result does not depend on timing.
    hipError_t product_split(int mode, const ProductArgs &args, int slices, void *partials, hipStream_t stream);

    // How many slices product_split should use for an m x n output over kp: 1 unless the output has
    // fewer tiles than the chip has CUs; then enough for about four workgroups per CU, each slice
    // keeping at least 8 k-tiles, at most 16 slices.
    int split_slices(int mode, int64_t m, int64_t n, int64_t kp);

    // Stores one element of a packed operand at hi[at] (and, for the split form, lo[at]): D is the
    // matrix-core input type; a bf16 high part is the value rounded to float and then to bf16, the
    // low part the rounded remainder, zero where the value is not finite.
    template <typename D, bool split, typename S>
    __device__ __forceinline__ void pack_element(S x, D *hi, D *lo, int64_t at) {
        if constexpr (sizeof(D) == 2) {
            const float f = (float)x;
            const __bf16 h = (__bf16)f;
            hi[at] = h;
            if constexpr (split) {
                const float rest = __builtin_isfinite(f) ? f - (float)h : 0.0f;
                lo[at] = (__bf16)rest;
            }
        } else {
            hi[at] = (D)x;
        }
    }

    // Packs a row-major source into the product's operand layout: out[i][k] = src[i][k] (transpose
    // false) or src[k][i] (transpose true), for i < rows and k < depth, zero for depth <= k < kp.
    // `source` is float or double (by `wide`); the output is in the mode's input type; kSplit also
    // writes the low parts to `lo`. Output rows are `stride` elements apart (0: kp), so that two
    // packs can fill the two halves of wider rows.
    hipError_t pack(int mode, const void *source, bool wide, int64_t rows, int64_t depth, int64_t ld, bool transpose,
                    void *hi, void *lo, int64_t kp, hipStream_t stream, int64_t stride = 0);

} // namespace scihip
