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
    };

    // Bytes of one k-tile row in LDS and the k extent it holds, per mode.
    constexpr int k_tile(int mode) { return mode == kFloat32 ? 32 : mode == kFloat64 ? 16 : mode == kSplit ? 32 : 64; }
    constexpr int element_bytes(int mode) { return mode == kFloat32 ? 4 : mode == kFloat64 ? 8 : 2; }

    // Launches the product of packed operands. Kp must be a multiple of k_tile(mode).
    hipError_t product(int mode, const ProductArgs &args, hipStream_t stream);

    // Packs a row-major source into the product's operand layout: out[i][k] = src[i][k] (transpose
    // false) or src[k][i] (transpose true), for i < rows and k < depth, zero for depth <= k < kp.
    // `source` is float or double (by `wide`); the output is in the mode's input type; kSplit also
    // writes the low parts to `lo`.
    hipError_t pack(int mode, const void *source, bool wide, int64_t rows, int64_t depth, int64_t ld, bool transpose,
                    void *hi, void *lo, int64_t kp, hipStream_t stream);

} // namespace scihip
