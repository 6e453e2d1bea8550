// LU with partial pivoting on the GPU, triangular solves and mixed-precision iterative refinement.
//
// The reference factors and solves on the host only (include/linear/decompose.hpp:29-72,
// include/linear/solve.hpp:25-76). The same row-major, in-place algorithm runs here blocked:
// for each panel of kPanel columns,
//   1. the panel is factored by one kernel that keeps each workgroup's rows in LDS: per column every
//      workgroup publishes its best pivot candidate (first maximum of |a|, the reference's rule)
//      with the candidate's panel row, every workgroup gathers them all and picks the same pivot,
//      and all workgroups swap and eliminate;
//   2. the panel's row swaps are applied to the columns outside it;
//   3. U12 = L11^-1 A12 by substitution in the reference's order, in the matrix's own precision;
//   4. A22 -= L21 U12 on the matrix cores (fp64, or fp32 / bf16x3 / bf16 for a float matrix): first
//      the next panel's columns, then that panel is factored on the main stream while the rest of
//      the update runs on a second stream (look-ahead), which the next step waits for. Steps go in
//      pairs: a step carries most of its update into the next one, which applies both steps'
//      updates as products of depth 2 kPanel, for half the traffic of C.
// Solves of an order that is a multiple of kBlock invert the kBlock x kBlock diagonal blocks of L and
// U once (all blocks in parallel) and link each block to the blocks one and two away, then solve each
// triangle in one launch that walks the blocks, four workgroups per block, each waiting on the blocks
// it depends on as they are published. Other orders run a sync-free blocked
// substitution: one workgroup per kSolveRows-row block, each waiting on the blocks before it with a
// bounded spin. Every spin in this file is bounded: past kSpinLimit the kernel records an error and
// exits, and the host reports it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace scihip {

    constexpr int kPanel = 128;       // panel width (columns factored per cooperative launch)
    constexpr int kPanelRows = 128;   // rows per workgroup (four waves, a thread per row) in the panel kernel
    constexpr int kSolveRows = 48;    // rows per workgroup in the triangular solves
    constexpr int kSolveColumns = 32; // right-hand sides per pass of the solve kernel
    constexpr int kBlock = 128;       // diagonal block of the blocked substitution

    // Whether an order takes the blocked substitution (whole blocks, rows 16-byte aligned).
    constexpr bool blocked_substitution(int64_t order) { return order >= 2 * kBlock && order % kBlock == 0; }

    // Device-side state the LU kernels share: barrier counters, error flag, publications.
    struct Workspace {
        unsigned *counter = nullptr; // grid barrier arrivals (zeroed before each launch)
        unsigned *error = nullptr;   // set when a bounded spin gave up
        unsigned *flags = nullptr;   // solve: per-block ready epochs
        void *publish = nullptr;     // panel: published candidate and diagonal rows (panel_publish_bytes)
        size_t publish_bytes = 0;
        unsigned launches = 0;       // panel launches so far, for the publications' tags
        size_t flag_count = 0;
        hipStream_t side = nullptr;         // updates beside the look-ahead panels (default priority)
        hipEvent_t columns_ready = nullptr; // main stream: the next panel's columns are up to date
        hipEvent_t panel_done = nullptr;    // side stream: the step's work there is done
        hipStream_t third = nullptr;        // diagonal-block inverses of finished blocks, beside the last panels
        hipEvent_t inverted = nullptr;      // third stream: those inverses are done
        hipEvent_t upper_inverted = nullptr; // third stream: U's inverses and links of the last blocks are done
        bool upper_pending = false;         // the next substitute() waits for upper_inverted before U's pass
        bool head_pending = false;          // factor() waits for panel_done (recorded by its caller on the side
                                            // stream) before its first step's work
    };

    // Factors the order x order row-major matrix in place (T = float or double); pivots[k] is the
    // row swapped with row k at step k (absolute). `mode` is the product mode of the trailing update
    // (kFloat64 for double; kFloat32, kSplit or kBfloat16 for float). Workspace buffers are passed in.
    // With `inverses` (blocked_substitution(order) only), also makes invert_blocks' output there: the
    // blocks that are final early (from kInvertTail panels before the end) on Workspace::third beside the
    // last panels, the others after the last one.
    template <typename T>
    int factor(T *matrix, int64_t order, int64_t *pivots, int mode, Workspace &ws, void *scratch, size_t scratch_bytes,
               hipStream_t stream, const char **error, T *inverses = nullptr);

    // Bytes of scratch factor() needs.
    size_t factor_scratch(int64_t order, int mode);

    // Bytes of Workspace::publish the panel kernel needs for an order (element: 4 or 8 bytes). The
    // buffer is zeroed when allocated.
    size_t panel_publish_bytes(int64_t order, size_t element);

    // Solves L U x = P b for `width` right-hand sides (row-major order x width), in place in `x`,
    // which holds P b on entry. `epoch` is the device's running launch count for the ready flags;
    // each launch takes the next value, so a flag never satisfies a wait of a later launch early.
    // With `inverses` (6 * order * kBlock elements from invert_blocks) and `scratch`
    // (substitute_scratch(order) bytes) and an order that takes it, the blocked substitution runs.
    template <typename T>
    int substitute(const T *factors, int64_t order, T *x, int64_t width, Workspace &ws, unsigned &epoch,
                   hipStream_t stream, const char **error, const T *inverses = nullptr, void *scratch = nullptr);

    // Bytes of scratch the blocked substitution needs for an order (tagged copies of the solved blocks,
    // and of the blocks' reduced right-hand sides that the four workgroups of a block exchange).
    size_t substitute_scratch(int64_t order);

    // The transposed inverses of the diagonal blocks of L (first order * kBlock elements) and of U
    // (the next order * kBlock) of packed LU factors, then the links inv(D_jj) L_j,j-1 and inv(D_jj)
    // U_j,j+1, then inv(D_jj) L_j,j-2 and inv(D_jj) U_j,j+2 (row-major, order * kBlock elements each),
    // for the blocked substitution: 6 * order * kBlock elements.
    template <typename T> hipError_t invert_blocks(const T *factors, int64_t order, T *inverses, hipStream_t stream);

    // r -= A x for a narrow x (order x width, row-major, width <= 32) and an order that takes the
    // blocked substitution: A (order x order, row-major fp64) is read in place once, on the fp64 matrix
    // cores, over kResidualSlices slices of its columns; `partials` holds kResidualSlices * order *
    // width doubles.
    constexpr int kResidualSlices = 8;
    bool narrow_residual_fits(int64_t order, int64_t width);
    hipError_t narrow_residual(const double *a, int64_t order, const double *x, int64_t width, double *r,
                               double *partials, hipStream_t stream);

    // Applies the row order of `pivots` to `source` (order x width) into `target`: target row k is
    // source row rows[k], where rows is the identity permuted by the swaps pivots[0..order).
    template <typename T, typename S>
    hipError_t permute(const S *source, T *target, const int64_t *rows, int64_t order, int64_t width,
                       hipStream_t stream);

} // namespace scihip
