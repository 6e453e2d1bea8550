// The fused forward + backward kernel of a chunk's adjoint, for gfx950.
//
// An adjoint program is chunk bytecode over a larger register file: the chunk's forward
// instructions, then the backward instructions that the host emits from them in reverse order
// (differentiate.hip). Each thread loads its elements of every input once, recomputes the forward
// values in registers, runs the backward instructions, stores the gradient of every input and,
// when asked, folds the loss (the sum of the chunk's result) as a reduction does. Nothing between
// the forward and backward pass touches memory: 4 B read per input and 4 B written per gradient
// per float element are the whole traffic.
//
// The program layout follows Program (interpreter.hpp): registers [0, arguments) are the inputs,
// [arguments, arguments + literals) the constants, then the values the instructions write. The
// instruction word and the third-source byte are Program's; codes kSignTimes, kRouteMinimum and
// kRouteMaximum appear only here.
#pragma once

#include "interpreter.hpp"

namespace scihip {

    constexpr int kAdjointLength = 256;      // instructions of an adjoint program at most
    constexpr int kAdjointInputs = kCapacity + 1;   // the chunk's inputs and the upstream gradient
    constexpr int kAdjointLiterals = kCapacity + 8; // the chunk's constants and the backward ones

    template <typename T> struct Adjoint {
        uint32_t length;     // instructions: the forward pass, then the backward pass
        uint32_t arguments;  // inputs; a seeded program's last input is the upstream gradient
        uint32_t literals;   // constants
        uint32_t gradients;  // gradient arrays: input i's gradient is register slot[i], stored to outputs[i]
        uint32_t loss;       // 1: fold the loss
        uint32_t contract;   // 1: the loss is the sum of left * right, each product exact in double
        uint32_t left;       // the loss register (or the left factor)
        uint32_t right;      // the right factor
        uint64_t count;      // elements
        uint32_t aligned;    // every pointer is 16-byte aligned
        uint32_t reserved;
        uint32_t word[kAdjointLength];  // code | target << 8 | first << 16 | second << 24
        uint8_t third[kAdjointLength];  // third source
        uint8_t slot[kCapacity];
        T constants[kAdjointLiterals];
        const T *inputs[kAdjointInputs];
        T *outputs[kCapacity];          // null: that gradient is not stored
    };

    template <typename T> struct AdjointLaunch {
        Adjoint<T> p;
        Partial *partials; // loss: one per workgroup (folded by fold_partials)
    };

    // The adjoint interpreter: the register file of R values per lane stays in VGPRs for R * sizeof(T)
    // of at most 32 dwords (File in interpreter.hpp); wider files run one lane per thread.
    template <typename T, int W, int R> __global__ __launch_bounds__(256) void differentiate(const AdjointLaunch<T> a) {
        const Adjoint<T> &p = a.p;
        SCIHIP_FILE(T, W, R);
        for (uint32_t i = 0; i < p.literals; ++i) regs.set(p.arguments + i, splat<T, W>(p.constants[i]));
        double total = 0.0;
        uint32_t poisoned = 0;
        const uint64_t steps = (p.count + W - 1) / W;
        const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
        for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < steps; v += stride) {
            const uint64_t at = v * W;
            const bool full = at + W <= p.count;
            for (uint32_t i = 0; i < p.arguments; ++i) regs.set(i, fetch<T, W>(p, i, at, full));
            run<T, W, R>(p, regs, p.length);
            if (p.loss) {
                if (p.contract) contract<T, W>(total, regs.get(p.left), regs.get(p.right), at, full, p.count);
                else consume<T, W>(0, total, poisoned, regs.get(p.left), at, full, p.count);
            }
            for (uint32_t g = 0; g < p.gradients; ++g)
                if (p.outputs[g]) store_to<T, W>(p.outputs[g], p.count, p.aligned != 0, at, full, regs.get(p.slot[g]));
        }
        if (p.loss) leave_partial(0, a.partials, total, poisoned);
    }

} // namespace scihip
