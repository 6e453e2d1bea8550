// The chunk bytecode interpreter for gfx950: one kernel template evaluates a chunk into an output
// array, another reduces its result. They replace the reference's Slang interpreter
// (shaders/interpreter.slang:11-96 with the entry points of shaders/common/entries.slang:17-80).
//
// Design (wave64, CDNA4):
//  - The program is a kernel argument. Every field is wave-uniform, so decoding an instruction
//    is scalar loads and scalar branches; the switch on the opcode never diverges.
//  - Inputs, constants, temporaries and the output share one register file (File below),
//    indexed by wave-uniform register numbers that the host assigns. Inputs are loaded for the
//    whole program before the first instruction runs, so all loads of a step are in flight at once.
//  - Each thread carries W consecutive elements (16 bytes of each input for W * sizeof(T) == 16),
//    so every load and store is one dwordx4 access and a wave moves 1 KiB per array per step.
//  - The last, partial step loads and stores lane by lane; nothing is read or written out of bounds.
#pragma once

#include "device_math.hpp"

namespace scihip {

    constexpr int kCapacity = 64;

    // One program as the device reads it. Register numbers: inputs [0, arguments), constants
    // [arguments, arguments + literals), then temporaries, then the output register.
    template <typename T> struct Program {
        uint32_t length;     // instructions to run
        uint32_t arguments;  // inputs
        uint32_t literals;   // constants
        uint32_t result;     // register of the value stored or reduced
        uint32_t fold;       // reduce: 0 sum, 1 minimum, 2 maximum
        uint32_t contract;   // reduce: 1 when the sum is a dot product of registers left / right
        uint32_t left;
        uint32_t right;
        uint64_t count;      // elements
        uint32_t aligned;    // every pointer is 16-byte aligned
        uint32_t reserved;
        uint32_t word[kCapacity];  // code | target << 8 | first << 16 | second << 24
        uint8_t third[kCapacity];  // third source
        T constants[kCapacity];
        const T *inputs[kCapacity];
        T *output;
    };

    template <typename T, int W> struct Lanes {
        T v[W];
    };

    // The register file: one array of R values per lane, each its own local array, so that the
    // compiler keeps it in VGPRs and reads it with wave-uniform indices (s_set_gpr_idx + v_mov)
    // instead of scratch memory. R * sizeof(T) must stay within 32 dwords per lane for that.
    template <typename T, int W, int R> struct File {
        T (&b0)[R];
        T (&b1)[R];
        T (&b2)[R];
        T (&b3)[R];
        __device__ __forceinline__ Lanes<T, W> get(uint32_t i) const {
            Lanes<T, W> l;
            l.v[0] = b0[i];
            if (W > 1) l.v[1 % W] = b1[i];
            if (W > 2) l.v[2 % W] = b2[i];
            if (W > 3) l.v[3 % W] = b3[i];
            return l;
        }
        __device__ __forceinline__ void set(uint32_t i, const Lanes<T, W> &l) const {
            b0[i] = l.v[0];
            if (W > 1) b1[i] = l.v[1 % W];
            if (W > 2) b2[i] = l.v[2 % W];
            if (W > 3) b3[i] = l.v[3 % W];
        }
    };

#define SCIHIP_FILE(T, W, R)                                                                                          \
    T bank0[R], bank1[W > 1 ? R : 1], bank2[W > 2 ? R : 1], bank3[W > 3 ? R : 1];                                     \
    const File<T, W, R> regs{bank0, *reinterpret_cast<T(*)[R]>(bank1), *reinterpret_cast<T(*)[R]>(bank2),             \
                             *reinterpret_cast<T(*)[R]>(bank3)}

    // One code over W lanes; `C` is a compile-time code, so apply() folds to that code's body.
    template <uint32_t C, typename T, int W>
    __device__ __forceinline__ Lanes<T, W> lanes(const Lanes<T, W> &a, const Lanes<T, W> &b, const Lanes<T, W> &c) {
        Lanes<T, W> o;
#pragma unroll
        for (int k = 0; k < W; ++k) o.v[k] = apply<T>(C, a.v[k], b.v[k], c.v[k]);
        return o;
    }

#define SCIHIP_CASE(C)                                                                                                 \
    case C:                                                                                                            \
        return lanes<C, T, W>(regs.get(first), regs.get(second), regs.get(third));

    // The runtime dispatch: a uniform switch, each case a straight-line body over W lanes that reads
    // only the registers its code uses (the reads of unused operands fold away with the code).
    template <typename T, int W, int R>
    __device__ __forceinline__ Lanes<T, W> dispatch(uint32_t code, const File<T, W, R> &regs, uint32_t first,
                                                    uint32_t second, uint32_t third) {
        switch (code) {
            SCIHIP_CASE(0) SCIHIP_CASE(1) SCIHIP_CASE(2) SCIHIP_CASE(3) SCIHIP_CASE(4) SCIHIP_CASE(5)
            SCIHIP_CASE(6) SCIHIP_CASE(7) SCIHIP_CASE(8) SCIHIP_CASE(9) SCIHIP_CASE(10) SCIHIP_CASE(11)
            SCIHIP_CASE(12) SCIHIP_CASE(13) SCIHIP_CASE(14) SCIHIP_CASE(15) SCIHIP_CASE(16) SCIHIP_CASE(17)
            SCIHIP_CASE(18) SCIHIP_CASE(19) SCIHIP_CASE(20) SCIHIP_CASE(21) SCIHIP_CASE(22) SCIHIP_CASE(23)
            SCIHIP_CASE(24) SCIHIP_CASE(25) SCIHIP_CASE(26) SCIHIP_CASE(27) SCIHIP_CASE(28) SCIHIP_CASE(29)
            SCIHIP_CASE(30) SCIHIP_CASE(31) SCIHIP_CASE(32) SCIHIP_CASE(33) SCIHIP_CASE(34) SCIHIP_CASE(35)
            SCIHIP_CASE(37) SCIHIP_CASE(38) SCIHIP_CASE(39)
        default:
            return lanes<kFuse, T, W>(regs.get(first), regs.get(second), regs.get(third));
        }
    }
#undef SCIHIP_CASE

    template <typename T, int W> struct Vector;
    template <> struct Vector<float, 4> {
        typedef float type __attribute__((ext_vector_type(4)));
    };
    template <> struct Vector<double, 2> {
        typedef double type __attribute__((ext_vector_type(2)));
    };

    // W elements at `at` of input `i` of a program (Program or Adjoint: both have inputs, count and
    // aligned); the specialised kernels' load.
    template <typename T, int W, typename P>
    __device__ __forceinline__ Lanes<T, W> fetch(const P &p, uint32_t i, uint64_t at, bool full) {
        const T *source = p.inputs[i] + at;
        Lanes<T, W> l;
        if constexpr (W * sizeof(T) == 16) {
            if (full && p.aligned) {
                typedef typename Vector<T, W>::type V;
                const V x = __builtin_nontemporal_load(reinterpret_cast<const V *>(source));
                __builtin_memcpy(&l, &x, sizeof(l));
                return l;
            }
        }
#pragma unroll
        for (int k = 0; k < W; ++k) l.v[k] = (full || at + k < p.count) ? source[k] : T(0);
        return l;
    }

    // The same as one 16-byte load, for W elements that lie inside an aligned input.
    template <typename T, int W, typename P>
    __device__ __forceinline__ Lanes<T, W> fetch_vector(const P &p, uint32_t i, uint64_t at) {
        static_assert(W * sizeof(T) == 16, "one 16-byte vector");
        typedef typename Vector<T, W>::type V;
        const V x = __builtin_nontemporal_load(reinterpret_cast<const V *>(p.inputs[i] + at));
        Lanes<T, W> l;
        __builtin_memcpy(&l, &x, sizeof(l));
        return l;
    }

    template <typename T, int W> __device__ __forceinline__ Lanes<T, W> splat(T value) {
        Lanes<T, W> l;
#pragma unroll
        for (int k = 0; k < W; ++k) l.v[k] = value;
        return l;
    }

    // Broadcasts the constants into registers [arguments, arguments + literals).
    template <typename T, int W, int R>
    __device__ __forceinline__ void constants(const Program<T> &p, const File<T, W, R> &regs) {
        for (uint32_t i = 0; i < p.literals; ++i) regs.set(p.arguments + i, splat<T, W>(p.constants[i]));
    }

    // Loads W elements at `at` of every input into registers [0, arguments).
    template <typename T, int W, int R>
    __device__ __forceinline__ void load(const Program<T> &p, const File<T, W, R> &regs, uint64_t at, bool full) {
        for (uint32_t i = 0; i < p.arguments; ++i) regs.set(i, fetch<T, W>(p, i, at, full));
    }

    // Runs the first `steps` instructions of a program (Program or Adjoint) on the register file.
    template <typename T, int W, int R, typename P>
    __device__ __forceinline__ void run(const P &p, const File<T, W, R> &regs, uint32_t steps) {
        // The third sources are read four to a dword: a byte load would be a vector load, whose value
        // the compiler cannot prove uniform (every register read through it would become a loop over
        // the lanes' indices), where a dword of the program is a scalar load.
        const uint32_t *thirds = reinterpret_cast<const uint32_t *>(p.third);
        for (uint32_t s = 0; s < steps; ++s) {
            const uint32_t word = p.word[s];
            const uint32_t third = (thirds[s >> 2] >> (8 * (s & 3))) & 0xFFu;
            regs.set((word >> 8) & 0xFFu, dispatch<T, W, R>(word & 0xFFu, regs, (word >> 16) & 0xFFu, word >> 24, third));
        }
    }

    // Writes W elements at `at` of an array of `count` elements (`aligned`: 16-byte aligned base).
    template <typename T, int W>
    __device__ __forceinline__ void store_to(T *base, uint64_t count, bool aligned, uint64_t at, bool full,
                                             const Lanes<T, W> &out) {
        T *target = base + at;
        if constexpr (W * sizeof(T) == 16) {
            if (full && aligned) {
                typedef typename Vector<T, W>::type V;
                V x;
                __builtin_memcpy(&x, &out, sizeof(x));
                __builtin_nontemporal_store(x, reinterpret_cast<V *>(target));
                return;
            }
        }
#pragma unroll
        for (int k = 0; k < W; ++k)
            if (full || at + k < count) target[k] = out.v[k];
    }

    // Writes W elements at `at` of the output.
    template <typename T, int W>
    __device__ __forceinline__ void store(const Program<T> &p, uint64_t at, bool full, const Lanes<T, W> &out) {
        store_to<T, W>(p.output, p.count, p.aligned != 0, at, full, out);
    }

    // The same as one 16-byte store, for W elements that lie inside an aligned output.
    template <typename T, int W>
    __device__ __forceinline__ void store_vector(const Program<T> &p, uint64_t at, const Lanes<T, W> &out) {
        static_assert(W * sizeof(T) == 16, "one 16-byte vector");
        typedef typename Vector<T, W>::type V;
        V x;
        __builtin_memcpy(&x, &out, sizeof(x));
        __builtin_nontemporal_store(x, reinterpret_cast<V *>(p.output + at));
    }

    // A reduction partial: the sum, or the extreme of the non-NaN values, and whether a NaN was seen.
    struct Partial {
        double value;
        uint32_t poisoned;
        uint32_t reserved;
    };

    // Everything a launch needs, one kernel argument for the interpreter and specialised kernels alike.
    template <typename T> struct Launch {
        Program<T> p;
        Partial *partials; // reduce: one per workgroup (folded by a second kernel, fold_partials)
    };

    // A chunk folded along one axis of a rows x columns row-major matrix (every input of the chunk is
    // such a matrix): output[r] folds row r (axis 1), output[c] folds column c (axis 0); one argument
    // for the interpreter and specialised kernels alike.
    template <typename T> struct AxisLaunch {
        Program<T> p;         // p.count = rows * columns; p.aligned: every row starts 16-byte aligned
        T *output;
        Partial *partials;    // axis 0: slices x columns
        uint64_t rows, columns;
    };

    __device__ __forceinline__ double combine(uint32_t fold, double a, double b) {
        return fold == 0 ? a + b : (fold == 1 ? (b < a ? b : a) : (b > a ? b : a));
    }

    __device__ __forceinline__ double identity(uint32_t fold) {
        return fold == 0 ? 0.0 : (fold == 1 ? __builtin_huge_val() : -__builtin_huge_val());
    }

    // Folds one value into a running sum / extreme. Sums propagate NaN by arithmetic; minimum and
    // maximum skip NaN and raise the flag, which makes the result NaN. `fold` is wave-uniform; the NaN
    // test is a select, not a branch, so lanes never diverge.
    __device__ __forceinline__ void fold_in(uint32_t fold, double &total, uint32_t &poisoned, double x) {
        if (fold == 0) {
            total += x;
            return;
        }
        const bool nan = unordered(x);
        poisoned |= nan ? 1u : 0u;
        const double y = nan ? total : x;
        total = fold == 1 ? (y < total ? y : total) : (y > total ? y : total);
    }

    // Folds W lanes of a result into a thread's running sum / extreme.
    template <typename T, int W>
    __device__ __forceinline__ void consume(uint32_t fold, double &total, uint32_t &poisoned, const Lanes<T, W> &r,
                                            uint64_t at, bool full, uint64_t count) {
#pragma unroll
        for (int k = 0; k < W; ++k)
            if (full || at + k < count) fold_in(fold, total, poisoned, (double)r.v[k]);
    }

    // A contracted sum: sum(a * b) with each product exact in double (two floats' product always is).
    template <typename T, int W>
    __device__ __forceinline__ void contract(double &total, const Lanes<T, W> &a, const Lanes<T, W> &b, uint64_t at,
                                             bool full, uint64_t count) {
#pragma unroll
        for (int k = 0; k < W; ++k)
            if (full || at + k < count) total = fused((double)a.v[k], (double)b.v[k], total);
    }

    // Folds a value across the workgroup: DPP / swizzle shuffles within each wave, then the waves'
    // values in LDS in wave order. The result is valid in thread 0.
    __device__ __forceinline__ Partial fold_block(uint32_t fold, double value, uint32_t poisoned) {
#pragma unroll
        for (int offset = 32; offset > 0; offset >>= 1) {
            value = combine(fold, value, __shfl_xor(value, offset, 64));
            poisoned |= __shfl_xor(poisoned, offset, 64);
        }
        __shared__ double values[16];
        __shared__ uint32_t flags[16];
        const uint32_t wave = threadIdx.x >> 6;
        __syncthreads();
        if ((threadIdx.x & 63) == 0) {
            values[wave] = value;
            flags[wave] = poisoned;
        }
        __syncthreads();
        Partial out{0.0, 0u, 0u};
        if (threadIdx.x == 0) {
            const uint32_t waves = (blockDim.x + 63) >> 6;
            double v = values[0];
            uint32_t f = flags[0];
            for (uint32_t w = 1; w < waves; ++w) {
                v = combine(fold, v, values[w]);
                f |= flags[w];
            }
            out.value = v;
            out.poisoned = f;
        }
        return out;
    }

    // Ends a reduction kernel: each workgroup leaves its Partial at its index. A second kernel
    // (fold_partials) folds them in workgroup order; the grid is fixed for a given size, so the result
    // is deterministic. (A last-arriving workgroup folding them instead needs a release fence and an
    // atomic per workgroup, which measured 40 us on a 2^28-element dot product.)
    __device__ __forceinline__ void leave_partial(uint32_t fold, Partial *partials, double total, uint32_t poisoned) {
        const Partial mine = fold_block(fold, total, poisoned);
        if (threadIdx.x == 0) partials[blockIdx.x] = mine;
    }

    // The fold of `count` partials: thread t takes partials t, t + 256, ... in order, then the
    // workgroup folds across threads. Valid in thread 0.
    __device__ __forceinline__ Partial fold_partials_block(uint32_t fold, const Partial *partials, uint32_t count) {
        double value = identity(fold);
        uint32_t flag = 0;
        for (uint32_t b = threadIdx.x; b < count; b += blockDim.x) {
            const Partial other = partials[b];
            value = combine(fold, value, other.value);
            flag |= other.poisoned;
        }
        return fold_block(fold, value, flag);
    }

    template <typename T, int W, int R> __global__ __launch_bounds__(256) void evaluate(const Launch<T> a) {
        const Program<T> &p = a.p;
        SCIHIP_FILE(T, W, R);
        constants<T, W, R>(p, regs);
        const uint64_t steps = (p.count + W - 1) / W;
        const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
        for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < steps; v += stride) {
            const uint64_t at = v * W;
            const bool full = at + W <= p.count;
            load<T, W, R>(p, regs, at, full);
            run<T, W, R>(p, regs, p.length);
            store<T, W>(p, at, full, regs.get(p.result));
        }
    }

    template <typename T, int W, int R> __global__ __launch_bounds__(256) void reduce(const Launch<T> a) {
        const Program<T> &p = a.p;
        SCIHIP_FILE(T, W, R);
        constants<T, W, R>(p, regs);
        double total = identity(p.fold);
        uint32_t poisoned = 0;
        const uint64_t steps = (p.count + W - 1) / W;
        const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
        const uint32_t length = p.contract ? p.length - 1 : p.length;
        for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < steps; v += stride) {
            const uint64_t at = v * W;
            const bool full = at + W <= p.count;
            load<T, W, R>(p, regs, at, full);
            run<T, W, R>(p, regs, length);
            if (p.contract) contract<T, W>(total, regs.get(p.left), regs.get(p.right), at, full, p.count);
            else consume<T, W>(p.fold, total, poisoned, regs.get(p.result), at, full, p.count);
        }
        leave_partial(p.fold, a.partials, total, poisoned);
    }

} // namespace scihip
