// Harness over the reference's own compiled CPU kernels (oracle/_ref). TEST INFRASTRUCTURE ONLY.
//
// oracle/Makefile compiles, from the sources where they lie under /root/reference, the kernel
// tables src/core/drivers/processor/kernels/{avx512,fallback}.cpp and the thread pool
// src/core/drivers/processor/{pool,context}.cpp, and this file against the reference's headers.
// The reference's Chunk (include/core/bytecode/chunk.hpp), Accumulator
// (include/core/bytecode/accumulator.hpp) and Plan (include/core/drivers/processor/plan.hpp) are
// used as they are. Dispatch (src/core/drivers/processor/dispatch.cpp) cannot be compiled here (it
// needs asmjit through include/core/drivers/processor/compiler.hpp:11), so `run` and `fold` below
// drive the reference's kernel tables in the order dispatch.cpp:42-165 does with the JIT off
// (SCI_JIT=0: src/core/drivers/processor/compiler.cpp:50).
//
// A chunk is recorded from a postfix program, exactly as Node::record walks an expression tree
// (include/array/node.hpp:75-88): operands first, left to right, then the operation.

#include "core/bytecode/accumulator.hpp"
#include "core/bytecode/chunk.hpp"
#include "core/drivers/processor/kernels/avx512.hpp"
#include "core/drivers/processor/kernels/fallback.hpp"
#include "core/drivers/processor/kernels/table.hpp"
#include "core/drivers/processor/plan.hpp"
#include "core/drivers/processor/pool.hpp"

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "scihip.h"

using namespace sci::core;
using namespace sci::core::drivers::processor;

// The C ABI's chunk is the reference's Chunk for every field a binding reads
// (include/core/bytecode/chunk.hpp:178-196): the same types, sizes and offsets, field by field, and
// the same numbering of codes, value-id tags, folds and precisions
// (include/core/bytecode/code.hpp:18-61, instruction.hpp:11-35, fold.hpp:11, precision.hpp:12). A
// binding therefore copies the public fields (or the whole public prefix) with memcpy and re-encodes
// nothing; this file does not compile if that stops being true. The reference's Chunk has private
// members after the public ones (chunk.hpp:226-230), so only the public prefix is compared;
// offsetof on it is conditionally supported (mixed access control), and clang supports it.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winvalid-offsetof"
#define SAME_FIELD(field)                                                                                            \
    static_assert(offsetof(Chunk, field) == offsetof(scihip_chunk, field), "offset of " #field);                   \
    static_assert(sizeof(Chunk::field) == sizeof(scihip_chunk::field), "size of " #field)
SAME_FIELD(precision);
SAME_FIELD(length);
SAME_FIELD(arguments);
SAME_FIELD(literals);
SAME_FIELD(temporaries);
SAME_FIELD(result);
SAME_FIELD(instructions);
SAME_FIELD(inputs);
SAME_FIELD(constants);
#undef SAME_FIELD
static_assert(offsetof(Chunk, constants) + sizeof(Chunk::constants) == sizeof(scihip_chunk),
              "scihip_chunk is exactly Chunk's public prefix");
static_assert(offsetof(Instruction, code) == offsetof(scihip_instruction, code) &&
                      offsetof(Instruction, target) == offsetof(scihip_instruction, target) &&
                      offsetof(Instruction, sources) == offsetof(scihip_instruction, sources) &&
                      sizeof(Instruction) == sizeof(scihip_instruction) &&
                      alignof(Instruction) == alignof(scihip_instruction),
              "Instruction layout");
#pragma clang diagnostic pop
static_assert(Chunk::capacity == SCIHIP_CAPACITY, "capacity");
static_assert(slot::input == SCIHIP_SLOT_INPUT && slot::literal == SCIHIP_SLOT_LITERAL &&
                      slot::temporary == SCIHIP_SLOT_TEMPORARY && slot::output == SCIHIP_SLOT_OUTPUT &&
                      slot::none == SCIHIP_SLOT_NONE && slot::types == SCIHIP_SLOT_TYPES,
              "value-id tags");
static_assert(static_cast<int>(Fold::sum) == SCIHIP_FOLD_SUM && static_cast<int>(Fold::minimum) == SCIHIP_FOLD_MINIMUM &&
                      static_cast<int>(Fold::maximum) == SCIHIP_FOLD_MAXIMUM,
              "folds");
static_assert(static_cast<int>(Precision::single) == SCIHIP_SINGLE && static_cast<int>(Precision::wide) == SCIHIP_WIDE,
              "precisions");
#define SAME_CODE(name, abi) static_assert(static_cast<int>(Code::name) == abi, "code " #name)
SAME_CODE(copy, SCIHIP_COPY); SAME_CODE(negate, SCIHIP_NEGATE); SAME_CODE(absolute, SCIHIP_ABSOLUTE);
SAME_CODE(root, SCIHIP_ROOT); SAME_CODE(exponential, SCIHIP_EXPONENTIAL); SAME_CODE(logarithm, SCIHIP_LOGARITHM);
SAME_CODE(sin, SCIHIP_SIN); SAME_CODE(cos, SCIHIP_COS); SAME_CODE(tan, SCIHIP_TAN); SAME_CODE(asin, SCIHIP_ASIN);
SAME_CODE(acos, SCIHIP_ACOS); SAME_CODE(atan, SCIHIP_ATAN); SAME_CODE(csc, SCIHIP_CSC); SAME_CODE(sec, SCIHIP_SEC);
SAME_CODE(cot, SCIHIP_COT); SAME_CODE(acsc, SCIHIP_ACSC); SAME_CODE(asec, SCIHIP_ASEC); SAME_CODE(acot, SCIHIP_ACOT);
SAME_CODE(sinh, SCIHIP_SINH); SAME_CODE(cosh, SCIHIP_COSH); SAME_CODE(tanh, SCIHIP_TANH);
SAME_CODE(asinh, SCIHIP_ASINH); SAME_CODE(acosh, SCIHIP_ACOSH); SAME_CODE(atanh, SCIHIP_ATANH);
SAME_CODE(csch, SCIHIP_CSCH); SAME_CODE(sech, SCIHIP_SECH); SAME_CODE(coth, SCIHIP_COTH);
SAME_CODE(acsch, SCIHIP_ACSCH); SAME_CODE(asech, SCIHIP_ASECH); SAME_CODE(acoth, SCIHIP_ACOTH);
SAME_CODE(add, SCIHIP_ADD); SAME_CODE(subtract, SCIHIP_SUBTRACT); SAME_CODE(multiply, SCIHIP_MULTIPLY);
SAME_CODE(divide, SCIHIP_DIVIDE); SAME_CODE(minimum, SCIHIP_MINIMUM); SAME_CODE(maximum, SCIHIP_MAXIMUM);
SAME_CODE(fuse, SCIHIP_FUSE); SAME_CODE(count, SCIHIP_CODES);
#undef SAME_CODE

namespace {

    // One step of a postfix program: kind 0 = input (value = input number, recorded by a distinct
    // pointer per number), 1 = constant (value), 2 = operation (code, arity).
    struct Step {
        int32_t kind;
        int32_t code;
        int32_t arity;
        int32_t input;
        double value;
    };

    // Distinct dummy addresses stand for the input arrays while recording; only identity matters.
    char addresses[Chunk::capacity * 2];

    bool record(const Step *steps, int count, int finish, Chunk &chunk) {
        std::vector<uint8> stack;
        for (int i = 0; i < count; ++i) {
            const Step &s = steps[i];
            if (s.kind == 0) {
                stack.push_back(chunk.input(addresses + s.input));
            } else if (s.kind == 1) {
                stack.push_back(chunk.constant(s.value));
            } else {
                if (static_cast<int>(stack.size()) < s.arity) return false;
                uint8 a[3] = {slot::none, slot::none, slot::none};
                for (int k = s.arity - 1; k >= 0; --k) {
                    a[k] = stack.back();
                    stack.pop_back();
                }
                stack.push_back(chunk.emit(static_cast<Code>(s.code), a[0], a[1], a[2]));
            }
        }
        if (stack.size() != 1) return false;
        if (finish == 0) chunk.output(stack.back());
        else chunk.reduce(stack.back());
        return true;
    }

    const kernels::Table *table(int isa) {
        if (isa == 2 && kernels::avx512) return kernels::avx512;
        return kernels::fallback;
    }

    constexpr size block = Plan::block;

    template <typename type, typename visitor>
    void sweep(const kernels::Routines<type> &routines, const Chunk &chunk, type *const *inputs, type *output,
               size first, size last, uint8 steps, visitor &&visit) {
        type *values[256];
        values[slot::none] = nullptr;
        values[slot::output] = output;
        std::vector<type> space((chunk.literals + chunk.temporaries) * block);
        for (size i = 0; i < chunk.literals; ++i) {
            std::fill_n(space.data() + i * block, block, static_cast<type>(chunk.constants[i]));
            values[slot::literal | i] = space.data() + i * block;
        }
        for (size i = 0; i < chunk.temporaries; ++i)
            values[slot::temporary | i] = space.data() + (chunk.literals + i) * block;
        for (size start = first; start < last; start += block) {
            const size n = std::min(block, last - start);
            for (size i = 0; i < chunk.arguments; ++i) values[slot::input | i] = inputs[i] + start;
            if (output) values[slot::output] = output + start;
            for (uint8 s = 0; s < steps; ++s) {
                const Instruction &step = chunk.instructions[s];
                const type *sources[3] = {values[step.sources[0]], values[step.sources[1]], values[step.sources[2]]};
                routines.kernels[static_cast<size>(step.code)](sources, values[step.target], n);
            }
            visit(values, n);
        }
    }

    // Maps a recorded chunk's input ids back to program input numbers.
    template <typename type> std::vector<type *> bind(const Chunk &chunk, type *const *arrays) {
        std::vector<type *> inputs(chunk.arguments);
        for (size i = 0; i < chunk.arguments; ++i)
            inputs[i] = arrays[static_cast<const char *>(chunk.inputs[i]) - addresses];
        return inputs;
    }

    template <typename type>
    void run(const kernels::Routines<type> &routines, const Chunk &chunk, type *const *arrays, type *output,
             size count) {
        const std::vector<type *> inputs = bind(chunk, arrays);
        const Plan plan(chunk, count);
        auto work = [&](const size part) {
            const Range r = plan.range(part);
            if (r.first < r.last)
                sweep(routines, chunk, inputs.data(), output, r.first, r.last, chunk.length, [](type *const *, size) {});
        };
        if (plan.parts == 1) work(0);
        else Pool::global().run(plan.parts, work, plan.seats);
    }

    template <typename type>
    double fold(const kernels::Routines<type> &routines, const Chunk &chunk, type *const *arrays, Fold how,
                size count) {
        if (count == 0) return how == Fold::sum ? 0.0 : std::numeric_limits<double>::quiet_NaN();
        const std::vector<type *> inputs = bind(chunk, arrays);
        const Instruction *product = chunk.length > 0 ? &chunk.instructions[chunk.length - 1] : nullptr;
        const bool contract = chunk.contracted(how);
        const uint8 steps = contract ? chunk.length - 1 : chunk.length;
        const auto reduction = routines.folds[static_cast<size>(how)];
        const Plan plan(chunk, count);
        auto segment = [&](const size part) {
            const Range r = plan.range(part);
            Accumulator accumulator(how);
            sweep(routines, chunk, inputs.data(), static_cast<type *>(nullptr), r.first, r.last, steps,
                  [&](type *const *values, size n) {
                      accumulator.add(static_cast<double>(
                              contract ? routines.dot(values[product->sources[0]], values[product->sources[1]], n)
                                       : reduction(values[chunk.result], n)));
                  });
            return accumulator.result();
        };
        if (plan.parts == 1) return segment(0);
        std::vector<double> partials(plan.parts);
        Pool::global().run(plan.parts, [&](const size part) { partials[part] = segment(part); }, plan.seats);
        Accumulator accumulator(how);
        for (const double p : partials) accumulator.add(p);
        return accumulator.result();
    }

} // namespace

extern "C" {

// Sets SCI_THREADS before the reference's pool is first built (src/core/drivers/processor/pool.cpp:62-73).
int ref_threads(int threads) {
    static const bool once = [&] {
        setenv("SCI_THREADS", std::to_string(threads > 0 ? threads : 1).c_str(), 1);
        return true;
    }();
    (void)once;
    return static_cast<int>(Pool::global().width());
}

int ref_has_avx512(void) { return kernels::avx512 != nullptr; }

// Records a chunk from a postfix program with the reference's Chunk and writes out its public fields:
// header = {precision, length, arguments, literals, temporaries, result, valid, cost}, code = 5 bytes per
// instruction, inputs = the program input number of each chunk input, constants = its constants.
// Returns 0, or -1 on a malformed program.
int ref_record(const Step *steps, int count, int precision, int finish, int32_t *header, uint8_t *code,
               int32_t *inputs, double *constants) {
    Chunk chunk(precision ? Precision::wide : Precision::single);
    if (!record(steps, count, finish, chunk)) return -1;
    header[0] = static_cast<int32_t>(chunk.precision);
    header[1] = chunk.length;
    header[2] = chunk.arguments;
    header[3] = chunk.literals;
    header[4] = chunk.temporaries;
    header[5] = chunk.result;
    header[6] = chunk.valid() ? 1 : 0;
    header[7] = static_cast<int32_t>(chunk.cost());
    for (int i = 0; i < chunk.length; ++i) {
        code[5 * i] = static_cast<uint8_t>(chunk.instructions[i].code);
        code[5 * i + 1] = chunk.instructions[i].target;
        for (int k = 0; k < 3; ++k) code[5 * i + 2 + k] = chunk.instructions[i].sources[k];
    }
    for (int i = 0; i < chunk.arguments; ++i)
        inputs[i] = static_cast<int32_t>(static_cast<const char *>(chunk.inputs[i]) - addresses);
    for (int i = 0; i < chunk.literals; ++i) constants[i] = chunk.constants[i];
    return 0;
}

// Evaluates a postfix program over `count` elements of the program's input arrays (indexed by input number).
int ref_execute(const Step *steps, int count, int precision, void *const *arrays, void *output, size_t elements,
                int isa) {
    Chunk chunk(precision ? Precision::wide : Precision::single);
    if (!record(steps, count, 0, chunk) || !chunk.valid()) return -1;
    if (precision) run(table(isa)->wide, chunk, reinterpret_cast<double *const *>(arrays), static_cast<double *>(output), elements);
    else run(table(isa)->single, chunk, reinterpret_cast<float *const *>(arrays), static_cast<float *>(output), elements);
    return 0;
}

// Reduces a postfix program's result over `count` elements.
double ref_reduce(const Step *steps, int count, int precision, void *const *arrays, int how, size_t elements,
                  int isa) {
    Chunk chunk(precision ? Precision::wide : Precision::single);
    if (!record(steps, count, 1, chunk) || !chunk.valid()) return std::numeric_limits<double>::quiet_NaN();
    if (precision)
        return fold(table(isa)->wide, chunk, reinterpret_cast<double *const *>(arrays), static_cast<Fold>(how), elements);
    return fold(table(isa)->single, chunk, reinterpret_cast<float *const *>(arrays), static_cast<Fold>(how), elements);
}

// One code applied elementwise by the reference's kernel table (the element kernels of generic/compose.hpp:35-79).
int ref_kernel(int code, int precision, const void *a, const void *b, const void *c, void *out, size_t n, int isa) {
    if (precision) {
        const double *sources[3] = {static_cast<const double *>(a), static_cast<const double *>(b),
                                    static_cast<const double *>(c)};
        table(isa)->wide.kernels[code](sources, static_cast<double *>(out), n);
    } else {
        const float *sources[3] = {static_cast<const float *>(a), static_cast<const float *>(b),
                                   static_cast<const float *>(c)};
        table(isa)->single.kernels[code](sources, static_cast<float *>(out), n);
    }
    return 0;
}

} // extern "C"
