/*
 * scihip.h - the C ABI of the MI355X (gfx950) HIP backend for libsci.
 *
 * This is the drop-in boundary: every entry point replaces one member of the reference's
 * device boundary, `sci::core::drivers::Device` (include/core/drivers/device.hpp:27-89), or one
 * host-only routine of include/linear/ that the backend moves onto the GPU. Plain pointers and
 * sizes only: no C++ or torch types cross it.
 *
 * Conventions.
 *  - Every entry point returns an int status: SCIHIP_OK (0) on success, SCIHIP_DECLINED (1)
 *    when the GPU cannot run the request as given (an operand that is not device memory of this
 *    device, an unsupported precision or size); the caller then runs its own CPU path, as the
 *    reference's Device does after a wait (src/core/drivers/device.cpp:260-262, 313-315,
 *    349-351). Negative values are errors (SCIHIP_E*). scihip_error() gives the last message of
 *    the calling thread.
 *  - Arrays are row-major and contiguous. Device pointers come from scihip_allocate; interior
 *    pointers are accepted wherever a device pointer is (as the reference's Allocator::locate,
 *    src/core/drivers/vulkan/allocator.cpp:81-89).
 *  - scihip_execute, scihip_product and the copies are asynchronous and ordered on the device's
 *    one stream, like the reference's Vulkan stream (src/core/drivers/vulkan/stream.cpp:71-83).
 *    scihip_reduce and every call that returns a value to the host are synchronous.
 *  - Each device serialises its callers with a mutex, as the reference's `guard` does
 *    (src/core/drivers/device.cpp:127).
 */
#ifndef SCIHIP_H
#define SCIHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCIHIP_OK 0
#define SCIHIP_DECLINED 1
#define SCIHIP_EINVAL (-1)
#define SCIHIP_ENOMEM (-2)
#define SCIHIP_EDEVICE (-3)
#define SCIHIP_ECOMPILE (-4)

/* Chunk capacity: instructions, inputs and constants (include/core/bytecode/chunk.hpp:31). */
#define SCIHIP_CAPACITY 64

/* Value ids: a 2-bit tag and a 6-bit index (include/core/bytecode/instruction.hpp:11-25). */
#define SCIHIP_SLOT_INPUT 0x00u
#define SCIHIP_SLOT_LITERAL 0x40u
#define SCIHIP_SLOT_TEMPORARY 0x80u
#define SCIHIP_SLOT_OUTPUT 0xC0u
#define SCIHIP_SLOT_NONE 0xFFu
#define SCIHIP_SLOT_TYPES 0xC0u

/* Operation codes, in the order of core::Code (include/core/bytecode/code.hpp:18-61). */
enum scihip_code {
    SCIHIP_COPY, SCIHIP_NEGATE, SCIHIP_ABSOLUTE,
    SCIHIP_ROOT, SCIHIP_EXPONENTIAL, SCIHIP_LOGARITHM,
    SCIHIP_SIN, SCIHIP_COS, SCIHIP_TAN, SCIHIP_ASIN, SCIHIP_ACOS, SCIHIP_ATAN,
    SCIHIP_CSC, SCIHIP_SEC, SCIHIP_COT, SCIHIP_ACSC, SCIHIP_ASEC, SCIHIP_ACOT,
    SCIHIP_SINH, SCIHIP_COSH, SCIHIP_TANH, SCIHIP_ASINH, SCIHIP_ACOSH, SCIHIP_ATANH,
    SCIHIP_CSCH, SCIHIP_SECH, SCIHIP_COTH, SCIHIP_ACSCH, SCIHIP_ASECH, SCIHIP_ACOTH,
    SCIHIP_ADD, SCIHIP_SUBTRACT, SCIHIP_MULTIPLY, SCIHIP_DIVIDE, SCIHIP_MINIMUM, SCIHIP_MAXIMUM,
    SCIHIP_FUSE,
    SCIHIP_CODES
};

/* Folds (include/core/bytecode/fold.hpp:11). */
enum scihip_fold { SCIHIP_FOLD_SUM = 0, SCIHIP_FOLD_MINIMUM = 1, SCIHIP_FOLD_MAXIMUM = 2 };

/* Element precision (include/core/bytecode/precision.hpp:12). */
enum scihip_precision { SCIHIP_SINGLE = 0, SCIHIP_WIDE = 1 };

/* One instruction: target = code(sources...) (include/core/bytecode/instruction.hpp:28-35). */
typedef struct scihip_instruction {
    uint8_t code;
    uint8_t target;
    uint8_t sources[3];
} scihip_instruction;

/*
 * A recorded chunk: the public fields of core::Chunk (include/core/bytecode/chunk.hpp:178-196),
 * in the same order and with the same meaning. Only the first `length` instructions, `arguments`
 * inputs and `literals` constants are read. Inputs are device pointers (interior pointers
 * allowed). The chunk is borrowed for the call only.
 */
typedef struct scihip_chunk {
    uint8_t precision;
    uint8_t length;
    uint8_t arguments;
    uint8_t literals;
    uint8_t temporaries;
    uint8_t result;
    scihip_instruction instructions[SCIHIP_CAPACITY];
    const void *inputs[SCIHIP_CAPACITY];
    double constants[SCIHIP_CAPACITY];
} scihip_chunk;

typedef struct scihip_device scihip_device;

/* The last error message of the calling thread, or "" (never null). */
const char *scihip_error(void);

/* Opens GPU `ordinal`; replaces Device::Device(Target::graphics, index)
 * (include/core/drivers/device.hpp:32, src/core/drivers/device.cpp:186-199). Returns
 * SCIHIP_DECLINED when no such GPU exists, so the caller can become a CPU device. */
int scihip_open(uint32_t ordinal, scihip_device **device);

/* Waits for queued work and releases the device (src/core/drivers/device.cpp:201). */
int scihip_close(scihip_device *device);

/* The device's name (include/core/drivers/device.hpp:46). Valid until scihip_close. */
const char *scihip_name(const scihip_device *device);

/* Blocks until all queued work has finished (include/core/drivers/device.hpp:49). */
int scihip_synchronize(scihip_device *device);

/* Device memory of `bytes` bytes, 256-byte aligned (include/core/drivers/device.hpp:55,
 * src/core/drivers/device.cpp:212-224). Reuses a cached block when one fits. */
int scihip_allocate(scihip_device *device, size_t bytes, void **pointer);

/* Returns memory scihip_allocate gave; the block is reused or freed only after the work queued
 * so far has finished (include/core/drivers/device.hpp:58, src/core/drivers/device.cpp:226-233). */
int scihip_dispose(scihip_device *device, void *pointer);

/* 1 when `pointer` lies inside a live allocation of this device, else 0
 * (src/core/drivers/vulkan/allocator.cpp:81-89). */
int scihip_locate(scihip_device *device, const void *pointer);

/* Stream-ordered copies between host and device memory, or within the device. The reference
 * copies through the mapped pointer (include/array/array.hpp:223-237); MI355X HBM is not
 * host-mapped, so Array::to / host() go through these. Upload and download wait for the copy. */
int scihip_upload(scihip_device *device, void *destination, const void *source, size_t bytes);
int scihip_download(scihip_device *device, void *destination, const void *source, size_t bytes);
int scihip_copy(scihip_device *device, void *destination, const void *source, size_t bytes);

/* Evaluates a chunk over `count` elements into `output` (include/core/drivers/device.hpp:63,
 * src/core/drivers/device.cpp:235-263). Asynchronous. A chunk whose specialised kernel is ready
 * runs that kernel; otherwise the bytecode interpreter runs it and a specialised kernel is
 * compiled on a background thread (src/core/drivers/vulkan/compiler.cpp:169-222). */
int scihip_execute(scihip_device *device, const scihip_chunk *chunk, void *output, size_t count);

/* Sum, minimum or maximum of the chunk's result over `count` elements, in double
 * (include/core/drivers/device.hpp:69, src/core/drivers/device.cpp:265-316). Synchronous.
 * NaN anywhere gives NaN; the minimum and maximum of nothing are NaN and the sum of nothing 0.
 * A sum whose chunk ends in a multiply into its result is contracted into a dot product
 * (include/core/bytecode/chunk.hpp:151-154). */
int scihip_reduce(scihip_device *device, const scihip_chunk *chunk, uint32_t fold, size_t count, double *result);

/* scihip_reduce with the result left in device memory, for callers that cannot allocate: code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. No counterpart in
 * the reference. It runs the same kernels as scihip_reduce but calls no hipMalloc, hipFree or other
 * allocating routine, never compiles (a specialised kernel runs only when one is already loaded,
 * e.g. by scihip_prepare; otherwise the interpreter runs), never waits and copies nothing to the
 * host: `result` is device memory of this device (8 bytes) and receives the value asynchronously.
 * The per-workgroup partials go to a static scratch array of 4096 partials (64 KiB) on each GPU, held
 * until the process ends; a launch that would need more partials is refused with SCIHIP_EINVAL
 * (MI355X launches at most 16 per compute unit, 4096). The scratch keeps the last call's contents and
 * nothing in it needs to start at zero. Every call on a GPU shares that GPU's scratch: calls on one
 * device are ordered on its stream, but calls through two devices opened on one GPU race. */
int scihip_reduce_static(scihip_device *device, const scihip_chunk *chunk, uint32_t fold, size_t count,
                         double *result);

/* Product modes. */
enum scihip_product_mode {
    SCIHIP_PRODUCT_FLOAT32 = 0,  /* float operands and result, fp32 matrix cores */
    SCIHIP_PRODUCT_BFLOAT16 = 1, /* float operands rounded to bf16, fp32 accumulation */
    SCIHIP_PRODUCT_SPLIT = 2,    /* float operands as bf16 hi + lo parts, three bf16 products (bf16x3) */
    SCIHIP_PRODUCT_FLOAT64 = 3   /* double operands and result, fp64 matrix cores */
};

/* Row-major output = left * right with left rows x depth and right depth x columns
 * (include/core/drivers/device.hpp:79-80, src/core/drivers/device.cpp:318-352). The reference's
 * Precision::single is SCIHIP_PRODUCT_FLOAT32 and Precision::wide SCIHIP_PRODUCT_FLOAT64.
 * Asynchronous. depth = 0 writes zeros (src/core/drivers/processor/dispatch.cpp:270-273). */
int scihip_product(scihip_device *device, uint32_t mode, const void *left, const void *right, void *output,
                   size_t rows, size_t columns, size_t depth);

/* Factorisation modes: the element type of the matrix and the precision of its trailing updates. */
enum scihip_factor_mode {
    SCIHIP_FACTOR_FLOAT64 = 0,  /* double matrix, fp64 matrix-core updates */
    SCIHIP_FACTOR_FLOAT32 = 1,  /* float matrix, fp32 matrix-core updates */
    SCIHIP_FACTOR_SPLIT = 2,    /* float matrix, bf16x3 matrix-core updates */
    SCIHIP_FACTOR_BFLOAT16 = 3  /* float matrix, bf16 matrix-core updates */
};

/* LU with partial pivoting of the order x order row-major matrix, in place: unit lower L below the
 * diagonal, U on and above it; rows[i] receives the original index of the row now at position i.
 * Replaces linear::factor (include/linear/decompose.hpp:29-72): the same pivot rule (the first row
 * with the largest |a|), row swaps and zero-pivot skip; blocked, with the trailing updates on the
 * matrix cores. `rows` is host memory of `order` entries. Synchronous. */
int scihip_factor(scihip_device *device, uint32_t mode, void *matrix, size_t order, int64_t *rows);

/* Forward substitution through unit L in the order `rows`, then back substitution through U:
 * solution = U^-1 L^-1 (right permuted by rows), `width` right-hand sides, all row-major
 * (include/linear/solve.hpp:25-52). `factors` and `rows` as scihip_factor gives them (rows in host
 * memory); precision SCIHIP_SINGLE for float, SCIHIP_WIDE for double. Asynchronous. */
int scihip_substitute(scihip_device *device, uint32_t precision, const void *factors, const int64_t *rows,
                      const void *right, void *solution, size_t order, size_t width);

typedef struct scihip_solve_stats {
    int32_t iterations;     /* refinement steps taken */
    int32_t converged;      /* 1 when the backward error reached the tolerance */
    double backward_error;  /* max over columns of |b - Ax|_inf / (|A|_inf |x|_inf + |b|_inf) */
    double factor_ms;       /* wall time of the factorisation */
    double solve_ms;        /* wall time of the whole call */
} scihip_solve_stats;

/* Solves A X = B for double A (order x order) and B (order x width), all device memory: a copy of A
 * is factored in `mode` precision, and X is refined in double (residual r = B - A X on the fp64
 * matrix cores, correction solved with the factors) until the backward error is at most
 * `tolerance` or `max_iterations` refinements were made. Replaces linear::solve
 * (include/linear/solve.hpp:60-76), which factors in the matrix's own precision on the host.
 * Synchronous; `stats` may be null. */
int scihip_solve(scihip_device *device, uint32_t mode, const double *matrix, const double *right, double *solution,
                 size_t order, size_t width, double tolerance, int32_t max_iterations, scihip_solve_stats *stats);

/* Dispatch policy: 0 = automatic (interpreter until the specialised kernel is ready),
 * 1 = interpreter only, 2 = specialised kernel only (compiles and waits on first use).
 * Mirrors SCI_JIT=0 (src/core/drivers/vulkan/compiler.cpp:153-158); the SCIHIP_JIT environment
 * variable sets the initial policy ("0" = 1, "force" = 2). */
int scihip_set_policy(scihip_device *device, int policy);

/* Compiles the specialised kernel of a chunk now and waits for it (the background compile that
 * scihip_execute starts, done in the foreground). SCIHIP_ECOMPILE when compilation fails. */
int scihip_prepare(scihip_device *device, const scihip_chunk *chunk);

/* Counts of launches since open: interpreter and specialised kernels, and completed compiles. */
int scihip_statistics(scihip_device *device, uint64_t *interpreted, uint64_t *specialised, uint64_t *compiled);

/* The device's HIP stream (a hipStream_t), on which every call above queues its work. For callers
 * that order or time their own work against it (HIP events, graphs). No reference counterpart: the
 * reference's Vulkan stream is internal (src/core/drivers/vulkan/stream.cpp:25-156). */
int scihip_stream(scihip_device *device, void **stream);

/* Seeded synthetic data, generated in place on the device and identical bit for bit to the
 * oracle's CPU generator: element i = lower + (upper - lower) * u(seed, i), u = top 24 bits of a
 * SplitMix64 hash of (seed, i) times 2^-24 (single) or top 53 bits times 2^-53 (wide); then every
 * element whose hash under `nan_seed` falls below nan_fraction * 2^32 becomes NaN (nan_fraction 0
 * disables it). No reference counterpart: it feeds benchmarks and tests without host traffic. */
int scihip_fill_uniform(scihip_device *device, void *output, uint32_t precision, size_t count, uint64_t seed,
                        double lower, double upper, uint64_t nan_seed, double nan_fraction);

/*
 * Reverse-mode adjoint of a chunk, forward and backward fused in one pass over the data: for every
 * input i with gradients[i] non-null, gradients[i][e] = seed[e] * d result[e] / d input_i[e], where
 * result is the chunk's result value and seed is an upstream gradient array (null: all ones, which
 * makes the gradients those of the loss sum(result)). With `loss` non-null, the sum of the result is
 * also returned, contracted as scihip_reduce does (synchronous then; asynchronous otherwise). Forward
 * values are recomputed in registers, never stored. Derivatives follow the closed forms of each code;
 * minimum and maximum pass the adjoint to the operand they return, half to each on a tie, and
 * |x| passes sign(x) (0 at 0). Gradient arrays may alias inputs. No reference counterpart: the
 * reference has no differentiation; the chunk is include/core/bytecode/chunk.hpp:28-232's.
 */
int scihip_adjoint(scihip_device *device, const scihip_chunk *chunk, const void *seed, void *const *gradients,
                   size_t count, double *loss);

/* The adjoint program scihip_adjoint runs for a chunk: the chunk's forward instructions and the
 * backward ones in the interpreter's register numbering (inputs, the seed last when seeded; then
 * constants; then values), codes 37-39 being a * sign(b), minimum routing and maximum routing. */
typedef struct scihip_adjoint_program {
    uint32_t length;        /* instructions */
    uint32_t arguments;     /* inputs, the seed included */
    uint32_t literals;      /* constants */
    uint32_t registers;     /* registers the program uses */
    uint32_t gradients;     /* the chunk's inputs */
    uint32_t loss;          /* 1: the loss is folded */
    uint32_t contract;      /* 1: the loss is the sum of left * right */
    uint32_t left, right;   /* loss registers */
    uint32_t word[256];     /* code | target << 8 | first << 16 | second << 24 */
    uint8_t third[256];
    uint8_t slot[SCIHIP_CAPACITY]; /* register holding input i's gradient */
    double constants[SCIHIP_CAPACITY + 8];
} scihip_adjoint_program;

/* Fills `program` with the adjoint program of `chunk` (seeded: an upstream gradient is given;
 * loss: the loss is folded). Needs no device. No reference counterpart. */
int scihip_adjoint_describe(const scihip_chunk *chunk, int seeded, int loss, scihip_adjoint_program *program);

/* Adjoint of output = left * right (rows x depth times depth x columns, row-major, `mode` as for
 * scihip_product): left_gradient = seed * right^T (rows x depth) and right_gradient = left^T * seed
 * (depth x columns) for the upstream gradient `seed` (rows x columns), each as one matrix-core
 * product; a null gradient is skipped. Asynchronous. For the loss sum((left * right - C)^2),
 * seed = 2 (left * right - C), the adjoint of that chunk. No reference counterpart: the product is
 * include/linear/product.hpp:22-65's. */
int scihip_product_adjoint(scihip_device *device, uint32_t mode, const void *left, const void *right, const void *seed,
                           void *left_gradient, void *right_gradient, size_t rows, size_t columns, size_t depth);

/* output (columns x rows) = input^T (rows x columns), both row-major device memory of the element
 * type `precision`; input and output must not overlap. Replaces the host index remap of transpose
 * (include/operations/manipulations.hpp:36-62), which returns a host array. Asynchronous. */
int scihip_transpose(scihip_device *device, uint32_t precision, const void *input, void *output, size_t rows,
                     size_t columns);

/* Folds the chunk's result along one axis of a rows x columns row-major matrix (every input of the
 * chunk is such a matrix): axis 1 gives output[r] = fold of row r (rows values), axis 0 gives
 * output[c] = fold of column c (columns values), in the element type. Each fold means what
 * scihip_reduce means for that row or column alone, cast to the element type
 * (include/array/expression.hpp:198): NaN anywhere gives NaN, a fold of nothing is 0 for a sum and
 * NaN otherwise, a contracted sum folds exact products. No reference counterpart (the reference has
 * no axis reductions, README.md:129). Asynchronous. */
int scihip_reduce_axis(scihip_device *device, const scihip_chunk *chunk, uint32_t fold, size_t rows, size_t columns,
                       uint32_t axis, void *output);

/*
 * Graphs of repeated work (HIP graphs in place of the reference's recorded Vulkan command batches,
 * src/core/drivers/vulkan/stream.cpp:25-156). Between scihip_capture_begin and scihip_capture_end
 * the device's asynchronous calls (scihip_execute, scihip_product, scihip_adjoint without a loss,
 * scihip_product_adjoint, scihip_copy, scihip_fill_uniform) are recorded into a graph instead of
 * run; calls that wait for the device or return a value (synchronize, reduce, host copies, factor,
 * substitute, solve, replay) fail with SCIHIP_EINVAL meanwhile. scihip_replay queues the whole graph
 * on the device's stream with one launch; it reads and writes the same arrays every time, so change
 * their contents, not the arrays. Scratch memory the recorded calls used stays with the graph until
 * scihip_graph_release, which may be called while replays are still queued. No reference counterpart
 * as an API.
 */
typedef struct scihip_graph scihip_graph;
int scihip_capture_begin(scihip_device *device);
int scihip_capture_end(scihip_device *device, scihip_graph **graph);
int scihip_replay(scihip_device *device, const scihip_graph *graph);
int scihip_graph_release(scihip_device *device, scihip_graph *graph);

#ifdef __cplusplus
}
#endif

#endif
