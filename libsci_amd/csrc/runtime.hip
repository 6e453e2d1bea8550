// The C ABI (include/scihip.h): device, allocator, copies, chunk execute / reduce.
#include "runtime.hpp"
#include "axis.hpp"
#include "compiler.hpp"
#include "kernels.hpp"
#include "product.hpp"
#include "linear.hpp"
#include "lu.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>

// The reduction partials of scihip_reduce_static, the one entry point that allocates nothing: one
// Partial per workgroup of its launch, folded into the caller's result by a second kernel. It is
// the only device memory that entry point uses beyond the caller's inputs and result. Each GPU holds
// its own copy (64 KiB) until the process ends; host code reaches it only through the address that
// scihip_open takes with hipGetSymbolAddress (scihip_device::static_partials), never by this name.
// Every call on a GPU shares it, so such calls on two devices opened on the same GPU (two streams)
// race; calls on one device are ordered on its stream.
static __device__ __align__(256) scihip::Partial scratch[scihip::kStaticPartials];

namespace scihip {

    thread_local std::string last_error;

    void fail(const std::string &message) { last_error = message; }

    int status(hipError_t error, const char *what) {
        if (error == hipSuccess) return SCIHIP_OK;
        last_error = std::string(what) + ": " + hipGetErrorString(error);
        return error == hipErrorOutOfMemory ? SCIHIP_ENOMEM : SCIHIP_EDEVICE;
    }

    // ------------------------------------------------------------------------------------ Allocator

    Allocator::~Allocator() {
        (void)hipStreamSynchronize(stream);
        reclaim(true);
        trim(0);
        for (auto &block : live) (void)hipFree(reinterpret_cast<void *>(block.first));
    }

    void Allocator::reclaim(bool all) {
        size_t kept = 0;
        for (Retired &r : retired) {
            if (all || hipEventQuery(r.event) == hipSuccess) {
                (void)hipEventDestroy(r.event);
                cache.emplace(r.bytes, r.pointer);
                cached += r.bytes;
            } else {
                retired[kept++] = r;
            }
        }
        retired.resize(kept);
        if (cached > limit) trim(limit / 2);
    }

    void Allocator::trim(size_t keep) {
        while (cached > keep && !cache.empty()) {
            auto largest = std::prev(cache.end());
            (void)hipFree(largest->second);
            cached -= largest->first;
            cache.erase(largest);
        }
    }

    int Allocator::allocate(size_t bytes, void **pointer) {
        reclaim(false);
        const size_t rounded = std::max<size_t>((bytes + 4095) / 4096 * 4096, 4096);
        // A cached block at most 25% larger than asked for (allocator.cpp:44-54).
        auto found = cache.lower_bound(rounded);
        if (found != cache.end() && found->first <= rounded + rounded / 4) {
            *pointer = found->second;
            live[reinterpret_cast<uintptr_t>(*pointer)] = found->first;
            cached -= found->first;
            cache.erase(found);
            return SCIHIP_OK;
        }
        hipError_t error = hipMalloc(pointer, rounded);
        if (error == hipErrorOutOfMemory) {
            // Out of memory: drop the cache and retry once (allocator.cpp:57-65).
            (void)hipGetLastError();
            (void)hipStreamSynchronize(stream);
            reclaim(true);
            trim(0);
            error = hipMalloc(pointer, rounded);
        }
        if (error != hipSuccess) {
            *pointer = nullptr;
            return status(error, "hipMalloc");
        }
        live[reinterpret_cast<uintptr_t>(*pointer)] = rounded;
        return SCIHIP_OK;
    }

    int Allocator::dispose(void *pointer) {
        auto found = live.find(reinterpret_cast<uintptr_t>(pointer));
        if (found == live.end()) {
            fail("scihip_dispose: not an allocation of this device");
            return SCIHIP_EINVAL;
        }
        if (holding) {
            holding->push_back(pointer);
            return SCIHIP_OK;
        }
        Retired r{pointer, found->second, nullptr};
        live.erase(found);
        hipError_t error = hipEventCreateWithFlags(&r.event, hipEventDisableTiming);
        if (error == hipSuccess) error = hipEventRecord(r.event, stream);
        if (error != hipSuccess) {
            (void)hipStreamSynchronize(stream);
            if (r.event) (void)hipEventDestroy(r.event);
            cache.emplace(r.bytes, r.pointer);
            cached += r.bytes;
            return SCIHIP_OK;
        }
        retired.push_back(r);
        return SCIHIP_OK;
    }

    bool Allocator::locate(const void *pointer, size_t bytes) const {
        const uintptr_t at = reinterpret_cast<uintptr_t>(pointer);
        auto found = live.upper_bound(at);
        if (found == live.begin()) return false;
        --found;
        return at >= found->first && at + bytes <= found->first + found->second;
    }

    // ------------------------------------------------------------------------------------ encoding

    // The register number of a value id: inputs, then constants, then temporaries, then the output.
    struct Numbering {
        uint32_t arguments, literals, temporaries;
        uint32_t operator()(uint8_t id) const {
            const uint32_t tag = id & SCIHIP_SLOT_TYPES, index = id & 0x3Fu;
            if (id == SCIHIP_SLOT_NONE) return 0;
            if (tag == SCIHIP_SLOT_INPUT) return index;
            if (tag == SCIHIP_SLOT_LITERAL) return arguments + index;
            if (tag == SCIHIP_SLOT_TEMPORARY) return arguments + literals + index;
            return arguments + literals + temporaries;
        }
        uint32_t registers() const { return arguments + literals + temporaries + 1; }
    };

    bool contracted(const scihip_chunk &chunk, uint32_t fold) { // chunk.hpp:151-154
        return fold == SCIHIP_FOLD_SUM && chunk.length > 0 &&
               chunk.instructions[chunk.length - 1].code == SCIHIP_MULTIPLY &&
               chunk.instructions[chunk.length - 1].target == chunk.result;
    }

    // Checks a chunk's structure: every id it uses names a slot it has, codes are known.
    bool wellformed(const scihip_chunk &chunk) {
        if (chunk.length > SCIHIP_CAPACITY || chunk.arguments > SCIHIP_CAPACITY || chunk.literals > SCIHIP_CAPACITY ||
            chunk.temporaries > SCIHIP_CAPACITY)
            return false;
        auto valid = [&](uint8_t id, bool output_ok) {
            if (id == SCIHIP_SLOT_NONE) return true;
            const uint8_t tag = id & SCIHIP_SLOT_TYPES, index = id & 0x3F;
            if (tag == SCIHIP_SLOT_INPUT) return index < chunk.arguments;
            if (tag == SCIHIP_SLOT_LITERAL) return index < chunk.literals;
            if (tag == SCIHIP_SLOT_TEMPORARY) return index < chunk.temporaries;
            return output_ok && id == SCIHIP_SLOT_OUTPUT;
        };
        for (int s = 0; s < chunk.length; ++s) {
            const scihip_instruction &in = chunk.instructions[s];
            if (in.code >= SCIHIP_CODES) return false;
            if (!valid(in.target, true) || in.target == SCIHIP_SLOT_NONE) return false;
            if ((in.target & SCIHIP_SLOT_TYPES) == SCIHIP_SLOT_INPUT || (in.target & SCIHIP_SLOT_TYPES) == SCIHIP_SLOT_LITERAL)
                return false;
            const int arity = in.code == SCIHIP_FUSE ? 3 : in.code >= SCIHIP_ADD ? 2 : 1;
            for (int k = 0; k < 3; ++k) {
                if (!valid(in.sources[k], false)) return false;
                if (k < arity && in.sources[k] == SCIHIP_SLOT_NONE) return false;
            }
        }
        return valid(chunk.result, true) && chunk.result != SCIHIP_SLOT_NONE;
    }

    template <typename T>
    int encode(scihip_device *device, const scihip_chunk &chunk, T *output, size_t count, uint32_t fold, bool reduce,
               Launch<T> &launch, int &registers) {
        if (!wellformed(chunk)) {
            fail("malformed chunk");
            return SCIHIP_EINVAL;
        }
        Program<T> &p = launch.p;
        std::memset(&p, 0, sizeof(p));
        const Numbering number{chunk.arguments, chunk.literals, chunk.temporaries};
        const bool contract = reduce && contracted(chunk, fold);
        p.length = chunk.length;
        p.arguments = chunk.arguments;
        p.literals = chunk.literals;
        p.result = number(chunk.result);
        p.fold = fold;
        p.contract = contract ? 1u : 0u;
        if (contract) {
            p.left = number(chunk.instructions[chunk.length - 1].sources[0]);
            p.right = number(chunk.instructions[chunk.length - 1].sources[1]);
        }
        p.count = count;
        bool aligned = (reinterpret_cast<uintptr_t>(output) & 15u) == 0;
        for (int s = 0; s < chunk.length; ++s) {
            const scihip_instruction &in = chunk.instructions[s];
            p.word[s] = uint32_t(in.code) | number(in.target) << 8 | number(in.sources[0]) << 16 |
                        number(in.sources[1]) << 24;
            p.third[s] = static_cast<uint8_t>(number(in.sources[2]));
        }
        // Constants are narrowed to the element type, as the CPU does (dispatch.cpp:52-56).
        for (int i = 0; i < chunk.literals; ++i) p.constants[i] = static_cast<T>(chunk.constants[i]);
        for (int i = 0; i < chunk.arguments; ++i) {
            if (!device->allocator->locate(chunk.inputs[i], count * sizeof(T))) {
                fail("an input of the chunk is not device memory of this device");
                return SCIHIP_DECLINED;
            }
            p.inputs[i] = static_cast<const T *>(chunk.inputs[i]);
            aligned = aligned && (reinterpret_cast<uintptr_t>(chunk.inputs[i]) & 15u) == 0;
        }
        p.aligned = aligned ? 1u : 0u;
        p.output = output;
        launch.partials = static_cast<Partial *>(device->partials);
        const uint32_t needed = number.registers();
        if (needed > uint32_t(kLarge)) {
            fail("chunk needs more registers than the interpreter has");
            return SCIHIP_EINVAL;
        }
        // Programs of up to 16 registers take the 16-register file: its wave-uniform indexing compiles to
        // one indexed move per bank (s_set_gpr_idx), where the 8-register file's compiles to chains of
        // selects.
        registers = needed <= kMedium ? kMedium : kLarge;
        return SCIHIP_OK;
    }

    // Workgroups for `count` elements at `width` lanes and `unroll` vectors per thread per sweep,
    // each sweeping the array grid-stride. Stores: up to 64 per CU (16384 on MI355X; the C2 store
    // measured 1.5% faster than with 32 per CU). Reductions: 16 per CU (4096 on MI355X, four waves
    // per SIMD), one partial each for the fold kernel: with 4 per CU a wave's transcendentals did not
    // overlap its loads (C2 sum 5.45 -> 6.01 TB/s, adjoint forward 3.18 -> 3.63 TB/s; the dot, all
    // loads, is the same at either).
    unsigned blocks_for(scihip_device *device, size_t count, int width, int unroll, int mode) {
        const size_t steps = (count + width - 1) / width;
        const size_t wanted = (steps + 256 * size_t(unroll) - 1) / (256 * size_t(unroll));
        const size_t fill = size_t(device->units) * (mode == 0 ? 64 : 16);
        return static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>({wanted, fill, kMaxBlocks})));
    }

    std::string key_of(const scihip_chunk &chunk, int mode, bool contract, uint32_t fold = 0) {
        std::string key;
        key.reserve(9 + 5 * chunk.length);
        key += char(mode);
        key += char(contract);
        key += char(fold);
        key += char(chunk.precision);
        key += char(chunk.arguments);
        key += char(chunk.literals);
        key += char(chunk.temporaries);
        key += char(chunk.result);
        for (int s = 0; s < chunk.length; ++s) {
            const scihip_instruction &in = chunk.instructions[s];
            key += char(in.code);
            key += char(in.target);
            key += char(in.sources[0]);
            key += char(in.sources[1]);
            key += char(in.sources[2]);
        }
        return key;
    }

    // Where a reduction leaves its partials and its result, and whether the call may compile.
    struct Sink {
        double *result;      // where the fold kernel writes the folded result
        unsigned capacity;   // partials arguments.partials holds
        bool compile;        // false: only a kernel already loaded runs (loading a module allocates)
        Partial *single = nullptr; // when set, a one-workgroup reduction leaves its partial here and
                                   // runs no fold kernel (*folded tells the caller)
        bool *folded = nullptr;
    };

    // What the fold kernel makes of a single partial (fold_partials_block, then the NaN flag): the
    // fold's identity combined with the value, so a sum of -0 gives +0 as it does there.
    double fold_single(uint32_t fold, const Partial &partial) {
        if (partial.poisoned) return std::numeric_limits<double>::quiet_NaN();
        const double v = partial.value;
        if (fold == 0) return 0.0 + v;
        const double inf = std::numeric_limits<double>::infinity();
        return fold == 1 ? (v < inf ? v : inf) : (v > -inf ? v : -inf);
    }

    // Runs a chunk: its specialised kernel when ready (or forced), else the interpreter. A reduction
    // folds its partials into sink.result.
    template <typename T>
    int launch(scihip_device *device, const scihip_chunk &chunk, Launch<T> &arguments, int registers, int mode,
               bool contract, const Sink &sink) {
        const size_t count = arguments.p.count;
        const Kernel *kernel = nullptr;
        if (device->policy != 1) {
            const uint32_t fold = mode == 1 && !contract ? arguments.p.fold : 0;
            const std::string key = key_of(chunk, mode, contract, fold);
            if (sink.compile) {
                kernel = device->compiler->find(key, [&] { return Compiler::generate(chunk, mode, contract, fold); },
                                                device->policy == 2);
                if (!kernel && device->policy == 2) {
                    fail("specialised kernel failed to compile");
                    return SCIHIP_ECOMPILE;
                }
            } else {
                kernel = device->compiler->loaded(key);
            }
        }
        // A specialised store covers the arrays once with one vector per thread; the other kernels
        // sweep them grid-stride on a grid sized to the device.
        const bool one_pass = kernel && mode == 0;
        const uint64_t vectors = (count + Traits<T>::width - 1) / Traits<T>::width;
        const unsigned blocks =
                one_pass ? unsigned(std::min<uint64_t>((vectors + Compiler::kStoreThreads - 1) / Compiler::kStoreThreads,
                                                       Compiler::kStoreGrid))
                : kernel ? blocks_for(device, count, Traits<T>::width, Compiler::kUnroll, mode)
                         : blocks_for(device, count, registers == kLarge ? 1 : Traits<T>::width, 1, mode);
        if (mode == 1 && blocks > sink.capacity) {
            fail("the reduction needs " + std::to_string(blocks) + " partials; its buffer holds " +
                 std::to_string(sink.capacity));
            return SCIHIP_EINVAL;
        }
        int rc = SCIHIP_OK;
        const bool single = mode == 1 && blocks == 1 && sink.single;
        if (single) arguments.partials = sink.single;
        if (sink.folded) *sink.folded = !single;
        if (kernel) {
            void *parameters[] = {&arguments}; // copied at launch (and into a graph node under capture)
            ++device->specialised;
            const unsigned threads = one_pass ? unsigned(Compiler::kStoreThreads) : 256u;
            rc = status(hipModuleLaunchKernel(kernel->function, blocks, 1, 1, threads, 1, 1, 0, device->stream,
                                              parameters, nullptr),
                        "specialised launch");
        } else {
            ++device->interpreted;
            rc = status(interpret(arguments, registers, mode, blocks, device->stream), "interpreter launch");
        }
        if (rc || mode != 1 || single) return rc;
        return status(fold_partials(arguments.p.fold, arguments.partials, blocks, sink.result, device->stream),
                      "reduction fold");
    }

    template <typename T>
    int execute(scihip_device *device, const scihip_chunk &chunk, void *output, size_t count) {
        if (!device->allocator->locate(output, count * sizeof(T))) {
            fail("the output is not device memory of this device");
            return SCIHIP_DECLINED;
        }
        Launch<T> arguments;
        int registers = 0;
        if (const int rc = encode<T>(device, chunk, static_cast<T *>(output), count, 0, false, arguments, registers))
            return rc;
        if (chunk.result != SCIHIP_SLOT_OUTPUT) {
            fail("a stored chunk ends in its output");
            return SCIHIP_EINVAL;
        }
        return launch<T>(device, chunk, arguments, registers, 0, false, Sink{nullptr, 0, true});
    }

    template <typename T>
    int reduce(scihip_device *device, const scihip_chunk &chunk, uint32_t fold, size_t count, double *result) {
        Launch<T> arguments;
        int registers = 0;
        if (const int rc = encode<T>(device, chunk, nullptr, count, fold, true, arguments, registers)) return rc;
        const bool contract = contracted(chunk, fold);
        // The fold kernel writes the value straight into the pinned host word (no copy launch); a
        // reduction small enough for one workgroup leaves its partial in pinned host memory instead,
        // and the host folds it (no fold launch either).
        bool folded = true;
        Sink sink{device->staging, kMaxBlocks, true};
        sink.single = device->staging_partial;
        sink.folded = &folded;
        if (const int rc = launch<T>(device, chunk, arguments, registers, 1, contract, sink)) return rc;
        if (const int rc = status(hipStreamSynchronize(device->stream), "reduce")) return rc;
        *result = folded ? *device->staging : fold_single(arguments.p.fold, *device->staging_partial);
        return SCIHIP_OK;
    }

    // scihip_reduce's kernels with the partials in the static scratch and the result left in device
    // memory: no allocation, no compile, no wait.
    template <typename T>
    int reduce_static(scihip_device *device, const scihip_chunk &chunk, uint32_t fold, size_t count, double *result) {
        Launch<T> arguments;
        int registers = 0;
        if (const int rc = encode<T>(device, chunk, nullptr, count, fold, true, arguments, registers)) return rc;
        arguments.partials = device->static_partials;
        return launch<T>(device, chunk, arguments, registers, 1, contracted(chunk, fold),
                         Sink{result, kStaticPartials, false});
    }


    int multiply(scihip_device *device, int mode, const void *left, int64_t left_ld, bool left_transposed,
                 const void *right, int64_t right_ld, bool right_transposed, void *output, int64_t rows,
                 int64_t columns, int64_t depth) {
        const int64_t tile = k_tile(mode);
        const int64_t kp = (depth + tile - 1) / tile * tile;
        const size_t eb = element_bytes(mode);
        const bool split = mode == kSplit, wide = mode == kFloat64;
        void *a_hi = nullptr, *a_lo = nullptr, *b_hi = nullptr, *b_lo = nullptr;
        int rc = SCIHIP_OK;
        if (const int slab = slab_slices(mode, rows, columns, kp)) {
            // the large-tile kernel: operands packed in slabs, the shared dimension maybe sliced
            void *partials = nullptr;
            rc = device->allocator->allocate(slab_bytes(rows, kp), &a_hi);
            if (!rc) rc = device->allocator->allocate(slab_bytes(columns, kp), &b_hi);
            if (!rc && split) rc = device->allocator->allocate(slab_bytes(rows, kp), &a_lo);
            if (!rc && split) rc = device->allocator->allocate(slab_bytes(columns, kp), &b_lo);
            if (!rc && slab > 1) rc = device->allocator->allocate(size_t(slab) * rows * columns * 4, &partials);
            if (!rc)
                rc = status(pack_slabs(left, false, rows, depth, left_ld, left_transposed, a_hi, a_lo, kp,
                                       device->stream),
                            "pack");
            if (!rc)
                rc = status(pack_slabs(right, false, columns, depth, right_ld, right_transposed, b_hi, b_lo, kp,
                                       device->stream),
                            "pack");
            if (!rc) {
                ProductArgs args{a_hi, a_lo, b_hi, b_lo, output, rows, columns, kp, kp, kp, columns, 1.0, 0.0};
                args.slabs = 1;
                rc = status(product_split(mode, args, slab, partials, device->stream), "product");
            }
            for (void *block : {a_hi, a_lo, b_hi, b_lo, partials})
                if (block) device->allocator->dispose(block);
            return rc;
        }
        // Packed rows a multiple of 1 KiB apart put the same k of every row on the same few L2
        // channels; 128 bytes more per row spread them (the padding is never read).
        const int64_t ld = kp + ((kp * int64_t(eb)) % 1024 == 0 ? 128 / int64_t(eb) : 0);
        const size_t a_bytes = std::max<size_t>(size_t(rows * ld) * eb, 16);
        const size_t b_bytes = std::max<size_t>(size_t(columns * ld) * eb, 16);
        if (!rc) rc = device->allocator->allocate(a_bytes, &a_hi);
        if (!rc) rc = device->allocator->allocate(b_bytes, &b_hi);
        if (!rc && split) rc = device->allocator->allocate(a_bytes, &a_lo);
        if (!rc && split) rc = device->allocator->allocate(b_bytes, &b_lo);
        const int slices = split_slices(mode, rows, columns, kp);
        void *partials = nullptr;
        if (!rc && slices > 1) rc = device->allocator->allocate(size_t(slices) * rows * columns * (wide ? 8 : 4), &partials);
        if (!rc)
            rc = status(pack(mode, left, wide, rows, depth, left_ld, left_transposed, a_hi, a_lo, kp, device->stream,
                             ld),
                        "pack");
        if (!rc)
            rc = status(pack(mode, right, wide, columns, depth, right_ld, right_transposed, b_hi, b_lo, kp,
                             device->stream, ld),
                        "pack");
        if (!rc) {
            ProductArgs args{a_hi, a_lo, b_hi, b_lo, output, rows, columns, kp, ld, ld, columns, 1.0, 0.0};
            rc = status(product_split(mode, args, slices, partials, device->stream), "product");
        }
        for (void *block : {a_hi, a_lo, b_hi, b_lo, partials})
            if (block) device->allocator->dispose(block);
        return rc;
    }

    // Calls that wait for the device, return values to the host or stamp their launches cannot be
    // recorded into a graph.
    int refuse_while_capturing(scihip_device *device, const char *what) {
        if (!device->capturing) return SCIHIP_OK;
        fail(std::string(what) + " cannot be recorded into a graph (it waits for the device or returns a value)");
        return SCIHIP_EINVAL;
    }

    // ------------------------------------------------------------------------------------ LU support

    const char *spin_failure();

    // Sizes the LU workspace for `order` (element bytes `element`): the barrier counter and error
    // flag, the solve's ready flags (zeroed when first made), the panel's publications.
    int ensure_lu(scihip_device *device, int64_t order, size_t element) {
        Workspace &ws = device->lu;
        if (!ws.side) {
            if (int rc = status(hipStreamCreateWithFlags(&ws.side, hipStreamNonBlocking), "hipStreamCreate")) return rc;
            if (int rc = status(hipEventCreateWithFlags(&ws.columns_ready, hipEventDisableTiming), "hipEventCreate"))
                return rc;
            if (int rc = status(hipEventCreateWithFlags(&ws.panel_done, hipEventDisableTiming), "hipEventCreate"))
                return rc;
        }
        if (!ws.third) {
            if (int rc = status(hipStreamCreateWithFlags(&ws.third, hipStreamNonBlocking), "hipStreamCreate")) return rc;
            if (int rc = status(hipEventCreateWithFlags(&ws.inverted, hipEventDisableTiming), "hipEventCreate"))
                return rc;
            if (int rc = status(hipEventCreateWithFlags(&ws.upper_inverted, hipEventDisableTiming), "hipEventCreate"))
                return rc;
        }
        if (!ws.counter) {
            void *block = nullptr;
            if (int rc = status(hipMalloc(&block, 256), "hipMalloc")) return rc;
            if (int rc = status(hipMemset(block, 0, 256), "hipMemset")) return rc;
            ws.counter = static_cast<unsigned *>(block);
            ws.error = ws.counter + 32;
        }
        const size_t blocks = size_t((order + kSolveRows - 1) / kSolveRows);
        if (blocks > ws.flag_count) {
            if (ws.flags) {
                (void)hipStreamSynchronize(device->stream);
                (void)hipFree(ws.flags);
                ws.flags = nullptr;
            }
            const size_t count = std::max<size_t>(blocks, 1024);
            void *block = nullptr;
            if (int rc = status(hipMalloc(&block, count * sizeof(unsigned)), "hipMalloc")) return rc;
            if (int rc = status(hipMemset(block, 0, count * sizeof(unsigned)), "hipMemset")) return rc;
            ws.flags = static_cast<unsigned *>(block);
            ws.flag_count = count;
        }
        const size_t bytes = std::max(panel_publish_bytes(order, 8), panel_publish_bytes(order, element));
        if (bytes > ws.publish_bytes) {
            if (ws.publish) {
                (void)hipStreamSynchronize(device->stream);
                (void)hipStreamSynchronize(ws.side);
                (void)hipFree(ws.publish);
                ws.publish = nullptr;
                ws.publish_bytes = 0;
            }
            if (int rc = status(hipMalloc(&ws.publish, bytes), "hipMalloc")) return rc;
            if (int rc = status(hipMemset(ws.publish, 0, bytes), "hipMemset")) return rc;
            ws.publish_bytes = bytes;
        }
        return SCIHIP_OK;
    }

    void release_lu(scihip_device *device) {
        Workspace &ws = device->lu;
        if (ws.side) {
            (void)hipStreamSynchronize(ws.side);
            (void)hipStreamDestroy(ws.side);
        }
        if (ws.third) {
            (void)hipStreamSynchronize(ws.third);
            (void)hipStreamDestroy(ws.third);
        }
        if (ws.inverted) (void)hipEventDestroy(ws.inverted);
        if (ws.upper_inverted) (void)hipEventDestroy(ws.upper_inverted);
        if (ws.columns_ready) (void)hipEventDestroy(ws.columns_ready);
        if (ws.panel_done) (void)hipEventDestroy(ws.panel_done);
        if (ws.counter) (void)hipFree(ws.counter);
        if (ws.flags) (void)hipFree(ws.flags);
        if (ws.publish) (void)hipFree(ws.publish);
        ws = Workspace{};
    }

    // Grows the pinned host buffers of the LU paths to `doubles` doubles and `indices` 64-bit entries.
    int ensure_lu_host(scihip_device *device, size_t doubles, size_t indices) {
        if (doubles > device->lu_host_count) {
            if (device->lu_host) {
                (void)hipStreamSynchronize(device->stream);
                (void)hipHostFree(device->lu_host);
                device->lu_host = nullptr;
                device->lu_host_count = 0;
            }
            if (int rc = status(hipHostMalloc(reinterpret_cast<void **>(&device->lu_host), doubles * sizeof(double)),
                                "hipHostMalloc"))
                return rc;
            device->lu_host_count = doubles;
        }
        if (indices > device->lu_index_count) {
            if (device->lu_index) {
                (void)hipStreamSynchronize(device->stream);
                (void)hipHostFree(device->lu_index);
                device->lu_index = nullptr;
                device->lu_index_count = 0;
            }
            if (int rc = status(hipHostMalloc(reinterpret_cast<void **>(&device->lu_index), indices * sizeof(int64_t)),
                                "hipHostMalloc"))
                return rc;
            device->lu_index_count = indices;
        }
        return SCIHIP_OK;
    }

    // Checks the error flag of the bounded waits after the queued work has finished.
    int check_spins(scihip_device *device) {
        if (int rc = status(hipStreamSynchronize(device->stream), "LU")) return rc;
        unsigned flag = 0;
        if (int rc = status(hipMemcpy(&flag, device->lu.error, sizeof(flag), hipMemcpyDeviceToHost), "hipMemcpy"))
            return rc;
        if (flag) {
            (void)hipMemset(device->lu.error, 0, sizeof(unsigned));
            fail(spin_failure());
            return SCIHIP_EDEVICE;
        }
        return SCIHIP_OK;
    }

    // A device block from the allocator, returned when the holder goes out of scope (deferred).
    struct Scratch {
        scihip_device *device;
        void *pointer = nullptr;
        Scratch(scihip_device *d, size_t bytes, int &rc) : device(d) {
            if (!rc) rc = d->allocator->allocate(std::max<size_t>(bytes, 16), &pointer);
        }
        ~Scratch() {
            if (pointer) device->allocator->dispose(pointer);
        }
        template <typename T> T *as() const { return static_cast<T *>(pointer); }
    };

    template <typename T>
    int reduce_axis(scihip_device *device, const scihip_chunk &chunk, uint32_t fold, size_t rows, size_t columns,
                    uint32_t axis, T *output) {
        const size_t outputs = axis == 1 ? rows : columns;
        if (!device->allocator->locate(output, outputs * sizeof(T))) {
            fail("the output of the axis reduction is not device memory of this device");
            return SCIHIP_DECLINED;
        }
        if (rows == 0 || columns == 0) { // folds of nothing: 0 for a sum, NaN otherwise (device.cpp:266-268)
            const T empty = fold == SCIHIP_FOLD_SUM ? T(0) : std::numeric_limits<T>::quiet_NaN();
            return status(fill_value<T>(output, outputs, empty, device->stream), "fill");
        }
        Launch<T> launch;
        int registers = 0;
        if (const int rc = encode<T>(device, chunk, nullptr, rows * columns, fold, true, launch, registers)) return rc;
        AxisLaunch<T> a;
        a.p = launch.p;
        a.p.aligned = a.p.aligned && columns % size_t(Traits<T>::width) == 0 ? 1u : 0u;
        a.output = output;
        a.rows = rows;
        a.columns = columns;
        a.partials = nullptr;
        // The chunk's specialised axis kernel when ready (or forced), else the interpreter.
        const bool contract = contracted(chunk, fold);
        const Kernel *kernel = nullptr;
        if (device->policy != 1) {
            const uint32_t kernel_fold = contract ? 0 : launch.p.fold;
            kernel = device->compiler->find(key_of(chunk, 3 + int(axis), contract, kernel_fold), [&] {
                return Compiler::generate_axis(chunk, int(axis), contract, kernel_fold);
            },
                                            device->policy == 2);
            if (!kernel && device->policy == 2) {
                fail("specialised kernel failed to compile");
                return SCIHIP_ECOMPILE;
            }
        }
        uint32_t slices = 1;
        int rc = SCIHIP_OK;
        std::unique_ptr<Scratch> partials;
        // Specialised column folds run one-wave workgroups, four times as many as the device holds at
        // once, so that workgroups finishing early on one XCD leave no compute units idle (measured
        // 4.7 -> 5.1 TB/s at 16384^2 against four-wave workgroups that all start together).
        const unsigned threads = kernel && axis == 0 ? 64 : 256;
        if (axis == 0) {
            const int w = (registers == kLarge && !kernel) ? 1 : Traits<T>::width;
            const size_t tiles = (columns + threads * size_t(w) - 1) / (threads * size_t(w));
            const size_t wanted = (size_t(device->units) * 8 * (256 / threads) + tiles - 1) / tiles;
            slices = uint32_t(std::max<size_t>(1, std::min<size_t>({wanted, rows, 65535})));
            partials = std::make_unique<Scratch>(device, size_t(slices) * columns * sizeof(Partial), rc);
            if (rc) return rc;
            a.partials = partials->as<Partial>();
        }
        if (kernel) {
            ++device->specialised;
            const unsigned grid_x = axis == 1 ? unsigned(rows)
                                              : unsigned((columns + threads * size_t(Traits<T>::width) - 1) /
                                                         (threads * size_t(Traits<T>::width)));
            void *parameters[] = {&a};
            if ((rc = status(hipModuleLaunchKernel(kernel->function, grid_x, axis == 1 ? 1 : slices, 1, threads, 1, 1, 0,
                                                   device->stream, parameters, nullptr),
                             "specialised axis launch")))
                return rc;
            return axis == 0 ? status(finish_axis<T>(a, slices, device->stream), "axis reduction") : SCIHIP_OK;
        }
        ++device->interpreted;
        return status(fold_axis<T>(a, registers, int(axis), slices, device->stream), "axis reduction");
    }

    // rows[i] = the original index of the row at position i after the swaps (decompose.hpp:31, 41).
    std::vector<int64_t> rows_of(const std::vector<int64_t> &pivots) {
        std::vector<int64_t> rows(pivots.size());
        for (size_t i = 0; i < rows.size(); ++i) rows[i] = int64_t(i);
        for (size_t k = 0; k < pivots.size(); ++k) std::swap(rows[k], rows[size_t(pivots[k])]);
        return rows;
    }

    template <typename T>
    int factor_on_device(scihip_device *device, T *matrix, int64_t order, int mode, std::vector<int64_t> &pivots,
                         T *inverses = nullptr) {
        int rc = ensure_lu(device, order, sizeof(T));
        Scratch pivot_block(device, size_t(order) * sizeof(int64_t), rc);
        Scratch work(device, factor_scratch(order, mode), rc);
        if (rc) return rc;
        const char *message = nullptr;
        const int e = factor<T>(matrix, order, pivot_block.as<int64_t>(), mode, device->lu, work.pointer,
                                factor_scratch(order, mode), device->stream, &message, inverses);
        if (e) {
            fail(std::string("LU factorisation: ") + (message ? message : hipGetErrorString(hipError_t(e))));
            return SCIHIP_EDEVICE;
        }
        // The pivots and the bounded waits' error flag into pinned host memory, then one wait.
        if ((rc = ensure_lu_host(device, 2, size_t(order)))) return rc;
        unsigned *flag = reinterpret_cast<unsigned *>(device->lu_host);
        *flag = 0;
        if ((rc = status(hipMemcpyAsync(device->lu_index, pivot_block.pointer, size_t(order) * sizeof(int64_t),
                                        hipMemcpyDeviceToHost, device->stream), "hipMemcpy")))
            return rc;
        if ((rc = status(hipMemcpyAsync(flag, device->lu.error, sizeof(unsigned), hipMemcpyDeviceToHost, device->stream),
                         "hipMemcpy")))
            return rc;
        if ((rc = status(hipStreamSynchronize(device->stream), "LU"))) return rc;
        if (*flag) {
            (void)hipMemset(device->lu.error, 0, sizeof(unsigned));
            fail(spin_failure());
            return SCIHIP_EDEVICE;
        }
        pivots.assign(device->lu_index, device->lu_index + order);
        return SCIHIP_OK;
    }

    // x = U^-1 L^-1 P source (source in S, solved in T), rows on the device.
    template <typename T, typename S>
    int solve_on_device(scihip_device *device, const T *factors, const int64_t *rows, const S *source, T *x,
                        int64_t order, int64_t width, const T *inverses = nullptr, void *packed = nullptr) {
        if (int rc = status(permute<T, S>(source, x, rows, order, width, device->stream), "permute")) return rc;
        const char *message = nullptr;
        const int e = substitute<T>(factors, order, x, width, device->lu, device->epoch, device->stream, &message,
                                    inverses, packed);
        if (e) {
            fail(std::string("substitution: ") + (message ? message : hipGetErrorString(hipError_t(e))));
            return SCIHIP_EDEVICE;
        }
        return SCIHIP_OK;
    }

} // namespace scihip

using namespace scihip;

extern "C" {

const char *scihip_error(void) { return last_error.c_str(); }

int scihip_open(uint32_t ordinal, scihip_device **out) {
    if (!out) return SCIHIP_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || int(ordinal) >= count) {
        (void)hipGetLastError();
        fail("no usable HIP device at this index");
        return SCIHIP_DECLINED;
    }
    auto device = std::make_unique<scihip_device>();
    device->ordinal = int(ordinal);
    if (int rc = status(hipSetDevice(device->ordinal), "hipSetDevice")) return rc;
    hipDeviceProp_t properties;
    if (int rc = status(hipGetDeviceProperties(&properties, device->ordinal), "hipGetDeviceProperties")) return rc;
    device->name = properties.name[0] ? properties.name : properties.gcnArchName;
    device->units = properties.multiProcessorCount;
    if (int rc = status(hipStreamCreateWithFlags(&device->stream, hipStreamNonBlocking), "hipStreamCreate")) return rc;
    device->allocator = std::make_unique<Allocator>(device->stream);
    if (int rc = status(hipMalloc(&device->partials, kMaxBlocks * sizeof(Partial)), "hipMalloc")) return rc;
    if (int rc = status(hipMalloc(reinterpret_cast<void **>(&device->result), 256), "hipMalloc")) return rc;
    if (int rc = status(hipHostMalloc(reinterpret_cast<void **>(&device->staging), sizeof(double)), "hipHostMalloc"))
        return rc;
    if (int rc = status(hipHostMalloc(reinterpret_cast<void **>(&device->staging_partial), sizeof(Partial)),
                        "hipHostMalloc"))
        return rc;
    // This GPU's copy of the static scratch (hipSetDevice above made it current).
    void *symbol = nullptr;
    if (int rc = status(hipGetSymbolAddress(&symbol, HIP_SYMBOL(scratch)), "hipGetSymbolAddress")) return rc;
    device->static_partials = static_cast<Partial *>(symbol);
    if (const char *text = std::getenv("SCIHIP_JIT")) {
        if (std::strcmp(text, "0") == 0) device->policy = 1;
        if (std::strcmp(text, "force") == 0) device->policy = 2;
    }
    device->compiler = std::make_unique<Compiler>(device->ordinal);
    *out = device.release();
    return SCIHIP_OK;
}

int scihip_close(scihip_device *device) {
    if (!device) return SCIHIP_OK;
    {
        std::lock_guard<std::mutex> lock(device->guard);
        if (device->capturing) { // an unfinished capture is abandoned
            hipGraph_t abandoned = nullptr;
            (void)hipStreamEndCapture(device->stream, &abandoned);
            if (abandoned) (void)hipGraphDestroy(abandoned);
            device->capturing = false;
            device->allocator->hold(nullptr);
            for (void *block : device->held) device->allocator->dispose(block);
        }
        (void)hipStreamSynchronize(device->stream);
        device->compiler.reset();
        release_lu(device);
        device->allocator.reset();
        (void)hipFree(device->partials);
        (void)hipFree(device->result);
        (void)hipHostFree(device->staging);
        (void)hipHostFree(device->staging_partial);
        if (device->lu_host) (void)hipHostFree(device->lu_host);
        if (device->lu_index) (void)hipHostFree(device->lu_index);
        (void)hipStreamDestroy(device->stream);
    }
    delete device;
    return SCIHIP_OK;
}

const char *scihip_name(const scihip_device *device) { return device ? device->name.c_str() : "cpu"; }

int scihip_synchronize(scihip_device *device) {
    if (!device) return SCIHIP_EINVAL;
    std::lock_guard<std::mutex> lock(device->guard);
    if (int rc = refuse_while_capturing(device, "scihip_synchronize")) return rc;
    const int rc = status(hipStreamSynchronize(device->stream), "hipStreamSynchronize");
    device->allocator->reclaim(true);
    return rc;
}

int scihip_allocate(scihip_device *device, size_t bytes, void **pointer) {
    if (!device || !pointer) return SCIHIP_EINVAL;
    std::lock_guard<std::mutex> lock(device->guard);
    return device->allocator->allocate(bytes, pointer);
}

int scihip_dispose(scihip_device *device, void *pointer) {
    if (!device) return SCIHIP_EINVAL;
    if (!pointer) return SCIHIP_OK;
    std::lock_guard<std::mutex> lock(device->guard);
    return device->allocator->dispose(pointer);
}

int scihip_locate(scihip_device *device, const void *pointer) {
    if (!device) return 0;
    std::lock_guard<std::mutex> lock(device->guard);
    return device->allocator->locate(pointer) ? 1 : 0;
}

static int copy(scihip_device *device, void *destination, const void *source, size_t bytes, hipMemcpyKind kind,
                bool wait) {
    if (!device) return SCIHIP_EINVAL;
    if (bytes == 0) return SCIHIP_OK;
    std::lock_guard<std::mutex> lock(device->guard);
    const bool in = kind == hipMemcpyHostToDevice || kind == hipMemcpyDeviceToDevice;
    const bool out = kind == hipMemcpyDeviceToHost || kind == hipMemcpyDeviceToDevice;
    if (wait) {
        if (int rc = refuse_while_capturing(device, "a copy to or from the host")) return rc;
    }
    if ((in && !device->allocator->locate(destination, bytes)) || (out && !device->allocator->locate(source, bytes))) {
        fail("a copy names memory that is not a device allocation of this size");
        return SCIHIP_EINVAL;
    }
    if (int rc = status(hipMemcpyAsync(destination, source, bytes, kind, device->stream), "hipMemcpyAsync")) return rc;
    return wait ? status(hipStreamSynchronize(device->stream), "copy") : SCIHIP_OK;
}

int scihip_upload(scihip_device *device, void *destination, const void *source, size_t bytes) {
    return copy(device, destination, source, bytes, hipMemcpyHostToDevice, true);
}

int scihip_download(scihip_device *device, void *destination, const void *source, size_t bytes) {
    return copy(device, destination, source, bytes, hipMemcpyDeviceToHost, true);
}

int scihip_copy(scihip_device *device, void *destination, const void *source, size_t bytes) {
    return copy(device, destination, source, bytes, hipMemcpyDeviceToDevice, false);
}

int scihip_execute(scihip_device *device, const scihip_chunk *chunk, void *output, size_t count) {
    if (!device || !chunk || chunk->precision > SCIHIP_WIDE) return SCIHIP_EINVAL;
    if (count == 0) return SCIHIP_OK;
    std::lock_guard<std::mutex> lock(device->guard);
    return chunk->precision == SCIHIP_SINGLE ? execute<float>(device, *chunk, output, count)
                                             : execute<double>(device, *chunk, output, count);
}

int scihip_reduce(scihip_device *device, const scihip_chunk *chunk, uint32_t fold, size_t count, double *result) {
    if (!device || !chunk || !result || fold > SCIHIP_FOLD_MAXIMUM || chunk->precision > SCIHIP_WIDE)
        return SCIHIP_EINVAL;
    if (count == 0) { // device.cpp:266-268
        *result = fold == SCIHIP_FOLD_SUM ? 0.0 : std::numeric_limits<double>::quiet_NaN();
        return SCIHIP_OK;
    }
    std::lock_guard<std::mutex> lock(device->guard);
    if (int rc = refuse_while_capturing(device, "scihip_reduce")) return rc;
    return chunk->precision == SCIHIP_SINGLE ? reduce<float>(device, *chunk, fold, count, result)
                                             : reduce<double>(device, *chunk, fold, count, result);
}

int scihip_reduce_static(scihip_device *device, const scihip_chunk *chunk, uint32_t fold, size_t count,
                         double *result) {
    if (!device || !chunk || !result || fold > SCIHIP_FOLD_MAXIMUM || chunk->precision > SCIHIP_WIDE)
        return SCIHIP_EINVAL;
    std::lock_guard<std::mutex> lock(device->guard);
    if (!device->allocator->locate(result, sizeof(double))) {
        fail("the result of scihip_reduce_static is not device memory of this device");
        return SCIHIP_DECLINED;
    }
    if (!device->static_partials) {
        fail("the static reduction scratch has no device address");
        return SCIHIP_EDEVICE;
    }
    if (count == 0) { // device.cpp:266-268, written by a fill kernel (no memset: callers may capture this)
        const double empty = fold == SCIHIP_FOLD_SUM ? 0.0 : std::numeric_limits<double>::quiet_NaN();
        return status(fill_value<double>(result, 1, empty, device->stream), "fill");
    }
    return chunk->precision == SCIHIP_SINGLE ? reduce_static<float>(device, *chunk, fold, count, result)
                                             : reduce_static<double>(device, *chunk, fold, count, result);
}

int scihip_set_policy(scihip_device *device, int policy) {
    if (!device || policy < 0 || policy > 2) return SCIHIP_EINVAL;
    std::lock_guard<std::mutex> lock(device->guard);
    device->policy = policy;
    return SCIHIP_OK;
}

int scihip_prepare(scihip_device *device, const scihip_chunk *chunk) {
    if (!device || !chunk) return SCIHIP_EINVAL;
    if (!wellformed(*chunk)) {
        fail("malformed chunk");
        return SCIHIP_EINVAL;
    }
    std::lock_guard<std::mutex> lock(device->guard);
    const bool stored = chunk->result == SCIHIP_SLOT_OUTPUT;
    const int mode = stored ? 0 : 1;
    // A stored chunk has one kernel. A reduced one has a kernel per fold (the fold is a constant of
    // each), plus the dot-product kernel of a contracted sum.
    const bool contract = !stored && contracted(*chunk, SCIHIP_FOLD_SUM);
    for (uint32_t fold = 0; fold <= (stored ? 0u : uint32_t(SCIHIP_FOLD_MAXIMUM)); ++fold) {
        if (!device->compiler->find(key_of(*chunk, mode, false, fold),
                                    [&] { return Compiler::generate(*chunk, mode, false, fold); }, true)) {
            fail("specialised kernel failed to compile");
            return SCIHIP_ECOMPILE;
        }
    }
    if (contract &&
        !device->compiler->find(key_of(*chunk, mode, true), [&] { return Compiler::generate(*chunk, mode, true); },
                                true)) {
        fail("specialised kernel failed to compile");
        return SCIHIP_ECOMPILE;
    }
    return SCIHIP_OK;
}

int scihip_statistics(scihip_device *device, uint64_t *interpreted, uint64_t *specialised, uint64_t *compiled) {
    if (!device) return SCIHIP_EINVAL;
    std::lock_guard<std::mutex> lock(device->guard);
    if (interpreted) *interpreted = device->interpreted;
    if (specialised) *specialised = device->specialised;
    if (compiled) *compiled = device->compiler->compiled();
    return SCIHIP_OK;
}

int scihip_product(scihip_device *device, uint32_t mode, const void *left, const void *right, void *output,
                   size_t rows, size_t columns, size_t depth) {
    if (!device || mode > SCIHIP_PRODUCT_FLOAT64) return SCIHIP_EINVAL;
    if (rows == 0 || columns == 0) return SCIHIP_OK; // device.cpp:320
    std::lock_guard<std::mutex> lock(device->guard);
    const bool wide = mode == SCIHIP_PRODUCT_FLOAT64;
    const size_t element = wide ? 8 : 4;
    if (!device->allocator->locate(left, rows * depth * element) ||
        !device->allocator->locate(right, depth * columns * element) ||
        !device->allocator->locate(output, rows * columns * element)) {
        fail("an operand of the product is not device memory of this device");
        return SCIHIP_DECLINED;
    }
    return multiply(device, int(mode), left, int64_t(depth), false, right, int64_t(columns), true, output,
                    int64_t(rows), int64_t(columns), int64_t(depth));
}


int scihip_factor(scihip_device *device, uint32_t mode, void *matrix, size_t order, int64_t *rows) {
    if (!device || !rows || mode > SCIHIP_FACTOR_BFLOAT16) return SCIHIP_EINVAL;
    if (order == 0) return SCIHIP_OK;
    std::lock_guard<std::mutex> lock(device->guard);
    if (int rc = refuse_while_capturing(device, "scihip_factor")) return rc;
    const bool wide = mode == SCIHIP_FACTOR_FLOAT64;
    if (!device->allocator->locate(matrix, order * order * (wide ? 8 : 4))) {
        fail("the matrix is not device memory of this device");
        return SCIHIP_DECLINED;
    }
    std::vector<int64_t> pivots;
    const int product_mode = wide ? kFloat64 : mode == SCIHIP_FACTOR_FLOAT32 ? kFloat32
                                              : mode == SCIHIP_FACTOR_SPLIT   ? kSplit
                                                                              : kBfloat16;
    const int rc = wide ? factor_on_device<double>(device, static_cast<double *>(matrix), int64_t(order), product_mode, pivots)
                        : factor_on_device<float>(device, static_cast<float *>(matrix), int64_t(order), product_mode, pivots);
    if (rc) return rc;
    const std::vector<int64_t> order_of_rows = rows_of(pivots);
    std::copy(order_of_rows.begin(), order_of_rows.end(), rows);
    return SCIHIP_OK;
}

int scihip_substitute(scihip_device *device, uint32_t precision, const void *factors, const int64_t *rows,
                      const void *right, void *solution, size_t order, size_t width) {
    if (!device || !rows || precision > SCIHIP_WIDE) return SCIHIP_EINVAL;
    if (order == 0 || width == 0) return SCIHIP_OK;
    std::lock_guard<std::mutex> lock(device->guard);
    if (int rc = refuse_while_capturing(device, "scihip_substitute")) return rc;
    const size_t eb = precision == SCIHIP_WIDE ? 8 : 4;
    if (!device->allocator->locate(factors, order * order * eb) || !device->allocator->locate(right, order * width * eb) ||
        !device->allocator->locate(solution, order * width * eb)) {
        fail("an operand of the substitution is not device memory of this device");
        return SCIHIP_DECLINED;
    }
    if (right == solution) {
        fail("the solution must not alias the right-hand side");
        return SCIHIP_EINVAL;
    }
    int rc = ensure_lu(device, int64_t(order), eb);
    Scratch device_rows(device, order * sizeof(int64_t), rc);
    if (rc) return rc;
    if ((rc = status(hipMemcpyAsync(device_rows.pointer, rows, order * sizeof(int64_t), hipMemcpyHostToDevice,
                                    device->stream),
                     "hipMemcpyAsync")))
        return rc;
    // The blocked substitution needs the inverses of the diagonal blocks of these factors.
    const bool blocked = blocked_substitution(int64_t(order));
    Scratch inverses(device, blocked ? 6 * order * kBlock * eb : 16, rc);
    Scratch packed(device, blocked ? substitute_scratch(order) : 16, rc);
    if (rc) return rc;
    if (precision == SCIHIP_WIDE) {
        const double *f = static_cast<const double *>(factors);
        if (blocked && (rc = status(invert_blocks<double>(f, int64_t(order), inverses.as<double>(), device->stream),
                                    "invert blocks")))
            return rc;
        rc = solve_on_device<double, double>(device, f, device_rows.as<int64_t>(), static_cast<const double *>(right),
                                             static_cast<double *>(solution), int64_t(order), int64_t(width),
                                             blocked ? inverses.as<double>() : nullptr,
                                             blocked ? packed.pointer : nullptr);
    } else {
        const float *f = static_cast<const float *>(factors);
        if (blocked && (rc = status(invert_blocks<float>(f, int64_t(order), inverses.as<float>(), device->stream),
                                    "invert blocks")))
            return rc;
        rc = solve_on_device<float, float>(device, f, device_rows.as<int64_t>(), static_cast<const float *>(right),
                                           static_cast<float *>(solution), int64_t(order), int64_t(width),
                                           blocked ? inverses.as<float>() : nullptr,
                                           blocked ? packed.pointer : nullptr);
    }
    if (rc) return rc;
    // The host rows buffer was copied asynchronously from pageable memory: finish before returning it.
    return check_spins(device);
}

int scihip_solve(scihip_device *device, uint32_t mode, const double *matrix, const double *right, double *solution,
                 size_t order, size_t width, double tolerance, int32_t max_iterations, scihip_solve_stats *stats) {
    if (!device || mode > SCIHIP_FACTOR_BFLOAT16) return SCIHIP_EINVAL;
    scihip_solve_stats local{0, 0, 0.0, 0.0, 0.0};
    if (!stats) stats = &local;
    *stats = local;
    if (order == 0 || width == 0) {
        stats->converged = 1;
        return SCIHIP_OK;
    }
    std::lock_guard<std::mutex> lock(device->guard);
    if (int rc = refuse_while_capturing(device, "scihip_solve")) return rc;
    const int64_t n = int64_t(order), w = int64_t(width);
    if (!device->allocator->locate(matrix, order * order * 8) || !device->allocator->locate(right, order * width * 8) ||
        !device->allocator->locate(solution, order * width * 8)) {
        fail("an operand of the solve is not device memory of this device");
        return SCIHIP_DECLINED;
    }
    const auto start = std::chrono::steady_clock::now();
    // SCIHIP_TRACE=lu: the wall time of every stage of the solve on stderr (the stream is waited at
    // each stage boundary, so the stages do not overlap while tracing).
    const char *trace_text = std::getenv("SCIHIP_TRACE");
    const bool tracing = trace_text && std::strstr(trace_text, "lu");
    auto last_mark = start;
    std::string trace_line;
    auto mark = [&](const char *stage) {
        if (!tracing) return;
        (void)hipStreamSynchronize(device->stream);
        const auto now = std::chrono::steady_clock::now();
        char text[64];
        std::snprintf(text, sizeof(text), " %s %.1f", stage, std::chrono::duration<double, std::milli>(now - last_mark).count());
        trace_line += text;
        last_mark = now;
    };
    const bool wide = mode == SCIHIP_FACTOR_FLOAT64;
    const size_t eb = wide ? 8 : 4;
    int rc = SCIHIP_OK;
    Scratch factors(device, order * order * eb, rc);
    Scratch device_rows(device, order * sizeof(int64_t), rc);
    Scratch correction(device, order * width * eb, rc);
    Scratch residual(device, order * width * 8, rc);
    const int64_t kp = (n + 15) / 16 * 16;
    Scratch packed_x(device, size_t(w) * size_t(kp) * 8, rc);
    Scratch sums(device, order * 8, rc);
    const int64_t chunks = column_max_chunks(n);
    // column maxima of the residual, the solution and the right-hand sides, then of |A|'s row sums
    Scratch partial(device, size_t(chunks * w * 3 + chunks) * 8, rc);
    const bool aligned = n % 16 == 0;
    Scratch packed_a(device, aligned ? 16 : size_t(n) * size_t(kp) * 8, rc);
    const bool blocked = blocked_substitution(n);
    Scratch inverses(device, blocked ? 6 * order * kBlock * eb : 16, rc);
    Scratch packed(device, blocked ? substitute_scratch(order) : 16, rc);
    // The residual product has few output tiles (n / 128 x 1) and a long shared dimension.
    const int slices = split_slices(kFloat64, n, w, kp);
    // The residual: the narrow kernel that reads A in place when the shape allows, else the general
    // product over packed operands, its shared dimension split.
    const bool narrow = narrow_residual_fits(n, w);
    const size_t slice_bytes = narrow ? size_t(kResidualSlices) * order * width * 8
                                      : (slices > 1 ? size_t(slices) * order * width * 8 : 16);
    Scratch residual_slices(device, slice_bytes, rc);
    if (rc) return rc;
    if ((rc = ensure_lu(device, n, eb))) return rc;

    // Factor a copy of A in the chosen precision; the copy also gives |A|'s row sums. Beyond one panel,
    // the first panel's columns are copied first, and the rest of the copy and the row sums run on the
    // LU side stream beside that panel (factor() waits for them before its first step's work).
    auto copy_in = [&](auto *target) -> int {
        Workspace &ws = device->lu;
        const bool split = n > kPanel;
        hipStream_t s = split ? ws.side : device->stream;
        if (split) {
            if (int e = status(copy_columns(matrix, n, n, kPanel, target, device->stream), "copy")) return e;
            if (int e = status(hipEventRecord(ws.columns_ready, device->stream), "event")) return e;
            if (int e = status(hipStreamWaitEvent(ws.side, ws.columns_ready, 0), "event")) return e;
        }
        if (int e = status(copy_with_row_sums(matrix, n, n, target, sums.as<double>(), s, split ? kPanel : 0), "copy"))
            return e;
        if (int e = status(column_max(sums.as<double>(), n, 1, partial.as<double>() + chunks * w * 3, s), "norm"))
            return e;
        if (split) {
            if (int e = status(hipEventRecord(ws.panel_done, ws.side), "event")) return e;
            ws.head_pending = true;
        }
        return SCIHIP_OK;
    };
    std::vector<int64_t> pivots;
    if (wide) {
        if ((rc = copy_in(factors.as<double>()))) return rc;
        rc = factor_on_device<double>(device, factors.as<double>(), n, kFloat64, pivots,
                                      blocked ? inverses.as<double>() : nullptr);
    } else {
        if ((rc = copy_in(factors.as<float>()))) return rc;
        const int product_mode = mode == SCIHIP_FACTOR_FLOAT32 ? kFloat32 : mode == SCIHIP_FACTOR_SPLIT ? kSplit : kBfloat16;
        rc = factor_on_device<float>(device, factors.as<float>(), n, product_mode, pivots,
                                     blocked ? inverses.as<float>() : nullptr);
    }
    if (rc) return rc;
    stats->factor_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
    mark("factor");
    // The row order goes up from pinned memory without a wait (nothing writes lu_index again before the
    // waits below).
    {
        const std::vector<int64_t> rows = rows_of(pivots);
        std::copy(rows.begin(), rows.end(), device->lu_index);
        if ((rc = status(hipMemcpyAsync(device_rows.pointer, device->lu_index, order * sizeof(int64_t),
                                        hipMemcpyHostToDevice, device->stream), "hipMemcpy")))
            return rc;
    }
    // (the factorisation also made the diagonal blocks' inverses and links when `blocked`)

    // |A|_inf: folded with the first residual's column maxima (no wait here).
    double norm = 0.0;

    // x = solve(b)
    auto solve_into = [&](const double *source, bool add) -> int {
        int e;
        if (wide) e = solve_on_device<double, double>(device, factors.as<double>(), device_rows.as<int64_t>(), source,
                                                      correction.as<double>(), n, w,
                                                      blocked ? inverses.as<double>() : nullptr,
                                                      blocked ? packed.pointer : nullptr);
        else e = solve_on_device<float, double>(device, factors.as<float>(), device_rows.as<int64_t>(), source,
                                                correction.as<float>(), n, w, blocked ? inverses.as<float>() : nullptr,
                                                blocked ? packed.pointer : nullptr);
        if (e) return e;
        if (!add) {
            if (wide) return status(hipMemcpyAsync(solution, correction.pointer, order * width * 8,
                                                   hipMemcpyDeviceToDevice, device->stream), "copy");
            return status(convert(correction.as<float>(), solution, n * w, device->stream), "convert");
        }
        return wide ? status(accumulate(solution, correction.as<double>(), n * w, device->stream), "accumulate")
                    : status(accumulate(solution, correction.as<float>(), n * w, device->stream), "accumulate");
    };
    if ((rc = solve_into(right, false))) return rc;
    mark("solve");

    const double *a_operand = matrix;
    if (!aligned) {
        if ((rc = status(pack(kFloat64, matrix, true, n, n, n, false, packed_a.pointer, nullptr, kp, device->stream),
                         "pack")))
            return rc;
        a_operand = packed_a.as<double>();
    }
    for (int iteration = 0;; ++iteration) {
        // r = b - A x on the fp64 matrix cores
        if ((rc = status(hipMemcpyAsync(residual.pointer, right, order * width * 8, hipMemcpyDeviceToDevice,
                                        device->stream), "copy")))
            return rc;
        if (narrow) {
            if ((rc = status(narrow_residual(matrix, n, solution, w, residual.as<double>(),
                                             residual_slices.as<double>(), device->stream),
                             "residual")))
                return rc;
        } else {
            if ((rc = status(pack(kFloat64, solution, true, w, n, w, true, packed_x.pointer, nullptr, kp,
                                  device->stream),
                             "pack")))
                return rc;
            ProductArgs args{a_operand, nullptr, packed_x.pointer, nullptr, residual.pointer, n, w, kp, kp, kp, w,
                             -1.0, 1.0};
            if ((rc = status(product_split(kFloat64, args, slices, residual_slices.pointer, device->stream),
                             "residual")))
                return rc;
        }
        double *partials = partial.as<double>();
        if ((rc = status(column_max(residual.as<double>(), n, w, partials, device->stream), "norm"))) return rc;
        if ((rc = status(column_max(solution, n, w, partials + chunks * w, device->stream), "norm"))) return rc;
        if ((rc = status(column_max(right, n, w, partials + 2 * chunks * w, device->stream), "norm"))) return rc;
        // The column maxima folded on the device straight into pinned host memory, with |A|_inf on the
        // first pass and the bounded waits' error flag: one wait, no copy.
        if ((rc = ensure_lu_host(device, size_t(3 * w + 2), 0))) return rc;
        double *folded = device->lu_host;
        if ((rc = status(fold_norms(partials, chunks, w, 3, iteration == 0 ? chunks : 0, device->lu.error, folded,
                                    device->stream), "norm")))
            return rc;
        if ((rc = status(hipStreamSynchronize(device->stream), "LU"))) return rc;
        if (folded[3 * w + 1] != 0.0) {
            (void)hipMemset(device->lu.error, 0, sizeof(unsigned));
            fail(spin_failure());
            return SCIHIP_EDEVICE;
        }
        mark("residual");
        if (iteration == 0) norm = folded[3 * w];
        double worst = 0.0;
        for (int64_t c = 0; c < w; ++c) {
            const double r = folded[c], x = folded[w + c], b = folded[2 * w + c];
            const double denominator = norm * x + b;
            const double e = denominator > 0.0 ? r / denominator : r;
            worst = std::isnan(e) || std::isnan(worst) ? std::numeric_limits<double>::quiet_NaN() : std::max(worst, e);
        }
        stats->backward_error = worst;
        stats->iterations = iteration;
        if (worst <= tolerance) {
            stats->converged = 1;
            break;
        }
        if (iteration >= max_iterations || std::isnan(worst)) break;
        if ((rc = solve_into(residual.as<double>(), true))) return rc;
        mark("correct");
    }
    stats->solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
    if (tracing) std::fprintf(stderr, "[scihip] solve stages ms:%s\n", trace_line.c_str());
    return SCIHIP_OK;
}

int scihip_stream(scihip_device *device, void **stream) {
    if (!device || !stream) return SCIHIP_EINVAL;
    *stream = device->stream;
    return SCIHIP_OK;
}

int scihip_fill_uniform(scihip_device *device, void *output, uint32_t precision, size_t count, uint64_t seed,
                        double lower, double upper, uint64_t nan_seed, double nan_fraction) {
    if (!device || precision > SCIHIP_WIDE) return SCIHIP_EINVAL;
    std::lock_guard<std::mutex> lock(device->guard);
    const size_t bytes = count * (precision == SCIHIP_SINGLE ? 4 : 8);
    if (count && !device->allocator->locate(output, bytes)) {
        fail("the output is not device memory of this device");
        return SCIHIP_DECLINED;
    }
    return status(fill_uniform(output, int(precision), count, seed, lower, upper, nan_seed, nan_fraction,
                               device->stream),
                  "fill_uniform");
}

int scihip_transpose(scihip_device *device, uint32_t precision, const void *input, void *output, size_t rows,
                     size_t columns) {
    if (!device || precision > SCIHIP_WIDE || (input == output && rows * columns > 0)) return SCIHIP_EINVAL;
    if (rows == 0 || columns == 0) return SCIHIP_OK;
    std::lock_guard<std::mutex> lock(device->guard);
    const size_t bytes = rows * columns * (precision == SCIHIP_WIDE ? 8 : 4);
    if (!device->allocator->locate(input, bytes) || !device->allocator->locate(output, bytes)) {
        fail("an operand of the transpose is not device memory of this device");
        return SCIHIP_DECLINED;
    }
    const hipError_t e = precision == SCIHIP_WIDE
                                 ? transpose<double>(static_cast<const double *>(input), static_cast<double *>(output),
                                                     int64_t(rows), int64_t(columns), device->stream)
                                 : transpose<float>(static_cast<const float *>(input), static_cast<float *>(output),
                                                    int64_t(rows), int64_t(columns), device->stream);
    return status(e, "transpose");
}

int scihip_reduce_axis(scihip_device *device, const scihip_chunk *chunk, uint32_t fold, size_t rows, size_t columns,
                       uint32_t axis, void *output) {
    if (!device || !chunk || fold > SCIHIP_FOLD_MAXIMUM || axis > 1 || chunk->precision > SCIHIP_WIDE)
        return SCIHIP_EINVAL;
    std::lock_guard<std::mutex> lock(device->guard);
    return chunk->precision == SCIHIP_SINGLE
                   ? reduce_axis<float>(device, *chunk, fold, rows, columns, axis, static_cast<float *>(output))
                   : reduce_axis<double>(device, *chunk, fold, rows, columns, axis, static_cast<double *>(output));
}

int scihip_capture_begin(scihip_device *device) {
    if (!device) return SCIHIP_EINVAL;
    std::lock_guard<std::mutex> lock(device->guard);
    if (device->capturing) {
        fail("a graph is already being captured on this device");
        return SCIHIP_EINVAL;
    }
    if (int rc = status(hipStreamBeginCapture(device->stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture"))
        return rc;
    device->capturing = true;
    device->held.clear();
    device->allocator->hold(&device->held);
    return SCIHIP_OK;
}

int scihip_capture_end(scihip_device *device, scihip_graph **graph) {
    if (!device || !graph) return SCIHIP_EINVAL;
    *graph = nullptr;
    std::lock_guard<std::mutex> lock(device->guard);
    if (!device->capturing) {
        fail("no graph is being captured on this device");
        return SCIHIP_EINVAL;
    }
    device->capturing = false;
    device->allocator->hold(nullptr);
    auto out = std::make_unique<scihip_graph>();
    out->held.swap(device->held);
    int rc = status(hipStreamEndCapture(device->stream, &out->graph), "hipStreamEndCapture");
    if (!rc) rc = status(hipGraphInstantiate(&out->exec, out->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (rc) {
        if (out->graph) (void)hipGraphDestroy(out->graph);
        for (void *block : out->held) device->allocator->dispose(block);
        return rc;
    }
    *graph = out.release();
    return SCIHIP_OK;
}

int scihip_replay(scihip_device *device, const scihip_graph *graph) {
    if (!device || !graph) return SCIHIP_EINVAL;
    std::lock_guard<std::mutex> lock(device->guard);
    if (int rc = refuse_while_capturing(device, "scihip_replay")) return rc;
    return status(hipGraphLaunch(graph->exec, device->stream), "hipGraphLaunch");
}

int scihip_graph_release(scihip_device *device, scihip_graph *graph) {
    if (!device) return SCIHIP_EINVAL;
    if (!graph) return SCIHIP_OK;
    std::lock_guard<std::mutex> lock(device->guard);
    if (int rc = refuse_while_capturing(device, "scihip_graph_release")) return rc;
    // Replays already queued keep running; the graph's scratch is returned only after them.
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    for (void *block : graph->held) device->allocator->dispose(block);
    delete graph;
    return SCIHIP_OK;
}

} // extern "C"
