// Host side of the HIP backend: the device object behind the C ABI (include/scihip.h).
#pragma once

#include "lu.hpp"
#include "scihip.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace scihip {

    void fail(const std::string &message);
    int status(hipError_t error, const char *what);

    // Chunk checks shared by the entry points (runtime.hip).
    bool wellformed(const scihip_chunk &chunk);
    bool contracted(const scihip_chunk &chunk, uint32_t fold);

    constexpr unsigned kMaxBlocks = 16384;     // reduction partials buffer capacity
    constexpr unsigned kStaticPartials = 4096; // partials of scihip_reduce_static's static scratch

    // Device memory with the reference allocator's contract (src/core/drivers/vulkan/allocator.cpp:16-136):
    // blocks are cached and reused when one at most 25% larger fits; `locate` finds the block holding
    // any interior pointer; a disposed block returns to the cache only once the work queued before
    // its disposal has finished (src/core/drivers/vulkan/stream.cpp:117-130).
    class Allocator {
    public:
        explicit Allocator(hipStream_t stream) : stream(stream) {}
        ~Allocator();
        int allocate(size_t bytes, void **pointer);
        int dispose(void *pointer);
        // While a graph is being captured, disposed blocks go to `held` instead (the graph's kernels
        // keep using them on every replay); null ends that.
        void hold(std::vector<void *> *held) { holding = held; }
        bool locate(const void *pointer, size_t bytes = 1) const;
        void reclaim(bool all);

    private:
        struct Retired {
            void *pointer;
            size_t bytes;
            hipEvent_t event;
        };
        void trim(size_t keep);
        hipStream_t stream;
        std::map<uintptr_t, size_t> live;        // start -> bytes
        std::multimap<size_t, void *> cache;     // bytes -> block
        std::vector<Retired> retired;
        size_t cached = 0;
        std::vector<void *> *holding = nullptr;
        static constexpr size_t limit = size_t{16} << 30; // cache at most 16 GiB of 288
    };

    class Compiler;
    struct Workspace;

    // Workgroups for `count` elements at `width` lanes and `unroll` vectors per thread per sweep
    // (mode 0 stores, 1 reduces).
    unsigned blocks_for(scihip_device *device, size_t count, int width, int unroll, int mode);

    // Packs both operands and multiplies on the matrix cores (product.hpp): output (rows x columns)
    // = L * R with L rows x depth and R depth x columns. Each operand is given as a row-major source
    // with its row stride: L itself (R^T) when not transposed, L^T (R) when transposed.
    int multiply(scihip_device *device, int mode, const void *left, int64_t left_ld, bool left_transposed,
                 const void *right, int64_t right_ld, bool right_transposed, void *output, int64_t rows,
                 int64_t columns, int64_t depth);

    struct Kernel;  // a specialised, loaded kernel
    struct Partial; // a reduction partial (interpreter.hpp)

} // namespace scihip

struct scihip_device {
    int ordinal = 0;
    hipStream_t stream = nullptr;
    std::string name;
    std::mutex guard;
    std::unique_ptr<scihip::Allocator> allocator;
    std::unique_ptr<scihip::Compiler> compiler;
    int units = 0;              // compute units
    void *partials = nullptr;   // reduction partials, one per workgroup
    scihip::Partial *static_partials = nullptr; // this GPU's static scratch (scihip_reduce_static)
    double *result = nullptr;   // reduction result on the device
    double *staging = nullptr;  // pinned host word the fold kernel writes a synchronous result into
    scihip::Partial *staging_partial = nullptr; // pinned host partial of a one-workgroup reduction
    double *lu_host = nullptr;  // pinned host: a solve's folded column maxima and the LU error flag
    size_t lu_host_count = 0;   // doubles in lu_host
    int64_t *lu_index = nullptr; // pinned host: a factorisation's pivots, then the solve's row order
    size_t lu_index_count = 0;  // entries in lu_index
    int policy = 0;             // 0 automatic, 1 interpreter only, 2 specialised only
    scihip::Workspace lu;       // LU barrier counter, error flag, solve flags, panel publications
    unsigned epoch = 0;         // launches of the solve kernels so far (their ready-flag values)
    uint64_t interpreted = 0, specialised = 0;
    bool capturing = false;     // between scihip_capture_begin and scihip_capture_end
    std::vector<void *> held;   // scratch blocks the graph being captured uses
};

struct scihip_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void *> held;   // scratch blocks its kernels use, returned on release
};
