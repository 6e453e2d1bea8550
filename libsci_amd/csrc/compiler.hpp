// Specialised chunk kernels: each distinct chunk signature is turned into straight-line HIP source,
// compiled with hipRTC on a background thread and loaded as a module. Replaces the reference's
// background Slang JIT (src/core/drivers/vulkan/compiler.cpp:35-222): the first sightings of a chunk
// run on the interpreter, later ones on its own kernel once that is ready.
#pragma once

#include "scihip.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>

namespace scihip {

    struct Kernel {
        hipModule_t module = nullptr;
        hipFunction_t function = nullptr;
        bool failed = false;
        std::string log;
    };

    class Compiler {
    public:
        explicit Compiler(int ordinal);
        ~Compiler();

        // The kernel for `key`, or null while it is being compiled. The first call for a key queues
        // a compile of the source `generate` returns (called only then); with `wait` the call
        // compiles in the foreground instead.
        const Kernel *find(const std::string &key, const std::function<std::string()> &generate, bool wait);

        // The kernel for `key` when it is compiled and loaded, else null. Never queues a compile.
        const Kernel *loaded(const std::string &key);

        // The symbol a key's kernel is compiled under: chunk_<kind>_<hash of the key>.
        static std::string symbol(const std::string &key);

        uint64_t compiled() const { return done; }

        // Straight-line source for a chunk. `mode` 0 evaluates, 1 reduces; `contract` marks a sum
        // whose last multiply is a dot product.
        static std::string generate(const scihip_chunk &chunk, int mode, bool contract, uint32_t fold = 0);

        // Straight-line source folding a chunk along one axis (axis 1: a workgroup per row; axis 0:
        // W columns per thread over a slice of the rows), with kUnroll vectors in flight per thread.
        static std::string generate_axis(const scihip_chunk &chunk, int axis, bool contract, uint32_t fold = 0);

        static constexpr size_t limit = 1024; // distinct chunks compiled at most (vulkan/compiler.cpp:27)
        static constexpr int kUnroll = 4;     // vectors per thread per sweep of a specialised reduction
        static constexpr int kStoreThreads = 64; // threads per workgroup of a specialised store (one vector each)
        static constexpr uint64_t kStoreGrid = uint64_t{1} << 24; // at most this many store workgroups

    private:
        void loop();
        void build(const std::string &key, const std::string &source, Kernel &kernel);

        int ordinal;
        std::mutex guard;
        std::condition_variable signal;
        std::unordered_map<std::string, std::unique_ptr<Kernel>> kernels;
        std::unordered_map<std::string, bool> ready;
        std::deque<std::pair<std::string, std::string>> queue;
        std::thread worker;
        bool stopping = false;
        std::atomic<uint64_t> done{0};
        bool trace = false;
    };

} // namespace scihip
