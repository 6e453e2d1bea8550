// Specialised chunk kernels compiled with hipRTC on a background thread (see compiler.hpp).
#include "compiler.hpp"
#include "runtime.hpp"

#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <vector>

namespace scihip {

    // device_math.hpp, interpreter.hpp and adjoint.hpp, embedded at build time (build.py writes embedded.inc).
#include "embedded.inc"


    std::string Compiler::generate(const scihip_chunk &chunk, int mode, bool contract, uint32_t fold) {
        const bool single = chunk.precision == SCIHIP_SINGLE;
        const char *type = single ? "float" : "double";
        const int width = single ? 4 : 2;
        std::ostringstream o;
        o << "extern \"C\" __global__ __launch_bounds__(" << (mode == 0 ? kStoreThreads : 256)
          << ") void scihip_specialised(const scihip::Launch<" << type << "> a) {\n"
          << "    using namespace scihip;\n"
          << "    typedef " << type << " T;\n"
          << "    constexpr int W = " << width << ";\n"
          << "    constexpr int U = " << kUnroll << ";\n"
          << "    typedef Lanes<T, W> L;\n"
          << "    const Program<T> &p = a.p;\n";
        for (int i = 0; i < chunk.literals; ++i)
            o << "    const L k" << i << " = splat<T, W>(p.constants[" << i << "]);\n";
        // The fold is a compile-time constant of the kernel, so each element's fold is straight-line code.
        if (mode == 1)
            o << "    constexpr uint32_t F = " << fold << ";\n    double total = identity(F);\n    uint32_t poisoned = 0;\n";
        // Each thread takes U vectors 256 apart per sweep; all of their loads are issued before any
        // of them is computed, so U * inputs 16-byte loads are in flight per thread.
        const int steps = contract ? chunk.length - 1 : chunk.length;
        // The chunk's instructions on vector u; `bulk`: the vector is inside the arrays, which are
        // aligned (no checks).
        auto body = [&](int u, bool bulk, const char *indent) {
            std::string temporaries[65];
            const std::string suffix = "_" + std::to_string(u);
            const std::string full = bulk ? "true" : "full" + std::to_string(u);
            auto name = [&](uint8_t id) -> std::string {
                const uint8_t tag = id & SCIHIP_SLOT_TYPES, index = id & 0x3F;
                if (tag == SCIHIP_SLOT_INPUT) return "i" + std::to_string(index) + suffix;
                if (tag == SCIHIP_SLOT_LITERAL) return "k" + std::to_string(index);
                if (tag == SCIHIP_SLOT_TEMPORARY) return temporaries[index];
                return temporaries[64];
            };
            for (int s = 0; s < steps; ++s) {
                const scihip_instruction &in = chunk.instructions[s];
                const std::string first = name(in.sources[0]);
                const std::string second = in.sources[1] == SCIHIP_SLOT_NONE ? first : name(in.sources[1]);
                const std::string third = in.sources[2] == SCIHIP_SLOT_NONE ? first : name(in.sources[2]);
                const std::string target = "s" + std::to_string(s) + suffix;
                o << indent << "const L " << target << " = lanes<" << unsigned(in.code) << "u, T, W>(" << first << ", "
                  << second << ", " << third << ");\n";
                if (in.target == SCIHIP_SLOT_OUTPUT) temporaries[64] = target;
                else temporaries[in.target & 0x3F] = target;
            }
            if (mode == 0) {
                if (bulk) o << indent << "store_vector<T, W>(p, at" << u << ", " << name(SCIHIP_SLOT_OUTPUT) << ");\n";
                else o << indent << "store<T, W>(p, at" << u << ", " << full << ", " << name(SCIHIP_SLOT_OUTPUT) << ");\n";
            } else if (contract) {
                const scihip_instruction &last = chunk.instructions[chunk.length - 1];
                o << indent << "contract<T, W>(total, " << name(last.sources[0]) << ", " << name(last.sources[1])
                  << ", at" << u << ", " << full << ", p.count);\n";
            } else {
                o << indent << "consume<T, W>(F, total, poisoned, " << name(chunk.result) << ", at" << u << ", " << full
                  << ", p.count);\n";
            }
        };
        if (mode == 0) {
            // A store: one 16-byte vector per thread in 64-thread workgroups, the grid covering the
            // arrays once in order (the loop strides only past the grid cap). The resident waves
            // then sweep memory together in a window a few MB wide; a fixed grid striding through
            // the arrays spread its accesses over hundreds of MB and streamed 10% slower.
            o << "    const uint64_t steps = (p.count + W - 1) / W;\n"
              << "    const uint64_t stride = (uint64_t)gridDim.x * " << kStoreThreads << ";\n"
              << "    for (uint64_t v = (uint64_t)blockIdx.x * " << kStoreThreads << " + threadIdx.x; v < steps; v += stride) {\n"
              << "        const uint64_t at0 = v * W;\n"
              << "        if (p.aligned && at0 + W <= p.count) {\n";
            for (int i = 0; i < chunk.arguments; ++i)
                o << "            const L i" << i << "_0 = fetch_vector<T, W>(p, " << i << ", at0);\n";
            body(0, true, "            ");
            o << "        } else {\n"
              << "            const bool full0 = at0 + W <= p.count;\n";
            for (int i = 0; i < chunk.arguments; ++i)
                o << "            const L i" << i << "_0 = fetch<T, W>(p, " << i << ", at0, full0);\n";
            body(0, false, "            ");
            o << "        }\n    }\n}\n";
            return o.str();
        }
        o << "    const uint64_t steps = (p.count + W - 1) / W;\n"
          << "    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * U;\n"
          << "    uint64_t base = (uint64_t)blockIdx.x * blockDim.x * U + threadIdx.x;\n";
        // Sweeps whose U vectors all lie inside aligned arrays: unconditional 16-byte loads and stores,
        // so no branch separates the loads from one another. The rest of the sweeps follow, checked.
        o << "    if (p.aligned) {\n"
          << "        const uint64_t whole = p.count / W;\n"
          << "        for (; base + " << 256 * (kUnroll - 1) << " < whole; base += stride) {\n";
        for (int u = 0; u < kUnroll; ++u) o << "            const uint64_t at" << u << " = (base + " << 256 * u << ") * W;\n";
        for (int u = 0; u < kUnroll; ++u)
            for (int i = 0; i < chunk.arguments; ++i)
                o << "            const L i" << i << "_" << u << " = fetch_vector<T, W>(p, " << i << ", at" << u << ");\n";
        for (int u = 0; u < kUnroll; ++u) body(u, true, "            ");
        o << "        }\n    }\n";
        o << "    for (; base < steps; base += stride) {\n";
        for (int u = 0; u < kUnroll; ++u)
            o << "        const uint64_t v" << u << " = base + " << 256 * u << ";\n"
              << "        const bool live" << u << " = v" << u << " < steps;\n"
              << "        const uint64_t at" << u << " = v" << u << " * W;\n"
              << "        const bool full" << u << " = at" << u << " + W <= p.count;\n";
        for (int u = 0; u < kUnroll; ++u)
            for (int i = 0; i < chunk.arguments; ++i)
                o << "        L i" << i << "_" << u << "{}; if (live" << u << ") i" << i << "_" << u
                  << " = fetch<T, W>(p, " << i << ", at" << u << ", full" << u << ");\n";
        for (int u = 0; u < kUnroll; ++u) {
            o << "        if (live" << u << ") {\n";
            body(u, false, "            ");
            o << "        }\n";
        }
        o << "    }\n";
        if (mode == 1) o << "    leave_partial(F, a.partials, total, poisoned);\n";
        o << "}\n";
        return o.str();
    }

    std::string Compiler::generate_axis(const scihip_chunk &chunk, int axis, bool contract, uint32_t fold) {
        const bool single = chunk.precision == SCIHIP_SINGLE;
        const char *type = single ? "float" : "double";
        const int width = single ? 4 : 2;
        std::ostringstream o;
        o << "extern \"C\" __global__ __launch_bounds__(256) void scihip_specialised(const scihip::AxisLaunch<" << type
          << "> a) {\n"
          << "    using namespace scihip;\n"
          << "    typedef " << type << " T;\n"
          << "    constexpr int W = " << width << ";\n"
          << "    typedef Lanes<T, W> L;\n"
          << "    const Program<T> &p = a.p;\n"
          << "    constexpr uint32_t F = " << fold << ";\n";
        for (int i = 0; i < chunk.literals; ++i)
            o << "    const L k" << i << " = splat<T, W>(p.constants[" << i << "]);\n";
        const int steps = contract ? chunk.length - 1 : chunk.length;
        // The chunk's instructions for vector u, as straight-line code over named values.
        auto body = [&](int u, const char *indent) {
            std::string temporaries[65];
            const std::string suffix = "_" + std::to_string(u);
            auto name = [&](uint8_t id) -> std::string {
                const uint8_t tag = id & SCIHIP_SLOT_TYPES, index = id & 0x3F;
                if (tag == SCIHIP_SLOT_INPUT) return "i" + std::to_string(index) + suffix;
                if (tag == SCIHIP_SLOT_LITERAL) return "k" + std::to_string(index);
                if (tag == SCIHIP_SLOT_TEMPORARY) return temporaries[index];
                return temporaries[64];
            };
            for (int s = 0; s < steps; ++s) {
                const scihip_instruction &in = chunk.instructions[s];
                const std::string first = name(in.sources[0]);
                const std::string second = in.sources[1] == SCIHIP_SLOT_NONE ? first : name(in.sources[1]);
                const std::string third = in.sources[2] == SCIHIP_SLOT_NONE ? first : name(in.sources[2]);
                const std::string target = "s" + std::to_string(s) + suffix;
                o << indent << "const L " << target << " = lanes<" << unsigned(in.code) << "u, T, W>(" << first << ", "
                  << second << ", " << third << ");\n";
                if (in.target == SCIHIP_SLOT_OUTPUT) temporaries[64] = target;
                else temporaries[in.target & 0x3F] = target;
            }
            if (contract) {
                const scihip_instruction &last = chunk.instructions[chunk.length - 1];
                return std::make_pair(name(last.sources[0]), name(last.sources[1]));
            }
            return std::make_pair(name(chunk.result), std::string());
        };
        if (axis == 1) {
            o << "    double total = identity(F);\n    uint32_t poisoned = 0;\n"
              << "    const uint64_t first = (uint64_t)blockIdx.x * a.columns, end = first + a.columns;\n"
              << "    const uint64_t vectors = (a.columns + W - 1) / W;\n"
              << "    for (uint64_t base = threadIdx.x; base < vectors; base += 256 * " << kUnroll << ") {\n";
            for (int u = 0; u < kUnroll; ++u)
                o << "        const uint64_t v" << u << " = base + " << 256 * u << ";\n"
                  << "        const bool live" << u << " = v" << u << " < vectors;\n"
                  << "        const uint64_t at" << u << " = first + v" << u << " * W;\n"
                  << "        const bool full" << u << " = at" << u << " + W <= end;\n";
            for (int u = 0; u < kUnroll; ++u)
                for (int i = 0; i < chunk.arguments; ++i)
                    o << "        L i" << i << "_" << u << "{}; if (live" << u << ") i" << i << "_" << u
                      << " = fetch<T, W>(p, " << i << ", at" << u << ", full" << u << ");\n";
            for (int u = 0; u < kUnroll; ++u) {
                o << "        if (live" << u << ") {\n";
                const auto result = body(u, "            ");
                if (contract)
                    o << "            contract<T, W>(total, " << result.first << ", " << result.second << ", at" << u
                      << ", full" << u << ", end);\n";
                else
                    o << "            consume<T, W>(F, total, poisoned, " << result.first << ", at" << u << ", full"
                      << u << ", end);\n";
                o << "        }\n";
            }
            o << "    }\n"
              << "    const Partial folded = fold_block(F, total, poisoned);\n"
              << "    if (threadIdx.x == 0) a.output[blockIdx.x] = folded.poisoned ? T(__builtin_nan(\"\")) : T(folded.value);\n";
        } else {
            o << "    const uint64_t column = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * W;\n"
              << "    if (column >= a.columns) return;\n"
              << "    const uint64_t per_slice = (a.rows + gridDim.y - 1) / gridDim.y;\n"
              << "    const uint64_t begin = (uint64_t)blockIdx.y * per_slice;\n"
              << "    const uint64_t end = begin + per_slice < a.rows ? begin + per_slice : a.rows;\n"
              << "    const bool full = column + W <= a.columns;\n"
              << "    constexpr uint32_t fold = F;\n"
              << "    double total[W];\n    uint32_t poisoned[W];\n"
              << "#pragma unroll\n    for (int k = 0; k < W; ++k) { total[k] = identity(fold); poisoned[k] = 0; }\n"
              << "    for (uint64_t r = begin; r < end; r += " << kUnroll << ") {\n";
            for (int u = 0; u < kUnroll; ++u)
                o << "        const bool live" << u << " = r + " << u << " < end;\n"
                  << "        const uint64_t at" << u << " = (r + " << u << ") * a.columns + column;\n";
            for (int u = 0; u < kUnroll; ++u)
                for (int i = 0; i < chunk.arguments; ++i)
                    o << "        L i" << i << "_" << u << "{}; if (live" << u << ") i" << i << "_" << u
                      << " = fetch<T, W>(p, " << i << ", at" << u << ", full);\n";
            for (int u = 0; u < kUnroll; ++u) {
                o << "        if (live" << u << ") {\n";
                const auto result = body(u, "            ");
                if (contract)
                    o << "#pragma unroll\n            for (int k = 0; k < W; ++k) total[k] = fused((double)" << result.first
                      << ".v[k], (double)" << result.second << ".v[k], total[k]);\n";
                else
                    o << "#pragma unroll\n            for (int k = 0; k < W; ++k) fold_in(fold, total[k], poisoned[k], (double)"
                      << result.first << ".v[k]);\n";
                o << "        }\n";
            }
            o << "    }\n"
              << "    Partial *slice = a.partials + (uint64_t)blockIdx.y * a.columns;\n"
              << "#pragma unroll\n    for (int k = 0; k < W; ++k)\n"
              << "        if (full || column + k < a.columns) slice[column + k] = Partial{total[k], poisoned[k], 0u};\n";
        }
        o << "}\n";
        return o.str();
    }

    Compiler::Compiler(int ordinal) : ordinal(ordinal) {
        if (const char *text = std::getenv("SCIHIP_JIT")) trace = std::strcmp(text, "trace") == 0;
        worker = std::thread([this] { loop(); });
    }

    Compiler::~Compiler() {
        {
            std::lock_guard<std::mutex> lock(guard);
            stopping = true;
        }
        signal.notify_all();
        if (worker.joinable()) worker.join();
        for (auto &entry : kernels)
            if (entry.second->module) (void)hipModuleUnload(entry.second->module);
    }

    const Kernel *Compiler::find(const std::string &key, const std::function<std::string()> &generate, bool wait) {
        std::unique_lock<std::mutex> lock(guard);
        auto found = kernels.find(key);
        if (found == kernels.end()) {
            if (kernels.size() >= limit) return nullptr;
            kernels.emplace(key, std::make_unique<Kernel>());
            ready[key] = false;
            const std::string source = generate();
            if (wait) {
                Kernel *kernel = kernels[key].get();
                lock.unlock();
                build(key, source, *kernel);
                lock.lock();
                ready[key] = true;
                signal.notify_all();
                return kernel->failed ? nullptr : kernel;
            }
            queue.emplace_back(key, source);
            signal.notify_all();
            return nullptr;
        }
        if (!ready[key]) {
            if (!wait) return nullptr;
            signal.wait(lock, [&] { return ready[key]; });
        }
        const Kernel *kernel = found->second.get();
        return kernel->failed ? nullptr : kernel;
    }

    void Compiler::loop() {
        (void)hipSetDevice(ordinal);
        std::unique_lock<std::mutex> lock(guard);
        for (;;) {
            signal.wait(lock, [&] { return stopping || !queue.empty(); });
            if (stopping) return;
            auto job = std::move(queue.front());
            queue.pop_front();
            Kernel *kernel = kernels[job.first].get();
            lock.unlock();
            build(job.first, job.second, *kernel);
            lock.lock();
            ready[job.first] = true;
            signal.notify_all();
        }
    }

    // The kernel's symbol: its kind (the key's first byte: 0 store, 1 reduce, 2 adjoint, 3 column
    // folds, 4 row folds) and a 64-bit FNV-1a hash of the whole key, so that a profile tells the
    // kernels of different chunks apart.
    std::string Compiler::symbol(const std::string &key) {
        static const char *const kinds[] = {"store", "reduce", "adjoint", "columns", "rows"};
        const unsigned kind = key.empty() ? 0u : static_cast<unsigned char>(key[0]);
        uint64_t hash = 0xcbf29ce484222325ull;
        for (const char c : key) {
            hash ^= static_cast<unsigned char>(c);
            hash *= 0x100000001b3ull;
        }
        char text[64];
        std::snprintf(text, sizeof(text), "chunk_%s_%016llx", kind < 5 ? kinds[kind] : "kernel",
                      static_cast<unsigned long long>(hash));
        return text;
    }

    const Kernel *Compiler::loaded(const std::string &key) {
        std::lock_guard<std::mutex> lock(guard);
        const auto found = kernels.find(key);
        if (found == kernels.end() || !ready[key] || found->second->failed) return nullptr;
        return found->second.get();
    }

    void Compiler::build(const std::string &key, const std::string &source, Kernel &kernel) {
        (void)hipSetDevice(ordinal);
        // Every generated source names its kernel scihip_specialised; the macro renames it.
        const std::string name = symbol(key);
        const std::string text = std::string(kEmbeddedSource) + "\n#define scihip_specialised " + name + "\n" + source;
        hiprtcProgram program;
        if (hiprtcCreateProgram(&program, text.c_str(), "chunk.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
            kernel.failed = true;
            return;
        }
        const char *options[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-DSCIHIP_RTC",
                                 "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
        const hiprtcResult compiled = hiprtcCompileProgram(program, sizeof(options) / sizeof(options[0]), options);
        size_t size = 0;
        hiprtcGetProgramLogSize(program, &size);
        if (size > 1) {
            kernel.log.resize(size);
            hiprtcGetProgramLog(program, kernel.log.data());
        }
        if (compiled != HIPRTC_SUCCESS) {
            if (trace) std::fprintf(stderr, "[scihip] specialised compile failed:\n%s\n", kernel.log.c_str());
            hiprtcDestroyProgram(&program);
            kernel.failed = true;
            return;
        }
        hiprtcGetCodeSize(program, &size);
        std::vector<char> code(size);
        hiprtcGetCode(program, code.data());
        hiprtcDestroyProgram(&program);
        if (hipModuleLoadData(&kernel.module, code.data()) != hipSuccess ||
            hipModuleGetFunction(&kernel.function, kernel.module, name.c_str()) != hipSuccess) {
            kernel.failed = true;
            return;
        }
        ++done;
        if (trace) std::fprintf(stderr, "[scihip] compiled %s (%zu bytes)\n", name.c_str(), code.size());
    }

} // namespace scihip
