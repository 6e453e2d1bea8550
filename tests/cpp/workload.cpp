// The C++ host mirror (include/sci/hip.hpp) driven the way a libsci user would drive sci::Device.
//
//   workload record           prints, one per line, the bytecode the mirror's Chunk records for the
//                             workload expressions (no GPU needed)
//   workload run DIRECTORY    runs the small CPU-runnable workload on GPU 0 and writes its results
//                             as raw little-endian files into DIRECTORY for the test to check:
//                             c = a*b + 2 with a = linspace(0, 1, 1e6), b = ones (c.f32), the total
//                             sum(sin(c)) (total.f64), the gradient of sum(sinh(a*b + c) * exp(-(d*d)))
//                             on 1000 elements of the inputs in inputs.f32 (gradients.f32, loss.f64),
//                             a 64 x 48 x 40 product (product.f32) and the 3 x 3 solve KAT (solve.f64).
//   workload split PROGRAM DIRECTORY COUNT INPUTS
//                             runs an expression too large for one chunk (a postfix program) on GPU 0:
//                             prints its split into chunks and writes its result and folds.
#include "sci/hip.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace sci::hip;

namespace {

    void print(const char *name, const Chunk &chunk) {
        const scihip_chunk &c = chunk.data();
        std::printf("%s %u %u %u %u %u", name, c.length, c.arguments, c.literals, c.temporaries, c.result);
        for (int s = 0; s < c.length; ++s) {
            const scihip_instruction &in = c.instructions[s];
            std::printf(" %u:%u:%u:%u:%u", in.code, in.target, in.sources[0], in.sources[1], in.sources[2]);
        }
        for (int i = 0; i < c.literals; ++i) std::printf(" L%.17g", c.constants[i]);
        std::printf("\n");
    }

    template <typename T> void write(const std::string &path, const std::vector<T> &values) {
        FILE *f = std::fopen(path.c_str(), "wb");
        if (!f || std::fwrite(values.data(), sizeof(T), values.size(), f) != values.size()) {
            std::fprintf(stderr, "cannot write %s\n", path.c_str());
            std::exit(1);
        }
        std::fclose(f);
    }

    // float(start + i * step) with a double step (include/operations/creation.hpp:65-84).
    std::vector<float> linspace(float start, float stop, size_t count) {
        std::vector<float> out(count);
        const double step = static_cast<double>(stop - start) / (static_cast<double>(count) - 1.0);
        for (size_t i = 0; i < count; ++i) out[i] = static_cast<float>(static_cast<double>(start) + double(i) * step);
        return out;
    }

    int record() {
        // Distinct dummy addresses stand for the arrays; nothing is dereferenced. Every call is its own
        // statement: operands are recorded left to right, as Node::record does (node.hpp:75-88).
        std::vector<const void *> at;
        for (int k = 0; k < 4; ++k) at.push_back(reinterpret_cast<const void *>(uintptr_t(0x1000 + 64 * k)));
        {
            Chunk c; // a * b + 2.0f
            const uint8_t a = c.input(at[0]);
            const uint8_t b = c.input(at[1]);
            const uint8_t product = c.emit(Code::multiply, a, b);
            const uint8_t two = c.constant(2.0f);
            c.output(c.emit(Code::add, product, two));
            print("c1_store", c);
        }
        {
            Chunk c; // sum(sin(c))
            const uint8_t x = c.input(at[0]);
            c.reduce(c.emit(Code::sin, x));
            print("c1_reduce", c);
        }
        {
            Chunk c; // exponential(a) * 0.5f + logarithm(absolute(b) + 1.0f) * c
            const uint8_t a = c.input(at[0]);
            const uint8_t ea = c.emit(Code::exponential, a);
            const uint8_t half = c.constant(0.5f);
            const uint8_t left = c.emit(Code::multiply, ea, half);
            const uint8_t b = c.input(at[1]);
            const uint8_t ab = c.emit(Code::absolute, b);
            const uint8_t one = c.constant(1.0f);
            const uint8_t shifted = c.emit(Code::add, ab, one);
            const uint8_t log = c.emit(Code::logarithm, shifted);
            const uint8_t x = c.input(at[2]);
            const uint8_t right = c.emit(Code::multiply, log, x);
            c.output(c.emit(Code::add, left, right));
            print("c2_store", c);
        }
        {
            Chunk c; // sin(a) * cos(c) + tanh(b)
            const uint8_t a = c.input(at[0]);
            const uint8_t s = c.emit(Code::sin, a);
            const uint8_t x = c.input(at[2]);
            const uint8_t k = c.emit(Code::cos, x);
            const uint8_t m = c.emit(Code::multiply, s, k);
            const uint8_t b = c.input(at[1]);
            const uint8_t t = c.emit(Code::tanh, b);
            c.reduce(c.emit(Code::add, m, t));
            print("c2_reduce", c);
        }
        {
            Chunk c; // sinh(a * b + c) * exponential(-(d * d))
            const uint8_t a = c.input(at[0]);
            const uint8_t b = c.input(at[1]);
            const uint8_t ab = c.emit(Code::multiply, a, b);
            const uint8_t x = c.input(at[2]);
            const uint8_t z = c.emit(Code::add, ab, x);
            const uint8_t s = c.emit(Code::sinh, z);
            const uint8_t d = c.input(at[3]);
            const uint8_t dd = c.emit(Code::multiply, d, d);
            const uint8_t n = c.emit(Code::negate, dd);
            const uint8_t e = c.emit(Code::exponential, n);
            c.reduce(c.emit(Code::multiply, s, e));
            print("c5_loss", c);
        }
        return 0;
    }

    int run(const std::string &directory) {
        Device device(0);
        std::printf("device %s\n", device.name());

        const size_t n = 1000000;
        Array<float> a = Array<float>::from(device, linspace(0.0f, 1.0f, n));
        Array<float> b = Array<float>::from(device, std::vector<float>(n, 1.0f));
        Array<float> c = evaluate(a * b + 2.0f);
        const double total = reduce_raw(sin(c), Fold::sum);
        write(directory + "/c.f32", c.host());
        write(directory + "/total.f64", std::vector<double>{total});

        const size_t m = 1000;
        std::vector<float> inputs(4 * m);
        for (size_t i = 0; i < inputs.size(); ++i) inputs[i] = -2.0f + 4.0f * float((i * 2654435761u) % 1000) / 1000.0f;
        std::vector<Array<float>> x;
        for (int k = 0; k < 4; ++k)
            x.push_back(Array<float>::from(device, std::vector<float>(inputs.begin() + long(k * m),
                                                                      inputs.begin() + long((k + 1) * m))));
        double loss = 0.0;
        const std::vector<Array<float>> g =
                gradient(sinh(x[0] * x[1] + x[2]) * exponential(-(x[3] * x[3])), x, &loss);
        std::vector<float> gradients;
        for (const Array<float> &each : g) {
            const std::vector<float> h = each.host();
            gradients.insert(gradients.end(), h.begin(), h.end());
        }
        write(directory + "/inputs.f32", inputs);
        write(directory + "/gradients.f32", gradients);
        write(directory + "/loss.f64", std::vector<double>{loss});

        std::vector<float> left(64 * 40), right(40 * 48);
        for (size_t i = 0; i < left.size(); ++i) left[i] = float(int(i % 7) - 3) * 0.25f;
        for (size_t i = 0; i < right.size(); ++i) right[i] = float(int(i % 5) - 2) * 0.5f;
        Array<float> l = Array<float>::from(device, left, {64, 40});
        Array<float> r = Array<float>::from(device, right, {40, 48});
        const Array<float> p = product(l, r, Mode::float32);
        write(directory + "/product.f32", p.host());
        // .npy out and back in (sci/npy.hpp): the file numpy reads, and the array loaded from it
        p.save(directory + "/product.npy");
        const Array<float> reloaded = Array<float>::load(device, directory + "/product.npy");
        std::printf("reloaded shape %zu x %zu\n", reloaded.shape()[0], reloaded.shape()[1]);
        write(directory + "/reloaded.f32", reloaded.host());

        // include/linear tests: A x = b with x = [1, -2, 3] (tests/linear/test_solve.cpp:8-25)
        Array<double> A = Array<double>::from(device, {4, -2, 1, -2, 4, -2, 1, -2, 4}, {3, 3});
        Array<double> B = Array<double>::from(device, {11, -16, 17}, {3, 1});
        Array<double> X(device, {3, 1});
        const scihip_solve_stats stats = device.solve(Factoring::float32, A.data(), B.data(), X.data(), 3, 1);
        std::printf("solve: %d refinements, backward error %.3g\n", stats.iterations, stats.backward_error);
        write(directory + "/solve.f64", X.host());

        // Element access through the host copy (include/array/array.hpp:186-217): reads after device
        // work see its result, writes reach the next device use, copies share the elements.
        Array<float> e = Array<float>::from(device, std::vector<float>{1, 2, 3, 4, 5, 6}, {2, 3});
        std::vector<float> seen;
        seen.push_back(e.at(1, 2));       // 6
        Array<float> alias = e;           // shares the allocation and its host copy
        e[0] = 10.0f;
        alias.at(1, 1) = 20.0f;
        seen.push_back(alias[0]);         // 10
        Array<float> twice = evaluate(e * 2.0f);
        for (size_t i = 0; i < twice.size(); ++i) seen.push_back(twice[i]); // 20 4 6 8 40 12
        seen.push_back(twice.at(1, 1));   // 40
        assign(twice, Expr<float>(e) + 1.0f); // a device write into an array whose copy was read
        seen.push_back(twice[0]);         // 11
        seen.push_back(twice.at(1, 1));   // 21
        const std::vector<float> host = e.host();
        seen.insert(seen.end(), host.begin(), host.end()); // 10 2 3 4 20 6
        write(directory + "/elements.f32", seen);
        device.synchronize();
        return 0;
    }

    template <typename T> std::vector<T> read(const std::string &path, size_t count) {
        std::vector<T> values(count);
        FILE *f = std::fopen(path.c_str(), "rb");
        if (!f || std::fread(values.data(), sizeof(T), count, f) != count) {
            std::fprintf(stderr, "cannot read %s\n", path.c_str());
            std::exit(1);
        }
        std::fclose(f);
        return values;
    }

    // An expression too large for one chunk: the postfix program in PROGRAM (lines "in K",
    // "const V", "op CODE") over INPUTS arrays of COUNT floats read from DIRECTORY/inputK.f32. Prints
    // the split of its store and of its reduction (one chunk per stage), then writes the stored result
    // (split.f32) and its sum, minimum and maximum (folds.f64) into DIRECTORY.
    int split_run(const std::string &path, const std::string &directory, size_t count, int inputs) {
        Device device(0);
        std::vector<Array<float>> arrays;
        for (int k = 0; k < inputs; ++k)
            arrays.push_back(Array<float>::from(device, read<float>(directory + "/input" + std::to_string(k) + ".f32",
                                                                    count)));
        std::vector<Expr<float>> stack;
        FILE *f = std::fopen(path.c_str(), "r");
        if (!f) return 1;
        char kind[16];
        double value;
        while (std::fscanf(f, "%15s %lf", kind, &value) == 2) {
            if (std::strcmp(kind, "in") == 0) stack.emplace_back(arrays.at(size_t(value)));
            else if (std::strcmp(kind, "const") == 0) stack.emplace_back(float(value));
            else {
                const Code code = static_cast<Code>(int(value));
                const size_t arity = code == Code::fuse ? 3 : int(code) >= int(Code::add) ? 2 : 1;
                std::vector<Expr<float>> operands(stack.end() - long(arity), stack.end());
                stack.erase(stack.end() - long(arity), stack.end());
                stack.emplace_back(code, operands);
            }
        }
        std::fclose(f);
        const Expr<float> e = stack.at(0);
        for (const bool reduce : {false, true}) {
            const Split<float> parts = split(e, reduce);
            for (const Expr<float> &stage : parts.stages) print(reduce ? "reduce_stage" : "store_stage", stage.chunk(false));
            print(reduce ? "reduce_last" : "store_last", parts.last.chunk(reduce));
        }
        write(directory + "/split.f32", evaluate(e).host());
        write(directory + "/folds.f64", std::vector<double>{reduce_raw(e, Fold::sum), reduce_raw(e, Fold::minimum),
                                                            reduce_raw(e, Fold::maximum)});
        return 0;
    }

} // namespace

int main(int argc, char **argv) {
    if (argc >= 6 && std::strcmp(argv[1], "split") == 0) {
        try {
            return split_run(argv[2], argv[3], std::strtoull(argv[4], nullptr, 10), std::atoi(argv[5]));
        } catch (const Error &e) {
            std::fprintf(stderr, "error: %s\n", e.what());
            return 1;
        }
    }
    if (argc >= 2 && std::strcmp(argv[1], "record") == 0) return record();
    if (argc >= 3 && std::strcmp(argv[1], "run") == 0) {
        try {
            return run(argv[2]);
        } catch (const Error &e) {
            std::fprintf(stderr, "error: %s\n", e.what());
            return 1;
        }
    }
    std::fprintf(stderr, "usage: %s record | run DIRECTORY | split PROGRAM DIRECTORY COUNT INPUTS\n", argv[0]);
    return 2;
}
