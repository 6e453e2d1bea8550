// Host-only test program for sci/npy.hpp (driven by tests/test_npy.py):
//   npy_roundtrip copy <in.npy> <out.npy> f4|f8   reads a file as float / double and writes it back
//   npy_roundtrip make <out.npy>                  writes a 3 x 4 float array with a[i][j] = 0.5 * (4i + j) - 1
// Prints the shape read or written; on a format error prints it to stderr and exits with status 2.
#include "sci/npy.hpp"

#include <cstdio>
#include <string>
#include <vector>

template <typename T> static int copy(const char *in, const char *out) {
    std::vector<size_t> shape;
    const std::vector<T> values = sci::npy::read<T>(in, shape);
    sci::npy::write(out, values, shape);
    std::printf("shape");
    for (size_t d : shape) std::printf(" %zu", d);
    std::printf("\n");
    return 0;
}

int main(int argc, char **argv) {
    try {
        const std::string mode = argc > 1 ? argv[1] : "";
        if (mode == "copy" && argc == 5) {
            const std::string type = argv[4];
            if (type == "f4") return copy<float>(argv[2], argv[3]);
            if (type == "f8") return copy<double>(argv[2], argv[3]);
        } else if (mode == "make" && argc == 3) {
            std::vector<float> values(12);
            for (size_t k = 0; k < values.size(); ++k) values[k] = 0.5f * float(k) - 1.0f;
            sci::npy::write(argv[2], values, {3, 4});
            std::printf("shape 3 4\n");
            return 0;
        }
        std::fprintf(stderr, "usage: npy_roundtrip copy <in> <out> f4|f8 | make <out>\n");
        return 1;
    } catch (const sci::npy::Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
}
