// sci/npy.hpp - reading and writing NumPy's .npy files (format versions 1.0, 2.0 and 3.0) for float
// and double arrays, header-only and host-only. The reference lists .npy I/O as planned
// (README.md:155, "Reading and writing .npy files"); here it carries arrays between C++ programs and
// Python (numpy.save / numpy.load) without a text format in between.
//
// The format: the magic string "\x93NUMPY", a major and a minor version byte, the header's length
// (2 bytes little-endian in version 1.0, 4 bytes in 2.0 and 3.0), then the header, a Python dict
// literal such as {'descr': '<f4', 'fortran_order': False, 'shape': (3, 4), } padded with spaces to
// end in '\n' at a multiple of 64 bytes from the start of the file, then the elements in C order.
// Written files are version 1.0 unless the header needs more than 65535 bytes. Element data is
// little-endian, which is the byte order of the hosts this library runs on (x86-64, AArch64).
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace sci::npy {

    class Error : public std::runtime_error {
    public:
        using std::runtime_error::runtime_error;
    };

    template <typename T> constexpr const char *descr() {
        static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>, "float or double elements");
        return std::is_same_v<T, float> ? "<f4" : "<f8";
    }

    namespace detail {

        struct File {
            std::FILE *f = nullptr;
            File(const std::string &path, const char *mode) : f(std::fopen(path.c_str(), mode)) {
                if (!f) throw Error("cannot open " + path);
            }
            ~File() {
                if (f) std::fclose(f);
            }
            File(const File &) = delete;
            File &operator=(const File &) = delete;
        };

        inline std::string shape_text(const std::vector<size_t> &shape) {
            std::string s = "(";
            for (size_t i = 0; i < shape.size(); ++i) {
                s += std::to_string(shape[i]);
                if (shape.size() == 1 || i + 1 < shape.size()) s += ",";
                if (i + 1 < shape.size()) s += " ";
            }
            return s + ")";
        }

        // The value text of `key` in a header dict: from after "'key':" to the next top-level ',' or '}'.
        inline std::string field(const std::string &header, const std::string &key) {
            const std::string quoted = "'" + key + "'";
            size_t at = header.find(quoted);
            if (at == std::string::npos) throw Error("the .npy header has no " + quoted);
            at = header.find(':', at + quoted.size());
            if (at == std::string::npos) throw Error("malformed .npy header");
            ++at;
            int depth = 0;
            size_t end = at;
            for (; end < header.size(); ++end) {
                const char c = header[end];
                if (c == '(') ++depth;
                else if (c == ')') --depth;
                else if ((c == ',' || c == '}') && depth == 0) break;
            }
            std::string value = header.substr(at, end - at);
            const size_t first = value.find_first_not_of(" \t");
            const size_t last = value.find_last_not_of(" \t");
            return first == std::string::npos ? std::string() : value.substr(first, last - first + 1);
        }

        inline std::vector<size_t> parse_shape(const std::string &text) {
            if (text.size() < 2 || text.front() != '(' || text.back() != ')') throw Error("malformed .npy shape " + text);
            std::vector<size_t> shape;
            size_t i = 1;
            while (i + 1 < text.size()) {
                while (i + 1 < text.size() && (text[i] == ' ' || text[i] == ',')) ++i;
                if (i + 1 >= text.size()) break;
                if (text[i] < '0' || text[i] > '9') throw Error("malformed .npy shape " + text);
                size_t v = 0;
                while (i + 1 < text.size() && text[i] >= '0' && text[i] <= '9') v = v * 10 + size_t(text[i++] - '0');
                shape.push_back(v);
            }
            return shape;
        }

    } // namespace detail

    // Writes `data` (C order, the product of `shape` elements) to `path`.
    template <typename T> void write(const std::string &path, const T *data, const std::vector<size_t> &shape) {
        size_t count = 1;
        for (size_t d : shape) count *= d;
        std::string header = std::string("{'descr': '") + descr<T>() + "', 'fortran_order': False, 'shape': " +
                             detail::shape_text(shape) + ", }";
        size_t prefix = 10; // magic, version, 2-byte length
        if (header.size() + 1 + 64 > 65535) prefix = 12;
        const size_t total = (prefix + header.size() + 1 + 63) / 64 * 64;
        header.append(total - prefix - header.size() - 1, ' ');
        header += '\n';
        detail::File file(path, "wb");
        unsigned char lead[12] = {0x93, 'N', 'U', 'M', 'P', 'Y', 0, 0, 0, 0, 0, 0};
        const uint32_t length = uint32_t(header.size());
        lead[6] = prefix == 10 ? 1 : 2;
        for (size_t i = 0; i < prefix - 8; ++i) lead[8 + i] = (unsigned char)(length >> (8 * i));
        if (std::fwrite(lead, 1, prefix, file.f) != prefix || std::fwrite(header.data(), 1, header.size(), file.f) != header.size() ||
            (count && std::fwrite(data, sizeof(T), count, file.f) != count))
            throw Error("cannot write " + path);
    }

    template <typename T> void write(const std::string &path, const std::vector<T> &data, const std::vector<size_t> &shape) {
        size_t count = 1;
        for (size_t d : shape) count *= d;
        if (count != data.size()) throw Error("the values do not fill the shape");
        write(path, data.data(), shape);
    }

    // Reads `path`, which must hold T elements ('<f4' or '<f8') in C order; the shape goes to `shape`.
    template <typename T> std::vector<T> read(const std::string &path, std::vector<size_t> &shape) {
        detail::File file(path, "rb");
        unsigned char lead[8];
        if (std::fread(lead, 1, 8, file.f) != 8 || std::memcmp(lead, "\x93NUMPY", 6) != 0)
            throw Error(path + " is not a .npy file");
        const int major = lead[6];
        if (major < 1 || major > 3) throw Error(path + ": unsupported .npy version " + std::to_string(major));
        const size_t width = major == 1 ? 2 : 4;
        unsigned char bytes[4] = {0, 0, 0, 0};
        if (std::fread(bytes, 1, width, file.f) != width) throw Error(path + ": truncated .npy header");
        uint32_t length = 0;
        for (size_t i = 0; i < width; ++i) length |= uint32_t(bytes[i]) << (8 * i);
        std::string header(length, '\0');
        if (length && std::fread(&header[0], 1, length, file.f) != length) throw Error(path + ": truncated .npy header");
        const std::string kind = detail::field(header, "descr");
        if (kind != std::string("'") + descr<T>() + "'")
            throw Error(path + ": elements are " + kind + ", not '" + descr<T>() + "'");
        if (detail::field(header, "fortran_order") != "False") throw Error(path + ": Fortran order is not supported");
        shape = detail::parse_shape(detail::field(header, "shape"));
        size_t count = 1;
        for (size_t d : shape) count *= d;
        std::vector<T> values(count);
        if (count && std::fread(values.data(), sizeof(T), count, file.f) != count) throw Error(path + ": truncated data");
        return values;
    }

} // namespace sci::npy
