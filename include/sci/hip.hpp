// sci/hip.hpp - a header-only C++17 host mirror of libsci's user API over the HIP backend's C ABI
// (scihip.h), for programs that use the MI355X backend without the reference library.
//
//  - Chunk records an expression with the reference's rules (include/core/bytecode/chunk.hpp:40-118):
//    inputs deduplicated by pointer, constants by bits, the lowest free temporary as target with the
//    sources released first, multiply-then-add contracted into fuse, output() retargeting the last
//    instruction.
//  - Device has the members of sci::core::drivers::Device (include/core/drivers/device.hpp:27-89)
//    plus the linear-algebra and adjoint entry points of the backend.
//  - Array<T> is a device array; Expr<T> a lazy expression over arrays and scalars that records one
//    Chunk when it is assigned, evaluated or reduced (include/array/expression.hpp:158-222).
//    Array<T>::save / load read and write .npy files (sci/npy.hpp).
//
// Errors throw sci::hip::Error (sci::hip::Declined when the backend cannot run the request as given).
#pragma once

#include "scihip.h"
#include "sci/npy.hpp"

#include <cstdint>
#include <cstring>
#include <functional>
#include <initializer_list>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace sci::hip {

    enum class Code : uint8_t {
        copy, negate, absolute, root, exponential, logarithm,
        sin, cos, tan, asin, acos, atan, csc, sec, cot, acsc, asec, acot,
        sinh, cosh, tanh, asinh, acosh, atanh, csch, sech, coth, acsch, asech, acoth,
        add, subtract, multiply, divide, minimum, maximum, fuse
    };
    enum class Fold : uint32_t { sum = SCIHIP_FOLD_SUM, minimum = SCIHIP_FOLD_MINIMUM, maximum = SCIHIP_FOLD_MAXIMUM };
    enum class Precision : uint8_t { single = SCIHIP_SINGLE, wide = SCIHIP_WIDE };
    enum class Mode : uint32_t {
        float32 = SCIHIP_PRODUCT_FLOAT32, bfloat16 = SCIHIP_PRODUCT_BFLOAT16,
        split = SCIHIP_PRODUCT_SPLIT, float64 = SCIHIP_PRODUCT_FLOAT64
    };
    enum class Factoring : uint32_t {
        float64 = SCIHIP_FACTOR_FLOAT64, float32 = SCIHIP_FACTOR_FLOAT32,
        split = SCIHIP_FACTOR_SPLIT, bfloat16 = SCIHIP_FACTOR_BFLOAT16
    };

    class Error : public std::runtime_error {
    public:
        Error(int status, const std::string &message) : std::runtime_error(message), status(status) {}
        int status;
    };

    class Declined : public Error {
    public:
        using Error::Error;
    };

    inline void check(int status, const char *what) {
        if (status == SCIHIP_OK) return;
        const std::string message = std::string(what) + ": " + scihip_error();
        if (status == SCIHIP_DECLINED) throw Declined(status, message);
        throw Error(status, message);
    }

    // ---------------------------------------------------------------------------------- Chunk

    class Chunk {
    public:
        static constexpr uint8_t none = SCIHIP_SLOT_NONE;

        explicit Chunk(Precision precision = Precision::single) {
            std::memset(&chunk, 0, sizeof(chunk));
            chunk.precision = static_cast<uint8_t>(precision);
            chunk.result = none;
        }

        uint8_t input(const void *pointer) {
            for (uint8_t i = 0; i < chunk.arguments; ++i)
                if (chunk.inputs[i] == pointer) return static_cast<uint8_t>(SCIHIP_SLOT_INPUT | i);
            if (chunk.arguments == SCIHIP_CAPACITY) return fail();
            chunk.inputs[chunk.arguments] = pointer;
            return static_cast<uint8_t>(SCIHIP_SLOT_INPUT | chunk.arguments++);
        }

        uint8_t constant(double value) {
            for (uint8_t i = 0; i < chunk.literals; ++i)
                if (std::memcmp(&chunk.constants[i], &value, sizeof(value)) == 0)
                    return static_cast<uint8_t>(SCIHIP_SLOT_LITERAL | i);
            if (chunk.literals == SCIHIP_CAPACITY) return fail();
            chunk.constants[chunk.literals] = value;
            return static_cast<uint8_t>(SCIHIP_SLOT_LITERAL | chunk.literals++);
        }

        uint8_t emit(Code code, uint8_t first, uint8_t second = none, uint8_t third = none) {
            if (overflow) return none;
            if (code == Code::add) { // a product recorded just before an add becomes one fuse
                if (const uint8_t merged = contract(second, first); merged != none) return merged;
                if (const uint8_t merged = contract(first, second); merged != none) return merged;
            }
            if (chunk.length == SCIHIP_CAPACITY) return fail();
            release(first);
            release(second);
            release(third);
            if (busy == ~uint64_t{0}) return fail();
            uint8_t index = 0;
            while ((busy >> index) & 1u) ++index;
            busy |= uint64_t{1} << index;
            if (index + 1 > chunk.temporaries) chunk.temporaries = static_cast<uint8_t>(index + 1);
            const uint8_t target = static_cast<uint8_t>(SCIHIP_SLOT_TEMPORARY | index);
            chunk.instructions[chunk.length++] = {static_cast<uint8_t>(code), target, {first, second, third}};
            return target;
        }

        void output(uint8_t value) {
            if (overflow) return;
            scihip_instruction *last = chunk.length ? &chunk.instructions[chunk.length - 1] : nullptr;
            if (last && last->target == value && (value & SCIHIP_SLOT_TYPES) == SCIHIP_SLOT_TEMPORARY) {
                last->target = SCIHIP_SLOT_OUTPUT;
            } else if (chunk.length < SCIHIP_CAPACITY) {
                chunk.instructions[chunk.length++] = {static_cast<uint8_t>(Code::copy), SCIHIP_SLOT_OUTPUT,
                                                      {value, none, none}};
            } else {
                fail();
            }
            chunk.result = SCIHIP_SLOT_OUTPUT;
        }

        void reduce(uint8_t value) { chunk.result = value; }
        bool valid() const { return !overflow; }
        const scihip_chunk &data() const { return chunk; }

    private:
        uint8_t fail() {
            overflow = true;
            return none;
        }
        void release(uint8_t value) {
            if ((value & SCIHIP_SLOT_TYPES) == SCIHIP_SLOT_TEMPORARY && value != none)
                busy &= ~(uint64_t{1} << (value & 0x3F));
        }
        uint8_t contract(uint8_t product, uint8_t addend) {
            if (chunk.length == 0 || (product & SCIHIP_SLOT_TYPES) != SCIHIP_SLOT_TEMPORARY || product == none)
                return none;
            scihip_instruction &last = chunk.instructions[chunk.length - 1];
            if (last.code != static_cast<uint8_t>(Code::multiply) || last.target != product) return none;
            last.code = static_cast<uint8_t>(Code::fuse);
            last.sources[2] = addend;
            release(addend);
            return product;
        }

        scihip_chunk chunk;
        uint64_t busy = 0;
        bool overflow = false;
    };

    // ---------------------------------------------------------------------------------- Device

    class Device {
    public:
        explicit Device(uint32_t index = 0) { check(scihip_open(index, &handle), "scihip_open"); }
        ~Device() { scihip_close(handle); }
        Device(const Device &) = delete;
        Device &operator=(const Device &) = delete;

        const char *name() const noexcept { return scihip_name(handle); }
        void synchronize() { check(scihip_synchronize(handle), "scihip_synchronize"); }

        void *allocate(size_t bytes) {
            void *pointer = nullptr;
            check(scihip_allocate(handle, bytes, &pointer), "scihip_allocate");
            return pointer;
        }
        void dispose(void *pointer) noexcept { (void)scihip_dispose(handle, pointer); }
        void upload(void *device_pointer, const void *host, size_t bytes) {
            check(scihip_upload(handle, device_pointer, host, bytes), "scihip_upload");
        }
        void download(void *host, const void *device_pointer, size_t bytes) {
            check(scihip_download(handle, host, device_pointer, bytes), "scihip_download");
        }

        void execute(const Chunk &chunk, void *output, size_t count) {
            check(scihip_execute(handle, &chunk.data(), output, count), "scihip_execute");
        }
        double reduce(const Chunk &chunk, Fold fold, size_t count) {
            double result = 0.0;
            check(scihip_reduce(handle, &chunk.data(), static_cast<uint32_t>(fold), count, &result), "scihip_reduce");
            return result;
        }
        void product(Mode mode, const void *left, const void *right, void *output, size_t rows, size_t columns,
                     size_t depth) {
            check(scihip_product(handle, static_cast<uint32_t>(mode), left, right, output, rows, columns, depth),
                  "scihip_product");
        }
        std::vector<int64_t> factor(Factoring mode, void *matrix, size_t order) {
            std::vector<int64_t> rows(order);
            check(scihip_factor(handle, static_cast<uint32_t>(mode), matrix, order, rows.data()), "scihip_factor");
            return rows;
        }
        scihip_solve_stats solve(Factoring mode, const double *matrix, const double *right, double *solution,
                                 size_t order, size_t width, double tolerance = 1e-14, int32_t iterations = 30) {
            scihip_solve_stats stats{};
            check(scihip_solve(handle, static_cast<uint32_t>(mode), matrix, right, solution, order, width, tolerance,
                               iterations, &stats),
                  "scihip_solve");
            return stats;
        }
        double adjoint(const Chunk &chunk, const void *seed, void *const *gradients, size_t count, bool loss = true) {
            double total = 0.0;
            check(scihip_adjoint(handle, &chunk.data(), seed, gradients, count, loss ? &total : nullptr),
                  "scihip_adjoint");
            return total;
        }
        void product_adjoint(Mode mode, const void *left, const void *right, const void *seed, void *left_gradient,
                             void *right_gradient, size_t rows, size_t columns, size_t depth) {
            check(scihip_product_adjoint(handle, static_cast<uint32_t>(mode), left, right, seed, left_gradient,
                                         right_gradient, rows, columns, depth),
                  "scihip_product_adjoint");
        }
        scihip_device *raw() const noexcept { return handle; }

    private:
        scihip_device *handle = nullptr;
    };

    // ---------------------------------------------------------------------------------- Array

    template <typename T> struct precision_of;
    template <> struct precision_of<float> {
        static constexpr Precision value = Precision::single;
    };
    template <> struct precision_of<double> {
        static constexpr Precision value = Precision::wide;
    };

    // A row-major device array. Copies share the allocation (like views); it is disposed with the last.
    //
    // Element access (a[i], a.at(i, j)) reads and writes a host copy of the elements, since HBM is not
    // mapped into the host as the reference's device memory is (include/array/array.hpp:186-217,
    // include/array/storage.hpp:100-102). The copy is downloaded at the first access after device work
    // (which the download waits for), shared by the array's copies, and handed back at the next device
    // use: data(), the pointer every device operation takes, uploads it first if an element was written
    // and drops it, since the operation may write the array.
    template <typename T> class Array {
    public:
        static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>, "float or double elements");

        Array(Device &device, std::vector<size_t> shape)
            : owner(&device), dims(std::move(shape)), mirror(std::make_shared<Mirror>()) {
            size_t n = 1;
            for (size_t d : dims) n *= d;
            count = n;
            void *pointer = device.allocate(n * sizeof(T));
            Device *d = &device;
            storage = std::shared_ptr<void>(pointer, [d](void *p) { d->dispose(p); });
        }

        static Array from(Device &device, const std::vector<T> &values, std::vector<size_t> shape = {}) {
            if (shape.empty()) shape = {values.size()};
            Array out(device, std::move(shape));
            if (out.size() != values.size()) throw Error(SCIHIP_EINVAL, "values do not fill the shape");
            device.upload(out.data(), values.data(), values.size() * sizeof(T));
            return out;
        }

        std::vector<T> host() const {
            std::vector<T> values(count);
            owner->download(values.data(), data(), count * sizeof(T));
            return values;
        }

        // a[i]: an element by its flat index, through the host copy (written back at the next device use).
        T &operator[](size_t index) {
            T &value = values().at(index);
            mirror->dirty = true;
            return value;
        }
        const T &operator[](size_t index) const { return values().at(index); }

        // a.at(i, j, ...): an element by its index along each dimension (row-major).
        template <typename... Indices> T &at(Indices... indices) {
            T &value = values().at(offset({static_cast<size_t>(indices)...}));
            mirror->dirty = true;
            return value;
        }
        template <typename... Indices> const T &at(Indices... indices) const {
            return values().at(offset({static_cast<size_t>(indices)...}));
        }

        // Writes the array to a .npy file (its shape and elements; see sci/npy.hpp).
        void save(const std::string &path) const { npy::write(path, host(), dims); }

        // An array on `device` holding a .npy file's elements, which must be of type T.
        static Array load(Device &device, const std::string &path) {
            std::vector<size_t> shape;
            const std::vector<T> values = npy::read<T>(path, shape);
            if (shape.empty()) shape = {1};
            return from(device, values, shape);
        }

        // The device pointer, for device work: elements written through the host copy go up first, and
        // the copy is dropped (the work may write the array).
        T *data() const {
            if (mirror->valid) {
                if (mirror->dirty) owner->upload(storage.get(), mirror->values.data(), count * sizeof(T));
                mirror->values = std::vector<T>();
                mirror->valid = mirror->dirty = false;
            }
            return static_cast<T *>(storage.get());
        }
        size_t size() const noexcept { return count; }
        const std::vector<size_t> &shape() const noexcept { return dims; }
        Device &device() const noexcept { return *owner; }

    private:
        struct Mirror {
            std::vector<T> values;
            bool valid = false; // values hold the elements as of the last device use
            bool dirty = false; // an element was written since
        };

        std::vector<T> &values() const {
            if (!mirror->valid) {
                mirror->values.resize(count);
                owner->download(mirror->values.data(), storage.get(), count * sizeof(T));
                mirror->valid = true;
                mirror->dirty = false;
            }
            return mirror->values;
        }

        size_t offset(std::initializer_list<size_t> indices) const {
            if (indices.size() != dims.size()) throw Error(SCIHIP_EINVAL, "one index per dimension");
            size_t flat = 0, k = 0;
            for (size_t index : indices) {
                if (index >= dims[k]) throw Error(SCIHIP_EINVAL, "index out of range");
                flat = flat * dims[k++] + index;
            }
            return flat;
        }

        Device *owner;
        std::vector<size_t> dims;
        size_t count = 0;
        std::shared_ptr<void> storage;
        std::shared_ptr<Mirror> mirror;
    };

    // ---------------------------------------------------------------------------------- expressions

    template <typename T> class Expr {
    public:
        Expr(const Array<T> &array) : term(std::make_shared<Term>(Term{Kind::leaf, Code::copy, 0.0, array, {}, 0})) {}
        Expr(T value) : term(std::make_shared<Term>(Term{Kind::literal, Code::copy, double(value), {}, {}, 0})) {}
        Expr(Code code, std::vector<Expr> children)
            : term(std::make_shared<Term>(Term{Kind::operation, code, 0.0, {}, std::move(children), 0})) {}

        // A leaf standing for the result of stage `index` of a split expression (see split below),
        // shaped like `like`. While the split is planned it records as an input of its own.
        static Expr staged(size_t index, const Array<T> &like) {
            Expr e(like);
            e.term = std::make_shared<Term>(Term{Kind::staged, Code::copy, 0.0, like, {}, index});
            return e;
        }

        // Records operands first, left to right, then the operation (include/array/node.hpp:75-88).
        uint8_t record(Chunk &chunk) const {
            switch (term->kind) {
            case Kind::leaf: return chunk.input(term->array->data());
            case Kind::staged: return chunk.input(reinterpret_cast<const void *>(uintptr_t(0x10 + 16 * term->index)));
            case Kind::literal: return chunk.constant(term->value);
            default: {
                uint8_t slots[3] = {Chunk::none, Chunk::none, Chunk::none};
                for (size_t k = 0; k < term->children.size(); ++k) slots[k] = term->children[k].record(chunk);
                return chunk.emit(term->code, slots[0], slots[1], slots[2]);
            }
            }
        }

        // The first array operand: its device, shape and count stand for the expression's.
        const Array<T> *anchor() const {
            if (term->kind == Kind::leaf || term->kind == Kind::staged) return &*term->array;
            for (const Expr &child : term->children)
                if (const Array<T> *found = child.anchor()) return found;
            return nullptr;
        }

        // The chunk as recorded, valid() false when the expression outgrew it.
        Chunk recorded(bool reduce) const {
            Chunk c(precision_of<T>::value);
            const uint8_t value = record(c);
            if (reduce) c.reduce(value);
            else c.output(value);
            return c;
        }

        Chunk chunk(bool reduce) const {
            Chunk c = recorded(reduce);
            if (!c.valid()) throw Error(SCIHIP_EINVAL, "the expression outgrew a chunk (64 instructions, inputs or constants)");
            return c;
        }

        bool operation() const noexcept { return term->kind == Kind::operation; }
        bool is_staged() const noexcept { return term->kind == Kind::staged; }
        size_t stage() const noexcept { return term->index; }
        Code code() const noexcept { return term->code; }
        const std::vector<Expr> &children() const noexcept { return term->children; }
        size_t operations() const {
            size_t n = operation() ? 1 : 0;
            for (const Expr &child : term->children) n += child.operations();
            return n;
        }

    private:
        enum class Kind { leaf, literal, operation, staged };
        struct Term {
            Kind kind;
            Code code;
            double value;
            std::optional<Array<T>> array;
            std::vector<Expr> children;
            size_t index;
        };
        std::shared_ptr<const Term> term;
    };

    // An expression that outgrows one chunk (64 instructions, inputs, constants or temporaries,
    // include/core/bytecode/chunk.hpp:31) split into chunks run one after another on the device, each
    // later one reading earlier results from device temporaries; the reference runs such an
    // expression element by element on the host instead (include/array/expression.hpp:178-184,
    // 204-221). stages[k] is a store whose staged leaves name earlier stages; `last` is stored or
    // reduced. The rule is libsci_amd.array.stages's: a node that fits one chunk stays whole;
    // otherwise each child is made to fit on its own, then, while the node still outgrows a chunk, its
    // operation child with the most operations (the leftmost of equals) becomes a stage of its own.
    template <typename T> struct Split {
        std::vector<Expr<T>> stages;
        Expr<T> last;
    };

    template <typename T> Expr<T> cut(const Expr<T> &node, bool reduce, std::vector<Expr<T>> &stages) {
        if (!node.operation() || node.recorded(reduce).valid()) return node;
        std::vector<Expr<T>> children;
        for (const Expr<T> &child : node.children()) children.push_back(cut(child, false, stages));
        for (;;) {
            Expr<T> candidate(node.code(), children);
            if (candidate.recorded(reduce).valid()) return candidate;
            size_t pick = children.size();
            for (size_t k = 0; k < children.size(); ++k)
                if (children[k].operation() &&
                    (pick == children.size() || children[k].operations() > children[pick].operations()))
                    pick = k;
            stages.push_back(children[pick]);
            children[pick] = Expr<T>::staged(stages.size() - 1, *children[pick].anchor());
        }
    }

    template <typename T> Split<T> split(const Expr<T> &e, bool reduce) {
        std::vector<Expr<T>> stages;
        Expr<T> last = cut(e, reduce, stages);
        return Split<T>{std::move(stages), std::move(last)};
    }

    // The expression with each staged leaf replaced by results[stage].
    template <typename T> Expr<T> bind_stages(const Expr<T> &e, const std::vector<Array<T>> &results) {
        if (e.is_staged()) return Expr<T>(results.at(e.stage()));
        if (!e.operation()) return e;
        std::vector<Expr<T>> children;
        for (const Expr<T> &child : e.children()) children.push_back(bind_stages(child, results));
        return Expr<T>(e.code(), std::move(children));
    }

    template <typename T> const Array<T> &anchor_of(const Expr<T> &e) {
        const Array<T> *found = e.anchor();
        if (!found) throw Error(SCIHIP_EINVAL, "an expression needs an array operand");
        return *found;
    }

#define SCI_HIP_BINARY(op, code)                                                                                      \
    template <typename T> Expr<T> operator op(const Expr<T> &l, const Expr<T> &r) { return Expr<T>(code, {l, r}); } \
    template <typename T> Expr<T> operator op(const Array<T> &l, const Array<T> &r) {                                \
        return Expr<T>(code, {Expr<T>(l), Expr<T>(r)});                                                              \
    }                                                                                                                \
    template <typename T> Expr<T> operator op(const Expr<T> &l, const Array<T> &r) { return Expr<T>(code, {l, Expr<T>(r)}); } \
    template <typename T> Expr<T> operator op(const Array<T> &l, const Expr<T> &r) { return Expr<T>(code, {Expr<T>(l), r}); } \
    template <typename T> Expr<T> operator op(const Expr<T> &l, T r) { return Expr<T>(code, {l, Expr<T>(r)}); }      \
    template <typename T> Expr<T> operator op(T l, const Expr<T> &r) { return Expr<T>(code, {Expr<T>(l), r}); }      \
    template <typename T> Expr<T> operator op(const Array<T> &l, T r) { return Expr<T>(code, {Expr<T>(l), Expr<T>(r)}); } \
    template <typename T> Expr<T> operator op(T l, const Array<T> &r) { return Expr<T>(code, {Expr<T>(l), Expr<T>(r)}); }

    SCI_HIP_BINARY(+, Code::add)
    SCI_HIP_BINARY(-, Code::subtract)
    SCI_HIP_BINARY(*, Code::multiply)
    SCI_HIP_BINARY(/, Code::divide)
#undef SCI_HIP_BINARY

    template <typename T> Expr<T> operator-(const Expr<T> &x) { return Expr<T>(Code::negate, {x}); }
    template <typename T> Expr<T> operator-(const Array<T> &x) { return Expr<T>(Code::negate, {Expr<T>(x)}); }

#define SCI_HIP_UNARY(name)                                                                                           \
    template <typename T> Expr<T> name(const Expr<T> &x) { return Expr<T>(Code::name, {x}); }                        \
    template <typename T> Expr<T> name(const Array<T> &x) { return Expr<T>(Code::name, {Expr<T>(x)}); }

    SCI_HIP_UNARY(absolute) SCI_HIP_UNARY(root) SCI_HIP_UNARY(exponential) SCI_HIP_UNARY(logarithm)
    SCI_HIP_UNARY(sin) SCI_HIP_UNARY(cos) SCI_HIP_UNARY(tan) SCI_HIP_UNARY(asin) SCI_HIP_UNARY(acos)
    SCI_HIP_UNARY(atan) SCI_HIP_UNARY(csc) SCI_HIP_UNARY(sec) SCI_HIP_UNARY(cot) SCI_HIP_UNARY(acsc)
    SCI_HIP_UNARY(asec) SCI_HIP_UNARY(acot) SCI_HIP_UNARY(sinh) SCI_HIP_UNARY(cosh) SCI_HIP_UNARY(tanh)
    SCI_HIP_UNARY(asinh) SCI_HIP_UNARY(acosh) SCI_HIP_UNARY(atanh) SCI_HIP_UNARY(csch) SCI_HIP_UNARY(sech)
    SCI_HIP_UNARY(coth) SCI_HIP_UNARY(acsch) SCI_HIP_UNARY(asech) SCI_HIP_UNARY(acoth)
#undef SCI_HIP_UNARY

    // Elementwise minimum / maximum (include/operations/binary.hpp:60-66).
    template <typename T> Expr<T> minimum(const Expr<T> &l, const Expr<T> &r) { return Expr<T>(Code::minimum, {l, r}); }
    template <typename T> Expr<T> maximum(const Expr<T> &l, const Expr<T> &r) { return Expr<T>(Code::maximum, {l, r}); }

    template <typename T> Array<T> evaluate(const Expr<T> &e);

    // The expression as one chunk can run it: unchanged when it fits, else its earlier stages are
    // queued on the device (asynchronously, each into a temporary) and the last stage returned.
    template <typename T> Expr<T> runnable(const Expr<T> &e, bool reduce) {
        if (!e.operation() || e.recorded(reduce).valid()) return e;
        const Split<T> parts = split(e, reduce);
        std::vector<Array<T>> results;
        for (const Expr<T> &stage : parts.stages) results.push_back(evaluate(bind_stages(stage, results)));
        return bind_stages(parts.last, results);
    }

    // Runs the expression into `target` (include/array/expression.hpp:158-185).
    template <typename T> void assign(Array<T> &target, const Expr<T> &e) {
        const Expr<T> whole = runnable(e, false); // holds the stages' temporaries until the call is queued
        target.device().execute(whole.chunk(false), target.data(), target.size());
    }

    template <typename T> Array<T> evaluate(const Expr<T> &e) {
        const Array<T> &a = anchor_of(e);
        Array<T> out(a.device(), a.shape());
        assign(out, e);
        return out;
    }

    // The fold in double, before the cast to the element type.
    template <typename T> double reduce_raw(const Expr<T> &e, Fold fold) {
        const Array<T> &a = anchor_of(e);
        const Expr<T> whole = runnable(e, true); // holds the stages' temporaries until the call is queued
        return a.device().reduce(whole.chunk(true), fold, a.size());
    }

    template <typename T> T sum(const Expr<T> &e) { return static_cast<T>(reduce_raw(e, Fold::sum)); }
    template <typename T> T sum(const Array<T> &a) { return sum(Expr<T>(a)); }
    template <typename T> T minimum(const Expr<T> &e) { return static_cast<T>(reduce_raw(e, Fold::minimum)); }
    template <typename T> T maximum(const Expr<T> &e) { return static_cast<T>(reduce_raw(e, Fold::maximum)); }

    // float(sum) / float(count) in the element type (include/operations/reduce.hpp:24-28).
    template <typename T> T mean(const Expr<T> &e) {
        const size_t n = anchor_of(e).size();
        return n == 0 ? T(0) : static_cast<T>(sum(e) / static_cast<T>(n));
    }

    // sum(a * b), contracted into a dot product (include/linear/product.hpp:72-76).
    template <typename T> T dot(const Array<T> &a, const Array<T> &b) { return sum(a * b); }

    // left (rows x depth) @ right (depth x columns) on the matrix cores.
    template <typename T> Array<T> product(const Array<T> &left, const Array<T> &right, Mode mode) {
        const size_t rows = left.shape().at(0), depth = left.shape().at(1), columns = right.shape().at(1);
        if (right.shape().at(0) != depth) throw Error(SCIHIP_EINVAL, "product: the shapes do not conform");
        Array<T> out(left.device(), {rows, columns});
        left.device().product(mode, left.data(), right.data(), out.data(), rows, columns, depth);
        return out;
    }

    // The gradient of sum(e) with respect to each of `arrays` (zeros for one e does not use), forward
    // and backward fused in one pass; the loss is returned through `loss` when it is non-null.
    template <typename T>
    std::vector<Array<T>> gradient(const Expr<T> &e, const std::vector<Array<T>> &arrays, double *loss = nullptr) {
        const Array<T> &a = anchor_of(e);
        const Chunk c = e.chunk(true);
        const scihip_chunk &raw = c.data();
        std::vector<Array<T>> out;
        std::vector<void *> table(raw.arguments, nullptr);
        for (const Array<T> &x : arrays) {
            int input = -1;
            for (uint8_t i = 0; i < raw.arguments; ++i)
                if (raw.inputs[i] == x.data()) input = i;
            if (input >= 0 && table[size_t(input)]) { // the same array asked for twice
                for (size_t k = 0; k < out.size(); ++k)
                    if (out[k].data() == table[size_t(input)]) out.push_back(out[k]);
                continue;
            }
            Array<T> g(a.device(), a.shape());
            if (input >= 0) table[size_t(input)] = g.data();
            else assign(g, Expr<T>(T(0))); // an array the expression does not read: zeros
            out.push_back(g);
        }
        const double total = a.device().adjoint(c, nullptr, table.data(), a.size(), loss != nullptr);
        if (loss) *loss = total;
        return out;
    }

} // namespace sci::hip
