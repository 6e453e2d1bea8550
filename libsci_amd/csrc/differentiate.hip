// Reverse-mode adjoints of recorded chunks and of the matrix product: the scihip_adjoint* and
// scihip_product_adjoint entry points.
//
// A chunk computes y = f(x_0, ..., x_{n-1}) element by element. Its adjoint program recomputes the
// forward values and then walks the instructions backwards, adding each instruction's local
// derivative times the adjoint of its result into the adjoints of its operands (the chain rule).
// The adjoint of the result is 1 (the gradient of the loss, the sum of y) or an upstream gradient
// array (a vector-Jacobian product). Local derivatives are written with the chunk codes themselves
// (d sin x = cos x, d tanh x = 1 - tanh^2 x, ...) plus three codes that only adjoints use (a * sign b
// and the minimum / maximum routing). The program then loses every value that no gradient needs,
// and its values are packed into the fewest registers by a linear scan over their last uses.
#include "adjoint.hpp"
#include "compiler.hpp"
#include "differentiate.hpp"
#include "kernels.hpp"
#include "product.hpp"
#include "runtime.hpp"

#include <algorithm>
#include <cstring>
#include <limits>
#include <map>
#include <sstream>

namespace scihip {

    namespace {

        enum Kind : uint8_t { kInputValue, kLiteralValue, kTemporaryValue, kAbsent };

        struct Val {
            Kind kind = kAbsent;
            uint32_t index = 0;
            bool present() const { return kind != kAbsent; }
            uint64_t key() const { return uint64_t(kind) << 32 | index; }
        };

        struct Step {
            uint32_t code;
            Val a, b, c;
        };

        // An adjoint: zero (never touched), exactly one, or a value.
        struct Adj {
            int state = 0;
            Val v;
        };

        constexpr uint32_t kSkipped = 0xFFFFFFFFu; // the result of a contracted chunk's final multiply

        class Emitter {
        public:
            explicit Emitter(bool single) : single(single) {}

            std::vector<Step> steps;
            std::vector<double> constants;

            Val literal(double x) {
                if (single) x = double(float(x));
                for (size_t i = 0; i < constants.size(); ++i)
                    if (std::memcmp(&constants[i], &x, sizeof(x)) == 0) return {kLiteralValue, uint32_t(i)};
                constants.push_back(x);
                return {kLiteralValue, uint32_t(constants.size() - 1)};
            }

            Val emit(uint32_t code, Val a, Val b = {}, Val c = {}) {
                steps.push_back({code, a, b, c});
                return {kTemporaryValue, uint32_t(steps.size() - 1)};
            }

            Val value(const Adj &g) { return g.state == 1 ? literal(1.0) : g.v; }

            // x += g
            void add(Adj *x, const Adj &g) {
                if (!x || g.state == 0) return;
                if (x->state == 0) *x = g;
                else *x = {2, emit(kAdd, value(*x), value(g))};
            }
            // x -= g
            void subtract(Adj *x, const Adj &g) {
                if (!x || g.state == 0) return;
                if (x->state == 0) *x = {2, g.state == 1 ? literal(-1.0) : emit(kNegate, g.v)};
                else *x = {2, emit(kSubtract, value(*x), value(g))};
            }
            // x += g * f
            void add_scaled(Adj *x, const Adj &g, Val f) {
                if (!x) return;
                if (g.state == 1) add(x, {2, f});
                else if (x->state == 0) *x = {2, emit(kMultiply, g.v, f)};
                else *x = {2, emit(kFuse, g.v, f, value(*x))};
            }
            // x -= g * f
            void subtract_scaled(Adj *x, const Adj &g, Val f) {
                if (!x) return;
                subtract(x, {2, g.state == 1 ? f : emit(kMultiply, g.v, f)});
            }
            // x += g / f and x -= g / f
            void add_quotient(Adj *x, const Adj &g, Val f) {
                if (x) add(x, {2, emit(kDivide, value(g), f)});
            }
            void subtract_quotient(Adj *x, const Adj &g, Val f) {
                if (x) subtract(x, {2, emit(kDivide, value(g), f)});
            }

        private:
            bool single;
        };

        // The backward step of one instruction y = code(a, b, c) with result adjoint g: adds the
        // instruction's contribution to the adjoints da, db, dc (null: the operand is a constant).
        void backward(Emitter &e, uint32_t code, Val a, Val b, Val y, const Adj &g, Adj *da, Adj *db, Adj *dc) {
            auto one = [&] { return e.literal(1.0); };
            auto square = [&](Val x) { return e.emit(kMultiply, x, x); };
            switch (code) {
            case kCopy: e.add(da, g); break;
            case kNegate: e.subtract(da, g); break;
            case kAbsolute:
                if (da) e.add(da, {2, e.emit(kSignTimes, e.value(g), a)});
                break;
            case kRoot: // d sqrt a = 0.5 / y
                if (da) e.add_scaled(da, g, e.emit(kDivide, e.literal(0.5), y));
                break;
            case kExponential: e.add_scaled(da, g, y); break;
            case kLogarithm: e.add_quotient(da, g, a); break;
            case kSin:
                if (da) e.add_scaled(da, g, e.emit(kCos, a));
                break;
            case kCos:
                if (da) e.subtract_scaled(da, g, e.emit(kSin, a));
                break;
            case kTan: // 1 + y^2
                if (da) e.add_scaled(da, g, e.emit(kFuse, y, y, one()));
                break;
            case kAsin: // 1 / sqrt(1 - a^2)
            case kAcos:
                if (da) {
                    const Val t = e.emit(kRoot, e.emit(kSubtract, one(), square(a)));
                    if (code == kAsin) e.add_quotient(da, g, t);
                    else e.subtract_quotient(da, g, t);
                }
                break;
            case kAtan: // 1 / (1 + a^2)
            case kAcot:
                if (da) {
                    const Val t = e.emit(kFuse, a, a, one());
                    if (code == kAtan) e.add_quotient(da, g, t);
                    else e.subtract_quotient(da, g, t);
                }
                break;
            case kCsc: // -y cot a
                if (da) e.subtract_scaled(da, g, e.emit(kMultiply, y, e.emit(kCot, a)));
                break;
            case kSec: // y tan a
                if (da) e.add_scaled(da, g, e.emit(kMultiply, y, e.emit(kTan, a)));
                break;
            case kCot: // -(1 + y^2)
                if (da) e.subtract_scaled(da, g, e.emit(kFuse, y, y, one()));
                break;
            case kAcsc: // -1 / (|a| sqrt(a^2 - 1)); asec the opposite sign
            case kAsec:
                if (da) {
                    const Val r = e.emit(kRoot, e.emit(kSubtract, square(a), one()));
                    const Val t = e.emit(kMultiply, e.emit(kAbsolute, a), r);
                    if (code == kAsec) e.add_quotient(da, g, t);
                    else e.subtract_quotient(da, g, t);
                }
                break;
            case kSinh:
                if (da) e.add_scaled(da, g, e.emit(kCosh, a));
                break;
            case kCosh:
                if (da) e.add_scaled(da, g, e.emit(kSinh, a));
                break;
            case kTanh: // 1 - y^2, also coth
            case kCoth:
                if (da) e.add_scaled(da, g, e.emit(kSubtract, one(), square(y)));
                break;
            case kAsinh: // 1 / sqrt(a^2 + 1)
                if (da) e.add_quotient(da, g, e.emit(kRoot, e.emit(kFuse, a, a, one())));
                break;
            case kAcosh: // 1 / sqrt(a^2 - 1)
                if (da) e.add_quotient(da, g, e.emit(kRoot, e.emit(kSubtract, square(a), one())));
                break;
            case kAtanh: // 1 / (1 - a^2), also acoth
            case kAcoth:
                if (da) e.add_quotient(da, g, e.emit(kSubtract, one(), square(a)));
                break;
            case kCsch: // -y coth a
                if (da) e.subtract_scaled(da, g, e.emit(kMultiply, y, e.emit(kCoth, a)));
                break;
            case kSech: // -y tanh a
                if (da) e.subtract_scaled(da, g, e.emit(kMultiply, y, e.emit(kTanh, a)));
                break;
            case kAcsch: // -1 / (|a| sqrt(1 + a^2))
                if (da) {
                    const Val r = e.emit(kRoot, e.emit(kFuse, a, a, one()));
                    e.subtract_quotient(da, g, e.emit(kMultiply, e.emit(kAbsolute, a), r));
                }
                break;
            case kAsech: // -1 / (a sqrt(1 - a^2))
                if (da) {
                    const Val r = e.emit(kRoot, e.emit(kSubtract, one(), square(a)));
                    e.subtract_quotient(da, g, e.emit(kMultiply, a, r));
                }
                break;
            case kAdd:
                e.add(da, g);
                e.add(db, g);
                break;
            case kSubtract:
                e.add(da, g);
                e.subtract(db, g);
                break;
            case kMultiply:
                e.add_scaled(da, g, b);
                e.add_scaled(db, g, a);
                break;
            case kDivide: // da += g / b, db -= (g / b) * y
                if (da || db) {
                    const Val t = e.emit(kDivide, e.value(g), b);
                    e.add(da, {2, t});
                    if (db) e.subtract(db, {2, e.emit(kMultiply, t, y)});
                }
                break;
            case kMinimum:
            case kMaximum: {
                const uint32_t route = code == kMinimum ? kRouteMinimum : kRouteMaximum;
                if (da) e.add(da, {2, e.emit(route, e.value(g), a, b)});
                if (db) e.add(db, {2, e.emit(route, e.value(g), b, a)});
                break;
            }
            default: // fuse: a * b + c
                e.add_scaled(da, g, b);
                e.add_scaled(db, g, a);
                e.add(dc, g);
                break;
            }
        }

        int arity(uint32_t code) { return code == kFuse ? 3 : code >= kAdd ? 2 : 1; }

    } // namespace

    int plan_adjoint(const scihip_chunk &chunk, bool seeded, bool loss, AdjointPlan &plan, std::string &error) {
        const bool single = chunk.precision == SCIHIP_SINGLE;
        Emitter e(single);
        // The chunk's constants keep their indices (narrowed to the element type, as for execute).
        for (int i = 0; i < chunk.literals; ++i) e.constants.push_back(single ? double(float(chunk.constants[i]))
                                                                              : chunk.constants[i]);
        const bool contract = loss && !seeded && contracted(chunk, SCIHIP_FOLD_SUM);

        // Forward pass: the chunk's instructions over single-assignment values.
        struct Forward {
            uint32_t code;
            Val a, b, c, y;
        };
        std::vector<Forward> forward;
        Val temporaries[SCIHIP_CAPACITY];
        Val output;
        auto value_of = [&](uint8_t id) -> Val {
            const uint8_t tag = id & SCIHIP_SLOT_TYPES, index = id & 0x3F;
            if (tag == SCIHIP_SLOT_INPUT) return {kInputValue, index};
            if (tag == SCIHIP_SLOT_LITERAL) return {kLiteralValue, index};
            if (tag == SCIHIP_SLOT_TEMPORARY) return temporaries[index];
            return output;
        };
        for (int s = 0; s < chunk.length; ++s) {
            const scihip_instruction &in = chunk.instructions[s];
            const int n = arity(in.code);
            const Val a = value_of(in.sources[0]);
            const Val b = n > 1 ? value_of(in.sources[1]) : Val{};
            const Val c = n > 2 ? value_of(in.sources[2]) : Val{};
            const bool last = s == chunk.length - 1;
            const Val y = (contract && last) ? Val{kTemporaryValue, kSkipped} : e.emit(in.code, a, b, c);
            forward.push_back({in.code, a, b, c, y});
            if (in.target == SCIHIP_SLOT_OUTPUT) output = y;
            else temporaries[in.target & 0x3F] = y;
        }
        const Val result = value_of(chunk.result);
        const uint32_t inputs = chunk.arguments + (seeded ? 1u : 0u);

        // Backward pass, in reverse order; every consumer of a value precedes it here, so its adjoint
        // is complete when its own instruction is reached.
        std::map<uint64_t, Adj> adjoint;
        adjoint[result.key()] = seeded ? Adj{2, Val{kInputValue, chunk.arguments}} : Adj{1, {}};
        auto slot_of = [&](Val v) -> Adj * { return v.kind == kLiteralValue || !v.present() ? nullptr : &adjoint[v.key()]; };
        for (int s = chunk.length - 1; s >= 0; --s) {
            const Forward &f = forward[size_t(s)];
            const auto found = adjoint.find(f.y.key());
            if (found == adjoint.end() || found->second.state == 0) continue;
            const Adj g = found->second;
            backward(e, f.code, f.a, f.b, f.y, g, slot_of(f.a), slot_of(f.b), slot_of(f.c));
        }

        // The gradient of each input (constant 0 or 1 when that is all it is), then the loss operands.
        std::vector<Val> slots;
        for (int i = 0; i < chunk.arguments; ++i) {
            const auto found = adjoint.find(Val{kInputValue, uint32_t(i)}.key());
            const Adj g = found == adjoint.end() ? Adj{} : found->second;
            slots.push_back(g.state == 0 ? e.literal(0.0) : g.state == 1 ? e.literal(1.0) : g.v);
        }
        Val left, right;
        if (contract) {
            left = forward.back().a;
            right = forward.back().b;
        } else if (loss) {
            left = result;
        }

        // Keep what the gradients and the loss need.
        const size_t n = e.steps.size();
        std::vector<char> needed(n, 0);
        auto need = [&](Val v) {
            if (v.kind == kTemporaryValue && v.index != kSkipped) needed[v.index] = 1;
        };
        for (const Val &v : slots) need(v);
        need(left);
        need(right);
        for (size_t j = n; j-- > 0;)
            if (needed[j]) {
                need(e.steps[j].a);
                need(e.steps[j].b);
                need(e.steps[j].c);
            }
        std::vector<uint32_t> renumber(n, kSkipped);
        std::vector<Step> kept;
        for (size_t j = 0; j < n; ++j)
            if (needed[j]) {
                renumber[j] = uint32_t(kept.size());
                kept.push_back(e.steps[j]);
            }
        auto moved = [&](Val v) {
            if (v.kind == kTemporaryValue) v.index = renumber[v.index];
            return v;
        };
        for (Step &s : kept) {
            s.a = moved(s.a);
            s.b = moved(s.b);
            s.c = moved(s.c);
        }
        for (Val &v : slots) v = moved(v);
        left = moved(left);
        right = moved(right);

        if (kept.size() > size_t(kAdjointLength)) {
            error = "the adjoint program needs more instructions than the kernel holds";
            return SCIHIP_EINVAL;
        }
        if (e.constants.size() > size_t(kAdjointLiterals)) {
            error = "the adjoint program needs more constants than the kernel holds";
            return SCIHIP_EINVAL;
        }

        // Registers: inputs, constants, then the values, each freed after its last use.
        const size_t m = kept.size();
        std::vector<size_t> last(m, 0);
        auto use = [&](Val v, size_t at) {
            if (v.kind == kTemporaryValue) last[v.index] = std::max(last[v.index], at);
        };
        for (size_t j = 0; j < m; ++j) {
            use(kept[j].a, j);
            use(kept[j].b, j);
            use(kept[j].c, j);
        }
        for (const Val &v : slots) use(v, m);
        use(left, m);
        use(right, m);
        const uint32_t base = inputs + uint32_t(e.constants.size());
        std::vector<uint32_t> physical(m, 0);
        std::vector<char> busy;
        uint32_t highest = 0;
        for (size_t j = 0; j < m; ++j) {
            for (const Val *v : {&kept[j].a, &kept[j].b, &kept[j].c})
                if (v->kind == kTemporaryValue && last[v->index] == j) busy[physical[v->index]] = 0;
            uint32_t r = 0;
            while (r < busy.size() && busy[r]) ++r;
            if (r == busy.size()) busy.push_back(0);
            busy[r] = 1;
            physical[j] = r;
            highest = std::max(highest, r + 1);
        }
        auto reg = [&](Val v) -> uint32_t {
            if (v.kind == kInputValue) return v.index;
            if (v.kind == kLiteralValue) return inputs + v.index;
            if (v.kind == kTemporaryValue) return base + physical[v.index];
            return 0;
        };
        plan = AdjointPlan{};
        plan.precision = chunk.precision;
        plan.seeded = seeded;
        plan.loss = loss;
        plan.contract = contract;
        plan.arguments = inputs;
        plan.gradients = chunk.arguments;
        plan.registers = base + highest;
        if (plan.registers > 256) {
            error = "the adjoint program needs more registers than the kernel has";
            return SCIHIP_EINVAL;
        }
        for (size_t j = 0; j < m; ++j) {
            const Step &s = kept[j];
            plan.word.push_back(s.code | (base + physical[j]) << 8 | reg(s.a) << 16 | reg(s.b) << 24);
            plan.third.push_back(uint8_t(reg(s.c)));
        }
        for (const Val &v : slots) plan.slot.push_back(uint8_t(reg(v)));
        plan.left = reg(left);
        plan.right = reg(right);
        plan.constants = e.constants;
        return SCIHIP_OK;
    }

    std::string adjoint_key(const AdjointPlan &plan) {
        std::string key;
        key += char(2); // distinct from the chunk kernels' keys (mode 0 / 1)
        key += char(plan.precision);
        key += char(plan.loss);
        key += char(plan.contract);
        key += char(plan.seeded);
        key += char(plan.arguments);
        key += char(plan.gradients);
        key += char(plan.constants.size());
        key += char(plan.left);
        key += char(plan.right);
        key.append(reinterpret_cast<const char *>(plan.word.data()), plan.word.size() * sizeof(uint32_t));
        key.append(reinterpret_cast<const char *>(plan.third.data()), plan.third.size());
        key.append(reinterpret_cast<const char *>(plan.slot.data()), plan.slot.size());
        return key;
    }

    std::string generate_adjoint(const AdjointPlan &plan, int unroll) {
        const bool single = plan.precision == SCIHIP_SINGLE;
        const char *type = single ? "float" : "double";
        const uint32_t literals = uint32_t(plan.constants.size());
        std::ostringstream o;
        o << "extern \"C\" __global__ __launch_bounds__(256) void scihip_specialised(const scihip::AdjointLaunch<"
          << type << "> a) {\n"
          << "    using namespace scihip;\n"
          << "    typedef " << type << " T;\n"
          << "    constexpr int W = " << (single ? 4 : 2) << ";\n"
          << "    constexpr int U = " << unroll << ";\n"
          << "    typedef Lanes<T, W> L;\n"
          << "    const Adjoint<T> &p = a.p;\n";
        for (uint32_t i = 0; i < literals; ++i) o << "    const L k" << i << " = splat<T, W>(p.constants[" << i << "]);\n";
        if (plan.loss) o << "    double total = 0.0;\n    uint32_t poisoned = 0;\n";
        o << "    const uint64_t steps = (p.count + W - 1) / W;\n"
          << "    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * U;\n"
          << "    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x * U + threadIdx.x; base < steps; base += stride) {\n";
        for (int u = 0; u < unroll; ++u)
            o << "        const uint64_t v" << u << " = base + " << 256 * u << ";\n"
              << "        const bool live" << u << " = v" << u << " < steps;\n"
              << "        const uint64_t at" << u << " = v" << u << " * W;\n"
              << "        const bool full" << u << " = at" << u << " + W <= p.count;\n";
        for (int u = 0; u < unroll; ++u)
            for (uint32_t i = 0; i < plan.arguments; ++i)
                o << "        L i" << i << "_" << u << "{}; if (live" << u << ") i" << i << "_" << u
                  << " = fetch<T, W>(p, " << i << ", at" << u << ", full" << u << ");\n";
        for (int u = 0; u < unroll; ++u) {
            const std::string suffix = "_" + std::to_string(u);
            std::vector<std::string> names(plan.registers);
            for (uint32_t r = 0; r < plan.registers; ++r) {
                if (r < plan.arguments) names[r] = "i" + std::to_string(r) + suffix;
                else if (r < plan.arguments + literals) names[r] = "k" + std::to_string(r - plan.arguments);
                else names[r] = "k0"; // never read before written
            }
            o << "        if (live" << u << ") {\n";
            for (size_t s = 0; s < plan.word.size(); ++s) {
                const uint32_t word = plan.word[s];
                const uint32_t code = word & 0xFFu, target = (word >> 8) & 0xFFu;
                const int n = code == kFuse || code == kRouteMinimum || code == kRouteMaximum ? 3
                              : (code >= kAdd) ? 2
                                               : 1;
                const std::string first = names[(word >> 16) & 0xFFu];
                const std::string second = n > 1 ? names[word >> 24] : first;
                const std::string third = n > 2 ? names[plan.third[s]] : first;
                const std::string value = "s" + std::to_string(s) + suffix;
                o << "            const L " << value << " = lanes<" << code << "u, T, W>(" << first << ", " << second
                  << ", " << third << ");\n";
                names[target] = value;
            }
            if (plan.loss) {
                if (plan.contract)
                    o << "            contract<T, W>(total, " << names[plan.left] << ", " << names[plan.right] << ", at"
                      << u << ", full" << u << ", p.count);\n";
                else
                    o << "            consume<T, W>(0, total, poisoned, " << names[plan.left] << ", at" << u << ", full"
                      << u << ", p.count);\n";
            }
            for (uint32_t g = 0; g < plan.gradients; ++g)
                o << "            if (p.outputs[" << g << "]) store_to<T, W>(p.outputs[" << g
                  << "], p.count, p.aligned != 0, at" << u << ", full" << u << ", " << names[plan.slot[g]] << ");\n";
            o << "        }\n";
        }
        o << "    }\n";
        if (plan.loss) o << "    leave_partial(0, a.partials, total, poisoned);\n";
        o << "}\n";
        return o.str();
    }

    namespace {

        constexpr int kAdjointSmall = 16;  // registers of the VGPR-resident variants
        constexpr int kAdjointMedium = 32; // float only: 32 dwords per lane bank
        constexpr int kAdjointLarge = 256; // scratch-resident, one lane per thread
        constexpr int kAdjointUnroll = 2;  // vectors per thread per sweep of a specialised adjoint

        template <typename T> constexpr int lanes_of() { return sizeof(T) == 4 ? 4 : 2; }

        template <typename T>
        hipError_t interpret_adjoint(const AdjointLaunch<T> &launch, int registers, unsigned blocks, hipStream_t stream) {
            constexpr int W = lanes_of<T>();
            const dim3 grid(blocks), block(256);
            if (registers == kAdjointSmall) {
                hipLaunchKernelGGL((differentiate<T, W, kAdjointSmall>), grid, block, 0, stream, launch);
            } else if (registers == kAdjointMedium) {
                if constexpr (sizeof(T) == 4)
                    hipLaunchKernelGGL((differentiate<T, W, kAdjointMedium>), grid, block, 0, stream, launch);
            } else {
                hipLaunchKernelGGL((differentiate<T, 1, kAdjointLarge>), grid, block, 0, stream, launch);
            }
            return hipGetLastError();
        }

        template <typename T>
        int adjoint(scihip_device *device, const scihip_chunk &chunk, const void *seed, void *const *gradients,
                    size_t count, double *loss) {
            AdjointPlan plan;
            std::string error;
            if (const int rc = plan_adjoint(chunk, seed != nullptr, loss != nullptr, plan, error)) {
                fail(error);
                return rc;
            }
            AdjointLaunch<T> launch;
            std::memset(&launch, 0, sizeof(launch));
            Adjoint<T> &p = launch.p;
            p.length = uint32_t(plan.word.size());
            p.arguments = plan.arguments;
            p.literals = uint32_t(plan.constants.size());
            p.gradients = plan.gradients;
            p.loss = plan.loss;
            p.contract = plan.contract;
            p.left = plan.left;
            p.right = plan.right;
            p.count = count;
            std::copy(plan.word.begin(), plan.word.end(), p.word);
            std::copy(plan.third.begin(), plan.third.end(), p.third);
            std::copy(plan.slot.begin(), plan.slot.end(), p.slot);
            for (size_t i = 0; i < plan.constants.size(); ++i) p.constants[i] = static_cast<T>(plan.constants[i]);
            bool aligned = true;
            const size_t bytes = count * sizeof(T);
            for (int i = 0; i < chunk.arguments; ++i) {
                if (!device->allocator->locate(chunk.inputs[i], bytes)) {
                    fail("an input of the chunk is not device memory of this device");
                    return SCIHIP_DECLINED;
                }
                p.inputs[i] = static_cast<const T *>(chunk.inputs[i]);
                aligned = aligned && (reinterpret_cast<uintptr_t>(chunk.inputs[i]) & 15u) == 0;
            }
            if (seed) {
                if (!device->allocator->locate(seed, bytes)) {
                    fail("the upstream gradient is not device memory of this device");
                    return SCIHIP_DECLINED;
                }
                p.inputs[chunk.arguments] = static_cast<const T *>(seed);
                aligned = aligned && (reinterpret_cast<uintptr_t>(seed) & 15u) == 0;
            }
            for (int i = 0; gradients && i < chunk.arguments; ++i) {
                if (!gradients[i]) continue;
                if (!device->allocator->locate(gradients[i], bytes)) {
                    fail("a gradient array is not device memory of this device");
                    return SCIHIP_DECLINED;
                }
                p.outputs[i] = static_cast<T *>(gradients[i]);
                aligned = aligned && (reinterpret_cast<uintptr_t>(gradients[i]) & 15u) == 0;
            }
            p.aligned = aligned ? 1u : 0u;
            launch.partials = static_cast<Partial *>(device->partials);
            const int mode = plan.loss ? 1 : 0;

            const Kernel *kernel = nullptr;
            if (device->policy != 1)
                kernel = device->compiler->find(adjoint_key(plan), [&] { return generate_adjoint(plan, kAdjointUnroll); },
                                                device->policy == 2);
            if (!kernel && device->policy == 2) {
                fail("specialised adjoint kernel failed to compile");
                return SCIHIP_ECOMPILE;
            }
            int rc;
            unsigned blocks = 0;
            if (kernel) {
                blocks = blocks_for(device, count, lanes_of<T>(), kAdjointUnroll, mode);
                void *parameters[] = {&launch}; // copied at launch (and into a graph node under capture)
                ++device->specialised;
                rc = status(hipModuleLaunchKernel(kernel->function, blocks, 1, 1, 256, 1, 1, 0, device->stream,
                                                  parameters, nullptr),
                            "specialised adjoint launch");
            } else {
                const int registers = plan.registers <= uint32_t(kAdjointSmall) ? kAdjointSmall
                                      : (sizeof(T) == 4 && plan.registers <= uint32_t(kAdjointMedium)) ? kAdjointMedium
                                                                                                   : kAdjointLarge;
                const int width = registers == kAdjointLarge ? 1 : lanes_of<T>();
                ++device->interpreted;
                blocks = blocks_for(device, count, width, 1, mode);
                rc = status(interpret_adjoint<T>(launch, registers, blocks, device->stream), "adjoint launch");
            }
            // the loss goes straight into the pinned host word (no copy launch)
            if (!rc && plan.loss)
                rc = status(fold_partials(0, launch.partials, blocks, loss ? device->staging : device->result,
                                          device->stream),
                            "loss fold");
            if (rc || !loss) return rc;
            if ((rc = status(hipStreamSynchronize(device->stream), "adjoint"))) return rc;
            *loss = *device->staging;
            return SCIHIP_OK;
        }

        int zero(scihip_device *device, void *output, size_t bytes) {
            if (device->capturing) { // a memset node replays wrongly: keep it out of graphs
                fail("a degenerate product adjoint cannot be recorded into a graph");
                return SCIHIP_EINVAL;
            }
            return status(hipMemsetAsync(output, 0, bytes, device->stream), "hipMemsetAsync");
        }

    } // namespace

} // namespace scihip

using namespace scihip;

extern "C" {

int scihip_adjoint(scihip_device *device, const scihip_chunk *chunk, const void *seed, void *const *gradients,
                   size_t count, double *loss) {
    if (!device || !chunk || chunk->precision > SCIHIP_WIDE) return SCIHIP_EINVAL;
    if (!wellformed(*chunk)) {
        fail("malformed chunk");
        return SCIHIP_EINVAL;
    }
    if (count == 0) {
        if (loss) *loss = 0.0;
        return SCIHIP_OK;
    }
    std::lock_guard<std::mutex> lock(device->guard);
    if (loss && device->capturing) {
        fail("scihip_adjoint with a loss cannot be recorded into a graph (it returns a value)");
        return SCIHIP_EINVAL;
    }
    return chunk->precision == SCIHIP_SINGLE ? adjoint<float>(device, *chunk, seed, gradients, count, loss)
                                             : adjoint<double>(device, *chunk, seed, gradients, count, loss);
}

int scihip_adjoint_describe(const scihip_chunk *chunk, int seeded, int loss, scihip_adjoint_program *program) {
    if (!chunk || !program || chunk->precision > SCIHIP_WIDE) return SCIHIP_EINVAL;
    if (!wellformed(*chunk)) {
        fail("malformed chunk");
        return SCIHIP_EINVAL;
    }
    AdjointPlan plan;
    std::string error;
    if (const int rc = plan_adjoint(*chunk, seeded != 0, loss != 0, plan, error)) {
        fail(error);
        return rc;
    }
    std::memset(program, 0, sizeof(*program));
    program->length = uint32_t(plan.word.size());
    program->arguments = plan.arguments;
    program->literals = uint32_t(plan.constants.size());
    program->registers = plan.registers;
    program->gradients = plan.gradients;
    program->loss = plan.loss;
    program->contract = plan.contract;
    program->left = plan.left;
    program->right = plan.right;
    std::copy(plan.word.begin(), plan.word.end(), program->word);
    std::copy(plan.third.begin(), plan.third.end(), program->third);
    std::copy(plan.slot.begin(), plan.slot.end(), program->slot);
    std::copy(plan.constants.begin(), plan.constants.end(), program->constants);
    return SCIHIP_OK;
}

int scihip_product_adjoint(scihip_device *device, uint32_t mode, const void *left, const void *right, const void *seed,
                           void *left_gradient, void *right_gradient, size_t rows, size_t columns, size_t depth) {
    if (!device || mode > SCIHIP_PRODUCT_FLOAT64) return SCIHIP_EINVAL;
    if (depth == 0) return SCIHIP_OK; // both gradients are empty
    std::lock_guard<std::mutex> lock(device->guard);
    const size_t element = mode == SCIHIP_PRODUCT_FLOAT64 ? 8 : 4;
    auto located = [&](const void *p, size_t n) { return !p || device->allocator->locate(p, n * element); };
    if (!located(left, rows * depth) || !located(right, depth * columns) || !located(seed, rows * columns) ||
        !located(left_gradient, rows * depth) || !located(right_gradient, depth * columns) ||
        (left_gradient && !right) || (right_gradient && !left) || ((left_gradient || right_gradient) && !seed)) {
        fail("an operand of the product adjoint is missing or not device memory of this device");
        return SCIHIP_DECLINED;
    }
    const int m = int(mode);
    const int64_t r = int64_t(rows), c = int64_t(columns), d = int64_t(depth);
    int rc = SCIHIP_OK;
    // d left = G right^T: G (rows x columns) times right^T, whose transpose is right itself.
    if (left_gradient)
        rc = rows == 0 ? SCIHIP_OK
             : columns == 0 ? zero(device, left_gradient, rows * depth * element)
                            : multiply(device, m, seed, c, false, right, c, false, left_gradient, r, d, c);
    // d right = left^T G: left^T (depth x rows) times G (rows x columns).
    if (!rc && right_gradient)
        rc = columns == 0 ? SCIHIP_OK
             : rows == 0  ? zero(device, right_gradient, depth * columns * element)
                          : multiply(device, m, left, d, true, seed, c, true, right_gradient, d, c, r);
    return rc;
}

} // extern "C"
