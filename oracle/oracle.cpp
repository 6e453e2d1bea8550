// CPU oracle: a restatement of the reference's CPU path (libsci 0.2.0-alpha.1), used as the
// checker of the HIP backend. TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it; the backend itself never does.
//
// What it restates, each with the reference file:line it follows:
//  - the chunk executor: sweep / evaluate / accumulate over 1,024-element blocks
//    (src/core/drivers/processor/dispatch.cpp:42-165) and the part split of Plan
//    (include/core/drivers/processor/plan.hpp:28-53), with the double Accumulator
//    (include/core/bytecode/accumulator.hpp:21-38);
//  - the element kernels of every code (include/core/drivers/processor/kernels/generic/
//    operation.hpp:25-110) with libsci's own exp / log / sin / cos
//    (generic/exponential.hpp:24-61, generic/logarithm.hpp:26-66, generic/trigonometry.hpp:29-80,
//    constants in generic/constants.hpp:16-119, polynomials in generic/polynomial.hpp:17-61),
//    glibc for the library codes (generic/library.hpp:22-34) and a true fma (generic/scalar.hpp:69-89
//    with __FMA__, or the AVX-512 pack);
//  - the folds sum / dot / extreme with the lane order of a pack of `width` lanes
//    (generic/reductions.hpp:14-124, pack semantics src/core/drivers/processor/kernels/avx512.cpp:17-83);
//  - the register-blocked product (src/core/drivers/processor/dispatch.cpp:171-224);
//  - LU with partial pivoting and the substitutions (include/linear/decompose.hpp:29-72,
//    include/linear/solve.hpp:25-52);
//  - linspace (include/operations/creation.hpp:65-84).
// And one function with no reference counterpart (the reference has no differentiation):
//  - adjoint: the gradient of sum(result) (or sum(seed * result)) of a chunk, element by element,
//    by the chain rule over the chunk's instructions in reverse, each local derivative its textbook
//    closed form evaluated in double at the forward values (which are computed in the chunk's
//    precision by the element kernels above). It shares no code with the backend's backward
//    bytecode; central differences in the tests pin it.
// Built with -ffp-contract=off so that every separate multiply and add rounds as written.

#include "scihip.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <bit>
#include <cmath>
#include <cstdint>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <math.h>
#include <numbers>
#include <thread>
#include <vector>

namespace {

    constexpr size_t block = 1024;            // plan.hpp:23
    constexpr size_t threshold = size_t{1} << 17; // plan.hpp:59

    // ---------------------------------------------------------------- constants (constants.hpp)

    template <typename T> struct K;

    template <> struct K<float> {
        using U = uint32_t;
        static constexpr int mantissa = 23;
        static constexpr int width = 32;
        static constexpr U fraction = 0x007fffffu;
        static constexpr U one = 0x3f800000u;
        static constexpr float round = 12582912.0f;
        static constexpr float magnitude = 8388608.0f;
        static constexpr float exponent = 8388735.0f;
        static constexpr float binary = std::numbers::log2e_v<float>;
        static constexpr float natural = 355.0f / 512.0f;
        static constexpr float remainder = -2.12194440e-4f;
        static constexpr float lower = -104.0f;
        static constexpr float upper = 89.0f;
        static constexpr float normal = 126.0f;
        static constexpr float tiny = 1.17549435e-38f;
        static constexpr float lift = 33554432.0f;
        static constexpr float shift = 25.0f;
        static constexpr float root = std::numbers::sqrt2_v<float>;
        static constexpr float quadrant = 0.636619772367581343f;
        static constexpr float quarter[4] = {1.5703125f, 4.837512969970703125e-4f, 7.549790126404332e-08f,
                                             -1.7151245100058819e-15f};
        static constexpr float reach = 1048576.0f;
        static constexpr size_t exponential = 7, logarithm = 4, sine = 5, cosine = 6;
    };

    template <> struct K<double> {
        using U = uint64_t;
        static constexpr int mantissa = 52;
        static constexpr int width = 64;
        static constexpr U fraction = 0x000fffffffffffffull;
        static constexpr U one = 0x3ff0000000000000ull;
        static constexpr double round = 6755399441055744.0;
        static constexpr double magnitude = 4503599627370496.0;
        static constexpr double exponent = 4503599627371519.0;
        static constexpr double binary = std::numbers::log2e;
        static constexpr double natural =
                std::bit_cast<double>(std::bit_cast<uint64_t>(std::numbers::ln2) & 0xFFFFFFFFFFE00000ULL);
        static constexpr double remainder = 1.90821492927058770002e-10;
        static constexpr double lower = -746.0;
        static constexpr double upper = 710.0;
        static constexpr double normal = 1022.0;
        static constexpr double tiny = 2.2250738585072014e-308;
        static constexpr double lift = 18014398509481984.0;
        static constexpr double shift = 54.0;
        static constexpr double root = std::numbers::sqrt2;
        static constexpr double quadrant = 6.36619772367581382433e-01;
        static constexpr double quarter[4] = {1.57079632673412561417e+00, 6.07710050630396597660e-11,
                                              2.02226624871116645580e-21, 8.47842766036889956997e-32};
        static constexpr double reach = 1048576.0;
        static constexpr size_t exponential = 13, logarithm = 10, sine = 9, cosine = 10;
    };

    // ---------------------------------------------------------------- polynomials (polynomial.hpp)

    // Coefficients of a polynomial, c[k] multiplying x^k, built in double then narrowed (polynomial.hpp:17-50).
    template <typename T, size_t N> struct Poly {
        T c[N]{};
        static constexpr Poly factorials() {
            Poly p{};
            double f = 1.0;
            for (size_t i = 0; i < N; ++i) {
                if (i > 0) f /= static_cast<double>(i);
                p.c[i] = static_cast<T>(f);
            }
            return p;
        }
        static constexpr Poly reciprocals() {
            Poly p{};
            for (size_t i = 0; i < N; ++i) p.c[i] = static_cast<T>(2.0 / static_cast<double>(2 * i + 3));
            return p;
        }
        template <size_t offset> static constexpr Poly alternating() {
            Poly p{};
            for (size_t i = 0; i < N; ++i) {
                double f = 1.0;
                for (size_t s = 2; s <= 2 * i + offset; ++s) f *= static_cast<double>(s);
                p.c[i] = static_cast<T>((i % 2 == 0 ? 1.0 : -1.0) / f);
            }
            return p;
        }
        // sum c[k] x^(k - from), k in [from, N), Horner with fma (polynomial.hpp:55-61).
        template <size_t from> T horner(T x) const {
            T r = c[N - 1];
            for (size_t i = N - 1; i-- > from;) r = std::fma(r, x, c[i]);
            return r;
        }
    };

    template <typename T> using Bits = typename K<T>::U;
    template <typename T> inline Bits<T> raw(T v) { return std::bit_cast<Bits<T>>(v); }
    template <typename T> inline T cast(Bits<T> v) { return std::bit_cast<T>(v); }

    // NaN-propagating minimum / maximum of the one-lane pack (scalar.hpp:112-127).
    template <typename T> inline T lesser(T l, T r) {
        if (std::isnan(l) || std::isnan(r)) return l + r;
        return r < l ? r : l;
    }
    template <typename T> inline T greater(T l, T r) {
        if (std::isnan(l) || std::isnan(r)) return l + r;
        return r > l ? r : l;
    }
    // The vector forms: blend(isnan(l), min_ps(l, r), l) (avx512.cpp:43-49); min_ps gives r on ties and NaN.
    template <typename T> inline T lesser_vector(T l, T r) {
        if (std::isnan(l)) return l;
        return l < r ? l : r;
    }
    template <typename T> inline T greater_vector(T l, T r) {
        if (std::isnan(l)) return l;
        return l > r ? l : r;
    }

    // ---------------------------------------------------------------- e^x (exponential.hpp:24-61)

    template <typename T> T exponential(T x) {
        using L = K<T>;
        static constexpr auto series = Poly<T, L::exponential + 1>::factorials();
        // pack::maximum(pack::minimum(x, upper), lower): NaN stays NaN.
        const T low = lesser_vector(x, L::upper);
        const T c = greater_vector(low, L::lower);
        const T k = std::fma(c, L::binary, L::round) - L::round;
        const T negative = -k;
        T r = std::fma(negative, L::natural, c);
        r = std::fma(negative, L::remainder, r);
        const T significand = std::fma(r, series.template horner<1>(r), T{1});
        auto scale = [](T e) { return cast<T>(raw<T>(e + L::exponent) << L::mantissa); };
        if (!(std::fabs(k) > L::normal)) return significand * scale(k);
        const T half = std::fma(k, T{0.5}, L::round) - L::round;
        const T rest = k - half;
        return (significand * scale(half)) * scale(rest);
    }

    // ---------------------------------------------------------------- ln x (logarithm.hpp:26-66)

    template <typename T> T logarithm(T x) {
        using L = K<T>;
        static constexpr auto series = Poly<T, L::logarithm>::reciprocals();
        const bool small = x < L::tiny;
        const T scaled = small ? x * L::lift : x;
        const Bits<T> pattern = raw(scaled);
        T exponent = cast<T>((pattern >> L::mantissa) | raw<T>(L::magnitude)) - L::exponent;
        T mantissa = cast<T>((pattern & L::fraction) | L::one);
        const bool large = mantissa > L::root;
        if (large) {
            mantissa = mantissa * T{0.5};
            exponent = exponent + T{1};
        }
        if (small) exponent = exponent - L::shift;
        const T f = mantissa - T{1};
        const T s = f / (f + T{2});
        const T z = s * s;
        const T tail = z * series.template horner<0>(z);
        const T half = (f * f) * T{0.5};
        const T inner = std::fma(s, half + tail, exponent * L::remainder);
        T result = exponent * L::natural - ((half - inner) - f);
        constexpr T infinity = std::numeric_limits<T>::infinity();
        if (x == infinity) result = x;
        if (x == T{0}) result = -infinity;
        if (x < T{0}) result = std::numeric_limits<T>::quiet_NaN();
        if (std::isnan(x)) result = x;
        return result;
    }

    // ---------------------------------------------------------------- sin / cos (trigonometry.hpp:29-80)

    template <typename T, bool phase> T trigonometry(T x) {
        using L = K<T>;
        using U = Bits<T>;
        static constexpr auto odd = Poly<T, L::sine + 1>::template alternating<1>();
        static constexpr auto even = Poly<T, L::cosine + 1>::template alternating<0>();
        if (!(x <= L::reach && x >= -L::reach)) {
            if constexpr (sizeof(T) == 4) return phase ? ::cosf(x) : ::sinf(x);
            else return phase ? ::cos(x) : ::sin(x);
        }
        const T shifted = std::fma(x, L::quadrant, L::round);
        const T quarters = shifted - L::round;
        const T negative = -quarters;
        T r = std::fma(negative, L::quarter[0], x);
        r = std::fma(negative, L::quarter[1], r);
        r = std::fma(negative, L::quarter[2], r);
        r = std::fma(negative, L::quarter[3], r);
        const T square = r * r;
        const T sine = std::fma(r * square, odd.template horner<1>(square), r);
        const T cosine = std::fma(square, even.template horner<1>(square), T{1});
        U quadrant = raw(shifted);
        if constexpr (phase) quadrant = quadrant ^ (quadrant << 1);
        const U flip = (quadrant & U{2}) << (L::width - 2);
        const bool swap = (raw(shifted) & U{1}) != 0;
        const T result = phase ? (swap ? sine : cosine) : (swap ? cosine : sine);
        return cast<T>(raw(result) ^ flip);
    }

    template <typename T> T library(int code, T x) {
        const bool f = sizeof(T) == 4;
        switch (code) {
        case SCIHIP_ASIN: return f ? ::asinf(x) : ::asin(x);
        case SCIHIP_ACOS: return f ? ::acosf(x) : ::acos(x);
        case SCIHIP_ATAN: return f ? ::atanf(x) : ::atan(x);
        case SCIHIP_SINH: return f ? ::sinhf(x) : ::sinh(x);
        case SCIHIP_COSH: return f ? ::coshf(x) : ::cosh(x);
        case SCIHIP_TANH: return f ? ::tanhf(x) : ::tanh(x);
        case SCIHIP_ASINH: return f ? ::asinhf(x) : ::asinh(x);
        case SCIHIP_ACOSH: return f ? ::acoshf(x) : ::acosh(x);
        default: return f ? ::atanhf(x) : ::atanh(x);
        }
    }

    // What one code does to one element (operation.hpp:25-110).
    template <typename T> T apply(int code, T a, T b, T c) {
        using U = Bits<T>;
        constexpr U sign = U{1} << (sizeof(T) * 8 - 1);
        const T one{1};
        switch (code) {
        case SCIHIP_COPY: return a;
        case SCIHIP_NEGATE: return cast<T>(raw(a) ^ sign);
        case SCIHIP_ABSOLUTE: return cast<T>(raw(a) & ~sign);
        case SCIHIP_ROOT: return std::sqrt(a);
        case SCIHIP_EXPONENTIAL: return exponential(a);
        case SCIHIP_LOGARITHM: return logarithm(a);
        case SCIHIP_SIN: return trigonometry<T, false>(a);
        case SCIHIP_COS: return trigonometry<T, true>(a);
        case SCIHIP_TAN: return trigonometry<T, false>(a) / trigonometry<T, true>(a);
        case SCIHIP_ASIN: case SCIHIP_ACOS: case SCIHIP_ATAN:
        case SCIHIP_SINH: case SCIHIP_COSH: case SCIHIP_TANH:
        case SCIHIP_ASINH: case SCIHIP_ACOSH: case SCIHIP_ATANH:
            return library<T>(code, a);
        case SCIHIP_CSC: return one / trigonometry<T, false>(a);
        case SCIHIP_SEC: return one / trigonometry<T, true>(a);
        case SCIHIP_COT: return one / (trigonometry<T, false>(a) / trigonometry<T, true>(a));
        case SCIHIP_ACSC: return library<T>(SCIHIP_ASIN, one / a);
        case SCIHIP_ASEC: return library<T>(SCIHIP_ACOS, one / a);
        case SCIHIP_ACOT: return library<T>(SCIHIP_ATAN, one / a);
        case SCIHIP_CSCH: return one / library<T>(SCIHIP_SINH, a);
        case SCIHIP_SECH: return one / library<T>(SCIHIP_COSH, a);
        case SCIHIP_COTH: return one / library<T>(SCIHIP_TANH, a);
        case SCIHIP_ACSCH: return library<T>(SCIHIP_ASINH, one / a);
        case SCIHIP_ASECH: return library<T>(SCIHIP_ACOSH, one / a);
        case SCIHIP_ACOTH: return library<T>(SCIHIP_ATANH, one / a);
        case SCIHIP_ADD: return a + b;
        case SCIHIP_SUBTRACT: return a - b;
        case SCIHIP_MULTIPLY: return a * b;
        case SCIHIP_DIVIDE: return a / b;
        case SCIHIP_MINIMUM: return lesser(a, b);
        case SCIHIP_MAXIMUM: return greater(a, b);
        default: return std::fma(a, b, c);
        }
    }

    // ---------------------------------------------------------------- packs (elementwise.hpp:14-83)

    // The reference runs each code over a block a pack at a time: every full pack through the pack's
    // operations, the ragged tail element by element through the one-lane pack
    // (include/core/drivers/processor/kernels/generic/elementwise.hpp:14-83). Here a pack is 64 bytes
    // (16 floats or 8 doubles, the AVX-512 width) as a clang vector, and the tail runs apply().
    // A lane's result never depends on its pack: each lane takes the same IEEE operations in the same
    // order as apply() does, and the two pack-wide branches give every lane the value it would get
    // alone (the exponential's split scaling multiplies by powers of two whose product is the single
    // scale, rounding once where the single scale rounds; the trigonometric functions send exactly
    // the lanes beyond the Cody-Waite reach to the C library). So packs change the speed, not a bit.
    template <typename T> struct Lanes {
        static constexpr size_t width = 64 / sizeof(T);
        typedef T V __attribute__((ext_vector_type(64 / sizeof(T))));
        typedef Bits<T> B __attribute__((ext_vector_type(64 / sizeof(T))));
    };
    template <typename T> using PV = typename Lanes<T>::V;
    template <typename T> using PB = typename Lanes<T>::B;

#define ORACLE_INLINE inline __attribute__((always_inline))

    template <typename T> ORACLE_INLINE PV<T> vload(const T *p) {
        PV<T> v;
        std::memcpy(&v, p, sizeof(v));
        return v;
    }
    template <typename T> ORACLE_INLINE void vstore(T *p, PV<T> v) { std::memcpy(p, &v, sizeof(v)); }
    template <typename T> ORACLE_INLINE PV<T> vfma(PV<T> a, PV<T> b, PV<T> c) {
        return __builtin_elementwise_fma(a, b, c);
    }
    template <typename T> ORACLE_INLINE PB<T> vraw(PV<T> v) { return __builtin_bit_cast(PB<T>, v); }
    template <typename T> ORACLE_INLINE PV<T> vcast(PB<T> v) { return __builtin_bit_cast(PV<T>, v); }
    template <typename T, size_t N, size_t from> ORACLE_INLINE PV<T> vhorner(const Poly<T, N> &p, PV<T> x) {
        PV<T> r = p.c[N - 1];
        for (size_t i = N - 1; i-- > from;) r = vfma<T>(r, x, PV<T>(p.c[i]));
        return r;
    }
    // The vector minimum / maximum with a constant: a NaN first operand stays (exponential's clamp).
    template <typename T> ORACLE_INLINE PV<T> vlesser_vector(PV<T> l, T r) { return l != l ? l : (l < r ? l : PV<T>(r)); }
    template <typename T> ORACLE_INLINE PV<T> vgreater_vector(PV<T> l, T r) { return l != l ? l : (l > r ? l : PV<T>(r)); }

    template <typename T> ORACLE_INLINE PV<T> vexponential(PV<T> x) { // exponential.hpp:24-49
        using L = K<T>;
        using V = PV<T>;
        static constexpr auto series = Poly<T, L::exponential + 1>::factorials();
        const V c = vgreater_vector<T>(vlesser_vector<T>(x, L::upper), L::lower);
        const V k = vfma<T>(c, V(L::binary), V(L::round)) - L::round;
        const V negative = -k;
        V r = vfma<T>(negative, V(L::natural), c);
        r = vfma<T>(negative, V(L::remainder), r);
        const V significand = vfma<T>(r, vhorner<T, L::exponential + 1, 1>(series, r), V(T{1}));
        auto scale = [](V e) { return vcast<T>(vraw<T>(e + L::exponent) << L::mantissa); };
        const V magnitude = vcast<T>(vraw<T>(k) & PB<T>(~(Bits<T>{1} << (sizeof(T) * 8 - 1))));
        if (!__builtin_reduce_or(magnitude > L::normal)) return significand * scale(k);
        const V half = vfma<T>(k, V(T{0.5}), V(L::round)) - L::round;
        const V rest = k - half;
        return (significand * scale(half)) * scale(rest);
    }

    template <typename T> ORACLE_INLINE PV<T> vlogarithm(PV<T> x) { // logarithm.hpp:26-61
        using L = K<T>;
        using V = PV<T>;
        using B = PB<T>;
        static constexpr auto series = Poly<T, L::logarithm>::reciprocals();
        const auto small = x < L::tiny;
        const V scaled = small ? x * L::lift : x;
        const B pattern = vraw<T>(scaled);
        V exponent = vcast<T>((pattern >> L::mantissa) | raw<T>(L::magnitude)) - L::exponent;
        V mantissa = vcast<T>((pattern & L::fraction) | L::one);
        const auto large = mantissa > L::root;
        mantissa = large ? mantissa * T{0.5} : mantissa;
        exponent = large ? exponent + T{1} : exponent;
        exponent = small ? exponent - L::shift : exponent;
        const V f = mantissa - T{1};
        const V s = f / (f + T{2});
        const V z = s * s;
        const V tail = z * vhorner<T, L::logarithm, 0>(series, z);
        const V half = (f * f) * T{0.5};
        const V inner = vfma<T>(s, half + tail, exponent * L::remainder);
        V result = exponent * L::natural - ((half - inner) - f);
        constexpr T infinity = std::numeric_limits<T>::infinity();
        result = x == infinity ? x : result;
        result = x == T{0} ? V(-infinity) : result;
        result = x < T{0} ? V(std::numeric_limits<T>::quiet_NaN()) : result;
        return x != x ? x : result;
    }

    template <typename T, bool phase> ORACLE_INLINE PV<T> vtrigonometry(PV<T> x) { // trigonometry.hpp:29-74
        using L = K<T>;
        using V = PV<T>;
        using B = PB<T>;
        static constexpr auto odd = Poly<T, L::sine + 1>::template alternating<1>();
        static constexpr auto even = Poly<T, L::cosine + 1>::template alternating<0>();
        const V shifted = vfma<T>(x, V(L::quadrant), V(L::round));
        const V quarters = shifted - L::round;
        const V negative = -quarters;
        V r = vfma<T>(negative, V(L::quarter[0]), x);
        r = vfma<T>(negative, V(L::quarter[1]), r);
        r = vfma<T>(negative, V(L::quarter[2]), r);
        r = vfma<T>(negative, V(L::quarter[3]), r);
        const V square = r * r;
        const V sine = vfma<T>(r * square, vhorner<T, L::sine + 1, 1>(odd, square), r);
        const V cosine = vfma<T>(square, vhorner<T, L::cosine + 1, 1>(even, square), V(T{1}));
        B quadrant = vraw<T>(shifted);
        if constexpr (phase) quadrant = quadrant ^ (quadrant << 1);
        const B flip = (quadrant & Bits<T>{2}) << (L::width - 2);
        const auto swap = (vraw<T>(shifted) & Bits<T>{1}) != Bits<T>{0};
        V result = phase ? (swap ? sine : cosine) : (swap ? cosine : sine);
        result = vcast<T>(vraw<T>(result) ^ flip);
        const auto inside = (x <= L::reach) & (x >= -L::reach);
        if (__builtin_reduce_and(inside) == 0) { // some lane beyond the reach, infinite or NaN
            for (size_t l = 0; l < Lanes<T>::width; ++l)
                if (!inside[l]) result[l] = trigonometry<T, phase>(x[l]);
        }
        return result;
    }

    template <typename T> ORACLE_INLINE PV<T> vlibrary(int code, PV<T> x) { // library.hpp:22-34
        PV<T> out;
        for (size_t l = 0; l < Lanes<T>::width; ++l) out[l] = library<T>(code, x[l]);
        return out;
    }

    // One code over one full pack (operation.hpp:25-110 on the AVX-512 pack).
    template <typename T> ORACLE_INLINE PV<T> vapply(int code, PV<T> a, PV<T> b, PV<T> c) {
        using V = PV<T>;
        constexpr Bits<T> sign = Bits<T>{1} << (sizeof(T) * 8 - 1);
        const V one = T{1};
        switch (code) {
        case SCIHIP_COPY: return a;
        case SCIHIP_NEGATE: return vcast<T>(vraw<T>(a) ^ sign);
        case SCIHIP_ABSOLUTE: return vcast<T>(vraw<T>(a) & ~sign);
        case SCIHIP_ROOT: return __builtin_elementwise_sqrt(a);
        case SCIHIP_EXPONENTIAL: return vexponential<T>(a);
        case SCIHIP_LOGARITHM: return vlogarithm<T>(a);
        case SCIHIP_SIN: return vtrigonometry<T, false>(a);
        case SCIHIP_COS: return vtrigonometry<T, true>(a);
        case SCIHIP_TAN: return vtrigonometry<T, false>(a) / vtrigonometry<T, true>(a);
        case SCIHIP_ASIN: case SCIHIP_ACOS: case SCIHIP_ATAN:
        case SCIHIP_SINH: case SCIHIP_COSH: case SCIHIP_TANH:
        case SCIHIP_ASINH: case SCIHIP_ACOSH: case SCIHIP_ATANH:
            return vlibrary<T>(code, a);
        case SCIHIP_CSC: return one / vtrigonometry<T, false>(a);
        case SCIHIP_SEC: return one / vtrigonometry<T, true>(a);
        case SCIHIP_COT: return one / (vtrigonometry<T, false>(a) / vtrigonometry<T, true>(a));
        case SCIHIP_ACSC: return vlibrary<T>(SCIHIP_ASIN, one / a);
        case SCIHIP_ASEC: return vlibrary<T>(SCIHIP_ACOS, one / a);
        case SCIHIP_ACOT: return vlibrary<T>(SCIHIP_ATAN, one / a);
        case SCIHIP_CSCH: return one / vlibrary<T>(SCIHIP_SINH, a);
        case SCIHIP_SECH: return one / vlibrary<T>(SCIHIP_COSH, a);
        case SCIHIP_COTH: return one / vlibrary<T>(SCIHIP_TANH, a);
        case SCIHIP_ACSCH: return vlibrary<T>(SCIHIP_ASINH, one / a);
        case SCIHIP_ASECH: return vlibrary<T>(SCIHIP_ACOSH, one / a);
        case SCIHIP_ACOTH: return vlibrary<T>(SCIHIP_ATANH, one / a);
        case SCIHIP_ADD: return a + b;
        case SCIHIP_SUBTRACT: return a - b;
        case SCIHIP_MULTIPLY: return a * b;
        case SCIHIP_DIVIDE: return a / b;
        case SCIHIP_MINIMUM: return ((a != a) | (b != b)) ? a + b : (b < a ? b : a);
        case SCIHIP_MAXIMUM: return ((a != a) | (b != b)) ? a + b : (b > a ? b : a);
        default: return vfma<T>(a, b, c);
        }
    }

    // One code over n elements of a block: full packs, then the tail element by element. Each case
    // is its own loop, so the code's pack body is straight-line inside it.
    template <typename T, int C>
    ORACLE_INLINE void vsweep(const T *a, const T *b, const T *c, T *t, size_t n) {
        constexpr size_t W = Lanes<T>::width;
        size_t e = 0;
        for (; e + W <= n; e += W) {
            const PV<T> x = vload(a + e);
            const PV<T> y = b ? vload(b + e) : PV<T>(T{});
            const PV<T> z = c ? vload(c + e) : PV<T>(T{});
            vstore(t + e, vapply<T>(C, x, y, z));
        }
        for (; e < n; ++e) t[e] = apply<T>(C, a[e], b ? b[e] : T{}, c ? c[e] : T{});
    }

    template <typename T> ORACLE_INLINE void run_code_impl(int code, const T *a, const T *b, const T *c, T *t, size_t n) {
        switch (code) {
#define ORACLE_CASE(k) case k: vsweep<T, k>(a, b, c, t, n); return;
            ORACLE_CASE(0) ORACLE_CASE(1) ORACLE_CASE(2) ORACLE_CASE(3) ORACLE_CASE(4) ORACLE_CASE(5)
            ORACLE_CASE(6) ORACLE_CASE(7) ORACLE_CASE(8) ORACLE_CASE(9) ORACLE_CASE(10) ORACLE_CASE(11)
            ORACLE_CASE(12) ORACLE_CASE(13) ORACLE_CASE(14) ORACLE_CASE(15) ORACLE_CASE(16) ORACLE_CASE(17)
            ORACLE_CASE(18) ORACLE_CASE(19) ORACLE_CASE(20) ORACLE_CASE(21) ORACLE_CASE(22) ORACLE_CASE(23)
            ORACLE_CASE(24) ORACLE_CASE(25) ORACLE_CASE(26) ORACLE_CASE(27) ORACLE_CASE(28) ORACLE_CASE(29)
            ORACLE_CASE(30) ORACLE_CASE(31) ORACLE_CASE(32) ORACLE_CASE(33) ORACLE_CASE(34) ORACLE_CASE(35)
#undef ORACLE_CASE
        default: vsweep<T, SCIHIP_FUSE>(a, b, c, t, n); return;
        }
    }

    // Compiled for AVX-512, for AVX2 + FMA and for baseline x86-64 (where the pack's fma becomes a
    // call of fmaf / fma, still correctly rounded); the loader picks the one the host supports.
    __attribute__((target_clones("avx512f", "arch=haswell", "default")))
    void run_code_f32(int code, const float *a, const float *b, const float *c, float *t, size_t n) {
        run_code_impl<float>(code, a, b, c, t, n);
    }
    __attribute__((target_clones("avx512f", "arch=haswell", "default")))
    void run_code_f64(int code, const double *a, const double *b, const double *c, double *t, size_t n) {
        run_code_impl<double>(code, a, b, c, t, n);
    }
    inline void run_code(int code, const float *a, const float *b, const float *c, float *t, size_t n) {
        run_code_f32(code, a, b, c, t, n);
    }
    inline void run_code(int code, const double *a, const double *b, const double *c, double *t, size_t n) {
        run_code_f64(code, a, b, c, t, n);
    }

    // ---------------------------------------------------------------- folds (reductions.hpp:14-124)

    // A pack of W lanes is modelled lane by lane (W a compile-time width, so the lane loops are
    // vector code); the vector bodies use the vector min / max.
    // The folds at the full 64-byte width, four pack accumulators in registers; the same lane
    // order as the array forms below (accumulator j, lane l takes element 4Wk + jW + l).
    template <typename T> ORACLE_INLINE T fold_lanes(PV<T> a0, PV<T> a1, PV<T> a2, PV<T> a3) {
        T result{0};
        for (size_t l = 0; l < Lanes<T>::width; ++l) result += (a0[l] + a1[l]) + (a2[l] + a3[l]);
        return result;
    }

    template <typename T> ORACLE_INLINE T fold_sum_packs(const T *src, size_t count) {
        constexpr size_t W = Lanes<T>::width;
        PV<T> a0 = T{0}, a1 = T{0}, a2 = T{0}, a3 = T{0};
        size_t i = 0;
        for (; i + 4 * W <= count; i += 4 * W) {
            a0 = a0 + vload(src + i);
            a1 = a1 + vload(src + i + W);
            a2 = a2 + vload(src + i + 2 * W);
            a3 = a3 + vload(src + i + 3 * W);
        }
        for (; i + W <= count; i += W) a0 = a0 + vload(src + i);
        T result = fold_lanes<T>(a0, a1, a2, a3);
        for (; i < count; ++i) result += src[i];
        return result;
    }

    template <typename T> ORACLE_INLINE T fold_dot_packs(const T *a, const T *b, size_t count) {
        constexpr size_t W = Lanes<T>::width;
        PV<T> a0 = T{0}, a1 = T{0}, a2 = T{0}, a3 = T{0};
        size_t i = 0;
        for (; i + 4 * W <= count; i += 4 * W) {
            a0 = vfma<T>(vload(a + i), vload(b + i), a0);
            a1 = vfma<T>(vload(a + i + W), vload(b + i + W), a1);
            a2 = vfma<T>(vload(a + i + 2 * W), vload(b + i + 2 * W), a2);
            a3 = vfma<T>(vload(a + i + 3 * W), vload(b + i + 3 * W), a3);
        }
        for (; i + W <= count; i += W) a0 = vfma<T>(vload(a + i), vload(b + i), a0);
        T result = fold_lanes<T>(a0, a1, a2, a3);
        for (; i < count; ++i) result = std::fma(a[i], b[i], result);
        return result;
    }

    template <typename T, size_t W> ORACLE_INLINE T fold_sum(const T *src, size_t count) {
        if constexpr (W == Lanes<T>::width) return fold_sum_packs<T>(src, count);
        T acc[4 * W] = {};
        size_t i = 0;
        for (; i + 4 * W <= count; i += 4 * W)
            for (size_t j = 0; j < 4 * W; ++j) acc[j] = acc[j] + src[i + j];
        for (; i + W <= count; i += W)
            for (size_t l = 0; l < W; ++l) acc[l] = acc[l] + src[i + l];
        T result{0};
        for (size_t l = 0; l < W; ++l) result += (acc[l] + acc[W + l]) + (acc[2 * W + l] + acc[3 * W + l]);
        for (; i < count; ++i) result += src[i];
        return result;
    }

    template <typename T, size_t W> ORACLE_INLINE T fold_dot(const T *a, const T *b, size_t count) {
        if constexpr (W == Lanes<T>::width) return fold_dot_packs<T>(a, b, count);
        T acc[4 * W] = {};
        size_t i = 0;
        for (; i + 4 * W <= count; i += 4 * W)
            for (size_t j = 0; j < 4 * W; ++j) acc[j] = std::fma(a[i + j], b[i + j], acc[j]);
        for (; i + W <= count; i += W)
            for (size_t l = 0; l < W; ++l) acc[l] = std::fma(a[i + l], b[i + l], acc[l]);
        T result{0};
        for (size_t l = 0; l < W; ++l) result += (acc[l] + acc[W + l]) + (acc[2 * W + l] + acc[3 * W + l]);
        for (; i < count; ++i) result = std::fma(a[i], b[i], result);
        return result;
    }

    template <typename T, size_t W> ORACLE_INLINE T fold_extreme(const T *src, size_t count, bool minimum) {
        auto vec = [&](T l, T r) { return minimum ? lesser_vector(l, r) : greater_vector(l, r); };
        auto one = [&](T l, T r) { return minimum ? lesser(l, r) : greater(l, r); };
        // A one-lane pack is the Scalar pack: its "vector" operation is the scalar one.
        auto body = [&](T l, T r) { return W == 1 ? one(l, r) : vec(l, r); };
        T result = src[0];
        size_t i = 0;
        if (count >= 4 * W) {
            T acc[4 * W];
            for (size_t j = 0; j < 4 * W; ++j) acc[j] = src[j];
            for (i = 4 * W; i + 4 * W <= count; i += 4 * W)
                for (size_t j = 0; j < 4 * W; ++j) acc[j] = body(acc[j], src[i + j]);
            for (size_t l = 0; l < W; ++l) {
                const T first = body(body(acc[l], acc[W + l]), body(acc[2 * W + l], acc[3 * W + l]));
                result = one(result, first);
            }
        }
        for (; i < count; ++i) result = one(result, src[i]);
        return result;
    }

    // The fold of one block (reductions.hpp:14-124) at the pack width of the chosen table.
    template <typename T, size_t W>
    ORACLE_INLINE T fold_block(int fold, bool contract, const T *a, const T *b, size_t n) {
        if (contract) return fold_dot<T, W>(a, b, n);
        if (fold == SCIHIP_FOLD_SUM) return fold_sum<T, W>(a, n);
        return fold_extreme<T, W>(a, n, fold == SCIHIP_FOLD_MINIMUM);
    }
    template <typename T>
    ORACLE_INLINE T fold_block_impl(int fold, bool contract, const T *a, const T *b, size_t n, size_t width) {
        switch (width) {
        case 16: return fold_block<T, 16>(fold, contract, a, b, n);
        case 8: return fold_block<T, 8>(fold, contract, a, b, n);
        case 4: return fold_block<T, 4>(fold, contract, a, b, n);
        default: return fold_block<T, 1>(fold, contract, a, b, n);
        }
    }
    __attribute__((target_clones("avx512f", "arch=haswell", "default")))
    float fold_block_f32(int fold, bool contract, const float *a, const float *b, size_t n, size_t width) {
        return fold_block_impl<float>(fold, contract, a, b, n, width);
    }
    __attribute__((target_clones("avx512f", "arch=haswell", "default")))
    double fold_block_f64(int fold, bool contract, const double *a, const double *b, size_t n, size_t width) {
        return fold_block_impl<double>(fold, contract, a, b, n, width);
    }
    inline float fold_block(int fold, bool contract, const float *a, const float *b, size_t n, size_t width) {
        return fold_block_f32(fold, contract, a, b, n, width);
    }
    inline double fold_block(int fold, bool contract, const double *a, const double *b, size_t n, size_t width) {
        return fold_block_f64(fold, contract, a, b, n, width);
    }

    // ---------------------------------------------------------------- accumulator (accumulator.hpp:21-38)

    struct Accumulator {
        int fold;
        double total;
        explicit Accumulator(int f)
            : fold(f), total(f == SCIHIP_FOLD_SUM ? 0.0
                             : f == SCIHIP_FOLD_MINIMUM ? std::numeric_limits<double>::infinity()
                                                   : -std::numeric_limits<double>::infinity()) {}
        void add(double v) {
            if (fold == SCIHIP_FOLD_SUM || std::isnan(total) || std::isnan(v)) total += v;
            else total = fold == SCIHIP_FOLD_MINIMUM ? std::min(total, v) : std::max(total, v);
        }
    };

    // ---------------------------------------------------------------- chunk executor (dispatch.cpp:42-165)

    uint32_t cost(const scihip_chunk &chunk) { // chunk.hpp:127-144
        uint32_t total = chunk.arguments + 1u;
        for (uint8_t i = 0; i < chunk.length; ++i) {
            const int op = chunk.instructions[i].code;
            if (op == SCIHIP_DIVIDE || op == SCIHIP_ROOT) total += 4;
            else if (op == SCIHIP_EXPONENTIAL || op == SCIHIP_LOGARITHM) total += 12;
            else if (op >= SCIHIP_SIN && op <= SCIHIP_TAN) total += 16;
            else if (op >= SCIHIP_ASIN && op <= SCIHIP_ACOTH) total += 24;
            else total += 1;
        }
        return total;
    }

    bool contracted(const scihip_chunk &chunk, int fold) { // chunk.hpp:151-154
        return fold == SCIHIP_FOLD_SUM && chunk.length > 0 &&
               chunk.instructions[chunk.length - 1].code == SCIHIP_MULTIPLY &&
               chunk.instructions[chunk.length - 1].target == chunk.result;
    }

    struct Plan { // plan.hpp:28-50
        size_t count, parts = 1;
        Plan(const scihip_chunk &chunk, size_t n, size_t threads) : count(n) {
            const size_t blocks = (n + block - 1) / block;
            if (blocks < 2 || n * cost(chunk) < threshold) return;
            if (threads < 2) return;
            parts = std::min(blocks, threads * 8);
        }
        std::pair<size_t, size_t> range(size_t part) const {
            const size_t blocks = (count + block - 1) / block;
            return {blocks * part / parts * block, std::min(blocks * (part + 1) / parts * block, count)};
        }
    };

    // A fork-join pool on threads that live for the process, as the reference's Pool
    // (src/core/drivers/processor/pool.cpp:44-157): the caller takes parts too, helpers claim parts
    // from a shared counter, and the call returns when every part is done. Calls are serialised.
    class Pool {
    public:
        static Pool &global() {
            static Pool pool;
            return pool;
        }

        void run(size_t parts, size_t threads, const std::function<void(size_t)> &work) {
            std::lock_guard<std::mutex> call(serial);
            const size_t helpers = std::min(threads, parts) - 1;
            std::unique_lock<std::mutex> lock(guard);
            while (workers.size() < helpers) workers.emplace_back([this, k = workers.size()] { loop(k); });
            job = &work;
            total = parts;
            next.store(0);
            wanted = helpers;
            busy = helpers;
            ++generation;
            wake.notify_all();
            lock.unlock();
            for (size_t p; (p = next.fetch_add(1)) < parts;) work(p);
            lock.lock();
            finished.wait(lock, [&] { return busy == 0; });
            job = nullptr;
        }

        ~Pool() {
            {
                std::lock_guard<std::mutex> lock(guard);
                stopping = true;
            }
            wake.notify_all();
            for (auto &t : workers) t.join();
        }

    private:
        void loop(size_t index) {
            uint64_t seen = 0;
            std::unique_lock<std::mutex> lock(guard);
            for (;;) {
                wake.wait(lock, [&] { return stopping || (generation != seen && index < wanted); });
                if (stopping) return;
                seen = generation;
                const std::function<void(size_t)> *work = job;
                const size_t parts = total;
                lock.unlock();
                for (size_t p; (p = next.fetch_add(1)) < parts;) (*work)(p);
                lock.lock();
                if (--busy == 0) finished.notify_all();
            }
        }

        std::mutex serial, guard;
        std::condition_variable wake, finished;
        std::vector<std::thread> workers;
        const std::function<void(size_t)> *job = nullptr;
        std::atomic<size_t> next{0};
        size_t total = 0, wanted = 0, busy = 0;
        uint64_t generation = 0;
        bool stopping = false;
    };

    template <typename F> void fork(size_t parts, size_t threads, F &&work) {
        if (parts <= 1 || threads <= 1) {
            for (size_t p = 0; p < parts; ++p) work(p);
            return;
        }
        const std::function<void(size_t)> function = std::ref(work);
        Pool::global().run(parts, threads, function);
    }

    // Runs the first `steps` instructions over [first, last) block by block and hands each block's
    // value table to `visit` (dispatch.cpp:42-77). Temporaries and splatted constants are one block each.
    template <typename T, typename V>
    void sweep(const scihip_chunk &chunk, T *output, size_t first, size_t last, uint8_t steps, V &&visit) {
        std::vector<T> space((chunk.literals + chunk.temporaries) * block);
        T *values[256] = {};
        for (size_t i = 0; i < chunk.literals; ++i) {
            std::fill_n(space.data() + i * block, block, static_cast<T>(chunk.constants[i]));
            values[SCIHIP_SLOT_LITERAL | i] = space.data() + i * block;
        }
        for (size_t i = 0; i < chunk.temporaries; ++i)
            values[SCIHIP_SLOT_TEMPORARY | i] = space.data() + (chunk.literals + i) * block;
        for (size_t start = first; start < last; start += block) {
            const size_t n = std::min(block, last - start);
            for (size_t i = 0; i < chunk.arguments; ++i)
                values[SCIHIP_SLOT_INPUT | i] = const_cast<T *>(static_cast<const T *>(chunk.inputs[i])) + start;
            if (output) values[SCIHIP_SLOT_OUTPUT] = output + start;
            for (uint8_t s = 0; s < steps; ++s) {
                const scihip_instruction &in = chunk.instructions[s];
                const T *a = values[in.sources[0]], *b = values[in.sources[1]], *c = values[in.sources[2]];
                run_code(in.code, a, b, c, values[in.target], n);
            }
            visit(values, n);
        }
    }

    template <typename T> void evaluate(const scihip_chunk &chunk, T *output, size_t count, size_t threads) {
        if (count == 0) return;
        // Elementwise: the split never changes a result, so every thread takes an equal run of blocks.
        const size_t blocks = (count + block - 1) / block;
        const size_t parts = threads > 1 ? std::min(blocks, threads * 8) : 1;
        fork(parts, threads, [&](size_t p) {
            const size_t first = blocks * p / parts * block;
            const size_t last = std::min(blocks * (p + 1) / parts * block, count);
            if (first < last) sweep<T>(chunk, output, first, last, chunk.length, [](T *const *, size_t) {});
        });
    }

    template <typename T>
    double accumulate(const scihip_chunk &chunk, int fold, size_t count, size_t threads, size_t width) {
        if (count == 0) return fold == SCIHIP_FOLD_SUM ? 0.0 : std::numeric_limits<double>::quiet_NaN();
        const bool contract = contracted(chunk, fold);
        const uint8_t steps = contract ? chunk.length - 1 : chunk.length;
        const scihip_instruction *product = chunk.length > 0 ? &chunk.instructions[chunk.length - 1] : nullptr;
        const Plan plan(chunk, count, threads);
        std::vector<double> partials(plan.parts);
        fork(plan.parts, threads, [&](size_t part) {
            const auto [first, last] = plan.range(part);
            Accumulator acc(fold);
            sweep<T>(chunk, static_cast<T *>(nullptr), first, last, steps, [&](T *const *values, size_t n) {
                const T v = contract ? fold_block(fold, true, values[product->sources[0]],
                                                  values[product->sources[1]], n, width)
                                     : fold_block(fold, false, values[chunk.result], static_cast<const T *>(nullptr),
                                                  n, width);
                acc.add(static_cast<double>(v));
            });
            partials[part] = acc.total;
        });
        if (plan.parts == 1) return partials[0];
        Accumulator acc(fold);
        for (double p : partials) acc.add(p);
        return acc.total;
    }

    // ---------------------------------------------------------------- product (dispatch.cpp:171-224)

    template <typename T>
    void multiply(const T *left, const T *right, T *output, size_t rows, size_t columns, size_t depth,
                  size_t threads) {
        constexpr size_t height = 4, line = 64 / sizeof(T), band = 32;
        if (depth == 0) {
            std::fill_n(output, rows * columns, T{0});
            return;
        }
        const size_t bands = (rows + band - 1) / band;
        auto work = [&](size_t task) {
            const size_t top = task * band, bottom = std::min(rows, top + band);
            for (size_t column = 0; column < columns; column += line) {
                const size_t span = std::min(line, columns - column);
                for (size_t row = top; row < bottom; row += height) {
                    const size_t tall = std::min(height, bottom - row);
                    if (tall == height && span == line) {
                        T sums[height][line]{};
                        for (size_t inner = 0; inner < depth; ++inner) {
                            const T *panel = right + inner * columns + column;
                            for (size_t down = 0; down < height; ++down) {
                                const T scale = left[(row + down) * depth + inner];
                                for (size_t across = 0; across < line; ++across)
                                    sums[down][across] += scale * panel[across];
                            }
                        }
                        for (size_t down = 0; down < height; ++down)
                            for (size_t across = 0; across < line; ++across)
                                output[(row + down) * columns + column + across] = sums[down][across];
                    } else {
                        for (size_t down = 0; down < tall; ++down)
                            for (size_t across = 0; across < span; ++across) {
                                T sum{};
                                for (size_t inner = 0; inner < depth; ++inner)
                                    sum += left[(row + down) * depth + inner] * right[inner * columns + column + across];
                                output[(row + down) * columns + column + across] = sum;
                            }
                    }
                }
            }
        };
        fork(bands, bands > 1 && rows * columns * depth >= (size_t{1} << 18) ? threads : 1, work);
    }

    // ---------------------------------------------------------------- LU (decompose.hpp:29-72, solve.hpp:25-52)

    template <typename T> void factor(T *matrix, int64_t *rows, size_t order, size_t threads) {
        for (size_t i = 0; i < order; ++i) rows[i] = static_cast<int64_t>(i);
        for (size_t step = 0; step < order; ++step) {
            size_t pivot = step;
            for (size_t row = step + 1; row < order; ++row)
                if (std::fabs(matrix[row * order + step]) > std::fabs(matrix[pivot * order + step])) pivot = row;
            if (pivot != step) {
                std::swap_ranges(matrix + step * order, matrix + (step + 1) * order, matrix + pivot * order);
                std::swap(rows[step], rows[pivot]);
            }
            const T diagonal = matrix[step * order + step];
            if (diagonal == T{0}) continue;
            const T *top = matrix + step * order;
            auto eliminate = [&](size_t row) {
                T *target = matrix + row * order;
                const T scale = target[step] /= diagonal;
                for (size_t column = step + 1; column < order; ++column) target[column] -= scale * top[column];
            };
            const size_t below = order - step - 1;
            if (below * (order - step) >= (size_t{1} << 16) && threads > 1) {
                const size_t parts = std::min(below, threads * 4);
                fork(parts, threads, [&](size_t part) {
                    const size_t first = step + 1 + below * part / parts;
                    const size_t last = step + 1 + below * (part + 1) / parts;
                    for (size_t row = first; row < last; ++row) eliminate(row);
                });
            } else {
                for (size_t row = step + 1; row < order; ++row) eliminate(row);
            }
        }
    }

    template <typename T>
    void substitute(const T *factors, const int64_t *rows, const T *right, T *solution, size_t order, size_t width) {
        for (size_t row = 0; row < order; ++row) {
            const T *source = right + static_cast<size_t>(rows[row]) * width;
            T *target = solution + row * width;
            for (size_t c = 0; c < width; ++c) target[c] = source[c];
            for (size_t inner = 0; inner < row; ++inner) {
                const T scale = factors[row * order + inner];
                const T *known = solution + inner * width;
                for (size_t c = 0; c < width; ++c) target[c] -= scale * known[c];
            }
        }
        for (size_t row = order; row-- > 0;) {
            T *target = solution + row * width;
            for (size_t inner = row + 1; inner < order; ++inner) {
                const T scale = factors[row * order + inner];
                const T *known = solution + inner * width;
                for (size_t c = 0; c < width; ++c) target[c] -= scale * known[c];
            }
            const T diagonal = factors[row * order + row];
            for (size_t c = 0; c < width; ++c) target[c] /= diagonal;
        }
    }

    // ---------------------------------------------------------------- synthetic data

    inline uint64_t mix(uint64_t seed, uint64_t index) { // SplitMix64 finaliser of seed * phi + index
        uint64_t z = seed * 0x9E3779B97F4A7C15ull + index;
        z ^= z >> 30;
        z *= 0xBF58476D1CE4E5B9ull;
        z ^= z >> 27;
        z *= 0x94D049BB133111EBull;
        z ^= z >> 31;
        return z;
    }

    size_t pack_width(int isa, bool single) { // avx512: 16 / 8 lanes; avx2: 8 / 4; fallback: 1
        if (isa == 2) return single ? 16 : 8;
        if (isa == 1) return single ? 8 : 4;
        return 1;
    }

    // d y / d operand k of y = code(a, b, c), as closed forms in double at the forward values.
    // minimum / maximum split a tie evenly between the operands; |x| has derivative sign(x), 0 at 0.
    void partials(int code, double a, double b, double y, double d[3]) {
        d[0] = d[1] = d[2] = 0.0;
        switch (code) {
        case SCIHIP_COPY: d[0] = 1.0; break;
        case SCIHIP_NEGATE: d[0] = -1.0; break;
        case SCIHIP_ABSOLUTE: d[0] = a > 0 ? 1.0 : a < 0 ? -1.0 : 0.0; break;
        case SCIHIP_ROOT: d[0] = 0.5 / std::sqrt(a); break;
        case SCIHIP_EXPONENTIAL: d[0] = std::exp(a); break;
        case SCIHIP_LOGARITHM: d[0] = 1.0 / a; break;
        case SCIHIP_SIN: d[0] = std::cos(a); break;
        case SCIHIP_COS: d[0] = -std::sin(a); break;
        case SCIHIP_TAN: d[0] = 1.0 / (std::cos(a) * std::cos(a)); break;
        case SCIHIP_ASIN: d[0] = 1.0 / std::sqrt(1.0 - a * a); break;
        case SCIHIP_ACOS: d[0] = -1.0 / std::sqrt(1.0 - a * a); break;
        case SCIHIP_ATAN: d[0] = 1.0 / (1.0 + a * a); break;
        case SCIHIP_CSC: d[0] = -std::cos(a) / (std::sin(a) * std::sin(a)); break;
        case SCIHIP_SEC: d[0] = std::sin(a) / (std::cos(a) * std::cos(a)); break;
        case SCIHIP_COT: d[0] = -1.0 / (std::sin(a) * std::sin(a)); break;
        case SCIHIP_ACSC: d[0] = -1.0 / (std::fabs(a) * std::sqrt(a * a - 1.0)); break;
        case SCIHIP_ASEC: d[0] = 1.0 / (std::fabs(a) * std::sqrt(a * a - 1.0)); break;
        case SCIHIP_ACOT: d[0] = -1.0 / (1.0 + a * a); break;
        case SCIHIP_SINH: d[0] = std::cosh(a); break;
        case SCIHIP_COSH: d[0] = std::sinh(a); break;
        case SCIHIP_TANH: d[0] = 1.0 / (std::cosh(a) * std::cosh(a)); break;
        case SCIHIP_ASINH: d[0] = 1.0 / std::sqrt(a * a + 1.0); break;
        case SCIHIP_ACOSH: d[0] = 1.0 / std::sqrt(a * a - 1.0); break;
        case SCIHIP_ATANH: d[0] = 1.0 / (1.0 - a * a); break;
        case SCIHIP_CSCH: d[0] = -std::cosh(a) / (std::sinh(a) * std::sinh(a)); break;
        case SCIHIP_SECH: d[0] = -std::sinh(a) / (std::cosh(a) * std::cosh(a)); break;
        case SCIHIP_COTH: d[0] = -1.0 / (std::sinh(a) * std::sinh(a)); break;
        case SCIHIP_ACSCH: d[0] = -1.0 / (std::fabs(a) * std::sqrt(1.0 + a * a)); break;
        case SCIHIP_ASECH: d[0] = -1.0 / (a * std::sqrt(1.0 - a * a)); break;
        case SCIHIP_ACOTH: d[0] = 1.0 / (1.0 - a * a); break;
        case SCIHIP_ADD: d[0] = 1.0; d[1] = 1.0; break;
        case SCIHIP_SUBTRACT: d[0] = 1.0; d[1] = -1.0; break;
        case SCIHIP_MULTIPLY: d[0] = b; d[1] = a; break;
        case SCIHIP_DIVIDE: d[0] = 1.0 / b; d[1] = -a / (b * b); break;
        case SCIHIP_MINIMUM:
        case SCIHIP_MAXIMUM: {
            const bool first = code == SCIHIP_MINIMUM ? a < b : a > b;
            const bool second = code == SCIHIP_MINIMUM ? b < a : b > a;
            d[0] = first ? 1.0 : a == b ? 0.5 : second ? 0.0 : 1.0;
            d[1] = second ? 1.0 : a == b ? 0.5 : first ? 0.0 : 1.0;
            break;
        }
        default: d[0] = b; d[1] = a; d[2] = 1.0; break; // fuse
        }
        (void)y;
    }

    template <typename T>
    void adjoint(const scihip_chunk &chunk, size_t count, const T *seed, T *const *gradients, double *loss) {
        const int length = chunk.length;
        std::vector<T> value(size_t(length) + 1);       // forward value of each instruction
        std::vector<double> bar(size_t(length) + 1);    // adjoint of each instruction's value
        std::vector<double> input_bar(chunk.arguments);
        // Which instruction (or input / constant) each operand of each instruction reads.
        enum { kFromInput, kFromLiteral, kFromStep };
        struct Source { int kind = kFromLiteral, index = 0; };
        std::vector<std::array<Source, 3>> sources(static_cast<size_t>(length));
        Source temporaries[SCIHIP_CAPACITY], output;
        auto resolve = [&](uint8_t id) {
            const uint8_t tag = id & SCIHIP_SLOT_TYPES, index = id & 0x3F;
            if (tag == SCIHIP_SLOT_INPUT) return Source{kFromInput, index};
            if (tag == SCIHIP_SLOT_LITERAL) return Source{kFromLiteral, index};
            if (tag == SCIHIP_SLOT_TEMPORARY) return temporaries[index];
            return output;
        };
        for (int s = 0; s < length; ++s) {
            const scihip_instruction &in = chunk.instructions[s];
            for (int k = 0; k < 3; ++k)
                sources[size_t(s)][size_t(k)] = in.sources[k] == SCIHIP_SLOT_NONE ? Source{} : resolve(in.sources[k]);
            if (in.target == SCIHIP_SLOT_OUTPUT) output = Source{kFromStep, s};
            else temporaries[in.target & 0x3F] = Source{kFromStep, s};
        }
        const Source result = resolve(chunk.result);
        const bool contract = !seed && contracted(chunk, SCIHIP_FOLD_SUM);
        double total = 0.0;
        for (size_t e = 0; e < count; ++e) {
            auto read = [&](const Source &src) -> T {
                if (src.kind == kFromInput) return static_cast<const T *>(chunk.inputs[src.index])[e];
                if (src.kind == kFromLiteral) return static_cast<T>(chunk.constants[src.index]);
                return value[size_t(src.index)];
            };
            for (int s = 0; s < length; ++s) {
                const auto &src = sources[size_t(s)];
                value[size_t(s)] = apply<T>(chunk.instructions[s].code, read(src[0]), read(src[1]), read(src[2]));
            }
            if (contract) {
                const auto &last = sources[size_t(length - 1)];
                total += static_cast<double>(read(last[0])) * static_cast<double>(read(last[1]));
            } else {
                total += static_cast<double>(read(result));
            }
            std::fill(bar.begin(), bar.end(), 0.0);
            std::fill(input_bar.begin(), input_bar.end(), 0.0);
            const double g0 = seed ? static_cast<double>(seed[e]) : 1.0;
            if (result.kind == kFromStep) bar[size_t(result.index)] = g0;
            else if (result.kind == kFromInput) input_bar[size_t(result.index)] += g0;
            for (int s = length - 1; s >= 0; --s) {
                const double g = bar[size_t(s)];
                if (g == 0.0) continue;
                const auto &src = sources[size_t(s)];
                double d[3];
                partials(chunk.instructions[s].code, static_cast<double>(read(src[0])),
                         static_cast<double>(read(src[1])), static_cast<double>(value[size_t(s)]), d);
                for (int k = 0; k < 3; ++k) {
                    if (chunk.instructions[s].sources[k] == SCIHIP_SLOT_NONE) continue;
                    if (src[size_t(k)].kind == kFromStep) bar[size_t(src[size_t(k)].index)] += g * d[k];
                    else if (src[size_t(k)].kind == kFromInput) input_bar[size_t(src[size_t(k)].index)] += g * d[k];
                }
            }
            for (int i = 0; i < chunk.arguments; ++i)
                if (gradients[i]) gradients[i][e] = static_cast<T>(input_bar[size_t(i)]);
        }
        if (loss) *loss = total;
    }

} // namespace

extern "C" {

// The gradient of sum(result) (seed null) or sum(seed * result) of a chunk with respect to each input
// whose gradients[i] is non-null, and the loss sum(result) (contracted as a dot product when the
// chunk ends in a multiply into its result and no seed is given). No reference counterpart.
int oracle_adjoint(const scihip_chunk *chunk, size_t count, const void *seed, void *const *gradients, double *loss) {
    if (chunk->precision == SCIHIP_SINGLE)
        adjoint<float>(*chunk, count, static_cast<const float *>(seed), reinterpret_cast<float *const *>(gradients),
                       loss);
    else
        adjoint<double>(*chunk, count, static_cast<const double *>(seed),
                        reinterpret_cast<double *const *>(gradients), loss);
    return 0;
}

int oracle_execute(const scihip_chunk *chunk, void *output, size_t count, int threads) {
    const size_t t = threads > 0 ? static_cast<size_t>(threads) : 1;
    if (chunk->precision == SCIHIP_SINGLE) evaluate<float>(*chunk, static_cast<float *>(output), count, t);
    else evaluate<double>(*chunk, static_cast<double *>(output), count, t);
    return 0;
}

double oracle_reduce(const scihip_chunk *chunk, int fold, size_t count, int threads, int isa) {
    const size_t t = threads > 0 ? static_cast<size_t>(threads) : 1;
    if (chunk->precision == SCIHIP_SINGLE) return accumulate<float>(*chunk, fold, count, t, pack_width(isa, true));
    return accumulate<double>(*chunk, fold, count, t, pack_width(isa, false));
}

uint32_t oracle_cost(const scihip_chunk *chunk) { return cost(*chunk); }

float oracle_apply_f32(int code, float a, float b, float c) { return apply<float>(code, a, b, c); }
double oracle_apply_f64(int code, double a, double b, double c) { return apply<double>(code, a, b, c); }

void oracle_map_f32(int code, const float *a, const float *b, const float *c, float *out, size_t n, int threads) {
    const size_t t = threads > 0 ? static_cast<size_t>(threads) : 1;
    const size_t parts = t > 1 ? std::min<size_t>(t * 8, (n + block - 1) / block) : 1;
    fork(parts, t, [&](size_t p) {
        const size_t first = n * p / parts, last = n * (p + 1) / parts;
        for (size_t start = first; start < last; start += block) {
            const size_t m = std::min(block, last - start);
            run_code(code, a + start, b ? b + start : nullptr, c ? c + start : nullptr, out + start, m);
        }
    });
}

void oracle_product(int precision, const void *left, const void *right, void *output, size_t rows, size_t columns,
                    size_t depth, int threads) {
    const size_t t = threads > 0 ? static_cast<size_t>(threads) : 1;
    if (precision == SCIHIP_SINGLE)
        multiply(static_cast<const float *>(left), static_cast<const float *>(right), static_cast<float *>(output),
                 rows, columns, depth, t);
    else
        multiply(static_cast<const double *>(left), static_cast<const double *>(right), static_cast<double *>(output),
                 rows, columns, depth, t);
}

void oracle_factor(int precision, void *matrix, int64_t *rows, size_t order, int threads) {
    const size_t t = threads > 0 ? static_cast<size_t>(threads) : 1;
    if (precision == SCIHIP_SINGLE) factor(static_cast<float *>(matrix), rows, order, t);
    else factor(static_cast<double *>(matrix), rows, order, t);
}

void oracle_substitute(int precision, const void *factors, const int64_t *rows, const void *right, void *solution,
                       size_t order, size_t width) {
    if (precision == SCIHIP_SINGLE)
        substitute(static_cast<const float *>(factors), rows, static_cast<const float *>(right),
                   static_cast<float *>(solution), order, width);
    else
        substitute(static_cast<const double *>(factors), rows, static_cast<const double *>(right),
                   static_cast<double *>(solution), order, width);
}

// include/operations/creation.hpp:65-84: x_i = float(start + i * step), step = double(stop - start) / (count - 1).
void oracle_linspace_f32(float start, float stop, size_t count, float *out) {
    if (count == 0) return;
    if (count == 1) {
        out[0] = start;
        return;
    }
    const double step = static_cast<double>(stop - start) / (static_cast<double>(count) - 1.0);
    for (size_t i = 0; i < count; ++i)
        out[i] = static_cast<float>(static_cast<double>(start) + static_cast<double>(i) * step);
}

void oracle_fill_uniform(void *output, int precision, size_t count, uint64_t seed, double lower, double upper,
                         uint64_t nan_seed, double nan_fraction) {
    const uint64_t bound = static_cast<uint64_t>(nan_fraction * 4294967296.0);
    for (size_t i = 0; i < count; ++i) {
        const uint64_t h = mix(seed, i);
        const bool hole = nan_fraction > 0.0 && (mix(nan_seed, i) >> 32) < bound;
        if (precision == SCIHIP_SINGLE) {
            const float u = static_cast<float>(h >> 40) * 0x1p-24f;
            const float lo = static_cast<float>(lower), span = static_cast<float>(upper) - lo;
            static_cast<float *>(output)[i] = hole ? std::numeric_limits<float>::quiet_NaN() : lo + span * u;
        } else {
            const double u = static_cast<double>(h >> 11) * 0x1p-53;
            static_cast<double *>(output)[i] =
                    hole ? std::numeric_limits<double>::quiet_NaN() : lower + (upper - lower) * u;
        }
    }
}

} // extern "C"
