// Element functions of every chunk code for gfx950, shared by the bytecode interpreter and the
// kernels specialised per chunk at run time (hipRTC compiles this same text).
//
// exp, log, sin and cos follow libsci's own algorithms so that they give the CPU path's bits:
// the Cody-Waite / Taylor exponential (include/core/drivers/processor/kernels/generic/
// exponential.hpp:24-61), fdlibm's logarithm form (generic/logarithm.hpp:26-66) and the
// four-part quadrant reduction with odd / even Taylor polynomials (generic/trigonometry.hpp:29-80),
// with the constants of generic/constants.hpp:16-119 and the polynomials of
// generic/polynomial.hpp:17-61. Every step that the CPU rounds separately is a separate operation
// here (built with -ffp-contract=off); every fused step is __builtin_fma. Arguments beyond 2^20 in
// sin / cos, and the codes the CPU sends to the C library (asin ... atanh, generic/operation.hpp:45-87),
// use the device library (OCML) instead of glibc.
#pragma once

#ifndef SCIHIP_RTC
#include <hip/hip_runtime.h>
#include <stdint.h>
#else
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
#endif

namespace scihip {

    enum : uint32_t {
        kCopy, kNegate, kAbsolute, kRoot, kExponential, kLogarithm,
        kSin, kCos, kTan, kAsin, kAcos, kAtan, kCsc, kSec, kCot, kAcsc, kAsec, kAcot,
        kSinh, kCosh, kTanh, kAsinh, kAcosh, kAtanh, kCsch, kSech, kCoth, kAcsch, kAsech, kAcoth,
        kAdd, kSubtract, kMultiply, kDivide, kMinimum, kMaximum, kFuse, kCodes,
        // Codes that only adjoint programs use (never in a recorded chunk): a * sign(b); the share
        // of g that flows to the first operand of minimum(x, y) or maximum(x, y), half of it on a tie.
        kSignTimes = kCodes, kRouteMinimum, kRouteMaximum, kAdjointCodes
    };

    template <typename T> struct Limits;

    template <> struct Limits<float> {
        typedef uint32_t U;
        static constexpr int mantissa = 23;
        static constexpr int width = 32;
        static constexpr U fraction = 0x007fffffu;
        static constexpr U one = 0x3f800000u;
        static constexpr U sign = 0x80000000u;
        static constexpr float round = 12582912.0f;    // 1.5 * 2^23
        static constexpr float magnitude = 8388608.0f; // 2^23
        static constexpr float exponent = 8388735.0f;  // 2^23 + 127
        static constexpr float binary = 1.44269502162933349609375f; // log2(e) rounded to float
        static constexpr float natural = 0.693359375f;  // 355 / 512
        static constexpr float remainder = -2.12194440e-4f;
        static constexpr float lower = -104.0f;
        static constexpr float upper = 89.0f;
        static constexpr float normal = 126.0f;
        static constexpr float tiny = 1.17549435e-38f;
        static constexpr float lift = 33554432.0f; // 2^25
        static constexpr float shift = 25.0f;
        static constexpr float root = 1.41421353816986083984375f; // sqrt(2) rounded to float
        static constexpr float quadrant = 0.636619772367581343f;
        static constexpr float quarter0 = 1.5703125f;
        static constexpr float quarter1 = 4.837512969970703125e-4f;
        static constexpr float quarter2 = 7.549790126404332e-08f;
        static constexpr float quarter3 = -1.7151245100058819e-15f;
        static constexpr float reach = 1048576.0f;
        static constexpr int exp_terms = 8, log_terms = 4, sin_terms = 6, cos_terms = 7;
    };

    template <> struct Limits<double> {
        typedef uint64_t U;
        static constexpr int mantissa = 52;
        static constexpr int width = 64;
        static constexpr U fraction = 0x000fffffffffffffull;
        static constexpr U one = 0x3ff0000000000000ull;
        static constexpr U sign = 0x8000000000000000ull;
        static constexpr double round = 6755399441055744.0;
        static constexpr double magnitude = 4503599627370496.0;
        static constexpr double exponent = 4503599627371519.0;
        static constexpr double binary = 1.4426950408889634074;
        static constexpr double natural = 0x1.62e42feep-1; // ln 2 with its low 21 bits cleared
        static constexpr double remainder = 1.90821492927058770002e-10;
        static constexpr double lower = -746.0;
        static constexpr double upper = 710.0;
        static constexpr double normal = 1022.0;
        static constexpr double tiny = 2.2250738585072014e-308;
        static constexpr double lift = 18014398509481984.0; // 2^54
        static constexpr double shift = 54.0;
        static constexpr double root = 1.4142135623730950488;
        static constexpr double quadrant = 6.36619772367581382433e-01;
        static constexpr double quarter0 = 1.57079632673412561417e+00;
        static constexpr double quarter1 = 6.07710050630396597660e-11;
        static constexpr double quarter2 = 2.02226624871116645580e-21;
        static constexpr double quarter3 = 8.47842766036889956997e-32;
        static constexpr double reach = 1048576.0;
        static constexpr int exp_terms = 14, log_terms = 10, sin_terms = 10, cos_terms = 11;
    };

    __device__ __forceinline__ uint32_t bits(float v) { return __float_as_uint(v); }
    __device__ __forceinline__ uint64_t bits(double v) { return (uint64_t)__double_as_longlong(v); }
    __device__ __forceinline__ float value(uint32_t v) { return __uint_as_float(v); }
    __device__ __forceinline__ double value(uint64_t v) { return __longlong_as_double((long long)v); }
    __device__ __forceinline__ float fused(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
    __device__ __forceinline__ double fused(double a, double b, double c) { return __builtin_fma(a, b, c); }
    __device__ __forceinline__ bool unordered(float v) { return __builtin_isnan(v); }
    __device__ __forceinline__ bool unordered(double v) { return __builtin_isnan(v); }

    // Taylor coefficient tables, computed in double and rounded once to T, as the CPU's constexpr
    // Polynomial builders do (polynomial.hpp:17-50).
    template <typename T, int N> struct Series {
        T c[N];
    };

    template <typename T, int N> __device__ __forceinline__ constexpr Series<T, N> factorials() {
        Series<T, N> s{};
        double f = 1.0;
        for (int i = 0; i < N; ++i) {
            if (i > 0) f /= (double)i;
            s.c[i] = (T)f;
        }
        return s;
    }

    template <typename T, int N> __device__ __forceinline__ constexpr Series<T, N> reciprocals() {
        Series<T, N> s{};
        for (int i = 0; i < N; ++i) s.c[i] = (T)(2.0 / (double)(2 * i + 3));
        return s;
    }

    template <typename T, int N, int offset> __device__ __forceinline__ constexpr Series<T, N> alternating() {
        Series<T, N> s{};
        for (int i = 0; i < N; ++i) {
            double f = 1.0;
            for (int k = 2; k <= 2 * i + offset; ++k) f *= (double)k;
            s.c[i] = (T)((i % 2 == 0 ? 1.0 : -1.0) / f);
        }
        return s;
    }

    // sum c[k] x^(k - from) for k in [from, N), Horner's rule with fma (polynomial.hpp:55-61).
    template <int from, typename T, int N> __device__ __forceinline__ T horner(const Series<T, N> &s, T x) {
        T r = s.c[N - 1];
#pragma unroll
        for (int i = N - 2; i >= from; --i) r = fused(r, x, s.c[i]);
        return r;
    }

    // NaN in either gives NaN, as numpy's minimum / maximum (scalar.hpp:112-127).
    template <typename T> __device__ __forceinline__ T lesser(T l, T r) {
        return (unordered(l) || unordered(r)) ? l + r : (r < l ? r : l);
    }
    template <typename T> __device__ __forceinline__ T greater(T l, T r) {
        return (unordered(l) || unordered(r)) ? l + r : (r > l ? r : l);
    }

    // e^x (exponential.hpp:24-49). The clamp to [lower, upper] is two compare-selects that let a NaN
    // through unchanged, which gives the same values as the CPU's NaN-propagating minimum / maximum
    // (a NaN then flows through the arithmetic to the result).
    template <typename T> __device__ __forceinline__ T exponential(T x) {
        typedef Limits<T> L;
        constexpr Series<T, L::exp_terms> series = factorials<T, L::exp_terms>();
        const T below = x > L::upper ? L::upper : x;
        const T clamped = below < L::lower ? L::lower : below;
        const T k = fused(clamped, L::binary, L::round) - L::round;
        const T negative = -k;
        T r = fused(negative, L::natural, clamped);
        r = fused(negative, L::remainder, r);
        const T significand = fused(r, horner<1>(series, r), T(1));
        if (__builtin_expect(!(__builtin_fabs(k) > L::normal), 1))
            return significand * value((bits(T(k + L::exponent))) << L::mantissa);
        const T half = fused(k, T(0.5), L::round) - L::round;
        const T rest = k - half;
        return (significand * value(bits(T(half + L::exponent)) << L::mantissa)) *
               value(bits(T(rest + L::exponent)) << L::mantissa);
    }

    // f / d for the logarithm's f / (f + 2), f in [sqrt(1/2) - 1, sqrt(2) - 1]. For float this is the
    // correctly rounded division's own Newton-Raphson sequence (hardware reciprocal, one refinement
    // of it, two corrections of the quotient) without its range scaling and special-case fix-up, which
    // change nothing for operands in this range: the same bits as f / d with two fewer steps.
    template <typename T> __device__ __forceinline__ T ratio(T f, T d) { return f / d; }
    template <> __device__ __forceinline__ float ratio(float f, float d) {
        float y = __builtin_amdgcn_rcpf(d);
        y = fused(fused(-d, y, 1.0f), y, y);
        float q = f * y;
        q = fused(fused(-d, q, f), y, q);
        return fused(fused(-d, q, f), y, q);
    }

    // ln x (logarithm.hpp:26-66). `ordinary`: x is a positive normal number or NaN, so the CPU's
    // scaling of subnormals and its special cases (0, +inf, negative) do not apply, and a NaN is
    // returned as it came.
    template <bool ordinary, typename T> __device__ __forceinline__ T logarithm_of(T x) {
        typedef Limits<T> L;
        typedef typename L::U U;
        constexpr Series<T, L::log_terms> series = reciprocals<T, L::log_terms>();
        const bool small = !ordinary && x < L::tiny;
        const T scaled = small ? x * L::lift : x;
        const U pattern = bits(scaled);
        T exponent = value((U)((pattern >> L::mantissa) | bits(L::magnitude))) - L::exponent;
        T mantissa = value((U)((pattern & L::fraction) | L::one));
        const bool large = mantissa > L::root;
        mantissa = large ? mantissa * T(0.5) : mantissa;
        exponent = large ? exponent + T(1) : exponent;
        exponent = small ? exponent - L::shift : exponent;
        const T f = mantissa - T(1);
        const T s = ordinary ? ratio(f, f + T(2)) : f / (f + T(2));
        const T z = s * s;
        const T tail = z * horner<0>(series, z);
        const T half = (f * f) * T(0.5);
        const T inner = fused(s, half + tail, exponent * L::remainder);
        T result = exponent * L::natural - ((half - inner) - f);
        if (!ordinary) {
            const T infinity = T(__builtin_huge_val());
            result = x == infinity ? x : result;
            result = x == T(0) ? -infinity : result;
            result = x < T(0) ? T(__builtin_nan("")) : result;
        }
        return unordered(x) ? x : result;
    }

    // Floating-point classes for __builtin_amdgcn_class: signalling and quiet NaN, positive normal.
    constexpr int kNanOrPositiveNormal = 0x001 | 0x002 | 0x100;
    __device__ __forceinline__ bool nan_or_positive_normal(float x) {
        return __builtin_amdgcn_classf(x, kNanOrPositiveNormal);
    }
    __device__ __forceinline__ bool nan_or_positive_normal(double x) {
        return __builtin_amdgcn_class(x, kNanOrPositiveNormal);
    }

    template <typename T> __device__ __forceinline__ T logarithm(T x) {
        if (__builtin_expect(nan_or_positive_normal(x), 1)) return logarithm_of<true>(x);
        return logarithm_of<false>(x);
    }

    template <typename T> __device__ __forceinline__ T library_sin(T x);
    template <> __device__ __forceinline__ float library_sin(float x) { return ::sinf(x); }
    template <> __device__ __forceinline__ double library_sin(double x) { return ::sin(x); }
    template <typename T> __device__ __forceinline__ T library_cos(T x);
    template <> __device__ __forceinline__ float library_cos(float x) { return ::cosf(x); }
    template <> __device__ __forceinline__ double library_cos(double x) { return ::cos(x); }

    // Sine (phase false) or cosine (phase true) (trigonometry.hpp:29-80).
    template <bool phase, typename T> __device__ __forceinline__ T trigonometry(T x) {
        typedef Limits<T> L;
        typedef typename L::U U;
        constexpr Series<T, L::sin_terms> odd = alternating<T, L::sin_terms, 1>();
        constexpr Series<T, L::cos_terms> even = alternating<T, L::cos_terms, 0>();
        const T shifted = fused(x, L::quadrant, L::round);
        const T quarters = shifted - L::round;
        const T negative = -quarters;
        T r = fused(negative, L::quarter0, x);
        r = fused(negative, L::quarter1, r);
        r = fused(negative, L::quarter2, r);
        r = fused(negative, L::quarter3, r);
        const T square = r * r;
        const T sine = fused(r * square, horner<1>(odd, square), r);
        const T cosine = fused(square, horner<1>(even, square), T(1));
        U quadrant = bits(shifted);
        if (phase) quadrant = quadrant ^ (quadrant << 1);
        const U flip = (quadrant & U(2)) << (L::width - 2);
        const bool swap = (bits(shifted) & U(1)) != 0;
        const T picked = phase ? (swap ? sine : cosine) : (swap ? cosine : sine);
        T result = value((U)(bits(picked) ^ flip));
        if (__builtin_expect(!(x <= L::reach && x >= -L::reach), 0))
            result = phase ? library_cos(x) : library_sin(x);
        return result;
    }

    template <typename T> __device__ __forceinline__ T library(uint32_t code, T x);

    template <> __device__ __forceinline__ float library(uint32_t code, float x) {
        switch (code) {
        case kAsin: return ::asinf(x);
        case kAcos: return ::acosf(x);
        case kAtan: return ::atanf(x);
        case kSinh: return ::sinhf(x);
        case kCosh: return ::coshf(x);
        case kTanh: return ::tanhf(x);
        case kAsinh: return ::asinhf(x);
        case kAcosh: return ::acoshf(x);
        default: return ::atanhf(x);
        }
    }

    template <> __device__ __forceinline__ double library(uint32_t code, double x) {
        switch (code) {
        case kAsin: return ::asin(x);
        case kAcos: return ::acos(x);
        case kAtan: return ::atan(x);
        case kSinh: return ::sinh(x);
        case kCosh: return ::cosh(x);
        case kTanh: return ::tanh(x);
        case kAsinh: return ::asinh(x);
        case kAcosh: return ::acosh(x);
        default: return ::atanh(x);
        }
    }

    template <typename T> __device__ __forceinline__ T root(T x);
    template <> __device__ __forceinline__ float root(float x) { return __builtin_sqrtf(x); }
    template <> __device__ __forceinline__ double root(double x) { return __builtin_sqrt(x); }

    // What one code does to one element (generic/operation.hpp:25-110). `code` is wave-uniform in
    // the interpreter and a compile-time constant in specialised kernels.
    template <typename T> __device__ __forceinline__ T apply(uint32_t code, T a, T b, T c) {
        typedef typename Limits<T>::U U;
        const T one = T(1);
        switch (code) {
        case kCopy: return a;
        case kNegate: return value((U)(bits(a) ^ Limits<T>::sign));
        case kAbsolute: return value((U)(bits(a) & ~Limits<T>::sign));
        case kRoot: return root(a);
        case kExponential: return exponential(a);
        case kLogarithm: return logarithm(a);
        case kSin: return trigonometry<false>(a);
        case kCos: return trigonometry<true>(a);
        case kTan: return trigonometry<false>(a) / trigonometry<true>(a);
        case kAsin: case kAcos: case kAtan: case kSinh: case kCosh: case kTanh:
        case kAsinh: case kAcosh: case kAtanh:
            return library<T>(code, a);
        case kCsc: return one / trigonometry<false>(a);
        case kSec: return one / trigonometry<true>(a);
        case kCot: return one / (trigonometry<false>(a) / trigonometry<true>(a));
        case kAcsc: return library<T>(kAsin, one / a);
        case kAsec: return library<T>(kAcos, one / a);
        case kAcot: return library<T>(kAtan, one / a);
        case kCsch: return one / library<T>(kSinh, a);
        case kSech: return one / library<T>(kCosh, a);
        case kCoth: return one / library<T>(kTanh, a);
        case kAcsch: return library<T>(kAsinh, one / a);
        case kAsech: return library<T>(kAcosh, one / a);
        case kAcoth: return library<T>(kAtanh, one / a);
        case kAdd: return a + b;
        case kSubtract: return a - b;
        case kMultiply: return a * b;
        case kDivide: return a / b;
        case kMinimum: return lesser(a, b);
        case kMaximum: return greater(a, b);
        case kSignTimes: return a * (b > T(0) ? one : (b < T(0) ? -one : b));
        case kRouteMinimum: return a * (b < c ? one : (b == c ? T(0.5) : (b > c ? T(0) : one)));
        case kRouteMaximum: return a * (b > c ? one : (b == c ? T(0.5) : (b < c ? T(0) : one)));
        default: return fused(a, b, c);
        }
    }

} // namespace scihip
