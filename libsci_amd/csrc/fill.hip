// Seeded synthetic data generated in place on the device, the fold of reduction partials, and small
// helpers shared by the runtime.
#include "kernels.hpp"

namespace scihip {

    const void *interpreter_function_f32(int registers, int mode);
    const void *interpreter_function_f64(int registers, int mode);

    const void *interpreter_function(int precision, int registers, int mode) {
        return precision == SCIHIP_SINGLE ? interpreter_function_f32(registers, mode)
                                          : interpreter_function_f64(registers, mode);
    }

    __global__ __launch_bounds__(256) void fold_partials_kernel(uint32_t fold, const Partial *partials, uint32_t count,
                                                                double *result) {
        const Partial folded = fold_partials_block(fold, partials, count);
        if (threadIdx.x == 0) *result = folded.poisoned ? __builtin_nan("") : folded.value;
    }

    hipError_t fold_partials(uint32_t fold, const Partial *partials, unsigned count, double *result,
                             hipStream_t stream) {
        hipLaunchKernelGGL(fold_partials_kernel, dim3(1), dim3(256), 0, stream, fold, partials, count, result);
        return hipGetLastError();
    }

    // SplitMix64's finaliser of seed * phi + index: a counter-based hash, so any element can be made
    // on its own, on the CPU or the GPU, with the same bits.
    __device__ __forceinline__ uint64_t mix(uint64_t seed, uint64_t index) {
        uint64_t z = seed * 0x9E3779B97F4A7C15ull + index;
        z ^= z >> 30;
        z *= 0xBF58476D1CE4E5B9ull;
        z ^= z >> 27;
        z *= 0x94D049BB133111EBull;
        z ^= z >> 31;
        return z;
    }

    // Four consecutive elements per thread, one 16-byte store for float when aligned.
    template <typename T>
    __global__ __launch_bounds__(256) void uniform(T *output, uint64_t count, uint64_t seed, T lower, T span,
                                                   uint64_t nan_seed, uint64_t bound) {
        const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
            const uint64_t h = mix(seed, i);
            T x;
            if (sizeof(T) == 4) x = lower + span * (T)((float)(h >> 40) * 0x1p-24f);
            else x = lower + span * (T)((double)(h >> 11) * 0x1p-53);
            if (bound && (mix(nan_seed, i) >> 32) < bound) x = (T)__builtin_nan("");
            output[i] = x;
        }
    }

    hipError_t fill_uniform(void *output, int precision, size_t count, uint64_t seed, double lower, double upper,
                            uint64_t nan_seed, double nan_fraction, hipStream_t stream) {
        if (count == 0) return hipSuccess;
        const uint64_t bound = nan_fraction > 0.0 ? (uint64_t)(nan_fraction * 4294967296.0) : 0;
        const unsigned blocks = (unsigned)std::min<uint64_t>((count + 255) / 256, 8192);
        if (precision == SCIHIP_SINGLE) {
            const float lo = (float)lower, span = (float)upper - lo;
            hipLaunchKernelGGL(uniform<float>, dim3(blocks), dim3(256), 0, stream, (float *)output, (uint64_t)count,
                               seed, lo, span, nan_seed, bound);
        } else {
            hipLaunchKernelGGL(uniform<double>, dim3(blocks), dim3(256), 0, stream, (double *)output,
                               (uint64_t)count, seed, lower, upper - lower, nan_seed, bound);
        }
        return hipGetLastError();
    }

} // namespace scihip
