// Interpreter kernels over float elements (see interpreter.hpp).
#include "kernels.hpp"

namespace scihip {

    hipError_t interpret(const Launch<float> &launch, int registers, int mode, unsigned blocks, hipStream_t stream) {
        constexpr int W = Traits<float>::width;
        const dim3 grid(blocks), block(256);
        if (mode == 0) {
            if (registers == kSmall) hipLaunchKernelGGL((evaluate<float, W, kSmall>), grid, block, 0, stream, launch);
            else if (registers == kMedium) hipLaunchKernelGGL((evaluate<float, W, kMedium>), grid, block, 0, stream, launch);
            else hipLaunchKernelGGL((evaluate<float, 1, kLarge>), grid, block, 0, stream, launch);
        } else {
            if (registers == kSmall) hipLaunchKernelGGL((reduce<float, W, kSmall>), grid, block, 0, stream, launch);
            else if (registers == kMedium) hipLaunchKernelGGL((reduce<float, W, kMedium>), grid, block, 0, stream, launch);
            else hipLaunchKernelGGL((reduce<float, 1, kLarge>), grid, block, 0, stream, launch);
        }
        return hipGetLastError();
    }

    const void *interpreter_function_f32(int registers, int mode) {
        constexpr int W = Traits<float>::width;
        if (mode == 0) {
            if (registers == kSmall) return reinterpret_cast<const void *>(&evaluate<float, W, kSmall>);
            if (registers == kMedium) return reinterpret_cast<const void *>(&evaluate<float, W, kMedium>);
            return reinterpret_cast<const void *>(&evaluate<float, 1, kLarge>);
        }
        if (registers == kSmall) return reinterpret_cast<const void *>(&reduce<float, W, kSmall>);
        if (registers == kMedium) return reinterpret_cast<const void *>(&reduce<float, W, kMedium>);
        return reinterpret_cast<const void *>(&reduce<float, 1, kLarge>);
    }

} // namespace scihip
