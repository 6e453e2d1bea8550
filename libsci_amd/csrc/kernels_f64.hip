// Interpreter kernels over double elements (see interpreter.hpp).
#include "kernels.hpp"

namespace scihip {

    hipError_t interpret(const Launch<double> &launch, int registers, int mode, unsigned blocks, hipStream_t stream) {
        constexpr int W = Traits<double>::width;
        const dim3 grid(blocks), block(256);
        if (mode == 0) {
            if (registers == kSmall) hipLaunchKernelGGL((evaluate<double, W, kSmall>), grid, block, 0, stream, launch);
            else if (registers == kMedium) hipLaunchKernelGGL((evaluate<double, W, kMedium>), grid, block, 0, stream, launch);
            else hipLaunchKernelGGL((evaluate<double, 1, kLarge>), grid, block, 0, stream, launch);
        } else {
            if (registers == kSmall) hipLaunchKernelGGL((reduce<double, W, kSmall>), grid, block, 0, stream, launch);
            else if (registers == kMedium) hipLaunchKernelGGL((reduce<double, W, kMedium>), grid, block, 0, stream, launch);
            else hipLaunchKernelGGL((reduce<double, 1, kLarge>), grid, block, 0, stream, launch);
        }
        return hipGetLastError();
    }

    const void *interpreter_function_f64(int registers, int mode) {
        constexpr int W = Traits<double>::width;
        if (mode == 0) {
            if (registers == kSmall) return reinterpret_cast<const void *>(&evaluate<double, W, kSmall>);
            if (registers == kMedium) return reinterpret_cast<const void *>(&evaluate<double, W, kMedium>);
            return reinterpret_cast<const void *>(&evaluate<double, 1, kLarge>);
        }
        if (registers == kSmall) return reinterpret_cast<const void *>(&reduce<double, W, kSmall>);
        if (registers == kMedium) return reinterpret_cast<const void *>(&reduce<double, W, kMedium>);
        return reinterpret_cast<const void *>(&reduce<double, 1, kLarge>);
    }

} // namespace scihip
