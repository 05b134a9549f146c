// Screen-mode level kernels, fp32 record storage (float accumulation), radius 1..7.
#include "screen_impl.hpp"

namespace nlpf {

hipError_t screen_level_fp32(const ScreenArgs<float>& a, int radius, hipStream_t st) {
    switch (radius) {
        case 1: return launch_screen_level<float, float, 1, 4, 4>(a, st);
        case 2: return launch_screen_level<float, float, 2, 4, 4>(a, st);
        case 3: return launch_screen_level<float, float, 3, 4, 4>(a, st);
        case 4: return launch_screen_level<float, float, 4, 4, 4>(a, st);
        case 5: return launch_screen_level<float, float, 5, 4, 4>(a, st);
        case 6: return launch_screen_level<float, float, 6, 4, 4>(a, st);
        case 7: return launch_screen_level<float, float, 7, 4, 4>(a, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t screen_film_fp32(const float* out, uint32_t S, uint32_t out_levels, uint32_t slot0, uint64_t plane,
                          float* film, hipStream_t st) {
    const uint32_t bs = 256;
    const uint64_t nb = (plane + bs - 1) / bs;
    hipLaunchKernelGGL(screen_film_kernel<float>, dim3((uint32_t)nb), dim3(bs), 0, st, out, S, out_levels, slot0,
                       plane, film);
    return hipGetLastError();
}

}  // namespace nlpf
