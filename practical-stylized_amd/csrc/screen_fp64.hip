// Screen-mode level kernels, fp64 record storage (double accumulation), radius 1..7.
#include "screen_impl.hpp"

namespace nlpf {

hipError_t screen_level_fp64(const ScreenArgs<double>& a, int radius, hipStream_t st) {
    switch (radius) {
        case 1: return launch_screen_level<double, double, 1, 4, 2>(a, st);
        case 2: return launch_screen_level<double, double, 2, 4, 2>(a, st);
        case 3: return launch_screen_level<double, double, 3, 4, 2>(a, st);
        case 4: return launch_screen_level<double, double, 4, 4, 2>(a, st);
        case 5: return launch_screen_level<double, double, 5, 4, 2>(a, st);
        case 6: return launch_screen_level<double, double, 6, 4, 2>(a, st);
        case 7: return launch_screen_level<double, double, 7, 4, 2>(a, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t screen_film_fp64(const double* out, uint32_t S, uint32_t out_levels, uint32_t slot0, uint64_t plane,
                          double* film, hipStream_t st) {
    const uint32_t bs = 256;
    const uint64_t nb = (plane + bs - 1) / bs;
    hipLaunchKernelGGL(screen_film_kernel<double>, dim3((uint32_t)nb), dim3(bs), 0, st, out, S, out_levels, slot0,
                       plane, film);
    return hipGetLastError();
}

}  // namespace nlpf
