// Screen-mode level kernels, fp64 record storage (double accumulation), radius 1..7,
// tile variants (waves per workgroup, rows per lane) selected at run time.
#include "screen_impl.hpp"

namespace nlpf {

template <int RAD>
static hipError_t dispatch_fp64(const ScreenArgs<double>& a, int variant, hipStream_t st) {
    constexpr int K = 2 * RAD + 1;
    if (a.material) return launch_screen_level<double, double, RAD, 4, (2 < K ? 2 : K), true>(a, st);
    switch (variant) {
        case 0: return launch_screen_level<double, double, RAD, 4, (2 < K ? 2 : K)>(a, st);
        default: return launch_screen_level<double, double, RAD, 4, (2 < K ? 2 : K)>(a, st);
    }
}

int screen_variants_fp64() { return 1; }

hipError_t screen_level_fp64(const ScreenArgs<double>& a, int radius, int variant, hipStream_t st) {
    switch (radius) {
        case 1: return dispatch_fp64<1>(a, variant, st);
        case 2: return dispatch_fp64<2>(a, variant, st);
        case 3: return dispatch_fp64<3>(a, variant, st);
        case 4: return dispatch_fp64<4>(a, variant, st);
        case 5: return dispatch_fp64<5>(a, variant, st);
        case 6: return dispatch_fp64<6>(a, variant, st);
        case 7: return dispatch_fp64<7>(a, variant, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t screen_film_fp64(const double* out, uint32_t S, uint32_t out_levels, uint32_t slot0, uint64_t plane,
                          double* film, hipStream_t st, uint32_t s_total, int accumulate, int finalize) {
    const uint32_t bs = 256;
    const uint64_t nb = (plane + bs - 1) / bs;
    hipLaunchKernelGGL(screen_film_kernel<double>, dim3((uint32_t)nb), dim3(bs), 0, st, out, S, out_levels, slot0,
                       plane, film, s_total ? s_total : S, accumulate, finalize);
    return hipGetLastError();
}

}  // namespace nlpf
