// Screen-mode level kernels, fp16 record storage (float accumulation), radius 1..7,
// tile variants (waves per workgroup, rows per lane) selected at run time.
#include "screen_impl.hpp"

namespace nlpf {

template <int RAD>
static hipError_t dispatch_fp16(const ScreenArgs<float>& a, int variant, hipStream_t st) {
    constexpr int K = 2 * RAD + 1;
    switch (variant) {
        case 1: return launch_screen_level<__half, float, RAD, 4, (4 < K ? 4 : K)>(a, st);
        case 0: return launch_screen_level<__half, float, RAD, 4, (2 < K ? 2 : K)>(a, st);
        case 2: return launch_screen_level<__half, float, RAD, 8, (2 < K ? 2 : K)>(a, st);
        case 3: return launch_screen_level<__half, float, RAD, 2, (4 < K ? 4 : K)>(a, st);
        default: return launch_screen_level<__half, float, RAD, 4, (2 < K ? 2 : K)>(a, st);
    }
}

int screen_variants_fp16() { return 4; }

hipError_t screen_level_fp16(const ScreenArgs<float>& a, int radius, int variant, hipStream_t st) {
    switch (radius) {
        case 1: return dispatch_fp16<1>(a, variant, st);
        case 2: return dispatch_fp16<2>(a, variant, st);
        case 3: return dispatch_fp16<3>(a, variant, st);
        case 4: return dispatch_fp16<4>(a, variant, st);
        case 5: return dispatch_fp16<5>(a, variant, st);
        case 6: return dispatch_fp16<6>(a, variant, st);
        case 7: return dispatch_fp16<7>(a, variant, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace nlpf
