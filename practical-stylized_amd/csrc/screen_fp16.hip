// Screen-mode level kernels, fp16 record storage (float accumulation), radius 1..7,
// tile variants (waves per workgroup, rows per lane) selected at run time; variant -1
// takes the radius's default.
#include "screen_strip.hpp"

namespace nlpf {

// Variants: 4 waves per workgroup with 2 (variant 0, 64x8 tiles), 4 (1) or 3 (2) rows per
// lane, or column-strip segments (screen_strip.hpp) with 2 (3) or 3 (4) rows per lane. Default
// per radius: the fastest at 1920x1080x16 spp x 5 levels on MI355X (also timed and never
// fastest: 8 or 2 waves with 2 rows, 2 waves with 4). The strips win at every radius but two:
// FP32 records at radius 1 (3-row tiles, 1 % ahead; the strips win with FP16 records there) and
// at radius 5 (4-row tiles, 2 % ahead of the 3-row strips). The 2-row tiles run out of VGPRs at
// radius 5 and 6 (one wave per SIMD); the 3-row strips need 83 KB of LDS at radius 7 (one
// workgroup per CU).
static constexpr int kDefaultVariant[8] = {0, 3, 3, 3, 3, 4, 4, 3};
static constexpr int kTileVariant[8] = {0, 2, 0, 0, 0, 1, 2, 0};   // the fastest tile shape per radius
static constexpr int kStripVariant[8] = {3, 3, 3, 3, 3, 4, 4, 3};  // the fastest strip shape per radius

template <int RAD>
static hipError_t dispatch_fp16(const ScreenArgs<float>& a, int variant, hipStream_t st) {
    constexpr int K = 2 * RAD + 1;
    if (a.material) {  // per-vertex shading: 2-row strips, or 2-row 64x8 tiles for a tile variant
        const bool tiles = variant == kScreenTiles || (variant >= 0 && variant < kScreenStrips);
        if (!tiles && screen_strip_eligible<__half>(a)) return launch_screen_strip<__half, float, RAD, 4, 2, true>(a, st);
        return launch_screen_level<__half, float, RAD, 4, 2, true>(a, st);
    }
    if (variant == kScreenTiles) variant = kTileVariant[RAD];
    if (variant == kScreenStripsAuto) variant = kStripVariant[RAD];
    if (variant < 0) variant = kDefaultVariant[RAD];
    switch (variant) {
        case 1: return launch_screen_level<__half, float, RAD, 4, (4 < K ? 4 : K)>(a, st);
        case 2: return launch_screen_level<__half, float, RAD, 4, 3>(a, st);
        case kScreenStrips:  // column strips, 2 rows per lane; the 64x8 tiles where not eligible
            if (screen_strip_eligible<__half>(a)) return launch_screen_strip<__half, float, RAD, 4, 2>(a, st);
            return launch_screen_level<__half, float, RAD, 4, 2>(a, st);
        case kScreenStrips + 1:  // column strips, 3 rows per lane
            if (screen_strip_eligible<__half>(a)) return launch_screen_strip<__half, float, RAD, 4, 3>(a, st);
            return launch_screen_level<__half, float, RAD, 4, 3>(a, st);
        default: return launch_screen_level<__half, float, RAD, 4, 2>(a, st);
    }
}

int screen_variants_fp16() { return 5; }

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
