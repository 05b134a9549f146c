// Film accumulation and resolve (crates/akari_render/src/film.rs).
//
// A film is one device buffer of 7*W*H floats laid out as the reference's Film::data
// (film.rs:66-67, 88-93, 108-110): | pixels: rgb per pixel | splat: rgb per pixel | weights |.
// The reference adds samples with float atomics (film.rs:180-184, 210-219), so its sums depend
// on scheduling. Here every pixel's samples are summed in input order instead: a stable radix
// sort groups the samples by pixel, and one thread per pixel run adds them to the film in order.
// The result is then the sequential sum, bit for bit, on every run.
#include <hipcub/hipcub.hpp>
#include "common.hpp"

namespace nlpf {

namespace {

constexpr uint32_t kBlock = 256;

inline dim3 grid_for(uint64_t n) { return dim3((uint32_t)((n + kBlock - 1) / kBlock)); }

__global__ void film_fill_kernel(float* p, uint64_t n, float v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void film_zero_u64_kernel(unsigned long long* p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0ull;
}

// Film::linear_index (film.rs:150-160): floor of the sample position. The reference only
// asserts in debug builds; a sample outside the film (or at a NaN position) gets the key
// W*H, which sorts after every pixel and is counted as dropped.
__global__ void film_keys_kernel(const float* xy, uint64_t n, uint32_t W, uint32_t H, uint32_t* keys,
                                 uint32_t* idx, unsigned long long* dropped) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = floorf(xy[2 * i]), y = floorf(xy[2 * i + 1]);
    const bool in = x >= 0.0f && y >= 0.0f && x < (float)W && y < (float)H;  // false for NaN
    keys[i] = in ? (uint32_t)x + (uint32_t)y * W : W * H;
    idx[i] = (uint32_t)i;
    if (!in) atomicAdd(dropped, 1ull);
}

// Color::remove_nan (color.rs:319-327): a colour with any NaN channel becomes zero.
__device__ inline void remove_nan(float& r, float& g, float& b) {
    if (isnan(r) || isnan(g) || isnan(b)) r = g = b = 0.0f;
}

// One thread per run of equal keys in the sorted order: the run's first thread adds the
// run's samples, in input order, onto the film's current values.
//   SPLAT = 0: Film::add_sample (film.rs:187-222): pixels += remove_nan(c) * w, weights += w
//   SPLAT = 1: Film::add_splat (film.rs:162-186): splat += remove_nan(c) * w with a NaN channel
//              of the product replaced by zero
template <int SPLAT>
__global__ void film_runs_kernel(const uint32_t* keys, const uint32_t* idx, uint64_t n, uint32_t npix,
                                 const float* rgb, const float* weight, float* film) {
#pragma clang fp contract(off)
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t k = keys[s];
    if (k >= npix || (s > 0 && keys[s - 1] == k)) return;
    float* px = film + (SPLAT ? 3ull * npix : 0ull) + 3ull * k;
    float ar = px[0], ag = px[1], ab = px[2];
    float aw = SPLAT ? 0.0f : film[6ull * npix + k];
    for (uint64_t j = s; j < n && keys[j] == k; ++j) {
        const uint64_t i = idx[j];
        float r = rgb[3 * i], g = rgb[3 * i + 1], b = rgb[3 * i + 2];
        const float w = weight ? weight[i] : 1.0f;
        remove_nan(r, g, b);
        r = r * w;
        g = g * w;
        b = b * w;
        if (SPLAT) {
            if (isnan(r)) r = 0.0f;
            if (isnan(g)) g = 0.0f;
            if (isnan(b)) b = 0.0f;
        } else {
            aw = aw + w;
        }
        ar = ar + r;
        ag = ag + g;
        ab = ab + b;
    }
    px[0] = ar;
    px[1] = ag;
    px[2] = ab;
    if (!SPLAT) film[6ull * npix + k] = aw;
}

// linear_to_srgb (color.rs:536-539)
__device__ inline float linear_to_srgb1(float v) {
#pragma clang fp contract(off)
    return v <= 0.0031308f ? v * 12.92f : powf(v, 1.0f / 2.4f) * 1.055f - 0.055f;
}

// The copy_to_rgba_image kernel (film.rs:120-147): pixels / weight (a zero weight divides by
// one) plus splat * splat_scale, optionally sRGB-encoded, alpha 1.
__global__ void film_resolve_kernel(const float* film, uint32_t npix, float splat_scale, int hdr, float4* rgba) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    const float* p = film + 3ull * i;
    const float* s = film + 3ull * npix + 3ull * i;
    const float w = film[6ull * npix + i];
    const float d = w == 0.0f ? 1.0f : w;
    float r = p[0] / d + s[0] * splat_scale;
    float g = p[1] / d + s[1] * splat_scale;
    float b = p[2] / d + s[2] * splat_scale;
    if (!hdr) {
        r = linear_to_srgb1(r);
        g = linear_to_srgb1(g);
        b = linear_to_srgb1(b);
    }
    rgba[i] = make_float4(r, g, b, 1.0f);
}

int film_check(nlpf_handle* h, const float* film, uint32_t W, uint32_t H, const char* what) {
    if (!film) NLPF_FAIL(h, NLPF_E_INVALID_ARG, std::string(what) + ": NULL film");
    if (W == 0 || H == 0) NLPF_FAIL(h, NLPF_E_INVALID_ARG, std::string(what) + ": width and height must be >= 1");
    if ((uint64_t)W * H >= (1ull << 31))
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, std::string(what) + ": width * height must be < 2^31");
    return NLPF_OK;
}

int film_add(nlpf_handle* h, float* film, uint32_t W, uint32_t H, const float* xy, const float* rgb,
             const float* weight, uint64_t n, uint64_t* dropped, bool splat, const char* what) {
    if (!h) return NLPF_E_INVALID_ARG;
    if (int rc = film_check(h, film, W, H, what)) return rc;
    if (n >= (1ull << 31)) NLPF_FAIL(h, NLPF_E_INVALID_ARG, std::string(what) + ": more than 2^31 - 1 samples");
    if (n > 0 && (!xy || !rgb)) NLPF_FAIL(h, NLPF_E_INVALID_ARG, std::string(what) + ": NULL samples");
    if (dropped) *dropped = 0;
    if (n == 0) return NLPF_OK;
    const uint32_t npix = W * H;
    int bits = 1;
    while (bits < 32 && (1ull << bits) <= npix) ++bits;  // keys are 0..npix
    hipStream_t st = h->stream;
    size_t cub_bytes = 0;
    NLPF_HIP(h, hipcub::DeviceRadixSort::SortPairs(nullptr, cub_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                                   (uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, bits, st));
    const size_t a = (4 * n + 63) & ~size_t(63);
    if (int rc = ensure_scratch(h, 4 * a + 256 + cub_bytes)) return rc;
    char* base = (char*)h->scratch;
    uint32_t* keys = (uint32_t*)base;
    uint32_t* keys_s = (uint32_t*)(base + a);
    uint32_t* idx = (uint32_t*)(base + 2 * a);
    uint32_t* idx_s = (uint32_t*)(base + 3 * a);
    unsigned long long* cnt = (unsigned long long*)(base + 4 * a);
    void* tmp = base + 4 * a + 256;
    hipLaunchKernelGGL(film_zero_u64_kernel, dim3(1), dim3(64), 0, st, cnt);
    hipLaunchKernelGGL(film_keys_kernel, grid_for(n), dim3(kBlock), 0, st, xy, n, W, H, keys, idx, cnt);
    if (int rc = check_launch(h, what)) return rc;
    NLPF_HIP(h, hipcub::DeviceRadixSort::SortPairs(tmp, cub_bytes, keys, keys_s, idx, idx_s, (int)n, 0, bits, st));
    if (splat)
        hipLaunchKernelGGL(film_runs_kernel<1>, grid_for(n), dim3(kBlock), 0, st, keys_s, idx_s, n, npix, rgb,
                           weight, film);
    else
        hipLaunchKernelGGL(film_runs_kernel<0>, grid_for(n), dim3(kBlock), 0, st, keys_s, idx_s, n, npix, rgb,
                           weight, film);
    if (int rc = check_launch(h, what)) return rc;
    if (dropped) {
        unsigned long long v = 0;
        NLPF_HIP(h, hipMemcpyAsync(&v, cnt, sizeof(v), hipMemcpyDeviceToHost, st));
        NLPF_HIP(h, hipStreamSynchronize(st));
        *dropped = v;
    }
    return NLPF_OK;
}

}  // namespace

}  // namespace nlpf

using namespace nlpf;

extern "C" {

int nlpf_film_clear(nlpf_handle* h, float* film, uint32_t width, uint32_t height) {
    if (!h) return NLPF_E_INVALID_ARG;
    if (int rc = film_check(h, film, width, height, "nlpf_film_clear")) return rc;
    const uint64_t n = 7ull * width * height;
    hipLaunchKernelGGL(film_fill_kernel, grid_for(n), dim3(kBlock), 0, h->stream, film, n, 0.0f);
    return check_launch(h, "nlpf_film_clear");
}

int nlpf_film_add_samples(nlpf_handle* h, float* film, uint32_t width, uint32_t height, const float* xy,
                          const float* rgb, const float* weight, uint64_t n, uint64_t* dropped) {
    return film_add(h, film, width, height, xy, rgb, weight, n, dropped, false, "nlpf_film_add_samples");
}

int nlpf_film_add_splats(nlpf_handle* h, float* film, uint32_t width, uint32_t height, const float* xy,
                         const float* rgb, const float* weight, uint64_t n, uint64_t* dropped) {
    return film_add(h, film, width, height, xy, rgb, weight, n, dropped, true, "nlpf_film_add_splats");
}

int nlpf_film_resolve(nlpf_handle* h, const float* film, uint32_t width, uint32_t height, float splat_scale,
                      int hdr, float* rgba) {
    if (!h) return NLPF_E_INVALID_ARG;
    if (int rc = film_check(h, film, width, height, "nlpf_film_resolve")) return rc;
    if (!rgba) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_film_resolve: NULL rgba");
    const uint32_t npix = width * height;
    hipLaunchKernelGGL(film_resolve_kernel, grid_for(npix), dim3(kBlock), 0, h->stream, film, npix, splat_scale,
                       hdr, (float4*)rgba);
    return check_launch(h, "nlpf_film_resolve");
}

}  // extern "C"
