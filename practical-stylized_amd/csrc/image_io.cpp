// Image output (crates/akari_render/src/util/mod.rs:57-121): a path ending in ".exr" gets a
// linear float RGB OpenEXR file, any other path an 8-bit sRGB PNG. Both encoders are written
// here from the published file formats. The EXR file is uncompressed scanline (OpenEXR 2.0,
// single part). The PNG file is RGB8 with its zlib stream in stored deflate blocks.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>
#include "common.hpp"

namespace nlpf {

namespace {

void put_u32le(std::vector<uint8_t>& o, uint32_t v) {
    for (int i = 0; i < 4; ++i) o.push_back((uint8_t)(v >> (8 * i)));
}
void put_u64le(std::vector<uint8_t>& o, uint64_t v) {
    for (int i = 0; i < 8; ++i) o.push_back((uint8_t)(v >> (8 * i)));
}
void put_u32be(std::vector<uint8_t>& o, uint32_t v) {
    for (int i = 3; i >= 0; --i) o.push_back((uint8_t)(v >> (8 * i)));
}
void put_f32(std::vector<uint8_t>& o, float f) {
    uint32_t v;
    memcpy(&v, &f, 4);
    put_u32le(o, v);
}
void put_str(std::vector<uint8_t>& o, const char* s) {
    while (*s) o.push_back((uint8_t)*s++);
    o.push_back(0);
}
// one header attribute: name, type name, byte size, value
void attr(std::vector<uint8_t>& o, const char* name, const char* type, const std::vector<uint8_t>& value) {
    put_str(o, name);
    put_str(o, type);
    put_u32le(o, (uint32_t)value.size());
    o.insert(o.end(), value.begin(), value.end());
}

// write_image_hdr (util/mod.rs:95-121): the R, G, B channels of the image as 32-bit floats.
std::vector<uint8_t> encode_exr(const float* rgba, uint32_t W, uint32_t H) {
    std::vector<uint8_t> o;
    put_u32le(o, 20000630u);  // magic number
    put_u32le(o, 2u);         // version 2, single-part scanline, short names
    std::vector<uint8_t> v;
    for (const char* ch : {"B", "G", "R"}) {  // channel list in ascending name order
        put_str(v, ch);
        put_u32le(v, 2u);   // FLOAT
        put_u32le(v, 0u);   // pLinear 0 + 3 reserved bytes
        put_u32le(v, 1u);   // x sampling
        put_u32le(v, 1u);   // y sampling
    }
    v.push_back(0);
    attr(o, "channels", "chlist", v);
    attr(o, "compression", "compression", {0});  // NO_COMPRESSION
    v.clear();
    put_u32le(v, 0);
    put_u32le(v, 0);
    put_u32le(v, W - 1);
    put_u32le(v, H - 1);
    attr(o, "dataWindow", "box2i", v);
    attr(o, "displayWindow", "box2i", v);
    attr(o, "lineOrder", "lineOrder", {0});  // INCREASING_Y
    v.clear();
    put_f32(v, 1.0f);
    attr(o, "pixelAspectRatio", "float", v);
    v.clear();
    put_f32(v, 0.0f);
    put_f32(v, 0.0f);
    attr(o, "screenWindowCenter", "v2f", v);
    v.clear();
    put_f32(v, 1.0f);
    attr(o, "screenWindowWidth", "float", v);
    o.push_back(0);  // end of header
    // offset table: one chunk per scanline
    const uint64_t line_bytes = 3ull * 4 * W;
    const uint64_t table_at = o.size();
    const uint64_t first = table_at + 8ull * H;
    for (uint32_t y = 0; y < H; ++y) put_u64le(o, first + (uint64_t)y * (8 + line_bytes));
    o.reserve(o.size() + (size_t)H * (8 + line_bytes));
    for (uint32_t y = 0; y < H; ++y) {
        put_u32le(o, y);
        put_u32le(o, (uint32_t)line_bytes);
        for (int c : {2, 1, 0})  // B, G, R planes of the scanline
            for (uint32_t x = 0; x < W; ++x) put_f32(o, rgba[4 * ((size_t)y * W + x) + c]);
    }
    return o;
}

uint32_t crc32_update(uint32_t crc, const uint8_t* p, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    crc = ~crc;
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFF] ^ (crc >> 8);
    return ~crc;
}

void png_chunk(std::vector<uint8_t>& o, const char* type, const std::vector<uint8_t>& data) {
    put_u32be(o, (uint32_t)data.size());
    const size_t at = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), data.begin(), data.end());
    put_u32be(o, crc32_update(0, o.data() + at, 4 + data.size()));
}

// f32_linear_to_srgb1 (color.rs:590-596), with the host powf the reference calls
float linear_to_srgb1_host(float l) {
#pragma clang fp contract(off)
    return l <= 0.0031308f ? l * 12.92f : powf(l, 1.0f / 2.4f) * 1.055f - 0.055f;
}

// (x * 255).clamp(0, 255) as u8 (util/mod.rs:90-91): truncation, NaN to 0
uint8_t to_u8(float x) {
    const float v = x * 255.0f;
    if (!(v > 0.0f)) return 0;
    if (v >= 255.0f) return 255;
    return (uint8_t)v;
}

// write_image_ldr (util/mod.rs:63-94): sRGB-encode, quantise, RGB8.
std::vector<uint8_t> encode_png(const float* rgba, uint32_t W, uint32_t H) {
    std::vector<uint8_t> raw;
    raw.reserve((size_t)H * (1 + 3ull * W));
    for (uint32_t y = 0; y < H; ++y) {
        raw.push_back(0);  // filter type None
        for (uint32_t x = 0; x < W; ++x)
            for (int c = 0; c < 3; ++c) raw.push_back(to_u8(linear_to_srgb1_host(rgba[4 * ((size_t)y * W + x) + c])));
    }
    std::vector<uint8_t> z{0x78, 0x01};  // zlib header: deflate, 32 KiB window, no dictionary
    size_t at = 0;
    do {
        const size_t len = std::min<size_t>(65535, raw.size() - at);
        const bool last = at + len == raw.size();
        z.push_back(last ? 1 : 0);  // BFINAL, BTYPE 00 (stored)
        z.push_back((uint8_t)len);
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)~len);
        z.push_back((uint8_t)(~len >> 8));
        z.insert(z.end(), raw.begin() + at, raw.begin() + at + len);
        at += len;
    } while (at < raw.size());
    uint32_t a = 1, b = 0;  // Adler-32
    for (uint8_t c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put_u32be(z, (b << 16) | a);
    std::vector<uint8_t> o{0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put_u32be(ihdr, W);
    put_u32be(ihdr, H);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit, truecolour, deflate, adaptive filters, no interlace
    png_chunk(o, "IHDR", ihdr);
    png_chunk(o, "IDAT", z);
    png_chunk(o, "IEND", {});
    return o;
}

bool ends_with(const std::string& s, const char* suf) {
    const size_t n = strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

// encodes and writes; creates the parent directory like the reference (util/mod.rs:83-84)
int write_image(const float* rgba, uint32_t W, uint32_t H, const char* path, std::string& err) {
    const std::string p(path);
    std::vector<uint8_t> bytes = ends_with(p, ".exr") ? encode_exr(rgba, W, H) : encode_png(rgba, W, H);
    std::error_code ec;
    const std::filesystem::path parent = std::filesystem::path(p).parent_path();
    if (!parent.empty()) std::filesystem::create_directories(parent, ec);
    if (ec) {
        err = "write_image: cannot create " + parent.string() + ": " + ec.message();
        return NLPF_E_IO;
    }
    FILE* f = fopen(path, "wb");
    if (!f) {
        err = "write_image: cannot open " + p;
        return NLPF_E_IO;
    }
    const size_t w = fwrite(bytes.data(), 1, bytes.size(), f);
    const int c = fclose(f);
    if (w != bytes.size() || c != 0) {
        err = "write_image: short write to " + p;
        return NLPF_E_IO;
    }
    return NLPF_OK;
}

bool image_args_ok(const float* rgba, uint32_t W, uint32_t H, const char* path, std::string& err) {
    if (!rgba || !path || !*path) {
        err = "write_image: NULL image or empty path";
        return false;
    }
    if (W == 0 || H == 0 || (uint64_t)W * H >= (1ull << 31)) {
        err = "write_image: width and height must be >= 1 with width * height < 2^31";
        return false;
    }
    return true;
}

}  // namespace

}  // namespace nlpf

using namespace nlpf;

extern "C" {

int nlpf_write_image_host(const float* rgba, uint32_t width, uint32_t height, const char* path) {
    std::string err;
    if (!image_args_ok(rgba, width, height, path, err)) {
        set_global_error(err);
        return NLPF_E_INVALID_ARG;
    }
    int rc;
    try {
        rc = write_image(rgba, width, height, path, err);
    } catch (const std::bad_alloc&) {
        err = "write_image: out of host memory";
        rc = NLPF_E_OOM;
    }
    if (rc) set_global_error(err);
    return rc;
}

static int write_image_device(nlpf_handle* h, const float* rgba, uint32_t width, uint32_t height, const char* path) {
    std::string err;
    if (!image_args_ok(rgba, width, height, path, err)) NLPF_FAIL(h, NLPF_E_INVALID_ARG, err);
    std::vector<float> host(4ull * width * height);
    NLPF_HIP(h, hipMemcpyAsync(host.data(), rgba, host.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    const int rc = write_image(host.data(), width, height, path, err);
    if (rc) h->err = err;
    return rc;
}

int nlpf_write_image(nlpf_handle* h, const float* rgba, uint32_t width, uint32_t height, const char* path) {
    if (!h) return NLPF_E_INVALID_ARG;
    try {
        return write_image_device(h, rgba, width, height, path);
    } catch (const std::bad_alloc&) {
        NLPF_FAIL(h, NLPF_E_OOM, "write_image: out of host memory");
    }
}

}  // extern "C"
