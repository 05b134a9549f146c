// Learned inner-expectation cache: a 96(84)-64-64-64-64-64-16(3) ReLU MLP on BF16 MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation).
//
// Layout convention.  Activations are kept "feature-major" in MFMA tiles: Z = W X with the
// 16x16x32 A operand = weight rows (out features) and the B operand = activations (in
// features x samples).  A D tile of layer l leaves lane (g = lane/16, n = lane%16) holding
// features 16t + 4g + i (i = 0..3) of sample n; the next layer's B fragment for K-step s is
// taken from the lane's own D values of tiles 2s and 2s+1, i.e. its K order is permuted:
//   logical k = 8g + j  <->  physical feature 32s + (j < 4 ? 4g + j : 16 + 4g + j - 4).
// The weight A fragments are loaded with the same permutation, so no data crosses lanes
// between layers.  The input layer uses the same order for the 96 encoded features.
#include "common.hpp"
#include "scene.hpp"

namespace nlpf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MLP_IN = 96, MLP_HID = 64, MLP_OUT = 16, MLP_LAYERS = 6;
constexpr int MLP_FREQ = 12;
__host__ __device__ constexpr int mlp_in(int l) { return l == 0 ? MLP_IN : MLP_HID; }
__host__ __device__ constexpr int mlp_out(int l) { return l == MLP_LAYERS - 1 ? MLP_OUT : MLP_HID; }
// parameter offsets: W_l [out][in] then b_l [out], layer after layer
__host__ __device__ constexpr int mlp_woff(int l) {
    return l == 0 ? 0 : mlp_woff(l - 1) + mlp_in(l - 1) * mlp_out(l - 1) + mlp_out(l - 1);
}
__host__ __device__ constexpr int mlp_boff(int l) { return mlp_woff(l) + mlp_in(l) * mlp_out(l); }
constexpr int MLP_NPARAM = mlp_woff(MLP_LAYERS);

__device__ __forceinline__ int phys_feature(int s, int g, int j) {
    return 32 * s + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
}

// 84 input features of a depth-d vertex (+12 zero pads): clamped position, 12-octave sin/cos
// of each coordinate, normal, albedo, view direction (towards the parent vertex; the camera
// at d=0).  Column P of the encoded input (= column P of W_0) is computed by lane group
// g = (P >> 2) & 3 of the MFMA B operand, as its slot u = 8 (P / 32) + j (j as in
// phys_feature); the columns are assigned so every lane group runs the same code:
//   slots u < 18    sin (u even) / cos (u odd) of 2^k pi p_c, k = 3 g + u / 6, c = (u % 6) / 2
//                   (octave 3 g from the hardware sine / cosine, the next two by doubling)
//   slots 18..23    g = 0: clamped p, n;  g = 1: albedo, view;  g = 2, 3: zero padding
struct VertexIn {
    float p[3], n[3], a[3], v[3];
};
__host__ __device__ constexpr int col_group(int P) { return (P >> 2) & 3; }
__host__ __device__ constexpr int col_slot(int P) { return 8 * (P >> 5) + ((P & 31) >= 16 ? 4 : 0) + (P & 3); }
__host__ __device__ constexpr bool col_is_pad(int P) { return col_group(P) >= 2 && col_slot(P) >= 18; }

// sin / cos of 2^k pi x = sin / cos of 2 pi fr, fr = the exactly reduced fraction of
// 2^(k-1) x in [0, 1), on the hardware sine / cosine (v_sin_f32 / v_cos_f32 take the angle in
// revolutions and need |arg| <= 256, which the reduction guarantees)
__device__ __forceinline__ void octave_sincos(float x, int k, float& sn, float& cs) {
    const float t = ldexpf(x, k - 1);
    const float fr = t - floorf(t);
    sn = __builtin_amdgcn_sinf(fr);
    cs = __builtin_amdgcn_cosf(fr);
}

// the next octave from the current one: sin 2x = (2 sin x) cos x, cos 2x = 1 - (2 sin x) sin x
__device__ __forceinline__ void octave_double(float& sn, float& cs) {
    const float t = 2.0f * sn;
    const float s2 = t * cs;
    cs = fmaf(-t, sn, 1.0f);
    sn = s2;
}

// feature of column P (reference path: the encode-only and FP64 reference kernels); octave
// 3 g + i comes from octave 3 g and i doublings, exactly as encode_frags computes it
__device__ inline float mlp_feature(const VertexIn& x, int P) {
    const int g = col_group(P), u = col_slot(P);
    if (u < 18) {
        const int c = (u % 6) >> 1;
        float sn, cs;
        octave_sincos(c == 0 ? x.p[0] : c == 1 ? x.p[1] : x.p[2], 3 * g, sn, cs);
        for (int i = 0; i < u / 6; ++i) octave_double(sn, cs);
        return (u & 1) ? cs : sn;
    }
    const int r = u - 18;
    if (g == 0) return r < 3 ? fminf(fmaxf(r == 0 ? x.p[0] : r == 1 ? x.p[1] : x.p[2], -2.0f), 2.0f)
                             : (r == 3 ? x.n[0] : r == 4 ? x.n[1] : x.n[2]);
    if (g == 1) return r < 3 ? (r == 0 ? x.a[0] : r == 1 ? x.a[1] : x.a[2]) : (r == 3 ? x.v[0] : r == 4 ? x.v[1] : x.v[2]);
    return 0.0f;
}

template <typename S>
__device__ __forceinline__ VertexIn load_vertex_at(const S* rec, const float* anchors, uint32_t W, uint32_t H,
                                                   uint32_t L, uint32_t d, uint64_t s, uint64_t px, uint32_t x,
                                                   uint32_t y) {
    const uint64_t plane = (uint64_t)W * H;
    const uint32_t abx = (W + 7) / 8, aby = (H + 7) / 8;
    auto pos = [&](uint32_t lv, float out[3]) {
        const S* r = rec + ((uint64_t)s * L + lv) * NLPF_SCREEN_PLANES * plane;
        for (int c = 0; c < 3; ++c) out[c] = (float)r[c * plane + px];
        if (anchors) {
            const float* q = anchors + (((uint64_t)s * L + lv) * abx * aby + (uint64_t)(y >> 3) * abx + (x >> 3)) * 3;
            for (int c = 0; c < 3; ++c) out[c] += q[c];
        }
    };
    VertexIn v;
    pos(d, v.p);
    const S* r = rec + ((uint64_t)s * L + d) * NLPF_SCREEN_PLANES * plane;
    for (int c = 0; c < 3; ++c) {
        v.n[c] = (float)r[(3 + c) * plane + px];
        v.a[c] = (float)r[(6 + c) * plane + px];
    }
    float q[3];
    if (d > 0) {
        pos(d - 1, q);
    } else {
        q[0] = 0.0f;
        q[1] = 0.0f;
        q[2] = SceneConst::cam_z;
    }
    float dv[3] = {q[0] - v.p[0], q[1] - v.p[1], q[2] - v.p[2]};
    const float len2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
    const float inv = len2 > 0.0f ? __builtin_amdgcn_rsqf(len2) : 0.0f;
    for (int c = 0; c < 3; ++c) v.v[c] = dv[c] * inv;
    return v;
}

template <typename S>
__device__ __forceinline__ VertexIn load_vertex(const S* rec, const float* anchors, uint32_t W, uint32_t H,
                                                uint32_t L, uint32_t d, uint64_t sample) {
    const uint64_t plane = (uint64_t)W * H;
    const uint64_t s = sample / plane, px = sample - s * plane;
    return load_vertex_at(rec, anchors, W, H, L, d, s, px, (uint32_t)(px % W), (uint32_t)(px / W));
}

// A fragment of W_l (rows = out features m0..m0+15, K-step s, permuted K order)
__device__ __forceinline__ bf16x8 load_wfrag(const __bf16* w, int in_dim, int m0, int s, int lane) {
    const int m = m0 + (lane & 15), g = lane >> 4;
    const __bf16* row = w + (size_t)m * in_dim + 32 * s + 4 * g;
    const bf16x4 lo = *(const bf16x4*)row;
    const bf16x4 hi = *(const bf16x4*)(row + 16);
    bf16x8 r;
    r[0] = lo[0]; r[1] = lo[1]; r[2] = lo[2]; r[3] = lo[3];
    r[4] = hi[0]; r[5] = hi[1]; r[6] = hi[2]; r[7] = hi[3];
    return r;
}

// B fragments of the next layer from this layer's D tiles (ReLU applied).  The ReLU runs
// on the rounded bf16 pairs as a signed 16-bit max with 0: negative bf16 values (sign bit set,
// -0 included) are negative int16, non-negative ones keep their order, and rounding commutes
// with max(., 0).
typedef short i16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
// two floats -> two bf16 (round to nearest even) in one v_cvt_pk_bf16_f32
__device__ __forceinline__ bf16x2 cvt2(float x, float y) { return __builtin_convertvector((f32x2){x, y}, bf16x2); }
__device__ __forceinline__ bf16x8 pack8(bf16x2 a, bf16x2 b, bf16x2 c, bf16x2 d) {
    return __builtin_shufflevector(__builtin_shufflevector(a, b, 0, 1, 2, 3), __builtin_shufflevector(c, d, 0, 1, 2, 3),
                                   0, 1, 2, 3, 4, 5, 6, 7);
}
__device__ __forceinline__ bf16x2 relu2(bf16x2 v) {
    i16x2 t;
    __builtin_memcpy(&t, &v, sizeof(t));
    t = __builtin_elementwise_max(t, (i16x2){0, 0});
    __builtin_memcpy(&v, &t, sizeof(t));
    return v;
}
__device__ __forceinline__ bf16x8 relu_pack(const f32x4& a, const f32x4& b) {
    return pack8(relu2(cvt2(a[0], a[1])), relu2(cvt2(a[2], a[3])), relu2(cvt2(b[0], b[1])), relu2(cvt2(b[2], b[3])));
}

struct MlpParamsDev {
    const __bf16* w[MLP_LAYERS];   // [out][in] bf16
    const __bf16* wt[MLP_LAYERS];  // [in][out] bf16 (transposed, for the backward pass)
    const float* b[MLP_LAYERS];    // FP32 biases
};

constexpr int MLP_QT = 4;  // sample tiles (of 16) per wave

// Fragments of the weights staged in LDS: fragment f = (layer, 16-row tile t, K-step s),
// 64 lanes x 16 B each, in the order layer 0 (t, s) 4x3, layers 1..4 4x2, layer 5 1x2.
constexpr int MLP_NFRAG = 12 + 4 * 8 + 2;
__host__ __device__ constexpr int frag_index(int l, int t, int s) {
    return l == 0 ? 3 * t + s : (l < MLP_LAYERS - 1 ? 12 + 8 * (l - 1) + 2 * t + s : 44 + s);
}
constexpr int MLP_NBIAS = 5 * MLP_HID + MLP_OUT;
__host__ __device__ constexpr int bias_index(int l) { return MLP_HID * l; }

// Forward through all layers for 16*MLP_QT samples per wave.  Input fragments come from
// `xin` (built by the caller); weights and biases from `wf(l, t, s)` / `bf(l, t)`.
template <typename WF, typename BF>
__device__ __forceinline__ void mlp_forward_wave(const WF& wf, const BF& bf, bf16x8 (&xin)[MLP_QT][3],
                                                 f32x4 (&y)[MLP_QT]) {
    f32x4 acc[MLP_QT][4];
    bf16x8 xb[MLP_QT][2];
    // each weight fragment is read from LDS one step ahead of the MFMAs that use it, so the read
    // is in flight while the previous fragment's MFMAs run
    bf16x8 w = wf(0, 0, 0);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const f32x4 bias = bf(0, t);
#pragma unroll
        for (int q = 0; q < MLP_QT; ++q) acc[q][t] = bias;
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const bf16x8 wn = s < 2 ? wf(0, t, s + 1) : (t < 3 ? wf(0, t + 1, 0) : wf(1, 0, 0));
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < MLP_QT; ++q) acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, xin[q][s], acc[q][t], 0, 0, 0);
            w = wn;
        }
    }
#pragma unroll
    for (int l = 1; l < MLP_LAYERS - 1; ++l) {
#pragma unroll
        for (int q = 0; q < MLP_QT; ++q) {
            xb[q][0] = relu_pack(acc[q][0], acc[q][1]);
            xb[q][1] = relu_pack(acc[q][2], acc[q][3]);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const f32x4 bias = bf(l, t);
#pragma unroll
            for (int q = 0; q < MLP_QT; ++q) acc[q][t] = bias;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const bf16x8 wn = s < 1 ? wf(l, t, 1) : (t < 3 ? wf(l, t + 1, 0) : wf(l + 1, 0, 0));
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int q = 0; q < MLP_QT; ++q)
                    acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, xb[q][s], acc[q][t], 0, 0, 0);
                w = wn;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < MLP_QT; ++q) {
        xb[q][0] = relu_pack(acc[q][0], acc[q][1]);
        xb[q][1] = relu_pack(acc[q][2], acc[q][3]);
    }
    const f32x4 bias = bf(MLP_LAYERS - 1, 0);
#pragma unroll
    for (int q = 0; q < MLP_QT; ++q) y[q] = bias;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const bf16x8 wn = s < 1 ? wf(MLP_LAYERS - 1, 0, 1) : w;
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < MLP_QT; ++q) y[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, xb[q][s], y[q], 0, 0, 0);
        w = wn;
    }
}

// slots 0..17 of lane group g: sin / cos of octaves 3g, 3g+1, 3g+2 of each coordinate
__device__ __forceinline__ void encode_trig(const float (&p)[3], int g, float (&u)[18]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float sn, cs;
        octave_sincos(p[c], 3 * g, sn, cs);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (i) octave_double(sn, cs);
            u[6 * i + 2 * c] = sn;
            u[6 * i + 2 * c + 1] = cs;
        }
    }
}

// the raw slots 18..23 as three bf16 pairs: (p0, p1), (p2, n0), (n1, n2) for lane group 0,
// (a0, a1), (a2, v0), (v1, v2) for lane group 1 (p clamped to [-2, 2])
__device__ __forceinline__ void raw_pairs(const VertexIn& v, bf16x2 (&r0)[3], bf16x2 (&r1)[3]) {
    const float c0 = fminf(fmaxf(v.p[0], -2.0f), 2.0f), c1 = fminf(fmaxf(v.p[1], -2.0f), 2.0f),
                c2 = fminf(fmaxf(v.p[2], -2.0f), 2.0f);
    r0[0] = cvt2(c0, c1);
    r0[1] = cvt2(c2, v.n[0]);
    r0[2] = cvt2(v.n[1], v.n[2]);
    r1[0] = cvt2(v.a[0], v.a[1]);
    r1[1] = cvt2(v.a[2], v.v[0]);
    r1[2] = cvt2(v.v[1], v.v[2]);
}

__device__ __forceinline__ void assemble_frags(const float (&u)[18], const bf16x2 (&raw)[3], bf16x8 (&x)[3]) {
    x[0] = pack8(cvt2(u[0], u[1]), cvt2(u[2], u[3]), cvt2(u[4], u[5]), cvt2(u[6], u[7]));
    x[1] = pack8(cvt2(u[8], u[9]), cvt2(u[10], u[11]), cvt2(u[12], u[13]), cvt2(u[14], u[15]));
    x[2] = pack8(cvt2(u[16], u[17]), raw[0], raw[1], raw[2]);
}

// encoded input fragments of the lane's sample for the 3 K-steps (slot u = 8 s + j)
__device__ __forceinline__ void encode_frags(const VertexIn& v, int lane, bf16x8 (&x)[3]) {
    const int g = lane >> 4;
    float u[18];
    encode_trig(v.p, g, u);
    bf16x2 r0[3], r1[3], raw[3];
    raw_pairs(v, r0, r1);
    const bf16x2 z = cvt2(0.0f, 0.0f);
#pragma unroll
    for (int k = 0; k < 3; ++k) raw[k] = g == 0 ? r0[k] : (g == 1 ? r1[k] : z);
    assemble_frags(u, raw, x);
}

struct MlpInferArgs {
    MlpParamsDev P;
    const void* rec;
    const float* anchors;
    int fp16;
    uint32_t W, H, L, d;
    const uint32_t* idx;  // sample list (slice * W*H + pixel) or nullptr for 0..n-1
    const float* feats;   // optional raw features [n][96] instead of records
    uint64_t n;
    float* out;           // [n][3], or planar [slice - out_slice0][out_levels][3][W*H] (slot out_slot)
    int planar;
    uint32_t out_levels, out_slot, out_slice0;
    uint64_t sample0;     // contiguous samples: sample0 .. sample0 + n - 1
};

// all weight fragments and biases into LDS (once per workgroup)
__device__ __forceinline__ void stage_weights(const MlpParamsDev& P, bf16x8 (*wl)[64], f32x4* bl) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int f = wv; f < MLP_NFRAG; f += 4) {
        const int l = f < 12 ? 0 : (f < 44 ? 1 + (f - 12) / 8 : MLP_LAYERS - 1);
        const int r = f < 12 ? f : (f < 44 ? (f - 12) % 8 : f - 44);
        const int ks = l == 0 ? 3 : 2;
        wl[f][lane] = load_wfrag(P.w[l], mlp_in(l), 16 * (r / ks), r % ks, lane);
    }
    for (int i = threadIdx.x; i < MLP_NBIAS; i += blockDim.x) {
        const int l = i / MLP_HID;
        ((float*)bl)[i] = P.b[l][i - MLP_HID * l];
    }
    __syncthreads();
}

// Persistent inference: the workgroup stages all weight fragments and biases in LDS once
// (47 KB), then its 4 waves loop over tiles of 16*MLP_QT samples.
__global__ void __launch_bounds__(256, 2) mlp_infer_kernel(MlpInferArgs a) {
    __shared__ bf16x8 wl[MLP_NFRAG][64];
    __shared__ f32x4 bl[MLP_NBIAS / 4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = lane >> 4;
    stage_weights(a.P, wl, bl);
    const auto wf = [&](int l, int t, int s) { return wl[frag_index(l, t, s)][lane]; };
    const auto bf = [&](int l, int t) { return bl[(bias_index(l) + 16 * t + 4 * g) >> 2]; };

    const uint64_t wave = (uint64_t)blockIdx.x * 4 + wv;
    const uint64_t nwaves = (uint64_t)gridDim.x * 4;
    const uint64_t per = 16 * MLP_QT;
    const uint64_t plane = (uint64_t)a.W * a.H;
    // contiguous samples on an image of >= 64 and < 2^24 pixels: the wave-uniform tile base is
    // split into (slice, pixel) once, each lane adds its offset (< 64: at most one slice
    // boundary), and the row comes from an FP32 reciprocal with one correction step (exact for
    // pixel indices below 2^24)
    const bool contiguous = !a.idx && !a.feats && plane >= 64 && plane < (1u << 24);
    const float inv_w = 1.0f / (float)a.W;
    for (uint64_t base = wave * per; base < a.n; base += nwaves * per) {
        // keeps the (loop-invariant) LDS weight reads inside the loop: hoisted, the 46
        // fragments would hold 184 VGPRs for the whole loop
        asm volatile("" ::: "memory");
        const uint64_t s_base = contiguous ? (a.sample0 + base) / plane : 0;
        const uint32_t px_base = contiguous ? (uint32_t)(a.sample0 + base - s_base * plane) : 0;
        bf16x8 xin[MLP_QT][3];
        if (contiguous) {
            // each lane loads its own sample (base + lane) once; the 16-lane row of tile q then
            // reads position and raw-feature pairs of its column's sample from row q
            float pc[3] = {0.0f, 0.0f, 0.0f};
            bf16x2 r0[3], r1[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) r0[k] = r1[k] = cvt2(0.0f, 0.0f);
            if (base + lane < a.n) {
                uint32_t px = px_base + lane, sl = (uint32_t)s_base;
                const bool wrap = px >= (uint32_t)plane;
                px = wrap ? px - (uint32_t)plane : px;
                sl = wrap ? sl + 1 : sl;
                int yy = (int)((float)px * inv_w);
                int xx = (int)px - yy * (int)a.W;
                const bool lo = xx < 0, hi = xx >= (int)a.W;
                yy = lo ? yy - 1 : (hi ? yy + 1 : yy);
                xx = lo ? xx + (int)a.W : (hi ? xx - (int)a.W : xx);
                const VertexIn v =
                    a.fp16 ? load_vertex_at((const __half*)a.rec, a.anchors, a.W, a.H, a.L, a.d, sl, px, xx, yy)
                           : load_vertex_at((const float*)a.rec, nullptr, a.W, a.H, a.L, a.d, sl, px, xx, yy);
                pc[0] = v.p[0];
                pc[1] = v.p[1];
                pc[2] = v.p[2];
                raw_pairs(v, r0, r1);
            }
            uint32_t rb[6];
            __builtin_memcpy(&rb[0], r0, sizeof(r0));
            __builtin_memcpy(&rb[3], r1, sizeof(r1));
#pragma unroll
            for (int q = 0; q < MLP_QT; ++q) {
                const int src = 16 * q + (lane & 15);
                float p[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) p[c] = __shfl(pc[c], src);
                uint32_t t[6];
#pragma unroll
                for (int k = 0; k < 6; ++k) t[k] = (uint32_t)__shfl((int)rb[k], src);
                float u[18];
                encode_trig(p, g, u);
                bf16x2 raw[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const uint32_t bits = g == 0 ? t[k] : (g == 1 ? t[3 + k] : 0u);
                    __builtin_memcpy(&raw[k], &bits, sizeof(bits));
                }
                assemble_frags(u, raw, xin[q]);
            }
        } else {
#pragma unroll
            for (int q = 0; q < MLP_QT; ++q) {
                const uint64_t i = base + 16 * q + (lane & 15);
                if (i < a.n) {
                    if (a.feats) {
                        for (int s = 0; s < 3; ++s)
                            for (int j = 0; j < 8; ++j)
                                xin[q][s][j] = (__bf16)a.feats[i * MLP_IN + phys_feature(s, g, j)];
                    } else {
                        const uint64_t sm = a.idx ? a.idx[i] : a.sample0 + i;
                        const VertexIn v = a.fp16
                                               ? load_vertex((const __half*)a.rec, a.anchors, a.W, a.H, a.L, a.d, sm)
                                               : load_vertex((const float*)a.rec, nullptr, a.W, a.H, a.L, a.d, sm);
                        encode_frags(v, lane, xin[q]);
                    }
                } else {
                    for (int s = 0; s < 3; ++s)
                        for (int j = 0; j < 8; ++j) xin[q][s][j] = (__bf16)0.0f;
                }
            }
        }
        f32x4 y[MLP_QT];
        mlp_forward_wave(wf, bf, xin, y);
        if (g == 0) {
#pragma unroll
            for (int q = 0; q < MLP_QT; ++q) {
                const uint64_t i = base + 16 * q + lane;
                if (i < a.n) {
                    if (a.planar) {
                        uint64_t sl, px;
                        if (contiguous) {
                            const uint32_t pq = px_base + 16 * q + lane;
                            const bool wrap = pq >= (uint32_t)plane;
                            px = wrap ? pq - (uint32_t)plane : pq;
                            sl = wrap ? s_base + 1 : s_base;
                        } else {
                            const uint64_t sm = a.idx ? a.idx[i] : a.sample0 + i;
                            sl = sm / plane;
                            px = sm - sl * plane;
                        }
                        float* o = a.out + ((uint64_t)(sl - a.out_slice0) * a.out_levels + a.out_slot) * 3 * plane + px;
                        o[0] = y[q][0];
                        o[plane] = y[q][1];
                        o[2 * plane] = y[q][2];
                    } else {
                        a.out[3 * i] = y[q][0];
                        a.out[3 * i + 1] = y[q][1];
                        a.out[3 * i + 2] = y[q][2];
                    }
                }
            }
        }
    }
}

}  // namespace nlpf

// ---------------------------------------------------------------------------------------
// training: forward (store activations), backward (dZ per layer), weight gradients, Adam
// ---------------------------------------------------------------------------------------
namespace nlpf {

struct MlpTrainBufs {
    __bf16* act[MLP_LAYERS];  // input of layer l, feature-major [in_l][B]
    __bf16* dz[MLP_LAYERS];   // dL/dZ_l, feature-major [out_l][B]
    float* y;                 // [B][3]
    float* loss_part;         // per-wave partial loss
    uint32_t B;
};

struct MlpTrainArgs {
    MlpParamsDev P;
    MlpTrainBufs T;
    const void* rec;
    const float* anchors;
    int fp16;
    uint32_t W, H, L, d;
    const uint32_t* idx;   // [B] sample ids (slice * W*H + pixel), or nullptr with feats
    const float* feats;    // optional raw features [B][96]
    const float* target;   // [B][3]
};

__device__ __forceinline__ void store_frag(__bf16* base, uint32_t B, uint64_t n, int s, int g, const bf16x8& v) {
#pragma unroll
    for (int j = 0; j < 8; ++j) base[(size_t)phys_feature(s, g, j) * B + n] = v[j];
}

// forward over one batch, 64 samples per wave, storing every layer's input (bf16) and the
// output; dZ of the output layer is 2 (y - target) / B on the three colour rows
__global__ void __launch_bounds__(256) mlp_train_fwd_kernel(MlpTrainArgs a) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t base = wave * 16 * MLP_QT;
    const uint32_t B = a.T.B;
    if (base >= B) return;
    bf16x8 xin[MLP_QT][3];
#pragma unroll
    for (int q = 0; q < MLP_QT; ++q) {
        const uint64_t i = base + 16 * q + (lane & 15);
        if (a.feats) {
            for (int s = 0; s < 3; ++s)
                for (int j = 0; j < 8; ++j) xin[q][s][j] = (__bf16)a.feats[i * MLP_IN + phys_feature(s, g, j)];
        } else {
            const uint64_t sm = a.idx[i];
            const VertexIn v = a.fp16 ? load_vertex((const __half*)a.rec, a.anchors, a.W, a.H, a.L, a.d, sm)
                                      : load_vertex((const float*)a.rec, nullptr, a.W, a.H, a.L, a.d, sm);
            encode_frags(v, lane, xin[q]);
        }
        for (int s = 0; s < 3; ++s) store_frag(a.T.act[0], B, i, s, g, xin[q][s]);
    }
    // forward, as mlp_forward_wave, with stores of each layer's input
    f32x4 acc[MLP_QT][4];
    bf16x8 xb[MLP_QT][2];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 bias;
        for (int i = 0; i < 4; ++i) bias[i] = a.P.b[0][16 * t + 4 * g + i];
#pragma unroll
        for (int q = 0; q < MLP_QT; ++q) acc[q][t] = bias;
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const bf16x8 w = load_wfrag(a.P.w[0], MLP_IN, 16 * t, s, lane);
#pragma unroll
            for (int q = 0; q < MLP_QT; ++q) acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, xin[q][s], acc[q][t], 0, 0, 0);
        }
    }
    for (int l = 1; l < MLP_LAYERS; ++l) {
#pragma unroll
        for (int q = 0; q < MLP_QT; ++q) {
            xb[q][0] = relu_pack(acc[q][0], acc[q][1]);
            xb[q][1] = relu_pack(acc[q][2], acc[q][3]);
            const uint64_t i = base + 16 * q + (lane & 15);
            store_frag(a.T.act[l], B, i, 0, g, xb[q][0]);
            store_frag(a.T.act[l], B, i, 1, g, xb[q][1]);
        }
        const int nt = l == MLP_LAYERS - 1 ? 1 : 4;
        for (int t = 0; t < nt; ++t) {
            f32x4 bias;
            for (int i = 0; i < 4; ++i) bias[i] = a.P.b[l][16 * t + 4 * g + i];
#pragma unroll
            for (int q = 0; q < MLP_QT; ++q) acc[q][t] = bias;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const bf16x8 w = load_wfrag(a.P.w[l], MLP_HID, 16 * t, s, lane);
#pragma unroll
                for (int q = 0; q < MLP_QT; ++q)
                    acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, xb[q][s], acc[q][t], 0, 0, 0);
            }
        }
    }
    // output rows 0..3 live in lanes g == 0 of tile 0
    float lsum = 0.0f;
#pragma unroll
    for (int q = 0; q < MLP_QT; ++q) {
        const uint64_t i = base + 16 * q + (lane & 15);
        for (int r = 0; r < 4; ++r) {
            const int row = 4 * g + r;
            float dz = 0.0f;
            if (row < 3) {
                const float yv = acc[q][0][r];
                const float d = yv - a.target[3 * i + row];
                a.T.y[3 * i + row] = yv;
                lsum += d * d;
                dz = 2.0f * d / (float)B;
            }
            a.T.dz[MLP_LAYERS - 1][(size_t)row * B + i] = (__bf16)dz;
        }
    }
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
    if (lane == 0) a.T.loss_part[wave] = lsum / (float)B;
}

// dZ_{l-1} = (W_l^T dZ_l) * [act_l > 0], l = 5..1, 64 samples per wave
__global__ void __launch_bounds__(256) mlp_train_bwd_kernel(MlpTrainArgs a) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t base = wave * 16 * MLP_QT;
    const uint32_t B = a.T.B;
    if (base >= B) return;
    bf16x8 db[MLP_QT][2];
    // dZ5 fragments: K = the 16 output rows (K-step 0 uses logical k = 8g + j, g < 2)
#pragma unroll
    for (int q = 0; q < MLP_QT; ++q) {
        const uint64_t i = base + 16 * q + (lane & 15);
        for (int j = 0; j < 8; ++j) {
            const int k = 8 * g + j;
            db[q][0][j] = k < 16 ? a.T.dz[MLP_LAYERS - 1][(size_t)k * B + i] : (__bf16)0.0f;
            db[q][1][j] = (__bf16)0.0f;
        }
    }
    f32x4 acc[MLP_QT][4];
    for (int l = MLP_LAYERS - 1; l >= 1; --l) {
        const int ksteps = l == MLP_LAYERS - 1 ? 1 : 2;
        for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int q = 0; q < MLP_QT; ++q) acc[q][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            for (int s = 0; s < ksteps; ++s) {
                bf16x8 w;
                if (l == MLP_LAYERS - 1) {
                    // W_5^T rows = the 64 inputs of layer 5, K = its 16 outputs in plain order
                    const int m = 16 * t + (lane & 15);
                    for (int j = 0; j < 8; ++j) {
                        const int k = 8 * g + j;
                        w[j] = k < 16 ? a.P.wt[l][(size_t)m * MLP_OUT + k] : (__bf16)0.0f;
                    }
                } else {
                    w = load_wfrag(a.P.wt[l], MLP_HID, 16 * t, s, lane);
                }
#pragma unroll
                for (int q = 0; q < MLP_QT; ++q)
                    acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, db[q][s], acc[q][t], 0, 0, 0);
            }
        }
        // ReLU mask from the stored input of layer l, store dZ_{l-1}, repack as B fragments
#pragma unroll
        for (int q = 0; q < MLP_QT; ++q) {
            const uint64_t i = base + 16 * q + (lane & 15);
            for (int t = 0; t < 4; ++t)
                for (int r = 0; r < 4; ++r) {
                    const int f = 16 * t + 4 * g + r;
                    const float x = (float)a.T.act[l][(size_t)f * B + i];
                    const float v = x > 0.0f ? acc[q][t][r] : 0.0f;
                    acc[q][t][r] = v;
                    a.T.dz[l - 1][(size_t)f * B + i] = (__bf16)v;
                }
            for (int s = 0; s < 2; ++s)
                for (int j = 0; j < 4; ++j) {
                    db[q][s][j] = (__bf16)acc[q][2 * s][j];
                    db[q][s][4 + j] = (__bf16)acc[q][2 * s + 1][j];
                }
        }
    }
}

// dW_l[o][i] = sum_n dZ_l[o][n] act_l[i][n], db_l[o] = sum_n dZ_l[o][n]; one workgroup per
// 16x16 tile of one layer, its 4 waves split the batch, fixed-order LDS reduction
struct MlpWgradArgs {
    MlpTrainBufs T;
    float* grad;  // [MLP_NPARAM]
};
__global__ void __launch_bounds__(256) mlp_wgrad_kernel(MlpWgradArgs a) {
    __shared__ float red[4][16][17];
    __shared__ float redb[4][16];
    int tile = blockIdx.x, l = 0;
    for (; l < MLP_LAYERS; ++l) {
        const int nt = (mlp_out(l) / 16) * (mlp_in(l) / 16);
        if (tile < nt) break;
        tile -= nt;
    }
    const int ti = mlp_in(l) / 16;
    const int ot = tile / ti, it = tile - ot * ti;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = lane >> 4;
    const uint32_t B = a.T.B;
    const uint32_t chunk = B / 4;
    const __bf16* dz = a.T.dz[l] + (size_t)(16 * ot + (lane & 15)) * B;
    const __bf16* x = a.T.act[l] + (size_t)(16 * it + (lane & 15)) * B;
    f32x4 acc{0.0f, 0.0f, 0.0f, 0.0f}, accb{0.0f, 0.0f, 0.0f, 0.0f};
    bf16x8 ones;
    for (int j = 0; j < 8; ++j) ones[j] = (__bf16)1.0f;
    for (uint32_t n0 = wv * chunk; n0 < (wv + 1) * chunk; n0 += 32) {
        const bf16x8 av = *(const bf16x8*)(dz + n0 + 8 * g);
        const bf16x8 bv = *(const bf16x8*)(x + n0 + 8 * g);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
        if (it == 0) accb = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, ones, accb, 0, 0, 0);
    }
    for (int r = 0; r < 4; ++r) red[wv][4 * g + r][lane & 15] = acc[r];
    if ((lane & 15) == 0)
        for (int r = 0; r < 4; ++r) redb[wv][4 * g + r] = accb[r];
    __syncthreads();
    const int o = threadIdx.x >> 4, i = threadIdx.x & 15;
    const float v = red[0][o][i] + red[1][o][i] + red[2][o][i] + red[3][o][i];
    a.grad[mlp_woff(l) + (16 * ot + o) * mlp_in(l) + 16 * it + i] = v;
    if (it == 0 && i == 0)
        a.grad[mlp_boff(l) + 16 * ot + o] = redb[0][o] + redb[1][o] + redb[2][o] + redb[3][o];
}

__global__ void mlp_loss_reduce_kernel(const float* part, uint32_t n, float* out) {
    if (threadIdx.x != 0) return;
    float s = 0.0f;
    for (uint32_t i = 0; i < n; ++i) s += part[i];
    *out = s;
}

// Adam (Kingma & Ba 2015) on FP32 master weights; refreshes the bf16 copies W and W^T.
// Padding (output rows 3..15 of the last layer, the 12 pad input columns of the first) stays 0.
struct MlpAdamArgs {
    float* param;
    float* m;
    float* v;
    const float* grad;
    __bf16* w16;   // per layer [out][in], packed like the params without biases
    __bf16* wt16;  // per layer [in][out]
    float lr, beta1, beta2, eps;
    int step;  // 1-based
    // optional exponential moving average of the parameters for inference (ema == nullptr: off):
    // ema_t updates so far including this one (0: reset), bias-corrected copy in ema_hat / w16_ema
    float* ema;
    float* ema_hat;
    __bf16* w16_ema;
    float ema_decay;
    int ema_t;
};
__device__ __forceinline__ bool mlp_locate(int p, int& l, int& o, int& i, bool& is_bias) {
    for (l = 0; l < MLP_LAYERS; ++l) {
        if (p < mlp_boff(l)) {
            const int q = p - mlp_woff(l);
            o = q / mlp_in(l);
            i = q - o * mlp_in(l);
            is_bias = false;
            return true;
        }
        if (p < mlp_woff(l + 1)) {
            o = p - mlp_boff(l);
            i = 0;
            is_bias = true;
            return true;
        }
    }
    return false;
}
__device__ __forceinline__ int mlp_w16off(int l) {
    int off = 0;
    for (int k = 0; k < l; ++k) off += mlp_in(k) * mlp_out(k);
    return off;
}
__global__ void mlp_adam_kernel(MlpAdamArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    int l, o, i;
    bool bias;
    mlp_locate(p, l, o, i, bias);
    const bool pad = (l == MLP_LAYERS - 1 && o >= 3) || (l == 0 && !bias && col_is_pad(i));
    float th = a.param[p];
    if (!pad && a.step > 0) {
        const float gr = a.grad[p];
        const float m = a.beta1 * a.m[p] + (1.0f - a.beta1) * gr;
        const float v = a.beta2 * a.v[p] + (1.0f - a.beta2) * gr * gr;
        a.m[p] = m;
        a.v[p] = v;
        const float mh = m / (1.0f - powf(a.beta1, (float)a.step));
        const float vh = v / (1.0f - powf(a.beta2, (float)a.step));
        th -= a.lr * mh / (sqrtf(vh) + a.eps);
        a.param[p] = th;
    }
    if (pad) th = 0.0f;
    if (!bias) {
        const int off = mlp_w16off(l);
        a.w16[off + o * mlp_in(l) + i] = (__bf16)th;
        a.wt16[off + i * mlp_out(l) + o] = (__bf16)th;
    }
    if (a.ema) {
        float eh = th;
        if (a.ema_t == 0) {
            a.ema[p] = 0.0f;
        } else {
            const float e = a.ema_decay * a.ema[p] + (1.0f - a.ema_decay) * th;
            a.ema[p] = e;
            eh = e / (1.0f - powf(a.ema_decay, (float)a.ema_t));
        }
        if (pad) eh = 0.0f;
        a.ema_hat[p] = eh;
        if (!bias) a.w16_ema[mlp_w16off(l) + o * mlp_in(l) + i] = (__bf16)eh;
    }
}

// FP64 VALU reference forward on the same (bf16-rounded) weights and inputs: the per-image
// error oracle of the learned path (FP64 is used only for this reference)
struct MlpRefArgs {
    const __bf16* w16;
    const float* param;
    const void* rec;
    const float* anchors;
    int fp16;
    uint32_t W, H, L, d;
    const uint32_t* idx;
    uint64_t n;
    double* out;
};
__global__ void mlp_ref_kernel(MlpRefArgs a) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint64_t sm = a.idx ? a.idx[i] : i;
    const VertexIn v = a.fp16 ? load_vertex((const __half*)a.rec, a.anchors, a.W, a.H, a.L, a.d, sm)
                              : load_vertex((const float*)a.rec, nullptr, a.W, a.H, a.L, a.d, sm);
    double x[MLP_IN], y[MLP_HID];
    for (int f = 0; f < MLP_IN; ++f) x[f] = (double)(float)(__bf16)mlp_feature(v, f);
    int in = MLP_IN;
    for (int l = 0; l < MLP_LAYERS; ++l) {
        const int outd = mlp_out(l);
        const __bf16* w = a.w16 + mlp_w16off(l);
        const float* b = a.param + mlp_boff(l);
        for (int o = 0; o < outd; ++o) {
            double s = b[o];
            for (int k = 0; k < in; ++k) s += (double)(float)w[o * in + k] * x[k];
            y[o] = s;
        }
        if (l == MLP_LAYERS - 1) break;
        // the MFMA path rounds each layer's ReLU output to bf16
        for (int o = 0; o < outd; ++o) x[o] = (double)(float)(__bf16)(float)(y[o] > 0.0 ? y[o] : 0.0);
        in = outd;
    }
    a.out[3 * i] = y[0];
    a.out[3 * i + 1] = y[1];
    a.out[3 * i + 2] = y[2];
}

}  // namespace nlpf

// ---------------------------------------------------------------------------------------
// encode-only kernel (tests of the feature encoder)
// ---------------------------------------------------------------------------------------
namespace nlpf {
__global__ void mlp_encode_kernel(MlpInferArgs a, float* feats) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint64_t sm = a.idx ? a.idx[i] : i;
    const VertexIn v = a.fp16 ? load_vertex((const __half*)a.rec, a.anchors, a.W, a.H, a.L, a.d, sm)
                              : load_vertex((const float*)a.rec, nullptr, a.W, a.H, a.L, a.d, sm);
    for (int f = 0; f < MLP_IN; ++f) feats[i * MLP_IN + f] = mlp_feature(v, f);
}
}  // namespace nlpf

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------
#include <random>

struct nlpf_mlp_state {
    float* param = nullptr;  // FP32 master [MLP_NPARAM]
    float* m = nullptr;
    float* v = nullptr;
    float* grad = nullptr;
    __bf16* w16 = nullptr;
    __bf16* wt16 = nullptr;
    int step = 0;
    // inference weights: an exponential moving average of the parameters when ema_decay > 0
    float* ema = nullptr;
    float* ema_hat = nullptr;
    __bf16* w16_ema = nullptr;
    float ema_decay = 0.0f;
    int ema_t = 0;
    // training workspace
    uint32_t cap = 0;
    __bf16* act = nullptr;
    __bf16* dz = nullptr;
    float* y = nullptr;
    float* loss_part = nullptr;
    float* loss = nullptr;
};

namespace nlpf {
nlpf_mlp_state* mlp_state(nlpf_handle* h, bool create) {
    if (!h->mlp && create) h->mlp = new nlpf_mlp_state();
    return h->mlp;
}
static void mlp_free_train(nlpf_mlp_state* s) {
    void* ps[] = {s->act, s->dz, s->y, s->loss_part, s->loss};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    s->act = nullptr;
    s->dz = nullptr;
    s->y = nullptr;
    s->loss_part = nullptr;
    s->loss = nullptr;
    s->cap = 0;
}
void mlp_release(nlpf_handle* h) {
    nlpf_mlp_state* s = h->mlp;
    if (!s) return;
    mlp_free_train(s);
    void* ps[] = {s->param, s->m, s->v, s->grad, s->w16, s->wt16, s->ema, s->ema_hat, s->w16_ema};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    delete s;
    h->mlp = nullptr;
}
static constexpr int MLP_NW16 = MLP_NPARAM - (5 * MLP_HID + MLP_OUT);

static MlpParamsDev mlp_dev_params(nlpf_mlp_state* s) {
    MlpParamsDev P;
    int off = 0;
    for (int l = 0; l < MLP_LAYERS; ++l) {
        P.w[l] = s->w16 + off;
        P.wt[l] = s->wt16 + off;
        P.b[l] = s->param + mlp_boff(l);
        off += mlp_in(l) * mlp_out(l);
    }
    return P;
}
// weights used by inference: the moving average when it is on
static MlpParamsDev mlp_infer_params(nlpf_mlp_state* s) {
    MlpParamsDev P = mlp_dev_params(s);
    if (s->ema_decay > 0.0f) {
        int off = 0;
        for (int l = 0; l < MLP_LAYERS; ++l) {
            P.w[l] = s->w16_ema + off;
            P.b[l] = s->ema_hat + mlp_boff(l);
            off += mlp_in(l) * mlp_out(l);
        }
    }
    return P;
}

static int mlp_refresh(nlpf_handle* h, nlpf_mlp_state* s, float lr, int step) {
    MlpAdamArgs a;
    a.param = s->param;
    a.m = s->m;
    a.v = s->v;
    a.grad = s->grad;
    a.w16 = s->w16;
    a.wt16 = s->wt16;
    a.lr = lr;
    a.beta1 = 0.9f;
    a.beta2 = 0.999f;
    a.eps = 1e-8f;
    a.step = step;
    a.ema = s->ema_decay > 0.0f ? s->ema : nullptr;
    a.ema_hat = s->ema_hat;
    a.w16_ema = s->w16_ema;
    a.ema_decay = s->ema_decay;
    if (step == 0) s->ema_t = 0;
    else s->ema_t += 1;
    a.ema_t = s->ema_t;
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, h->stream, a);
    return check_launch(h, "mlp adam");
}

static int mlp_need(nlpf_handle* h, nlpf_mlp_state** out) {
    nlpf_mlp_state* s = mlp_state(h, false);
    if (!s || !s->param) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "MLP not initialised (call nlpf_mlp_init)");
    *out = s;
    return NLPF_OK;
}

static MlpInferArgs infer_args(nlpf_mlp_state* s, const nlpf_screen_desc* desc, const void* rec, const float* anc,
                               uint32_t depth, const uint32_t* idx, const float* feats, uint64_t n, float* out) {
    MlpInferArgs a{};
    a.P = mlp_infer_params(s);
    a.rec = rec;
    a.anchors = desc && desc->precision == NLPF_FP16 ? anc : nullptr;
    a.fp16 = desc && desc->precision == NLPF_FP16;
    a.W = desc ? desc->width : 0;
    a.H = desc ? desc->height : 0;
    a.L = desc ? desc->levels : 0;
    a.d = depth;
    a.idx = idx;
    a.feats = feats;
    a.n = n;
    a.out = out;
    a.sample0 = 0;
    a.out_slice0 = 0;
    return a;
}

// persistent grid of at most 3 workgroups per CU (LDS-bound)
static void launch_infer(nlpf_handle* h, const MlpInferArgs& a) {
    const uint64_t plane = (uint64_t)a.W * a.H;
    const uint64_t per_block = 4 * 16 * MLP_QT;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((a.n + per_block - 1) / per_block, 768);
    (void)plane;
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(blocks), dim3(256), 0, h->stream, a);
}

static int check_rec_args(nlpf_handle* h, const nlpf_screen_desc* desc, const void* rec, const float* feats,
                          uint32_t depth) {
    if (feats) return NLPF_OK;
    if (!desc || !rec) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "MLP: records + desc or raw features required");
    if (desc->precision == NLPF_FP64) NLPF_FAIL(h, NLPF_E_UNSUPPORTED, "MLP reads FP32 or FP16 records");
    if (depth >= desc->levels) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "MLP: depth outside the records' levels");
    return NLPF_OK;
}

}  // namespace nlpf

using namespace nlpf;

extern "C" {

uint64_t nlpf_mlp_param_count(void) { return MLP_NPARAM; }

int nlpf_mlp_init(nlpf_handle* h, uint64_t seed) {
    if (!h) return NLPF_E_INVALID_ARG;
    nlpf_mlp_state* s = mlp_state(h, true);
    if (!s->param) {
        NLPF_HIP(h, hipMalloc(&s->param, sizeof(float) * MLP_NPARAM));
        NLPF_HIP(h, hipMalloc(&s->m, sizeof(float) * MLP_NPARAM));
        NLPF_HIP(h, hipMalloc(&s->v, sizeof(float) * MLP_NPARAM));
        NLPF_HIP(h, hipMalloc(&s->grad, sizeof(float) * MLP_NPARAM));
        NLPF_HIP(h, hipMalloc(&s->w16, sizeof(__bf16) * MLP_NW16));
        NLPF_HIP(h, hipMalloc(&s->wt16, sizeof(__bf16) * MLP_NW16));
        NLPF_HIP(h, hipMalloc(&s->ema, sizeof(float) * MLP_NPARAM));
        NLPF_HIP(h, hipMalloc(&s->ema_hat, sizeof(float) * MLP_NPARAM));
        NLPF_HIP(h, hipMalloc(&s->w16_ema, sizeof(__bf16) * MLP_NW16));
    }
    // He-uniform for the ReLU layers, half the range for the linear output layer; biases 0
    std::vector<float> p(MLP_NPARAM, 0.0f);
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    for (int l = 0; l < MLP_LAYERS; ++l) {
        const int fan_in = l == 0 ? 84 : mlp_in(l);
        const float a = (l == MLP_LAYERS - 1 ? 0.5f : 1.0f) * std::sqrt(6.0f / (float)fan_in);
        for (int o = 0; o < mlp_out(l); ++o)
            for (int i = 0; i < mlp_in(l); ++i) {
                const bool pad = (l == MLP_LAYERS - 1 && o >= 3) || (l == 0 && col_is_pad(i));
                p[mlp_woff(l) + o * mlp_in(l) + i] = pad ? 0.0f : a * u(rng);
            }
    }
    NLPF_HIP(h, hipMemcpy(s->param, p.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
    NLPF_HIP(h, hipMemset(s->m, 0, sizeof(float) * MLP_NPARAM));
    NLPF_HIP(h, hipMemset(s->v, 0, sizeof(float) * MLP_NPARAM));
    NLPF_HIP(h, hipMemset(s->grad, 0, sizeof(float) * MLP_NPARAM));
    s->step = 0;
    int rc = mlp_refresh(h, s, 0.0f, 0);
    if (rc) return rc;
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    return NLPF_OK;
}

int nlpf_mlp_set_ema(nlpf_handle* h, float decay) {
    if (!h) return NLPF_E_INVALID_ARG;
    nlpf_mlp_state* s;
    int rc = mlp_need(h, &s);
    if (rc) return rc;
    if (!(decay >= 0.0f && decay < 1.0f)) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "EMA decay must be in [0, 1)");
    s->ema_decay = decay;
    rc = mlp_refresh(h, s, 0.0f, 0);  // restarts the average at the current parameters
    if (rc) return rc;
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    return NLPF_OK;
}

int nlpf_mlp_get_params(nlpf_handle* h, float* host, uint64_t n) {
    if (!h || !host) return NLPF_E_INVALID_ARG;
    nlpf_mlp_state* s;
    int rc = mlp_need(h, &s);
    if (rc) return rc;
    if (n != (uint64_t)MLP_NPARAM) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "parameter count mismatch");
    NLPF_HIP(h, hipMemcpyAsync(host, s->param, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, h->stream));
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    return NLPF_OK;
}

int nlpf_mlp_set_params(nlpf_handle* h, const float* host, uint64_t n) {
    if (!h || !host) return NLPF_E_INVALID_ARG;
    nlpf_mlp_state* s;
    int rc = mlp_need(h, &s);
    if (rc) return rc;
    if (n != (uint64_t)MLP_NPARAM) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "parameter count mismatch");
    NLPF_HIP(h, hipMemcpyAsync(s->param, host, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, h->stream));
    NLPF_HIP(h, hipMemsetAsync(s->m, 0, sizeof(float) * MLP_NPARAM, h->stream));
    NLPF_HIP(h, hipMemsetAsync(s->v, 0, sizeof(float) * MLP_NPARAM, h->stream));
    s->step = 0;
    rc = mlp_refresh(h, s, 0.0f, 0);
    if (rc) return rc;
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    return NLPF_OK;
}

int nlpf_mlp_encode(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records, const float* anchors,
                    uint32_t depth, const uint32_t* idx, uint64_t n, float* feats) {
    if (!h) return NLPF_E_INVALID_ARG;
    if (!feats) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_mlp_encode: NULL output");
    int rc = check_rec_args(h, desc, records, nullptr, depth);
    if (rc) return rc;
    if (n == 0) return NLPF_OK;
    nlpf_mlp_state dummy;
    MlpInferArgs a = infer_args(&dummy, desc, records, anchors, depth, idx, nullptr, n, nullptr);
    hipLaunchKernelGGL(mlp_encode_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, h->stream, a, feats);
    NLPF_HIP(h, hipGetLastError());
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    return NLPF_OK;
}

int nlpf_mlp_infer(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records, const float* anchors,
                   uint32_t depth, const uint32_t* idx, const float* feats, uint64_t n, float* out_rgb) {
    if (!h) return NLPF_E_INVALID_ARG;
    nlpf_mlp_state* s;
    int rc = mlp_need(h, &s);
    if (rc) return rc;
    if (!out_rgb) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_mlp_infer: NULL output");
    rc = check_rec_args(h, desc, records, feats, depth);
    if (rc) return rc;
    if (n == 0) return NLPF_OK;
    MlpInferArgs a = infer_args(s, desc, records, anchors, depth, idx, feats, n, out_rgb);
    rc = record_event(h, 0);
    if (rc) return rc;
    launch_infer(h, a);
    NLPF_HIP(h, hipGetLastError());
    rc = record_event(h, 1);
    if (rc) return rc;
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    h->stats.total_ms = event_ms(h, 0, 1);
    return NLPF_OK;
}

}  // extern "C"

namespace nlpf {
// Enqueues one training step on the handle's stream (no host synchronisation); the batch's
// loss is left in device memory (mlp_loss_device).
int mlp_train_enqueue(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records, const float* anchors,
                      uint32_t depth, const uint32_t* idx, const float* feats, const float* target, uint32_t batch,
                      float lr, int apply_update, float* grad) {
    nlpf_mlp_state* s;
    int rc = mlp_need(h, &s);
    if (rc) return rc;
    if (!target) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_mlp_train_step: NULL target");
    if (batch == 0 || batch % 128 != 0) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "batch must be a positive multiple of 128");
    if (!feats && !idx) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_mlp_train_step: sample ids or features required");
    rc = check_rec_args(h, desc, records, feats, depth);
    if (rc) return rc;
    if (s->cap < batch) {
        NLPF_HIP(h, hipStreamSynchronize(h->stream));
        mlp_free_train(s);
        const size_t act_rows = MLP_IN + (MLP_LAYERS - 1) * MLP_HID;
        const size_t dz_rows = (MLP_LAYERS - 1) * MLP_HID + MLP_OUT;
        NLPF_HIP(h, hipMalloc(&s->act, sizeof(__bf16) * act_rows * batch));
        NLPF_HIP(h, hipMalloc(&s->dz, sizeof(__bf16) * dz_rows * batch));
        NLPF_HIP(h, hipMalloc(&s->y, sizeof(float) * 3 * batch));
        NLPF_HIP(h, hipMalloc(&s->loss_part, sizeof(float) * (batch / 64 + 1)));
        NLPF_HIP(h, hipMalloc(&s->loss, sizeof(float)));
        s->cap = batch;
    }
    MlpTrainArgs a{};
    a.P = mlp_dev_params(s);
    {
        size_t off = 0;
        for (int l = 0; l < MLP_LAYERS; ++l) {
            a.T.act[l] = s->act + off * batch;
            off += mlp_in(l);
        }
        off = 0;
        for (int l = 0; l < MLP_LAYERS; ++l) {
            a.T.dz[l] = s->dz + off * batch;
            off += mlp_out(l);
        }
    }
    a.T.y = s->y;
    a.T.loss_part = s->loss_part;
    a.T.B = batch;
    a.rec = records;
    a.anchors = desc && desc->precision == NLPF_FP16 ? anchors : nullptr;
    a.fp16 = desc && desc->precision == NLPF_FP16;
    a.W = desc ? desc->width : 0;
    a.H = desc ? desc->height : 0;
    a.L = desc ? desc->levels : 0;
    a.d = depth;
    a.idx = idx;
    a.feats = feats;
    a.target = target;
    const uint32_t waves = batch / 64, blocks = (waves + 3) / 4;
    hipLaunchKernelGGL(mlp_train_fwd_kernel, dim3(blocks), dim3(256), 0, h->stream, a);
    hipLaunchKernelGGL(mlp_loss_reduce_kernel, dim3(1), dim3(64), 0, h->stream, s->loss_part, waves, s->loss);
    hipLaunchKernelGGL(mlp_train_bwd_kernel, dim3(blocks), dim3(256), 0, h->stream, a);
    MlpWgradArgs wa;
    wa.T = a.T;
    wa.grad = s->grad;
    int ntiles = 0;
    for (int l = 0; l < MLP_LAYERS; ++l) ntiles += (mlp_out(l) / 16) * (mlp_in(l) / 16);
    hipLaunchKernelGGL(mlp_wgrad_kernel, dim3(ntiles), dim3(256), 0, h->stream, wa);
    if ((rc = check_launch(h, "mlp train"))) return rc;
    if (grad) NLPF_HIP(h, hipMemcpyAsync(grad, s->grad, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToDevice, h->stream));
    if (apply_update) {
        s->step += 1;
        rc = mlp_refresh(h, s, lr, s->step);
        if (rc) return rc;
    }
    return NLPF_OK;
}

const float* mlp_loss_device(nlpf_handle* h) {
    nlpf_mlp_state* s = mlp_state(h, false);
    return s ? s->loss : nullptr;
}
}  // namespace nlpf

extern "C" {

// One Adam step on a batch (batch % 128 == 0).  loss (host, optional) receives the batch's
// mean squared error before the update; grad (device, optional) receives dL/dparams.
int nlpf_mlp_train_step(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records, const float* anchors,
                        uint32_t depth, const uint32_t* idx, const float* feats, const float* target, uint32_t batch,
                        float lr, int apply_update, float* loss, float* grad) {
    if (!h) return NLPF_E_INVALID_ARG;
    int rc = mlp_train_enqueue(h, desc, records, anchors, depth, idx, feats, target, batch, lr, apply_update, grad);
    if (rc) return rc;
    if (loss) NLPF_HIP(h, hipMemcpyAsync(loss, mlp_loss_device(h), sizeof(float), hipMemcpyDeviceToHost, h->stream));
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    return NLPF_OK;
}

int nlpf_mlp_reference(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records, const float* anchors,
                       uint32_t depth, const uint32_t* idx, uint64_t n, double* out_rgb) {
    if (!h) return NLPF_E_INVALID_ARG;
    nlpf_mlp_state* s;
    int rc = mlp_need(h, &s);
    if (rc) return rc;
    if (!out_rgb) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_mlp_reference: NULL output");
    rc = check_rec_args(h, desc, records, nullptr, depth);
    if (rc) return rc;
    if (n == 0) return NLPF_OK;
    MlpRefArgs a;
    // the weights inference uses (the moving average when it is on)
    a.w16 = s->ema_decay > 0.0f ? s->w16_ema : s->w16;
    a.param = s->ema_decay > 0.0f ? s->ema_hat : s->param;
    a.rec = records;
    a.anchors = desc->precision == NLPF_FP16 ? anchors : nullptr;
    a.fp16 = desc->precision == NLPF_FP16;
    a.W = desc->width;
    a.H = desc->height;
    a.L = desc->levels;
    a.d = depth;
    a.idx = idx;
    a.n = n;
    a.out = out_rgb;
    hipLaunchKernelGGL(mlp_ref_kernel, dim3((uint32_t)((n + 127) / 128)), dim3(128), 0, h->stream, a);
    NLPF_HIP(h, hipGetLastError());
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    return NLPF_OK;
}

}  // extern "C"

namespace nlpf {
// internal entry points used by the learned-cache NL-PF pass (mlp_run.hip)
// estimates for samples sample0..sample0+n-1 written into slot `out_slot` of planar level data
// [slice - out_slice0][out_levels][3][W*H] (the samples' slices must be >= out_slice0)
int mlp_infer_planar(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records, const float* anchors,
                     uint32_t depth, uint64_t sample0, uint64_t n, float* planar_out, uint32_t out_levels,
                     uint32_t out_slot, uint32_t out_slice0) {
    nlpf_mlp_state* s;
    int rc = mlp_need(h, &s);
    if (rc) return rc;
    if (n == 0) return NLPF_OK;
    MlpInferArgs a = infer_args(s, desc, records, anchors, depth, nullptr, nullptr, n, planar_out);
    a.planar = 1;
    a.out_levels = out_levels;
    a.out_slot = out_slot;
    a.sample0 = sample0;
    a.out_slice0 = out_slice0;
    launch_infer(h, a);
    return check_launch(h, "mlp infer");
}
}  // namespace nlpf
