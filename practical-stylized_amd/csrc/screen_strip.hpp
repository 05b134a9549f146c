// Screen-mode NL-PF level kernel over column-strip segments.
//
// The tile kernel (screen_impl.hpp) stages a 64 x 8 output tile plus its halo for every tile,
// so at radius r each 8 output rows restage 8 + 2r rows, and every record load pays 64-bit
// address arithmetic.  Here a workgroup walks down a segment of a 64-column strip of one
// (slice, depth) image, 8 output rows per step, and keeps the step's window of 8 + 2r staged
// rows in an LDS ring: after the first step's window, a step stages only the 8 rows the next
// step adds, over the 8 ring rows it was the last to read.  Record planes are read through
// buffer descriptors (plane offset in an SGPR, one 32-bit lane offset), so a load costs no
// VALU address arithmetic.
//
// A segment is seg_steps steps (64 rows by default); consecutive workgroups take horizontally
// adjacent segments, so the segments resident at once cover a compact band of the image (the
// 2r halo rows are restaged once per segment).  The per-vertex arithmetic (features, weights,
// Kahan-folded row groups, epilogue) is the tile kernel's; only the number of rows per Kahan
// group differs: about 30 pairs per group at every radius.
#pragma once
#include <cstdlib>
#include "screen_impl.hpp"

namespace nlpf {

typedef float strip_f2 __attribute__((ext_vector_type(2)));

// raw buffer descriptor over `bytes` bytes at base (stride 0; last word: the gfx9 data format)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t strip_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float strip_ld1(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ strip_f2 strip_ld2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(strip_f2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ void strip_st1(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, (int)soff, 0);
}

// one record pair of one plane: FP32 records load as two floats, FP16 records as one half2 word
template <typename S>
struct StripPair;
template <>
struct StripPair<float> {
    using T = strip_f2;
    static __device__ __forceinline__ T load(__amdgpu_buffer_rsrc_t r, uint32_t elem, uint32_t soff) {
        return strip_ld2(r, elem * 4u, soff);
    }
    static __device__ __forceinline__ float lo(T v) { return v.x; }
    static __device__ __forceinline__ float hi(T v) { return v.y; }
};
template <>
struct StripPair<__half> {
    using T = uint32_t;
    static __device__ __forceinline__ T load(__amdgpu_buffer_rsrc_t r, uint32_t elem, uint32_t soff) {
        return __builtin_amdgcn_raw_buffer_load_b32(r, (int)(elem * 2u), (int)soff, 0);
    }
    static __device__ __forceinline__ float lo(T v) { return (float)__builtin_bit_cast(_Float16, (uint16_t)v); }
    static __device__ __forceinline__ float hi(T v) {
        return (float)__builtin_bit_cast(_Float16, (uint16_t)(v >> 16));
    }
};

// the loads of one staged record pair, in flight between issue and the LDS write
template <typename S>
struct StripLoads {
    typename StripPair<S>::T v[NLPF_SCREEN_PLANES];
    strip_f2 ch[3];
    float an[3];
};

template <typename S, typename A, int RAD, int NW, int R, bool SHADE = false>
__global__ void __launch_bounds__(64 * NW) screen_strip_kernel(ScreenArgs<A> a) {
    static_assert(std::is_same<A, float>::value, "strip kernel: FP32 accumulation");
    static_assert(R <= 2 * RAD + 1, "rows per lane must not exceed the window height");
    static_assert(64 * NW >= NLPF_CHAIN_MAX * 32, "one lane per ramp-table entry");
    constexpr int TW = 64, TH = NW * R, HT = TH + 2 * RAD, NT = 64 * NW;
    constexpr int K = 2 * RAD + 1, NR = 2 * RAD + R;  // window width, input rows per lane
    // staged columns start at an even x (tx0 - RAD - OFF), so a record pair is one aligned load
    constexpr int OFF = RAD & 1, WS = TW + 2 * RAD + 2 * OFF, WP = WS / 2;
    constexpr int NPRO = HT * WP, NITP = (NPRO + NT - 1) / NT;  // pairs staged when a strip starts
    constexpr int NPF = TH * WP;                                 // pairs staged per step
    static_assert(NPF >= NT && NPF <= 2 * NT, "a step stages one or two pairs per thread");
    constexpr int G = 30 / K > 1 ? 30 / K : 1;                  // input rows per Kahan group
    constexpr bool FULL = NR * K <= 100;                         // rows fully unrolled
    constexpr bool HALF = std::is_same<S, __half>::value;
    using PL = StripPair<S>;
    __shared__ V4<A> shA[HT * WS];
    __shared__ V4<A> shB[HT * WS];
    __shared__ V2<A> shC[HT * WS];
    __shared__ float ramp_tabs[NLPF_CHAIN_MAX * 32];
    __shared__ nlpf_style_node shade_nodes[SHADE ? kShadeLdsNodes : 1];  // referenced by SHADE only

    // block -> segment: strips fastest, then bands, then slices, so the segments resident at
    // once are neighbours in the image and share their halo columns in L2
    const uint32_t t = xcd_remap(blockIdx.x, a.nblocks);
    const uint32_t per_slice = a.tiles_x * a.bands;
    uint32_t s = t / per_slice;
    const uint32_t rem = t - s * per_slice;
    s += a.hole1 && s + 1 >= a.hole1 ? 1u : 0u;
    const uint32_t band = rem / a.tiles_x, tx = rem - band * a.tiles_x;
    const uint32_t sy0 = band * a.seg_steps;
    const uint32_t n = min(a.seg_steps, a.steps_y - sy0);
    if (threadIdx.x < NLPF_CHAIN_MAX * 32) {  // visible to the epilogue after the first barrier
        const int i = threadIdx.x >> 5;
        if (a.chain.valid && i < a.chain.n && a.chain.node[i].type == NLPF_STY_RAMP)
            ramp_tabs[threadIdx.x] = a.chain.node[i].p[threadIdx.x & 31];
    }
    if constexpr (SHADE) {  // the style table for the epilogue's walks, 64 words per node
        const uint32_t* src = (const uint32_t*)a.nodes;
        uint32_t* dst = (uint32_t*)shade_nodes;
        for (uint32_t w = threadIdx.x; w < a.lds_nodes * 64u; w += 64 * NW) dst[w] = src[w];
    }
    const int W = (int)a.W, H = (int)a.H;
    const uint32_t plane = (uint32_t)a.plane;
    const uint32_t pbytes = plane * (uint32_t)sizeof(S), obytes = plane * 4u;
    const int lc = threadIdx.x & 63;
    const int r0 = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) * R;
    const A kn = a.kn;
    const bool has_child = a.has_child != 0;

    {
        // ---- the segment: steps sy0 .. sy0 + n - 1 of strip tx in slice s ----
        const S* rec = (const S*)a.rec + ((size_t)s * a.L + a.d) * NLPF_SCREEN_PLANES * plane;
        const auto rr = strip_rsrc(rec, NLPF_SCREEN_PLANES * pbytes);
        const A* child = a.out + ((size_t)s * a.out_levels + (has_child ? a.child_slot : a.out_slot)) * 3 * plane;
        const auto rc = strip_rsrc(child, 3 * obytes);
        const uint32_t abytes = a.abx * a.aby * 12u;
        // FP32 records have no anchors: the descriptor then covers 4 bytes of the output, never read
        const float* anc = HALF ? a.anchors + ((size_t)s * a.L + a.d) * a.abx * a.aby * 3 : (const float*)a.out;
        const auto ra = strip_rsrc(anc, HALF ? abytes : 4u);
        const auto ro = strip_rsrc(a.out + ((size_t)s * a.out_levels + a.out_slot) * 3 * plane, 3 * obytes);
        const int tx0 = (int)tx * TW, xs = tx0 - RAD - OFF;
        const bool cols_in = xs >= 0 && xs + WS <= W;

        // issue the loads of the pair at staged row gy, staged column gx (even); positions
        // outside the image load the nearest in-image pair and are replaced by the sentinel
        auto issue = [&](int gy, int gx, StripLoads<S>& L) __attribute__((always_inline)) {
            const int gxc = min(max(gx, 0), W - 2), gyc = min(max(gy, 0), H - 1);
            const uint32_t o = (uint32_t)gyc * (uint32_t)W + (uint32_t)gxc;
            // the plane offsets are formed here (scalar multiplies) rather than held across the
            // loop: kept live, the 18 of them spill SGPRs
            uint32_t pb = pbytes, ob = obytes;
            asm volatile("" : "+s"(pb), "+s"(ob));
#pragma unroll
            for (int c = 0; c < NLPF_SCREEN_PLANES; ++c) L.v[c] = PL::load(rr, o, (uint32_t)c * pb);
#pragma unroll
            for (int c = 0; c < 3; ++c) L.ch[c] = has_child ? strip_ld2(rc, o * 4u, (uint32_t)c * ob) : strip_f2{0.f, 0.f};
            const uint32_t ab = ((uint32_t)(gyc >> 3) * a.abx + (uint32_t)(gxc >> 3)) * 12u;
#pragma unroll
            for (int c = 0; c < 3; ++c) L.an[c] = HALF ? strip_ld1(ra, ab + 4u * c, 0u) : 0.f;
        };
        // features and level input of a loaded pair into LDS slots idx, idx + 1
        auto put = [&](int idx, bool ok, const StripLoads<S>& L) __attribute__((always_inline)) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                auto f = [&](int c) { return j ? PL::hi(L.v[c]) : PL::lo(L.v[c]); };
                A px = f(0), py = f(1), pz = f(2);
                if (HALF) {
                    px += L.an[0];
                    py += L.an[1];
                    pz += L.an[2];
                }
                A ir = f(9) * a.rscale, ig = f(10) * a.rscale, ib = f(11) * a.rscale;
                if (has_child) {
                    ir = fma(f(12), j ? L.ch[0].y : L.ch[0].x, ir);
                    ig = fma(f(13), j ? L.ch[1].y : L.ch[1].x, ig);
                    ib = fma(f(14), j ? L.ch[2].y : L.ch[2].x, ib);
                }
                V4<A> fa{px * a.kp, py * a.kp, pz * a.kp, lum709(v3<A>{f(6), f(7), f(8)}) * a.kl};
                V4<A> fb{f(3), f(4), f(5), ir};
                V2<A> fc{ig, ib};
                if (!ok) {  // outside the image: a far-away guide position gives weight exactly 0
                    fa = V4<A>{(A)1e18, (A)1e18, (A)1e18, (A)0};
                    fb = V4<A>{(A)0, (A)0, (A)0, (A)0};
                    fc = V2<A>{(A)0, (A)0};
                }
                shA[idx + j] = fa;
                shB[idx + j] = fb;
                shC[idx + j] = fc;
            }
        };

        // ---- stage the first step's window: image rows y0 - RAD .. y0 + TH + RAD - 1 ----
        // ring slot of image row y: (y + RAD) mod HT; b = slot of the window's first row
        uint32_t b = (sy0 * TH) % HT;
        {
            const int yw = (int)sy0 * TH - RAD;
            // two pairs in flight per thread: the window is staged once per strip
#pragma unroll 1
            for (int it0 = 0; it0 < NITP; it0 += 2) {
                StripLoads<S> L[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) {  // every lane loads (lanes past the end: the last pair)
                    const int p = min((it0 + u) * NT + (int)threadIdx.x, NPRO - 1);
                    const int jr = p / WP, cp = p - jr * WP;
                    issue(yw + jr, xs + 2 * cp, L[u]);
                }
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int p = (it0 + u) * NT + (int)threadIdx.x;
                    if (p < NPRO) {
                        const int jr = p / WP, cp = p - jr * WP;
                        const int gy = yw + jr, gx = xs + 2 * cp;
                        int sl = (int)b + jr;
                        sl -= sl >= HT ? HT : 0;
                        put(sl * WS + 2 * cp, gy >= 0 && gy < H && gx >= 0 && gx + 1 < W, L[u]);
                    }
                }
            }
        }
        __syncthreads();

        for (uint32_t i = 0; i < n; ++i) {
            const int y0 = (int)(sy0 + i) * TH;
            const bool more = i + 1 < n;
            // ---- gather over the ring ----
            A cpx[R], cpy[R], cpz[R], cnx[R], cny[R], cnz[R], cl[R];
            A Wt[R], Sx[R], Sy[R], Sz[R], cW[R], cX[R], cY[R], cZ[R];
            A rw[R], rx[R], ry[R], rz[R];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                int sl = (int)b + r0 + k + RAD;
                sl -= sl >= HT ? HT : 0;
                const V4<A> ca = shA[sl * WS + lc + RAD + OFF];
                const V4<A> cb = shB[sl * WS + lc + RAD + OFF];
                cpx[k] = ca.x;
                cpy[k] = ca.y;
                cpz[k] = ca.z;
                cl[k] = ca.w;
                cnx[k] = kn * cb.x;
                cny[k] = kn * cb.y;
                cnz[k] = kn * cb.z;
                Wt[k] = Sx[k] = Sy[k] = Sz[k] = cW[k] = cX[k] = cY[k] = cZ[k] = (A)0;
                rw[k] = rx[k] = ry[k] = rz[k] = (A)0;
            }
            // input row r of the lane's strip feeds outputs k with 0 <= r - k <= 2 RAD
            auto row = [&](int r, int klo, int khi) __attribute__((always_inline)) {
                int sl = (int)b + r0 + r;
                sl -= sl >= HT ? HT : 0;
                const V4<A>* pa = shA + sl * WS + lc + OFF;
                const V4<A>* pb = shB + sl * WS + lc + OFF;
                const V2<A>* pc = shC + sl * WS + lc + OFF;
#pragma unroll
                for (int dx = 0; dx < K; ++dx) {
                    const V4<A> na = pa[dx];
                    const V4<A> nb = pb[dx];
                    const V2<A> nc = pc[dx];
#pragma unroll
                    for (int k = 0; k < R; ++k) {
                        if (k >= klo && k <= khi) {
                            const A ex = cpx[k] - na.x, ey = cpy[k] - na.y, ez = cpz[k] - na.z;
                            A e = ex * ex;
                            e = fma(ey, ey, e);
                            e = fma(ez, ez, e);
                            const A tt = fma(-cnz[k], nb.z, fma(-cny[k], nb.y, fma(-cnx[k], nb.x, kn)));
                            e = fma(tt, tt, e);
                            const A dl = cl[k] - na.w;
                            e = fma(dl, dl, e);
                            const A w = exp2_neg(e);
                            rw[k] += w;
                            rx[k] = fma(w, nb.w, rx[k]);
                            ry[k] = fma(w, nc.x, ry[k]);
                            rz[k] = fma(w, nc.y, rz[k]);
                        }
                    }
                }
            };
            auto flush = [&]() __attribute__((always_inline)) {
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    kahan_add(Wt[k], cW[k], rw[k]);
                    kahan_add(Sx[k], cX[k], rx[k]);
                    kahan_add(Sy[k], cY[k], ry[k]);
                    kahan_add(Sz[k], cZ[k], rz[k]);
                    rw[k] = rx[k] = ry[k] = rz[k] = (A)0;
                }
            };
            if constexpr (FULL) {
                // groups of G input rows, each folded into the Kahan totals
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    row(r, r - 2 * RAD > 0 ? r - 2 * RAD : 0, r < R - 1 ? r : R - 1);
                    if ((r + 1) % G == 0 || r == NR - 1) flush();
                }
            } else {
                // edge rows unrolled; the middle rows, which feed every output, one per loop
                // iteration with a Kahan fold after every second (a two-row body, whose rows
                // may wrap around the ring, was scheduled into register spills)
#pragma unroll
                for (int r = 0; r < R - 1; ++r) row(r, 0, r);
                if (R > 1) flush();
#pragma unroll 1
                for (int r = R - 1; r <= 2 * RAD; ++r) {
                    row(r, 0, R - 1);
                    if (((r - (R - 1)) & 1) || r == 2 * RAD) flush();
                }
#pragma unroll
                for (int r = 0; r < R - 1; ++r) row(2 * RAD + 1 + r, r + 1, R - 1);
                if (R > 1) flush();
            }

            // ---- epilogue: normalise, stylize, store (rolled over the lane's outputs) ----
            v3<A> Fo[R];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const A inv = m_div((A)1, Wt[k]);  // Wt includes the centre's own weight
                Fo[k] = v3<A>{Sx[k] * inv, Sy[k] * inv, Sz[k] * inv};
            }
            const int ox = tx0 + lc;
#pragma unroll 1
            for (int k = 0; k < R; ++k) {
                const int oy = y0 + r0 + k;
                v3<A> F = Fo[0];
#pragma unroll
                for (int j = 1; j < R; ++j)
                    if (k == j) F = Fo[j];
                if (ox < W && oy < H) {
                    const uint32_t o = (uint32_t)oy * (uint32_t)W + (uint32_t)ox;
                    v3<A> p{(A)0, (A)0, (A)0};
                    if (a.need_pos) {
                        const uint32_t e = o * (uint32_t)sizeof(S);
                        if (HALF) {
                            auto h = [&](uint32_t soff) {
                                const uint32_t w2 = __builtin_amdgcn_raw_buffer_load_b32(rr, (int)(e & ~3u), (int)soff, 0);
                                return (e & 2u) ? PL::hi(w2) : PL::lo(w2);
                            };
                            const uint32_t ab = ((uint32_t)(oy >> 3) * a.abx + (uint32_t)(ox >> 3)) * 12u;
                            p = v3<A>{h(0) + strip_ld1(ra, ab, 0u), h(pbytes) + strip_ld1(ra, ab + 4u, 0u),
                                      h(2 * pbytes) + strip_ld1(ra, ab + 8u, 0u)};
                        } else {
                            p = v3<A>{strip_ld1(rr, e, 0u), strip_ld1(rr, e, pbytes), strip_ld1(rr, e, 2 * pbytes)};
                        }
                    }
                    F = screen_stylize<A, SHADE>(a, s, o, F, p, ramp_tabs, shade_nodes);
                    strip_st1(ro, o * 4u, 0u, F.x);
                    strip_st1(ro, o * 4u, obytes, F.y);
                    strip_st1(ro, o * 4u, 2 * obytes, F.z);
                }
            }
            if (more) {
                // ---- stage the next step's new rows y0 + TH + RAD .. y0 + 2 TH + RAD - 1 ----
                // into the slots of this window's first TH rows; the pairs beyond the first NT
                // go to one wave, a different one every step
                const int rot = (int)(i & (NW - 1)) * 64;
                __syncthreads();  // every lane is done with the ring rows being replaced
                // pair p of the step's NPF: row jr, pair column cp; its ring slot row follows b
                auto put_pair = [&](int p, const StripLoads<S>& L) __attribute__((always_inline)) {
                    const int jr = p / WP, cp = p - jr * WP;
                    const int gy = y0 + TH + RAD + jr, gx = xs + 2 * cp;
                    int sl = (int)b + jr;
                    sl -= sl >= HT ? HT : 0;
                    put(sl * WS + 2 * cp, (cols_in || (gx >= 0 && gx + 1 < W)) && gy < H, L);
                };
                auto issue_pair = [&](int p, StripLoads<S>& L) __attribute__((always_inline)) {
                    const int jr = p / WP, cp = p - jr * WP;
                    issue(y0 + TH + RAD + jr, xs + 2 * cp, L);
                };
                const int p0 = ((int)threadIdx.x + rot) & (NT - 1);
                StripLoads<S> L0;
                issue_pair(p0, L0);
                if (p0 + NT < NPF) {  // the pairs past the first NT: part of one wave
                    StripLoads<S> L1;
                    issue_pair(p0 + NT, L1);
                    put_pair(p0 + NT, L1);
                }
                put_pair(p0, L0);
                b += TH;
                b -= b >= (uint32_t)HT ? (uint32_t)HT : 0u;
                __syncthreads();  // the new rows are visible to the next step's gather
            } else {
                __syncthreads();  // the next strip's window may overwrite the ring
            }
        }
    }
}

// buffer offsets are 32-bit: the strip kernel takes even widths whose (slice, depth) record
// block, output slot and anchor block each stay below 2^31 bytes; masked launches keep the
// tile kernel (masks describe its 64 x 8 tiles)
template <typename S, typename A>
inline bool screen_strip_eligible(const ScreenArgs<A>& a) {
    const uint64_t plane = a.plane;
    const uint64_t lim = 1ull << 31;
    return !a.tile_mask && (a.W & 1) == 0 && a.W >= 2 && plane * NLPF_SCREEN_PLANES * sizeof(S) < lim &&
           plane * 3 * sizeof(A) < lim && (uint64_t)a.abx * a.aby * 12 < lim;
}

// steps (8-row blocks) per workgroup: 8, i.e. 64 output rows; NLPF_STRIP_SEGMENT overrides
// it (1..64, for tuning)
inline uint32_t strip_segment_steps() {
    static int v = 0;
    if (v == 0) {
        const char* e = getenv("NLPF_STRIP_SEGMENT");
        const int x = e ? atoi(e) : 8;
        v = x >= 1 && x <= 64 ? x : 8;
    }
    return (uint32_t)v;
}

template <typename S, typename A, int RAD, int NW, int R, bool SHADE = false>
hipError_t launch_screen_strip(const ScreenArgs<A>& a0, hipStream_t st) {
    constexpr int TW = 64, TH = NW * R;
    if (!screen_strip_eligible<S>(a0)) return hipErrorInvalidValue;
    if (std::is_same<S, __half>::value && !a0.anchors) return hipErrorInvalidValue;
    ScreenArgs<A> a = a0;
    a.tiles_x = (a.W + TW - 1) / TW;
    a.steps_y = (a.H + TH - 1) / TH;
    a.seg_steps = strip_segment_steps();
    a.bands = (a.steps_y + a.seg_steps - 1) / a.seg_steps;
    const uint64_t nb = (uint64_t)a.tiles_x * a.bands * a.S;
    if (nb >= (1ull << 31)) return hipErrorInvalidValue;
    a.nblocks = (uint32_t)nb;
    hipLaunchKernelGGL((screen_strip_kernel<S, A, RAD, NW, R, SHADE>), dim3(a.nblocks), dim3(64 * NW), 0, st, a);
    return hipGetLastError();
}

}  // namespace nlpf
