// Screen-mode NL-PF level kernel: LDS-tiled (2r+1)^2 cross-bilateral gather with the
// stylization epilogue fused.
//
// One launch filters depth d of every spp slice.  A workgroup owns a 64 x (NW*R) output tile
// of one (slice, depth) image: lanes map to the 64 columns, wave w owns rows [w*R, w*R+R),
// and each lane keeps R vertically adjacent outputs in registers so that every neighbour it
// reads from LDS is reused by up to R outputs.  The tile plus a RAD-pixel halo is staged in
// LDS once, as pre-scaled guide features (p*kp, n, lum(albedo)*kl) and the level input
// in = radiance + throughput * out_{d+1}.  Weights are exp2 of a negated quadratic form whose
// log2(e)/(2 sigma^2) factors are folded into kp, kn, kl, so one v_exp_f32 per pair remains.
// Row partial sums (2r+1 terms) are folded into Kahan-compensated totals.
#pragma once
#include <type_traits>
#include "common.hpp"
#include "style.hpp"

namespace nlpf {

// level-kernel variants: kScreenAuto takes each radius's fastest, kScreenTiles each radius's
// fastest tile shape, kScreenStripsAuto its fastest strip shape; 0..2 are tile shapes,
// kScreenStrips and kScreenStrips + 1 the column-strip kernel with 2 and 3 rows per lane
constexpr int kScreenAuto = -1, kScreenTiles = -2, kScreenStripsAuto = -3, kScreenStrips = 3;
int screen_variant_of(const Config& cfg);

template <typename A>
struct ScreenArgs {
    const void* rec;
    const float* anchors;
    A* out;  // [S][out_levels][3][H][W]
    const nlpf_style_node* nodes;
    int root;
    int need_pos;
    uint32_t W, H, S, L, d;
    uint32_t hole1;  // 1 + a slice the launch passes over (later slices move up by one); 0: none
    uint32_t out_levels, out_slot, child_slot;
    int has_child;
    uint64_t plane;
    uint32_t tiles_x, tiles_y, nblocks;
    uint32_t steps_y, seg_steps, bands;  // strip kernel: steps per strip, steps per segment, segments per strip
    uint32_t abx, aby;  // anchor blocks per row / column (fp16)
    A kp, kn, kl;
    A rscale;  // radiance decode factor (FP16 records store radiance / NLPF_FP16_RADIANCE_SCALE)
    const uint8_t* tile_mask;  // optional [ceil(H/8)][ceil(W/64)]: 0 = skip the 64x8 tile (all slices)
    StyleChain chain;          // root's chain for flags 0 by value (not valid: walk nodes from root)
    // optional per-vertex shading, [S][L][H][W] like one record plane: vertex i is stylized by
    // material[i]'s tree (roots[material[i]], identity for materials >= n_materials or root -1)
    // with flags[i] (0 when NULL); NULL material: `root` / `chain` with flags 0 for every vertex
    const uint16_t* material;
    const uint8_t* flags;
    const int32_t* roots;
    uint32_t n_materials;
    uint32_t lds_nodes;  // SHADE instances: style-table nodes staged in LDS (0: walk `nodes` in memory)
};

// style-table nodes the SHADE instances stage in LDS (4 KB); larger tables are walked in memory
constexpr uint32_t kShadeLdsNodes = 16;

// the epilogue's stylization of vertex o of slice s: per-vertex material and flags in the
// SHADE instances (launched only when the launch has a material plane), else the launch's one
// tree (resolved chain or node walk) with flags 0.  A compile-time switch: a run-time branch on
// a.material costs the 15x15 kernels 2 % even when it is not taken.
template <typename A, bool SHADE>
__device__ __forceinline__ v3<A> screen_stylize(const ScreenArgs<A>& a, uint32_t s, size_t o, v3<A> F, v3<A> p,
                                                const float* ramp_tabs, const nlpf_style_node* lds_nodes) {
    if constexpr (SHADE) {
        const size_t q = ((size_t)s * a.L + a.d) * a.plane + o;
        const uint32_t m = a.material[q];
        const int root = m < a.n_materials ? a.roots[m] : -1;
        const uint32_t fl = a.flags ? (uint32_t)a.flags[q] : 0u;
        if (root >= 0) {
            // neighbouring vertices walk different trees: per-lane reads, from LDS when staged
            if (a.lds_nodes)
                F = apply_style<A>(lds_nodes, root, a.d, fl, F, p);
            else
                F = apply_style<A>(a.nodes, root, a.d, fl, F, p);
        }
    } else if (a.chain.valid) {
        F = apply_chain<A>(a.chain, a.d, F, p, ramp_tabs);
    } else if (a.root >= 0) {
        F = apply_style<A>(a.nodes, a.root, a.d, 0u, F, p);
    }
    return F;
}

__device__ __forceinline__ float ld_rec(const float* p, size_t i) { return p[i]; }
__device__ __forceinline__ float ld_rec(const __half* p, size_t i) { return __half2float(p[i]); }
__device__ __forceinline__ double ld_rec(const double* p, size_t i) { return p[i]; }

// two adjacent elements (i even, so the pair is 2-element aligned)
__device__ __forceinline__ void ld_rec2(const float* p, size_t i, float& a, float& b) {
    const float2 t = *(const float2*)(p + i);
    a = t.x;
    b = t.y;
}
__device__ __forceinline__ void ld_rec2(const __half* p, size_t i, float& a, float& b) {
    const __half2 t = *(const __half2*)(p + i);
    a = __low2float(t);
    b = __high2float(t);
}
__device__ __forceinline__ void ld_rec2(const double* p, size_t i, double& a, double& b) {
    const double2 t = *(const double2*)(p + i);
    a = t.x;
    b = t.y;
}

__device__ __forceinline__ float exp2_neg(float e) { return __builtin_amdgcn_exp2f(-e); }
__device__ __forceinline__ double exp2_neg(double e) { return exp2(-e); }

template <typename T>
__device__ __forceinline__ void kahan_add(T& sum, T& comp, T v) {
    T y = v - comp;
    T t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// block -> tile map: the dispatcher hands block b to XCD (b mod 8); give every XCD a
// contiguous run of tiles so neighbouring tiles (which share halos) meet in one L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb) {
    const uint32_t per = nb >> 3;
    return b < (per << 3) ? (b & 7u) * per + (b >> 3) : b;
}

template <typename A>
struct alignas(4 * sizeof(A)) V4 {
    A x, y, z, w;
};
template <typename A>
struct alignas(2 * sizeof(A)) V2 {
    A x, y;
};

template <typename S, typename A, int RAD, int NW, int R, bool SHADE = false>
__global__ void __launch_bounds__(64 * NW) screen_level_kernel(ScreenArgs<A> a) {
    static_assert(R <= 2 * RAD + 1, "rows per lane must not exceed the window height");
    static_assert(64 * NW >= NLPF_CHAIN_MAX * 32, "one lane per ramp-table entry");
    constexpr int TW = 64, TH = NW * R, HT = TH + 2 * RAD, NT = 64 * NW;
    constexpr int K = 2 * RAD + 1;
    // staged columns start at an even x (pairs of records load as one aligned 2-element
    // vector when W is even): XS0 = tx0 - RAD - OFF, the window column tx0 - RAD sits at OFF
    constexpr int OFF = RAD & 1, WS = TW + 2 * RAD + 2 * OFF, WP = WS / 2;
    // guide features and level input of the tile + halo: (p*kp, lum*kl), (n, in.r), (in.g, in.b)
    constexpr int NPAIR = HT * WP, NIT = (NPAIR + NT - 1) / NT;
    constexpr bool HALF = std::is_same<S, __half>::value;  // FP16 records carry position anchors
    __shared__ V4<A> shA[HT * WS + 2];
    __shared__ V4<A> shB[HT * WS + 2];
    __shared__ V2<A> shC[HT * WS + 2];
    __shared__ float ramp_tabs[NLPF_CHAIN_MAX * 32];  // the chain's ramp stops, for per-lane lookups
    __shared__ nlpf_style_node shade_nodes[SHADE ? kShadeLdsNodes : 1];  // referenced by SHADE only

    const uint32_t t = xcd_remap(blockIdx.x, a.nblocks);
    const uint32_t tiles = a.tiles_x * a.tiles_y;
    uint32_t s = t / tiles;
    const uint32_t tile = t - s * tiles;
    s += a.hole1 && s + 1 >= a.hole1 ? 1u : 0u;
    if (a.tile_mask && !a.tile_mask[tile]) return;  // uniform per workgroup, before any barrier
    const int tx0 = (int)(tile % a.tiles_x) * TW, ty0 = (int)(tile / a.tiles_x) * TH;
    const int W = (int)a.W, H = (int)a.H;
    const size_t plane = a.plane;
    const S* rec = (const S*)a.rec + (size_t)((size_t)s * a.L + a.d) * NLPF_SCREEN_PLANES * plane;
    const A* child = a.has_child ? a.out + ((size_t)s * a.out_levels + a.child_slot) * 3 * plane : nullptr;
    const float* anc = a.anchors ? a.anchors + (size_t)((size_t)s * a.L + a.d) * a.abx * a.aby * 3 : nullptr;
    const bool even_w = (W & 1) == 0;
    if (threadIdx.x < NLPF_CHAIN_MAX * 32) {  // visible to the epilogue after the staging barrier
        const int i = threadIdx.x >> 5;
        if (a.chain.valid && i < a.chain.n && a.chain.node[i].type == NLPF_STY_RAMP)
            ramp_tabs[threadIdx.x] = a.chain.node[i].p[threadIdx.x & 31];
    }
    if constexpr (SHADE) {  // the style table for the epilogue's walks, 64 words per node
        const uint32_t* src = (const uint32_t*)a.nodes;
        uint32_t* dst = (uint32_t*)shade_nodes;
        for (uint32_t w = threadIdx.x; w < a.lds_nodes * 64u; w += 64 * NW) dst[w] = src[w];
    }

    // ---- stage tile + halo in LDS, two horizontally adjacent records per thread ----
    // put(): guide features and level input of one loaded record pair into LDS slots
    // idx, idx + 1; records outside the image get a far-away guide position, which drives
    // their weight to exactly 0
    auto put = [&](int idx, bool ok0, bool ok1, int gy, int gx, const A (&v)[NLPF_SCREEN_PLANES][2],
                   const A (&ch)[3][2], bool has_child) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            V4<A> fa, fb;
            V2<A> fc;
            if (j == 0 ? ok0 : ok1) {
                A px = v[0][j], py = v[1][j], pz = v[2][j];
                if (HALF) {
                    const float* q = anc + ((size_t)(gy >> 3) * a.abx + ((gx + j) >> 3)) * 3;
                    px += (A)q[0];
                    py += (A)q[1];
                    pz += (A)q[2];
                }
                A ir = v[9][j] * a.rscale, ig = v[10][j] * a.rscale, ib = v[11][j] * a.rscale;
                if (has_child) {
                    ir = fma(v[12][j], ch[0][j], ir);
                    ig = fma(v[13][j], ch[1][j], ig);
                    ib = fma(v[14][j], ch[2][j], ib);
                }
                fa = V4<A>{px * a.kp, py * a.kp, pz * a.kp, lum709(v3<A>{v[6][j], v[7][j], v[8][j]}) * a.kl};
                fb = V4<A>{v[3][j], v[4][j], v[5][j], ir};
                fc = V2<A>{ig, ib};
            } else {
                fa = V4<A>{(A)1e18, (A)1e18, (A)1e18, (A)0};
                fb = V4<A>{(A)0, (A)0, (A)0, (A)0};
                fc = V2<A>{(A)0, (A)0};
            }
            shA[idx + j] = fa;
            shB[idx + j] = fb;
            shC[idx + j] = fc;
        }
    };
    if (HALF && even_w) {
        // FP16 records, even width and even staged columns: the two records of a pair are both
        // inside or both outside the image. Every pair loads from its position clamped into
        // the image (one aligned 2-element load per plane) and out-of-image pairs take the
        // sentinel, so the loop body is branch-free and the loads of consecutive iterations are
        // in flight together. Threads past the last pair write two spare LDS slots. (FP16
        // passes run 1.5-3 % faster this way; the same form with FP32 records ran 4-11 %
        // slower than the loop below, so FP32 and FP64 keep that loop.)
        auto stage_even = [&](auto has_child_c) __attribute__((always_inline)) {
            constexpr bool HC = decltype(has_child_c)::value;
#pragma unroll
            for (int it = 0; it < NIT; ++it) {
                const int i0 = (int)threadIdx.x + it * NT;
                const bool live = NPAIR % NT == 0 || i0 < NPAIR;
                const int i = live ? i0 : NPAIR - 1;
                const int r = i / WP, cp = i - r * WP;
                const int gx = tx0 - RAD - OFF + 2 * cp, gy = ty0 - RAD + r;
                const bool ok = gy >= 0 && gy < H && gx >= 0 && gx < W;
                const int gxc = min(max(gx, 0), W - 2), gyc = min(max(gy, 0), H - 1);
                const size_t o = (size_t)gyc * W + gxc;
                A v[NLPF_SCREEN_PLANES][2], ch[3][2] = {};
#pragma unroll
                for (int c = 0; c < NLPF_SCREEN_PLANES; ++c) ld_rec2(rec, c * plane + o, v[c][0], v[c][1]);
                if (HC) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) ld_rec2(child, c * plane + o, ch[c][0], ch[c][1]);
                }
                put(live ? r * WS + 2 * cp : HT * WS, ok, ok, gyc, gxc, v, ch, HC);
            }
        };
        if (child)
            stage_even(std::true_type{});
        else
            stage_even(std::false_type{});
    } else {
        for (int i = threadIdx.x; i < NPAIR; i += NT) {
            const int r = i / WP, cp = i - r * WP;
            const int gx = tx0 - RAD - OFF + 2 * cp, gy = ty0 - RAD + r;
            const bool row_ok = gy >= 0 && gy < H;
            const bool ok0 = row_ok && gx >= 0 && gx < W, ok1 = row_ok && gx + 1 >= 0 && gx + 1 < W;
            A v[NLPF_SCREEN_PLANES][2], ch[3][2];
            if (ok0 && ok1 && even_w) {
                const size_t o = (size_t)gy * W + gx;
#pragma unroll
                for (int c = 0; c < NLPF_SCREEN_PLANES; ++c) ld_rec2(rec, c * plane + o, v[c][0], v[c][1]);
                if (child) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) ld_rec2(child, c * plane + o, ch[c][0], ch[c][1]);
                }
            } else {
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    if (j == 0 ? ok0 : ok1) {
                        const size_t o = (size_t)gy * W + (gx + j);
#pragma unroll
                        for (int c = 0; c < NLPF_SCREEN_PLANES; ++c) v[c][j] = (A)ld_rec(rec, c * plane + o);
                        if (child) {
#pragma unroll
                            for (int c = 0; c < 3; ++c) ch[c][j] = child[c * plane + o];
                        }
                    }
                }
            }
            put(r * WS + 2 * cp, ok0, ok1, gy, gx, v, ch, child != nullptr);
        }
    }
    __syncthreads();

    // ---- gather ----
    // Each lane owns R vertically adjacent outputs; input row r of the lane's strip feeds the
    // outputs k with 0 <= r - k <= 2 RAD.  The edge rows (r < R-1 or r > 2 RAD) run in fully
    // unrolled loops where that test is a compile-time constant; the middle rows feed every
    // output, so no branch sits between the R independent dependency chains.  Row partials
    // are plain FP32 sums over up to two rows, folded into Kahan-compensated totals.
    const int lc = threadIdx.x & 63;
    const int r0 = (threadIdx.x >> 6) * R;
    const A kn = a.kn;
    // centre normals pre-scaled by kn: kn (1 - n_c.n_j) = kn - (kn n_c).n_j
    A cpx[R], cpy[R], cpz[R], cnx[R], cny[R], cnz[R], cl[R];
    A Wt[R], Sx[R], Sy[R], Sz[R], cW[R], cX[R], cY[R], cZ[R];
    A rw[R], rx[R], ry[R], rz[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const V4<A> ca = shA[(r0 + k + RAD) * WS + lc + RAD + OFF];
        const V4<A> cb = shB[(r0 + k + RAD) * WS + lc + RAD + OFF];
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
    auto row = [&](int r, int klo, int khi) __attribute__((always_inline)) {
        const V4<A>* pa = shA + (r0 + r) * WS + lc + OFF;
        const V4<A>* pb = shB + (r0 + r) * WS + lc + OFF;
        const V2<A>* pc = shC + (r0 + r) * WS + lc + OFF;
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
#pragma unroll
    for (int r = 0; r < R - 1; ++r) row(r, 0, r);
    if (R > 1) flush();
    for (int r = R - 1; r + 1 <= 2 * RAD; r += 2) {
        row(r, 0, R - 1);
        row(r + 1, 0, R - 1);
        flush();
    }
    if ((2 * RAD + 2 - R) % 2) {
        row(2 * RAD, 0, R - 1);
        flush();
    }
#pragma unroll
    for (int i = 0; i < R - 1; ++i) row(2 * RAD + 1 + i, i + 1, R - 1);
    if (R > 1) flush();

    // ---- epilogue: normalise, stylize, store ----
    // The outputs are normalised first; one rolled loop then stylizes and stores them, taking
    // output k by selects, so the stylization code exists once and no register array is
    // indexed dynamically.
    A* outp = a.out + ((size_t)s * a.out_levels + a.out_slot) * 3 * plane;
    v3<A> Fo[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const A inv = m_div((A)1, Wt[k]);  // Wt includes the centre's own weight, 1 for unit normals
        Fo[k] = v3<A>{Sx[k] * inv, Sy[k] * inv, Sz[k] * inv};
    }
#pragma unroll 1
    for (int k = 0; k < R; ++k) {
        const int ox = tx0 + lc, oy = ty0 + r0 + k;
        v3<A> F = Fo[0];
#pragma unroll
        for (int j = 1; j < R; ++j)
            if (k == j) F = Fo[j];
        if (ox < W && oy < H) {
            const size_t o = (size_t)oy * W + ox;
            v3<A> p{(A)0, (A)0, (A)0};
            if (a.need_pos) {
                p = v3<A>{(A)ld_rec(rec, o), (A)ld_rec(rec, plane + o), (A)ld_rec(rec, 2 * plane + o)};
                if (anc) {
                    const float* q = anc + ((size_t)(oy >> 3) * a.abx + (ox >> 3)) * 3;
                    p.x += (A)q[0];
                    p.y += (A)q[1];
                    p.z += (A)q[2];
                }
            }
            F = screen_stylize<A, SHADE>(a, s, o, F, p, ramp_tabs, shade_nodes);
            outp[o] = F.x;
            outp[plane + o] = F.y;
            outp[2 * plane + o] = F.z;
        }
    }
}

// film[y][x][c] = mean over slices of out_0 (deterministic slice order).  The slices can come
// in consecutive groups: `accumulate` continues the running sums the previous group left in
// film, `finalize` scales them by 1 / s_total; the sums are then the ones a single launch over
// all slices forms.
template <typename A>
__global__ void screen_film_kernel(const A* out, uint32_t S, uint32_t out_levels, uint32_t slot0,
                                   uint64_t plane, A* film, uint32_t s_total, int accumulate, int finalize) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= plane) return;
    A acc[3] = {(A)0, (A)0, (A)0};
    if (accumulate) {
        acc[0] = film[3 * i];
        acc[1] = film[3 * i + 1];
        acc[2] = film[3 * i + 2];
    }
    for (uint32_t s = 0; s < S; ++s) {
        const A* o = out + ((size_t)s * out_levels + slot0) * 3 * plane;
        acc[0] += o[i];
        acc[1] += o[plane + i];
        acc[2] += o[2 * plane + i];
    }
    const A inv = finalize ? (A)1 / (A)s_total : (A)1;
    film[3 * i] = acc[0] * inv;
    film[3 * i + 1] = acc[1] * inv;
    film[3 * i + 2] = acc[2] * inv;
}

template <typename S, typename A, int RAD, int NW, int R, bool SHADE = false>
hipError_t launch_screen_level(const ScreenArgs<A>& a0, hipStream_t st) {
    constexpr int TW = 64, TH = NW * R;
    if (a0.tile_mask && TH != 8) return hipErrorInvalidValue;  // masks cover 64x8 tiles
    if (std::is_same<S, __half>::value && !a0.anchors) return hipErrorInvalidValue;
    ScreenArgs<A> a = a0;
    a.tiles_x = (a.W + TW - 1) / TW;
    a.tiles_y = (a.H + TH - 1) / TH;
    a.nblocks = a.tiles_x * a.tiles_y * a.S;
    hipLaunchKernelGGL((screen_level_kernel<S, A, RAD, NW, R, SHADE>), dim3(a.nblocks), dim3(64 * NW), 0, st, a);
    return hipGetLastError();
}

}  // namespace nlpf
