// Adjoint of the screen-mode NL-PF with respect to the filter bandwidths (sigma_pos,
// sigma_normal, sigma_lum) and the stop values of every colour-ramp node, given dLoss/dfilm.
//
// Forward, per level d and vertex i:  in_j = rad_j + thr_j * out_{d+1,j},
//   w_ij = 2^-(e_p + e_n + e_l),  e_k = log2(e) * term_k / (2 sigma_k^2),
//   F_i = sum_j w_ij in_j / W_i,  out_i = S_d(F_i).
// Reverse, level 0 -> D:
//   (A) per i: recompute F_i and the sigma moments A_k = sum_j w e_k, B_k = sum_j w e_k in_j
//       (weights are recomputed per tile, never stored); h_i = J_S(F_i)^T g_i;
//       dL/dsigma_k += (2 ln2 / sigma_k) h_i . (B_k - F_i A_k) / W_i;  ramp stops from J_S;
//       u_i = h_i / W_i.
//   (B) per j: g_{d+1,j} = thr_j * sum_i w_ij u_i   (w symmetric, windows symmetric).
// Parameter gradients are reduced per workgroup into fixed slots and summed in a fixed order,
// so the result is bitwise reproducible.
#include "common.hpp"
#include "dual.hpp"
#include "screen_impl.hpp"

namespace nlpf {

#define ADJ_MAX_RAMPS 2
#define ADJ_NP (3 + 16 * ADJ_MAX_RAMPS)
#define ADJ_MAX_CHAIN 8

// the stylization chain of material 0 with flags 0, resolved on the host and passed by value
// (node[0] applies first); slot[k]: ramp ordinal of node k, or -1
struct AdjChain {
    int32_t n;
    int32_t slot[ADJ_MAX_CHAIN];
    nlpf_style_node node[ADJ_MAX_CHAIN];
};

struct BwdArgs {
    const float* rec;       // FP32 record planes [S][L][15][H][W]
    const float* level_out; // forward outputs [S][L][3][H][W]
    const float* g_in;      // dL/dout_d   [S][3][H][W]
    float* g_out;           // dL/dout_{d+1} [S][3][H][W] (pass B)
    float* u;               // h / W        [S][3][H][W]
    float* partial;         // [nblocks][ADJ_NP] accumulated over levels
    int32_t np;             // live parameters: 3 + 16 * (ramp nodes); the rest stay 0
    uint32_t W, H, S, L, d;
    uint64_t plane;
    uint32_t tiles_x, tiles_y, nblocks;
    float kp, kn, kl;
    float sig[3];
    int has_child;
    AdjChain chain;
};

// h = J_S(F)^T g; ramp stop gradients accumulated into pg[3 + 16 * slot + k].  tabs: the chain
// nodes' 32 parameters in LDS ([ADJ_MAX_CHAIN][32]) for the per-lane ramp lookups.
__device__ __forceinline__ void style_vjp(const BwdArgs& a, v3<float> F, v3<float> g, float* pg, v3<float>& h,
                                          const float* tabs) {
    const int n = a.chain.n;
    const v3<float> p0{0.0f, 0.0f, 0.0f};
    for (int k = ADJ_MAX_CHAIN - 1; k >= 0; --k) {
        if (k < n) {
            const nlpf_style_node& nd = a.chain.node[k];
            const float* tab = tabs + 32 * k;
            // node k's input, recomputed from F through nodes 0..k-1 (a per-node array of the
            // inputs, indexed by the loop counter, would live in scratch memory)
            v3<float> xi = F;
            for (int j = 0; j < k; ++j) xi = apply_node<float>(a.chain.node[j], a.d, xi, p0, tabs + 32 * j);
            if (nd.type == NLPF_STY_RAMP && depth_active(nd.depth_mask, a.d)) {
                const int slot = a.chain.slot[k];
                const v3<float> c = clamp3(xi, 0.0f, 100000.0f);
                const float l = lum709(c);
                if (slot >= 0 && l > 0.0f) {
                    const float q = (g.x * c.x + g.y * c.y + g.z * c.z) / l;
                    const uint32_t cnt = nd.count < NLPF_RAMP_STOPS ? nd.count : NLPF_RAMP_STOPS;
                    const uint32_t kk = ramp_segment(tab, cnt, l);  // the forward's segment
                    float w0 = 0.0f, w1 = 0.0f;
                    int i0 = 0, i1 = 0;
                    if (kk == 0) {
                        w0 = 1.0f;
                    } else if (kk == cnt) {
                        i0 = (int)cnt - 1;
                        w0 = 1.0f;
                    } else {
                        const float t = (l - tab[kk - 1]) / (tab[kk] - tab[kk - 1]);
                        i0 = (int)kk - 1;
                        i1 = (int)kk;
                        w0 = 1.0f - t;
                        w1 = t;
                    }
#pragma unroll
                    for (int r = 0; r < ADJ_MAX_RAMPS; ++r) {
                        if (r == slot) {
#pragma unroll
                            for (int m = 0; m < 16; ++m)
                                pg[3 + 16 * r + m] += (m == i0 ? w0 * q : 0.0f) + (m == i1 ? w1 * q : 0.0f);
                        }
                    }
                }
            }
            const v3<Dual3> xd{Dual3::var(xi.x, 0), Dual3::var(xi.y, 1), Dual3::var(xi.z, 2)};
            const v3<Dual3> pd{Dual3(0.0f), Dual3(0.0f), Dual3(0.0f)};
            const v3<Dual3> yd = apply_node<Dual3>(nd, a.d, xd, pd, tab);
            v3<float> gn;
            gn.x = g.x * yd.x.d[0] + g.y * yd.y.d[0] + g.z * yd.z.d[0];
            gn.y = g.x * yd.x.d[1] + g.y * yd.y.d[1] + g.z * yd.z.d[1];
            gn.z = g.x * yd.x.d[2] + g.y * yd.y.d[2] + g.z * yd.z.d[2];
            g = gn;
        }
    }
    h = g;
}

// Tile + halo of one (slice, depth) image in LDS, as in the forward level kernel: two
// horizontally adjacent records per thread (one aligned 2-element load per plane when W is
// even), packed as (p*kp, lum*kl), (n, v.x), (v.y, v.z).  Pass A stages v = in = radiance +
// throughput * out_{d+1}; pass B stages v = u.
template <int RAD, int NW, int R, bool PASS_A>
__device__ __forceinline__ void bwd_stage(const BwdArgs& a, uint32_t s, int tx0, int ty0, V4<float>* shA,
                                          V4<float>* shB, V2<float>* shC) {
    constexpr int TW = 64, TH = NW * R, HT = TH + 2 * RAD, NT = 64 * NW;
    constexpr int OFF = RAD & 1, WS = TW + 2 * RAD + 2 * OFF, WP = WS / 2;
    const int W = (int)a.W, H = (int)a.H;
    const size_t plane = a.plane;
    const float* rec = a.rec + (size_t)((size_t)s * a.L + a.d) * NLPF_SCREEN_PLANES * plane;
    const float* child = PASS_A && a.has_child ? a.level_out + ((size_t)s * a.L + a.d + 1) * 3 * plane : nullptr;
    const float* up = PASS_A ? nullptr : a.u + (size_t)s * 3 * plane;
    const bool even_w = (W & 1) == 0;
    constexpr int NPL = PASS_A ? NLPF_SCREEN_PLANES : 9;  // planes read from the records
    for (int i = threadIdx.x; i < HT * WP; i += NT) {
        const int r = i / WP, cp = i - r * WP;
        const int gx = tx0 - RAD - OFF + 2 * cp, gy = ty0 - RAD + r;
        const bool row_ok = gy >= 0 && gy < H;
        const bool ok0 = row_ok && gx >= 0 && gx < W, ok1 = row_ok && gx + 1 >= 0 && gx + 1 < W;
        float v[NLPF_SCREEN_PLANES][2], ch[3][2];
        if (ok0 && ok1 && even_w) {
            const size_t o = (size_t)gy * W + gx;
#pragma unroll
            for (int c = 0; c < NPL; ++c) ld_rec2(rec, c * plane + o, v[c][0], v[c][1]);
            if (PASS_A) {
                if (child) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) ld_rec2(child, c * plane + o, ch[c][0], ch[c][1]);
                }
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) ld_rec2(up, c * plane + o, ch[c][0], ch[c][1]);
            }
        } else {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (j == 0 ? ok0 : ok1) {
                    const size_t o = (size_t)gy * W + (gx + j);
#pragma unroll
                    for (int c = 0; c < NPL; ++c) v[c][j] = rec[c * plane + o];
                    const float* src = PASS_A ? child : up;
                    if (src) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) ch[c][j] = src[c * plane + o];
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            V4<float> fa, fb;
            V2<float> fc;
            if (j == 0 ? ok0 : ok1) {
                float vx, vy, vz;
                if (PASS_A) {
                    vx = v[9][j];
                    vy = v[10][j];
                    vz = v[11][j];
                    if (child) {
                        vx = fmaf(v[12][j], ch[0][j], vx);
                        vy = fmaf(v[13][j], ch[1][j], vy);
                        vz = fmaf(v[14][j], ch[2][j], vz);
                    }
                } else {
                    vx = ch[0][j];
                    vy = ch[1][j];
                    vz = ch[2][j];
                }
                fa = V4<float>{v[0][j] * a.kp, v[1][j] * a.kp, v[2][j] * a.kp,
                               lum709(v3<float>{v[6][j], v[7][j], v[8][j]}) * a.kl};
                fb = V4<float>{v[3][j], v[4][j], v[5][j], vx};
                fc = V2<float>{vy, vz};
            } else {
                fa = V4<float>{1e18f, 1e18f, 1e18f, 0.0f};
                fb = V4<float>{0.0f, 0.0f, 0.0f, 0.0f};
                fc = V2<float>{0.0f, 0.0f};
            }
            shA[r * WS + 2 * cp + j] = fa;
            shB[r * WS + 2 * cp + j] = fb;
            shC[r * WS + 2 * cp + j] = fc;
        }
    }
}

// Pass A.  Lanes own R vertically adjacent outputs; the middle rows of the window feed every
// output without a branch, the edge rows are unrolled with compile-time output ranges.
template <int RAD, int NW, int R>
__global__ void __launch_bounds__(64 * NW) screen_bwd_a_kernel(BwdArgs a) {
    constexpr int TW = 64, TH = NW * R, HT = TH + 2 * RAD, K = 2 * RAD + 1;
    constexpr int OFF = RAD & 1, WS = TW + 2 * RAD + 2 * OFF;
    __shared__ V4<float> shA[HT * WS];
    __shared__ V4<float> shB[HT * WS];
    __shared__ V2<float> shC[HT * WS];
    __shared__ float red[NW][ADJ_NP];
    __shared__ float tabs[ADJ_MAX_CHAIN * 32];  // chain parameters for per-lane ramp lookups
    static_assert(64 * NW >= ADJ_MAX_CHAIN * 32, "one lane per table entry");
    if (threadIdx.x < ADJ_MAX_CHAIN * 32 && (int)(threadIdx.x >> 5) < a.chain.n)
        tabs[threadIdx.x] = a.chain.node[threadIdx.x >> 5].p[threadIdx.x & 31];  // read after the staging barrier

    const uint32_t t = xcd_remap(blockIdx.x, a.nblocks);
    const uint32_t tiles = a.tiles_x * a.tiles_y;
    const uint32_t s = t / tiles, tile = t - s * tiles;
    const int tx0 = (int)(tile % a.tiles_x) * TW, ty0 = (int)(tile / a.tiles_x) * TH;
    const int W = (int)a.W, H = (int)a.H;
    const size_t plane = a.plane;
    bwd_stage<RAD, NW, R, true>(a, s, tx0, ty0, shA, shB, shC);
    __syncthreads();

    const int lc = threadIdx.x & 63;
    const int r0 = (threadIdx.x >> 6) * R;
    const float kn = a.kn;
    float cpx[R], cpy[R], cpz[R], cnx[R], cny[R], cnz[R], cl[R];
    float Wt[R], Sx[R], Sy[R], Sz[R], cW[R], cX[R], cY[R], cZ[R];
    float rw[R], rx[R], ry[R], rz[R];
    float Am[3][R], Bm[3][3][R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const V4<float> ca = shA[(r0 + k + RAD) * WS + lc + RAD + OFF];
        const V4<float> cb = shB[(r0 + k + RAD) * WS + lc + RAD + OFF];
        cpx[k] = ca.x;
        cpy[k] = ca.y;
        cpz[k] = ca.z;
        cl[k] = ca.w;
        cnx[k] = kn * cb.x;
        cny[k] = kn * cb.y;
        cnz[k] = kn * cb.z;
        Wt[k] = Sx[k] = Sy[k] = Sz[k] = cW[k] = cX[k] = cY[k] = cZ[k] = 0.0f;
        rw[k] = rx[k] = ry[k] = rz[k] = 0.0f;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            Am[q][k] = 0.0f;
            Bm[q][0][k] = Bm[q][1][k] = Bm[q][2][k] = 0.0f;
        }
    }
    auto row = [&](int r, int klo, int khi) __attribute__((always_inline)) {
        const V4<float>* pa = shA + (r0 + r) * WS + lc + OFF;
        const V4<float>* pb = shB + (r0 + r) * WS + lc + OFF;
        const V2<float>* pc = shC + (r0 + r) * WS + lc + OFF;
#pragma unroll
        for (int dx = 0; dx < K; ++dx) {
            const V4<float> na = pa[dx];
            const V4<float> nb = pb[dx];
            const V2<float> nc = pc[dx];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                if (k >= klo && k <= khi) {
                    const float ex = cpx[k] - na.x, ey = cpy[k] - na.y, ez = cpz[k] - na.z;
                    float ep = ex * ex;
                    ep = fmaf(ey, ey, ep);
                    ep = fmaf(ez, ez, ep);
                    const float tt = fmaf(-cnz[k], nb.z, fmaf(-cny[k], nb.y, fmaf(-cnx[k], nb.x, kn)));
                    const float en = tt * tt;
                    const float dl = cl[k] - na.w;
                    const float el = dl * dl;
                    const float w = exp2_neg(ep + en + el);
                    rw[k] += w;
                    rx[k] = fmaf(w, nb.w, rx[k]);
                    ry[k] = fmaf(w, nc.x, ry[k]);
                    rz[k] = fmaf(w, nc.y, rz[k]);
                    const float wp = w * ep, wn = w * en, wl = w * el;
                    Am[0][k] += wp;
                    Am[1][k] += wn;
                    Am[2][k] += wl;
                    Bm[0][0][k] = fmaf(wp, nb.w, Bm[0][0][k]);
                    Bm[0][1][k] = fmaf(wp, nc.x, Bm[0][1][k]);
                    Bm[0][2][k] = fmaf(wp, nc.y, Bm[0][2][k]);
                    Bm[1][0][k] = fmaf(wn, nb.w, Bm[1][0][k]);
                    Bm[1][1][k] = fmaf(wn, nc.x, Bm[1][1][k]);
                    Bm[1][2][k] = fmaf(wn, nc.y, Bm[1][2][k]);
                    Bm[2][0][k] = fmaf(wl, nb.w, Bm[2][0][k]);
                    Bm[2][1][k] = fmaf(wl, nc.x, Bm[2][1][k]);
                    Bm[2][2][k] = fmaf(wl, nc.y, Bm[2][2][k]);
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
            rw[k] = rx[k] = ry[k] = rz[k] = 0.0f;
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

    float pg[ADJ_NP];
#pragma unroll
    for (int q = 0; q < ADJ_NP; ++q) pg[q] = 0.0f;
    const float* gin = a.g_in + (size_t)s * 3 * plane;
    float* upw = a.u + (size_t)s * 3 * plane;
    // The loop over the R outputs stays rolled (the chain's VJP is large) and takes output k's
    // sums by selects: indexed by the loop counter, the per-output arrays would live in scratch.
#pragma unroll 1
    for (int k = 0; k < R; ++k) {
        float wt = Wt[0], sx = Sx[0], sy = Sy[0], sz = Sz[0], am[3], bm[3][3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            am[q] = Am[q][0];
#pragma unroll
            for (int c = 0; c < 3; ++c) bm[q][c] = Bm[q][c][0];
        }
#pragma unroll
        for (int j = 1; j < R; ++j) {
            if (k == j) {
                wt = Wt[j];
                sx = Sx[j];
                sy = Sy[j];
                sz = Sz[j];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    am[q] = Am[q][j];
#pragma unroll
                    for (int c = 0; c < 3; ++c) bm[q][c] = Bm[q][c][j];
                }
            }
        }
        const int ox = tx0 + lc, oy = ty0 + r0 + k;
        if (ox < W && oy < H) {
            const size_t o = (size_t)oy * W + ox;
            const float inv = 1.0f / wt;
            const v3<float> F{sx * inv, sy * inv, sz * inv};
            const v3<float> g{gin[o], gin[plane + o], gin[2 * plane + o]};
            v3<float> hh;
            style_vjp(a, F, g, pg, hh, tabs);
            const float Fv[3] = {F.x, F.y, F.z};
            const float hv[3] = {hh.x, hh.y, hh.z};
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                float acc = 0.0f;
#pragma unroll
                for (int c = 0; c < 3; ++c) acc = fmaf(hv[c], bm[q][c] - Fv[c] * am[q], acc);
                pg[q] += acc * inv * (2.0f * 0.69314718056f / a.sig[q]);
            }
            upw[o] = hh.x * inv;
            upw[plane + o] = hh.y * inv;
            upw[2 * plane + o] = hh.z * inv;
        }
    }
    // deterministic block reduction: wave butterflies, then waves in order
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < ADJ_NP; ++q) {
        if (q < a.np) {  // uniform: the ramp slots the chain does not use are skipped
            float v = pg[q];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if (lane == 0) red[wv][q] = v;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < a.np) {
        float v = 0.0f;
        for (int w2 = 0; w2 < NW; ++w2) v += red[w2][threadIdx.x];
        a.partial[(size_t)blockIdx.x * ADJ_NP + threadIdx.x] += v;
    }
}

// Pass B: g_{d+1, j} = thr_j * sum_i w_ij u_i (the weights are symmetric and so are the windows)
template <int RAD, int NW, int R>
__global__ void __launch_bounds__(64 * NW, 2) screen_bwd_b_kernel(BwdArgs a) {
    constexpr int TW = 64, TH = NW * R, HT = TH + 2 * RAD, K = 2 * RAD + 1;
    constexpr int OFF = RAD & 1, WS = TW + 2 * RAD + 2 * OFF;
    __shared__ V4<float> shA[HT * WS];
    __shared__ V4<float> shB[HT * WS];
    __shared__ V2<float> shC[HT * WS];
    const uint32_t t = xcd_remap(blockIdx.x, a.nblocks);
    const uint32_t tiles = a.tiles_x * a.tiles_y;
    const uint32_t s = t / tiles, tile = t - s * tiles;
    const int tx0 = (int)(tile % a.tiles_x) * TW, ty0 = (int)(tile / a.tiles_x) * TH;
    const int W = (int)a.W, H = (int)a.H;
    const size_t plane = a.plane;
    bwd_stage<RAD, NW, R, false>(a, s, tx0, ty0, shA, shB, shC);
    __syncthreads();
    const int lc = threadIdx.x & 63;
    const int r0 = (threadIdx.x >> 6) * R;
    const float kn = a.kn;
    float cpx[R], cpy[R], cpz[R], cnx[R], cny[R], cnz[R], cl[R];
    float Sx[R], Sy[R], Sz[R], cX[R], cY[R], cZ[R], rx[R], ry[R], rz[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const V4<float> ca = shA[(r0 + k + RAD) * WS + lc + RAD + OFF];
        const V4<float> cb = shB[(r0 + k + RAD) * WS + lc + RAD + OFF];
        cpx[k] = ca.x;
        cpy[k] = ca.y;
        cpz[k] = ca.z;
        cl[k] = ca.w;
        cnx[k] = kn * cb.x;
        cny[k] = kn * cb.y;
        cnz[k] = kn * cb.z;
        Sx[k] = Sy[k] = Sz[k] = cX[k] = cY[k] = cZ[k] = rx[k] = ry[k] = rz[k] = 0.0f;
    }
    auto row = [&](int r, int klo, int khi) __attribute__((always_inline)) {
        const V4<float>* pa = shA + (r0 + r) * WS + lc + OFF;
        const V4<float>* pb = shB + (r0 + r) * WS + lc + OFF;
        const V2<float>* pc = shC + (r0 + r) * WS + lc + OFF;
#pragma unroll
        for (int dx = 0; dx < K; ++dx) {
            const V4<float> na = pa[dx];
            const V4<float> nb = pb[dx];
            const V2<float> nc = pc[dx];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                if (k >= klo && k <= khi) {
                    const float ex = cpx[k] - na.x, ey = cpy[k] - na.y, ez = cpz[k] - na.z;
                    float e = ex * ex;
                    e = fmaf(ey, ey, e);
                    e = fmaf(ez, ez, e);
                    const float tt = fmaf(-cnz[k], nb.z, fmaf(-cny[k], nb.y, fmaf(-cnx[k], nb.x, kn)));
                    e = fmaf(tt, tt, e);
                    const float dl = cl[k] - na.w;
                    e = fmaf(dl, dl, e);
                    const float w = exp2_neg(e);
                    rx[k] = fmaf(w, nb.w, rx[k]);
                    ry[k] = fmaf(w, nc.x, ry[k]);
                    rz[k] = fmaf(w, nc.y, rz[k]);
                }
            }
        }
    };
    auto flush = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < R; ++k) {
            kahan_add(Sx[k], cX[k], rx[k]);
            kahan_add(Sy[k], cY[k], ry[k]);
            kahan_add(Sz[k], cZ[k], rz[k]);
            rx[k] = ry[k] = rz[k] = 0.0f;
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

    const float* rec = a.rec + (size_t)((size_t)s * a.L + a.d) * NLPF_SCREEN_PLANES * plane;
    float* gout = a.g_out + (size_t)s * 3 * plane;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int ox = tx0 + lc, oy = ty0 + r0 + k;
        if (ox < W && oy < H) {
            const size_t o = (size_t)oy * W + ox;
            gout[o] = rec[12 * plane + o] * Sx[k];
            gout[plane + o] = rec[13 * plane + o] * Sy[k];
            gout[2 * plane + o] = rec[14 * plane + o] * Sz[k];
        }
    }
}

__global__ void bwd_seed_kernel(const float* dfilm, float* g0, uint32_t S, uint64_t plane) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= plane * S) return;
    const uint64_t s = i / plane, px = i - s * plane;
    const float inv = 1.0f / (float)S;
    for (int c = 0; c < 3; ++c) g0[(s * 3 + c) * plane + px] = dfilm[3 * px + c] * inv;
}

// Fixed-order sum of the per-workgroup partials: one workgroup per parameter, thread t
// Kahan-sums partials t, t+256, ... in order, then a fixed pairwise tree over the threads.
__global__ void bwd_reduce_kernel(const float* partial, uint32_t nblocks, float* out) {
    __shared__ float sh[256];
    const int q = blockIdx.x;
    float s = 0.0f, c = 0.0f;
    for (uint32_t b = threadIdx.x; b < nblocks; b += 256) kahan_add(s, c, partial[(size_t)b * ADJ_NP + q]);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = sh[0];
}

template <int RAD>
hipError_t launch_bwd(BwdArgs a, bool pass_b, hipStream_t st) {
    constexpr int NW = 4, R = 2, TW = 64, TH = NW * R;
    a.tiles_x = (a.W + TW - 1) / TW;
    a.tiles_y = (a.H + TH - 1) / TH;
    a.nblocks = a.tiles_x * a.tiles_y * a.S;
    if (pass_b)
        hipLaunchKernelGGL((screen_bwd_b_kernel<RAD, NW, R>), dim3(a.nblocks), dim3(64 * NW), 0, st, a);
    else
        hipLaunchKernelGGL((screen_bwd_a_kernel<RAD, NW, R>), dim3(a.nblocks), dim3(64 * NW), 0, st, a);
    return hipGetLastError();
}

hipError_t screen_bwd_level(const BwdArgs& a, int radius, bool pass_b, hipStream_t st) {
    switch (radius) {
        case 1: return launch_bwd<1>(a, pass_b, st);
        case 2: return launch_bwd<2>(a, pass_b, st);
        case 3: return launch_bwd<3>(a, pass_b, st);
        case 4: return launch_bwd<4>(a, pass_b, st);
        case 5: return launch_bwd<5>(a, pass_b, st);
        case 6: return launch_bwd<6>(a, pass_b, st);
        case 7: return launch_bwd<7>(a, pass_b, st);
        default: return hipErrorInvalidValue;
    }
}

uint32_t screen_bwd_blocks(uint32_t W, uint32_t H, uint32_t S) {
    return ((W + 63) / 64) * ((H + 7) / 8) * S;
}

}  // namespace nlpf

using namespace nlpf;

extern "C" int nlpf_screen_backward(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records,
                                    const void* level_out, const float* dL_dfilm, float* grad_params) {
    if (!h) return NLPF_E_INVALID_ARG;
    if (!desc || !records || !level_out || !dL_dfilm || !grad_params)
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_screen_backward: NULL argument");
    if (desc->precision != NLPF_FP32) NLPF_FAIL(h, NLPF_E_UNSUPPORTED, "adjoint runs on FP32 records");
    if (desc->width == 0 || desc->height == 0 || desc->spp == 0 || desc->levels == 0 || desc->levels > 31 ||
        desc->radius < 1 || desc->radius > 7)
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_screen_backward: bad screen desc");
    // ramp ordinals, chain length limits
    std::vector<int32_t> slots(h->h_nodes.size(), -1);
    int nr = 0;
    for (size_t i = 0; i < h->h_nodes.size(); ++i)
        if (h->h_nodes[i].type == NLPF_STY_RAMP) slots[i] = nr++;
    if (nr > ADJ_MAX_RAMPS) NLPF_FAIL(h, NLPF_E_UNSUPPORTED, "adjoint supports at most 2 ramp nodes");
    const int root = h->h_roots.empty() ? -1 : h->h_roots[0];
    {
        int n = 0, cur = root;
        while (cur >= 0 && n <= NLPF_STYLE_STACK) {
            const nlpf_style_node& nd = h->h_nodes[cur];
            if (nd.type == NLPF_STY_ENABLE_IF) cur = nd.idx == 0 ? nd.inner : nd.else_;
            else {
                if (nd.type != NLPF_STY_SET_IDX && nd.type != NLPF_STY_BASE) ++n;
                cur = nd.inner;
            }
        }
        if (n > ADJ_MAX_CHAIN) NLPF_FAIL(h, NLPF_E_UNSUPPORTED, "adjoint supports stylization chains of <= 8 nodes");
    }
    const uint64_t plane = (uint64_t)desc->width * desc->height;
    const uint32_t S = desc->spp, L = desc->levels;
    const uint32_t nb = screen_bwd_blocks(desc->width, desc->height, S);
    // scratch: g (2 slots), u, partials
    const size_t gbytes = (size_t)S * 3 * plane * sizeof(float);
    const size_t need = 3 * gbytes + (size_t)nb * ADJ_NP * sizeof(float) + 256;
    int rc = ensure_scratch(h, need);
    if (rc) return rc;
    char* base = (char*)h->scratch;
    float* g0 = (float*)base;
    float* g1 = (float*)(base + gbytes);
    float* u = (float*)(base + 2 * gbytes);
    float* partial = (float*)(base + 3 * gbytes);
    hipStream_t st = h->stream;
    NLPF_HIP(h, hipMemsetAsync(partial, 0, (size_t)nb * ADJ_NP * sizeof(float), st));
    hipLaunchKernelGGL(bwd_seed_kernel, dim3((uint32_t)((plane * S + 255) / 256)), dim3(256), 0, st, dL_dfilm, g0, S,
                       plane);
    const double log2e = 1.4426950408889634;
    BwdArgs a{};
    a.rec = (const float*)records;
    a.level_out = (const float*)level_out;
    a.partial = partial;
    a.np = 3 + 16 * nr;
    a.u = u;
    a.W = desc->width;
    a.H = desc->height;
    a.S = S;
    a.L = L;
    a.plane = plane;
    a.kp = (float)sqrt(log2e / (2.0 * (double)h->cfg.sigma_pos * h->cfg.sigma_pos));
    a.kn = (float)sqrt(log2e / (2.0 * (double)h->cfg.sigma_normal * h->cfg.sigma_normal));
    a.kl = (float)sqrt(log2e / (2.0 * (double)h->cfg.sigma_lum * h->cfg.sigma_lum));
    a.sig[0] = h->cfg.sigma_pos;
    a.sig[1] = h->cfg.sigma_normal;
    a.sig[2] = h->cfg.sigma_lum;
    {
        int seq[NLPF_STYLE_STACK];
        const int n = root >= 0 ? resolve_style(h->h_nodes.data(), root, 0u, seq) : 0;
        a.chain.n = n;  // n <= ADJ_MAX_CHAIN, checked above
        for (int k = 0; k < n; ++k) {
            a.chain.node[k] = h->h_nodes[seq[n - 1 - k]];
            a.chain.slot[k] = slots[seq[n - 1 - k]];
        }
    }
    rc = record_event(h, 0);
    if (rc) return rc;
    float* gin = g0;
    float* gout = g1;
    for (uint32_t d = 0; d < L; ++d) {
        a.d = d;
        a.has_child = d + 1 < L;
        a.g_in = gin;
        a.g_out = gout;
        NLPF_HIP(h, screen_bwd_level(a, desc->radius, false, st));
        if (d + 1 < L) NLPF_HIP(h, screen_bwd_level(a, desc->radius, true, st));
        std::swap(gin, gout);
    }
    // the caller's array holds 3 + 16 * (number of ramp nodes) floats
    hipLaunchKernelGGL(bwd_reduce_kernel, dim3(3 + 16 * nr), dim3(256), 0, st, partial, nb, grad_params);
    if ((rc = check_launch(h, "backward reduce"))) return rc;
    rc = record_event(h, 1);
    if (rc) return rc;
    NLPF_HIP(h, hipStreamSynchronize(st));
    h->stats.total_ms = event_ms(h, 0, 1);
    return NLPF_OK;
}
