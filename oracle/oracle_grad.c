/*
 * TEST INFRASTRUCTURE ONLY — FP64 reference gradient of the screen-mode NL-PF (CPU).  Loaded
 * only by tests/ as the checker of the library's adjoint kernels.
 *
 * Same forward definition as oracle_screen_core.inc.  The reverse pass differs in construction
 * from the library on purpose: weights are exact exp() of the physical distances, the
 * stylization Jacobian and the ramp-stop derivatives come from FP64 central differences of
 * or_apply_style over the whole chain, and sums are plain FP64.
 *   dL/dsigma_k = sum_i g_i^T J_S(F_i) sum_j (dw_ij/dsigma_k)(in_j - F_i) / W_i,
 *   dw/dsigma = w * term / sigma^3,  term in {|dp|^2, (1-n.n)^2, dl^2};
 *   dL/din_j = sum_i w_ij J_S(F_i)^T g_i / W_i;  g_{d+1} = thr * dL/din.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static void style_at(const or_node* nodes, int root, uint32_t d, const double x[3], double y[3]) {
    if (root < 0) {
        memcpy(y, x, 3 * sizeof(double));
        return;
    }
    or_apply_style(nodes, root, d, 0u, x, NULL, y);
}

int or_screen_grad_f64(uint32_t W, uint32_t H, uint32_t S, uint32_t L, int radius, const double sig[3],
                       const float* rec, const or_node* nodes_in, uint32_t n_nodes, int root, const double* dfilm,
                       double* grad, double* loss_unused) {
    (void)loss_unused;
    const uint64_t plane = (uint64_t)W * H;
    double* lv = (double*)malloc(sizeof(double) * S * L * 3 * plane);
    double* film = (double*)malloc(sizeof(double) * plane * 3);
    or_node* nodes = (or_node*)malloc(sizeof(or_node) * (n_nodes ? n_nodes : 1));
    int* ramp_of = (int*)malloc(sizeof(int) * (n_nodes ? n_nodes : 1));
    if (!lv || !film || !nodes || !ramp_of) return -1;
    if (n_nodes) memcpy(nodes, nodes_in, sizeof(or_node) * n_nodes);
    int nr = 0;
    for (uint32_t i = 0; i < n_nodes; ++i) ramp_of[i] = nodes[i].type == OR_RAMP ? nr++ : -1;
    const int np = 3 + 16 * nr;
    for (int q = 0; q < np; ++q) grad[q] = 0;
    if (or_screen_run_f64(W, H, S, L, radius, sig[0], sig[1], sig[2], rec, nodes, root, film, lv, 0)) return -1;
    double* g = (double*)malloc(sizeof(double) * S * 3 * plane);
    double* gn = (double*)calloc(S * 3 * plane, sizeof(double));
    double* u = (double*)malloc(sizeof(double) * 3 * plane);
    double* inb = (double*)malloc(sizeof(double) * 3 * plane);
    double* lum = (double*)malloc(sizeof(double) * plane);
    if (!g || !gn || !u || !inb || !lum) return -1;
    for (uint32_t s = 0; s < S; ++s)
        for (uint64_t i = 0; i < plane; ++i)
            for (int c = 0; c < 3; ++c) g[(s * 3 + c) * plane + i] = dfilm[3 * i + c] / (double)S;
    const double a[3] = {1.0 / (2 * sig[0] * sig[0]), 1.0 / (2 * sig[1] * sig[1]), 1.0 / (2 * sig[2] * sig[2])};
    for (uint32_t d = 0; d < L; ++d) {
        memset(gn, 0, sizeof(double) * S * 3 * plane);
        for (uint32_t s = 0; s < S; ++s) {
            const float* r = rec + ((uint64_t)s * L + d) * 15 * plane;
            const double* ch = d + 1 < L ? lv + ((uint64_t)s * L + d + 1) * 3 * plane : NULL;
            for (uint64_t i = 0; i < plane; ++i) {
                for (int c = 0; c < 3; ++c)
                    inb[c * plane + i] = r[(9 + c) * plane + i] + (ch ? r[(12 + c) * plane + i] * ch[c * plane + i] : 0);
                lum[i] = 0.2126 * r[6 * plane + i] + 0.7152 * r[7 * plane + i] + 0.0722 * r[8 * plane + i];
            }
            for (int64_t y = 0; y < (int64_t)H; ++y)
                for (int64_t x = 0; x < (int64_t)W; ++x) {
                    const uint64_t i = (uint64_t)y * W + x;
                    double Wt = 0, Sv[3] = {0, 0, 0}, A[3] = {0, 0, 0}, B[3][3] = {{0}};
                    for (int64_t yy = y - radius; yy <= y + radius; ++yy)
                        for (int64_t xx = x - radius; xx <= x + radius; ++xx) {
                            if (yy < 0 || xx < 0 || yy >= (int64_t)H || xx >= (int64_t)W) continue;
                            const uint64_t j = (uint64_t)yy * W + xx;
                            double t[3];
                            double dp = 0, nn = 0;
                            for (int c = 0; c < 3; ++c) {
                                double dd = (double)r[c * plane + i] - r[c * plane + j];
                                dp += dd * dd;
                                nn += (double)r[(3 + c) * plane + i] * r[(3 + c) * plane + j];
                            }
                            t[0] = dp;
                            t[1] = (1 - nn) * (1 - nn);
                            t[2] = (lum[i] - lum[j]) * (lum[i] - lum[j]);
                            const double w = exp(-(a[0] * t[0] + a[1] * t[1] + a[2] * t[2]));
                            Wt += w;
                            for (int c = 0; c < 3; ++c) Sv[c] += w * inb[c * plane + j];
                            for (int k = 0; k < 3; ++k) {
                                const double dw = w * t[k] / (sig[k] * sig[k] * sig[k]);
                                A[k] += dw;
                                for (int c = 0; c < 3; ++c) B[k][c] += dw * inb[c * plane + j];
                            }
                        }
                    const double F[3] = {Sv[0] / Wt, Sv[1] / Wt, Sv[2] / Wt};
                    const double gi[3] = {g[(s * 3 + 0) * plane + i], g[(s * 3 + 1) * plane + i],
                                          g[(s * 3 + 2) * plane + i]};
                    /* h = J^T g by central differences of the whole chain */
                    double hvec[3];
                    for (int k = 0; k < 3; ++k) {
                        const double hstep = 1e-7 * (fabs(F[k]) > 1 ? fabs(F[k]) : 1.0);
                        double xp[3] = {F[0], F[1], F[2]}, xm[3] = {F[0], F[1], F[2]}, yp[3], ym[3];
                        xp[k] += hstep;
                        xm[k] -= hstep;
                        style_at(nodes, root, d, xp, yp);
                        style_at(nodes, root, d, xm, ym);
                        hvec[k] = 0;
                        for (int c = 0; c < 3; ++c) hvec[k] += gi[c] * (yp[c] - ym[c]) / (2 * hstep);
                    }
                    /* ramp stop values */
                    for (uint32_t nd = 0; nd < n_nodes; ++nd) {
                        if (ramp_of[nd] < 0) continue;
                        for (uint32_t m = 0; m < nodes[nd].count && m < 16; ++m) {
                            const float v0 = nodes[nd].p[16 + m];
                            const double hs = 1e-3;
                            double yp[3], ym[3];
                            nodes[nd].p[16 + m] = (float)(v0 + hs);
                            const double vp = nodes[nd].p[16 + m];
                            style_at(nodes, root, d, F, yp);
                            nodes[nd].p[16 + m] = (float)(v0 - hs);
                            const double vm = nodes[nd].p[16 + m];
                            style_at(nodes, root, d, F, ym);
                            nodes[nd].p[16 + m] = v0;
                            double acc = 0;
                            for (int c = 0; c < 3; ++c) acc += gi[c] * (yp[c] - ym[c]) / (vp - vm);
                            grad[3 + 16 * ramp_of[nd] + m] += acc;
                        }
                    }
                    for (int k = 0; k < 3; ++k) {
                        double acc = 0;
                        for (int c = 0; c < 3; ++c) acc += hvec[c] * (B[k][c] - F[c] * A[k]) / Wt;
                        grad[k] += acc;
                    }
                    for (int c = 0; c < 3; ++c) u[c * plane + i] = hvec[c] / Wt;
                }
            if (d + 1 < L) {
                for (int64_t y = 0; y < (int64_t)H; ++y)
                    for (int64_t x = 0; x < (int64_t)W; ++x) {
                        const uint64_t j = (uint64_t)y * W + x;
                        double G[3] = {0, 0, 0};
                        for (int64_t yy = y - radius; yy <= y + radius; ++yy)
                            for (int64_t xx = x - radius; xx <= x + radius; ++xx) {
                                if (yy < 0 || xx < 0 || yy >= (int64_t)H || xx >= (int64_t)W) continue;
                                const uint64_t i = (uint64_t)yy * W + xx;
                                double dp = 0, nn = 0;
                                for (int c = 0; c < 3; ++c) {
                                    double dd = (double)r[c * plane + i] - r[c * plane + j];
                                    dp += dd * dd;
                                    nn += (double)r[(3 + c) * plane + i] * r[(3 + c) * plane + j];
                                }
                                const double dl = lum[i] - lum[j];
                                const double w = exp(-(a[0] * dp + a[1] * (1 - nn) * (1 - nn) + a[2] * dl * dl));
                                for (int c = 0; c < 3; ++c) G[c] += w * u[c * plane + i];
                            }
                        for (int c = 0; c < 3; ++c) gn[(s * 3 + c) * plane + j] = r[(12 + c) * plane + j] * G[c];
                    }
            }
        }
        double* t = g;
        g = gn;
        gn = t;
    }
    free(lv);
    free(film);
    free(nodes);
    free(ramp_of);
    free(g);
    free(gn);
    free(u);
    free(inb);
    free(lum);
    return np;
}
