/*
 * dp_check.c — exact dynamic-programming scores (TEST INFRASTRUCTURE ONLY).
 *
 * Independent of the wavefront recurrences: textbook Gotoh DP for
 * gap-affine (gap of length l costs o + l*e), its 2-piece variant (a gap
 * costs min over the two pieces, each gap one piece), Levenshtein edit
 * distance and indel (LCS) distance. With match = 0 the WFA score is the
 * optimum of exactly these costs, so these pin the oracle's and the
 * engine's scores on random pairs (the reference's own DP oracle,
 * tests/corrected_endsfree.rs:56-107, is the same Gotoh recurrence). The
 * ends-free variant restates that model (free starts and ends, a match
 * reward allowed).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define BIG (INT32_MAX / 4)

static inline int32_t mn(int32_t a, int32_t b) { return a < b ? a : b; }

int32_t dp_affine_score(const uint8_t* P, int32_t plen, const uint8_t* T, int32_t tlen, int32_t x,
                        int32_t o1, int32_t e1, int32_t o2, int32_t e2, int32_t two_piece) {
    const int W = tlen + 1;
    int32_t* M = (int32_t*)malloc(sizeof(int32_t) * 2 * W);
    int32_t* I1 = (int32_t*)malloc(sizeof(int32_t) * 2 * W);
    int32_t* I2 = (int32_t*)malloc(sizeof(int32_t) * 2 * W);
    int32_t* D1 = (int32_t*)malloc(sizeof(int32_t) * 2 * W);
    int32_t* D2 = (int32_t*)malloc(sizeof(int32_t) * 2 * W);
    /* row 0 */
    for (int j = 0; j <= tlen; ++j) {
        D1[j] = D2[j] = BIG;
        if (j == 0) {
            M[0] = 0; I1[0] = I2[0] = BIG;
        } else {
            I1[j] = o1 + j * e1;
            I2[j] = two_piece ? o2 + j * e2 : BIG;
            M[j] = mn(I1[j], I2[j]);
        }
    }
    for (int i = 1; i <= plen; ++i) {
        int c = i & 1, p = c ^ 1;
        int32_t *Mc = M + c * W, *Mp = M + p * W;
        int32_t *I1c = I1 + c * W, *I2c = I2 + c * W;
        int32_t *D1c = D1 + c * W, *D1p = D1 + p * W, *D2c = D2 + c * W, *D2p = D2 + p * W;
        D1c[0] = mn(Mp[0] + o1 + e1, D1p[0] + e1);
        D2c[0] = two_piece ? mn(Mp[0] + o2 + e2, D2p[0] + e2) : BIG;
        I1c[0] = I2c[0] = BIG;
        Mc[0] = mn(D1c[0], D2c[0]);
        for (int j = 1; j <= tlen; ++j) {
            D1c[j] = mn(Mp[j] + o1 + e1, D1p[j] + e1);
            I1c[j] = mn(Mc[j - 1] + o1 + e1, I1c[j - 1] + e1);
            if (two_piece) {
                D2c[j] = mn(Mp[j] + o2 + e2, D2p[j] + e2);
                I2c[j] = mn(Mc[j - 1] + o2 + e2, I2c[j - 1] + e2);
            } else {
                D2c[j] = I2c[j] = BIG;
            }
            int32_t diag = Mp[j - 1] + (P[i - 1] == T[j - 1] ? 0 : x);
            Mc[j] = mn(diag, mn(mn(I1c[j], I2c[j]), mn(D1c[j], D2c[j])));
        }
    }
    int32_t r = M[(plen & 1) * W + tlen];
    free(M); free(I1); free(I2); free(D1); free(D2);
    return r;
}

/* Optimal ends-free penalty (lower is better; matches cost `match`, <= 0)
 * under the wavefront model: an alignment starts at (v, 0) with
 * v <= pb or at (0, h) with h <= tb, and ends at (v, |T|) with |P| - v <= pe
 * or at (|P|, h) with |T| - h <= te; a gap of length l costs
 * min(o1 + l*e1, o2 + l*e2) (one piece when two_piece is 0). Full matrices:
 * for short sequences only. */
int64_t dp_endsfree_penalty(const uint8_t* P, int32_t n, const uint8_t* T, int32_t m,
                            int32_t match, int32_t x, int32_t o1, int32_t e1, int32_t o2,
                            int32_t e2, int32_t two_piece, int32_t pb, int32_t pe, int32_t tb,
                            int32_t te) {
    const int64_t INF = INT64_MAX / 4;
    const int W = m + 1, np = two_piece ? 2 : 1;
    const int64_t go[2] = {o1, o2}, ge[2] = {e1, e2};
    int64_t* H = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1) * W);
    int64_t* I = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1) * W * 2);
    int64_t* D = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1) * W * 2);
    for (int v = 0; v <= n; ++v) {
        for (int j = 0; j <= m; ++j) {
            int64_t best = INF;
            if ((j == 0 && v <= pb) || (v == 0 && j <= tb)) best = 0;
            if (v > 0 && j > 0 && H[(v - 1) * W + j - 1] < INF) {
                int64_t c = H[(v - 1) * W + j - 1] + (P[v - 1] == T[j - 1] ? match : x);
                if (c < best) best = c;
            }
            for (int q = 0; q < 2; ++q) {
                int64_t* iq = &I[((size_t)v * W + j) * 2 + q];
                int64_t* dq = &D[((size_t)v * W + j) * 2 + q];
                *iq = INF;
                *dq = INF;
                if (q >= np) continue;
                if (j > 0) {
                    int64_t a = H[v * W + j - 1] + go[q] + ge[q];
                    int64_t b = I[((size_t)v * W + j - 1) * 2 + q] + ge[q];
                    *iq = a < b ? a : b;
                    if (*iq < best) best = *iq;
                }
                if (v > 0) {
                    int64_t a = H[(v - 1) * W + j] + go[q] + ge[q];
                    int64_t b = D[((size_t)(v - 1) * W + j) * 2 + q] + ge[q];
                    *dq = a < b ? a : b;
                    if (*dq < best) best = *dq;
                }
            }
            H[v * W + j] = best;
        }
    }
    int64_t best = INF;
    for (int v = 0; v <= n; ++v)
        if (n - v <= pe && H[v * W + m] < best) best = H[v * W + m];
    for (int j = 0; j <= m; ++j)
        if (m - j <= te && H[n * W + j] < best) best = H[n * W + j];
    free(H); free(I); free(D);
    return best;
}

int32_t dp_edit_score(const uint8_t* P, int32_t plen, const uint8_t* T, int32_t tlen,
                      int32_t indel_only) {
    int32_t* prev = (int32_t*)malloc(sizeof(int32_t) * (size_t)(tlen + 1));
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * (size_t)(tlen + 1));
    for (int j = 0; j <= tlen; ++j) prev[j] = j;
    for (int i = 1; i <= plen; ++i) {
        cur[0] = i;
        for (int j = 1; j <= tlen; ++j) {
            int32_t sub = prev[j - 1] + (P[i - 1] == T[j - 1] ? 0 : (indel_only ? BIG : 1));
            cur[j] = mn(sub, mn(prev[j] + 1, cur[j - 1] + 1));
        }
        int32_t* t = prev; prev = cur; cur = t;
    }
    int32_t r = prev[tlen];
    free(prev); free(cur);
    return r;
}
