/*
 * san_check.c — the oracle and the workload generator under AddressSanitizer
 * and UndefinedBehaviorSanitizer (TEST INFRASTRUCTURE ONLY; `make -C oracle
 * san`).
 *
 * Runs the oracle over every configuration it accepts (metric x memory mode
 * x span x heuristic x compat x scope, penalties with and without a match
 * reward) on small synthetic pairs from csrc/synth.c plus empty and
 * one-sided pairs, and checks each completed CIGAR with the reference's
 * check_alignment restatement. The sanitizers catch out-of-bounds reads and
 * writes, use after free, leaks and undefined arithmetic on the way.
 * Exit status 0 when every check passes.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "wfa_oracle.h"

typedef struct {
    uint64_t seed;
    int32_t len_mode;
    int32_t len_lo, len_hi;
    double div_lo, div_hi;
    double indel_mean;
    double run_frac;
    double run_mean;
} wfa_synth_cfg;

void wfa_synth_lengths(const wfa_synth_cfg* c, int64_t n, int32_t* plen, int32_t* tlen,
                       int32_t nthreads);
void wfa_synth_fill(const wfa_synth_cfg* c, int64_t n, const int32_t* plen, const int32_t* tlen,
                    int32_t fmt, uint8_t* out, const int64_t* poff, const int64_t* toff,
                    int32_t nthreads);

#define NPAIRS 12

static void set_pen(wfa_attr_t* a, int metric, int reward) {
    memset(a, 0, sizeof(*a));
    a->penalties.distance_metric = metric;
    switch (metric) {
    case WFA_METRIC_GAP_LINEAR:
        a->penalties.match = reward ? -1 : 0;
        a->penalties.mismatch = 4;
        a->penalties.gap_opening1 = 2;
        break;
    case WFA_METRIC_GAP_AFFINE:
        a->penalties.match = reward ? -3 : 0;
        a->penalties.mismatch = reward ? 5 : 4;
        a->penalties.gap_opening1 = reward ? 3 : 6;
        a->penalties.gap_extension1 = 2;
        break;
    case WFA_METRIC_GAP_AFFINE_2P:
        a->penalties.match = reward ? -1 : 0;
        a->penalties.mismatch = 4;
        a->penalties.gap_opening1 = 6;
        a->penalties.gap_extension1 = 2;
        a->penalties.gap_opening2 = 24;
        a->penalties.gap_extension2 = 1;
        break;
    default:
        break;
    }
    a->max_alignment_steps = INT32_MAX;
}

int main(void) {
    /* pairs: synthetic (1-250 bp, 0-30% divergence, indel runs), then
     * empty and one-sided ones */
    wfa_synth_cfg c = {7, 1, 1, 250, 0.0, 0.30, 2.0, 0.2, 8.0};
    int32_t plen[NPAIRS + 4], tlen[NPAIRS + 4];
    int64_t poff[NPAIRS + 4], toff[NPAIRS + 4];
    wfa_synth_lengths(&c, NPAIRS, plen, tlen, 2);
    int64_t total = 0;
    for (int i = 0; i < NPAIRS; ++i) {
        poff[i] = total;
        total += plen[i];
        toff[i] = total;
        total += tlen[i];
    }
    uint8_t* seq = (uint8_t*)malloc((size_t)total + 1);
    wfa_synth_fill(&c, NPAIRS, plen, tlen, 0 /* ASCII */, seq, poff, toff, 2);
    const uint8_t* P[NPAIRS + 4];
    const uint8_t* T[NPAIRS + 4];
    for (int i = 0; i < NPAIRS; ++i) {
        P[i] = seq + poff[i];
        T[i] = seq + toff[i];
    }
    static const uint8_t acgt[] = "ACGTTGCA";
    const int extra[4][2] = {{0, 0}, {0, 5}, {7, 0}, {1, 1}};
    for (int j = 0; j < 4; ++j) {
        P[NPAIRS + j] = acgt;
        T[NPAIRS + j] = acgt + 1;
        plen[NPAIRS + j] = extra[j][0];
        tlen[NPAIRS + j] = extra[j][1];
    }
    const int npairs = NPAIRS + 4;

    static const int metrics[] = {WFA_METRIC_INDEL, WFA_METRIC_EDIT, WFA_METRIC_GAP_LINEAR,
                                  WFA_METRIC_GAP_AFFINE, WFA_METRIC_GAP_AFFINE_2P};
    static const int heur[][4] = {{WFA_HEURISTIC_NONE, 0, 0, 0},
                                  {WFA_HEURISTIC_WFADAPTIVE, 10, 50, 1},
                                  {WFA_HEURISTIC_XDROP, 20, 1, 0},
                                  {WFA_HEURISTIC_ZDROP, 20, 1, 0},
                                  {WFA_HEURISTIC_BANDED_STATIC, -20, 20, 0},
                                  {WFA_HEURISTIC_BANDED_ADAPTIVE, -10, 10, 1}};
    long aligned = 0, checked = 0, refused = 0, failed = 0;
    for (unsigned m = 0; m < sizeof(metrics) / sizeof(metrics[0]); ++m)
        for (int reward = 0; reward < 2; ++reward) {
            if (reward && metrics[m] <= WFA_METRIC_EDIT) continue;
            for (int mem = WFA_MEMORY_HIGH; mem <= WFA_MEMORY_ULTRALOW; ++mem)
                for (int span = 0; span < 3; ++span)
                    for (unsigned h = 0; h < sizeof(heur) / sizeof(heur[0]); ++h)
                        for (int compat = 0; compat < 2; ++compat)
                            for (int scope = 0; scope < 2; ++scope) {
                                wfa_attr_t a;
                                set_pen(&a, metrics[m], reward);
                                a.memory_mode = mem;
                                a.alignment_scope = scope ? WFA_SCOPE_ALIGNMENT : WFA_SCOPE_SCORE;
                                a.compat = compat;
                                a.heuristic = heur[h][0];
                                a.heuristic_p[0] = heur[h][1];
                                a.heuristic_p[1] = heur[h][2];
                                a.heuristic_p[2] = heur[h][3];
                                if (span > 0) {
                                    a.span = WFA_SPAN_ENDSFREE;
                                    a.extension = span == 2;
                                    a.pattern_begin_free = a.pattern_end_free = 3;
                                    a.text_begin_free = a.text_end_free = 3;
                                }
                                int32_t err = 0;
                                wfo_aligner* al = wfo_aligner_new(&a, &err);
                                if (!al) {
                                    refused++;
                                    continue;
                                }
                                for (int i = 0; i < npairs; ++i) {
                                    int32_t s;
                                    if (mem == WFA_MEMORY_ULTRALOW)
                                        s = scope ? wfo_align_biwfa(al, P[i], plen[i], T[i], tlen[i])
                                                  : wfo_align_biwfa_score(al, P[i], plen[i], T[i],
                                                                          tlen[i]);
                                    else if (span > 0)
                                        s = wfo_align_endsfree(al, P[i], plen[i], T[i], tlen[i]);
                                    else
                                        s = wfo_align_end2end(al, P[i], plen[i], T[i], tlen[i]);
                                    (void)s;
                                    aligned++;
                                    if (!scope || wfo_status(al) != WFA_STATUS_COMPLETED ||
                                        span > 0)
                                        continue;
                                    const uint8_t* ops = NULL;
                                    int32_t n = wfo_cigar(al, &ops);
                                    checked++;
                                    if (wfo_cigar_check(ops, n, P[i], plen[i], T[i], tlen[i]) != 0) {
                                        failed++;
                                        fprintf(stderr,
                                                "invalid CIGAR: metric %d reward %d mem %d heur %d "
                                                "compat %d pair %d\n",
                                                metrics[m], reward, mem, heur[h][0], compat, i);
                                    }
                                }
                                wfo_aligner_free(al);
                            }
        }
    free(seq);
    printf("san_check: %ld alignments, %ld CIGARs checked, %ld configurations refused, "
           "%ld failures\n",
           aligned, checked, refused, failed);
    return failed != 0;
}
