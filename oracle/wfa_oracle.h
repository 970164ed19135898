/*
 * wfa_oracle.h — CPU restatement of the reference aligner (TEST
 * INFRASTRUCTURE ONLY).
 *
 * This is the checker the GPU engine is compared against, and the CPU
 * baseline bench.py times. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the engine never links it.
 *
 * It restates, in plain C, the semantics of HPCBio/wfa2lib-rs for the
 * paths it covers, following the reference file:line cited at each
 * function in wfa_oracle.c.
 *
 * Parity status per function (how each is pinned):
 *   - gap-affine / gap-affine-2p / edit / indel end-to-end scores:
 *     the reference's literal vectors (tests/cli_output.rs,
 *     src/aligner.rs unit tests) + an exact Gotoh DP (oracle/dp_check.c)
 *     on random pairs.
 *   - Singletrack CIGARs (high memory, affine/affine2p): the reference's
 *     literal CIGARs (tests/cli_output.rs:77,216-230) + CIGAR validity +
 *     rescoring == score. Co-optimal tie-breaks beyond those vectors are
 *     pinned only by the literal vectors ("parity partially pinned").
 */
#ifndef WFA_ORACLE_H
#define WFA_ORACLE_H

#include <stdint.h>
#include "../include/wfa_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wfo_aligner wfo_aligner;

wfo_aligner* wfo_aligner_new(const wfa_attr_t* attr, int32_t* err);
void wfo_aligner_free(wfo_aligner* a);
int32_t wfo_align_end2end(wfo_aligner* a, const uint8_t* pattern, int32_t plen,
                          const uint8_t* text, int32_t tlen);
/* align_endsfree (src/aligner.rs:707-831): the attribute's free ends, or
 * extension when attr->extension is set */
int32_t wfo_align_endsfree(wfo_aligner* a, const uint8_t* pattern, int32_t plen,
                           const uint8_t* text, int32_t tlen);
/* BiWFA (src/aligner.rs:2563-2690): CIGAR (align_biwfa) and score only
 * (align_biwfa_score); memory_mode ULTRALOW selects them in
 * wfo_align_batch */
int32_t wfo_align_biwfa(wfo_aligner* a, const uint8_t* pattern, int32_t plen,
                        const uint8_t* text, int32_t tlen);
int32_t wfo_align_biwfa_score(wfo_aligner* a, const uint8_t* pattern, int32_t plen,
                              const uint8_t* text, int32_t tlen);
int32_t wfo_status(const wfo_aligner* a);
/* alignment_end_pos (src/aligner.rs:146-204) after the last alignment:
 * {score, k, offset}; a Z-drop end is the best cell seen */
void wfo_end_position(const wfo_aligner* a, int32_t* out);
int32_t wfo_cigar(const wfo_aligner* a, const uint8_t** ops);
int32_t wfo_cigar_score(const wfo_aligner* a);

/* Scores the reference's Cigar computes (src/cigar.rs:118-202), on the
 * original (user) penalties. */
int32_t wfo_cigar_score_edit(const uint8_t* ops, int32_t n);
int32_t wfo_cigar_score_gap_linear(const uint8_t* ops, int32_t n, int32_t match,
                                   int32_t mismatch, int32_t indel);
int32_t wfo_cigar_score_gap_affine(const uint8_t* ops, int32_t n, int32_t match,
                                   int32_t mismatch, int32_t o, int32_t e);
int32_t wfo_cigar_score_gap_affine2p(const uint8_t* ops, int32_t n, int32_t match,
                                     int32_t mismatch, int32_t o1, int32_t e1,
                                     int32_t o2, int32_t e2);
/* Cigar::check_alignment (src/cigar.rs:446-518): 0 if valid. */
int32_t wfo_cigar_check(const uint8_t* ops, int32_t n, const uint8_t* pattern,
                        int32_t plen, const uint8_t* text, int32_t tlen);

/* Many pairs at once (ASCII or 2-bit packed input, same layout as
 * wfa_batch_t but in host memory), sharded contiguously over `nthreads`
 * threads, one aligner per thread. `ops_off[i]` reserves plen+tlen bytes.
 * Returns 0, or a negative WFA_ERR_*. */
int32_t wfo_align_batch(const wfa_attr_t* attr, int32_t n, int32_t seq_format,
                        const uint8_t* seq, const int64_t* poff, const int32_t* plen,
                        const int64_t* toff, const int32_t* tlen,
                        int32_t* score, int32_t* status,
                        uint8_t* ops, const int64_t* ops_off, int32_t* ops_len,
                        int32_t nthreads);


/* Test hooks: the step primitives on caller arrays (tests/ only), so the
 * reference's unit literals can be checked against them. A row is n cells
 * for the diagonals [base, base + n - 1]; lohi = {lo, hi} in/out; st = the
 * six HeuristicState fields {steps_wait, max_sw_score, max_sw_k,
 * max_sw_off, max_wf_score, zlast}; drop = {match, endsfree,
 * pattern_begin_free, text_begin_free, scope}. */
int32_t wfo_t_wf_distance(int32_t off, int32_t k, int32_t plen, int32_t tlen);
void wfo_t_wfadaptive_reduce(int32_t* lohi, const int32_t* d, int32_t dbase, int32_t min_d,
                             int32_t thr, int32_t min_k, int32_t max_k);
void wfo_t_wfadaptive_cutoff(int32_t* st, int32_t* mem, int32_t base, int32_t n, int32_t* lohi,
                             int32_t* d, int32_t plen, int32_t tlen, int32_t min_len,
                             int32_t thr, int32_t steps);
void wfo_t_banded_static(int32_t* lohi, int32_t min_k, int32_t max_k);
void wfo_t_banded_adaptive(int32_t* st, int32_t* mem, int32_t base, int32_t n, int32_t* lohi,
                           int32_t plen, int32_t tlen, int32_t min_k, int32_t max_k,
                           int32_t steps);
void wfo_t_hstate_reset(int32_t* st, int32_t heuristic, const int32_t* hp);
void wfo_t_compute_sw_scores(int32_t* mem, int32_t base, int32_t n, int32_t lo, int32_t hi,
                             int32_t* sw, const int32_t* drop, int32_t score, int32_t* out);
void wfo_t_xdrop_cutoff(int32_t* st, int32_t* mem, int32_t base, int32_t n, int32_t* lohi,
                        int32_t* sw, const int32_t* drop, int32_t score, int32_t xdrop,
                        int32_t steps);
int32_t wfo_t_zdrop_cutoff(int32_t* st, int32_t* mem, int32_t base, int32_t n, int32_t* lohi,
                           int32_t* sw, const int32_t* drop, int32_t score, int32_t zdrop,
                           int32_t steps);
void wfo_t_heuristic_equate(int32_t* dst, const int32_t* src);
void wfo_t_affine_cells(int32_t plen, int32_t tlen, int32_t* const* rows, int32_t base,
                        int32_t n, int32_t lo, int32_t hi, int32_t use2);
void wfo_t_edit_cells(int32_t plen, int32_t tlen, int32_t* prev, int32_t* cur, int32_t base,
                      int32_t n, int32_t lo, int32_t hi, int32_t indel);
void wfo_t_breakpoint_scan(int32_t* m0, int32_t base0, int32_t n0, int32_t lo0, int32_t hi0,
                           int32_t* m1, int32_t base1, int32_t n1, int32_t lo1, int32_t hi1,
                           int32_t plen, int32_t tlen, int32_t s0, int32_t s1, int32_t gap_open,
                           int32_t comp, int32_t fwd, int32_t* bp);

#ifdef __cplusplus
}
#endif
#endif
