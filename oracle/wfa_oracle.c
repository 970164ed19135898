/*
 * wfa_oracle.c — CPU restatement of HPCBio/wfa2lib-rs's wavefront aligner.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine and the CPU
 * baseline of bench.py. Never linked into the engine.
 *
 * The structure follows the reference closely on purpose, because the
 * point is to reproduce its results bit for bit, including C-compatible
 * edge behaviour: wavefronts are objects with a diagonal range [lo, hi], an
 * "initialized" range [init_min, init_max] and an allocation; components
 * are score-indexed (high memory) or modular (score-only); each step does
 * the input lookup, range union, NULL-halo fill, kernel and independent
 * trims exactly as src/aligner.rs does. Citations are to files of the
 * reference (paths relative to its root).
 *
 * Deliberate, documented difference: sequence ends are hard bounds in the
 * extension. The reference instead relies on the sentinels '!' after the
 * pattern and '?' after the text (src/sequences.rs:16-19,178-191), which
 * gives the same result for every input whose text holds no '!' and whose
 * pattern holds no '?'.
 */
#include "wfa_oracle.h"

#include <limits.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OFFSET_NULL (INT32_MIN / 2) /* src/offset.rs:15 */
#define DIAG_NULL INT32_MAX         /* src/offset.rs:43 */

/* ------------------------------------------------------------------ */
/* Penalties: WavefrontPenalties::new_* (src/penalties.rs:122-296)      */
/* ------------------------------------------------------------------ */
typedef struct {
    int metric;
    int match, mismatch, o1, e1, o2, e2; /* internal (Eizenga-adjusted) */
    wfa_penalties_t user;
} pen_t;

static int pen_init(pen_t* p, const wfa_penalties_t* u) {
    memset(p, 0, sizeof(*p));
    p->user = *u;
    p->metric = u->distance_metric;
    switch (u->distance_metric) {
    case WFA_METRIC_INDEL: /* src/penalties.rs:122-134 */
        p->match = 0; p->mismatch = -1; p->o1 = 1; p->e1 = -1; p->o2 = -1; p->e2 = -1;
        return 0;
    case WFA_METRIC_EDIT: /* src/penalties.rs:136-150 */
        p->match = 0; p->mismatch = 1; p->o1 = 1; p->e1 = -1; p->o2 = -1; p->e2 = -1;
        return 0;
    case WFA_METRIC_GAP_LINEAR: /* src/penalties.rs:156-190 */
        if (u->match > 0 || u->mismatch <= 0 || u->gap_opening1 <= 0) return -1;
        if (u->match < 0) {
            p->match = u->match;
            p->mismatch = 2 * u->mismatch - 2 * u->match;
            p->o1 = 2 * u->gap_opening1 - u->match;
        } else {
            p->match = 0; p->mismatch = u->mismatch; p->o1 = u->gap_opening1;
        }
        p->e1 = -1; p->o2 = -1; p->e2 = -1;
        return 0;
    case WFA_METRIC_GAP_AFFINE: /* src/penalties.rs:196-238 */
        if (u->match > 0 || u->mismatch <= 0 || u->gap_opening1 < 0 || u->gap_extension1 <= 0)
            return -1;
        if (u->match < 0) {
            p->match = u->match;
            p->mismatch = 2 * u->mismatch - 2 * u->match;
            p->o1 = 2 * u->gap_opening1;
            p->e1 = 2 * u->gap_extension1 - u->match;
        } else {
            p->match = 0; p->mismatch = u->mismatch;
            p->o1 = u->gap_opening1; p->e1 = u->gap_extension1;
        }
        p->o2 = -1; p->e2 = -1;
        return 0;
    case WFA_METRIC_GAP_AFFINE_2P: /* src/penalties.rs:243-296 */
        if (u->match > 0 || u->mismatch <= 0 || u->gap_opening1 < 0 || u->gap_extension1 <= 0 ||
            u->gap_opening2 < 0 || u->gap_extension2 <= 0)
            return -1;
        if (u->match < 0) {
            p->match = u->match;
            p->mismatch = 2 * u->mismatch - 2 * u->match;
            p->o1 = 2 * u->gap_opening1;
            p->e1 = 2 * u->gap_extension1 - u->match;
            p->o2 = 2 * u->gap_opening2;
            p->e2 = 2 * u->gap_extension2 - u->match;
        } else {
            p->match = 0; p->mismatch = u->mismatch;
            p->o1 = u->gap_opening1; p->e1 = u->gap_extension1;
            p->o2 = u->gap_opening2; p->e2 = u->gap_extension2;
        }
        return 0;
    default:
        return -1;
    }
}

/* ------------------------------------------------------------------ */
/* Wavefront (src/wavefront.rs:43-490)                                 */
/* ------------------------------------------------------------------ */
typedef struct wf {
    int32_t* mem; /* cells [base, base + nalloc - 1] */
    uint32_t* pc; /* piggyback: packed CIGAR per cell (med/low memory) */
    uint32_t* pv; /* piggyback: previous backtrace block per cell */
    int occ;      /* bt_occupancy_max */
    int base, nalloc;
    int lo, hi;
    int init_min, init_max;
    int null;
} wf_t;

static wf_t* wf_new(int alloc_lo, int alloc_hi, int bt) {
    wf_t* w = (wf_t*)malloc(sizeof(wf_t));
    w->base = alloc_lo;
    w->nalloc = alloc_hi - alloc_lo + 1;
    w->mem = (int32_t*)malloc(sizeof(int32_t) * (size_t)w->nalloc);
    w->pc = bt ? (uint32_t*)calloc((size_t)w->nalloc, sizeof(uint32_t)) : NULL;
    w->pv = bt ? (uint32_t*)calloc((size_t)w->nalloc, sizeof(uint32_t)) : NULL;
    w->occ = 0;
    /* Wavefront::init (src/wavefront.rs:307-328) */
    w->null = 0;
    w->lo = 1; w->hi = -1;
    w->init_min = 0; w->init_max = 0;
    return w;
}
static void wf_free(wf_t* w) {
    if (w) { free(w->mem); free(w->pc); free(w->pv); free(w); }
}
static inline int wf_allocated(const wf_t* w, int k) {
    return k >= w->base && k < w->base + w->nalloc;
}
static inline int32_t wf_get(const wf_t* w, int k) {
    if (!wf_allocated(w, k)) {
        fprintf(stderr, "oracle: read outside allocation k=%d [%d,%d]\n", k, w->base,
                w->base + w->nalloc - 1);
        abort();
    }
    return w->mem[k - w->base];
}
static inline void wf_set(wf_t* w, int k, int32_t v) {
    if (!wf_allocated(w, k)) {
        fprintf(stderr, "oracle: write outside allocation k=%d\n", k);
        abort();
    }
    w->mem[k - w->base] = v;
}
/* zero_bt (src/wavefront.rs:570-580): halo cells carry an empty chain */
static void wf_zero_bt(wf_t* w, int k) {
    if (w->pc) { w->pc[k - w->base] = 0; w->pv[k - w->base] = 0; }
}
/* set_limits (src/wavefront.rs:360-368) */
static void wf_set_limits(wf_t* w, int lo, int hi) {
    w->lo = lo; w->hi = hi; w->init_min = lo; w->init_max = hi;
}
/* init_ends_lower (src/wavefront.rs:422-437) */
static void wf_init_ends_lower(wf_t* w, int min_lo) {
    if (w->init_min <= min_lo) return;
    int min_init = w->init_min < w->lo ? w->init_min : w->lo;
    int start = min_lo > w->base ? min_lo : w->base;
    for (int k = start; k < min_init; ++k) { wf_set(w, k, OFFSET_NULL); wf_zero_bt(w, k); }
    w->init_min = start;
}
/* init_ends_higher (src/wavefront.rs:442-457) */
static void wf_init_ends_higher(wf_t* w, int max_hi) {
    if (w->init_max >= max_hi) return;
    int max_init = w->init_max > w->hi ? w->init_max : w->hi;
    int alloc_max = w->base + w->nalloc - 1;
    int end = max_hi < alloc_max ? max_hi : alloc_max;
    for (int k = max_init + 1; k <= end; ++k) { wf_set(w, k, OFFSET_NULL); wf_zero_bt(w, k); }
    w->init_max = end;
}
/* A read of an input wavefront inside a kernel: the null wavefront reads
 * OFFSET_NULL everywhere (src/components.rs:109-112). Reads are always in
 * the initialized range after the init_ends calls; anything else would be
 * a restatement error, so it aborts. */
static inline int32_t rd(const wf_t* w, int k) {
    if (!w) return OFFSET_NULL;
    if (k < w->init_min || k > w->init_max) {
        fprintf(stderr, "oracle: read outside init range k=%d [%d,%d]\n", k, w->init_min,
                w->init_max);
        abort();
    }
    return wf_get(w, k);
}

/* trim_ends (src/compute/mod.rs:14-49) */
static void trim_ends(int plen, int tlen, wf_t* w) {
    int k = w->hi;
    while (k >= w->lo) {
        int32_t off = wf_get(w, k);
        if ((uint32_t)off <= (uint32_t)tlen && (uint32_t)(off - k) <= (uint32_t)plen) break;
        --k;
    }
    w->hi = k; w->init_max = k;
    k = w->lo;
    while (k <= w->hi) {
        int32_t off = wf_get(w, k);
        if ((uint32_t)off <= (uint32_t)tlen && (uint32_t)(off - k) <= (uint32_t)plen) break;
        ++k;
    }
    w->lo = k; w->init_min = k;
    w->null = w->lo > w->hi;
}

/* ------------------------------------------------------------------ */
/* Aligner state (src/aligner.rs:146-204)                              */
/* ------------------------------------------------------------------ */
enum { C_M = 0, C_I1, C_D1, C_I2, C_D2, NCOMP };

/* HeuristicState (src/heuristic.rs:60-143): the fields in this order, six
 * int32 (the test hooks below take them as an array) */
typedef struct {
    int32_t steps_wait, max_sw_score, max_sw_k, max_sw_off, max_wf_score, zlast;
} hstate_t;

struct wfo_aligner {
    wfa_attr_t attr;
    pen_t pen;
    /* sequences (unpadded copies) */
    uint8_t* pattern; int plen;
    uint8_t* text; int tlen;
    size_t seq_cap;
    /* components */
    int modular;
    int scope;          /* max_score_scope */
    int num_wavefronts; /* flat-array capacity in high memory */
    int nslots;
    wf_t** comp[NCOMP];
    int hist_lo, hist_hi;   /* historic_min_lo / historic_max_hi */
    int alloc_lo, alloc_hi; /* alloc_min_lo / alloc_max_hi */
    /* every wavefront allocated in this alignment (freed at the next one) */
    wf_t** pool; int npool, cap_pool;
    /* state */
    int num_null_steps;
    int status;
    int end_score, end_k, end_offset;
    /* cigar */
    uint8_t* cig; int cig_cap, cig_begin, cig_end, cig_score;
    /* piggyback backtrace (med/low memory): BacktraceBuffer
     * (src/bt_buffer.rs:34-209) as flat arrays */
    int piggyback, low;
    /* heuristic state (src/heuristic.rs:60-143) */
    hstate_t hs;
    int32_t* dist; int cap_dist;
    /* Compat::Corrected (src/aligner.rs:53-94, 1111-1360): the best ends-free
     * end candidate so far (EndsfreeCandidate) with its piggyback chain; the
     * begin-free dominance record (furthest offset per component and
     * diagonal); the Z-drop best cell's chain */
    int cand_valid, cand_total, cand_score, cand_k, cand_off;
    uint32_t cand_pc, cand_pv;
    int dom_on, dom_base, dom_len, cap_dom;
    int32_t* dom;
    uint32_t zbest_pc, zbest_pv;
    /* BiWFA sub-aligners (forward, reverse: score-only; base: CIGAR),
     * created on first use (src/aligner.rs:2721-2740) */
    struct wfo_aligner *bi_fwd, *bi_rev, *bi_base;
    /* span (src/aligner.rs:163-170): ends-free limits of this alignment */
    int endsfree, pb_free, pe_free, tb_free, te_free;
    uint32_t *blk_pc, *blk_prev; int nblk, cap_blk;
    int32_t *init_v, *init_h; int ninit, cap_init;
    uint32_t* packed; int cap_packed;
};

static void pool_add(wfo_aligner* a, wf_t* w) {
    if (a->npool == a->cap_pool) {
        a->cap_pool = a->cap_pool ? 2 * a->cap_pool : 1024;
        a->pool = (wf_t**)realloc(a->pool, sizeof(wf_t*) * (size_t)a->cap_pool);
    }
    a->pool[a->npool++] = w;
}
static void pool_clear(wfo_aligner* a) {
    for (int i = 0; i < a->npool; ++i) wf_free(a->pool[i]);
    a->npool = 0;
}

/* compute_dimensions (src/components.rs:435-504) */
static void compute_dimensions(const pen_t* p, int plen, int tlen, int modular, int* scope,
                               int* nwf) {
    int absdiff = plen > tlen ? plen - tlen : tlen - plen;
    int mn = plen < tlen ? plen : tlen;
    switch (p->metric) {
    case WFA_METRIC_INDEL:
    case WFA_METRIC_EDIT:
        *scope = 2;
        if (modular) *nwf = 2;
        else if (p->metric == WFA_METRIC_INDEL) *nwf = plen + tlen + 1;
        else *nwf = (plen > tlen ? plen : tlen) + 1;
        return;
    case WFA_METRIC_GAP_LINEAR:
        *scope = (p->mismatch > p->o1 ? p->mismatch : p->o1) + 1;
        *nwf = modular ? *scope : mn * p->mismatch + p->o1 * absdiff + 1;
        return;
    case WFA_METRIC_GAP_AFFINE: {
        int si = p->o1 + p->e1;
        *scope = (si > p->mismatch ? si : p->mismatch) + 1;
        *nwf = modular ? *scope : mn * p->mismatch + p->o1 + absdiff * p->e1 + 1;
        return;
    }
    default: {
        int si = p->o1 + p->e1, s2 = p->o2 + p->e2;
        if (s2 > si) si = s2;
        *scope = (si > p->mismatch ? si : p->mismatch) + 1;
        if (modular) {
            *nwf = *scope;
        } else {
            int i1 = p->o1 + absdiff * p->e1, i2 = p->o2 + absdiff * p->e2;
            *nwf = mn * p->mismatch + (i1 < i2 ? i1 : i2) + 1;
        }
        return;
    }
    }
}

static int ncomp_of(int metric) {
    return metric == WFA_METRIC_GAP_AFFINE ? 3 : metric == WFA_METRIC_GAP_AFFINE_2P ? 5 : 1;
}

static inline int slot_of(const wfo_aligner* a, int score) {
    return a->modular ? score % a->scope : score;
}
static inline wf_t* get(const wfo_aligner* a, int c, int score) {
    return a->comp[c][slot_of(a, score)];
}
static inline void put(wfo_aligner* a, int c, int score, wf_t* w) {
    a->comp[c][slot_of(a, score)] = w;
}
/* has_slot (src/components.rs:307-309) */
static inline int has_slot(const wfo_aligner* a, int score) {
    return a->modular || score < a->num_wavefronts;
}
/* is_null_wf / live_or_none / null_wf_bounds (src/aligner.rs:382-405) */
static inline int is_null_wf(const wf_t* w) { return !w || w->null; }
static inline wf_t* live_or_none(wf_t* w) { return is_null_wf(w) ? NULL : w; }
static inline int nlo(const wf_t* w) { return w ? w->lo : 1; }
static inline int nhi(const wf_t* w) { return w ? w->hi : -1; }

/* output_alloc_range (src/aligner.rs:3325-3342), end-to-end part. */
static void output_alloc_range(wfo_aligner* a, int lo, int hi, int* alo, int* ahi) {
    int pad = a->scope + 1;
    if (lo - pad < a->alloc_lo) a->alloc_lo = lo - pad;
    if (hi + pad > a->alloc_hi) a->alloc_hi = hi + pad;
    *alo = a->alloc_lo; *ahi = a->alloc_hi;
}

static wf_t* alloc_wf(wfo_aligner* a, int alo, int ahi) {
    wf_t* w = wf_new(alo, ahi, a->piggyback);
    pool_add(a, w);
    return w;
}

static void set_bt_init(wfo_aligner* a, wf_t* w, int k, int v, int h);

/* endsfree_begin_k (src/aligner.rs:3369-3386): with a match reward the
 * begin-free cells (h, 0) and (0, v), h = v = d, are reached at score
 * d * -match; d when `score` is such a score within a free range, else -1.
 * (With match == 0 they are all seeded at score 0.) */
static int endsfree_begin_k(const wfo_aligner* a, int score) {
    const int m = a->pen.match;
    if (m == 0 || !a->endsfree || (a->tb_free == 0 && a->pb_free == 0)) return -1;
    if (score % (-m) != 0) return -1;
    const int k = score / (-m);
    return a->tb_free >= k || a->pb_free >= k ? k : -1;
}

/* output_alloc_range with the ends-free widening of a begin-free score
 * (src/aligner.rs:3318-3342) */
static void output_alloc_range_at(wfo_aligner* a, int score, int lo, int hi, int* alo, int* ahi) {
    const int k = score >= 0 ? endsfree_begin_k(a, score) : -1;
    if (k >= 0) {
        if (-k < lo) lo = -k;
        if (k > hi) hi = k;
    }
    output_alloc_range(a, lo, hi, alo, ahi);
}

/* endsfree_init_offset (src/aligner.rs:3421-3437): a begin-free cell (v, h)
 * on diagonal k, starting a backtrace chain there */
static void endsfree_init_offset(wfo_aligner* a, wf_t* w, int k, int v, int h) {
    wf_set(w, k, h);
    if (a->piggyback) set_bt_init(a, w, k, v, h);
}
static void set_null_cell(wf_t* w, int k) {
    wf_set(w, k, OFFSET_NULL);
    wf_zero_bt(w, k);
}

/* endsfree_init (src/aligner.rs:3388-3419): the begin-free cells of score
 * d * -match added to a computed M before its ends are trimmed: (h = d, 0)
 * on diagonal d and (0, v = d) on -d, replacing a cell that reaches no
 * further, the range widened over NULL cells to reach them. */
static void endsfree_init(wfo_aligner* a, wf_t* wf, int k) {
    const int lo = wf->lo, hi = wf->hi;
    if (a->tb_free >= k) {
        if (hi >= k) {
            if (wf_get(wf, k) <= k) endsfree_init_offset(a, wf, k, 0, k);
        } else {
            for (int kk = hi + 1; kk < k; ++kk) set_null_cell(wf, kk);
            endsfree_init_offset(a, wf, k, 0, k);
            wf->hi = k;
        }
    }
    if (a->pb_free >= k) {
        if (lo <= -k) {
            if (wf_get(wf, -k) <= 0) endsfree_init_offset(a, wf, -k, k, 0);
        } else {
            endsfree_init_offset(a, wf, -k, k, 0);
            for (int kk = -k + 1; kk < lo; ++kk) set_null_cell(wf, kk);
            wf->lo = -k;
        }
    }
}

/* endsfree_null_m (src/aligner.rs:3439-3465): at a begin-free score whose
 * inputs are all null, M holds just the begin-free cells. */
static wf_t* endsfree_null_m(wfo_aligner* a, int k) {
    const int text = a->tb_free >= k, pattern = a->pb_free >= k;
    const int lo = pattern ? -k : text ? k : 0;
    const int hi = text ? k : pattern ? -k : 0;
    int alo, ahi;
    output_alloc_range(a, lo, hi, &alo, &ahi);
    wf_t* w = alloc_wf(a, alo, ahi);
    for (int kk = lo; kk <= hi; ++kk) set_null_cell(w, kk);
    if (text) endsfree_init_offset(a, w, k, 0, k);
    if (pattern) endsfree_init_offset(a, w, -k, k, 0);
    wf_set_limits(w, lo, hi);
    return w;
}

/* ------------------------------------------------------------------ */
/* Aligner construction                                                */
/* ------------------------------------------------------------------ */
wfo_aligner* wfo_aligner_new(const wfa_attr_t* attr, int32_t* err) {
    wfo_aligner* a = (wfo_aligner*)calloc(1, sizeof(wfo_aligner));
    a->attr = *attr;
    if (pen_init(&a->pen, &attr->penalties) != 0) {
        if (err) *err = WFA_ERR_INVALID_PENALTIES;
        free(a);
        return NULL;
    }
    /* BiWFA is end-to-end; Z-drop is refused with it (C since 7110ee3; the
     * reference's driver refuses X-drop too, src/bin/align_benchmark.rs:
     * 529-560, but its aligner runs X-drop sub-alignments,
     * tests/xdrop_steps.rs:57-92). */
    const int zdrop = attr->heuristic == WFA_HEURISTIC_ZDROP;
    if ((attr->memory_mode == WFA_MEMORY_ULTRALOW &&
         (attr->span != WFA_SPAN_END2END || zdrop))) {
        if (err) *err = WFA_ERR_UNSUPPORTED;
        free(a);
        return NULL;
    }
    if (err) *err = WFA_OK;
    a->status = WFA_STATUS_OK;
    a->cig_score = INT32_MIN;
    return a;
}

void wfo_aligner_free(wfo_aligner* a) {
    if (!a) return;
    pool_clear(a);
    free(a->pool);
    for (int c = 0; c < NCOMP; ++c) free(a->comp[c]);
    free(a->pattern);
    free(a->text);
    free(a->cig);
    free(a->blk_pc);
    free(a->blk_prev);
    free(a->init_v);
    free(a->init_h);
    free(a->packed);
    free(a->dist);
    free(a->dom);
    wfo_aligner_free(a->bi_fwd);
    wfo_aligner_free(a->bi_rev);
    wfo_aligner_free(a->bi_base);
    free(a);
}

int32_t wfo_status(const wfo_aligner* a) { return a->status; }
int32_t wfo_cigar(const wfo_aligner* a, const uint8_t** ops) {
    int n = a->cig_end - a->cig_begin;
    if (n <= 0) { if (ops) *ops = NULL; return 0; }
    if (ops) *ops = a->cig + a->cig_begin;
    return n;
}
int32_t wfo_cigar_score(const wfo_aligner* a) { return a->cig_score; }

/* ------------------------------------------------------------------ */
/* Backtrace buffer (src/bt_buffer.rs:113-209) and packed CIGARs       */
/* (src/pcigar.rs)                                                      */
/* ------------------------------------------------------------------ */
#define BT_NULL 0xFFFFFFFFu
#define PC_DEL 1u
#define PC_MISMS 2u
#define PC_INS 3u
#define PC_FULL_MASK 0x40000000u
#define PC_ALMOST_FULL_MASK 0x10000000u
#define PC_HALF_FULL_MASK 0x00010000u

static uint32_t store_block(wfo_aligner* a, uint32_t pcigar, uint32_t prev) {
    if (a->nblk == a->cap_blk) {
        a->cap_blk = a->cap_blk ? 2 * a->cap_blk : 4096;
        a->blk_pc = (uint32_t*)realloc(a->blk_pc, sizeof(uint32_t) * (size_t)a->cap_blk);
        a->blk_prev = (uint32_t*)realloc(a->blk_prev, sizeof(uint32_t) * (size_t)a->cap_blk);
    }
    a->blk_pc[a->nblk] = pcigar;
    a->blk_prev[a->nblk] = prev;
    return (uint32_t)a->nblk++;
}
/* init_block (src/bt_buffer.rs:135-139) */
static uint32_t init_block(wfo_aligner* a, int v, int h) {
    if (a->ninit == a->cap_init) {
        a->cap_init = a->cap_init ? 2 * a->cap_init : 64;
        a->init_v = (int32_t*)realloc(a->init_v, sizeof(int32_t) * (size_t)a->cap_init);
        a->init_h = (int32_t*)realloc(a->init_h, sizeof(int32_t) * (size_t)a->cap_init);
    }
    a->init_v[a->ninit] = v;
    a->init_h[a->ninit] = h;
    return store_block(a, (uint32_t)a->ninit++, BT_NULL);
}
/* set_bt_init (src/aligner.rs:216-225) */
static void set_bt_init(wfo_aligner* a, wf_t* w, int k, int v, int h) {
    w->pc[k - w->base] = 0;
    w->pv[k - w->base] = init_block(a, v, h);
}
static inline uint32_t rd_pc(const wf_t* w, int k) { return w ? w->pc[k - w->base] : 0; }
static inline uint32_t rd_pv(const wf_t* w, int k) { return w ? w->pv[k - w->base] : 0; }

/* offload_selective (src/compute/piggyback.rs:535-556) */
static void offload_selective(wfo_aligner* a, wf_t* w, int lo, int hi, uint32_t mask) {
    for (int k = lo; k <= hi; ++k) {
        uint32_t pc = w->pc[k - w->base], pv = w->pv[k - w->base];
        if (wf_get(w, k) >= 0 && pc >= mask) {
            uint32_t idx = store_block(a, pc, pv);
            w->pc[k - w->base] = 0;
            w->pv[k - w->base] = idx;
        }
    }
}
/* occupancy helpers (src/compute/piggyback.rs:591-620): a missing input is
 * the null wavefront, whose `null` flag is set */
static inline int occ(const wf_t* w) { return w ? w->occ : 0; }
static inline int occ_gap(const wf_t* open, const wf_t* ext) {
    int o = open ? open->occ + 1 : 0;
    if (ext && ext->occ + 1 > o) o = ext->occ + 1;
    return o;
}
/* offload_affine_component + offload_blocks_affine (src/compute/
 * piggyback.rs:576-590, 622-633) */
static void offload_affine_component(wfo_aligner* a, wf_t* w, int lo, int hi) {
    if (w->occ >= 16 - 1) {
        if (a->low) {
            offload_selective(a, w, lo, hi, PC_ALMOST_FULL_MASK);
            w->occ = 16 - 2;
        } else {
            offload_selective(a, w, lo, hi, PC_HALF_FULL_MASK);
            w->occ = 16 / 2;
        }
    }
}

static void heuristic_reset(wfo_aligner* a);

/* ------------------------------------------------------------------ */
/* init_alignment (src/aligner.rs:846-1000), end-to-end                */
/* ------------------------------------------------------------------ */
static void init_alignment_component(wfo_aligner* a, int comp);
static void init_alignment(wfo_aligner* a) {
    init_alignment_component(a, C_M);
    a->end_score = -1; a->end_k = DIAG_NULL; a->end_offset = OFFSET_NULL;
}
/* init_alignment_component (src/aligner.rs:3205-3266): as init_alignment,
 * with the score-0 cell (k = 0, offset 0) placed in component `comp`; the
 * end position starts as (0, 0, 0) there. */
static void init_alignment_component(wfo_aligner* a, int comp) {
    const int plen = a->plen, tlen = a->tlen;
    int score_only = a->attr.alignment_scope == WFA_SCOPE_SCORE;
    /* med/low alignments: modular storage with piggyback backtrace
     * (src/aligner.rs:855-866) */
    a->piggyback = !score_only && a->attr.memory_mode != WFA_MEMORY_HIGH;
    a->low = a->attr.memory_mode == WFA_MEMORY_LOW;
    a->modular = score_only || a->piggyback;
    pool_clear(a);
    a->nblk = 0;
    a->ninit = 0;
    /* Components are created for 1000 x 1000 and only grow
     * (src/aligner.rs:351-370, src/components.rs:190-213); this models a
     * freshly created aligner for every call. */
    int scope, nwf, s0, n0;
    compute_dimensions(&a->pen, 1000, 1000, a->modular, &s0, &n0);
    compute_dimensions(&a->pen, plen, tlen, a->modular, &scope, &nwf);
    if (n0 > nwf) nwf = n0;
    a->scope = scope;
    a->num_wavefronts = nwf;
    int nslots = a->modular ? scope : nwf;
    for (int c = 0; c < NCOMP; ++c) {
        a->comp[c] = (wf_t**)realloc(a->comp[c], sizeof(wf_t*) * (size_t)nslots);
        memset(a->comp[c], 0, sizeof(wf_t*) * (size_t)nslots);
    }
    a->nslots = nslots;
    if (!score_only) {
        int need = 2 * (plen + tlen) + 1; /* cigar_buffer_len, src/aligner.rs:378-380 */
        if (need > a->cig_cap) {
            a->cig = (uint8_t*)realloc(a->cig, (size_t)need);
            a->cig_cap = need;
        }
    }
    a->cig_begin = a->cig_end = 0;
    a->cig_score = INT32_MIN;
    a->num_null_steps = 0;
    a->status = WFA_STATUS_OK;
    heuristic_reset(a);
    a->end_score = 0; a->end_k = 0; a->end_offset = 0;
    a->hist_lo = -(plen + scope + 2);
    a->hist_hi = tlen + scope + 2;
    a->cand_valid = 0;
    a->zbest_pc = 0; a->zbest_pv = BT_NULL;
    /* begin-free dominance pruning (src/aligner.rs:925-945): corrected
     * ends-free with a match reward and begin-free gaps (at most 4 for the
     * affine kinds), no heuristic reading the wavefronts. Begin-free gaps
     * longer than their sequence (outside the reference's contract: its
     * driver refuses them) put seeds beyond the record's diagonals, which
     * then stays off. */
    a->dom_on = 0;
    if (a->attr.compat == WFA_COMPAT_CORRECTED && a->pen.match < 0 && a->endsfree &&
        a->pb_free <= a->plen && a->tb_free <= a->tlen) {
        const int bf = a->pb_free > a->tb_free ? a->pb_free : a->tb_free;
        const int heur = a->attr.heuristic;
        if (bf > 0 && (a->pen.metric == WFA_METRIC_GAP_LINEAR || bf <= 4) &&
            (heur == WFA_HEURISTIC_NONE || heur == WFA_HEURISTIC_BANDED_STATIC)) {
            a->dom_on = 1;
            a->dom_base = a->hist_lo;
            a->dom_len = a->hist_hi - a->hist_lo + 1;
            const int n = ncomp_of(a->pen.metric) * a->dom_len;
            if (n > a->cap_dom) {
                a->dom = (int32_t*)realloc(a->dom, sizeof(int32_t) * (size_t)n);
                a->cap_dom = n;
            }
            for (int i = 0; i < n; ++i) a->dom[i] = -1;
        }
    }
    int alo, ahi;
    int lo0 = 0, hi0 = 0;
    if (a->endsfree && a->pen.match == 0) { lo0 = -a->pb_free; hi0 = a->tb_free; }
    a->alloc_lo = 0; a->alloc_hi = 0;
    output_alloc_range(a, lo0, hi0, &alo, &ahi);
    wf_t* m0 = alloc_wf(a, alo, ahi);
    wf_set(m0, 0, 0);
    wf_set_limits(m0, 0, 0);
    if (a->piggyback) set_bt_init(a, m0, 0, 0, 0);
    /* ends-free with match == 0: the begin-free cells (h,0) and (0,v) are
     * reachable at score 0 (src/aligner.rs:966-998) */
    if (a->endsfree && a->pen.match == 0) {
        for (int h = 1; h <= a->tb_free; ++h) {
            wf_set(m0, h, h);
            if (a->piggyback) set_bt_init(a, m0, h, 0, h);
        }
        for (int v = 1; v <= a->pb_free; ++v) {
            wf_set(m0, -v, 0);
            if (a->piggyback) set_bt_init(a, m0, -v, v, 0);
        }
        m0->lo = -a->pb_free; m0->hi = a->tb_free;
        m0->init_min = -a->pb_free; m0->init_max = a->tb_free;
    }
    put(a, comp, 0, m0);
}

/* ------------------------------------------------------------------ */
/* Extension (src/extend.rs:47-69) and termination (src/termination.rs)*/
/* ------------------------------------------------------------------ */
static inline int lcp(const wfo_aligner* a, int v, int h) {
    int n = 0;
    while (v + n < a->plen && h + n < a->tlen && a->pattern[v + n] == a->text[h + n]) ++n;
    return n;
}

/* ------------------------------------------------------------------ */
/* Heuristics (src/heuristic.rs, src/aligner.rs:1644-1766)              */
/* ------------------------------------------------------------------ */
/* DropParams::new (src/heuristic.rs:75-105) */
typedef struct { int swg_match, pbf, tbf, max_penalty; } drop_t;

static drop_t drop_params(int match, int endsfree, int pbf, int tbf, int scope) {
    const int begin_free = match == 0 && endsfree;
    drop_t d;
    d.swg_match = match != 0 ? -match : 1;
    d.pbf = begin_free ? pbf : 0;
    d.tbf = begin_free ? tbf : 0;
    d.max_penalty = scope - 1;
    return d;
}
/* HeuristicState::reset (src/heuristic.rs:122-143) */
static void hstate_reset(hstate_t* st, int heuristic, const int* hp) {
    st->steps_wait = 0;
    if (heuristic == WFA_HEURISTIC_WFADAPTIVE || heuristic == WFA_HEURISTIC_BANDED_ADAPTIVE)
        st->steps_wait = hp[2];
    else if (heuristic == WFA_HEURISTIC_XDROP || heuristic == WFA_HEURISTIC_ZDROP)
        st->steps_wait = hp[1];
    st->max_sw_score = 0;
    st->max_sw_k = DIAG_NULL;
    st->max_sw_off = OFFSET_NULL;
    st->max_wf_score = 0;
    st->zlast = 0;
}
static void heuristic_reset(wfo_aligner* a) {
    hstate_reset(&a->hs, a->attr.heuristic, a->attr.heuristic_p);
}
/* wf_distance_end2end (src/heuristic.rs:150-160) */
static inline int32_t wf_distance(int32_t off, int k, int plen, int tlen) {
    if (off < 0) return INT32_MAX;
    int lv = plen - (off - k), lh = tlen - off;
    return lv > lh ? lv : lh;
}
static int32_t* dist_at(wfo_aligner* a, const wf_t* w) {
    /* scratch indexed by k - w->base over the allocation */
    if (w->nalloc > a->cap_dist) {
        a->cap_dist = w->nalloc * 2;
        a->dist = (int32_t*)realloc(a->dist, sizeof(int32_t) * (size_t)a->cap_dist);
    }
    return a->dist;
}
/* wfadaptive_reduce (src/heuristic.rs:303-339): drop diagonals from each
 * end while their distance is more than `thr` above the minimum, keeping
 * [min_k, max_k] */
static void wfadaptive_reduce(wf_t* wf, const int32_t* d, int dbase, int min_d, int thr,
                              int min_k, int max_k) {
    int top = max_k < wf->hi ? max_k : wf->hi;
    int lo_r = wf->lo;
    for (int k = wf->lo; k < top; ++k) {
        if ((int64_t)d[k - dbase] - min_d <= thr) break;
        lo_r++;
    }
    wf->lo = lo_r;
    int bottom = min_k > wf->lo ? min_k : wf->lo;
    int hi_r = wf->hi;
    for (int k = wf->hi; k > bottom; --k) {
        if ((int64_t)d[k - dbase] - min_d <= thr) break;
        hi_r--;
    }
    wf->hi = hi_r;
}
/* wfadaptive_cutoff (src/heuristic.rs:258-300) */
static void wfadaptive_cutoff(hstate_t* st, wf_t* wf, int32_t* d, int dbase, int plen, int tlen,
                              int min_len, int thr, int steps) {
    if (st->steps_wait > 0) return;
    if (wf->hi - wf->lo + 1 < min_len) return;
    int min_d = plen > tlen ? plen : tlen;
    for (int k = wf->lo; k <= wf->hi; ++k) {
        int32_t dd = wf_distance(wf_get(wf, k), k, plen, tlen);
        d[k - dbase] = dd;
        if (dd < min_d) min_d = dd;
    }
    const int ak = tlen - plen;
    wfadaptive_reduce(wf, d, dbase, min_d, thr, ak, ak);
    st->steps_wait = steps;
}
/* banded_adaptive_cutoff (src/heuristic.rs:355-394) */
static void banded_adaptive_cutoff(hstate_t* st, wf_t* wf, int plen, int tlen, int min_k,
                                   int max_k, int steps) {
    if (st->steps_wait > 0) return;
    int lo = wf->lo, hi = wf->hi, len = hi - lo + 1;
    if (len < 4) return;
    int maxlen = max_k - min_k + 1;
    if (len > maxlen) {
        int leeway = (len - maxlen) / 2, q = len / 4;
        int32_t d0 = wf_distance(wf_get(wf, lo), lo, plen, tlen);
        int32_t d1 = wf_distance(wf_get(wf, lo + q), lo + q, plen, tlen);
        int32_t d2 = wf_distance(wf_get(wf, lo + 2 * q), lo + 2 * q, plen, tlen);
        int32_t d3 = wf_distance(wf_get(wf, hi), hi, plen, tlen);
        int nlo = lo;
        if (d0 > d3) nlo += leeway;
        if (d1 > d2) nlo += leeway;
        wf->lo = nlo > lo ? nlo : lo;
        int nhi = nlo + maxlen - 1;
        wf->hi = nhi < hi ? nhi : hi;
    }
    st->steps_wait = steps;
}
/* banded_static_cutoff (src/heuristic.rs:396-403) */
static void banded_static_cutoff(wf_t* wf, int min_k, int max_k) {
    if (wf->lo < min_k) wf->lo = min_k;
    if (wf->hi > max_k) wf->hi = max_k;
}
/* compute_sw_scores (src/heuristic.rs:491-525): the Smith-Waterman score
 * of every non-null cell, free leading gaps left out; returns the first
 * cell with the largest one */
static void compute_sw_scores(const wf_t* wf, int32_t* sw, int swbase, const drop_t* dp, int score,
                              int* cmax, int* cmax_k, int* cmax_off) {
    int mx = INT32_MIN, mk = 0, mo = 0;
    for (int k = wf->lo; k <= wf->hi; ++k) {
        int32_t off = wf_get(wf, k);
        if (off < 0) continue;
        int fv = k < 0 ? ((-k) < dp->pbf ? -k : dp->pbf) : 0;
        int fh = k > 0 ? (k < dp->tbf ? k : dp->tbf) : 0;
        int v = off - k - fv, h = off - fh;
        int sws = (dp->swg_match * (v + h) - score) / 2;
        sw[k - swbase] = sws;
        if (mx < sws) { mx = sws; mk = k; mo = off; }
    }
    *cmax = mx; *cmax_k = mk; *cmax_off = mo;
}
/* xdrop_cutoff (src/heuristic.rs:410-448) */
static void xdrop_cutoff(hstate_t* st, wf_t* wf, int32_t* sw, int swbase, const drop_t* dp,
                         int score, int xdrop, int steps) {
    if (st->steps_wait > 0) return;
    int cmax, cmax_k, cmax_off;
    compute_sw_scores(wf, sw, swbase, dp, score, &cmax, &cmax_k, &cmax_off);
    if (st->max_sw_k != DIAG_NULL) {
        int k = wf->lo;
        while (k <= wf->hi && (wf_get(wf, k) < 0 || st->max_sw_score - sw[k - swbase] >= xdrop))
            k++;
        wf->lo = k;
        k = wf->hi;
        while (k >= wf->lo && (wf_get(wf, k) < 0 || st->max_sw_score - sw[k - swbase] >= xdrop))
            k--;
        wf->hi = k;
        if (cmax > st->max_sw_score) { st->max_sw_score = cmax; st->max_sw_k = cmax_k; }
    } else {
        st->max_sw_score = cmax;
        st->max_sw_k = cmax_k;
    }
    st->steps_wait = steps;
}
/* zdrop_cutoff (src/heuristic.rs:449-489): returns 1 when the drop has
 * lasted the largest single penalty */
static int zdrop_cutoff(hstate_t* st, wf_t* wf, int32_t* sw, int swbase, const drop_t* dp,
                        int score, int zdrop, int steps) {
    if (st->steps_wait > 0) return 0;
    int cmax, cmax_k, cmax_off;
    compute_sw_scores(wf, sw, swbase, dp, score, &cmax, &cmax_k, &cmax_off);
    if (st->max_sw_k == DIAG_NULL || cmax > st->max_sw_score) {
        /* a new best: its cell and the score it was reached at */
        st->max_sw_score = cmax;
        st->max_wf_score = score;
        st->max_sw_k = cmax_k;
        st->max_sw_off = cmax_off;
        st->zlast = score;
    } else if ((int64_t)st->max_sw_score - cmax <= zdrop) {
        st->zlast = score;
    } else if (score - st->zlast >= dp->max_penalty) {
        return 1;
    }
    st->steps_wait = steps;
    return 0;
}
/* heuristic_cutoff (src/heuristic.rs:167-255); narrows wf->lo/hi only (cut
 * cells keep their offsets and stay readable). Returns 1 when Z-drop ends
 * the alignment. */
static int heuristic_cutoff(wfo_aligner* a, wf_t* wf, int score) {
    if (wf->lo > wf->hi) return 0;
    a->hs.steps_wait -= 1;
    const int plen = a->plen, tlen = a->tlen;
    const int* hp = a->attr.heuristic_p;
    const drop_t dp = drop_params(a->pen.match, a->endsfree, a->pb_free, a->tb_free, a->scope);
    switch (a->attr.heuristic) {
    case WFA_HEURISTIC_WFADAPTIVE:
        wfadaptive_cutoff(&a->hs, wf, dist_at(a, wf), wf->base, plen, tlen, hp[0], hp[1], hp[2]);
        return 0;
    case WFA_HEURISTIC_BANDED_STATIC:
        banded_static_cutoff(wf, hp[0], hp[1]);
        return 0;
    case WFA_HEURISTIC_BANDED_ADAPTIVE:
        banded_adaptive_cutoff(&a->hs, wf, plen, tlen, hp[0], hp[1], hp[2]);
        return 0;
    case WFA_HEURISTIC_XDROP:
        xdrop_cutoff(&a->hs, wf, dist_at(a, wf), wf->base, &dp, score, hp[0], hp[1]);
        return 0;
    case WFA_HEURISTIC_ZDROP:
        return zdrop_cutoff(&a->hs, wf, dist_at(a, wf), wf->base, &dp, score, hp[0], hp[1]);
    default:
        return 0;
    }
}

/* heuristic_equate (src/heuristic.rs:526-538): an I/D row takes the
 * narrowed M bounds */
static void heuristic_equate(wf_t* d, const wf_t* wf) {
    if (wf->lo > d->lo) d->lo = wf->lo;
    if (wf->hi < d->hi) d->hi = wf->hi;
    if (d->lo > d->hi) { d->null = 1; d->init_min = d->lo; d->init_max = d->hi; }
}

/* apply_heuristic_cutoff (src/aligner.rs:1644-1766). Returns 1 when Z-drop
 * ends the alignment: the end position is then the best cell seen. */
static int apply_heuristic_cutoff(wfo_aligner* a, int score, wf_t* wf) {
    if (a->attr.heuristic == WFA_HEURISTIC_BANDED_STATIC &&
        wf->lo >= a->attr.heuristic_p[0] && wf->hi <= a->attr.heuristic_p[1])
        return 0;
    int lo_base = wf->lo, hi_base = wf->hi;
    const int max_wf_base = a->hs.max_wf_score;
    const int zdropped = heuristic_cutoff(a, wf, score);
    /* Corrected: keep the backtrace of a new Z-drop best cell, which the
     * end backtrace needs after its wavefront is gone (med/low,
     * src/aligner.rs:1676-1688) */
    if (a->attr.compat == WFA_COMPAT_CORRECTED && a->piggyback &&
        a->hs.max_wf_score == score && max_wf_base != score) {
        const int k = a->hs.max_sw_k;
        a->zbest_pc = wf->pc[k - wf->base];
        a->zbest_pv = wf->pv[k - wf->base];
    }
    if (zdropped) {
        a->end_score = a->hs.max_wf_score;
        a->end_k = a->hs.max_sw_k;
        a->end_offset = a->hs.max_sw_off;
        return 1;
    }
    if (lo_base > hi_base || (wf->lo == lo_base && wf->hi == hi_base)) return 0;
    if (wf->lo > wf->hi) {
        wf->null = 1;
        wf->init_min = wf->lo; wf->init_max = wf->hi; /* clear_init_range */
    }
    int nc = a->pen.metric == WFA_METRIC_GAP_AFFINE ? 3 : a->pen.metric == WFA_METRIC_GAP_AFFINE_2P ? 5 : 1;
    for (int c = 1; c < nc; ++c) {
        wf_t* d = get(a, c, score);
        if (d) heuristic_equate(d, wf);
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Compat::Corrected ends-free with a match reward                     */
/* ------------------------------------------------------------------ */
static inline int term_endsfree(const wfo_aligner* a, int k, int offset);

/* Wavefront::trim_nulled (src/wavefront.rs:459-490): narrow [lo, hi] to
 * [lo, hi] ∩ [blo, bhi], NULLing the cells dropped (the initialized range
 * is kept). */
static void trim_nulled(wf_t* w, int blo, int bhi) {
    if (!w || w->null || w->lo > w->hi || (blo <= w->lo && bhi >= w->hi)) return;
    const int olo = w->lo, ohi = w->hi;
    const int nlo = blo > olo ? blo : olo, nhi = bhi < ohi ? bhi : ohi;
    if (nlo > nhi) {
        for (int k = olo; k <= ohi; ++k) wf_set(w, k, OFFSET_NULL);
        w->null = 1;
    } else {
        for (int k = olo; k < nlo; ++k) wf_set(w, k, OFFSET_NULL);
        for (int k = nhi + 1; k <= ohi; ++k) wf_set(w, k, OFFSET_NULL);
    }
    w->lo = nlo;
    w->hi = nhi;
}

/* endsfree_prune (src/aligner.rs:1203-1245): every completion of a cell on
 * diagonal k at `score` totals at least score + -match * |k - k_end|; the
 * diagonals that cannot beat (tie) `total` are dropped from every
 * component at this score. */
static void endsfree_prune(wfo_aligner* a, int score, int total) {
    const int radius = (total - score) / -a->pen.match;
    const int k_end = a->tlen - a->plen;
    const int nc = ncomp_of(a->pen.metric);
    for (int c = 0; c < nc; ++c) trim_nulled(get(a, c, score), k_end - radius, k_end + radius);
}

/* endsfree_dominance_prune (src/aligner.rs:1247-1312): drop from the ends
 * of a component just computed the cells that reach no further than a cell
 * of that component did at an earlier score; `record` adds the kept ones
 * to the record (I/D; M is recorded after its extension). */
static void dominance_prune(wfo_aligner* a, wf_t* w, int c, int record) {
    if (!w || w->lo > w->hi) return;
    const int lo = w->lo, hi = w->hi;
    int32_t* best = a->dom + c * a->dom_len - a->dom_base;
    int first = lo, last = hi;
    while (first <= hi && wf_get(w, first) <= best[first]) ++first;
    while (last >= first && wf_get(w, last) <= best[last]) --last;
    if (record)
        for (int k = first; k <= last; ++k) {
            const int32_t o = wf_get(w, k);
            if (o > best[k]) best[k] = o;
        }
    for (int k = lo; k < first; ++k) wf_set(w, k, OFFSET_NULL);
    for (int k = last + 1; k <= hi; ++k) wf_set(w, k, OFFSET_NULL);
    if (first > last) {
        w->lo = 1; w->hi = 0; w->null = 1;
    } else {
        w->lo = first; w->hi = last;
    }
}

/* endsfree_dominance_update (src/aligner.rs:1314-1330) */
static void dominance_update(wfo_aligner* a, const wf_t* w, int c) {
    if (!w || w->null || w->lo > w->hi) return;
    int32_t* best = a->dom + c * a->dom_len - a->dom_base;
    for (int k = w->lo; k <= w->hi; ++k) {
        const int32_t o = wf_get(w, k);
        if (o > best[k]) best[k] = o;
    }
}

/* after a compute step: dominance pruning of every component (order and
 * recording as src/aligner.rs:2183-2190, 2540-2551, 1998-2001) */
static void dominance_prune_step(wfo_aligner* a, int score) {
    if (!a->dom_on) return;
    dominance_prune(a, get(a, C_M, score), 0, 0);
    if (a->pen.metric == WFA_METRIC_GAP_AFFINE || a->pen.metric == WFA_METRIC_GAP_AFFINE_2P) {
        dominance_prune(a, get(a, C_I1, score), 1, 1);
        dominance_prune(a, get(a, C_D1, score), 2, 1);
    }
    if (a->pen.metric == WFA_METRIC_GAP_AFFINE_2P) {
        dominance_prune(a, get(a, C_I2, score), 3, 1);
        dominance_prune(a, get(a, C_D2, score), 4, 1);
    }
}

/* endsfree_left (src/termination.rs:65-86): the bases a free end gap skips
 * from a cell at a free end (-1 if none) */
static int endsfree_left(const wfo_aligner* a, int k, int offset) {
    const int h = offset, v = offset - k;
    const int pl = a->plen - v, tl = a->tlen - h;
    const int by_text = h >= a->tlen && pl <= a->pe_free ? pl : -1;
    const int by_pattern = v >= a->plen && tl <= a->te_free ? tl : -1;
    if (by_text >= 0 && by_pattern >= 0) return by_text < by_pattern ? by_text : by_pattern;
    return by_text >= 0 ? by_text : by_pattern;
}

/* extend_endsfree_corrected (src/aligner.rs:1111-1200): a cell reaching a
 * free end at `score` with `left` bases to skip completes an alignment of
 * total score + -match * left; the best candidate ends the search once the
 * score reaches its total. */
static int extend_endsfree_corrected(wfo_aligner* a, int score) {
    wf_t* m = get(a, C_M, score);
    const int heur = a->attr.heuristic;
    const int prune = heur == WFA_HEURISTIC_NONE || heur == WFA_HEURISTIC_BANDED_STATIC;
    const int mr = -a->pen.match;
    if (prune && a->cand_valid && a->cand_total > score) endsfree_prune(a, score, a->cand_total);
    if (m) {
        /* extend_matches_packed_endsfree_best (src/extend.rs:179-232): the
         * first cell in ascending k with the fewest bases left */
        int bk = 0, boff = 0, bleft = -1;
        for (int k = m->lo; k <= m->hi; ++k) {
            int32_t off = wf_get(m, k);
            if (off == OFFSET_NULL) continue;
            off += lcp(a, off - k, off);
            wf_set(m, k, off);
            if (!term_endsfree(a, k, off)) continue;
            const int left = endsfree_left(a, k, off);
            if (left >= 0 && (bleft < 0 || left < bleft)) { bk = k; boff = off; bleft = left; }
        }
        if (bleft >= 0) {
            const int total = score + mr * bleft;
            if (!a->cand_valid || total < a->cand_total) {
                a->cand_valid = 1;
                a->cand_total = total;
                a->cand_score = score; a->cand_k = bk; a->cand_off = boff;
                /* bt_or_zero (src/wavefront.rs:405-416) */
                a->cand_pc = a->piggyback ? m->pc[bk - m->base] : 0;
                a->cand_pv = a->piggyback ? m->pv[bk - m->base] : BT_NULL;
                if (prune && total > score) endsfree_prune(a, score, total);
            }
        }
    }
    if (a->dom_on && !is_null_wf(m)) dominance_update(a, m, 0);
    if (a->cand_valid && a->cand_total <= score) {
        a->end_score = a->cand_score; a->end_k = a->cand_k; a->end_offset = a->cand_off;
        a->status = WFA_STATUS_END_REACHED;
        return 1;
    }
    int done;
    if (!m) /* with a candidate null steps do not end the search */
        done = !a->cand_valid && a->num_null_steps > a->scope;
    else
        done = heur != WFA_HEURISTIC_NONE && apply_heuristic_cutoff(a, score, m);
    if (done) {
        if (a->cand_valid) {
            a->end_score = a->cand_score; a->end_k = a->cand_k; a->end_offset = a->cand_off;
            a->status = WFA_STATUS_END_REACHED;
        } else {
            a->status = WFA_STATUS_END_UNREACHABLE;
        }
        return 1;
    }
    return 0;
}

/* extend_end2end (src/aligner.rs:1004-1049) */
static int extend_end2end(wfo_aligner* a, int score) {
    wf_t* m = get(a, C_M, score);
    if (!m) {
        if (a->num_null_steps > a->scope) {
            a->status = WFA_STATUS_END_UNREACHABLE;
            return 1;
        }
        return 0;
    }
    for (int k = m->lo; k <= m->hi; ++k) {
        int32_t off = wf_get(m, k);
        if (off >= 0) wf_set(m, k, off + lcp(a, off - k, off));
    }
    /* termination_end2end (src/termination.rs:14-23) */
    int k_end = a->tlen - a->plen;
    if (!(m->lo > k_end || k_end > m->hi) && wf_get(m, k_end) >= a->tlen) {
        a->end_score = score; a->end_k = k_end; a->end_offset = a->tlen;
        a->status = WFA_STATUS_END_REACHED;
        return 1;
    }
    if (a->attr.heuristic != WFA_HEURISTIC_NONE && apply_heuristic_cutoff(a, score, m)) {
        a->status = WFA_STATUS_END_UNREACHABLE; /* Z-dropped */
        return 1;
    }
    return 0;
}

/* termination_endsfree (src/termination.rs:31-59) */
static inline int term_endsfree(const wfo_aligner* a, int k, int offset) {
    int h = offset, v = offset - k;
    if (h >= a->tlen && a->plen - v <= a->pe_free) return 1;
    if (v >= a->plen && a->tlen - h <= a->te_free) return 1;
    return 0;
}

/* extend_endsfree (src/aligner.rs:1053-1102) with
 * extend_matches_packed_endsfree (src/extend.rs:112-170): the first
 * diagonal in ascending k that reaches a free end wins. */
static int extend_endsfree(wfo_aligner* a, int score) {
    if (a->attr.compat == WFA_COMPAT_CORRECTED && a->pen.match < 0)
        return extend_endsfree_corrected(a, score);
    wf_t* m = get(a, C_M, score);
    if (!m) {
        if (a->num_null_steps > a->scope) {
            a->status = WFA_STATUS_END_UNREACHABLE;
            return 1;
        }
        return 0;
    }
    for (int k = m->lo; k <= m->hi; ++k) {
        int32_t off = wf_get(m, k);
        if (off != OFFSET_NULL) {
            off += lcp(a, off - k, off);
            wf_set(m, k, off);
            if (term_endsfree(a, k, off)) {
                a->end_score = score; a->end_k = k; a->end_offset = off;
                a->status = WFA_STATUS_END_REACHED;
                return 1;
            }
        }
    }
    if (a->attr.heuristic != WFA_HEURISTIC_NONE && apply_heuristic_cutoff(a, score, m)) {
        a->status = WFA_STATUS_END_UNREACHABLE; /* Z-dropped */
        return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Compute steps                                                       */
/* ------------------------------------------------------------------ */
static void free_outputs(wfo_aligner* a, int score, int ncomp) {
    for (int c = 0; c < ncomp; ++c) put(a, c, score, NULL);
}

/* compute_affine_idm / compute_affine2p_idm (scalar semantics,
 * src/compute/affine.rs:318-351, src/compute/affine2p.rs:205-236):
 * I1 = max(Mo1[k-1], I1e[k-1]) + 1, D1 = max(Mo1[k+1], D1e[k+1]) (I2/D2
 * likewise from Mo2, I2e, D2e when `use2`), M = max(Mx[k] + 1, I*, D*),
 * NULL when out of the sequences (only M is bounds-checked). An absent
 * input reads OFFSET_NULL everywhere. */
static void affine_kernel(int plen, int tlen, const wf_t* mx, const wf_t* mo1, const wf_t* i1e,
                          const wf_t* d1e, const wf_t* mo2, const wf_t* i2e, const wf_t* d2e,
                          int use2, wf_t* m, wf_t* i1, wf_t* d1, wf_t* i2, wf_t* d2, int lo,
                          int hi) {
    for (int k = lo; k <= hi; ++k) {
        int32_t vi1 = rd(mo1, k - 1) > rd(i1e, k - 1) ? rd(mo1, k - 1) : rd(i1e, k - 1);
        vi1 += 1;
        int32_t vd1 = rd(mo1, k + 1) > rd(d1e, k + 1) ? rd(mo1, k + 1) : rd(d1e, k + 1);
        int32_t mm = rd(mx, k) + 1;
        int32_t ins = vi1, del = vd1;
        if (use2) {
            int32_t vi2 = (rd(mo2, k - 1) > rd(i2e, k - 1) ? rd(mo2, k - 1) : rd(i2e, k - 1)) + 1;
            int32_t vd2 = rd(mo2, k + 1) > rd(d2e, k + 1) ? rd(mo2, k + 1) : rd(d2e, k + 1);
            wf_set(i2, k, vi2);
            wf_set(d2, k, vd2);
            if (vi2 > ins) ins = vi2;
            if (vd2 > del) del = vd2;
        }
        int32_t mv = ins > del ? ins : del;
        if (mm > mv) mv = mm;
        if ((uint32_t)mv > (uint32_t)tlen || (uint32_t)(mv - k) > (uint32_t)plen) mv = OFFSET_NULL;
        wf_set(m, k, mv);
        wf_set(i1, k, vi1);
        wf_set(d1, k, vd1);
    }
}

/* compute_affine_step + compute_affine2p_step (src/aligner.rs:2006-2193,
 * 2201-2553) with the kernels of src/compute/affine.rs:318-351 and
 * src/compute/affine2p.rs:205-236. */
static void compute_affine_step(wfo_aligner* a, int score) {
    const int plen = a->plen, tlen = a->tlen;
    const pen_t* p = &a->pen;
    const int two = p->metric == WFA_METRIC_GAP_AFFINE_2P;
    const int x = p->mismatch, oe1 = p->o1 + p->e1, e1 = p->e1;
    const int oe2 = two ? p->o2 + p->e2 : 0, e2 = two ? p->e2 : 0;

    wf_t* mx = score - x >= 0 ? get(a, C_M, score - x) : NULL;
    wf_t* mo1 = score - oe1 >= 0 ? get(a, C_M, score - oe1) : NULL;
    wf_t* i1e = score - e1 >= 0 ? get(a, C_I1, score - e1) : NULL;
    wf_t* d1e = score - e1 >= 0 ? get(a, C_D1, score - e1) : NULL;
    wf_t *mo2 = NULL, *i2e = NULL, *d2e = NULL;
    if (two) {
        mo2 = score - oe2 >= 0 ? get(a, C_M, score - oe2) : NULL;
        i2e = score - e2 >= 0 ? get(a, C_I2, score - e2) : NULL;
        d2e = score - e2 >= 0 ? get(a, C_D2, score - e2) : NULL;
    }
    const int ncomp = two ? 5 : 3;
    const int bk = endsfree_begin_k(a, score);
    if (is_null_wf(mx) && is_null_wf(mo1) && is_null_wf(i1e) && is_null_wf(d1e) &&
        is_null_wf(mo2) && is_null_wf(i2e) && is_null_wf(d2e)) {
        a->num_null_steps++;
        free_outputs(a, score, ncomp);
        if (bk >= 0) put(a, C_M, score, endsfree_null_m(a, bk));
        return;
    }
    a->num_null_steps = 0;
    mx = live_or_none(mx); mo1 = live_or_none(mo1);
    i1e = live_or_none(i1e); d1e = live_or_none(d1e);
    mo2 = live_or_none(mo2); i2e = live_or_none(i2e); d2e = live_or_none(d2e);
    /* affine2p with no O2/E2 input runs the affine kernel and stores no
     * I2/D2 (src/aligner.rs:2261-2264, 2507-2524). */
    int use2 = two && (mo2 || i2e || d2e);

    int lo = nlo(mx), hi = nhi(mx);
#define LMIN(v) do { int _v = (v); if (_v < lo) lo = _v; } while (0)
#define HMAX(v) do { int _v = (v); if (_v > hi) hi = _v; } while (0)
    LMIN(nlo(mo1) - 1); LMIN(nlo(i1e) + 1); LMIN(nlo(d1e) - 1);
    HMAX(nhi(mo1) + 1); HMAX(nhi(i1e) + 1); HMAX(nhi(d1e) - 1);
    if (two) {
        LMIN(nlo(mo2) - 1); LMIN(nlo(i2e) + 1); LMIN(nlo(d2e) - 1);
        HMAX(nhi(mo2) + 1); HMAX(nhi(i2e) + 1); HMAX(nhi(d2e) - 1);
    }
#undef LMIN
#undef HMAX
    /* clamp_to_band (src/aligner.rs:3308-3316) */
    if (a->attr.heuristic == WFA_HEURISTIC_BANDED_STATIC ||
        a->attr.heuristic == WFA_HEURISTIC_BANDED_ADAPTIVE) {
        if (lo < a->attr.heuristic_p[0]) lo = a->attr.heuristic_p[0];
        if (hi > a->attr.heuristic_p[1]) hi = a->attr.heuristic_p[1];
    }
    if (lo > hi) {
        if (a->modular) free_outputs(a, score, ncomp);
        return;
    }
    if (lo < a->hist_lo + 1) lo = a->hist_lo + 1;
    if (hi > a->hist_hi - 1) hi = a->hist_hi - 1;
    if (lo > hi) {
        if (a->modular) free_outputs(a, score, ncomp);
        return;
    }
    int alo, ahi;
    output_alloc_range_at(a, score, lo, hi, &alo, &ahi);
    if (mx) { wf_init_ends_higher(mx, hi); wf_init_ends_lower(mx, lo); }
    if (mo1) { wf_init_ends_higher(mo1, hi + 1); wf_init_ends_lower(mo1, lo - 1); }
    if (mo2) { wf_init_ends_higher(mo2, hi + 1); wf_init_ends_lower(mo2, lo - 1); }
    if (i1e) { wf_init_ends_higher(i1e, hi); wf_init_ends_lower(i1e, lo - 1); }
    if (i2e) { wf_init_ends_higher(i2e, hi); wf_init_ends_lower(i2e, lo - 1); }
    if (d1e) { wf_init_ends_higher(d1e, hi + 1); wf_init_ends_lower(d1e, lo); }
    if (d2e) { wf_init_ends_higher(d2e, hi + 1); wf_init_ends_lower(d2e, lo); }

    wf_t* m = alloc_wf(a, alo, ahi);
    wf_t* i1 = alloc_wf(a, alo, ahi);
    wf_t* d1 = alloc_wf(a, alo, ahi);
    wf_t *i2 = NULL, *d2 = NULL;
    wf_set_limits(m, lo, hi); wf_set_limits(i1, lo, hi); wf_set_limits(d1, lo, hi);
    if (use2) {
        i2 = alloc_wf(a, alo, ahi); d2 = alloc_wf(a, alo, ahi);
        wf_set_limits(i2, lo, hi); wf_set_limits(d2, lo, hi);
    }
    if (a->piggyback) {
        /* compute_affine_idm / compute_affine2p_idm with piggyback
         * (src/compute/piggyback.rs:341-393, 438-528): a gap cell takes the
         * extension if ext >= open; M takes I1, I2, D1, D2 then misms, a
         * later source winning ties; M always pushes MISMS. */
        for (int k = lo; k <= hi; ++k) {
            int32_t oo, ee, ins1, del1, ins2 = 0, del2 = 0;
            uint32_t ipc1, ipv1, dpc1, dpv1, ipc2 = 0, ipv2 = 0, dpc2 = 0, dpv2 = 0;
            oo = rd(mo1, k - 1); ee = rd(i1e, k - 1);
            if (ee >= oo) { ins1 = ee; ipc1 = rd_pc(i1e, k - 1); ipv1 = rd_pv(i1e, k - 1); }
            else { ins1 = oo; ipc1 = rd_pc(mo1, k - 1); ipv1 = rd_pv(mo1, k - 1); }
            ins1 += 1; ipc1 = (ipc1 << 2) | PC_INS;
            oo = rd(mo1, k + 1); ee = rd(d1e, k + 1);
            if (ee >= oo) { del1 = ee; dpc1 = rd_pc(d1e, k + 1); dpv1 = rd_pv(d1e, k + 1); }
            else { del1 = oo; dpc1 = rd_pc(mo1, k + 1); dpv1 = rd_pv(mo1, k + 1); }
            dpc1 = (dpc1 << 2) | PC_DEL;
            wf_set(i1, k, ins1); i1->pc[k - i1->base] = ipc1; i1->pv[k - i1->base] = ipv1;
            wf_set(d1, k, del1); d1->pc[k - d1->base] = dpc1; d1->pv[k - d1->base] = dpv1;
            int32_t mis = rd(mx, k) + 1, mxv;
            uint32_t pc, pv;
            if (use2) {
                oo = rd(mo2, k - 1); ee = rd(i2e, k - 1);
                if (ee >= oo) { ins2 = ee; ipc2 = rd_pc(i2e, k - 1); ipv2 = rd_pv(i2e, k - 1); }
                else { ins2 = oo; ipc2 = rd_pc(mo2, k - 1); ipv2 = rd_pv(mo2, k - 1); }
                ins2 += 1; ipc2 = (ipc2 << 2) | PC_INS;
                oo = rd(mo2, k + 1); ee = rd(d2e, k + 1);
                if (ee >= oo) { del2 = ee; dpc2 = rd_pc(d2e, k + 1); dpv2 = rd_pv(d2e, k + 1); }
                else { del2 = oo; dpc2 = rd_pc(mo2, k + 1); dpv2 = rd_pv(mo2, k + 1); }
                dpc2 = (dpc2 << 2) | PC_DEL;
                wf_set(i2, k, ins2); i2->pc[k - i2->base] = ipc2; i2->pv[k - i2->base] = ipv2;
                wf_set(d2, k, del2); d2->pc[k - d2->base] = dpc2; d2->pv[k - d2->base] = dpv2;
                mxv = ins1;
                if (ins2 > mxv) mxv = ins2;
                if (del1 > mxv) mxv = del1;
                if (del2 > mxv) mxv = del2;
                if (mis > mxv) mxv = mis;
                pc = ipc1; pv = ipv1;
                if (mxv == ins2) { pc = ipc2; pv = ipv2; }
                if (mxv == del1) { pc = dpc1; pv = dpv1; }
                if (mxv == del2) { pc = dpc2; pv = dpv2; }
                if (mxv == mis) { pc = rd_pc(mx, k); pv = rd_pv(mx, k); }
            } else {
                mxv = mis > ins1 ? mis : ins1;
                if (del1 > mxv) mxv = del1;
                pc = ipc1; pv = ipv1;
                if (mxv == del1) { pc = dpc1; pv = dpv1; }
                if (mxv == mis) { pc = rd_pc(mx, k); pv = rd_pv(mx, k); }
            }
            m->pc[k - m->base] = (pc << 2) | PC_MISMS;
            m->pv[k - m->base] = pv;
            if ((uint32_t)mxv > (uint32_t)tlen || (uint32_t)(mxv - k) > (uint32_t)plen)
                mxv = OFFSET_NULL;
            wf_set(m, k, mxv);
        }
        /* offload_affine / offload_affine2p (src/compute/piggyback.rs:636-705) */
        int oi1 = occ_gap(mo1, i1e), od1 = occ_gap(mo1, d1e);
        int om = occ(mx);
        if (oi1 > om) om = oi1;
        if (od1 > om) om = od1;
        int oi2 = 0, od2 = 0;
        if (use2) {
            oi2 = occ_gap(mo2, i2e); od2 = occ_gap(mo2, d2e);
            if (oi2 > om) om = oi2;
            if (od2 > om) om = od2;
            i2->occ = oi2; d2->occ = od2;
        }
        i1->occ = oi1; d1->occ = od1; m->occ = om + 1;
        offload_affine_component(a, m, lo, hi);
        offload_affine_component(a, i1, lo, hi);
        offload_affine_component(a, d1, lo, hi);
        if (use2) {
            offload_affine_component(a, i2, lo, hi);
            offload_affine_component(a, d2, lo, hi);
        }
    } else {
        affine_kernel(plen, tlen, mx, mo1, i1e, d1e, mo2, i2e, d2e, use2, m, i1, d1, i2, d2, lo,
                      hi);
    }
    put(a, C_M, score, m); put(a, C_I1, score, i1); put(a, C_D1, score, d1);
    if (two) { put(a, C_I2, score, i2); put(a, C_D2, score, d2); }
    if (bk >= 0) endsfree_init(a, m, bk);
    trim_ends(plen, tlen, m);
    trim_ends(plen, tlen, i1);
    trim_ends(plen, tlen, d1);
    if (i2) trim_ends(plen, tlen, i2);
    if (d2) trim_ends(plen, tlen, d2);
    dominance_prune_step(a, score);
}

/* compute_edit_idm / compute_indel_idm (src/compute/edit.rs:104-118,
 * 218-231): edit M = max(prev[k+1], max(prev[k-1], prev[k]) + 1); indel
 * M = max(prev[k+1], prev[k-1] + 1); NULL when out of the sequences. */
static void edit_kernel(int plen, int tlen, const wf_t* prev, wf_t* cur, int lo, int hi,
                        int indel) {
    for (int k = lo; k <= hi; ++k) {
        int32_t ins = wf_get(prev, k - 1), del = wf_get(prev, k + 1), mx;
        if (indel) {
            ins += 1;
            mx = del > ins ? del : ins;
        } else {
            int32_t mis = wf_get(prev, k);
            int32_t t = (ins > mis ? ins : mis) + 1;
            mx = del > t ? del : t;
        }
        if ((uint32_t)mx > (uint32_t)tlen || (uint32_t)(mx - k) > (uint32_t)plen) mx = OFFSET_NULL;
        wf_set(cur, k, mx);
    }
}

/* compute_edit_step (src/aligner.rs:1769-1865) with compute_edit_idm /
 * compute_indel_idm (src/compute/edit.rs:104-118, 218-231). */
static void compute_edit_step(wfo_aligner* a, int score) {
    const int plen = a->plen, tlen = a->tlen;
    wf_t* prev = get(a, C_M, score - 1);
    if (is_null_wf(prev)) {
        a->num_null_steps++;
        put(a, C_M, score, NULL);
        return;
    }
    a->num_null_steps = 0;
    int lo = prev->lo - 1, hi = prev->hi + 1;
    if (lo < a->hist_lo + 1) lo = a->hist_lo + 1;
    if (hi > a->hist_hi - 1) hi = a->hist_hi - 1;
    if (lo > hi) { put(a, C_M, score, NULL); return; }
    /* NULL the cells read beyond prev's ends, leaving its init range as is */
    int ks[4] = {prev->lo - 2, prev->lo - 1, prev->hi + 1, prev->hi + 2};
    for (int i = 0; i < 4; ++i)
        if (wf_allocated(prev, ks[i])) { wf_set(prev, ks[i], OFFSET_NULL); wf_zero_bt(prev, ks[i]); }
    int alo, ahi;
    output_alloc_range(a, lo, hi, &alo, &ahi);
    wf_t* cur = alloc_wf(a, alo, ahi);
    wf_set_limits(cur, lo, hi);
    int indel = a->pen.metric == WFA_METRIC_INDEL;
    if (a->piggyback) {
        /* edit_kernel / indel_kernel (src/compute/piggyback.rs:195-285):
         * later checks win: ins, then del, then misms */
        for (int k = lo; k <= hi; ++k) {
            int32_t ins = wf_get(prev, k - 1) + 1, del = wf_get(prev, k + 1), mx;
            uint32_t pc = (prev->pc[k - 1 - prev->base] << 2) | PC_INS;
            uint32_t pv = prev->pv[k - 1 - prev->base];
            if (indel) {
                mx = del > ins ? del : ins;
                if (mx == del) {
                    pc = (prev->pc[k + 1 - prev->base] << 2) | PC_DEL;
                    pv = prev->pv[k + 1 - prev->base];
                }
            } else {
                int32_t mis = wf_get(prev, k) + 1;
                mx = ins > mis ? ins : mis;
                if (del > mx) mx = del;
                if (mx == del) {
                    pc = (prev->pc[k + 1 - prev->base] << 2) | PC_DEL;
                    pv = prev->pv[k + 1 - prev->base];
                }
                if (mx == mis) {
                    pc = (prev->pc[k - prev->base] << 2) | PC_MISMS;
                    pv = prev->pv[k - prev->base];
                }
            }
            if ((uint32_t)mx > (uint32_t)tlen || (uint32_t)(mx - k) > (uint32_t)plen) mx = OFFSET_NULL;
            wf_set(cur, k, mx);
            cur->pc[k - cur->base] = pc;
            cur->pv[k - cur->base] = pv;
        }
        /* offload every 16 scores (src/aligner.rs:1849-1852) */
        if (score % 16 == 0)
            offload_selective(a, cur, lo, hi, a->low ? PC_FULL_MASK : PC_HALF_FULL_MASK);
    } else
        edit_kernel(plen, tlen, prev, cur, lo, hi, indel);
    put(a, C_M, score, cur);
    trim_ends(plen, tlen, cur);
}

/* compute_linear_step (src/aligner.rs:1866-2002) with compute_linear_idm
 * (src/compute/linear.rs:20-131) and, for med/low memory, the piggyback
 * kernel (src/compute/piggyback.rs:279-302, the edit cell update over the
 * two inputs) and offload_linear (src/compute/piggyback.rs:556-613).
 * M[s,k] = max(M[s-x,k] + 1, M[s-o,k-1] + 1, M[s-o,k+1]). Unlike the edit
 * step, the inputs' ends are NULL-filled only beyond their initialized
 * ranges (init_ends), so cells a heuristic cut keep their offsets. Ends-free
 * with a match reward adds the begin-free cells of this score
 * (endsfree_init / endsfree_null_m). */
static void compute_linear_step(wfo_aligner* a, int score) {
    const int plen = a->plen, tlen = a->tlen;
    const pen_t* p = &a->pen;
    wf_t* mx = score - p->mismatch >= 0 ? get(a, C_M, score - p->mismatch) : NULL;
    wf_t* mg = score - p->o1 >= 0 ? get(a, C_M, score - p->o1) : NULL;
    const int bk = endsfree_begin_k(a, score);
    if (is_null_wf(mx) && is_null_wf(mg)) {
        a->num_null_steps++;
        put(a, C_M, score, bk >= 0 ? endsfree_null_m(a, bk) : NULL);
        return;
    }
    a->num_null_steps = 0;
    /* trimmed-empty inputs read as the null wavefront; its bounds (1, -1)
     * still enter the union */
    mx = live_or_none(mx);
    mg = live_or_none(mg);
    int lo = nlo(mx), hi = nhi(mx);
    if (nlo(mg) - 1 < lo) lo = nlo(mg) - 1;
    if (nhi(mg) + 1 > hi) hi = nhi(mg) + 1;
    /* clamp_to_band (src/aligner.rs:3308-3316) */
    if (a->attr.heuristic == WFA_HEURISTIC_BANDED_STATIC ||
        a->attr.heuristic == WFA_HEURISTIC_BANDED_ADAPTIVE) {
        if (lo < a->attr.heuristic_p[0]) lo = a->attr.heuristic_p[0];
        if (hi > a->attr.heuristic_p[1]) hi = a->attr.heuristic_p[1];
    }
    if (lo < a->hist_lo + 1) lo = a->hist_lo + 1;
    if (hi > a->hist_hi - 1) hi = a->hist_hi - 1;
    if (lo > hi) { put(a, C_M, score, NULL); return; }
    int alo, ahi;
    output_alloc_range_at(a, score, lo, hi, &alo, &ahi);
    if (mx) { wf_init_ends_higher(mx, hi); wf_init_ends_lower(mx, lo); }
    if (mg) { wf_init_ends_higher(mg, hi + 1); wf_init_ends_lower(mg, lo - 1); }
    wf_t* cur = alloc_wf(a, alo, ahi);
    wf_set_limits(cur, lo, hi);
    for (int k = lo; k <= hi; ++k) {
        const int32_t mis = rd(mx, k) + 1, ins = rd(mg, k - 1) + 1, del = rd(mg, k + 1);
        int32_t m = ins > mis ? ins : mis;
        if (del > m) m = del;
        if (a->piggyback) {
            /* later checks win: ins, then del, then misms */
            uint32_t pc = (rd_pc(mg, k - 1) << 2) | PC_INS, pv = rd_pv(mg, k - 1);
            if (m == del) { pc = (rd_pc(mg, k + 1) << 2) | PC_DEL; pv = rd_pv(mg, k + 1); }
            if (m == mis) { pc = (rd_pc(mx, k) << 2) | PC_MISMS; pv = rd_pv(mx, k); }
            cur->pc[k - cur->base] = pc;
            cur->pv[k - cur->base] = pv;
        }
        if ((uint32_t)m > (uint32_t)tlen || (uint32_t)(m - k) > (uint32_t)plen) m = OFFSET_NULL;
        wf_set(cur, k, m);
    }
    if (a->piggyback) {
        /* offload_linear: every cell holds at most one op more than its
         * inputs' fullest; full chains move to the block store */
        int o = occ(mg) > occ(mx) ? occ(mg) : occ(mx);
        cur->occ = o + 1;
        if (cur->occ >= 16) {
            offload_selective(a, cur, lo, hi, a->low ? PC_FULL_MASK : PC_HALF_FULL_MASK);
            cur->occ = a->low ? 16 - 1 : 16 / 2;
        }
    }
    put(a, C_M, score, cur);
    if (bk >= 0) endsfree_init(a, cur, bk);
    trim_ends(plen, tlen, cur);
    dominance_prune_step(a, score);
}

static void compute_step(wfo_aligner* a, int score) {
    switch (a->pen.metric) {
    case WFA_METRIC_EDIT:
    case WFA_METRIC_INDEL: compute_edit_step(a, score); break;
    case WFA_METRIC_GAP_LINEAR: compute_linear_step(a, score); break;
    default: compute_affine_step(a, score); break;
    }
}

/* ------------------------------------------------------------------ */
/* Backtraces                                                          */
/* ------------------------------------------------------------------ */
static inline void write_ops(wfo_aligner* a, uint8_t op, int n) {
    for (int i = 0; i < n; ++i) a->cig[a->cig_begin--] = op;
}

/* st_m_offset (src/backtrace.rs:645-663): 1 and *out if the M cell exists. */
static int st_m_offset(const wfo_aligner* a, int score, int k, int32_t* out) {
    if (score < 0 || !has_slot(a, score)) return 0;
    const wf_t* w = get(a, C_M, score);
    if (!w) return 0;
    if (k < w->init_min || w->init_max < k) return 0;
    int32_t off = wf_get(w, k);
    if (off == OFFSET_NULL) return 0;
    *out = off;
    return 1;
}

/* st_backward_matches (src/backtrace.rs:668-686) */
static int st_backward_matches(const wfo_aligner* a, int v, int h) {
    int mx = v < h ? v : h, n = 0;
    while (n < mx && a->pattern[v - n - 1] == a->text[h - n - 1]) ++n;
    return n;
}

/* st_gap (src/backtrace.rs:707-737) */
static int st_gap(wfo_aligner* a, int gap_score, int l, int offset_orig, int* k, int* offset) {
    int32_t m;
    if (st_m_offset(a, gap_score, *k - l, &m) && m + l >= offset_orig && m + l <= *offset) {
        write_ops(a, 'M', *offset - (m + l));
        write_ops(a, 'I', l);
        *k -= l;
        *offset = m;
        return 1;
    }
    if (st_m_offset(a, gap_score, *k + l, &m) && m >= offset_orig && m <= *offset) {
        write_ops(a, 'M', *offset - m);
        write_ops(a, 'D', l);
        *k += l;
        *offset = m;
        return 1;
    }
    return 0;
}

/* backtrace_affine_singletrack (src/backtrace.rs:751-858), component_end M,
 * end-to-end (begin-free 0). */
static void backtrace_singletrack(wfo_aligner* a, int score, int k, int offset) {
    const pen_t* p = &a->pen;
    const int two = p->metric == WFA_METRIC_GAP_AFFINE_2P;
    a->cig_end = a->cig_cap;
    a->cig_begin = a->cig_cap - 1;
    const int alignment_score = score;
    int offset_orig = offset;
    {
        int v = offset - k, h = offset;
        write_ops(a, 'D', a->plen - v);
        write_ops(a, 'I', a->tlen - h);
    }
    int in_m = 1, l = 0;
    while (score != 0) {
        if (in_m) {
            int v = offset - k, h = offset;
            int nm = st_backward_matches(a, v, h);
            offset_orig = offset - nm;
            int mis = score - p->mismatch;
            int32_t mo;
            if (st_m_offset(a, mis, k, &mo) && mo == offset_orig - 1) {
                write_ops(a, 'M', nm);
                write_ops(a, 'X', 1);
                offset = offset_orig - 1;
                score = mis;
            } else {
                in_m = 0;
                l = 0;
            }
        } else {
            l += 1;
            int indel1 = score - p->o1 - l * p->e1;
            int indel2 = two ? score - p->o2 - l * p->e2 : -1;
            if (st_gap(a, indel1, l, offset_orig, &k, &offset)) {
                score = indel1;
                in_m = 1;
            } else if (st_gap(a, indel2, l, offset_orig, &k, &offset)) {
                score = indel2;
                in_m = 1;
            } else if (indel1 < 0 && indel2 < 0) {
                in_m = (offset_orig - k) == 0 || offset_orig == 0;
                break;
            }
        }
    }
    int v = offset - k, h = offset;
    int nm = v < h ? v : h;
    v -= nm; h -= nm;
    write_ops(a, 'M', nm);
    write_ops(a, 'D', v);
    write_ops(a, 'I', h);
    a->cig_begin += 1;
    a->cig_score = alignment_score;
}

/* bt_misms / bt_ins1_open / bt_del1_open (src/backtrace.rs:40-105), packed
 * as (offset << 4) | type in 64 bits. */
#define BT_I1_OPEN 1
#define BT_D1_OPEN 5
#define BT_M 9
static int64_t bt_pack(int32_t off, int64_t t) { return (int64_t)off * 16 + t; }
static int64_t bt_misms(const wfo_aligner* a, int score, int k) {
    if (score < 0) return OFFSET_NULL;
    const wf_t* w = get(a, C_M, score);
    if (!w) return OFFSET_NULL;
    if (w->init_min <= k && k <= w->init_max) return bt_pack(wf_get(w, k) + 1, BT_M);
    return OFFSET_NULL;
}
static int64_t bt_ins1_open(const wfo_aligner* a, int score, int k) {
    if (score < 0) return OFFSET_NULL;
    const wf_t* w = get(a, C_M, score);
    if (!w) return OFFSET_NULL;
    if (w->init_min < k && k - 1 <= w->init_max) return bt_pack(wf_get(w, k - 1) + 1, BT_I1_OPEN);
    return OFFSET_NULL;
}
static int64_t bt_del1_open(const wfo_aligner* a, int score, int k) {
    if (score < 0) return OFFSET_NULL;
    const wf_t* w = get(a, C_M, score);
    if (!w) return OFFSET_NULL;
    if (w->init_min <= k + 1 && k < w->init_max) return bt_pack(wf_get(w, k + 1), BT_D1_OPEN);
    return OFFSET_NULL;
}
static void write_matches(wfo_aligner* a, int n) {
    if (n < 0) { a->cig_begin += -n; return; }
    write_ops(a, 'M', n);
}

/* backtrace_linear (src/backtrace.rs:132-258): edit / indel CIGAR. */
static void backtrace_linear(wfo_aligner* a, int score, int k, int offset) {
    const pen_t* p = &a->pen;
    a->cig_end = a->cig_cap;
    a->cig_begin = a->cig_cap - 1;
    const int alignment_score = score;
    int v = offset - k, h = offset;
    if (v < a->plen) write_ops(a, 'D', a->plen - v);
    if (h < a->tlen) write_ops(a, 'I', a->tlen - h);
    while (v > 0 && h > 0 && score > 0) {
        int mis_s = score - p->mismatch, gap_s = score - p->o1;
        int64_t mis = p->metric != WFA_METRIC_INDEL ? bt_misms(a, mis_s, k) : OFFSET_NULL;
        int64_t ins = bt_ins1_open(a, gap_s, k);
        int64_t del = bt_del1_open(a, gap_s, k);
        int64_t mx = ins > del ? ins : del;
        if (mis > mx) mx = mis;
        if (mx < 0) break;
        int32_t moff = (int32_t)(mx >> 4);
        write_matches(a, offset - moff);
        offset = moff;
        v = offset - k; h = offset;
        if (v <= 0 || h <= 0) break;
        switch (mx & 0xF) {
        case BT_M: score = mis_s; write_ops(a, 'X', 1); offset -= 1; break;
        case BT_I1_OPEN: score = gap_s; write_ops(a, 'I', 1); k -= 1; offset -= 1; break;
        case BT_D1_OPEN: score = gap_s; write_ops(a, 'D', 1); k += 1; break;
        default: abort();
        }
        v = offset - k; h = offset;
    }
    if (v > 0 && h > 0) {
        int nm = v < h ? v : h;
        write_matches(a, nm);
        v -= nm; h -= nm;
    }
    while (v > 0) { write_ops(a, 'D', 1); v--; }
    while (h > 0) { write_ops(a, 'I', 1); h--; }
    a->cig_begin += 1;
    a->cig_score = alignment_score;
}

/* backtrace_piggyback (src/aligner.rs:1443-1481) -> traceback_from
 * (src/bt_buffer.rs:198-209) -> unpack_piggyback (src/aligner.rs:232-311). */
static void backtrace_piggyback_bt(wfo_aligner* a, uint32_t pcigar, uint32_t prev, int score,
                                   int k, int offset);
static void backtrace_piggyback(wfo_aligner* a, int score, int k, int offset) {
    const wf_t* m = get(a, C_M, score);
    if (!m || !wf_allocated(m, k)) { a->cig_begin = a->cig_end = 0; return; }
    backtrace_piggyback_bt(a, m->pc[k - m->base], m->pv[k - m->base], score, k, offset);
}

/* backtrace_piggyback_bt (src/aligner.rs:1462-1487): unpack the chain
 * (pcigar, prev) of M cell (k, offset) at `score`. */
static void backtrace_piggyback_bt(wfo_aligner* a, uint32_t pcigar, uint32_t prev, int score,
                                   int k, int offset) {
    /* chain: collect pcigars end-first until the initial block */
    int np = 0;
    while (prev != BT_NULL) {
        if (np == a->cap_packed) {
            a->cap_packed = a->cap_packed ? 2 * a->cap_packed : 256;
            a->packed = (uint32_t*)realloc(a->packed, sizeof(uint32_t) * (size_t)a->cap_packed);
        }
        a->packed[np++] = pcigar;
        pcigar = a->blk_pc[prev];
        prev = a->blk_prev[prev];
    }
    int bv = a->init_v[pcigar], bh = a->init_h[pcigar];
    int ev = offset - k, eh = offset;
    const int affine = a->pen.metric == WFA_METRIC_GAP_AFFINE ||
                       a->pen.metric == WFA_METRIC_GAP_AFFINE_2P;
    uint8_t* ops = a->cig;
    int n = 0, v = bv, h = bh;
    for (int i = 0; i < h; ++i) ops[n++] = 'I';
    for (int i = 0; i < v; ++i) ops[n++] = 'D';
    int in_m = 1;
    for (int b = np - 1; b >= 0; --b) {
        uint32_t pc = a->packed[b];
        int length = 16;
        if (pc < PC_FULL_MASK) {
            int free_slots = pc ? __builtin_clz(pc) / 2 : 16;
            length -= free_slots;
            pc = free_slots >= 16 ? 0 : pc << (2 * free_slots);
        }
        for (int j = 0; j < length; ++j) {
            if (in_m || !affine) {
                while (v < a->plen && h < a->tlen && a->pattern[v] == a->text[h]) {
                    ops[n++] = 'M'; v++; h++;
                }
            }
            uint32_t op = pc >> 30;
            pc <<= 2;
            uint8_t ch; int dv, dh, mm;
            if (op == PC_DEL) { ch = 'D'; dv = 1; dh = 0; mm = 0; }
            else if (op == PC_MISMS) { ch = 'X'; dv = 1; dh = 1; mm = 1; }
            else if (op == PC_INS) { ch = 'I'; dv = 0; dh = 1; mm = 0; }
            else { ch = '?'; dv = 0; dh = 0; mm = 1; }
            if (affine && !in_m && ch == 'X') { in_m = 1; continue; }
            ops[n++] = ch; v += dv; h += dh; in_m = mm;
        }
    }
    int nm = (ev - v) < (eh - h) ? (ev - v) : (eh - h);
    for (int i = 0; i < nm; ++i) ops[n++] = 'M';
    v += nm; h += nm;
    while (h < a->tlen) { ops[n++] = 'I'; h++; }
    while (v < a->plen) { ops[n++] = 'D'; v++; }
    a->cig_begin = 0;
    a->cig_end = n;
    a->cig_score = score;
}

/* backtrace_from (src/aligner.rs:1414-1437). */
static void backtrace_from(wfo_aligner* a, int score, int k, int offset) {
    if (a->piggyback) { backtrace_piggyback(a, score, k, offset); return; }
    if (a->pen.metric == WFA_METRIC_GAP_AFFINE || a->pen.metric == WFA_METRIC_GAP_AFFINE_2P)
        backtrace_singletrack(a, score, k, offset);
    else
        backtrace_linear(a, score, k, offset);
}

/* backtrace_zdrop_end (src/aligner.rs:1393-1406): the best cell, backtraced
 * from the score it was found at (Compat::C: high memory; medium/low memory
 * read the best cell's diagonal in the M wavefront of the current score,
 * whatever it holds there: an unallocated cell gives an empty CIGAR, one
 * never written the zeroed chain of the first block (backtrace_piggyback,
 * src/aligner.rs:1443-1460, HPCBio/wfa2lib-rs#22). Corrected: in every
 * memory mode, med/low from the chain kept when the cell was found). */
static void backtrace_zdrop_end(wfo_aligner* a, int score) {
    if (a->attr.compat == WFA_COMPAT_CORRECTED && a->piggyback) {
        backtrace_piggyback_bt(a, a->zbest_pc, a->zbest_pv, a->end_score, a->end_k,
                               a->end_offset);
        return;
    }
    const int from = !a->piggyback && a->end_score >= 0 ? a->end_score : score;
    backtrace_from(a, from, a->end_k, a->end_offset);
}

/* ------------------------------------------------------------------ */
/* align_end2end (src/aligner.rs:620-698)                              */
/* ------------------------------------------------------------------ */
static void set_sequences(wfo_aligner* a, const uint8_t* p, int plen, const uint8_t* t,
                          int tlen) {
    size_t need = (size_t)(plen > tlen ? plen : tlen) + 1;
    if (need > a->seq_cap) {
        a->pattern = (uint8_t*)realloc(a->pattern, need);
        a->text = (uint8_t*)realloc(a->text, need);
        a->seq_cap = need;
    }
    if (plen) memcpy(a->pattern, p, (size_t)plen);
    if (tlen) memcpy(a->text, t, (size_t)tlen);
    a->plen = plen;
    a->tlen = tlen;
}

/* maxtrim_gap_affine / _affine2p (src/cigar.rs:272-419) over ops[b, e):
 * the maximal-scoring prefix; returns the new end (b when cleared) and sets
 * *trimmed. */
static int maxtrim(const wfo_aligner* a, const uint8_t* ops, int b, int e, int* trimmed) {
    const wfa_penalties_t* u = &a->pen.user;
    *trimmed = 0;
    if (b >= e) return e;
    int mscore = u->match != 0 ? u->match : -1;
    int max_score = 0, max_off = b, score = 0;
    if (a->pen.metric == WFA_METRIC_GAP_AFFINE_2P) {
        uint8_t last = 0;
        int len = 0;
        for (int i = b; i < e; ++i) {
            if (ops[i] != last && last != 0) {
                int d;
                if (last == 'M') d = mscore * len;
                else if (last == 'X') d = u->mismatch * len;
                else {
                    int s1 = u->gap_opening1 + u->gap_extension1 * len;
                    int s2 = u->gap_opening2 + u->gap_extension2 * len;
                    d = s1 < s2 ? s1 : s2;
                }
                score -= d;
                len = 0;
                if (max_score < score) { max_score = score; max_off = i - 1; }
            }
            last = ops[i];
            len++;
        }
        int d;
        if (last == 'M') d = mscore * len;
        else if (last == 'X') d = u->mismatch * len;
        else {
            int s1 = u->gap_opening1 + u->gap_extension1 * len;
            int s2 = u->gap_opening2 + u->gap_extension2 * len;
            d = s1 < s2 ? s1 : s2;
        }
        score -= d;
        if (max_score < score) { max_score = score; max_off = e - 1; }
    } else {
        uint8_t last = 0;
        for (int i = b; i < e; ++i) {
            uint8_t op = ops[i];
            if (op == 'M') score -= mscore;
            else if (op == 'X') score -= u->mismatch;
            else if (a->pen.metric == WFA_METRIC_GAP_LINEAR) score -= u->gap_opening1;
            else score -= u->gap_extension1 + (last == op ? 0 : u->gap_opening1);
            last = op;
            if (max_score < score) { max_score = score; max_off = i; }
        }
    }
    *trimmed = max_off != e - 1;
    if (max_score == 0) return b; /* cleared */
    return max_off + 1;
}

/* cigar_wfa_score (src/aligner.rs:2700-2719) of ops[0, n): their classic
 * score on the user penalties, in the WFA scale over the bases they span. */
static int wf_score_ops(const wfo_aligner* a, const uint8_t* ops, int n) {
    const wfa_penalties_t* u = &a->pen.user;
    int v = 0, h = 0, classic;
    for (int i = 0; i < n; ++i) { v += ops[i] != 'I'; h += ops[i] != 'D'; }
    switch (a->pen.metric) {
    case WFA_METRIC_INDEL:
    case WFA_METRIC_EDIT: return wfo_cigar_score_edit(ops, n);
    case WFA_METRIC_GAP_LINEAR:
        classic = wfo_cigar_score_gap_linear(ops, n, u->match, u->mismatch, u->gap_opening1);
        break;
    case WFA_METRIC_GAP_AFFINE:
        classic = wfo_cigar_score_gap_affine(ops, n, u->match, u->mismatch, u->gap_opening1,
                                             u->gap_extension1);
        break;
    default:
        classic = wfo_cigar_score_gap_affine2p(ops, n, u->match, u->mismatch, u->gap_opening1,
                                               u->gap_extension1, u->gap_opening2,
                                               u->gap_extension2);
        break;
    }
    return a->pen.match < 0 ? -a->pen.match * (v + h) - 2 * classic : -classic;
}

/* maxtrim_cigar (src/aligner.rs:1545-1583): the kept CIGAR's score is its
 * WFA score, a free leading gap excluded. */
static int maxtrim_cigar(wfo_aligner* a) {
    if (a->pen.metric == WFA_METRIC_EDIT || a->pen.metric == WFA_METRIC_INDEL) return 0;
    int begin_v = 0, begin_h = 0;
    if (a->endsfree) {
        while (begin_h < a->tb_free && a->cig_begin < a->cig_end && a->cig[a->cig_begin] == 'I') {
            a->cig_begin++; begin_h++;
        }
        if (begin_h == 0)
            while (begin_v < a->pb_free && a->cig_begin < a->cig_end &&
                   a->cig[a->cig_begin] == 'D') {
                a->cig_begin++; begin_v++;
            }
    }
    if (a->cig_begin >= a->cig_end) {
        a->cig_begin = a->cig_end = 0;
        a->cig_score = INT32_MIN;
        return 1;
    }
    int trimmed;
    int ne = maxtrim(a, a->cig, a->cig_begin, a->cig_end, &trimmed);
    if (ne == a->cig_begin) {
        /* Cigar::clear */
        a->cig_begin = a->cig_end = 0;
        a->cig_score = INT32_MIN;
        return trimmed;
    }
    a->cig_end = ne;
    a->cig_score = wf_score_ops(a, a->cig + a->cig_begin, ne - a->cig_begin);
    a->cig_begin -= begin_v + begin_h;
    return trimmed;
}

/* rescore_heuristic_cigar (src/aligner.rs:1605-1640) with cigar_wfa_score
 * (src/aligner.rs:2700-2719): with a heuristic, the gap-affine/2p CIGAR's
 * own score (free end gaps excluded) becomes the returned score. */
static int rescore_heuristic_cigar(wfo_aligner* a, int score, int end_v, int end_h) {
    if (a->attr.heuristic == WFA_HEURISTIC_NONE ||
        !(a->pen.metric == WFA_METRIC_GAP_AFFINE || a->pen.metric == WFA_METRIC_GAP_AFFINE_2P))
        return score;
    int b = a->cig_begin, e = a->cig_end;
    if (a->endsfree) {
        int trailing = (a->plen - end_v) + (a->tlen - end_h);
        /* end_offset.saturating_sub(trailing as usize): a negative trailing
         * count (an end beyond the sequences, after free begin gaps longer
         * than them) wraps to a huge one and leaves nothing */
        e = trailing < 0 || e - trailing < 0 ? 0 : e - trailing;
        uint8_t op = b < a->cig_end ? a->cig[b] : 0;
        int fr = op == 'I' ? a->tb_free : op == 'D' ? a->pb_free : 0;
        int len = 0;
        while (len < fr && b < e && a->cig[b] == op) { b++; len++; }
    }
    if (e < b) e = b;
    if (e == b) { a->cig_score = 0; return 0; }
    const wfa_penalties_t* u = &a->pen.user;
    int v = 0, h = 0;
    for (int i = b; i < e; ++i) { v += a->cig[i] != 'I'; h += a->cig[i] != 'D'; }
    int classic = a->pen.metric == WFA_METRIC_GAP_AFFINE
                      ? wfo_cigar_score_gap_affine(a->cig + b, e - b, u->match, u->mismatch,
                                                   u->gap_opening1, u->gap_extension1)
                      : wfo_cigar_score_gap_affine2p(a->cig + b, e - b, u->match, u->mismatch,
                                                     u->gap_opening1, u->gap_extension1,
                                                     u->gap_opening2, u->gap_extension2);
    int r = a->pen.match < 0 ? -a->pen.match * (v + h) - 2 * classic : -classic;
    a->cig_score = r;
    return r;
}

int32_t wfo_align_endsfree(wfo_aligner* a, const uint8_t* pattern, int32_t plen,
                           const uint8_t* text, int32_t tlen) {
    set_sequences(a, pattern, plen, text, tlen);
    a->endsfree = 1;
    if (a->attr.extension) {
        a->pb_free = 0; a->pe_free = plen; a->tb_free = 0; a->te_free = tlen;
    } else {
        a->pb_free = a->attr.pattern_begin_free; a->pe_free = a->attr.pattern_end_free;
        a->tb_free = a->attr.text_begin_free; a->te_free = a->attr.text_end_free;
    }
    init_alignment(a);
    int score = 0;
    for (;;) {
        if (extend_endsfree(a, score)) break;
        score += 1;
        if (!has_slot(a, score)) { a->status = WFA_STATUS_MAX_STEPS_REACHED; break; }
        compute_step(a, score);
        if (score >= a->attr.max_alignment_steps) {
            a->status = WFA_STATUS_MAX_STEPS_REACHED;
            break;
        }
    }
    int result = score;
    if (a->status == WFA_STATUS_END_REACHED || a->status == WFA_STATUS_END_UNREACHABLE) {
        int unreachable = a->status == WFA_STATUS_END_UNREACHABLE;
        /* Corrected: the end candidate's cell can be from an earlier score,
         * and its total is the alignment's score (src/aligner.rs:769-776) */
        const int candidate = a->cand_valid && !unreachable;
        if (candidate) result = a->cand_total;
        if (a->attr.alignment_scope == WFA_SCOPE_ALIGNMENT &&
            (a->status == WFA_STATUS_END_REACHED || a->end_offset != OFFSET_NULL)) {
            if (candidate) {
                if (a->piggyback)
                    backtrace_piggyback_bt(a, a->cand_pc, a->cand_pv, a->cand_score, a->cand_k,
                                           a->cand_off);
                else
                    backtrace_from(a, a->cand_score, a->cand_k, a->cand_off);
                a->cig_score = a->cand_total;
            } else if (unreachable) {
                backtrace_zdrop_end(a, score);
            } else {
                backtrace_from(a, score, a->end_k, a->end_offset);
            }
            if (a->attr.extension || unreachable) {
                int trimmed = maxtrim_cigar(a);
                a->status = (trimmed || unreachable) ? WFA_STATUS_PARTIAL : WFA_STATUS_COMPLETED;
            } else {
                if (!candidate)
                    result = rescore_heuristic_cigar(a, score, a->end_offset - a->end_k,
                                                     a->end_offset);
                a->status = WFA_STATUS_COMPLETED;
            }
        } else {
            a->status = unreachable ? WFA_STATUS_PARTIAL : WFA_STATUS_COMPLETED;
        }
    }
    return result;
}

int32_t wfo_align_end2end(wfo_aligner* a, const uint8_t* pattern, int32_t plen,
                          const uint8_t* text, int32_t tlen) {
    set_sequences(a, pattern, plen, text, tlen);
    a->endsfree = 0;
    a->pb_free = a->pe_free = a->tb_free = a->te_free = 0;
    init_alignment(a);
    int score = 0;
    for (;;) {
        if (extend_end2end(a, score)) break;
        score += 1;
        if (!has_slot(a, score)) { a->status = WFA_STATUS_MAX_STEPS_REACHED; break; }
        compute_step(a, score);
        if (score >= a->attr.max_alignment_steps) {
            a->status = WFA_STATUS_MAX_STEPS_REACHED;
            break;
        }
    }
    int result = score;
    if (a->status == WFA_STATUS_END_REACHED) {
        if (a->attr.alignment_scope == WFA_SCOPE_ALIGNMENT) {
            backtrace_from(a, score, a->end_k, a->end_offset);
            result = rescore_heuristic_cigar(a, score, a->plen, a->tlen);
        }
        a->status = WFA_STATUS_COMPLETED;
    } else if (a->status == WFA_STATUS_END_UNREACHABLE && a->end_offset != OFFSET_NULL) {
        /* Z-dropped: the partial alignment up to the best cell, max-trimmed
         * (src/aligner.rs:683-694) */
        if (a->attr.alignment_scope == WFA_SCOPE_ALIGNMENT) {
            backtrace_zdrop_end(a, score);
            maxtrim_cigar(a);
        }
        a->status = WFA_STATUS_PARTIAL;
    }
    return result;
}

void wfo_end_position(const wfo_aligner* a, int32_t* out) {
    out[0] = a->end_score;
    out[1] = a->end_k;
    out[2] = a->end_offset;
}

/* ------------------------------------------------------------------ */
/* BiWFA (src/aligner.rs:2563-3302, src/bialign.rs:1-520)              */
/* ------------------------------------------------------------------ */
/* BiStatus (src/bialign.rs) */
enum { BI_OK = 0, BI_END_REACHED, BI_UNREACHABLE, BI_MAX_STEPS, BI_UNATTAINABLE };
/* BIALIGN_* thresholds (src/bialign.rs:511-520) */
#define BI_FALLBACK_MIN_SCORE 250
#define BI_FALLBACK_MIN_LENGTH 100
#define BI_RECOVERY_MIN_SCORE 500

/* BiAlignBreakpoint (src/bialign.rs:20-57) */
typedef struct {
    int score, score_forward, score_reverse;
    int k_forward, k_reverse, offset_forward, offset_reverse;
    int component;
} bi_bp_t;

static void bp_init(bi_bp_t* bp) {
    memset(bp, 0, sizeof(*bp));
    bp->score = INT32_MAX;
    bp->component = C_M;
}

/* set_sequences with init_ascii(.., reverse = true) (src/sequences.rs:153-206) */
static void set_sequences_reversed(wfo_aligner* a, const uint8_t* p, int plen, const uint8_t* t,
                                   int tlen) {
    set_sequences(a, p, plen, t, tlen);
    for (int i = 0; i < plen; ++i) a->pattern[i] = p[plen - 1 - i];
    for (int i = 0; i < tlen; ++i) a->text[i] = t[tlen - 1 - i];
}

/* heuristic_band / with_band: the banded heuristics' (min_k, max_k) */
static int has_band(const wfa_attr_t* at) {
    return at->heuristic == WFA_HEURISTIC_BANDED_STATIC ||
           at->heuristic == WFA_HEURISTIC_BANDED_ADAPTIVE;
}

/* new_biwfa_sub (src/aligner.rs:2768-2800) and take_biwfa_subs
 * (:2721-2740): forward/reverse run score-only with modular storage; the
 * base aligner computes CIGARs with full storage. */
static wfo_aligner* bi_sub(wfo_aligner* a, wfo_aligner** slot, int alignment) {
    if (!*slot) {
        wfa_attr_t at = a->attr;
        at.alignment_scope = alignment ? WFA_SCOPE_ALIGNMENT : WFA_SCOPE_SCORE;
        at.memory_mode = WFA_MEMORY_HIGH;
        at.span = WFA_SPAN_END2END;
        at.extension = 0;
        at.max_alignment_steps = INT32_MAX;
        at.heuristic = WFA_HEURISTIC_NONE;
        int32_t err = 0;
        *slot = wfo_aligner_new(&at, &err);
    }
    return *slot;
}
static void bi_set_heuristic(wfo_aligner* sub, const wfa_attr_t* top, int min_k, int max_k) {
    sub->attr.heuristic = top->heuristic;
    sub->attr.heuristic_p[0] = top->heuristic_p[0];
    sub->attr.heuristic_p[1] = top->heuristic_p[1];
    sub->attr.heuristic_p[2] = top->heuristic_p[2];
    if (has_band(top)) {
        sub->attr.heuristic_p[0] = min_k;
        sub->attr.heuristic_p[1] = max_k;
    }
}

/* check_component_termination (src/aligner.rs:3269-3302) */
static int check_component_termination(wfo_aligner* a, int score, int comp) {
    const wf_t* w = get(a, comp, score);
    if (!w) return 0;
    int k_end = a->tlen - a->plen;
    if (k_end >= w->lo && k_end <= w->hi && wf_get(w, k_end) >= a->tlen) {
        a->end_score = score; a->end_k = k_end; a->end_offset = a->tlen;
        a->status = WFA_STATUS_END_REACHED;
        return 1;
    }
    return 0;
}

/* extend_biwfa_step (src/aligner.rs:3516-3554) with
 * extend_matches_packed_end2end[_max] (src/extend.rs:47-106): extend M,
 * then the end check on `comp_end`, then the heuristic. *max_ak is the
 * largest antidiagonal 2*offset - k of an extended cell (0 if none). */
static int extend_biwfa_step(wfo_aligner* a, int score, int comp_end, int* max_ak) {
    *max_ak = 0;
    wf_t* m = get(a, C_M, score);
    if (!m) {
        if (a->num_null_steps > a->scope) {
            a->status = WFA_STATUS_END_UNREACHABLE;
            return 1;
        }
        return 0;
    }
    int mx = 0;
    for (int k = m->lo; k <= m->hi; ++k) {
        int32_t off = wf_get(m, k);
        if (off >= 0) {
            off += lcp(a, off - k, off);
            wf_set(m, k, off);
            if (2 * off - k > mx) mx = 2 * off - k;
        }
    }
    if (check_component_termination(a, score, comp_end)) return 1;
    if (a->attr.heuristic != WFA_HEURISTIC_NONE) apply_heuristic_cutoff(a, score, m);
    *max_ak = mx;
    return 0;
}

static inline const wf_t* valid_wf(const wf_t* w) { return w && !w->null ? w : NULL; }
static inline int k_inverse(int k, int plen, int tlen) { return (tlen - plen) - k; }

/* breakpoint_m2m / breakpoint_indel2indel (src/bialign.rs:79-227) with
 * scan_overlap (:244-295): the first diagonal, ascending in wf_0's k, where
 * both directions' text positions add up to the text length. An indel
 * breakpoint also needs the forward cell inside the sequences. */
static void bi_breakpoint_scan(const wf_t* w0, const wf_t* w1, int plen, int tlen, int s0,
                               int s1, int gap_open, int comp, int fwd, bi_bp_t* bp) {
    if (s0 + s1 - gap_open >= bp->score) return;
    int lo1 = k_inverse(w1->hi, plen, tlen), hi1 = k_inverse(w1->lo, plen, tlen);
    if (hi1 < w0->lo || w0->hi < lo1) return;
    int lo = w0->lo > lo1 ? w0->lo : lo1, hi = w0->hi < hi1 ? w0->hi : hi1;
    for (int k0 = lo; k0 <= hi; ++k0) {
        int k1 = k_inverse(k0, plen, tlen);
        int32_t o0 = wf_get(w0, k0), o1 = wf_get(w1, k1);
        if (o0 + o1 < tlen) continue;
        if (comp != C_M) {
            int ck = fwd ? k0 : k1, ch = fwd ? o0 : o1;
            if (ch - ck > plen || ch > tlen) continue;
        }
        if (fwd) {
            bp->score_forward = s0; bp->score_reverse = s1;
            bp->k_forward = k0; bp->k_reverse = k1;
            bp->offset_forward = o0; bp->offset_reverse = o1;
        } else {
            bp->score_forward = s1; bp->score_reverse = s0;
            bp->k_forward = k1; bp->k_reverse = k0;
            bp->offset_forward = o1; bp->offset_reverse = o0;
        }
        bp->score = s0 + s1 - gap_open;
        bp->component = comp;
        return;
    }
}

/* bialign_overlap (src/bialign.rs:362-508): a0 at s0 against a1 at the last
 * `scope` scores from s1 down: D2, I2, D1, I1, then M. */
static void bi_overlap(const wfo_aligner* a0, const wfo_aligner* a1, int s0, int s1, int plen,
                       int tlen, int fwd, bi_bp_t* bp) {
    const pen_t* p = &a0->pen;
    const int affine = p->metric == WFA_METRIC_GAP_AFFINE || p->metric == WFA_METRIC_GAP_AFFINE_2P;
    const int two = p->metric == WFA_METRIC_GAP_AFFINE_2P;
    const wf_t* m0 = valid_wf(get(a0, C_M, s0));
    if (!m0) return;
    const wf_t* d1 = affine ? valid_wf(get(a0, C_D1, s0)) : NULL;
    const wf_t* i1 = affine ? valid_wf(get(a0, C_I1, s0)) : NULL;
    const wf_t* d2 = two ? valid_wf(get(a0, C_D2, s0)) : NULL;
    const wf_t* i2 = two ? valid_wf(get(a0, C_I2, s0)) : NULL;
    for (int i = 0; i < a1->scope; ++i) {
        int si = s1 - i;
        if (si < 0) break;
        if (two && s0 + si - p->o2 < bp->score) {
            const wf_t* w;
            if (d2 && (w = valid_wf(get(a1, C_D2, si))))
                bi_breakpoint_scan(d2, w, plen, tlen, s0, si, p->o2, C_D2, fwd, bp);
            if (i2 && (w = valid_wf(get(a1, C_I2, si))))
                bi_breakpoint_scan(i2, w, plen, tlen, s0, si, p->o2, C_I2, fwd, bp);
        }
        if (affine && s0 + si - p->o1 < bp->score) {
            const wf_t* w;
            if (d1 && (w = valid_wf(get(a1, C_D1, si))))
                bi_breakpoint_scan(d1, w, plen, tlen, s0, si, p->o1, C_D1, fwd, bp);
            if (i1 && (w = valid_wf(get(a1, C_I1, si))))
                bi_breakpoint_scan(i1, w, plen, tlen, s0, si, p->o1, C_I1, fwd, bp);
        }
        if (s0 + si >= bp->score) continue;
        const wf_t* w = valid_wf(get(a1, C_M, si));
        if (w) bi_breakpoint_scan(m0, w, plen, tlen, s0, si, 0, C_M, fwd, bp);
    }
}

/* bialign_find_breakpoint (src/aligner.rs:2958-3120). Returns a BI_*
 * status; *end_score is the end score of BI_END_REACHED. */
static int bi_find_breakpoint(const pen_t* pen, int max_steps, wfo_aligner* fwd,
                              wfo_aligner* rev, const uint8_t* P, int plen, const uint8_t* T,
                              int tlen, int cbeg, int cend, bi_bp_t* bp, int* end_score) {
    set_sequences(fwd, P, plen, T, tlen);
    init_alignment_component(fwd, cbeg);
    set_sequences_reversed(rev, P, plen, T, tlen);
    init_alignment_component(rev, cend);
    const int fwd_end = cend, rev_end = cbeg;
#define QUIT(A) \
    do { \
        if ((A)->status == WFA_STATUS_END_REACHED) { *end_score = (A)->end_score; return BI_END_REACHED; } \
        return BI_UNREACHABLE; \
    } while (0)
    int gap_open = pen->metric == WFA_METRIC_GAP_AFFINE ? pen->o1
                   : pen->metric == WFA_METRIC_GAP_AFFINE_2P ? (pen->o1 > pen->o2 ? pen->o1 : pen->o2)
                                                             : 0;
    const int max_ad = plen + tlen - 1;
    int s_fwd = 0, s_rev = 0, last_fwd = 0, ak;
    int fwd_max, rev_max;
    if (extend_biwfa_step(fwd, 0, fwd_end, &fwd_max)) QUIT(fwd);
    if (extend_biwfa_step(rev, 0, rev_end, &rev_max)) QUIT(rev);
    /* phase 1: alternate until the antidiagonals meet */
    while (fwd_max + rev_max < max_ad) {
        s_fwd += 1;
        compute_step(fwd, s_fwd);
        int done = extend_biwfa_step(fwd, s_fwd, fwd_end, &ak);
        if (ak > fwd_max) fwd_max = ak;
        last_fwd = 1;
        if (done) QUIT(fwd);
        if (fwd_max + rev_max >= max_ad) break;
        s_rev += 1;
        compute_step(rev, s_rev);
        done = extend_biwfa_step(rev, s_rev, rev_end, &ak);
        if (ak > rev_max) rev_max = ak;
        last_fwd = 0;
        if (done) QUIT(rev);
        if (s_fwd + s_rev >= max_steps) return BI_MAX_STEPS;
    }
    /* phase 2: overlap checks, then one step of the other direction */
    const int scope = fwd->scope;
    for (;;) {
        if (last_fwd) {
            int min_rev = s_rev > scope - 1 ? s_rev - (scope - 1) : 0;
            if (s_fwd + min_rev - gap_open >= bp->score) break;
            bi_overlap(fwd, rev, s_fwd, s_rev, plen, tlen, 1, bp);
            s_rev += 1;
            compute_step(rev, s_rev);
            if (extend_biwfa_step(rev, s_rev, rev_end, &ak)) QUIT(rev);
        }
        int min_fwd = s_fwd > scope - 1 ? s_fwd - (scope - 1) : 0;
        if (min_fwd + s_rev - gap_open >= bp->score) break;
        bi_overlap(rev, fwd, s_rev, s_fwd, plen, tlen, 0, bp);
        s_fwd += 1;
        compute_step(fwd, s_fwd);
        if (extend_biwfa_step(fwd, s_fwd, fwd_end, &ak)) QUIT(fwd);
        if (s_fwd + s_rev >= max_steps) return BI_MAX_STEPS;
        last_fwd = 1;
    }
#undef QUIT
    return BI_OK;
}

/* bt_* extend lookups (src/backtrace.rs:260-449): a gap component's cell
 * continued by one more gap op; M opens read M (bt_ins1_open etc.). */
#define BT_I1_EXT 2
#define BT_I2_OPEN 3
#define BT_I2_EXT 4
#define BT_D1_EXT 6
#define BT_D2_OPEN 7
#define BT_D2_EXT 8
static int64_t bt_ins(const wfo_aligner* a, int comp, int score, int k, int64_t type) {
    if (score < 0) return OFFSET_NULL;
    const wf_t* w = get(a, comp, score);
    if (!w) return OFFSET_NULL;
    if (w->init_min < k && k - 1 <= w->init_max) return bt_pack(wf_get(w, k - 1) + 1, type);
    return OFFSET_NULL;
}
static int64_t bt_del(const wfo_aligner* a, int comp, int score, int k, int64_t type) {
    if (score < 0) return OFFSET_NULL;
    const wf_t* w = get(a, comp, score);
    if (!w) return OFFSET_NULL;
    if (w->init_min <= k + 1 && k < w->init_max) return bt_pack(wf_get(w, k + 1), type);
    return OFFSET_NULL;
}
static inline int64_t max64(int64_t x, int64_t y) { return x > y ? x : y; }

/* backtrace_affine_c (src/backtrace.rs:492-635): gap-affine / 2-piece
 * CIGAR between components, every predecessor read from its source
 * wavefront, the larger packed (offset << 4 | type) winning. */
static void backtrace_affine_c(wfo_aligner* a, int score, int k, int offset, int comp_end) {
    const pen_t* p = &a->pen;
    const int two = p->metric == WFA_METRIC_GAP_AFFINE_2P;
    a->cig_end = a->cig_cap;
    a->cig_begin = a->cig_cap - 1;
    const int alignment_score = score;
    int matrix = comp_end;
    int v = offset - k, h = offset;
    if (comp_end == C_M) {
        if (a->plen - v > 0) write_ops(a, 'D', a->plen - v);
        if (a->tlen - h > 0) write_ops(a, 'I', a->tlen - h);
    }
    while (v > 0 && h > 0 && score > 0) {
        const int mis = score - p->mismatch;
        const int go1 = score - p->o1 - p->e1, go2 = score - p->o2 - p->e2;
        const int ge1 = score - p->e1, ge2 = score - p->e2;
        int64_t mx;
        switch (matrix) {
        case C_M: {
            int64_t misms = bt_misms(a, mis, k);
            int64_t ins1 = max64(bt_ins(a, C_M, go1, k, BT_I1_OPEN), bt_ins(a, C_I1, ge1, k, BT_I1_EXT));
            int64_t del1 = max64(bt_del(a, C_M, go1, k, BT_D1_OPEN), bt_del(a, C_D1, ge1, k, BT_D1_EXT));
            if (two) {
                int64_t ins2 = max64(bt_ins(a, C_M, go2, k, BT_I2_OPEN), bt_ins(a, C_I2, ge2, k, BT_I2_EXT));
                int64_t del2 = max64(bt_del(a, C_M, go2, k, BT_D2_OPEN), bt_del(a, C_D2, ge2, k, BT_D2_EXT));
                mx = max64(misms, max64(max64(ins1, ins2), max64(del1, del2)));
            } else {
                mx = max64(misms, max64(ins1, del1));
            }
            break;
        }
        case C_I1: mx = max64(bt_ins(a, C_M, go1, k, BT_I1_OPEN), bt_ins(a, C_I1, ge1, k, BT_I1_EXT)); break;
        case C_D1: mx = max64(bt_del(a, C_M, go1, k, BT_D1_OPEN), bt_del(a, C_D1, ge1, k, BT_D1_EXT)); break;
        case C_I2: mx = max64(bt_ins(a, C_M, go2, k, BT_I2_OPEN), bt_ins(a, C_I2, ge2, k, BT_I2_EXT)); break;
        default: mx = max64(bt_del(a, C_M, go2, k, BT_D2_OPEN), bt_del(a, C_D2, ge2, k, BT_D2_EXT)); break;
        }
        if (mx < 0) break;
        if (matrix == C_M) {
            int32_t moff = (int32_t)(mx >> 4);
            write_matches(a, offset - moff);
            offset = moff;
            v = offset - k; h = offset;
            if (v <= 0 || h <= 0) break;
        }
        int64_t t = mx & 0xF;
        switch (t) {
        case BT_M: score = mis; matrix = C_M; break;
        case BT_I1_OPEN: case BT_D1_OPEN: score = go1; matrix = C_M; break;
        case BT_I2_OPEN: case BT_D2_OPEN: score = go2; matrix = C_M; break;
        case BT_I1_EXT: score = ge1; matrix = C_I1; break;
        case BT_I2_EXT: score = ge2; matrix = C_I2; break;
        case BT_D1_EXT: score = ge1; matrix = C_D1; break;
        case BT_D2_EXT: score = ge2; matrix = C_D2; break;
        default: abort();
        }
        if (t == BT_M) { write_ops(a, 'X', 1); offset -= 1; }
        else if (t == BT_I1_OPEN || t == BT_I1_EXT || t == BT_I2_OPEN || t == BT_I2_EXT) {
            write_ops(a, 'I', 1); k -= 1; offset -= 1;
        } else { write_ops(a, 'D', 1); k += 1; }
        v = offset - k; h = offset;
    }
    if (matrix == C_M) {
        if (v > 0 && h > 0) {
            int nm = v < h ? v : h;
            write_matches(a, nm);
            v -= nm; h -= nm;
        }
        if (v > 0) write_ops(a, 'D', v);
        if (h > 0) write_ops(a, 'I', h);
    }
    a->cig_begin += 1;
    a->cig_score = alignment_score;
}

/* backtrace_components (src/aligner.rs:1489-1536) */
static void backtrace_components(wfo_aligner* a, int score, int k, int offset, int comp_end) {
    if (a->pen.metric == WFA_METRIC_GAP_AFFINE || a->pen.metric == WFA_METRIC_GAP_AFFINE_2P)
        backtrace_affine_c(a, score, k, offset, comp_end);
    else
        backtrace_linear(a, score, k, offset);
}

/* Cigar append (src/cigar.rs:80-111) into the top aligner's forward CIGAR */
static void cig_append(wfo_aligner* a, const uint8_t* ops, int n) {
    if (n > 0) memcpy(a->cig + a->cig_end, ops, (size_t)n);
    if (n > 0) a->cig_end += n;
}
static void cig_append_run(wfo_aligner* a, uint8_t op, int n) {
    for (int i = 0; i < n; ++i) a->cig[a->cig_end++] = op;
}

/* bialign_base (src/aligner.rs:3127-3202): unidirectional WFA from
 * component cbeg to cend on the sub-problem, its CIGAR appended whether or
 * not it completed. *score_out is the base score on success. */
static int bi_base(wfo_aligner* top, const uint8_t* P, const uint8_t* T, int pb, int pe, int tb,
                   int te, int cbeg, int cend, int* score_out) {
    wfo_aligner* b = bi_sub(top, &top->bi_base, 1);
    bi_set_heuristic(b, &top->attr, top->attr.heuristic_p[0] - (tb - pb),
                     top->attr.heuristic_p[1] - (tb - pb));
    set_sequences(b, P + pb, pe - pb, T + tb, te - tb);
    b->endsfree = 0;
    b->pb_free = b->pe_free = b->tb_free = b->te_free = 0;
    init_alignment_component(b, cbeg);
    int score = 0, ak;
    for (;;) {
        if (extend_biwfa_step(b, score, cend, &ak)) break;
        score += 1;
        if (!has_slot(b, score)) { b->status = WFA_STATUS_MAX_STEPS_REACHED; break; }
        compute_step(b, score);
        if (score >= b->attr.max_alignment_steps) { b->status = WFA_STATUS_MAX_STEPS_REACHED; break; }
    }
    int r;
    if (b->status == WFA_STATUS_END_REACHED) {
        backtrace_components(b, score, b->end_k, b->end_offset, cend);
        *score_out = score;
        r = BI_OK;
    } else if (b->status == WFA_STATUS_END_UNREACHABLE && b->end_offset != OFFSET_NULL) {
        int from = b->end_score >= 0 ? b->end_score : score;
        backtrace_components(b, from, b->end_k, b->end_offset, cend);
        maxtrim_cigar(b);
        r = BI_UNATTAINABLE;
    } else {
        b->cig_begin = b->cig_end = 0;
        r = BI_UNATTAINABLE;
    }
    cig_append(top, b->cig + b->cig_begin, b->cig_end - b->cig_begin);
    return r;
}

/* bialign_recursive (src/aligner.rs:2809-2950) */
static int bi_recursive(wfo_aligner* top, const uint8_t* P, const uint8_t* T, int pb, int pe,
                        int tb, int te, int cbeg, int cend, int score_remaining, int* score_out) {
    const int plen = pe - pb, tlen = te - tb;
    *score_out = 0;
    if (tlen == 0) { cig_append_run(top, 'D', plen); return BI_OK; }
    if (plen == 0) { cig_append_run(top, 'I', tlen); return BI_OK; }
    if (score_remaining <= BI_FALLBACK_MIN_SCORE)
        return bi_base(top, P, T, pb, pe, tb, te, cbeg, cend, score_out);
    wfo_aligner* fwd = bi_sub(top, &top->bi_fwd, 0);
    wfo_aligner* rev = bi_sub(top, &top->bi_rev, 0);
    /* wavefront_bialign_set_subsidiary_band: the band in each direction's
     * own diagonals */
    const int shift = tb - pb, diag = tlen - plen;
    const int fmin = top->attr.heuristic_p[0] - shift, fmax = top->attr.heuristic_p[1] - shift;
    bi_set_heuristic(fwd, &top->attr, fmin, fmax);
    bi_set_heuristic(rev, &top->attr, diag - fmax, diag - fmin);
    bi_bp_t bp;
    bp_init(&bp);
    int end_score = 0;
    int st = bi_find_breakpoint(&top->pen, top->attr.max_alignment_steps, fwd, rev, P + pb, plen,
                                T + tb, tlen, cbeg, cend, &bp, &end_score);
    if (st == BI_END_REACHED) {
        if (end_score <= BI_RECOVERY_MIN_SCORE)
            return bi_base(top, P, T, pb, pe, tb, te, cbeg, cend, score_out);
        return BI_UNREACHABLE;
    }
    if (st != BI_OK) return st;
    const int bp_h = bp.offset_forward, bp_v = bp.offset_forward - bp.k_forward;
    int s;
    st = bi_recursive(top, P, T, pb, pb + bp_v, tb, tb + bp_h, cbeg, bp.component,
                      bp.score_forward, &s);
    if (st != BI_OK) return st;
    st = bi_recursive(top, P, T, pb + bp_v, pe, tb + bp_h, te, bp.component, cend,
                      bp.score_reverse, &s);
    if (st != BI_OK) return st;
    *score_out = bp.score;
    return BI_OK;
}

/* cigar_wfa_score (src/aligner.rs:2700-2719) of the forward CIGAR */
static int cigar_wfa_score_all(const wfo_aligner* a, int plen, int tlen) {
    const wfa_penalties_t* u = &a->pen.user;
    const uint8_t* ops = a->cig + a->cig_begin;
    int n = a->cig_end - a->cig_begin;
    int classic;
    switch (a->pen.metric) {
    case WFA_METRIC_INDEL:
    case WFA_METRIC_EDIT: return wfo_cigar_score_edit(ops, n);
    case WFA_METRIC_GAP_LINEAR:
        classic = wfo_cigar_score_gap_linear(ops, n, u->match, u->mismatch, u->gap_opening1);
        break;
    case WFA_METRIC_GAP_AFFINE:
        classic = wfo_cigar_score_gap_affine(ops, n, u->match, u->mismatch, u->gap_opening1,
                                             u->gap_extension1);
        break;
    default:
        classic = wfo_cigar_score_gap_affine2p(ops, n, u->match, u->mismatch, u->gap_opening1,
                                               u->gap_extension1, u->gap_opening2,
                                               u->gap_extension2);
        break;
    }
    return a->pen.match < 0 ? -a->pen.match * (plen + tlen) - 2 * classic : -classic;
}

/* align_biwfa (src/aligner.rs:2563-2642) */
int32_t wfo_align_biwfa(wfo_aligner* a, const uint8_t* pattern, int32_t plen,
                        const uint8_t* text, int32_t tlen) {
    set_sequences(a, pattern, plen, text, tlen);
    int need = 2 * (plen + tlen) + 1;
    if (need > a->cig_cap) {
        a->cig = (uint8_t*)realloc(a->cig, (size_t)need);
        a->cig_cap = need;
    }
    a->cig_begin = a->cig_end = 0;
    a->cig_score = INT32_MIN;
    a->status = WFA_STATUS_OK;
    int score = 0;
    int st = bi_recursive(a, pattern, text, 0, plen, 0, tlen, C_M, C_M,
                          (plen > tlen ? plen : tlen) <= BI_FALLBACK_MIN_LENGTH ? 0 : INT32_MAX,
                          &score);
    if (st == BI_OK) {
        int heur = a->attr.heuristic != WFA_HEURISTIC_NONE && a->cig_begin < a->cig_end;
        if (plen == 0 || tlen == 0 || heur) score = cigar_wfa_score_all(a, plen, tlen);
        a->cig_score = score;
        a->status = WFA_STATUS_COMPLETED;
        return score;
    }
    a->status = st == BI_MAX_STEPS ? WFA_STATUS_MAX_STEPS_REACHED : WFA_STATUS_END_UNREACHABLE;
    return 0;
}

/* align_biwfa_score (src/aligner.rs:2650-2690) */
int32_t wfo_align_biwfa_score(wfo_aligner* a, const uint8_t* pattern, int32_t plen,
                              const uint8_t* text, int32_t tlen) {
    set_sequences(a, pattern, plen, text, tlen);
    a->cig_begin = a->cig_end = 0;
    a->cig_score = INT32_MIN;
    wfo_aligner* fwd = bi_sub(a, &a->bi_fwd, 0);
    wfo_aligner* rev = bi_sub(a, &a->bi_rev, 0);
    const int diag = tlen - plen;
    bi_set_heuristic(fwd, &a->attr, a->attr.heuristic_p[0], a->attr.heuristic_p[1]);
    bi_set_heuristic(rev, &a->attr, diag - a->attr.heuristic_p[1], diag - a->attr.heuristic_p[0]);
    bi_bp_t bp;
    bp_init(&bp);
    int end_score = 0;
    int st = bi_find_breakpoint(&a->pen, a->attr.max_alignment_steps, fwd, rev, pattern, plen,
                                text, tlen, C_M, C_M, &bp, &end_score);
    int score = st == BI_OK ? bp.score : st == BI_END_REACHED ? end_score : 0;
    a->status = (st == BI_OK || st == BI_END_REACHED) ? WFA_STATUS_COMPLETED
                : st == BI_MAX_STEPS ? WFA_STATUS_MAX_STEPS_REACHED
                                     : WFA_STATUS_END_UNREACHABLE;
    return score;
}

/* ------------------------------------------------------------------ */
/* CIGAR scores (src/cigar.rs:118-202) and validity (:446-518)          */
/* ------------------------------------------------------------------ */
int32_t wfo_cigar_score_edit(const uint8_t* ops, int32_t n) {
    int32_t s = 0;
    for (int i = 0; i < n; ++i) s += ops[i] == 'M' ? 0 : 1;
    return s;
}
int32_t wfo_cigar_score_gap_linear(const uint8_t* ops, int32_t n, int32_t match,
                                   int32_t mismatch, int32_t indel) {
    int32_t s = 0;
    for (int i = 0; i < n; ++i)
        s += ops[i] == 'M' ? -match : ops[i] == 'X' ? -mismatch : -indel;
    return s;
}
int32_t wfo_cigar_score_gap_affine(const uint8_t* ops, int32_t n, int32_t match,
                                   int32_t mismatch, int32_t o, int32_t e) {
    int32_t s = 0;
    uint8_t last = 0;
    for (int i = 0; i < n; ++i) {
        uint8_t op = ops[i];
        if (op == 'M') s -= match;
        else if (op == 'X') s -= mismatch;
        else if (op == 'D') s -= e + (last == 'D' ? 0 : o);
        else if (op == 'I') s -= e + (last == 'I' ? 0 : o);
        last = op;
    }
    return s;
}
static int32_t affine2p_op(uint8_t op, int32_t len, int32_t match, int32_t mismatch, int32_t o1,
                           int32_t e1, int32_t o2, int32_t e2) {
    if (op == 'M') return match * len;
    if (op == 'X') return mismatch * len;
    int32_t s1 = o1 + e1 * len, s2 = o2 + e2 * len;
    return s1 < s2 ? s1 : s2;
}
int32_t wfo_cigar_score_gap_affine2p(const uint8_t* ops, int32_t n, int32_t match,
                                     int32_t mismatch, int32_t o1, int32_t e1, int32_t o2,
                                     int32_t e2) {
    int32_t s = 0, len = 0;
    uint8_t last = 0;
    for (int i = 0; i < n; ++i) {
        if (ops[i] != last && last != 0) {
            s -= affine2p_op(last, len, match, mismatch, o1, e1, o2, e2);
            len = 0;
        }
        last = ops[i];
        len++;
    }
    if (last) s -= affine2p_op(last, len, match, mismatch, o1, e1, o2, e2);
    return s;
}
int32_t wfo_cigar_check(const uint8_t* ops, int32_t n, const uint8_t* pattern, int32_t plen,
                        const uint8_t* text, int32_t tlen) {
    int pp = 0, tp = 0;
    for (int i = 0; i < n; ++i) {
        switch (ops[i]) {
        case 'M':
            if (pp >= plen || tp >= tlen || pattern[pp] != text[tp]) return 1;
            pp++; tp++; break;
        case 'X':
            if (pp >= plen || tp >= tlen || pattern[pp] == text[tp]) return 2;
            pp++; tp++; break;
        case 'I': tp++; break;
        case 'D': pp++; break;
        default: return 3;
        }
    }
    if (pp != plen) return 4;
    if (tp != tlen) return 5;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Batch driver (pthreads, contiguous shards, one aligner per thread)   */
/* ------------------------------------------------------------------ */
typedef struct {
    const wfa_attr_t* attr;
    int32_t fmt;
    const uint8_t* seq;
    const int64_t *poff, *toff;
    const int32_t *plen, *tlen;
    int32_t *score, *status;
    uint8_t* ops;
    const int64_t* ops_off;
    int32_t* ops_len;
    int32_t begin, end;
    int32_t err;
} job_t;

/* decode_packed2bits (src/sequences.rs:34-51) */
static void unpack2(const uint8_t* src, int len, uint8_t* dst) {
    static const uint8_t dna[4] = {'A', 'C', 'G', 'T'};
    for (int i = 0; i < len; ++i) dst[i] = dna[(src[i >> 2] >> (2 * (i & 3))) & 3];
}

static void* job_run(void* arg) {
    job_t* j = (job_t*)arg;
    int32_t err = 0;
    wfo_aligner* a = wfo_aligner_new(j->attr, &err);
    if (!a) { j->err = err; return NULL; }
    uint8_t *pb = NULL, *tb = NULL;
    size_t cap = 0;
    for (int32_t i = j->begin; i < j->end; ++i) {
        int pl = j->plen[i], tl = j->tlen[i];
        const uint8_t *P = j->seq + j->poff[i], *T = j->seq + j->toff[i];
        if (j->fmt == WFA_SEQ_PACKED2BITS) {
            size_t need = (size_t)(pl > tl ? pl : tl) + 1;
            if (need > cap) {
                pb = (uint8_t*)realloc(pb, need);
                tb = (uint8_t*)realloc(tb, need);
                cap = need;
            }
            unpack2(P, pl, pb);
            unpack2(T, tl, tb);
            P = pb; T = tb;
        }
        if (j->attr->memory_mode == WFA_MEMORY_ULTRALOW)
            j->score[i] = j->attr->alignment_scope == WFA_SCOPE_ALIGNMENT
                              ? wfo_align_biwfa(a, P, pl, T, tl)
                              : wfo_align_biwfa_score(a, P, pl, T, tl);
        else if (j->attr->span == WFA_SPAN_ENDSFREE)
            j->score[i] = wfo_align_endsfree(a, P, pl, T, tl);
        else
            j->score[i] = wfo_align_end2end(a, P, pl, T, tl);
        j->status[i] = a->status;
        if (j->ops && j->ops_len) {
            const uint8_t* o = NULL;
            int n = wfo_cigar(a, &o);
            if (n > pl + tl) n = pl + tl;
            if (n > 0) memcpy(j->ops + j->ops_off[i], o, (size_t)n);
            j->ops_len[i] = n;
        }
    }
    free(pb);
    free(tb);
    wfo_aligner_free(a);
    return NULL;
}

int32_t wfo_align_batch(const wfa_attr_t* attr, int32_t n, int32_t seq_format,
                        const uint8_t* seq, const int64_t* poff, const int32_t* plen,
                        const int64_t* toff, const int32_t* tlen, int32_t* score,
                        int32_t* status, uint8_t* ops, const int64_t* ops_off, int32_t* ops_len,
                        int32_t nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > n) nthreads = n > 0 ? n : 1;
    job_t* jobs = (job_t*)calloc((size_t)nthreads, sizeof(job_t));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    for (int t = 0; t < nthreads; ++t) {
        job_t* j = &jobs[t];
        j->attr = attr; j->fmt = seq_format; j->seq = seq;
        j->poff = poff; j->toff = toff; j->plen = plen; j->tlen = tlen;
        j->score = score; j->status = status;
        j->ops = ops; j->ops_off = ops_off; j->ops_len = ops_len;
        j->begin = (int32_t)((int64_t)n * t / nthreads);
        j->end = (int32_t)((int64_t)n * (t + 1) / nthreads);
    }
    if (nthreads == 1) {
        job_run(&jobs[0]);
    } else {
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, job_run, &jobs[t]);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    }
    int32_t err = 0;
    for (int t = 0; t < nthreads; ++t)
        if (jobs[t].err) err = jobs[t].err;
    free(jobs);
    free(th);
    return err;
}

/* ------------------------------------------------------------------ */
/* Test hooks: the step primitives above on caller arrays, so the      */
/* reference's unit literals (cells, heuristic decisions, breakpoints) */
/* can be checked against them directly. A row is `n` cells for the    */
/* diagonals [base, base + n - 1], with [lo, hi] its range and every   */
/* allocated cell readable.                                            */
/* ------------------------------------------------------------------ */
static wf_t row_view(int32_t* mem, int base, int n, int lo, int hi) {
    wf_t w;
    memset(&w, 0, sizeof(w));
    w.mem = mem;
    w.base = base;
    w.nalloc = n;
    w.lo = lo;
    w.hi = hi;
    w.init_min = base;
    w.init_max = base + n - 1;
    return w;
}

int32_t wfo_t_wf_distance(int32_t off, int32_t k, int32_t plen, int32_t tlen) {
    return wf_distance(off, k, plen, tlen);
}

void wfo_t_wfadaptive_reduce(int32_t* lohi, const int32_t* d, int32_t dbase, int32_t min_d,
                             int32_t thr, int32_t min_k, int32_t max_k) {
    wf_t w = row_view(NULL, 0, 0, lohi[0], lohi[1]);
    wfadaptive_reduce(&w, d, dbase, min_d, thr, min_k, max_k);
    lohi[0] = w.lo;
    lohi[1] = w.hi;
}

void wfo_t_wfadaptive_cutoff(int32_t* st, int32_t* mem, int32_t base, int32_t n, int32_t* lohi,
                             int32_t* d, int32_t plen, int32_t tlen, int32_t min_len,
                             int32_t thr, int32_t steps) {
    wf_t w = row_view(mem, base, n, lohi[0], lohi[1]);
    wfadaptive_cutoff((hstate_t*)st, &w, d, base, plen, tlen, min_len, thr, steps);
    lohi[0] = w.lo;
    lohi[1] = w.hi;
}

void wfo_t_banded_static(int32_t* lohi, int32_t min_k, int32_t max_k) {
    wf_t w = row_view(NULL, 0, 0, lohi[0], lohi[1]);
    banded_static_cutoff(&w, min_k, max_k);
    lohi[0] = w.lo;
    lohi[1] = w.hi;
}

void wfo_t_banded_adaptive(int32_t* st, int32_t* mem, int32_t base, int32_t n, int32_t* lohi,
                           int32_t plen, int32_t tlen, int32_t min_k, int32_t max_k,
                           int32_t steps) {
    wf_t w = row_view(mem, base, n, lohi[0], lohi[1]);
    banded_adaptive_cutoff((hstate_t*)st, &w, plen, tlen, min_k, max_k, steps);
    lohi[0] = w.lo;
    lohi[1] = w.hi;
}

void wfo_t_hstate_reset(int32_t* st, int32_t heuristic, const int32_t* hp) {
    hstate_reset((hstate_t*)st, heuristic, hp);
}

/* drop = {match, endsfree, pattern_begin_free, text_begin_free, scope};
 * out = {max, k, offset} */
void wfo_t_compute_sw_scores(int32_t* mem, int32_t base, int32_t n, int32_t lo, int32_t hi,
                             int32_t* sw, const int32_t* drop, int32_t score, int32_t* out) {
    wf_t w = row_view(mem, base, n, lo, hi);
    drop_t dp = drop_params(drop[0], drop[1], drop[2], drop[3], drop[4]);
    compute_sw_scores(&w, sw, base, &dp, score, &out[0], &out[1], &out[2]);
}

void wfo_t_xdrop_cutoff(int32_t* st, int32_t* mem, int32_t base, int32_t n, int32_t* lohi,
                        int32_t* sw, const int32_t* drop, int32_t score, int32_t xdrop,
                        int32_t steps) {
    wf_t w = row_view(mem, base, n, lohi[0], lohi[1]);
    drop_t dp = drop_params(drop[0], drop[1], drop[2], drop[3], drop[4]);
    xdrop_cutoff((hstate_t*)st, &w, sw, base, &dp, score, xdrop, steps);
    lohi[0] = w.lo;
    lohi[1] = w.hi;
}

int32_t wfo_t_zdrop_cutoff(int32_t* st, int32_t* mem, int32_t base, int32_t n, int32_t* lohi,
                           int32_t* sw, const int32_t* drop, int32_t score, int32_t zdrop,
                           int32_t steps) {
    wf_t w = row_view(mem, base, n, lohi[0], lohi[1]);
    drop_t dp = drop_params(drop[0], drop[1], drop[2], drop[3], drop[4]);
    int r = zdrop_cutoff((hstate_t*)st, &w, sw, base, &dp, score, zdrop, steps);
    lohi[0] = w.lo;
    lohi[1] = w.hi;
    return r;
}

/* rows = {Mx, Mo1, I1e, D1e, Mo2, I2e, D2e} inputs (NULL = absent),
 * then {M, I1, D1, I2, D2} outputs, all over the same n diagonals from
 * `base`; cells [lo, hi] are computed */
void wfo_t_affine_cells(int32_t plen, int32_t tlen, int32_t* const* rows, int32_t base,
                        int32_t n, int32_t lo, int32_t hi, int32_t use2) {
    wf_t v[12];
    wf_t* p[12];
    for (int i = 0; i < 12; ++i) {
        v[i] = row_view(rows[i], base, n, lo, hi);
        p[i] = rows[i] ? &v[i] : NULL;
    }
    affine_kernel(plen, tlen, p[0], p[1], p[2], p[3], p[4], p[5], p[6], use2, p[7], p[8], p[9],
                  p[10], p[11], lo, hi);
}

void wfo_t_edit_cells(int32_t plen, int32_t tlen, int32_t* prev, int32_t* cur, int32_t base,
                      int32_t n, int32_t lo, int32_t hi, int32_t indel) {
    wf_t vp = row_view(prev, base, n, lo, hi), vc = row_view(cur, base, n, lo, hi);
    edit_kernel(plen, tlen, &vp, &vc, lo, hi, indel);
}

/* One breakpoint scan of a forward row (w0) against a reverse row (w1);
 * bp (in/out) = {score, score_forward, score_reverse, k_forward,
 * k_reverse, offset_forward, offset_reverse, component}: pass score =
 * INT32_MAX for none found yet */
void wfo_t_breakpoint_scan(int32_t* m0, int32_t base0, int32_t n0, int32_t lo0, int32_t hi0,
                           int32_t* m1, int32_t base1, int32_t n1, int32_t lo1, int32_t hi1,
                           int32_t plen, int32_t tlen, int32_t s0, int32_t s1, int32_t gap_open,
                           int32_t comp, int32_t fwd, int32_t* bp) {
    wf_t w0 = row_view(m0, base0, n0, lo0, hi0), w1 = row_view(m1, base1, n1, lo1, hi1);
    bi_bp_t b;
    b.score = bp[0]; b.score_forward = bp[1]; b.score_reverse = bp[2];
    b.k_forward = bp[3]; b.k_reverse = bp[4]; b.offset_forward = bp[5];
    b.offset_reverse = bp[6]; b.component = bp[7];
    bi_breakpoint_scan(&w0, &w1, plen, tlen, s0, s1, gap_open, comp, fwd, &b);
    bp[0] = b.score; bp[1] = b.score_forward; bp[2] = b.score_reverse;
    bp[3] = b.k_forward; bp[4] = b.k_reverse; bp[5] = b.offset_forward;
    bp[6] = b.offset_reverse; bp[7] = b.component;
}

/* lohi (in/out) = {lo, hi, null} of the I/D row; src = {lo, hi} of M */
void wfo_t_heuristic_equate(int32_t* dst, const int32_t* src) {
    wf_t d = row_view(NULL, 0, 0, dst[0], dst[1]), m = row_view(NULL, 0, 0, src[0], src[1]);
    d.null = dst[2];
    heuristic_equate(&d, &m);
    dst[0] = d.lo;
    dst[1] = d.hi;
    dst[2] = d.null;
}
