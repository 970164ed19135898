/*
 * synth.c — deterministic synthetic pattern/text pairs for the benchmark
 * and the parity tests.
 *
 * Random numbers come from the reference's own test LCG
 * (tests/singletrack.rs:12-26): state = state * 6364136223846793005 +
 * 1442695040888963407, output = state >> 33. Each pair has its own
 * stream, seeded from (seed, pair index), so pairs can be generated in
 * parallel and any pair regenerated alone.
 *
 * A pair: the pattern is uniform ACGT of length L (fixed, uniform or
 * log-uniform in [len_lo, len_hi]); a per-pair divergence d is uniform in
 * [div_lo, div_hi]. The text walks the pattern; at each base, with
 * probability d an edit event happens. With probability run_frac it is an
 * indel run (insertion or deletion, equally likely, length geometric with
 * mean run_mean); otherwise it is a substitution, insertion or deletion,
 * equally likely, where a substitution is one of the three other bases, an
 * insertion is G random bases followed by the pattern base, and a deletion
 * skips G pattern bases, G geometric >= 1 with mean indel_mean.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint64_t seed;
    int32_t len_mode; /* 0 fixed (len_lo), 1 uniform, 2 log-uniform */
    int32_t len_lo, len_hi;
    double div_lo, div_hi;
    double indel_mean;
    double run_frac;
    double run_mean;
} wfa_synth_cfg;

typedef struct { uint64_t s; } lcg_t;

static inline uint64_t lcg_next(lcg_t* r) {
    r->s = r->s * 6364136223846793005ULL + 1442695040888963407ULL;
    return r->s >> 33;
}
static inline uint64_t lcg_below(lcg_t* r, uint64_t n) { return lcg_next(r) % n; }
static inline double lcg_unif(lcg_t* r) { return (double)lcg_next(r) / 2147483648.0; }
static inline int lcg_geom(lcg_t* r, double mean) {
    double q = mean > 1.0 ? 1.0 - 1.0 / mean : 0.0;
    int n = 1;
    while (lcg_unif(r) < q) ++n;
    return n;
}

static void pair_stream(const wfa_synth_cfg* c, int64_t i, lcg_t* r) {
    r->s = c->seed * 0x9E3779B97F4A7C15ULL + (uint64_t)(i + 1) * 0xD1B54A32D192ED03ULL;
    lcg_next(r);
    lcg_next(r);
}

static int pair_len(const wfa_synth_cfg* c, lcg_t* r) {
    if (c->len_mode == 0) return c->len_lo;
    if (c->len_mode == 1) return c->len_lo + (int)lcg_below(r, (uint64_t)(c->len_hi - c->len_lo + 1));
    double a = log((double)c->len_lo), b = log((double)c->len_hi);
    int L = (int)llround(exp(a + (b - a) * lcg_unif(r)));
    if (L < c->len_lo) L = c->len_lo;
    if (L > c->len_hi) L = c->len_hi;
    return L;
}

/* Generate pair i. When pat/txt are NULL only the lengths are produced.
 * Symbols are 0..3 (A C G T). The pattern symbols are the first L draws
 * of the pair's stream; a copy of the stream replays them while the
 * original continues with the text events, so both passes agree. */
static void gen_pair(const wfa_synth_cfg* c, int64_t i, uint8_t* pat, uint8_t* txt, int32_t* plen,
                     int32_t* tlen) {
    lcg_t r;
    pair_stream(c, i, &r);
    const int L = pair_len(c, &r);
    const double d = c->div_lo + (c->div_hi - c->div_lo) * lcg_unif(&r);
    lcg_t rb = r; /* replays the pattern */
    for (int k = 0; k < L; ++k) {
        uint8_t s = (uint8_t)lcg_below(&r, 4);
        if (pat) pat[k] = s;
    }
    int n = 0;
    for (int k = 0; k < L; ++k) {
        const uint8_t base = (uint8_t)lcg_below(&rb, 4);
        if (lcg_unif(&r) < d) {
            int type; /* 0 sub, 1 ins, 2 del */
            int glen;
            if (c->run_frac > 0 && lcg_unif(&r) < c->run_frac) {
                type = 1 + (int)lcg_below(&r, 2);
                glen = lcg_geom(&r, c->run_mean);
            } else {
                type = (int)lcg_below(&r, 3);
                glen = type ? lcg_geom(&r, c->indel_mean) : 1;
            }
            if (type == 0) {
                uint8_t s = (uint8_t)((base + 1 + lcg_below(&r, 3)) & 3);
                if (txt) txt[n] = s;
                ++n;
            } else if (type == 1) {
                for (int g = 0; g < glen; ++g) {
                    uint8_t s = (uint8_t)lcg_below(&r, 4);
                    if (txt) txt[n] = s;
                    ++n;
                }
                if (txt) txt[n] = base;
                ++n;
            } else {
                /* delete this base and glen-1 more */
                for (int skip = glen - 1; skip > 0 && k + 1 < L; --skip) {
                    ++k;
                    (void)lcg_below(&rb, 4);
                }
            }
        } else {
            if (txt) txt[n] = base;
            ++n;
        }
    }
    *plen = L;
    *tlen = n;
}

typedef struct {
    const wfa_synth_cfg* c;
    int64_t begin, end;
    int32_t *plen, *tlen;
    int32_t fmt; /* 0 ascii, 1 packed */
    uint8_t* out;
    const int64_t *poff, *toff;
} sjob_t;

static void* len_job(void* arg) {
    sjob_t* j = (sjob_t*)arg;
    for (int64_t i = j->begin; i < j->end; ++i) gen_pair(j->c, i, NULL, NULL, &j->plen[i], &j->tlen[i]);
    return NULL;
}

static void emit(uint8_t* dst, const uint8_t* sym, int n, int fmt) {
    static const uint8_t dna[4] = {'A', 'C', 'G', 'T'};
    if (fmt == 0) {
        for (int k = 0; k < n; ++k) dst[k] = dna[sym[k]];
    } else {
        memset(dst, 0, (size_t)(n + 3) / 4);
        for (int k = 0; k < n; ++k) dst[k >> 2] |= (uint8_t)(sym[k] << (2 * (k & 3)));
    }
}

static void* fill_job(void* arg) {
    sjob_t* j = (sjob_t*)arg;
    size_t cap = 0;
    uint8_t *pb = NULL, *tb = NULL;
    for (int64_t i = j->begin; i < j->end; ++i) {
        size_t need = (size_t)(j->plen[i] > j->tlen[i] ? j->plen[i] : j->tlen[i]) + 1;
        if (need > cap) {
            cap = need * 2;
            pb = (uint8_t*)realloc(pb, cap);
            tb = (uint8_t*)realloc(tb, cap);
        }
        int32_t pl, tl;
        gen_pair(j->c, i, pb, tb, &pl, &tl);
        emit(j->out + j->poff[i], pb, pl, j->fmt);
        emit(j->out + j->toff[i], tb, tl, j->fmt);
    }
    free(pb);
    free(tb);
    return NULL;
}

static void run(sjob_t* base, int64_t n, int nthreads, void* (*fn)(void*)) {
    if (nthreads < 1) nthreads = 1;
    if (n < 1024) nthreads = 1;
    sjob_t* jobs = (sjob_t*)calloc((size_t)nthreads, sizeof(sjob_t));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    for (int t = 0; t < nthreads; ++t) {
        jobs[t] = *base;
        jobs[t].begin = n * t / nthreads;
        jobs[t].end = n * (t + 1) / nthreads;
    }
    if (nthreads == 1) {
        fn(&jobs[0]);
    } else {
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, fn, &jobs[t]);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    }
    free(jobs);
    free(th);
}

/* Pass 1: lengths of pairs [0, n). */
void wfa_synth_lengths(const wfa_synth_cfg* c, int64_t n, int32_t* plen, int32_t* tlen,
                       int32_t nthreads) {
    sjob_t b;
    memset(&b, 0, sizeof(b));
    b.c = c;
    b.plen = plen;
    b.tlen = tlen;
    run(&b, n, nthreads, len_job);
}

/* Pass 2: write pair i's pattern at out + poff[i] and text at out + toff[i]
 * (ASCII when fmt == 0, 2-bit packed LSB first when fmt == 1). */
void wfa_synth_fill(const wfa_synth_cfg* c, int64_t n, const int32_t* plen, const int32_t* tlen,
                    int32_t fmt, uint8_t* out, const int64_t* poff, const int64_t* toff,
                    int32_t nthreads) {
    sjob_t b;
    memset(&b, 0, sizeof(b));
    b.c = c;
    b.plen = (int32_t*)plen;
    b.tlen = (int32_t*)tlen;
    b.fmt = fmt;
    b.out = out;
    b.poff = poff;
    b.toff = toff;
    run(&b, n, nthreads, fill_job);
}
