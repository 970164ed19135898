/*
 * wfa_mi355x.h — C ABI of the MI355X wavefront-alignment engine.
 *
 * The reference (HPCBio/wfa2lib-rs) is a pure Rust library with no C ABI
 * (src/lib.rs:1-18). Its public aligner API is `WavefrontAligner<N>`
 * (src/aligner.rs:146-211): construct with penalties, configure with
 * setters, call an `align_*` method, read `status()` and `cigar()`. Every
 * entry point below replaces one of those methods one-to-one (the citation
 * sits beside each declaration), so a thin Rust `-sys` crate, a ctypes
 * module or a cgo stub can bind the engine exactly where the reference's
 * aligner is called today. The batch entry points have no single reference
 * counterpart: they replace the per-pair loop of the reference's benchmark
 * driver (src/bin/align_benchmark.rs:593-627) with one device launch.
 *
 * Conventions: plain pointers and sizes, no C++ or HIP types. Functions
 * that can fail return an int status (0 = WFA_OK, negative = WFA_ERR_*);
 * the reference panics on invalid penalties (src/penalties.rs:157-167,
 * 198-209, 245-262) and this ABI returns WFA_ERR_INVALID_PENALTIES instead.
 * Scores are WFA scores (penalty sums, >= 0), exactly as the reference's
 * align methods return them.
 */
#ifndef WFA_MI355X_H
#define WFA_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Distance metrics (src/penalties.rs:12-24, same order) ---- */
#define WFA_METRIC_INDEL          0
#define WFA_METRIC_EDIT           1
#define WFA_METRIC_GAP_LINEAR     2
#define WFA_METRIC_GAP_AFFINE     3
#define WFA_METRIC_GAP_AFFINE_2P  4

/* ---- Alignment scope (src/aligner.rs:26-31) ---- */
#define WFA_SCOPE_SCORE      0
#define WFA_SCOPE_ALIGNMENT  1

/* ---- Memory modes (src/aligner.rs:44-52; ULTRALOW = BiWFA, :2563) ---- */
#define WFA_MEMORY_HIGH      0
#define WFA_MEMORY_MED       1
#define WFA_MEMORY_LOW       2
#define WFA_MEMORY_ULTRALOW  3

/* ---- Span (src/aligner.rs:109-115) ---- */
#define WFA_SPAN_END2END     0
#define WFA_SPAN_ENDSFREE    1

/* ---- Compat (src/aligner.rs:53-94) ---- */
/* C: results identical to C WFA2-lib, its known bugs included (the
 * default). CORRECTED: ends-free with a match reward returns the optimal
 * alignment (candidate ends charged for the bases their free end gaps skip;
 * the score is the whole alignment's), and a Z-drop end is backtraced from
 * the best cell at the score it was found at in every memory mode. */
#define WFA_COMPAT_C          0
#define WFA_COMPAT_CORRECTED  1

/* ---- Heuristics (src/heuristic.rs, HeuristicStrategy) ---- */
#define WFA_HEURISTIC_NONE             0
#define WFA_HEURISTIC_WFADAPTIVE       1  /* p0 = min_wavefront_length, p1 = max_distance_threshold, p2 = steps_between_cutoffs */
#define WFA_HEURISTIC_XDROP            2  /* p0 = xdrop, p1 = steps_between_cutoffs */
#define WFA_HEURISTIC_ZDROP            3  /* p0 = zdrop, p1 = steps_between_cutoffs;
                                              not with BiWFA (WFA_ERR_UNSUPPORTED).
                                              A medium/low-memory CIGAR in
                                              WFA_COMPAT_C is read at the best
                                              cell's diagonal of the stop score,
                                              as the reference reads it
                                              (HPCBio/wfa2lib-rs#22) */
#define WFA_HEURISTIC_BANDED_STATIC    4  /* p0 = min_k, p1 = max_k */
#define WFA_HEURISTIC_BANDED_ADAPTIVE  5  /* p0 = min_k, p1 = max_k, p2 = steps_between_cutoffs */

/* ---- Alignment status, same ordinals as AlignStatus (src/aligner.rs:120-135) ---- */
#define WFA_STATUS_OK                 0
#define WFA_STATUS_END_REACHED        1
#define WFA_STATUS_END_UNREACHABLE    2
#define WFA_STATUS_MAX_STEPS_REACHED  3
#define WFA_STATUS_OOM                4
#define WFA_STATUS_COMPLETED          5
#define WFA_STATUS_PARTIAL            6

/* ---- Sequence formats of batch inputs ---- */
#define WFA_SEQ_ASCII       0  /* one byte per base, any byte values */
#define WFA_SEQ_PACKED2BITS 1  /* 4 bases per byte, LSB first, A=0 C=1 G=2 T=3
                                  (src/sequences.rs:30-74) */

/* ---- Error codes ---- */
#define WFA_OK                      0
#define WFA_ERR_INVALID_PENALTIES  -1
#define WFA_ERR_INVALID_ARGUMENT   -2
#define WFA_ERR_UNSUPPORTED        -3  /* configuration not available on the device path */
#define WFA_ERR_DEVICE             -4  /* HIP runtime error */
#define WFA_ERR_NO_DEVICE          -5
#define WFA_ERR_OUT_OF_MEMORY      -6
#define WFA_ERR_REJECTED           -7  /* combination the reference rejects
                                          (src/bin/align_benchmark.rs:531-560) */
#define WFA_ERR_LIMIT              -8  /* beyond a fixed limit of
                                          wfa_align_batch_noalloc (pairs,
                                          length, memory mode) */

/* User penalties, as passed to WavefrontPenalties::new_* (src/penalties.rs:
 * 136-296). `match` <= 0; a negative match is a reward and the engine
 * applies the same Eizenga transform. For GAP_LINEAR `gap_opening1` is the
 * indel penalty. Unused fields are ignored. */
typedef struct wfa_penalties_t {
    int32_t distance_metric;
    int32_t match;
    int32_t mismatch;
    int32_t gap_opening1;
    int32_t gap_extension1;
    int32_t gap_opening2;
    int32_t gap_extension2;
} wfa_penalties_t;

/* Full aligner configuration: the penalties plus everything the reference's
 * setters configure (src/aligner.rs:462-611). `wfa_attr_default` fills the
 * reference's defaults (score scope, high memory, end-to-end, no heuristic,
 * max steps INT32_MAX, src/aligner.rs:416-458). */
typedef struct wfa_attr_t {
    wfa_penalties_t penalties;
    int32_t alignment_scope;        /* WFA_SCOPE_* */
    int32_t memory_mode;            /* WFA_MEMORY_* */
    int32_t span;                   /* WFA_SPAN_* */
    int32_t extension;              /* 1 = extension alignment (src/aligner.rs:582) */
    int32_t pattern_begin_free;
    int32_t pattern_end_free;
    int32_t text_begin_free;
    int32_t text_end_free;
    int32_t heuristic;              /* WFA_HEURISTIC_* */
    int32_t heuristic_p[3];
    int32_t max_alignment_steps;
    int32_t compat;                 /* WFA_COMPAT_* (set_compat, src/aligner.rs:53-94) */
} wfa_attr_t;

/* Defaults of WavefrontAligner::new (src/aligner.rs:409-458) for `metric`,
 * with the benchmark driver's default penalties for that metric
 * (src/bin/align_benchmark.rs:148-209). */
void wfa_attr_default(wfa_attr_t* attr, int32_t distance_metric);

/* Validate penalties as WavefrontPenalties::new_* does
 * (src/penalties.rs:157-167, 198-209, 245-262): WFA_OK or
 * WFA_ERR_INVALID_PENALTIES. */
int32_t wfa_penalties_check(const wfa_penalties_t* penalties);

/* ======================================================================
 * Per-call aligner: WavefrontAligner<N> (src/aligner.rs:146-211)
 * ====================================================================== */
typedef struct wfa_aligner_t wfa_aligner_t;

/* WavefrontAligner::new(WavefrontPenalties) (src/aligner.rs:409) followed by
 * the setters the attribute struct covers. `device` is a HIP device
 * ordinal. Returns NULL and sets *err on failure. */
wfa_aligner_t* wfa_aligner_new(const wfa_attr_t* attr, int32_t device, int32_t* err);
/* Drop of WavefrontAligner. */
void wfa_aligner_free(wfa_aligner_t* aligner);

/* set_alignment_scope (src/aligner.rs:542) */
int32_t wfa_aligner_set_alignment_scope(wfa_aligner_t* aligner, int32_t scope);
/* set_memory_mode (src/aligner.rs:462); WFA_MEMORY_ULTRALOW selects BiWFA as
 * the CLI does (src/bin/align_benchmark.rs:510-519). */
int32_t wfa_aligner_set_memory_mode(wfa_aligner_t* aligner, int32_t mode);
/* set_compat (src/aligner.rs:606-608): WFA_COMPAT_C or WFA_COMPAT_CORRECTED */
int32_t wfa_aligner_set_compat(wfa_aligner_t* aligner, int32_t compat);
/* set_heuristic (src/aligner.rs:611) */
int32_t wfa_aligner_set_heuristic(wfa_aligner_t* aligner, int32_t heuristic,
                                  int32_t p0, int32_t p1, int32_t p2);
/* set_alignment_free_ends (src/aligner.rs:563) */
int32_t wfa_aligner_set_alignment_free_ends(wfa_aligner_t* aligner,
                                            int32_t pattern_begin_free, int32_t pattern_end_free,
                                            int32_t text_begin_free, int32_t text_end_free);
/* set_alignment_extension (src/aligner.rs:582) */
int32_t wfa_aligner_set_alignment_extension(wfa_aligner_t* aligner);
/* set_alignment_end_to_end (src/aligner.rs:595) */
int32_t wfa_aligner_set_alignment_end_to_end(wfa_aligner_t* aligner);
/* set_max_alignment_steps (src/aligner.rs:533) */
int32_t wfa_aligner_set_max_alignment_steps(wfa_aligner_t* aligner, int32_t max_steps);
/* set_max_memory_compact (src/aligner.rs:469): the backtrace-buffer size
 * above which the reference's medium/low modes compact that buffer, a step
 * that only frees host memory and changes no result. The device keeps each
 * pair's CIGAR history in its own planned pool, so the value is stored
 * (wfa_aligner_max_memory_compact reads it back) and changes nothing. */
int32_t wfa_aligner_set_max_memory_compact(wfa_aligner_t* aligner, uint64_t bytes);
/* The stored value: 256 MiB (the reference's MAX_MEMORY_COMPACT,
 * src/aligner.rs:356) until set. 0 for a null aligner. */
uint64_t wfa_aligner_max_memory_compact(const wfa_aligner_t* aligner);

/* align_end2end(&[u8], &[u8]) -> i32 (src/aligner.rs:620). Returns the WFA
 * score; INT32_MIN if the call itself failed (see wfa_aligner_error). */
int32_t wfa_align_end2end(wfa_aligner_t* aligner,
                          const uint8_t* pattern, int32_t pattern_length,
                          const uint8_t* text, int32_t text_length);
/* align_packed2bits(&[u8], usize, &[u8], usize) -> i32 (src/aligner.rs:2755):
 * 2-bit packed input, LSB first (src/sequences.rs:34-51). */
int32_t wfa_align_packed2bits(wfa_aligner_t* aligner,
                              const uint8_t* pattern, int32_t pattern_length,
                              const uint8_t* text, int32_t text_length);
/* align_lambda(match_funct, pattern_length, text_length) -> i32
 * (src/aligner.rs:2743-2752, src/sequences.rs:225-257) for the match
 * functions that compare symbols: pattern position i matches text position j
 * iff pattern_class[pattern[i]] == text_class[text[j]] (two 256-entry
 * tables). This covers every relation that is an equivalence between the
 * pattern's and the text's symbols (byte equality, the reference's own
 * tests/singletrack.rs:123-136, with identity tables; case-insensitive or
 * RNA U = T matching; masked symbols mapped to a class of their own). A
 * device cannot call a host closure, so closures over positions, and
 * symbol relations that are not equivalences (an N matching every base),
 * have no counterpart here. End-to-end, as align_lambda; the CIGAR's M and X
 * follow the classes. */
int32_t wfa_align_lambda_classes(wfa_aligner_t* aligner,
                                 const uint8_t* pattern, int32_t pattern_length,
                                 const uint8_t* text, int32_t text_length,
                                 const uint8_t* pattern_class, const uint8_t* text_class);
/* align_endsfree(&[u8], &[u8]) -> i32 (src/aligner.rs:707): ends-free with
 * the free ends of wfa_aligner_set_alignment_free_ends, or extension after
 * wfa_aligner_set_alignment_extension (all ends zero otherwise). */
int32_t wfa_align_endsfree(wfa_aligner_t* aligner,
                           const uint8_t* pattern, int32_t pattern_length,
                           const uint8_t* text, int32_t text_length);
/* align_biwfa(&[u8], &[u8]) -> i32 (src/aligner.rs:2563): BiWFA, CIGAR in
 * O(score) memory, end-to-end. */
int32_t wfa_align_biwfa(wfa_aligner_t* aligner,
                        const uint8_t* pattern, int32_t pattern_length,
                        const uint8_t* text, int32_t text_length);
/* align_biwfa_score(&[u8], &[u8]) -> i32 (src/aligner.rs:2650): the score
 * of the first breakpoint (or of the end, reached first), no CIGAR. */
int32_t wfa_align_biwfa_score(wfa_aligner_t* aligner,
                              const uint8_t* pattern, int32_t pattern_length,
                              const uint8_t* text, int32_t text_length);
/* status() -> AlignStatus (src/aligner.rs:834), WFA_STATUS_* ordinals. */
int32_t wfa_aligner_status(const wfa_aligner_t* aligner);
/* cigar() -> &Cigar (src/aligner.rs:839): operations M/X/I/D, one byte each,
 * valid until the next call on this aligner. Returns the op count; *ops may
 * be NULL when the count is 0. */
int32_t wfa_aligner_cigar(const wfa_aligner_t* aligner, const uint8_t** ops);
/* Cigar.score (src/cigar.rs:26): the WFA score the CIGAR was traced at,
 * INT32_MIN when there is no CIGAR. */
int32_t wfa_aligner_cigar_score(const wfa_aligner_t* aligner);
/* alignment_end_pos (src/aligner.rs:146-204, a public field there) after
 * the last unidirectional alignment: out[0..3) = {score, k, offset} of the
 * end cell or of a Z-drop's best cell; unset {-1, INT32_MAX, INT32_MIN/2}
 * when neither exists or after a BiWFA alignment. */
void wfa_aligner_end_position(const wfa_aligner_t* aligner, int32_t* out);
/* Error code of the last call on this aligner (WFA_OK if it succeeded). */
int32_t wfa_aligner_error(const wfa_aligner_t* aligner);

/* ======================================================================
 * Batch engine: one device launch for many independent pairs
 * (replaces the per-pair loop of src/bin/align_benchmark.rs:593-627)
 * ====================================================================== */
typedef struct wfa_engine_t wfa_engine_t;

/* A batch whose arrays all live in device memory (HBM) of the engine's
 * device. Pair i's pattern starts at byte pattern_offset[i] of seq_data and
 * has pattern_length[i] bases (ceil(len/4) bytes when PACKED2BITS), likewise
 * the text. Outputs: score[i] (WFA score), status[i] (WFA_STATUS_*), and in
 * alignment scope the CIGAR of pair i as cigar_length[i] op bytes starting
 * at cigar_ops + cigar_offset[i]; the caller reserves
 * pattern_length[i] + text_length[i] bytes there. */
typedef struct wfa_batch_t {
    int32_t num_pairs;
    int32_t seq_format;              /* WFA_SEQ_* */
    int32_t max_length;              /* max over pairs of max(pattern, text) length;
                                        0 = unknown (the engine reduces it on the device) */
    const uint8_t* seq_data;
    const int64_t* pattern_offset;
    const int32_t* pattern_length;
    const int64_t* text_offset;
    const int32_t* text_length;
    int32_t* score;
    int32_t* status;
    uint8_t* cigar_ops;              /* NULL in score scope */
    const int64_t* cigar_offset;     /* NULL in score scope */
    int32_t* cigar_length;           /* NULL in score scope */
    const struct wfa_pair_config_t* pair_config;  /* optional, one per pair; NULL =
                                        every pair as the attribute says */
    int32_t* end_position;           /* optional, 3 per pair: the aligner's
                                        alignment_end_pos {score, k, offset}
                                        (src/aligner.rs:146-204): the end
                                        cell, a Z-drop's best cell (found at
                                        that earlier score), else unset
                                        {-1, INT32_MAX, INT32_MIN/2}. Not
                                        written for WFA_MEMORY_ULTRALOW. */
} wfa_batch_t;

/* One pair's span and heuristic within a batch: what the reference's
 * setters set on that pair's aligner (set_alignment_free_ends /
 * set_alignment_extension / set_alignment_end_to_end / set_heuristic,
 * src/aligner.rs:563-611). The launch's wfa_attr_t is the envelope: a pair
 * runs ends-free or extension only when attr.span is WFA_SPAN_ENDSFREE, and
 * its heuristic only when attr.heuristic is not WFA_HEURISTIC_NONE (the
 * attribute's own free ends and heuristic parameters are then unused).
 * The caller validates each pair's values as wfa_attr_t values are
 * validated: free ends >= 0, a heuristic kind of WFA_HEURISTIC_* (any other
 * value runs without one). Every heuristic, Z-drop included, runs in every
 * memory mode and compat of the batch. Not for WFA_MEMORY_ULTRALOW
 * batches. */
typedef struct wfa_pair_config_t {
    int32_t span;                   /* WFA_SPAN_* */
    int32_t extension;              /* 1 = extension (span WFA_SPAN_ENDSFREE) */
    int32_t pattern_begin_free;
    int32_t pattern_end_free;
    int32_t text_begin_free;
    int32_t text_end_free;
    int32_t heuristic;              /* WFA_HEURISTIC_* */
    int32_t heuristic_p[3];
    int32_t reserved[2];            /* zero */
} wfa_pair_config_t;

/* Create an engine on HIP device `device` (owns the device work queue,
 * per-slot wavefront scratch and the launch configuration). */
wfa_engine_t* wfa_engine_new(int32_t device, int32_t* err);
void wfa_engine_free(wfa_engine_t* engine);

/* Align every pair of `batch` with the configuration `attr` in one
 * persistent launch on `hip_stream` (a hipStream_t, NULL = default stream).
 * Returns once the results are in the output arrays: the engine reads back
 * a retry counter after the launch, and pairs whose wavefronts outgrew the
 * first launch's per-pair storage are re-run by wider launches before it
 * returns. Scratch grows on demand (allocations happen only when a batch
 * needs more than any earlier one), up to the scratch cap per launch
 * (wfa_engine_set_scratch_budget); a pair whose
 * CIGAR history still does not fit after six ever wider passes gets
 * WFA_STATUS_OOM (score 0, no CIGAR) instead of failing the batch. */
int32_t wfa_engine_align_batch(wfa_engine_t* engine, const wfa_attr_t* attr,
                               const wfa_batch_t* batch, void* hip_stream);

/* The plain per-pair path: the same device code launched once per pair,
 * with one wavefront, synchronizing after each pair — the comparator the
 * batched launch is measured against. */
int32_t wfa_engine_align_batch_per_pair(wfa_engine_t* engine, const wfa_attr_t* attr,
                                        const wfa_batch_t* batch, void* hip_stream);

/* Cap the device scratch (wavefront rings, CIGAR history) one launch of
 * this engine may use, in bytes (at least 64 MiB). Until set, the cap is
 * 48 GiB for short reads and, for long reads (over 1024 bases) and BiWFA,
 * half of the device's free memory plus what the engine already holds,
 * between 48 GiB (or 90% of that, when less is free) and 128 GiB, taken
 * again at most once a second. Pairs whose CIGAR history does not fit the cap get
 * WFA_STATUS_OOM. The reference has no counterpart (its memory is the
 * host's); this is the engine's own limit. */
int32_t wfa_engine_set_scratch_budget(wfa_engine_t* engine, int64_t bytes);

/* ---- Launch shapes of long-read launches and their autotuner ----
 * No counterpart in the reference (its aligner runs one pair on one core).
 * A long-read launch (sequences over 1024 bases, no heuristic, not BiWFA)
 * gives each pair a workgroup of several waves whose wavefront rings live
 * in HBM; how many waves per pair (4, 8 or 16; 4 and 16 for score-only and
 * medium/low-memory launches of pairs up to 16 kbp) and how many
 * workgroups per CU run together decide how the rings share the caches.
 * The engine keeps one shape per (metric, scope and memory mode, length
 * bucket = ceil(log2(longest sequence))); without one it uses a fixed
 * plan (8 waves, as many workgroups as fit). Results never depend on the
 * shape. */
#define WFA_TUNE_MAX_TRIALS 16
typedef struct {
    int32_t waves_per_pair;
    int32_t groups_per_cu;
    float ms;                    /* the trial launch on the sample (HIP events) */
} wfa_tune_trial_t;
typedef struct {
    int32_t sample_pairs;        /* pairs each trial aligned */
    int32_t num_trials;
    wfa_tune_trial_t trials[WFA_TUNE_MAX_TRIALS];
    int32_t waves_per_pair;      /* the shape kept (0: nothing to tune) */
    int32_t groups_per_cu;
    int32_t default_waves_per_pair;  /* the fixed plan's shape */
    int32_t default_groups_per_cu;
} wfa_tune_report_t;

/* Time the candidate shapes for `batch` under `attr` on its
 * `sample_pairs` longest pairs (0: 32 per CU), longest first as a batch
 * runs, each trial with its retry passes, on `hip_stream`, and keep the
 * fastest in the engine for the batch's key. The trials write the sampled
 * pairs' results into the batch's outputs (a later wfa_engine_align_batch
 * rewrites them). Batches
 * with nothing to choose (short reads, heuristics, BiWFA) return WFA_OK
 * with waves_per_pair 0. `report` may be NULL. */
int32_t wfa_engine_autotune(wfa_engine_t* engine, const wfa_attr_t* attr,
                            const wfa_batch_t* batch, void* hip_stream, int32_t sample_pairs,
                            wfa_tune_report_t* report);
/* Set (or, with 0 and 0, clear) the shape for launches under `attr` whose
 * longest sequence falls in the bucket of `max_length`. */
int32_t wfa_engine_set_launch_shape(wfa_engine_t* engine, const wfa_attr_t* attr,
                                    int32_t max_length, int32_t waves_per_pair,
                                    int32_t groups_per_cu);
/* The shape kept for that key, 0 and 0 when the fixed plan applies. */
int32_t wfa_engine_get_launch_shape(wfa_engine_t* engine, const wfa_attr_t* attr,
                                    int32_t max_length, int32_t* waves_per_pair,
                                    int32_t* groups_per_cu);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code that must not call hipMalloc, such as tight serving loops
 * or harnesses that forbid it). Aligns `batch` exactly as
 * wfa_engine_align_batch does, on the current HIP device and `hip_stream`,
 * but allocates nothing: every scratch buffer comes from one fixed device
 * array of WFA_NOALLOC_SCRATCH_BYTES that each GPU holds until the process
 * ends. The caller passes inputs and outputs as for wfa_engine_align_batch.
 * Limits: at most WFA_NOALLOC_MAX_PAIRS pairs and WFA_NOALLOC_MAX_LENGTH
 * bases per sequence, memory modes HIGH/MED/LOW (not ULTRALOW); beyond them
 * WFA_ERR_LIMIT. Within the limits, a pair whose CIGAR history does not fit
 * the array after the wider passes gets WFA_STATUS_OOM. All calls on one
 * GPU share its array and leave their contents in it; calls in one process
 * are serialized per GPU and return after their kernels finish, so they
 * never use the array at the same time. */
#define WFA_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
#define WFA_NOALLOC_MAX_PAIRS     (1 << 20)
#define WFA_NOALLOC_MAX_LENGTH    16384
int32_t wfa_align_batch_noalloc(const wfa_attr_t* attr, const wfa_batch_t* batch,
                                void* hip_stream);

/* Launch statistics of the last batch call: number of pairs that needed a
 * wider wavefront or longer history than the first pass provides and were
 * re-run by a second, wider pass. */
int32_t wfa_engine_last_retries(const wfa_engine_t* engine);

/* HIP-event timing of the last batch call, measured on the stream it ran
 * on: the first (persistent) launch alone and all launches including the
 * retry passes, in milliseconds. */
int32_t wfa_engine_last_kernel_ms(const wfa_engine_t* engine, float* first_pass_ms,
                                  float* all_passes_ms);

/* No counterpart in the reference: the phase clocks of a profiling build
 * (python wfa2lib-rs_amd/build.py --phase-clocks, libwfa_mi355x_phases.so).
 * After a batch call whose first pass ran the short-read kernel: out[0]
 * pairs it aligned, out[1] cycles its waves spent staging sequences from
 * HBM into LDS, out[2] cycles extending M rows (the LCP over the staged
 * sequences), out[3] cycles per pair in all, out[4] cycles in the CIGAR
 * backtrace, summed over the waves (five values). WFA_ERR_UNSUPPORTED in a
 * release build. */
int32_t wfa_engine_last_phase_cycles(wfa_engine_t* engine, uint64_t* out);

/* ======================================================================
 * Serving loop: continuous batching of independent requests
 * (replaces one reference aligner per request, align_end2end /
 * align_endsfree / align_biwfa per call, src/aligner.rs:620, 707, 2563)
 * ====================================================================== */
typedef struct wfa_server_t wfa_server_t;

typedef struct wfa_server_config_t {
    int32_t device;              /* HIP device ordinal */
    int32_t num_engines;         /* launches that may run at once, each on its own
                                    stream with its own scratch (1..16, default 4) */
    int32_t max_batch_pairs;     /* most requests one launch takes (default 65536) */
    int64_t memory_budget;       /* predicted device bytes of the work in flight
                                    (default 64 GiB); each engine's launches may
                                    use up to twice its even share of it */
} wfa_server_config_t;

/* One request: its own configuration and sequences (host memory, copied by
 * wfa_server_submit). The span, heuristic, memory mode and scope of `attr`
 * act as the reference's setters would on a fresh aligner. */
typedef struct wfa_server_request_t {
    wfa_attr_t attr;
    int32_t seq_format;          /* WFA_SEQ_* */
    int32_t pattern_length;      /* bases */
    int32_t text_length;
    const uint8_t* pattern;
    const uint8_t* text;
    uint64_t tag;                /* the caller's id, returned with the result */
} wfa_server_request_t;

typedef struct wfa_server_result_t {
    uint64_t tag;
    int32_t error;               /* WFA_OK, or the WFA_ERR_* of the launch that ran it */
    int32_t score;               /* WFA score, as the reference's align call returns */
    int32_t status;              /* WFA_STATUS_* */
    int32_t cigar_length;        /* alignment scope: op count */
    const uint8_t* cigar_ops;    /* M/X/I/D bytes, valid until the next poll */
    double latency_us;           /* from submission to the result being ready */
} wfa_server_result_t;

/* Defaults: device 0, 4 engines, 65536 requests per launch, 64 GiB. */
void wfa_server_config_default(wfa_server_config_t* config);
/* Start the engines and their worker threads; NULL and *err on failure.
 * Each engine launches on its own HIP stream. HIP runs a process's streams
 * on at most GPU_MAX_HW_QUEUES hardware queues (4 unless the environment
 * sets it before HIP starts), and streams that share a queue run their
 * kernels in order, so a short request's launch can wait behind a long
 * one's: give a serving process at least num_engines + 4 queues. Long-read
 * launches (sequences over 1024 bases) may use half the engines at once. */
wfa_server_t* wfa_server_new(const wfa_server_config_t* config, int32_t* err);
/* Queue one request (validated as the aligner validates its configuration:
 * WFA_ERR_INVALID_PENALTIES, WFA_ERR_UNSUPPORTED, WFA_ERR_INVALID_ARGUMENT).
 * Requests are admitted to the device by class (configuration, format,
 * length bucket) in arrival order while their predicted wavefront memory
 * fits the budget; results do not depend on the batch they ran in. */
int32_t wfa_server_submit(wfa_server_t* server, const wfa_server_request_t* request);
/* Up to `max_results` finished requests, in completion order. Waits up to
 * `timeout_us` (0: no wait, negative: until a result is ready or nothing is
 * queued or running). Returns the count, or a negative WFA_ERR_*. */
int32_t wfa_server_poll(wfa_server_t* server, wfa_server_result_t* results,
                        int32_t max_results, int32_t timeout_us);
/* Requests queued, requests running, launches made and requests launched. */
int32_t wfa_server_stats(wfa_server_t* server, int64_t* queued, int64_t* running,
                         int64_t* launches, int64_t* launched_pairs);
/* Stop the workers (after the launches that run finish), drop queued
 * requests and release every engine. */
void wfa_server_free(wfa_server_t* server);

/* Library version string. */
const char* wfa_version(void);

#ifdef __cplusplus
}
#endif
#endif /* WFA_MI355X_H */
