/*
 * cosine_similarity_amd.h -- C ABI of the MI355X (gfx950) spectral-similarity
 * library (libcsamd.so).
 *
 * Plain C: device pointers, sizes and an opaque stream.  No exceptions cross
 * this boundary; every entry point returns a cs_status and leaves a message
 * for cs_last_error() on failure.  The caller owns every buffer (inputs,
 * outputs, workspace); the library allocates no device memory and never
 * synchronises the stream.  Calls on different streams are independent.
 *
 * Reference interfaces replaced (files in tornikeo/cosine-similarity):
 *   - the Numba-CUDA launcher `kernel(rspec, qspec, lens, out, stream)` built
 *     by cosine_greedy_kernel(), cudams/similarity/spectrum_similarity_
 *     functions.py:12-37 (factory) and :264-276 (launcher), called from
 *     cudams/similarity/CudaCosineGreedy.py:179;
 *   - the torch implementations of jaccard/dice/cosine_similarity_matrix,
 *     cudams/similarity/vector_similarity_functions.py:12-47, :50-91, :94-119,
 *     called from cudams/similarity/CudaFingerprintSimilarity.py:166-177.
 */
#ifndef COSINE_SIMILARITY_AMD_H
#define COSINE_SIMILARITY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cs_status {
    CS_OK = 0,
    CS_ERR_INVALID = 1,     /* bad argument (null pointer, size, layout) */
    CS_ERR_WORKSPACE = 2,   /* workspace smaller than the *_workspace_size result */
    CS_ERR_HIP = 3,         /* a HIP runtime call failed (launch, attribute) */
    CS_ERR_UNSUPPORTED = 4  /* input outside the supported envelope */
} cs_status;

/* Message of the last failing call on this host thread ("" if none). */
const char *cs_last_error(void);

/* Library ABI version (major * 10000 + minor * 100 + patch). */
int cs_version(void);

/* ------------------------------------------------------------------------ */
/* Spectra                                                                    */
/* ------------------------------------------------------------------------ */

/*
 * A set of spectra resident in device memory as ragged CSR.  Spectrum s owns
 * peaks [offsets[s], offsets[s+1]) of the peak arrays; its m/z are ascending.
 * offsets[0] need not be 0 (a slice of a larger set), and n_peaks must be at
 * least offsets[n_spectra] - offsets[0].  Replaces the padded
 * [2, batch, n_max_peaks] float32 tensor + lens of CudaCosineGreedy.py:63-87.
 *
 * The per-peak power factors are computed by the caller in float64 exactly as
 * matchms writes them (x ** power), so the products the library forms,
 * ((mz_pow_r * int_pow_r) * mz_pow_q) * int_pow_q, are bit-identical to
 * matchms collect_peak_pairs.
 */
typedef struct cs_spectra {
    int64_t n_spectra;
    int64_t n_peaks;
    const int64_t *offsets;  /* device [n_spectra + 1]                             */
    const float *mz;         /* device [n_peaks] m/z rounded to nearest float32   */
    const double *mz_exact;  /* device [n_peaks] m/z as given (float64); NULL when
                                every value of `mz` is exact                      */
    const double *mz_pow;    /* device [n_peaks] m/z ** mz_power; NULL means 1.0
                                (mz_power == 0)                                   */
    const double *int_pow;   /* device [n_peaks] intensity ** intensity_power     */
    const double *norm;      /* device [n_spectra] sqrt(sum((mz_pow*int_pow)^2))  */
    int32_t max_peaks;       /* host-known max peaks of one spectrum (<= 65535)   */
    float max_abs_mz;        /* host-known max |m/z| over all peaks               */
} cs_spectra;

/*
 * Fault bits of cs_validate_spectra / cs_cpu_cosine_greedy's input check.
 */
enum {
    CS_FAULT_OFFSETS = 1,   /* offsets[0] < 0, decreasing, or past offsets[0] + n_peaks */
    CS_FAULT_LENGTH = 2,    /* a spectrum longer than max_peaks                         */
    CS_FAULT_MZ_RANGE = 4,  /* an m/z not finite or above max_abs_mz in magnitude       */
    CS_FAULT_MZ_ORDER = 8   /* m/z not ascending within a spectrum                      */
};

/*
 * Checks a device-resident spectrum set before it is trusted:
 * *fault (a device uint32) receives the OR of the CS_FAULT_* bits over every
 * spectrum, 0 when the set is sound.  Asynchronous on `stream` (the flag is
 * cleared by a kernel first, so the call can sit in a captured graph); read
 * the flag after the stream is synchronised.  Peak arrays are read only
 * within [offsets[0], offsets[0] + n_peaks), and only for spectra whose own
 * offsets lie inside it.  The entry points that compute take max_peaks,
 * max_abs_mz, the offsets and the m/z order on trust (an understated
 * max_abs_mz narrows the probe window, unsorted m/z breaks the tile merge):
 * run this on any set the caller did not pack itself.  Replaces the checks
 * the reference leaves to matchms's Spectrum constructor (sorted float64
 * m/z) and to CudaCosineGreedy._spectra_peaks_to_tensor's truncation
 * (CudaCosineGreedy.py:63-87).
 */
cs_status cs_validate_spectra(const cs_spectra *spectra, uint32_t *fault, void *stream);

/*
 * norm[x] = sqrt(sum of (mz_pow[p] * int_pow[p])^2 over spectrum x's peaks,
 * in peak order), int_pow[p]^2 when mz_pow is NULL: the norms that
 * cs_spectra.norm must hold, computed on the device (`norm`: device
 * [n_spectra] float64; spectra->norm is not read).  A sequential float64 sum
 * per spectrum, as the host's left-to-right sum, so either may fill the
 * field.  Asynchronous on `stream`.  Replaces the norm part of the reference
 * kernel's preamble (spectrum_similarity_functions.py:82-110); the offsets
 * are taken on trust (validate first for a set the caller did not pack).
 */
cs_status cs_spectra_norms(const cs_spectra *spectra, double *norm, void *stream);

/*
 * CosineGreedy parameters (CudaCosineGreedy.py:24-56).  mz_power and
 * intensity_power are folded into cs_spectra.mz_pow / int_pow.
 */
typedef struct cs_cosine_params {
    double tolerance;        /* peaks match when |mz_q + shift - mz_r| <= tolerance
                                (matchms find_matches bounds, inclusive)          */
    double shift;            /* added to query m/z (matchms CosineGreedy: 0)      */
    int64_t match_limit;     /* overflow flag = (#candidates >= match_limit);
                                results stay exact beyond it                      */
    int32_t light_candidates;/* candidates one lane settles in registers before the
                                pair moves to the wave-cooperative path: one of
                                1, 4, 8, 16; 0 = library default                  */
    int32_t round_capacity;  /* candidates the cooperative path settles per round
                                (power of two, 4..1024); 0 = library default      */
    int32_t refs_per_block;  /* references one lane-kernel workgroup walks against
                                its query tile, 1..2048; 0 = library default      */
    int32_t stages;          /* 0 = whole pipeline, else a mask of CS_STAGE_*:
                                calls with PREPARE, then LIGHT, then HEAVY on one
                                stream and one workspace equal one call with 0.
                                PREPARE depends only on `queries`: once done, any
                                number of LIGHT|HEAVY calls with other `refs` may
                                reuse it (same workspace, same queries)           */
} cs_cosine_params;

enum {
    CS_STAGE_PREPARE = 1,    /* index the queries: 64-query tiles whose merged,
                                sorted peak list and bucket directory live in the
                                workspace                                         */
    CS_STAGE_LIGHT = 2,      /* probe every reference against every tile; pairs
                                with <= 64 candidates are settled here, the rest
                                are queued                                        */
    CS_STAGE_HEAVY = 4       /* one workgroup per queued pair (one wave per pair
                                beyond 4096 candidates)                           */
};

/* ------------------------------------------------------------------------ */
/* CosineGreedy all-pairs (dense output)                                      */
/* ------------------------------------------------------------------------ */

/* Bytes of device workspace cs_cosine_greedy needs for these shapes.  The
 * size depends on `queries` only (refs is accepted for symmetry). */
cs_status cs_cosine_greedy_workspace_size(const cs_spectra *refs, const cs_spectra *queries,
                                          size_t *bytes);

/*
 * score[r*ld + q], matches[r*ld + q], overflow[r*ld + q] for every reference r
 * and query q (row-major, ld >= queries->n_spectra).  Replaces the launcher at
 * spectrum_similarity_functions.py:264-276 plus the per-tile loop of
 * CudaCosineGreedy.matrix (CudaCosineGreedy.py:149-185): one call covers the
 * whole R x Q matrix, asynchronously on `stream` (a hipStream_t, NULL = the
 * null stream).  Semantics are matchms CosineGreedy.pair per pair:
 *   score    float32 of the float64 greedy score,
 *   matches  number of greedily accepted peak pairs (exact, never truncated),
 *   overflow 1 iff the pair had >= match_limit candidate peak pairs.
 */
cs_status cs_cosine_greedy(const cs_spectra *refs, const cs_spectra *queries,
                           const cs_cosine_params *params, float *score, int32_t *matches,
                           uint8_t *overflow, int64_t ld, void *workspace,
                           size_t workspace_bytes, void *stream);

/*
 * The same computation with the workspace split in two parts:
 *   index    the query index, written by CS_STAGE_PREPARE and only read by
 *            CS_STAGE_LIGHT / CS_STAGE_HEAVY;
 *   scratch  the per-call heavy-pair queue and counters (LIGHT / HEAVY).
 * Both sizes depend on `queries` only.  Calls that share one prepared index
 * and each use their own scratch may run concurrently on different streams:
 * a device-resident library (the `queries` side, indexed once) searched by
 * several in-flight batches of request spectra (the `refs` side).  Replaces
 * the same launcher as cs_cosine_greedy (spectrum_similarity_functions.py:
 * 264-276) for callers that keep one side resident, which the reference
 * re-uploads per tile pair (CudaCosineGreedy.py:157-163).
 */
cs_status cs_cosine_greedy_indexed_sizes(const cs_spectra *queries, size_t *index_bytes,
                                         size_t *scratch_bytes);

cs_status cs_cosine_greedy_indexed(const cs_spectra *refs, const cs_spectra *queries,
                                   const cs_cosine_params *params, float *score, int32_t *matches,
                                   uint8_t *overflow, int64_t ld, void *index, size_t index_bytes,
                                   void *scratch, size_t scratch_bytes, void *stream);

/*
 * The same computation as cs_cosine_greedy without any workspace argument and
 * without allocating: the workspace is a file-scope device array of
 * CS_NOALLOC_WORKSPACE_BYTES that every GPU the library runs on holds until
 * the process ends (its device address is looked up once per device, on the
 * first call with that device current).  The caller passes inputs and
 * outputs; the call makes no hipMalloc / hipFree or other allocating call.
 * Query sets whose cs_cosine_greedy_workspace_size exceeds the array are
 * rejected with CS_ERR_UNSUPPORTED.  Every call on one GPU shares that GPU's
 * array and leaves its contents behind (a PREPARE-only call's index stays for
 * later LIGHT | HEAVY calls until another call prepares): calls that overlap
 * on one GPU, on any streams, race.
 * No counterpart in the reference: this entry point is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation).
 */
#define CS_NOALLOC_WORKSPACE_BYTES (64u << 20)

cs_status cs_cosine_greedy_noalloc(const cs_spectra *refs, const cs_spectra *queries,
                                   const cs_cosine_params *params, float *score, int32_t *matches,
                                   uint8_t *overflow, int64_t ld, void *stream);

/* ------------------------------------------------------------------------ */
/* Batches of independent CosineGreedy matrices                               */
/* ------------------------------------------------------------------------ */

/* One matrix of a batch: the arguments of one cs_cosine_greedy call. */
typedef struct cs_cosine_job {
    cs_spectra refs;
    cs_spectra queries;
    double tolerance;
    double shift;
    int64_t match_limit;
    float *score;            /* device [refs.n_spectra x ld] */
    int32_t *matches;
    uint8_t *overflow;
    int64_t ld;
} cs_cosine_job;

/* Launch shape of a batch (what an autotuner chooses); 0 = library default. */
typedef struct cs_batch_params {
    int32_t light_candidates;  /* as cs_cosine_params.light_candidates          */
    int32_t round_capacity;    /* as cs_cosine_params.round_capacity            */
    int32_t refs_per_block;    /* references of one work item (x 64 queries), 1..2048;
                                  default 256                                     */
    int32_t grid;              /* persistent workgroups of the lane kernel;
                                  default one per compute unit                    */
} cs_batch_params;

typedef struct cs_cosine_batch cs_cosine_batch;  /* host-side plan, opaque */

/*
 * Plans a batch of n_jobs (1..4095) independent matrices of uneven sizes and
 * tolerances: validates every job as cs_cosine_greedy would, splits all of
 * them into work items (job, 64-query tile, block of references), orders the
 * items costliest first and sizes one workspace for everything.  Replaces the
 * reference's one launch per tile pair and host sync per tile
 * (CudaCosineGreedy.py:149-185), repeated per matrix; the reference has no
 * batch entry point.  The plan is host memory: free it with
 * cs_cosine_batch_destroy.
 */
cs_status cs_cosine_batch_create(const cs_cosine_job *jobs, int32_t n_jobs,
                                 const cs_batch_params *params, cs_cosine_batch **batch,
                                 size_t *workspace_bytes);

/* Copies the plan's tables into `workspace` (synchronises `stream`). */
cs_status cs_cosine_batch_upload(cs_cosine_batch *batch, void *workspace, size_t workspace_bytes,
                                 void *stream);

/*
 * Computes every job: query tiling of all jobs, one persistent lane-kernel
 * launch whose workgroups take work items from a shared counter, then one
 * launch each for the queued many-candidate pairs.  Only kernel launches on
 * `stream` (no copies, no memsets, no host sync), so it can be captured into
 * a HIP graph and replayed; results equal cs_cosine_greedy's per job.
 */
cs_status cs_cosine_batch_run(const cs_cosine_batch *batch, void *stream);

void cs_cosine_batch_destroy(cs_cosine_batch *batch);

/* ------------------------------------------------------------------------ */
/* Fingerprint similarity                                                     */
/* ------------------------------------------------------------------------ */

typedef enum cs_fingerprint_measure {
    CS_JACCARD = 0,  /* I / (|a| + |b| - I)      vector_similarity_functions.py:42-46   */
    CS_DICE = 1,     /* 2 * (I / (|a| + |b|))    vector_similarity_functions.py:87-90   */
    CS_COSINE = 2    /* I / sqrt(|a| * |b|)      vector_similarity_functions.py:116-118 */
} cs_fingerprint_measure;

/* Bytes of device workspace cs_fingerprint_similarity needs (currently 0:
 * the kernel reads the packed words directly; the workspace arguments are
 * kept so callers need not change if a future kernel needs scratch). */
cs_status cs_fingerprint_workspace_size(int64_t n_refs, int64_t n_queries, int64_t n_words,
                                        size_t *bytes);

/*
 * out[r*ld + q] = measure(ref r, query q) as float32; 0/0 gives 0 (the
 * reference's nan_to_num).  Fingerprints are bit-packed, n_words uint32 words
 * per row (bit b of row x is word b/32, bit b%32); at most 2^23 bits
 * (CS_ERR_UNSUPPORTED beyond).  Intersections and popcounts are exact (fp4
 * matrix-core products of 0/1 bit planes, accumulated in float32).  Rows
 * whose valid flag is 0 (missing fingerprint) get `empty_score` instead
 * (CudaFingerprintSimilarity.py:147-154); a NULL valid array means all valid.
 * If `intersections` is not NULL it receives the int32 counts [r*ld + q].
 */
cs_status cs_fingerprint_similarity(int measure, const uint32_t *ref_words, int64_t n_refs,
                                    const uint32_t *query_words, int64_t n_queries,
                                    int64_t n_words, const uint8_t *ref_valid,
                                    const uint8_t *query_valid, float empty_score, float *out,
                                    int32_t *intersections, int64_t ld, void *workspace,
                                    size_t workspace_bytes, void *stream);

/* ------------------------------------------------------------------------ */
/* Reductions over a dense score block                                        */
/* ------------------------------------------------------------------------ */

/*
 * Per row, the k (1..16) best columns by score descending, ties to the lower
 * column; NaN scores rank last.  top_score / top_index (/ top_matches if not
 * NULL) are [rows, k] row-major; slots beyond `cols` get NaN / -1.  Replaces
 * the host-side best-hits-per-query of matchms Scores.scores_by_query(sort=True)
 * that the reference's users apply (tests/usecase/test_tutorial_pesticide.py:
 * 38-41, 64-67).
 */
cs_status cs_topk_rows(const float *score, const int32_t *matches, int64_t rows, int64_t cols,
                       int64_t ld, int k, float *top_score, int32_t *top_index,
                       int32_t *top_matches, void *workspace, size_t workspace_bytes,
                       void *stream);

/* Bytes of workspace cs_topk_rows needs (0 when cols <= 4096: rows longer
 * than that are ranked in 4096-column chunks whose partial hits are merged). */
cs_status cs_topk_rows_workspace_size(int64_t rows, int64_t cols, int k, size_t *bytes);

/*
 * Threshold compaction (CudaCosineGreedy.py:186-200): entries with
 * score >= threshold, row-major order.  Two calls on one stream:
 *   1. out_row == NULL: row_counts[rows] and row_offsets[rows + 1] (exclusive
 *      prefix; row_offsets[rows] is the total) are written;
 *   2. with outputs of `capacity` entries: (row0 + row, col, score[, matches]
 *      [, overflow]) are written at row_offsets[row] onward (entries beyond
 *      capacity are dropped).
 */
cs_status cs_threshold_compact(const float *score, const int32_t *matches,
                               const uint8_t *overflow, int64_t rows, int64_t cols, int64_t ld,
                               float threshold, int64_t row0, int64_t *row_offsets,
                               int32_t *row_counts, int64_t capacity, int64_t *out_row,
                               int64_t *out_col, float *out_score, int32_t *out_matches,
                               uint8_t *out_overflow, void *stream);

/*
 * The dense result as packed 9-byte records, row-major without padding:
 * records[(r*cols + c)*9 + 0..3] = score (float32), + 4..7 = matches (int32),
 * + 8 = overflow (uint8), little-endian -- the memory of a numpy array of
 * dtype [("score", "<f4"), ("matches", "<i4"), ("overflow", "u1")], so one
 * device-to-host copy delivers the structured array.  `records` must be
 * 4-byte aligned.  Replaces the three copies and np.rec.fromarrays of
 * CudaCosineGreedy.py:201-205.
 */
cs_status cs_pack_records(const float *score, const int32_t *matches, const uint8_t *overflow,
                          int64_t rows, int64_t cols, int64_t ld, uint8_t *records, void *stream);

/* ------------------------------------------------------------------------ */
/* Host engine (libcsamd_cpu.so: g++ + OpenMP, no GPU needed)                 */
/* ------------------------------------------------------------------------ */

/* Message of the last failing cs_cpu_* call on this host thread. */
const char *cs_cpu_last_error(void);

/*
 * CosineGreedy for every (reference, query) pair on the host cores:
 * score[r*ld + q] (float32 of the float64 greedy score) and matches[r*ld + q],
 * matchms CosineGreedy.pair semantics as cs_cosine_greedy, without the
 * overflow plane (CPUParallelCosineGreedy.score_datatype has none).  Every
 * pointer in `refs` / `queries` is HOST memory; the whole CSR is validated
 * first (offsets non-decreasing and within n_peaks, lengths <= max_peaks,
 * m/z finite, ascending per spectrum and within max_abs_mz): a violation
 * returns CS_ERR_INVALID.  n_threads <= 0 uses OpenMP's default.
 * Replaces the numba prange kernel cpu_parallel_cosine_greedy_kernel
 * (cudams/similarity/spectrum_similarity_functions.py:281-320) called from
 * CPUParallelCosineGreedy (the kernel built at
 * cudams/similarity/CPUParallelCosineGreedy.py:42-49, called at :51-68).
 */
cs_status cs_cpu_cosine_greedy(const cs_spectra *refs, const cs_spectra *queries,
                               double tolerance, double shift, float *score, int32_t *matches,
                               int64_t ld, int n_threads);

#ifdef __cplusplus
}
#endif

#endif /* COSINE_SIMILARITY_AMD_H */
