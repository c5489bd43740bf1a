/*
 * cs_oracle.c -- CPU oracle for the pairwise spectral-similarity path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or as the timed CPU baseline.  The product path (the HIP library under
 * cosine-similarity_amd/) never calls into it.
 *
 * It restates, in plain C and float64, the algorithm the reference delegates
 * to matchms (>= 0.24.0, the lower bound in setup.py:35; matchms is not in the
 * reference tree and cannot be installed here):
 *
 *   CosineGreedy.pair
 *     = collect_peak_pairs  (find_matches window scan + power products)
 *     -> np.argsort(v, kind="mergesort")[::-1]   (stable ascending, reversed)
 *     -> score_best_matches (greedy unique assignment, then normalisation)
 *
 * The reference restates the same call sequence in
 * cudams/similarity/spectrum_similarity_functions.py:306-317 (CPU kernel) and
 * implements it on the GPU at :123-149 (window), :168-198 (tie order),
 * :208-218 (greedy) and :261 (normalisation).  The overflow flag follows the
 * reference rule "candidates >= match_limit" (:149-151).
 *
 * Fingerprint measures restate vector_similarity_functions.py:42-46 (jaccard),
 * :87-90 (dice) and :116-118 (cosine), including nan_to_num -> 0, on 0/1
 * fingerprints with exact integer counts.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   cso_cosine_greedy_pair : the four known-answer cases of
 *       tests/similarity/test_cosine_greedy.py:27-95 (expected matched index
 *       lists + closed-form score), the symmetry case :219-249, the nine
 *       matchms values printed in notebooks/samples/colab_tutorial_pesticide
 *       .ipynb (query scan 2161), brute force over all candidate orderings on
 *       tiny inputs, and a plant test (a wrong tie order must fail).
 *   cso_fingerprint_matrix : the closed-form values of
 *       tests/similarity/test_fingerprint_similarity.py:8-112 and a brute-force
 *       popcount check.
 *   cso_near_tie_pairs : constructed pairs whose two conflicting candidates
 *       differ below float32 resolution (flagged) or above it (not flagged),
 *       and a brute-force pure-Python restatement with float32 keys on tiny
 *       random inputs.
 * The norm sum runs left to right; numpy's pairwise summation can differ in
 * the last bit of the norm (1e-16 relative), far below the 1e-5 score bar.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int32_t i;  /* reference peak index */
    int32_t j;  /* query peak index */
    double v;   /* power product */
} cso_pair;

/* find_matches + collect_peak_pairs (matchms); the reference GPU kernel uses
 * the same scan at spectrum_similarity_functions.py:123-149.  Candidates are
 * appended in (i ascending, j ascending) order.  Returns the count. */
static int64_t collect_peak_pairs(const double *mz1, const double *in1, int64_t n1,
                                  const double *mz2, const double *in2, int64_t n2,
                                  double tol, double shift, double mzp, double inp,
                                  cso_pair *out, int64_t cap) {
    int64_t lowest = 0, n = 0;
    for (int64_t a = 0; a < n1; ++a) {
        const double low_bound = mz1[a] - tol;
        const double high_bound = mz1[a] + tol;
        for (int64_t b = lowest; b < n2; ++b) {
            const double mz2s = mz2[b] + shift;
            if (mz2s > high_bound) break;
            if (mz2s < low_bound) {
                lowest = b;
            } else {
                if (out && n < cap) {
                    /* left-to-right product, as written in matchms */
                    double v = pow(mz1[a], mzp) * pow(in1[a], inp);
                    v = v * pow(mz2[b], mzp);
                    v = v * pow(in2[b], inp);
                    out[n].i = (int32_t)a;
                    out[n].j = (int32_t)b;
                    out[n].v = v;
                }
                ++n;
            }
        }
    }
    return n;
}

/* Stable bottom-up merge sort, ascending by v (ties keep collection order),
 * i.e. numpy's argsort(kind="mergesort"). */
static void stable_sort_ascending(cso_pair *a, cso_pair *tmp, int64_t n) {
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            int64_t mid = lo + w < n ? lo + w : n;
            int64_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t x = lo, y = mid, k = lo;
            while (x < mid && y < hi) {
                if (a[y].v < a[x].v) tmp[k++] = a[y++];
                else tmp[k++] = a[x++];      /* equal: left run first (stable) */
            }
            while (x < mid) tmp[k++] = a[x++];
            while (y < hi) tmp[k++] = a[y++];
        }
        memcpy(a, tmp, (size_t)n * sizeof(cso_pair));
    }
}

static double power_norm(const double *mz, const double *in, int64_t n, double mzp, double inp) {
    double s = 0.0;
    for (int64_t k = 0; k < n; ++k) {
        double w = pow(mz[k], mzp) * pow(in[k], inp);
        s += w * w;
    }
    return sqrt(s);
}

/* score_best_matches over the reversed stable order.  Workspace: cap pairs
 * in `buf` and `tmp`, n1 + n2 bytes in `used`. */
static void score_pair(const double *mz1, const double *in1, int64_t n1,
                       const double *mz2, const double *in2, int64_t n2,
                       double tol, double shift, double mzp, double inp,
                       cso_pair **buf, cso_pair **tmp, int64_t *cap, unsigned char **used,
                       int64_t *used_cap, double *score, int32_t *matches, int64_t *ncand) {
    int64_t c = collect_peak_pairs(mz1, in1, n1, mz2, in2, n2, tol, shift, mzp, inp, NULL, 0);
    *ncand = c;
    if (c == 0) {
        *score = 0.0;
        *matches = 0;
        return;
    }
    if (c > *cap) {
        *cap = c;
        *buf = (cso_pair *)realloc(*buf, (size_t)c * sizeof(cso_pair));
        *tmp = (cso_pair *)realloc(*tmp, (size_t)c * sizeof(cso_pair));
    }
    if (n1 + n2 > *used_cap) {
        *used_cap = n1 + n2;
        *used = (unsigned char *)realloc(*used, (size_t)(n1 + n2));
    }
    collect_peak_pairs(mz1, in1, n1, mz2, in2, n2, tol, shift, mzp, inp, *buf, c);
    stable_sort_ascending(*buf, *tmp, c);
    unsigned char *u1 = *used, *u2 = *used + n1;
    memset(*used, 0, (size_t)(n1 + n2));
    double s = 0.0;
    int32_t m = 0;
    for (int64_t k = c - 1; k >= 0; --k) {   /* [::-1] */
        const cso_pair p = (*buf)[k];
        if (!u1[p.i] && !u2[p.j]) {
            s += p.v;
            u1[p.i] = 1;
            u2[p.j] = 1;
            ++m;
        }
    }
    const double den = power_norm(mz1, in1, n1, mzp, inp) * power_norm(mz2, in2, n2, mzp, inp);
    *score = s / den;
    *matches = m;
}

/* One pair.  Returns 0. */
int cso_cosine_greedy_pair(const double *mz1, const double *in1, int64_t n1,
                           const double *mz2, const double *in2, int64_t n2,
                           double tol, double shift, double mz_power, double int_power,
                           double *score, int32_t *matches, int64_t *n_candidates) {
    cso_pair *buf = NULL, *tmp = NULL;
    unsigned char *used = NULL;
    int64_t cap = 0, ucap = 0;
    score_pair(mz1, in1, n1, mz2, in2, n2, tol, shift, mz_power, int_power, &buf, &tmp, &cap,
               &used, &ucap, score, matches, n_candidates);
    free(buf);
    free(tmp);
    free(used);
    return 0;
}

/* All pairs of two CSR spectrum sets (row r = reference r, column q = query q),
 * OpenMP over reference rows with `nthreads` threads (<= 0: runtime default).
 * Outputs are row-major [R, Q]; any output pointer may be NULL. */
int cso_cosine_greedy_matrix(const double *rmz, const double *rin, const int64_t *roff, int64_t R,
                             const double *qmz, const double *qin, const int64_t *qoff, int64_t Q,
                             double tol, double shift, double mz_power, double int_power,
                             int64_t match_limit, double *score, int32_t *matches,
                             uint8_t *overflow, int64_t *n_candidates, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel
    {
        cso_pair *buf = NULL, *tmp = NULL;
        unsigned char *used = NULL;
        int64_t cap = 0, ucap = 0;
#pragma omp for schedule(dynamic, 1)
        for (int64_t r = 0; r < R; ++r) {
            for (int64_t q = 0; q < Q; ++q) {
                double s;
                int32_t m;
                int64_t c;
                score_pair(rmz + roff[r], rin + roff[r], roff[r + 1] - roff[r], qmz + qoff[q],
                           qin + qoff[q], qoff[q + 1] - qoff[q], tol, shift, mz_power, int_power,
                           &buf, &tmp, &cap, &used, &ucap, &s, &m, &c);
                const int64_t o = r * Q + q;
                if (score) score[o] = s;
                if (matches) matches[o] = m;
                if (overflow) overflow[o] = (uint8_t)(c >= match_limit);
                if (n_candidates) n_candidates[o] = c;
            }
        }
        free(buf);
        free(tmp);
        free(used);
    }
    return 0;
}

/* Greedy acceptance over candidates sorted ascending (stable), walked in
 * reverse: flags[k] = 1 iff sorted candidate k is accepted. */
static int32_t greedy_flags(const cso_pair *a, int64_t c, unsigned char *u1, unsigned char *u2,
                            unsigned char *flags) {
    int32_t m = 0;
    for (int64_t k = c - 1; k >= 0; --k) {
        flags[k] = 0;
        if (!u1[a[k].i] && !u2[a[k].j]) {
            u1[a[k].i] = 1;
            u2[a[k].j] = 1;
            flags[k] = 1;
            ++m;
        }
    }
    return m;
}

static int cmp_ij(const void *x, const void *y) {
    const cso_pair *a = (const cso_pair *)x, *b = (const cso_pair *)y;
    if (a->i != b->i) return a->i < b->i ? -1 : 1;
    if (a->j != b->j) return a->j < b->j ? -1 : 1;
    return 0;
}

/* Near-tie exposure of every pair: near_tie[r*Q + q] = 1 iff ordering the
 * candidates by their values rounded to float32 (as the reference GPU kernel
 * stores them, spectrum_similarity_functions.py:117,147; same stable sort and
 * reversal, i.e. ties to the later-collected candidate) makes the greedy accept
 * a different set of peak pairs than the float64 order of matchms.  These are
 * the pairs whose assignment hinges on values closer than float32 can tell
 * apart.  Returns the number of such pairs. */
int64_t cso_near_tie_pairs(const double *rmz, const double *rin, const int64_t *roff, int64_t R,
                           const double *qmz, const double *qin, const int64_t *qoff, int64_t Q,
                           double tol, double shift, double mz_power, double int_power,
                           uint8_t *near_tie, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int64_t total = 0;
#pragma omp parallel reduction(+ : total)
    {
        cso_pair *a = NULL, *b = NULL, *tmp = NULL;
        unsigned char *used = NULL, *fa = NULL, *fb = NULL;
        int64_t cap = 0, ucap = 0;
#pragma omp for schedule(dynamic, 1)
        for (int64_t r = 0; r < R; ++r) {
            const int64_t n1 = roff[r + 1] - roff[r];
            for (int64_t q = 0; q < Q; ++q) {
                const int64_t n2 = qoff[q + 1] - qoff[q];
                const double *m1 = rmz + roff[r], *i1 = rin + roff[r];
                const double *m2 = qmz + qoff[q], *i2 = qin + qoff[q];
                int64_t c = collect_peak_pairs(m1, i1, n1, m2, i2, n2, tol, shift, mz_power,
                                               int_power, NULL, 0);
                uint8_t flag = 0;
                if (c >= 2) {
                    if (c > cap) {
                        cap = c;
                        a = (cso_pair *)realloc(a, (size_t)c * sizeof(cso_pair));
                        b = (cso_pair *)realloc(b, (size_t)c * sizeof(cso_pair));
                        tmp = (cso_pair *)realloc(tmp, (size_t)c * sizeof(cso_pair));
                        fa = (unsigned char *)realloc(fa, (size_t)c);
                        fb = (unsigned char *)realloc(fb, (size_t)c);
                    }
                    if (n1 + n2 > ucap) {
                        ucap = n1 + n2;
                        used = (unsigned char *)realloc(used, (size_t)ucap);
                    }
                    collect_peak_pairs(m1, i1, n1, m2, i2, n2, tol, shift, mz_power, int_power,
                                       a, c);
                    memcpy(b, a, (size_t)c * sizeof(cso_pair));
                    for (int64_t k = 0; k < c; ++k) b[k].v = (double)(float)b[k].v;
                    stable_sort_ascending(a, tmp, c);
                    stable_sort_ascending(b, tmp, c);
                    memset(used, 0, (size_t)(n1 + n2));
                    greedy_flags(a, c, used, used + n1, fa);
                    memset(used, 0, (size_t)(n1 + n2));
                    greedy_flags(b, c, used, used + n1, fb);
                    /* compare the accepted sets: keep accepted (i, j), sorted */
                    int64_t na = 0, nb = 0;
                    for (int64_t k = 0; k < c; ++k)
                        if (fa[k]) a[na++] = a[k];
                    for (int64_t k = 0; k < c; ++k)
                        if (fb[k]) b[nb++] = b[k];
                    if (na != nb) {
                        flag = 1;
                    } else {
                        qsort(a, (size_t)na, sizeof(cso_pair), cmp_ij);
                        qsort(b, (size_t)nb, sizeof(cso_pair), cmp_ij);
                        for (int64_t k = 0; k < na && !flag; ++k)
                            flag = (a[k].i != b[k].i) || (a[k].j != b[k].j);
                    }
                }
                if (near_tie) near_tie[r * Q + q] = flag;
                total += flag;
            }
        }
        free(a);
        free(b);
        free(tmp);
        free(used);
        free(fa);
        free(fb);
    }
    return total;
}

/* Fingerprint measures on 0/1 byte fingerprints [rows, nbits], float64.
 * measure: 0 = jaccard, 1 = dice, 2 = cosine.  0/0 -> 0 (nan_to_num). */
int cso_fingerprint_matrix(const uint8_t *refs, int64_t R, const uint8_t *queries, int64_t Q,
                           int64_t nbits, int measure, double *out, int64_t *intersection,
                           int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < R; ++r) {
        const uint8_t *u = refs + r * nbits;
        int64_t a = 0;
        for (int64_t k = 0; k < nbits; ++k) a += (u[k] != 0);
        for (int64_t q = 0; q < Q; ++q) {
            const uint8_t *w = queries + q * nbits;
            int64_t b = 0, inter = 0;
            for (int64_t k = 0; k < nbits; ++k) {
                b += (w[k] != 0);
                inter += (u[k] != 0) & (w[k] != 0);
            }
            double s = 0.0;
            if (measure == 0) {            /* I / (a + b - I) */
                const int64_t uni = a + b - inter;
                s = uni ? (double)inter / (double)uni : 0.0;
            } else if (measure == 1) {     /* 2 * (I / (a + b)) */
                const int64_t den = a + b;
                s = den ? 2.0 * ((double)inter / (double)den) : 0.0;
            } else {                       /* I / sqrt(a * b) */
                s = (a && b) ? (double)inter / sqrt((double)a * (double)b) : 0.0;
            }
            out[r * Q + q] = s;
            if (intersection) intersection[r * Q + q] = inter;
        }
    }
    return 0;
}

/* Debug form of one pair: also returns the accepted (i, j) pairs in the order
 * the greedy accepted them (at most `cap`).  Used to compare against the
 * expected matched index lists of tests/similarity/test_cosine_greedy.py. */
int cso_cosine_greedy_pair_assign(const double *mz1, const double *in1, int64_t n1,
                                  const double *mz2, const double *in2, int64_t n2,
                                  double tol, double shift, double mz_power, double int_power,
                                  int32_t *acc_i, int32_t *acc_j, int64_t cap,
                                  double *score, int32_t *matches) {
    int64_t c = collect_peak_pairs(mz1, in1, n1, mz2, in2, n2, tol, shift, mz_power, int_power,
                                   NULL, 0);
    if (c == 0) {
        *score = 0.0;
        *matches = 0;
        return 0;
    }
    cso_pair *buf = (cso_pair *)malloc((size_t)c * sizeof(cso_pair));
    cso_pair *tmp = (cso_pair *)malloc((size_t)c * sizeof(cso_pair));
    unsigned char *used = (unsigned char *)calloc((size_t)(n1 + n2), 1);
    collect_peak_pairs(mz1, in1, n1, mz2, in2, n2, tol, shift, mz_power, int_power, buf, c);
    stable_sort_ascending(buf, tmp, c);
    double s = 0.0;
    int32_t m = 0;
    for (int64_t k = c - 1; k >= 0; --k) {
        const cso_pair p = buf[k];
        if (!used[p.i] && !used[n1 + p.j]) {
            s += p.v;
            used[p.i] = 1;
            used[n1 + p.j] = 1;
            if (m < cap) {
                acc_i[m] = p.i;
                acc_j[m] = p.j;
            }
            ++m;
        }
    }
    *score = s / (power_norm(mz1, in1, n1, mz_power, int_power) *
                  power_norm(mz2, in2, n2, mz_power, int_power));
    *matches = m;
    free(buf);
    free(tmp);
    free(used);
    return 0;
}
