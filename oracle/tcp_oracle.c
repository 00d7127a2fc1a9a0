/*
 * CPU oracle: a plain-C restatement of the TenCirPauli reference algorithms on
 * the matrix-free Pauli MVP path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only the test suite, the smoke check and the
 * benchmark's CPU-baseline leg load this library, and only as the checker or
 * as the timed CPU baseline.  The product (libtencirpauli_amd.so) never links,
 * loads or calls it.
 *
 * Every function names the reference file:line it restates (paths relative to
 * the reference repository; core/ = crates/tencir-pauli-core/src/).  The
 * reference is Rust and cannot be compiled here (no cargo/rustc, crates not
 * vendored), so this restatement is pinned by the reference's own known-answer
 * tests, by an independent dense Kronecker oracle and by closed forms.
 *
 * Conventions (core/word.rs:53-59, core/pauli_action.rs:65):
 *   codes 0=I 1=X 2=Y 3=Z; qubit q is bit (n-1-q) of a computational index
 *   (qubit 0 is the most significant bit).  Complex numbers are stored as
 *   interleaved (re, im) doubles.  n <= 63 for full-space paths, n <= 64 for
 *   sector paths (one packed limb).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Status codes shared with the product's C ABI (include/tencirpauli_amd.h). */
enum {
    ORC_OK = 0,
    ORC_INVALID_STRUCTURE_LENGTH = 1,
    ORC_INVALID_CODE = 2,
    ORC_NON_FINITE_COEFFICIENT = 3,
    ORC_MEMORY_LIMIT = 4,
    ORC_OVERFLOW = 5,
    ORC_INVALID_SECTOR = 6,
    ORC_SECTOR_LEAKAGE = 7,
    ORC_INVALID_INDEX = 8,
    ORC_INVALID_ARGUMENT = 9,
};

/* ------------------------------------------------------------------------ */
/* Complex helpers: textbook (a+bi)(c+di) as in num-complex 0.4.6.           */
/* ------------------------------------------------------------------------ */
typedef struct { double re, im; } cplx;

static inline cplx c_add(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cplx c_sub(cplx a, cplx b) { cplx r = {a.re - b.re, a.im - b.im}; return r; }
static inline cplx c_mul(cplx a, cplx b) {
    cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline cplx c_neg(cplx a) { cplx r = {-a.re, -a.im}; return r; }
static inline int c_is_zero(cplx a) { return a.re == 0.0 && a.im == 0.0; } /* core/scalar.rs:5-7 */

static inline uint64_t dbits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }

/* ------------------------------------------------------------------------ */
/* Canonicalisation: PauliOperator::from_terms (core/operator.rs:117-155).   */
/* Words ordered by code tuple from qubit 0 (core/word.rs:240-254); equal    */
/* words summed after sorting the contributions by their IEEE bit patterns;  */
/* only exact zeros dropped.                                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    const uint8_t *codes;
    uint32_t n;
} word_ctx;

static word_ctx g_word_ctx; /* qsort has no context argument; the oracle is single-threaded here */

static int cmp_word_index(const void *pa, const void *pb) {
    uint64_t a = *(const uint64_t *)pa, b = *(const uint64_t *)pb;
    const uint8_t *ca = g_word_ctx.codes + a * g_word_ctx.n;
    const uint8_t *cb = g_word_ctx.codes + b * g_word_ctx.n;
    for (uint32_t q = 0; q < g_word_ctx.n; ++q) {
        if (ca[q] != cb[q]) return ca[q] < cb[q] ? -1 : 1;
    }
    /* stable tie-break so equal words keep input order (order is irrelevant:
       contributions are re-sorted by bit pattern below) */
    return a < b ? -1 : (a > b);
}

static int cmp_value_bits(const void *pa, const void *pb) {
    const cplx *a = (const cplx *)pa, *b = (const cplx *)pb;
    uint64_t ar = dbits(a->re), br = dbits(b->re);
    if (ar != br) return ar < br ? -1 : 1;
    uint64_t ai = dbits(a->im), bi = dbits(b->im);
    if (ai != bi) return ai < bi ? -1 : 1;
    return 0;
}

/* Returns ORC_OK and writes *out_T canonical terms.  err_index receives the
   offending term (non-finite) or qubit (invalid code) index. */
int orc_canonicalize(uint32_t n, uint64_t T, const uint8_t *codes, const double *coef,
                     uint8_t *out_codes, double *out_coef, uint64_t *out_T, uint64_t *err_index) {
    *out_T = 0;
    for (uint64_t t = 0; t < T; ++t) {
        if (!isfinite(coef[2 * t]) || !isfinite(coef[2 * t + 1])) {
            if (err_index) *err_index = t;
            return ORC_NON_FINITE_COEFFICIENT;
        }
        for (uint32_t q = 0; q < n; ++q) {
            if (codes[t * n + q] > 3) {
                if (err_index) *err_index = q;
                return ORC_INVALID_CODE;
            }
        }
    }
    if (T == 0) return ORC_OK;
    uint64_t *idx = (uint64_t *)malloc(T * sizeof(uint64_t));
    cplx *vals = (cplx *)malloc(T * sizeof(cplx));
    for (uint64_t t = 0; t < T; ++t) idx[t] = t;
    g_word_ctx.codes = codes;
    g_word_ctx.n = n;
    qsort(idx, T, sizeof(uint64_t), cmp_word_index);
    uint64_t out = 0;
    uint64_t start = 0;
    while (start < T) {
        uint64_t end = start + 1;
        while (end < T && memcmp(codes + idx[end] * n, codes + idx[start] * n, n) == 0) ++end;
        uint64_t m = end - start;
        for (uint64_t j = 0; j < m; ++j) {
            vals[j].re = coef[2 * idx[start + j]];
            vals[j].im = coef[2 * idx[start + j] + 1];
        }
        qsort(vals, m, sizeof(cplx), cmp_value_bits);
        cplx sum = {0.0, 0.0};
        for (uint64_t j = 0; j < m; ++j) sum = c_add(sum, vals[j]);
        if (!c_is_zero(sum)) {
            memcpy(out_codes + out * n, codes + idx[start] * n, n);
            out_coef[2 * out] = sum.re;
            out_coef[2 * out + 1] = sum.im;
            ++out;
        }
        start = end;
    }
    free(idx);
    free(vals);
    *out_T = out;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Index-space action of one word: StatevectorPauliAction::from_packed       */
/* (core/pauli_action.rs:24-88) and matrix_term (core/hamiltonian.rs:564-571):*/
/* qubit q -> bit n-1-q; w = c * i^(#Y mod 4).                               */
/* ------------------------------------------------------------------------ */
void orc_matrix_terms(uint32_t n, uint64_t T, const uint8_t *codes, const double *coef,
                      uint64_t *x_mask, uint64_t *z_mask, double *w) {
    for (uint64_t t = 0; t < T; ++t) {
        uint64_t x = 0, z = 0;
        unsigned ycount = 0;
        for (uint32_t q = 0; q < n; ++q) {
            uint8_t c = codes[t * n + q];
            uint64_t bit = 1ull << (n - 1 - q);
            if (c == 1 || c == 2) x |= bit;
            if (c == 2 || c == 3) z |= bit;
            if (c == 2) ++ycount;
        }
        cplx phase;
        switch (ycount % 4) {
            case 0: phase.re = 1.0; phase.im = 0.0; break;
            case 1: phase.re = 0.0; phase.im = 1.0; break;
            case 2: phase.re = -1.0; phase.im = 0.0; break;
            default: phase.re = 0.0; phase.im = -1.0; break;
        }
        cplx c = {coef[2 * t], coef[2 * t + 1]};
        cplx r = c_mul(c, phase);
        x_mask[t] = x;
        z_mask[t] = z;
        w[2 * t] = r.re;
        w[2 * t + 1] = r.im;
    }
}

/* ------------------------------------------------------------------------ */
/* TermDirect apply: MvpPlan::apply_into (core/hamiltonian.rs:204-270).      */
/* Row-major gather; each row sums every canonical term in canonical order:  */
/*   col = row ^ x; c = w * v[col]; negate if popcount(z & col) odd.         */
/* The reference parallelises rows with rayon; here a pthread pool over      */
/* contiguous row blocks (row ownership, no numeric effect).                 */
/* ------------------------------------------------------------------------ */
static inline cplx termdirect_row(uint64_t row, uint64_t T, const uint64_t *xm, const uint64_t *zm,
                                  const double *w, const double *v) {
    cplx value = {0.0, 0.0};
    for (uint64_t t = 0; t < T; ++t) {
        uint64_t col = row ^ xm[t];
        cplx wt = {w[2 * t], w[2 * t + 1]};
        cplx s = {v[2 * col], v[2 * col + 1]};
        cplx contribution = c_mul(wt, s);
        if (__builtin_popcountll(zm[t] & col) & 1) contribution = c_neg(contribution);
        value = c_add(value, contribution);
    }
    return value;
}

typedef struct {
    uint64_t row_lo, row_hi, T;
    const uint64_t *xm, *zm, *rows;
    const double *w, *v;
    double *y;
} td_job;

static void *td_worker(void *arg) {
    td_job *j = (td_job *)arg;
    for (uint64_t r = j->row_lo; r < j->row_hi; ++r) {
        uint64_t row = j->rows ? j->rows[r] : r;
        cplx value = termdirect_row(row, j->T, j->xm, j->zm, j->w, j->v);
        j->y[2 * r] = value.re;
        j->y[2 * r + 1] = value.im;
    }
    return NULL;
}

static void td_run(uint64_t nrows, const uint64_t *rows, uint64_t T, const uint64_t *xm,
                   const uint64_t *zm, const double *w, const double *v, double *y, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if ((uint64_t)nthreads > nrows) nthreads = nrows ? (int)nrows : 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    td_job *jobs = (td_job *)malloc(sizeof(td_job) * nthreads);
    for (int i = 0; i < nthreads; ++i) {
        jobs[i].row_lo = nrows * i / nthreads;
        jobs[i].row_hi = nrows * (i + 1) / nthreads;
        jobs[i].T = T;
        jobs[i].xm = xm;
        jobs[i].zm = zm;
        jobs[i].rows = rows;
        jobs[i].w = w;
        jobs[i].v = v;
        jobs[i].y = y;
    }
    if (nthreads == 1) {
        td_worker(&jobs[0]);
    } else {
        for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, td_worker, &jobs[i]);
        for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    }
    free(th);
    free(jobs);
}

/* Full apply over all 2^n rows (output overwritten, core/tests.rs:7-23). */
int orc_termdirect_apply(uint32_t n, uint64_t T, const uint64_t *xm, const uint64_t *zm,
                         const double *w, const double *v, double *y, int nthreads) {
    if (n >= 64) return ORC_OVERFLOW; /* core/hamiltonian.rs:544-555 */
    td_run(1ull << n, NULL, T, xm, zm, w, v, y, nthreads);
    return ORC_OK;
}

/* Same gather for a chosen list of rows only (y_rows[r] = (H v)[rows[r]]). */
int orc_termdirect_rows(uint32_t n, uint64_t T, const uint64_t *xm, const uint64_t *zm,
                        const double *w, const double *v, const uint64_t *rows, uint64_t nrows,
                        double *y_rows, int nthreads) {
    if (n >= 64) return ORC_OVERFLOW;
    for (uint64_t r = 0; r < nrows; ++r)
        if (rows[r] >> n) return ORC_INVALID_INDEX;
    td_run(nrows, rows, T, xm, zm, w, v, y_rows, nthreads);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* U(1) sector: U1Sector (core/sector.rs:79-342, 1307-1362).                 */
/* Packed occupation word: qubit q at bit q (core/sector.rs:127-138).        */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t n, k, active;
    int complement;
    uint64_t dim;
    uint64_t *choose; /* (n+1) x (active+1), row-major C(row, col) */
} u1_sector;

/* choose_u64 (core/sector.rs:1342-1362) with overflow detection. */
static int choose_u64(uint32_t n, uint32_t k, uint64_t *out) {
    if (k > n) { *out = 0; return ORC_OK; }
    if (k > n - k) k = n - k;
    unsigned __int128 value = 1;
    for (uint32_t i = 1; i <= k; ++i) {
        value = value * (unsigned __int128)(n - k + i) / i;
        if (value > (unsigned __int128)UINT64_MAX) return ORC_OVERFLOW;
    }
    *out = (uint64_t)value;
    return ORC_OK;
}

/* U1Sector::new (core/sector.rs:79-99) + build_choose_table (:1307-1340). */
static int u1_new(uint32_t n, uint32_t k, u1_sector *s) {
    if (k > n) return ORC_INVALID_SECTOR;
    if (n > 64) return ORC_OVERFLOW;
    int st = choose_u64(n, k, &s->dim);
    if (st) return st;
    s->n = n;
    s->k = k;
    s->active = k < n - k ? k : n - k;
    s->complement = k > n - k;
    uint32_t width = s->active + 1;
    s->choose = (uint64_t *)calloc((size_t)(n + 1) * width, sizeof(uint64_t));
    for (uint32_t row = 0; row <= n; ++row) {
        s->choose[row * width] = 1;
        uint32_t limit = s->active < row ? s->active : row;
        for (uint32_t col = 1; col <= limit; ++col)
            s->choose[row * width + col] =
                col == row ? 1 : s->choose[(row - 1) * width + col - 1] + s->choose[(row - 1) * width + col];
    }
    return ORC_OK;
}

static void u1_free(u1_sector *s) { free(s->choose); }

/* U1Sector::choose_value (core/sector.rs:289-300). */
static uint64_t u1_choose(const u1_sector *s, uint32_t n, uint32_t k) {
    if (k > n) return 0;
    uint32_t norm = k < n - k ? k : n - k;
    return s->choose[n * (s->active + 1) + norm];
}

static inline uint64_t tail_mask(uint32_t n) { return n == 0 ? 0 : (n % 64 == 0 ? ~0ull : (1ull << n) - 1); }

/* rank_active_words (core/sector.rs:310-341) + rank_words_known_weight
   (:150-167): walk the active bits (particles, or holes when complemented)
   from qubit 0 and add C(n-p-1, needed). */
static int u1_rank_packed(const u1_sector *s, uint64_t word, uint64_t *rank) {
    if (word & ~tail_mask(s->n)) return ORC_INVALID_INDEX;
    if ((uint32_t)__builtin_popcountll(word) != s->k) return ORC_INVALID_INDEX;
    uint64_t bits = (s->complement ? ~word : word) & tail_mask(s->n);
    uint64_t r = 0;
    uint32_t selected = 0;
    while (bits) {
        uint32_t p = (uint32_t)__builtin_ctzll(bits);
        r += u1_choose(s, s->n - p - 1, s->active - selected);
        ++selected;
        bits &= bits - 1;
    }
    if (selected != s->active) return ORC_INVALID_INDEX;
    *rank = s->complement ? s->dim - 1 - r : r;
    return ORC_OK;
}

/* unrank_into (core/sector.rs:184-238). */
static int u1_unrank_packed(const u1_sector *s, uint64_t index, uint64_t *word) {
    if (index >= s->dim) return ORC_INVALID_INDEX;
    uint64_t a = s->complement ? s->dim - 1 - index : index;
    uint32_t remaining = s->active;
    uint64_t out = 0;
    for (uint32_t p = 0; p < s->n; ++p) {
        uint64_t zero_count = u1_choose(s, s->n - p - 1, remaining);
        if (a >= zero_count) {
            a -= zero_count;
            if (remaining == 0) return ORC_INVALID_INDEX;
            out |= 1ull << p;
            --remaining;
        }
    }
    if (remaining != 0 || a != 0) return ORC_INVALID_INDEX;
    if (s->complement) out = ~out & tail_mask(s->n);
    *word = out;
    return ORC_OK;
}

/* Integer (qubit 0 = MSB) <-> packed (qubit q at bit q) conversion,
   U1Sector::rank / unrank (core/sector.rs:127-138, 170-182). */
static uint64_t reverse_bits(uint64_t v, uint32_t n) {
    uint64_t r = 0;
    for (uint32_t q = 0; q < n; ++q)
        if ((v >> (n - 1 - q)) & 1) r |= 1ull << q;
    return r;
}

int orc_u1_dimension(uint32_t n, uint32_t k, uint64_t *dim) {
    if (k > n) return ORC_INVALID_SECTOR;
    return choose_u64(n, k, dim);
}

int orc_u1_rank(uint32_t n, uint32_t k, uint64_t bitstring, uint64_t *rank) {
    u1_sector s;
    int st = u1_new(n, k, &s);
    if (st) return st;
    if (n < 64 && (bitstring >> n)) { u1_free(&s); return ORC_INVALID_INDEX; }
    st = u1_rank_packed(&s, reverse_bits(bitstring, n), rank);
    u1_free(&s);
    return st;
}

int orc_u1_unrank(uint32_t n, uint32_t k, uint64_t index, uint64_t *bitstring) {
    u1_sector s;
    int st = u1_new(n, k, &s);
    if (st) return st;
    uint64_t word;
    st = u1_unrank_packed(&s, index, &word);
    if (!st) *bitstring = reverse_bits(word, n);
    u1_free(&s);
    return st;
}

/* Batched rank of many integers (bit-exact sweeps). */
int orc_u1_rank_many(uint32_t n, uint32_t k, const uint64_t *bitstrings, uint64_t count, uint64_t *ranks) {
    u1_sector s;
    int st = u1_new(n, k, &s);
    if (st) return st;
    for (uint64_t i = 0; i < count && !st; ++i)
        st = u1_rank_packed(&s, reverse_bits(bitstrings[i], n), &ranks[i]);
    u1_free(&s);
    return st;
}

int orc_u1_unrank_many(uint32_t n, uint32_t k, const uint64_t *indices, uint64_t count, uint64_t *bitstrings) {
    u1_sector s;
    int st = u1_new(n, k, &s);
    if (st) return st;
    for (uint64_t i = 0; i < count && !st; ++i) {
        uint64_t word;
        st = u1_unrank_packed(&s, indices[i], &word);
        if (!st) bitstrings[i] = reverse_bits(word, n);
    }
    u1_free(&s);
    return st;
}

/* ------------------------------------------------------------------------ */
/* U(1) lazy restricted apply: U1LazyMvpPlan::apply_into                     */
/* (core/sector.rs:579-642) with compile_terms (:926-1021), weighted         */
/* coefficients (:1034-1049) and aggregate_source (:1051-1127).              */
/* Source-major serial scatter: zero the output; for each source state (in   */
/* restricted order) and each x-group (first-appearance order), sum +-w over */
/* the group's terms by the parity of z & occupied; skip exact zeros; a      */
/* nonzero value whose destination has the wrong weight is SectorLeakage;    */
/* output[rank(src ^ x)] += value * state[src].                              */
/* Input: canonical terms as codes.                                          */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint64_t x;          /* packed, qubit q at bit q */
    uint64_t term_start, term_end;
} u1_group;

static void packed_terms(uint32_t n, uint64_t T, const uint8_t *codes, const double *coef,
                         uint64_t *xq, uint64_t *zq, cplx *w) {
    for (uint64_t t = 0; t < T; ++t) {
        uint64_t x = 0, z = 0;
        unsigned y = 0;
        for (uint32_t q = 0; q < n; ++q) {
            uint8_t c = codes[t * n + q];
            if (c == 1 || c == 2) x |= 1ull << q;
            if (c == 2 || c == 3) z |= 1ull << q;
            if (c == 2) ++y;
        }
        cplx ph;
        switch (y % 4) {
            case 0: ph.re = 1; ph.im = 0; break;
            case 1: ph.re = 0; ph.im = 1; break;
            case 2: ph.re = -1; ph.im = 0; break;
            default: ph.re = 0; ph.im = -1; break;
        }
        cplx c = {coef[2 * t], coef[2 * t + 1]};
        xq[t] = x;
        zq[t] = z;
        w[t] = c_mul(c, ph);
    }
}

/* Sources [src_begin, src_end) only (a bounded sample for timing; the full
   apply is the range [0, dim)).  The output is zero-filled in full. */
int orc_u1_lazy_apply_range(uint32_t n, uint32_t k, uint64_t T, const uint8_t *codes, const double *coef,
                            const double *state, double *output, uint64_t src_begin, uint64_t src_end,
                            uint64_t *leak_source, uint32_t *leak_weight) {
    u1_sector s;
    int st = u1_new(n, k, &s);
    if (st) return st;
    if (src_end > s.dim) src_end = s.dim;
    uint64_t *xq = (uint64_t *)malloc((T + 1) * sizeof(uint64_t));
    uint64_t *zq = (uint64_t *)malloc((T + 1) * sizeof(uint64_t));
    cplx *w = (cplx *)malloc((T + 1) * sizeof(cplx));
    packed_terms(n, T, codes, coef, xq, zq, w);
    /* compile_terms: groups in first-appearance order; terms of one group keep
       canonical order (stable counting placement, core/sector.rs:926-1021). */
    u1_group *groups = (u1_group *)malloc((T + 1) * sizeof(u1_group));
    uint64_t *tg = (uint64_t *)malloc((T + 1) * sizeof(uint64_t));
    uint64_t G = 0;
    for (uint64_t t = 0; t < T; ++t) {
        uint64_t g;
        for (g = 0; g < G; ++g)
            if (groups[g].x == xq[t]) break;
        if (g == G) {
            groups[G].x = xq[t];
            groups[G].term_start = 0;
            groups[G].term_end = 0; /* used as count */
            ++G;
        }
        tg[t] = g;
        groups[g].term_end++;
    }
    uint64_t off = 0;
    uint64_t *next = (uint64_t *)malloc((G + 1) * sizeof(uint64_t));
    for (uint64_t g = 0; g < G; ++g) {
        uint64_t cnt = groups[g].term_end;
        groups[g].term_start = off;
        off += cnt;
        groups[g].term_end = off;
        next[g] = groups[g].term_start;
    }
    uint64_t *gz = (uint64_t *)malloc((T + 1) * sizeof(uint64_t));
    cplx *gw = (cplx *)malloc((T + 1) * sizeof(cplx));
    for (uint64_t t = 0; t < T; ++t) {
        uint64_t pos = next[tg[t]]++;
        gz[pos] = zq[t];
        gw[pos] = w[t];
    }
    for (uint64_t i = 0; i < 2 * s.dim; ++i) output[i] = 0.0;
    uint64_t src_word = 0;
    for (uint64_t src = src_begin; src < src_end && !st; ++src) {
        st = u1_unrank_packed(&s, src, &src_word);
        if (st) break;
        cplx sv = {state[2 * src], state[2 * src + 1]};
        for (uint64_t g = 0; g < G; ++g) {
            cplx value = {0.0, 0.0};
            for (uint64_t t = groups[g].term_start; t < groups[g].term_end; ++t) {
                if (__builtin_popcountll(gz[t] & src_word) & 1)
                    value = c_sub(value, gw[t]);
                else
                    value = c_add(value, gw[t]);
            }
            if (c_is_zero(value)) continue;
            uint64_t dst_word = src_word ^ groups[g].x;
            uint32_t weight = (uint32_t)__builtin_popcountll(dst_word);
            if (weight != k) {
                if (leak_source) *leak_source = src;
                if (leak_weight) *leak_weight = weight;
                st = ORC_SECTOR_LEAKAGE;
                break;
            }
            uint64_t dst;
            st = u1_rank_packed(&s, dst_word, &dst);
            if (st) break;
            cplx contrib = c_mul(value, sv);
            output[2 * dst] += contrib.re;
            output[2 * dst + 1] += contrib.im;
        }
    }
    free(xq); free(zq); free(w); free(groups); free(tg); free(next); free(gz); free(gw);
    u1_free(&s);
    return st;
}

int orc_u1_lazy_apply(uint32_t n, uint32_t k, uint64_t T, const uint8_t *codes, const double *coef,
                      const double *state, double *output, uint64_t *leak_source, uint32_t *leak_weight) {
    return orc_u1_lazy_apply_range(n, k, T, codes, coef, state, output, 0, UINT64_MAX, leak_source, leak_weight);
}

/* ------------------------------------------------------------------------ */
/* Contiguous fixed-population blocks (the factorized charge-sector layout   */
/* with two or more components): FixedPopulationIndex rank/unrank of a       */
/* contiguous binary word (core/charge_sector.rs:973-1038) and the mixed-    */
/* radix strides (core/charge_sector.rs:193-217, 437-456): block 0 covers    */
/* qubits [0, s0), block 1 the next s1 qubits, ...; the last block has       */
/* stride 1.  A block's word has its first qubit as the MSB, i.e. the block  */
/* is a field of the integer index (qubit 0 = MSB).                          */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t nblocks;
    uint32_t size[64], pop[64], shift[64];
    uint64_t dim[64], stride[64];
    uint64_t total_dim;
    uint32_t n;
} blocks_sector;

static int blocks_new(uint32_t nblocks, const uint32_t *sizes, const uint32_t *pops, blocks_sector *b) {
    if (nblocks == 0 || nblocks > 64) return ORC_INVALID_SECTOR;
    uint32_t n = 0;
    for (uint32_t j = 0; j < nblocks; ++j) {
        if (pops[j] > sizes[j]) return ORC_INVALID_SECTOR;
        n += sizes[j];
    }
    if (n > 64) return ORC_OVERFLOW;
    b->nblocks = nblocks;
    b->n = n;
    uint32_t shift = n;
    for (uint32_t j = 0; j < nblocks; ++j) {
        b->size[j] = sizes[j];
        b->pop[j] = pops[j];
        shift -= sizes[j];
        b->shift[j] = shift;
        int st = choose_u64(sizes[j], pops[j], &b->dim[j]);
        if (st) return st;
    }
    unsigned __int128 stride = 1;
    for (int j = (int)nblocks - 1; j >= 0; --j) {
        if (stride > UINT64_MAX) return ORC_OVERFLOW;
        b->stride[j] = (uint64_t)stride;
        stride *= b->dim[j];
    }
    if (stride > UINT64_MAX) return ORC_OVERFLOW;
    b->total_dim = (uint64_t)stride;
    return ORC_OK;
}

/* rank_contiguous_binary_word (core/charge_sector.rs:973-1007). */
static int block_rank(uint32_t sites, uint32_t population, uint64_t word, uint64_t *rank) {
    if (sites < 64 && (word >> sites)) return ORC_INVALID_INDEX;
    uint32_t needed = population;
    unsigned __int128 r = 0;
    for (uint32_t axis = 0; axis < sites; ++axis) {
        uint64_t bit = 1ull << (sites - axis - 1);
        if (!(word & bit)) continue;
        if (needed == 0) return ORC_INVALID_SECTOR;
        uint64_t c;
        int st = choose_u64(sites - axis - 1, needed, &c);
        if (st) return st;
        r += c;
        --needed;
    }
    if (needed != 0) return ORC_INVALID_SECTOR;
    *rank = (uint64_t)r;
    return ORC_OK;
}

/* unrank_contiguous_binary_word (core/charge_sector.rs:1009-1038). */
static int block_unrank(uint32_t sites, uint32_t population, uint64_t dim, uint64_t rank, uint64_t *word) {
    if (rank >= dim) return ORC_INVALID_INDEX;
    uint64_t w = 0;
    uint64_t remaining = rank;
    uint32_t needed = population;
    for (uint32_t axis = 0; axis < sites; ++axis) {
        uint64_t zero_count;
        int st = choose_u64(sites - axis - 1, needed, &zero_count);
        if (st) return st;
        if (remaining < zero_count) continue;
        remaining -= zero_count;
        w |= 1ull << (sites - axis - 1);
        if (needed == 0) return ORC_INVALID_SECTOR;
        --needed;
    }
    *word = w;
    return ORC_OK;
}

static inline uint64_t field(uint64_t v, uint32_t shift, uint32_t size) {
    return size >= 64 ? v : (v >> shift) & ((1ull << size) - 1);
}

static int blocks_rank(const blocks_sector *b, uint64_t bitstring, uint64_t *rank) {
    unsigned __int128 r = 0;
    for (uint32_t j = 0; j < b->nblocks; ++j) {
        uint64_t local;
        int st = block_rank(b->size[j], b->pop[j], field(bitstring, b->shift[j], b->size[j]), &local);
        if (st) return st;
        r += (unsigned __int128)local * b->stride[j];
    }
    *rank = (uint64_t)r;
    return ORC_OK;
}

static int blocks_unrank(const blocks_sector *b, uint64_t index, uint64_t *bitstring) {
    if (index >= b->total_dim) return ORC_INVALID_INDEX;
    uint64_t v = 0;
    for (uint32_t j = 0; j < b->nblocks; ++j) {
        uint64_t local = index / b->stride[j];
        index %= b->stride[j];
        uint64_t w;
        int st = block_unrank(b->size[j], b->pop[j], b->dim[j], local, &w);
        if (st) return st;
        v |= b->size[j] >= 64 ? w : (w << b->shift[j]);
    }
    *bitstring = v;
    return ORC_OK;
}

int orc_blocks_dimension(uint32_t nblocks, const uint32_t *sizes, const uint32_t *pops, uint64_t *dim) {
    blocks_sector b;
    int st = blocks_new(nblocks, sizes, pops, &b);
    if (!st) *dim = b.total_dim;
    return st;
}

int orc_blocks_rank_many(uint32_t nblocks, const uint32_t *sizes, const uint32_t *pops,
                         const uint64_t *bitstrings, uint64_t count, uint64_t *ranks) {
    blocks_sector b;
    int st = blocks_new(nblocks, sizes, pops, &b);
    for (uint64_t i = 0; i < count && !st; ++i) st = blocks_rank(&b, bitstrings[i], &ranks[i]);
    return st;
}

int orc_blocks_unrank_many(uint32_t nblocks, const uint32_t *sizes, const uint32_t *pops,
                           const uint64_t *indices, uint64_t count, uint64_t *bitstrings) {
    blocks_sector b;
    int st = blocks_new(nblocks, sizes, pops, &b);
    for (uint64_t i = 0; i < count && !st; ++i) st = blocks_unrank(&b, indices[i], &bitstrings[i]);
    return st;
}

/* ------------------------------------------------------------------------ */
/* Generic restricted charge apply: apply_charge_mvp_from_prepared_plan_into */
/* (core/charge/generic.rs:949-1045, the aggregate branch that a non-        */
/* termwise-conserved Pauli operator takes), with the per-axis schedule of   */
/* apply_compiled_charge_schedule (core/charge/schedule.rs:645-726: X flips; */
/* Y flips with +i on bit 0 and -i on bit 1; Z negates on bit 1) and the     */
/* destination accumulator (schedule.rs:273-372: equal destinations summed   */
/* in schedule order).  Schedules are the canonical terms in order (the      */
/* BTreeMap over code tuples in core/charge/types.rs:31-80 sorts them the    */
/* same way).  Exact-zero sums are skipped; a nonzero value whose            */
/* destination is outside the sector is InvalidSector (leakage).            */
/* ------------------------------------------------------------------------ */
/* The sector is given by its rank/unrank pair (the plan's rank authority). */
typedef struct {
    const void *ctx;
    uint32_t n;
    uint64_t dim;
    int (*rank)(const void *ctx, uint64_t bitstring, uint64_t *rank);
    int (*unrank)(const void *ctx, uint64_t index, uint64_t *bitstring);
} sector_ops;

/* Sources [col_begin, col_end) only (a bounded sample for timing). */
static int generic_apply_range(const sector_ops *sec, uint64_t T, const uint8_t *codes, const double *coef,
                               const double *state, double *output, uint64_t col_begin, uint64_t col_end,
                               uint64_t *leak_source) {
    int st = ORC_OK;
    if (col_end > sec->dim) col_end = sec->dim;
    uint32_t n = sec->n;
    uint64_t *keys = (uint64_t *)malloc((T + 1) * sizeof(uint64_t));
    cplx *vals = (cplx *)malloc((T + 1) * sizeof(cplx));
    for (uint64_t i = 0; i < 2 * sec->dim; ++i) output[i] = 0.0;
    for (uint64_t col = col_begin; col < col_end && !st; ++col) {
        uint64_t src;
        st = sec->unrank(sec->ctx, col, &src);
        if (st) break;
        cplx sv = {state[2 * col], state[2 * col + 1]};
        uint64_t nk = 0;
        for (uint64_t t = 0; t < T; ++t) {
            cplx value = {coef[2 * t], coef[2 * t + 1]};
            uint64_t dst = src;
            for (uint32_t axis = 0; axis < n; ++axis) {
                uint64_t bitpos = (uint64_t)(n - 1 - axis);
                unsigned bit = (unsigned)((dst >> bitpos) & 1);
                switch (codes[t * n + axis]) {
                    case 0: break;
                    case 1: dst ^= 1ull << bitpos; break;
                    case 2: {
                        dst ^= 1ull << bitpos;
                        cplx ph = {0.0, bit == 0 ? 1.0 : -1.0};
                        value = c_mul(value, ph);
                        break;
                    }
                    default:
                        if (bit == 1) value = c_neg(value);
                        break;
                }
            }
            if (c_is_zero(value)) continue;
            uint64_t j;
            for (j = 0; j < nk; ++j)
                if (keys[j] == dst) break;
            if (j == nk) {
                keys[nk] = dst;
                vals[nk] = value;
                ++nk;
            } else {
                vals[j] = c_add(vals[j], value);
            }
        }
        for (uint64_t j = 0; j < nk; ++j) {
            if (c_is_zero(vals[j])) continue;
            uint64_t row;
            int rs = sec->rank(sec->ctx, keys[j], &row);
            if (rs) {
                if (leak_source) *leak_source = col;
                st = ORC_INVALID_SECTOR;
                break;
            }
            cplx contrib = c_mul(vals[j], sv);
            output[2 * row] += contrib.re;
            output[2 * row + 1] += contrib.im;
        }
    }
    free(keys);
    free(vals);
    return st;
}

static int blocks_rank_cb(const void *ctx, uint64_t bitstring, uint64_t *rank) {
    return blocks_rank((const blocks_sector *)ctx, bitstring, rank);
}
static int blocks_unrank_cb(const void *ctx, uint64_t index, uint64_t *bitstring) {
    return blocks_unrank((const blocks_sector *)ctx, index, bitstring);
}

int orc_blocks_generic_apply_range(uint32_t nblocks, const uint32_t *sizes, const uint32_t *pops,
                                   uint64_t T, const uint8_t *codes, const double *coef,
                                   const double *state, double *output, uint64_t col_begin, uint64_t col_end,
                                   uint64_t *leak_source) {
    blocks_sector b;
    int st = blocks_new(nblocks, sizes, pops, &b);
    if (st) return st;
    sector_ops sec = {&b, b.n, b.total_dim, blocks_rank_cb, blocks_unrank_cb};
    return generic_apply_range(&sec, T, codes, coef, state, output, col_begin, col_end, leak_source);
}

int orc_blocks_generic_apply(uint32_t nblocks, const uint32_t *sizes, const uint32_t *pops,
                             uint64_t T, const uint8_t *codes, const double *coef,
                             const double *state, double *output, uint64_t *leak_source) {
    return orc_blocks_generic_apply_range(nblocks, sizes, pops, T, codes, coef, state, output, 0, UINT64_MAX,
                                          leak_source);
}

/* ------------------------------------------------------------------------ */
/* General charge sector: ChargeSectorPlan (core/charge_sector.rs:8-16) as   */
/* built by build_charge_sector_plan (:1620-1722) for binary axes.  Axis a   */
/* (bit n-1-a) has the two candidate contributions levels[j][a][0|1] per     */
/* constraint j.  suffix[p] maps a remainder vector to the number of         */
/* assignments of axes p..n-1 whose contributions sum to it: suffix[n] =     */
/* {0: 1}; suffix[p][c + r] += suffix[p+1][r] for every candidate c of axis  */
/* p.  The reference keys hash maps; here each table is an array sorted by   */
/* remainder and searched by bisection (the lookup order is not observable). */
/* rank (:1442-1503): for each axis, the counts of the smaller candidate     */
/* (remaining - c0) are added when the axis is set; the final remainder must */
/* be zero (InvalidSector).  unrank (:1529-1578): candidate 0 is taken when  */
/* the index is below its count, else the index drops by that count.        */
/* ------------------------------------------------------------------------ */
#define ORC_MAX_CHARGES 4
typedef struct {
    int64_t key[ORC_MAX_CHARGES];
    uint64_t count;
} charge_entry;

typedef struct {
    uint32_t n, m;
    int64_t *levels; /* [m][n][2] */
    int64_t target[ORC_MAX_CHARGES];
    charge_entry **suffix; /* [n + 1] tables */
    uint64_t *len;         /* [n + 1] */
    uint64_t dim;
} charge_sector;

static int cmp_charge_entry(const void *pa, const void *pb) {
    const charge_entry *a = (const charge_entry *)pa, *b = (const charge_entry *)pb;
    for (int j = 0; j < ORC_MAX_CHARGES; ++j)
        if (a->key[j] != b->key[j]) return a->key[j] < b->key[j] ? -1 : 1;
    return 0;
}

static uint64_t charge_lookup(const charge_sector *s, uint32_t p, const int64_t *key) {
    charge_entry probe;
    memset(&probe, 0, sizeof probe);
    for (uint32_t j = 0; j < s->m; ++j) probe.key[j] = key[j];
    const charge_entry *hit =
        (const charge_entry *)bsearch(&probe, s->suffix[p], s->len[p], sizeof(charge_entry), cmp_charge_entry);
    return hit ? hit->count : 0;
}

static void charge_free(charge_sector *s) {
    if (s->suffix)
        for (uint32_t p = 0; p <= s->n; ++p) free(s->suffix[p]);
    free(s->suffix);
    free(s->len);
    free(s->levels);
}

static int charge_new(uint32_t n, uint32_t m, const int64_t *levels, const int64_t *targets, charge_sector *s) {
    memset(s, 0, sizeof *s);
    if (m == 0 || m > ORC_MAX_CHARGES || n > 63) return ORC_INVALID_SECTOR;
    s->n = n;
    s->m = m;
    s->levels = (int64_t *)malloc(sizeof(int64_t) * 2 * n * m + 1);
    memcpy(s->levels, levels, sizeof(int64_t) * 2 * n * m);
    for (uint32_t j = 0; j < m; ++j) s->target[j] = targets[j];
    s->suffix = (charge_entry **)calloc(n + 1, sizeof(charge_entry *));
    s->len = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
    s->suffix[n] = (charge_entry *)calloc(1, sizeof(charge_entry));
    s->len[n] = 1;
    s->suffix[n][0].count = 1;
    for (int p = (int)n - 1; p >= 0; --p) {
        const uint64_t prev = s->len[p + 1];
        charge_entry *t = (charge_entry *)malloc(sizeof(charge_entry) * 2 * prev + 1);
        uint64_t k = 0;
        for (uint32_t v = 0; v < 2; ++v)
            for (uint64_t i = 0; i < prev; ++i) {
                t[k] = s->suffix[p + 1][i];
                for (uint32_t j = 0; j < m; ++j) t[k].key[j] += levels[((uint64_t)j * n + p) * 2 + v];
                ++k;
            }
        qsort(t, k, sizeof(charge_entry), cmp_charge_entry);
        uint64_t w = 0;
        for (uint64_t i = 0; i < k; ++i) {
            if (w && cmp_charge_entry(&t[w - 1], &t[i]) == 0) {
                if (t[w - 1].count > UINT64_MAX - t[i].count) {
                    free(t);
                    charge_free(s);
                    return ORC_OVERFLOW;
                }
                t[w - 1].count += t[i].count;
            } else {
                t[w++] = t[i];
            }
        }
        s->suffix[p] = t;
        s->len[p] = w;
    }
    s->dim = charge_lookup(s, 0, s->target);
    return ORC_OK;
}

static int charge_rank(const charge_sector *s, uint64_t bitstring, uint64_t *rank) {
    const uint32_t n = s->n, m = s->m;
    if (n < 64 && (bitstring >> n)) return ORC_INVALID_INDEX;
    int64_t rem[ORC_MAX_CHARGES], cand[ORC_MAX_CHARGES];
    for (uint32_t j = 0; j < m; ++j) rem[j] = s->target[j];
    uint64_t r = 0;
    for (uint32_t a = 0; a < n; ++a) {
        const uint32_t value = (uint32_t)((bitstring >> (n - 1 - a)) & 1);
        if (value) {
            for (uint32_t j = 0; j < m; ++j) cand[j] = rem[j] - s->levels[((uint64_t)j * n + a) * 2 + 0];
            r += charge_lookup(s, a + 1, cand);
        }
        for (uint32_t j = 0; j < m; ++j) rem[j] -= s->levels[((uint64_t)j * n + a) * 2 + value];
    }
    for (uint32_t j = 0; j < m; ++j)
        if (rem[j] != 0) return ORC_INVALID_SECTOR;
    *rank = r;
    return ORC_OK;
}

static int charge_unrank(const charge_sector *s, uint64_t index, uint64_t *bitstring) {
    const uint32_t n = s->n, m = s->m;
    if (index >= s->dim) return ORC_INVALID_INDEX;
    int64_t rem[ORC_MAX_CHARGES], cand[ORC_MAX_CHARGES];
    for (uint32_t j = 0; j < m; ++j) rem[j] = s->target[j];
    uint64_t word = 0;
    for (uint32_t a = 0; a < n; ++a) {
        int selected = -1;
        for (uint32_t v = 0; v < 2 && selected < 0; ++v) {
            for (uint32_t j = 0; j < m; ++j) cand[j] = rem[j] - s->levels[((uint64_t)j * n + a) * 2 + v];
            const uint64_t c = charge_lookup(s, a + 1, cand);
            if (index < c) selected = (int)v;
            else index -= c;
        }
        if (selected < 0) return ORC_INVALID_SECTOR;
        if (selected) word |= 1ull << (n - 1 - a);
        for (uint32_t j = 0; j < m; ++j) rem[j] = cand[j];
    }
    *bitstring = word;
    return ORC_OK;
}

static int charge_rank_cb(const void *ctx, uint64_t bitstring, uint64_t *rank) {
    return charge_rank((const charge_sector *)ctx, bitstring, rank);
}
static int charge_unrank_cb(const void *ctx, uint64_t index, uint64_t *bitstring) {
    return charge_unrank((const charge_sector *)ctx, index, bitstring);
}

int orc_charge_dimension(uint32_t n, uint32_t m, const int64_t *levels, const int64_t *targets, uint64_t *dim) {
    charge_sector s;
    int st = charge_new(n, m, levels, targets, &s);
    if (st) return st;
    *dim = s.dim;
    charge_free(&s);
    return ORC_OK;
}

/* ranks[i] = UINT64_MAX where bitstrings[i] is outside the sector */
int orc_charge_rank_many(uint32_t n, uint32_t m, const int64_t *levels, const int64_t *targets,
                         const uint64_t *bitstrings, uint64_t count, uint64_t *ranks) {
    charge_sector s;
    int st = charge_new(n, m, levels, targets, &s);
    if (st) return st;
    for (uint64_t i = 0; i < count && !st; ++i) {
        st = charge_rank(&s, bitstrings[i], &ranks[i]);
        if (st == ORC_INVALID_SECTOR) {
            ranks[i] = UINT64_MAX;
            st = ORC_OK;
        }
    }
    charge_free(&s);
    return st;
}

int orc_charge_unrank_many(uint32_t n, uint32_t m, const int64_t *levels, const int64_t *targets,
                           const uint64_t *indices, uint64_t count, uint64_t *bitstrings) {
    charge_sector s;
    int st = charge_new(n, m, levels, targets, &s);
    if (st) return st;
    for (uint64_t i = 0; i < count && !st; ++i) st = charge_unrank(&s, indices[i], &bitstrings[i]);
    charge_free(&s);
    return st;
}

int orc_charge_generic_apply_range(uint32_t n, uint32_t m, const int64_t *levels, const int64_t *targets,
                                   uint64_t T, const uint8_t *codes, const double *coef, const double *state,
                                   double *output, uint64_t col_begin, uint64_t col_end, uint64_t *leak_source) {
    charge_sector s;
    int st = charge_new(n, m, levels, targets, &s);
    if (st) return st;
    sector_ops sec = {&s, n, s.dim, charge_rank_cb, charge_unrank_cb};
    st = generic_apply_range(&sec, T, codes, coef, state, output, col_begin, col_end, leak_source);
    charge_free(&s);
    return st;
}

/* ------------------------------------------------------------------------ */
/* Eager restricted apply: CsrGraph::apply_into (core/csr.rs:194-238).  Each */
/* output row is the row's entries in stored order, accumulated with the     */
/* ManualRealImag arithmetic (re += a.re b.re - a.im b.im, im += a.re b.im + */
/* a.im b.re); rows split over threads as rayon's par_windows does.  Rows    */
/* [row_begin, row_end) only (a bounded sample; the others are untouched).   */
/* ------------------------------------------------------------------------ */
typedef struct {
    const uint64_t *indptr, *cols;
    const double *vals, *state;
    double *out;
    uint64_t lo, hi;
} csr_job;

static void *csr_rows(void *arg) {
    const csr_job *j = (const csr_job *)arg;
    for (uint64_t r = j->lo; r < j->hi; ++r) {
        double re = 0.0, im = 0.0;
        for (uint64_t k = j->indptr[r]; k < j->indptr[r + 1]; ++k) {
            const double cr = j->vals[2 * k], ci = j->vals[2 * k + 1];
            const uint64_t c = j->cols[k];
            const double sr = j->state[2 * c], si = j->state[2 * c + 1];
            re += cr * sr - ci * si;
            im += cr * si + ci * sr;
        }
        j->out[2 * r] = re;
        j->out[2 * r + 1] = im;
    }
    return NULL;
}

int orc_csr_apply(uint64_t dim, const uint64_t *indptr, const uint64_t *cols, const double *vals,
                  const double *state, double *output, uint64_t row_begin, uint64_t row_end, int nthreads) {
    if (row_end > dim) row_end = dim;
    if (row_begin > row_end) row_begin = row_end;
    const uint64_t rows = row_end - row_begin;
    if (nthreads < 1) nthreads = 1;
    if ((uint64_t)nthreads > rows) nthreads = rows ? (int)rows : 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    csr_job *jobs = (csr_job *)malloc(sizeof(csr_job) * nthreads);
    for (int i = 0; i < nthreads; ++i) {
        jobs[i] = (csr_job){indptr, cols, vals, state, output, row_begin + rows * i / nthreads,
                            row_begin + rows * (i + 1) / nthreads};
        pthread_create(&th[i], NULL, csr_rows, &jobs[i]);
    }
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th);
    free(jobs);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Exact charge-conservation analysis: PauliOperator::analyze_charge         */
/* (core/operator.rs:396-445).  For every X/Y letter on a qubit whose two    */
/* charge levels differ, the commutator [H, Q] contributes the word with     */
/* that letter swapped (X<->Y) and coefficient c * i * (+-difference);       */
/* contributions to equal words are summed in IEEE-bit-pattern order and    */
/* the number of nonzero aggregated words is returned.                      */
/* levels: 2*n doubles, (level of |0>, level of |1>) per qubit.              */
/* ------------------------------------------------------------------------ */
int orc_analyze_charge(uint32_t n, uint64_t T, const uint8_t *codes, const double *coef,
                       const double *levels, uint64_t *nonzero) {
    uint64_t cap = T * (uint64_t)n + 1;
    uint8_t *wcodes = (uint8_t *)malloc(cap * (n ? n : 1));
    double *wcoef = (double *)malloc(cap * 2 * sizeof(double));
    uint64_t m = 0;
    for (uint64_t t = 0; t < T; ++t) {
        for (uint32_t q = 0; q < n; ++q) {
            uint8_t c = codes[t * n + q];
            if (c != 1 && c != 2) continue;
            double difference = levels[2 * q] - levels[2 * q + 1];
            if (difference == 0.0) continue;
            memcpy(wcodes + m * n, codes + t * n, n);
            wcodes[m * n + q] = c == 1 ? 2 : 1;
            double scale = c == 1 ? -difference : difference;
            cplx cc = {coef[2 * t], coef[2 * t + 1]};
            cplx s = {0.0, scale};
            cplx r = c_mul(cc, s);
            wcoef[2 * m] = r.re;
            wcoef[2 * m + 1] = r.im;
            ++m;
        }
    }
    uint8_t *oc = (uint8_t *)malloc((m + 1) * (n ? n : 1));
    double *of = (double *)malloc((m + 1) * 2 * sizeof(double));
    uint64_t outT = 0;
    int st = orc_canonicalize(n, m, wcodes, wcoef, oc, of, &outT, NULL);
    *nonzero = outT;
    free(wcodes); free(wcoef); free(oc); free(of);
    return st;
}
