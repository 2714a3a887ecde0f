/*
 * gmfa_oracle.c -- CPU restatement of the reference's classical alignment core.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke entry
 * point and bench.py's cpu_baseline leg may load it; the product
 * (graphmambaformer_amd) never does.  It is plain scalar C, written for
 * obviousness, in the same row / anchor order as the reference.
 *
 * What it restates (citations are files of sarakh1999/GraphMambaFormer):
 *   oracle_sw_banded      graphmambaformer/alignment/extension.py:78-204
 *                         (banded_affine_sw_batch, incl. the batch-wide x-drop
 *                          break at :190-193 and the first-strict-row /
 *                          first-column best cell at :178-182)
 *   oracle_sw_traceback   graphmambaformer/alignment/extension.py:207-302
 *                         (traceback_banded, incl. cell() = -inf outside the
 *                          stored band at :238-242)
 *   oracle_chain_dp_f64   graphmambaformer/alignment/chaining.py:120-180
 *                         (chain_dp_numpy; penalty LUT supplied by the caller,
 *                          computed with the reference's own numpy expression
 *                          chaining.py:120-126, so libm never enters)
 *   oracle_chain_dp_i32   the same DP on the integer profile (log_coeff == 0,
 *                         integer open/extend/weights), exact in int32
 *   oracle_chain_backtrack graphmambaformer/alignment/chaining.py:530-576
 *                         (AffineChainer._backtrack: argsort(-f, stable),
 *                          claimed-walk, partial-chain score)
 *   oracle_sw_align_batch independent pairs (one reference batch each),
 *                         forward + traceback, fanned out over host threads
 *                         as the reference's parallel_map does
 *                         (pipeline.py:355-363, accel/parallel.py:60-103):
 *                         the CPU comparator, with no per-pair host-language
 *                         overhead
 *
 * Number model.  The reference computes SW in float32 with -1e30 standing for
 * minus infinity (extension.py:56).  Every finite value is an exact integer of
 * magnitude < 2^24, so float32 arithmetic on them is exact, and -1e30 absorbs
 * any small addend.  The restatement uses int32 with ONEG = -2^30 for -1e30
 * and saturating add/sub (nadd) so that ONEG absorbs exactly like -1e30.
 *
 * Parity pins (tests/test_oracle.py): KAT-SW1 (tests/test_accel.py:306-334 of
 * the reference), identical-sequence case (test_alignment_stages.py:1029-1048),
 * unbanded full-matrix Gotoh (test_alignment_stages.py:977-1026, restated in
 * oracle/brute.py), CIGAR rescoring and span invariants, the float32
 * numpy transcription oracle/ref_semantics.py, golden vectors written by
 * tests/golden/make_golden.py from the reference itself, and for chaining
 * KAT-CH1/KAT-CH2 and the O(n^2) brute force (test_alignment_stages.py:372-416,
 * 502-517).
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ONEG (-(1 << 30))

static inline int is_neg(int32_t x) { return x <= ONEG / 2; }
/* x + k with -inf absorbing, like float32 -1e30 + small */
static inline int32_t nadd(int32_t x, int32_t k) { return is_neg(x) ? ONEG : x + k; }
static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
static inline int feq(int32_t a, int32_t b)
{
    /* abs(a - b) <= 1e-3 in float32: exact equality for integers, and two
       -1e30 sentinels compare equal */
    if (is_neg(a) || is_neg(b)) return is_neg(a) && is_neg(b);
    return a == b;
}

#define N_CODE 5

/* extension.py:146-151 -- match iff equal, query code >= 0 and != N */
static inline int is_match(int16_t qb, int16_t tb)
{
    return qb == tb && qb >= 0 && qb != N_CODE;
}

/*
 * Batched banded affine SW, extension.py:78-204.
 *   query  (B, M) int16, target (B, N) int16, row-major
 *   qlen, tlen, off: (B,)
 *   xdrop: < 0 disables; otherwise rows stop once every pair satisfies
 *          row_max_seen - row_best > xdrop.  Scores are integers, so the
 *          reference's float test `> x_drop` equals `> floor(x_drop)`; the
 *          caller passes floor(x_drop) when x_drop > 0 and -1 otherwise.
 *   Hs/Es/Fs: optional (B, M+1, W) stores exactly like H_store/E_store/F_store
 *   rows_done: optional, receives the number of rows processed
 */
int oracle_sw_banded(int64_t B, int64_t M, int64_t N,
                     const int16_t *query, const int16_t *target,
                     const int64_t *qlen, const int64_t *tlen, const int64_t *off,
                     int32_t hb, int32_t ma, int32_t mi, int32_t go, int32_t ge,
                     int64_t xdrop,
                     int32_t *score, int64_t *qend, int64_t *tend,
                     int32_t *Hs, int32_t *Es, int32_t *Fs, int64_t *rows_done)
{
    const int64_t W = 2 * (int64_t)hb + 1;
    int32_t *hprev = malloc(sizeof(int32_t) * B * (W + 1));
    int32_t *fprev = malloc(sizeof(int32_t) * B * (W + 1));
    int32_t *hrow = malloc(sizeof(int32_t) * (W + 1));
    int32_t *frow = malloc(sizeof(int32_t) * (W + 1));
    int32_t *mrow = malloc(sizeof(int32_t) * W);
    int32_t *erow = malloc(sizeof(int32_t) * W);
    int32_t *row_best = malloc(sizeof(int32_t) * B);
    int32_t *seen = malloc(sizeof(int32_t) * B);
    if (!hprev || !fprev || !hrow || !frow || !mrow || !erow || !row_best || !seen) return -1;

    for (int64_t b = 0; b < B; ++b) {
        for (int64_t d = 0; d <= W; ++d) {
            hprev[b * (W + 1) + d] = 0;      /* extension.py:125 */
            fprev[b * (W + 1) + d] = ONEG;   /* extension.py:126 */
        }
        score[b] = 0; qend[b] = 0; tend[b] = 0; seen[b] = 0;
    }
    if (Hs) {
        for (int64_t x = 0; x < B * (M + 1) * W; ++x) { Hs[x] = 0; Es[x] = ONEG; Fs[x] = ONEG; }
    }

    int64_t max_rows = 0;
    for (int64_t b = 0; b < B; ++b) if (qlen[b] > max_rows) max_rows = qlen[b];
    int64_t done = 0;

    for (int64_t i = 1; i <= max_rows; ++i) {
        for (int64_t b = 0; b < B; ++b) {
            int32_t *hp = hprev + b * (W + 1);
            int32_t *fp = fprev + b * (W + 1);
            const int16_t qb = query[b * M + (i - 1)];
            int32_t run = ONEG; /* inclusive cummax of m + d*ge */
            for (int64_t d = 0; d < W; ++d) {
                const int64_t j = i + off[b] - hb + d;
                const int inb = (j >= 1) && (j <= tlen[b]) && (i <= qlen[b]);
                int64_t tj = j - 1;
                if (tj < 0) tj = 0;
                if (tj > N - 1) tj = N - 1 < 0 ? 0 : N - 1;
                const int16_t tbase = target[b * N + tj];
                const int32_t sub = is_match(qb, tbase) ? ma : -mi;
                const int32_t diag = hp[d] + sub;
                const int32_t vert = imax(nadd(hp[d + 1], -go), nadd(fp[d + 1], -ge));
                int32_t m = imax(imax(diag, vert), 0);
                if (!inb) m = ONEG;
                mrow[d] = m;
                /* exclusive max-plus scan (extension.py:161-167) */
                const int32_t excl = run;
                const int32_t sc = nadd(m, (int32_t)(d * ge));
                run = imax(run, sc);
                int32_t e = nadd(excl, -(go + (int32_t)((d - 1) * ge)));
                if (!inb) e = ONEG;
                erow[d] = e;
                int32_t h = imax(imax(m, e), 0);
                if (!inb) h = 0;
                hrow[d] = h;
                frow[d] = inb ? vert : ONEG;
            }
            /* row best: first argmax, strict improvement (extension.py:178-183) */
            int32_t rb = hrow[0];
            int64_t slot = 0;
            for (int64_t d = 1; d < W; ++d) if (hrow[d] > rb) { rb = hrow[d]; slot = d; }
            row_best[b] = rb;
            if (rb > score[b]) {
                score[b] = rb;
                qend[b] = i;
                tend[b] = i + off[b] - hb + slot;
            }
            if (rb > seen[b]) seen[b] = rb;
            if (Hs) {
                int64_t base = (b * (M + 1) + i) * W;
                for (int64_t d = 0; d < W; ++d) {
                    Hs[base + d] = hrow[d];
                    Es[base + d] = erow[d];
                    Fs[base + d] = frow[d];
                }
            }
            for (int64_t d = 0; d < W; ++d) { hp[d] = hrow[d]; fp[d] = frow[d]; }
            hp[W] = 0;
            fp[W] = ONEG;
        }
        done = i;
        if (xdrop >= 0) {
            int all = 1;
            for (int64_t b = 0; b < B; ++b)
                if (!((int64_t)seen[b] - row_best[b] > xdrop)) { all = 0; break; }
            if (all) break;
        }
    }
    if (rows_done) *rows_done = done;
    free(hprev); free(fprev); free(hrow); free(frow); free(mrow); free(erow);
    free(row_best); free(seen);
    return 0;
}

/* single-letter op codes used by the oracle's CIGAR output */
#define OP_EQ 0
#define OP_X 1
#define OP_I 2
#define OP_D 3

/*
 * traceback_banded for one pair, extension.py:207-302.
 *   H/E/F: (M+1, W) band-relative matrices of this pair (as stored)
 *   q/t:   codes of this pair (length >= qend / tend)
 *   ops:   out, per-column ops in forward order (capacity >= qend + tend)
 *   returns the number of ops; *read_start / *ref_start receive i / j
 */
int64_t oracle_sw_traceback(const int32_t *H, const int32_t *E, const int32_t *F,
                            int64_t M, int32_t hb, int64_t off,
                            const int16_t *q, const int16_t *t,
                            int64_t qend, int64_t tend,
                            int32_t ma, int32_t mi, int32_t go, int32_t ge,
                            uint8_t *ops, int64_t *read_start, int64_t *ref_start)
{
    const int64_t W = 2 * (int64_t)hb + 1;
    const int64_t rows = M + 1;
#define CELL(X, ii, jj) \
    ((((ii) < 0) || ((ii) >= rows) || (((jj) - (ii) - off + hb) < 0) || \
      (((jj) - (ii) - off + hb) >= W)) ? ONEG : (X)[(ii) * W + ((jj) - (ii) - off + hb)])
    int64_t i = qend, j = tend, n = 0;
    int state = 0; /* 0 = H, 1 = F (vertical, I), 2 = E (horizontal, D) */
    while (i > 0 && j > 0) {
        if (state == 0) {
            const int32_t h = CELL(H, i, j);
            if (!is_neg(h) && h <= 0) break;
            if (is_neg(h)) break; /* -1e30 <= 0.0 */
            const int m = is_match(q[i - 1], t[j - 1]);
            const int32_t sub = m ? ma : -mi;
            const int32_t dh = CELL(H, i - 1, j - 1);
            const int32_t diag = (is_neg(dh) ? 0 : imax(dh, 0)) + sub;
            if (feq(h, diag)) {
                ops[n++] = m ? OP_EQ : OP_X;
                --i; --j;
            } else if (feq(h, CELL(F, i, j))) {
                state = 1;
            } else {
                state = 2;
            }
        } else if (state == 1) {
            const int32_t cur = CELL(F, i, j);
            ops[n++] = OP_I;
            --i;
            if (feq(cur, nadd(CELL(H, i, j), -go))) state = 0;
            else if (!feq(cur, nadd(CELL(F, i, j), -ge))) state = 0;
        } else {
            const int32_t cur = CELL(E, i, j);
            ops[n++] = OP_D;
            --j;
            if (feq(cur, nadd(CELL(H, i, j), -go))) state = 0;
            else if (!feq(cur, nadd(CELL(E, i, j), -ge))) state = 0;
        }
    }
#undef CELL
    /* ops were produced backwards (extension.py:285) */
    for (int64_t a = 0, z = n - 1; a < z; ++a, --z) { uint8_t tmp = ops[a]; ops[a] = ops[z]; ops[z] = tmp; }
    *read_start = i;
    *ref_start = j;
    return n;
}

/*
 * chain_dp_numpy, chaining.py:129-180, float64.
 *   pen_lut[g] for g in [0, max_gap) is the reference's _penalty(g)
 *   bonus: optional (n, L) row-major, column t = predecessor i - L + t
 */
void oracle_chain_dp_f64(int64_t n, const int64_t *read_end, const int64_t *ref_end,
                         const double *w, int32_t max_lookback, int64_t max_gap,
                         const double *pen_lut, const double *bonus,
                         double *f, int64_t *parent)
{
    const int64_t L = max_lookback > 1 ? max_lookback : 1;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t lo = i - L > 0 ? i - L : 0;
        f[i] = w[i];
        parent[i] = -1;
        if (lo == i) continue;
        int any = 0;
        double best = 0.0;
        int64_t bj = -1;
        for (int64_t jj = lo; jj < i; ++jj) {
            const int64_t dq = read_end[i] - read_end[jj];
            const int64_t dr = ref_end[i] - ref_end[jj];
            const int ok = dq > 0 && dr > 0 && dq <= max_gap && dr <= max_gap;
            double s;
            if (ok) {
                int64_t mn = dq < dr ? dq : dr;
                double adv = (double)mn < w[i] ? (double)mn : w[i];
                int64_t g = dr - dq;
                if (g < 0) g = -g;
                volatile double t1 = f[jj] + adv;
                s = t1 - pen_lut[g];
                if (bonus) s = s + bonus[i * L + (L - (i - lo)) + (jj - lo)];
                any = 1;
            } else {
                s = -1e30;
            }
            if (bj < 0 || s > best) { best = s; bj = jj; }
        }
        if (!any) continue;
        if (best > w[i]) { f[i] = best; parent[i] = bj; }
    }
}

/* The same DP on the integer profile: pen(g) = g > 0 ? go + ge*g : 0. */
void oracle_chain_dp_i32(int64_t n, const int64_t *read_end, const int64_t *ref_end,
                         const int32_t *w, int32_t max_lookback, int64_t max_gap,
                         int32_t go, int32_t ge, int32_t *f, int64_t *parent)
{
    const int64_t L = max_lookback > 1 ? max_lookback : 1;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t lo = i - L > 0 ? i - L : 0;
        f[i] = w[i];
        parent[i] = -1;
        int64_t best = 0, bj = -1;
        for (int64_t jj = lo; jj < i; ++jj) {
            const int64_t dq = read_end[i] - read_end[jj];
            const int64_t dr = ref_end[i] - ref_end[jj];
            if (!(dq > 0 && dr > 0 && dq <= max_gap && dr <= max_gap)) continue;
            int64_t mn = dq < dr ? dq : dr;
            if (mn > w[i]) mn = w[i];
            int64_t g = dr - dq;
            if (g < 0) g = -g;
            const int64_t s = (int64_t)f[jj] + mn - (g > 0 ? go + (int64_t)ge * g : 0);
            if (bj < 0 || s > best) { best = s; bj = jj; }
        }
        if (bj >= 0 && best > w[i]) { f[i] = (int32_t)best; parent[i] = bj; }
    }
}

/* stable order of -f (np.argsort(-f, kind="stable")) by merge sort */
static void msort(int64_t *idx, int64_t *tmp, const double *key, int64_t lo, int64_t hi)
{
    if (hi - lo < 2) return;
    int64_t mid = (lo + hi) / 2;
    msort(idx, tmp, key, lo, mid);
    msort(idx, tmp, key, mid, hi);
    int64_t a = lo, b = mid, k = lo;
    while (a < mid && b < hi) {
        /* ascending -f == descending f; ties keep the left (earlier) index */
        if (-key[idx[b]] < -key[idx[a]]) tmp[k++] = idx[b++];
        else tmp[k++] = idx[a++];
    }
    while (a < mid) tmp[k++] = idx[a++];
    while (b < hi) tmp[k++] = idx[b++];
    for (k = lo; k < hi; ++k) idx[k] = tmp[k];
}

/*
 * AffineChainer._backtrack, chaining.py:530-576.
 * Outputs chains as: chain_start[c] .. chain_start[c+1] index into members
 * (DP-order local indices, in read order, i.e. already reversed), and
 * chain_score[c].  Returns the number of chains.
 */
int64_t oracle_chain_backtrack(int64_t n, const double *f, const int64_t *parent,
                               int64_t min_anchors, double min_score,
                               int64_t *members, int64_t *chain_start, double *chain_score)
{
    int64_t *order = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    int64_t *tmp = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    uint8_t *claimed = calloc(n > 0 ? n : 1, 1);
    for (int64_t x = 0; x < n; ++x) order[x] = x;
    msort(order, tmp, f, 0, n);
    int64_t nc = 0, nm = 0;
    chain_start[0] = 0;
    for (int64_t r = 0; r < n; ++r) {
        const int64_t i = order[r];
        if (claimed[i]) continue;
        int64_t node = i, start = nm;
        while (node >= 0 && !claimed[node]) {
            claimed[node] = 1;
            members[nm++] = node;
            node = parent[node];
        }
        const double s = f[i] - (node >= 0 ? f[node] : 0.0);
        const int64_t cnt = nm - start;
        if (cnt < min_anchors || s < min_score) { nm = start; continue; }
        for (int64_t a = start, z = nm - 1; a < z; ++a, --z) {
            int64_t t2 = members[a]; members[a] = members[z]; members[z] = t2;
        }
        chain_score[nc] = s;
        chain_start[++nc] = nm;
    }
    free(order); free(tmp); free(claimed);
    return nc;
}

/* ------------------------------------------------------------------------
 * oracle_sw_align_batch: n independent pairs (pair p: query codes
 * q[q_off[p] .. + q_len[p]), target codes t[t_off[p] .. + t_len[p]), band
 * offset band_off[p]), each its own batch of one (no x-drop), forward +
 * traceback, on `threads` POSIX threads taking pairs from a shared counter.
 * Outputs per pair: score, 1-based ends, 0-based starts, number of CIGAR
 * columns.  Returns 0, or -1 if an allocation failed.
 * ---------------------------------------------------------------------- */
typedef struct {
    int64_t n;
    const int16_t *q, *t;
    const int64_t *q_off, *q_len, *t_off, *t_len, *band_off;
    int32_t hb, ma, mi, go, ge;
    int32_t *score;
    int64_t *qend, *tend, *rs, *ts, *n_cols;
    int64_t next;  /* shared work counter */
    int failed;
} align_batch;

static void *align_worker(void *arg)
{
    align_batch *b = (align_batch *)arg;
    const int64_t W = 2 * (int64_t)b->hb + 1;
    for (;;) {
        const int64_t p = __atomic_fetch_add(&b->next, 1, __ATOMIC_RELAXED);
        if (p >= b->n) break;
        const int64_t M = b->q_len[p], N = b->t_len[p];
        const int16_t *q = b->q + b->q_off[p], *t = b->t + b->t_off[p];
        const size_t cells = (size_t)(M + 1) * (size_t)W;
        int32_t *H = malloc(sizeof(int32_t) * cells);
        int32_t *E = malloc(sizeof(int32_t) * cells);
        int32_t *F = malloc(sizeof(int32_t) * cells);
        uint8_t *ops = malloc((size_t)(M + N + 1));
        int64_t qe = 0, te = 0, rows = 0, rs = 0, ts = 0;
        int32_t sc = 0;
        if (!H || !E || !F || !ops ||
            oracle_sw_banded(1, M, N, q, t, &M, &N, &b->band_off[p], b->hb, b->ma, b->mi, b->go,
                             b->ge, -1, &sc, &qe, &te, H, E, F, &rows) != 0) {
            __atomic_store_n(&b->failed, 1, __ATOMIC_RELAXED);
        } else {
            b->n_cols[p] = oracle_sw_traceback(H, E, F, M, b->hb, b->band_off[p], q, t, qe, te,
                                               b->ma, b->mi, b->go, b->ge, ops, &rs, &ts);
            b->score[p] = sc;
            b->qend[p] = qe;
            b->tend[p] = te;
            b->rs[p] = rs;
            b->ts[p] = ts;
        }
        free(H);
        free(E);
        free(F);
        free(ops);
    }
    return NULL;
}

int oracle_sw_align_batch(int64_t n, const int16_t *q, const int64_t *q_off, const int64_t *q_len,
                          const int16_t *t, const int64_t *t_off, const int64_t *t_len,
                          const int64_t *band_off, int32_t hb, int32_t ma, int32_t mi, int32_t go,
                          int32_t ge, int32_t threads, int32_t *score, int64_t *qend,
                          int64_t *tend, int64_t *rs, int64_t *ts, int64_t *n_cols)
{
    align_batch b = {n, q, t, q_off, q_len, t_off, t_len, band_off, hb, ma, mi, go, ge,
                     score, qend, tend, rs, ts, n_cols, 0, 0};
    if (threads < 1) threads = 1;
    pthread_t *tid = malloc(sizeof(pthread_t) * (size_t)threads);
    if (!tid) return -1;
    int started = 0;
    for (int k = 0; k < threads; ++k)
        if (pthread_create(&tid[k], NULL, align_worker, &b) == 0) ++started;
    if (started == 0) align_worker(&b);
    for (int k = 0; k < threads; ++k)
        if (k < started) pthread_join(tid[k], NULL);
    free(tid);
    return b.failed ? -1 : 0;
}
