// Batched unit-cost wavefront alignment (WFA) with CIGAR for MI355X.
//
// Reference semantics: graphmambaformer/alignment/extension.py:308-463
// (WavefrontAligner) of sarakh1999/GraphMambaFormer, and the distance kernel
// accel/cuda_kernels.py:290-357:
//   diagonal k = i - j, offset = query bases consumed; wavefront s covers
//   k in [-s, s]; next[k] = max(clip(prev[k] + 1), clip(prev[k-1] + 1),
//   clip(prev[k+1])) where clip rejects offsets outside [0, n] or columns
//   outside [0, m]; then each diagonal extends through its run of matches
//   (a query N never matches); done when diagonal n - m reaches n.
//   Traceback: from the last wavefront, the candidate with the largest
//   unclipped offset (ties X, I, D), the gap to the reach being '=' runs.
//
// One wavefront per pair; lanes own diagonals (in chunks of 64).  The match
// extension compares 16 bases per step: the 2-bit query and target words at
// the two positions are aligned with v_alignbit, XORed, the N masks ORed in,
// and the first nonzero 2-bit field gives the run length.
// Pass 1 (wfa_distance_kernel) keeps two wavefronts in global scratch and
// finds the distance; pass 2 (wfa_align_kernel) repeats it with the whole
// history (s^2 + s + k for wavefront s, diagonal k) and walks it into a
// CIGAR of BAM runs (len << 4 | op).
#include "common.h"

namespace gmfa {

struct WfaSeq {
    const uint32_t *qseq, *tseq, *qmask, *tmask;
    int64_t q_words, t_words;
};

__device__ __forceinline__ uint32_t word16(const uint32_t *w, int64_t nw, int64_t e)
{
    const int64_t wi = e >> 4;
    const uint32_t w0 = w[wi < nw ? wi : nw - 1];
    const uint32_t w1 = w[wi + 1 < nw ? wi + 1 : nw - 1];
    return __builtin_amdgcn_alignbit(w1, w0, (uint32_t)(e & 15) * 2u);
}

// advance offset i on diagonal k through matching bases (extension.py:325-345)
__device__ __forceinline__ int wfa_extend(const WfaSeq &S, int64_t qo, int64_t to, int n, int m,
                                          int i, int k)
{
    if (i < 0) return i;
    int j = i - k;
    while (i < n && j >= 0 && j < m) {
        const int lim = min(16, min(n - i, m - j));
        uint32_t d = word16(S.qseq, S.q_words, qo + i) ^ word16(S.tseq, S.t_words, to + j);
        if (S.qmask) d |= word16(S.qmask, S.q_words, qo + i) | word16(S.tmask, S.t_words, to + j);
        d = (d | (d >> 1)) & 0x55555555u;
        const int run = d ? (int)(__builtin_ctz(d) >> 1) : 16;
        const int step = min(run, lim);
        i += step;
        j += step;
        if (step < 16) break;
    }
    return i;
}

__device__ __forceinline__ int wfa_clip(int v, int k, int n, int m)
{
    const int c = v - k;
    return (v < 0 || v > n || c < 0 || c > m) ? -1 : v;
}

struct WfaArgs {
    WfaSeq S;
    const int64_t *q_off, *t_off;
    const int32_t *qlen, *tlen;
    int32_t n_pairs, max_distance, cap;
    int32_t *scratch;         // pass 1: 2 * (2 cap + 3) per pair
    int32_t *distance;        // pass 1 out, pass 2 in (-1: over budget)
    const int64_t *hist_off;  // pass 2: (d + 1)^2 ints per pair
    int32_t *hist;
    const int64_t *cig_off;
    uint32_t *cigar;
    int32_t *ncig, *counts, *status;
};

// One score step: cur[k] for k in [-s, s] from prev (index k + c for both).
__device__ __forceinline__ void wfa_step(const WfaArgs &a, int64_t qo, int64_t to, int n, int m,
                                         int s, const int32_t *prev, int pc, int32_t *cur, int cc)
{
    for (int k0 = -s; k0 <= s; k0 += 64) {
        const int k = k0 + (int)threadIdx.x;
        if (k <= s) {
            const int p0 = prev[k + pc], pm = prev[k - 1 + pc], pp = prev[k + 1 + pc];
            const int sub = wfa_clip(p0 >= 0 ? p0 + 1 : -1, k, n, m);
            const int ins = wfa_clip(pm >= 0 ? pm + 1 : -1, k, n, m);
            const int del = wfa_clip(pp, k, n, m);
            cur[k + cc] = wfa_extend(a.S, qo, to, n, m, max(sub, max(ins, del)), k);
        }
    }
}

__global__ __launch_bounds__(64) void wfa_distance_kernel(WfaArgs a)
{
    const int p = blockIdx.x;
    if (p >= a.n_pairs) return;
    const int n = a.qlen[p], m = a.tlen[p];
    const int64_t qo = a.q_off[p], to = a.t_off[p];
    if (n == 0 || m == 0) {
        // an empty side is all insertions or deletions (no budget check, as the reference)
        if (threadIdx.x == 0) a.distance[p] = max(n, m);
        return;
    }
    const int W = 2 * a.cap + 3, c = a.cap + 1;
    int32_t *b0 = a.scratch + (int64_t)p * 2 * W, *b1 = b0 + W;
    for (int x = threadIdx.x; x < W; x += 64) { b0[x] = -1; b1[x] = -1; }
    __syncthreads();
    if (threadIdx.x == 0) b0[c] = wfa_extend(a.S, qo, to, n, m, 0, 0);
    __syncthreads();
    const int fk = n - m;
    const int lim = min(a.max_distance, a.cap);
    int32_t *prev = b0, *cur = b1;
    for (int s = 0;; ++s) {
        const bool done = fk >= -s && fk <= s && prev[fk + c] >= n;
        if (done) {
            if (threadIdx.x == 0) a.distance[p] = s;
            return;
        }
        if (s >= lim) break;
        wfa_step(a, qo, to, n, m, s + 1, prev, c, cur, c);
        __syncthreads();
        int32_t *t = prev;
        prev = cur;
        cur = t;
    }
    if (threadIdx.x == 0) a.distance[p] = -1;
}

__global__ __launch_bounds__(64) void wfa_align_kernel(WfaArgs a)
{
    const int p = blockIdx.x;
    if (p >= a.n_pairs) return;
    const int D = a.distance[p];
    if (D < 0) return;
    const int n = a.qlen[p], m = a.tlen[p];
    const int64_t qo = a.q_off[p], to = a.t_off[p];
    uint32_t *out = a.cigar + a.cig_off[p];
    int32_t *cnt = a.counts + 4 * (int64_t)p;
    if (n == 0 || m == 0) {
        if (threadIdx.x == 0) {
            int nr = 0;
            if (n > 0) out[nr++] = ((uint32_t)n << 4) | GMFA_OP_I;
            if (m > 0) out[nr++] = ((uint32_t)m << 4) | GMFA_OP_D;
            a.ncig[p] = nr;
            cnt[0] = 0; cnt[1] = 0; cnt[2] = n; cnt[3] = m;
        }
        return;
    }
    int32_t *h = a.hist + a.hist_off[p];
    // wavefront s at h[s*s + (k + s)]; reads of k outside [-s, s] return -1
    if (threadIdx.x == 0) h[0] = wfa_extend(a.S, qo, to, n, m, 0, 0);
    __syncthreads();
    for (int s = 1; s <= D; ++s) {
        const int32_t *prev = h + (int64_t)(s - 1) * (s - 1);
        int32_t *cur = h + (int64_t)s * s;
        for (int k0 = -s; k0 <= s; k0 += 64) {
            const int k = k0 + (int)threadIdx.x;
            if (k <= s) {
                auto pv = [&](int kk) { return (kk >= -(s - 1) && kk <= s - 1) ? prev[kk + s - 1] : -1; };
                const int p0 = pv(k), pm = pv(k - 1), pp = pv(k + 1);
                const int sub = wfa_clip(p0 >= 0 ? p0 + 1 : -1, k, n, m);
                const int ins = wfa_clip(pm >= 0 ? pm + 1 : -1, k, n, m);
                const int del = wfa_clip(pp, k, n, m);
                cur[k + s] = wfa_extend(a.S, qo, to, n, m, max(sub, max(ins, del)), k);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    // traceback (extension.py:425-463), runs built backwards then reversed
    int nr = 0;
    int c4[4] = {0, 0, 0, 0};  // =, X, I, D
    auto emit = [&](uint32_t op, int len) {
        if (len <= 0) return;
        const int slot = op == GMFA_OP_EQ ? 0 : op == GMFA_OP_X ? 1 : op == GMFA_OP_I ? 2 : 3;
        c4[slot] += len;
        if (nr > 0 && (out[nr - 1] & 15u) == op) out[nr - 1] += (uint32_t)len << 4;
        else out[nr++] = ((uint32_t)len << 4) | op;
    };
    int k = n - m;
    for (int s = D; s >= 1; --s) {
        const int32_t *cur = h + (int64_t)s * s;
        const int32_t *prev = h + (int64_t)(s - 1) * (s - 1);
        const int reach = cur[k + s];
        auto pv = [&](int kk) {
            const int x = (kk >= -(s - 1) && kk <= s - 1) ? prev[kk + s - 1] : -1;
            return x >= 0 ? x : -1;
        };
        const int sub = pv(k), ins = pv(k - 1), del = pv(k + 1);
        uint32_t op = GMFA_OP_X;
        int start = sub >= 0 ? sub + 1 : -1, nk = k;
        const int ci = ins >= 0 ? ins + 1 : -1;
        if (ci > start) { op = GMFA_OP_I; start = ci; nk = k - 1; }
        if (del > start) { op = GMFA_OP_D; start = del; nk = k + 1; }
        emit(GMFA_OP_EQ, reach - start);
        emit(op, 1);
        k = nk;
    }
    if (k != 0 && k != -1) {
        atomicOr(a.status, 1);  // the reference's traceback would fail here
        a.ncig[p] = 0;
        return;
    }
    emit(GMFA_OP_EQ, h[0]);
    for (int x = 0, y = nr - 1; x < y; ++x, --y) {
        const uint32_t t = out[x];
        out[x] = out[y];
        out[y] = t;
    }
    a.ncig[p] = nr;
    cnt[0] = c4[0]; cnt[1] = c4[1]; cnt[2] = c4[2]; cnt[3] = c4[3];
}

static int wfa_common(const gmfa_sw_batch *b, WfaArgs &a)
{
    GMFA_REQUIRE(b, "NULL batch");
    GMFA_REQUIRE(b->n_pairs >= 0, "negative n_pairs");
    GMFA_REQUIRE(b->n_pairs == 0 || (b->qseq && b->tseq && b->q_off && b->t_off && b->qlen &&
                                     b->tlen && b->q_words > 0 && b->t_words > 0),
                 "NULL batch pointer");
    GMFA_REQUIRE(!b->qmask == !b->tmask, "qmask and tmask must both be given or both NULL");
    a.S = WfaSeq{b->qseq, b->tseq, b->qmask, b->tmask, b->q_words, b->t_words};
    a.q_off = b->q_off;
    a.t_off = b->t_off;
    a.qlen = b->qlen;
    a.tlen = b->tlen;
    a.n_pairs = b->n_pairs;
    return GMFA_OK;
}

}  // namespace gmfa

using namespace gmfa;

extern "C" int gmfa_wfa_distance(const gmfa_sw_batch *batch, int32_t max_distance, int32_t cap,
                                 int32_t *scratch, int32_t *distance, void *stream)
{
    WfaArgs a{};
    int rc = wfa_common(batch, a);
    if (rc != GMFA_OK) return rc;
    GMFA_REQUIRE(max_distance >= 0 && cap >= 0 && cap <= (1 << 24), "bad distance bounds");
    if (a.n_pairs == 0) return GMFA_OK;
    GMFA_REQUIRE(scratch && distance, "NULL pointer");
    a.max_distance = max_distance;
    a.cap = cap;
    a.scratch = scratch;
    a.distance = distance;
    hipLaunchKernelGGL(wfa_distance_kernel, dim3(a.n_pairs), dim3(64), 0, (hipStream_t)stream, a);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
}

extern "C" int gmfa_wfa_align(const gmfa_sw_batch *batch, const int32_t *distance,
                              const int64_t *hist_off, int32_t *hist, const int64_t *cigar_off,
                              uint32_t *cigar, int32_t *n_cigar, int32_t *counts, int32_t *status,
                              void *stream)
{
    WfaArgs a{};
    int rc = wfa_common(batch, a);
    if (rc != GMFA_OK) return rc;
    if (a.n_pairs == 0) return GMFA_OK;
    GMFA_REQUIRE(distance && hist_off && hist && cigar_off && cigar && n_cigar && counts && status,
                 "NULL pointer");
    a.distance = const_cast<int32_t *>(distance);
    a.hist_off = hist_off;
    a.hist = hist;
    a.cig_off = cigar_off;
    a.cigar = cigar;
    a.ncig = n_cigar;
    a.counts = counts;
    a.status = status;
    hipLaunchKernelGGL(wfa_align_kernel, dim3(a.n_pairs), dim3(64), 0, (hipStream_t)stream, a);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
}
