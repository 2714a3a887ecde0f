// Anchor-chaining DP for MI355X (gfx950).
//
// Reference semantics: graphmambaformer/alignment/chaining.py:129-180
// (chain_dp_numpy) of sarakh1999/GraphMambaFormer:
//   f[i] = max(w_i, max_{j in [i-L, i), ok} (f[j] + min(dq, dr, w_i)) - pen(|dr-dq|))
//   ok  <=> 0 < dq, dr <= max_gap;  ties -> smallest j;  taken only if > w_i.
//
// One wavefront per read (segment).  The DP is sequential in i, so the
// parallelism is the lookback window: lane l holds predecessor j = i-64+l
// (read_end, ref_end, f) in registers; each step evaluates all candidates,
// reduces (max score, first j) across the wave with DPP, writes f[i] and
// parent[i], and slides the window one lane left (DPP wave_shl:1) with anchor
// i entering at lane 63.  Lookbacks up to 256 take 2 or 4 such windows
// (chain_dp_wide_kernel).
//
// The float64 profile reproduces numpy exactly: the penalty comes from a
// caller-supplied table computed with the reference's own numpy expression
// (chaining.py:120-126) and the score is formed as (f[j] + adv) - pen in two
// IEEE double operations, the order numpy evaluates them in.
#include "common.h"

namespace gmfa {

struct ChainArgs {
    const int64_t *__restrict__ read_end;
    const int64_t *__restrict__ ref_end;
    const double *__restrict__ wd;
    const int32_t *__restrict__ wi;
    const int64_t *__restrict__ seg_off;
    int32_t n_segments, L;
    int64_t max_gap;
    const double *__restrict__ lut;
    const double *__restrict__ bonus;
    int32_t go, ge;
    double *fd;
    int32_t *fi;
    int64_t *parent;
};

__device__ __forceinline__ int64_t shl_i64(int64_t old, int64_t v)
{
    const int lo = from_next_lane((int)(uint32_t)old, (int)(uint32_t)v);
    const int hi = from_next_lane((int)(uint32_t)((uint64_t)old >> 32), (int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

__device__ __forceinline__ double shl_f64(double old, double v)
{
    return __longlong_as_double(shl_i64(__double_as_longlong(old), __double_as_longlong(v)));
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_f64(double old, double v)
{
    const int64_t o = __double_as_longlong(old), x = __double_as_longlong(v);
    const int lo = dpp<CTRL, ROW_MASK>((int)(uint32_t)o, (int)(uint32_t)x);
    const int hi = dpp<CTRL, ROW_MASK>((int)(uint32_t)((uint64_t)o >> 32), (int)(uint32_t)((uint64_t)x >> 32));
    return __longlong_as_double((int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// max over the 64 lanes, result valid in lane 63
__device__ __forceinline__ double wave_max_to63(double x)
{
    const double ninf = -__builtin_inf();
    x = fmax(x, dpp_f64<kRowShr1>(ninf, x));
    x = fmax(x, dpp_f64<kRowShr2>(ninf, x));
    x = fmax(x, dpp_f64<kRowShr4>(ninf, x));
    x = fmax(x, dpp_f64<kRowShr8>(ninf, x));
    x = fmax(x, dpp_f64<kRowBcast15, 0xa>(ninf, x));
    x = fmax(x, dpp_f64<kRowBcast31, 0xc>(ninf, x));
    return x;
}

__device__ __forceinline__ int64_t readlane_i64(int64_t x, int l)
{
    const int lo = __builtin_amdgcn_readlane((int)(uint32_t)x, l);
    const int hi = __builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)x >> 32), l);
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const int64_t x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(uint32_t)x, l);
    const int hi = __builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)x >> 32), l);
    return __longlong_as_double((int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

__device__ __forceinline__ int32_t shl_t(int32_t old, int32_t v) { return from_next_lane(old, v); }
__device__ __forceinline__ int64_t shl_t(int64_t old, int64_t v) { return shl_i64(old, v); }
__device__ __forceinline__ int32_t readlane_t(int32_t v, int l)
{
    return __builtin_amdgcn_readlane(v, l);
}
__device__ __forceinline__ int64_t readlane_t(int64_t v, int l) { return readlane_i64(v, l); }

// One read's DP (lookback <= 64) with coordinates of type T: int64 as given,
// or int32 after subtracting the read's smallest coordinates (the caller
// takes that form when both spans are below 2^30, so every difference the
// DP forms is exact in 32 bits; a 32-bit difference and compare per
// coordinate instead of a 64-bit pair).
template <bool F64, bool BONUS, typename T>
__device__ __forceinline__ void chain_segment(const ChainArgs &a, int64_t base, int n,
                                              int64_t re0, int64_t fe0, T G)
{
    const int lane = (int)threadIdx.x;
    const int L = a.L;
    T wre = 0, wfe = 0;
    double wfd = 0.0;
    int wfi = 0;
    // anchors 64 at a time (lane k: anchor i0 + k), the next block loaded a
    // block ahead: a step reads its anchor with readlane instead of waiting
    // on a load
    T bre = 0, bfe = 0, nre = 0, nfe = 0;
    double bwd = 0.0, nwd = 0.0;
    int bwi = 0, nwi = 0;
    auto load_blk = [&](int i0, T &re, T &fe, double &wd, int &wi) {
        const int k = i0 + lane;
        if (k < n) {
            re = (T)(a.read_end[base + k] - re0);
            fe = (T)(a.ref_end[base + k] - fe0);
            if constexpr (F64) wd = a.wd[base + k];
            else wi = a.wi[base + k];
        }
    };
    load_blk(0, bre, bfe, bwd, bwi);
    load_blk(64, nre, nfe, nwd, nwi);
    for (int i = 0; i < n; ++i) {
        const int t = i & 63;
        if (t == 0 && i > 0) {
            bre = nre;
            bfe = nfe;
            bwd = nwd;
            bwi = nwi;
            load_blk(i + 64, nre, nfe, nwd, nwi);
        }
        const T re_i = readlane_t(bre, t);
        const T fe_i = readlane_t(bfe, t);
        const int lo = i - L > 0 ? i - L : 0;
        const int nvalid = i - lo;  // predecessors lo..i-1 sit in lanes 64-nvalid..63
        const int j = i - 64 + lane;
        const T dq = re_i - wre;
        const T dr = fe_i - wfe;
        const bool ok = (lane >= 64 - nvalid) && dq > 0 && dr > 0 && dq <= G && dr <= G;
        const T mn = dq < dr ? dq : dr;
        T g = dr - dq;
        g = g < 0 ? -g : g;
        const uint64_t okmask = __builtin_amdgcn_ballot_w64(ok);
        int par = -1;
        if constexpr (F64) {
            const double w_i = readlane_f64(bwd, t);
            double s = -__builtin_inf();
            if (ok) {
                const double adv = (double)mn < w_i ? (double)mn : w_i;
                const double t1 = __dadd_rn(wfd, adv);
                s = __dsub_rn(t1, a.lut[g]);
                if (BONUS) s = __dadd_rn(s, a.bonus[(base + i) * L + (L - nvalid) + (j - lo)]);
            }
            const double best = readlane_f64(wave_max_to63(s), 63);
            double fi = w_i;
            if (okmask != 0ull) {
                const uint64_t hit = __builtin_amdgcn_ballot_w64(ok && s == best);
                const int bl = (int)__builtin_ctzll(hit);
                if (best > w_i) { fi = best; par = i - 64 + bl; }
            }
            if (lane == 0) { a.fd[base + i] = fi; a.parent[base + i] = par; }
            wfd = shl_f64(fi, wfd);
        } else {
            const int w_i = __builtin_amdgcn_readlane(bwi, t);
            int key = INT_MIN;
            if (ok) {
                const int adv = (int)(mn < (T)w_i ? mn : (T)w_i);
                const int pen = g > 0 ? a.go + a.ge * (int)g : 0;
                const int s = wfi + adv - pen;
                key = (int)(((uint32_t)s << 6) | (uint32_t)(63 - lane));
            }
            key = seg_scan_max<64>(key);
            key = __builtin_amdgcn_readlane(key, 63);
            int fi = w_i;
            if (okmask != 0ull) {
                const int best = key >> 6;
                if (best > w_i) { fi = best; par = i - 64 + (63 - (key & 63)); }
            }
            if (lane == 0) { a.fi[base + i] = fi; a.parent[base + i] = par; }
            wfi = from_next_lane(fi, wfi);
        }
        wre = shl_t(re_i, wre);
        wfe = shl_t(fe_i, wfe);
    }
}

// wave-wide min (MAX = false) or max of an int64, the same in every lane
template <bool MAX>
__device__ __forceinline__ int64_t wave_reduce_i64(int64_t x)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const int64_t y = (int64_t)(((uint64_t)(uint32_t)__shfl_xor((int)((uint64_t)x >> 32), o) << 32) |
                                    (uint32_t)__shfl_xor((int)(uint32_t)x, o));
        x = MAX ? (y > x ? y : x) : (y < x ? y : x);
    }
    return readlane_i64(x, 0);
}

template <bool F64, bool BONUS>
__global__ __launch_bounds__(64) void chain_dp_kernel(ChainArgs a)
{
    const int seg = blockIdx.x;
    if (seg >= a.n_segments) return;
    const int lane = (int)threadIdx.x;
    const int64_t base = a.seg_off[seg];
    const int n = (int)(a.seg_off[seg + 1] - base);
    // the read's coordinate ranges (one pass, 64 anchors per load)
    int64_t lo_re = INT64_MAX, lo_fe = INT64_MAX, hi_re = INT64_MIN, hi_fe = INT64_MIN;
    for (int k = lane; k < n; k += 64) {
        const int64_t re = a.read_end[base + k], fe = a.ref_end[base + k];
        lo_re = re < lo_re ? re : lo_re;
        lo_fe = fe < lo_fe ? fe : lo_fe;
        hi_re = re > hi_re ? re : hi_re;
        hi_fe = fe > hi_fe ? fe : hi_fe;
    }
    lo_re = wave_reduce_i64<false>(lo_re);
    lo_fe = wave_reduce_i64<false>(lo_fe);
    hi_re = wave_reduce_i64<true>(hi_re);
    hi_fe = wave_reduce_i64<true>(hi_fe);
    // (spans compared as differences of non-overflowing halves: coordinates
    // may be any int64)
    constexpr int64_t kSpan = (int64_t)1 << 30;
    const bool narrow = n > 0 && lo_re > INT64_MIN / 2 && hi_re < INT64_MAX / 2 &&
                        lo_fe > INT64_MIN / 2 && hi_fe < INT64_MAX / 2 &&
                        hi_re - lo_re < kSpan && hi_fe - lo_fe < kSpan;
    if (narrow) {
        const int32_t G32 = (int32_t)(a.max_gap < kSpan ? a.max_gap : kSpan);
        chain_segment<F64, BONUS, int32_t>(a, base, n, lo_re, lo_fe, G32);
    } else {
        chain_segment<F64, BONUS, int64_t>(a, base, n, 0, 0, a.max_gap);
    }
}

// Lookback beyond one wavefront (64 < L <= 64 K): K windows of 64
// predecessors, window w (0 the oldest) holding j = i - 64 (K - w) + lane.
// The windows slide together, lane 63 of window w taking lane 0 of window
// w+1 and the newest taking anchor i.  The best predecessor is the largest
// score, then the smallest j (chain_dp_numpy's argmax): the wave maximum,
// then the oldest window with a lane at it, its first such lane.
template <bool F64, bool BONUS, int K>
__global__ __launch_bounds__(64) void chain_dp_wide_kernel(ChainArgs a)
{
    const int seg = blockIdx.x;
    if (seg >= a.n_segments) return;
    const int lane = (int)threadIdx.x;
    const int64_t base = a.seg_off[seg];
    const int n = (int)(a.seg_off[seg + 1] - base);
    const int L = a.L;
    const int64_t G = a.max_gap;

    int64_t wre[K], wfe[K];
    double wfd[K];
    int wfi[K];
#pragma unroll
    for (int w = 0; w < K; ++w) {
        wre[w] = 0;
        wfe[w] = 0;
        wfd[w] = 0.0;
        wfi[w] = 0;
    }
    int64_t bre = 0, bfe = 0, nre = 0, nfe = 0;
    double bwd = 0.0, nwd = 0.0;
    int bwi = 0, nwi = 0;
    auto load_blk = [&](int i0, int64_t &re, int64_t &fe, double &wd, int &wi) {
        const int k = i0 + lane;
        if (k < n) {
            re = a.read_end[base + k];
            fe = a.ref_end[base + k];
            if constexpr (F64) wd = a.wd[base + k];
            else wi = a.wi[base + k];
        }
    };
    load_blk(0, bre, bfe, bwd, bwi);
    load_blk(64, nre, nfe, nwd, nwi);
    for (int i = 0; i < n; ++i) {
        const int t = i & 63;
        if (t == 0 && i > 0) {
            bre = nre;
            bfe = nfe;
            bwd = nwd;
            bwi = nwi;
            load_blk(i + 64, nre, nfe, nwd, nwi);
        }
        const int64_t re_i = readlane_i64(bre, t);
        const int64_t fe_i = readlane_i64(bfe, t);
        const int lo = i - L > 0 ? i - L : 0;
        const int nvalid = i - lo;
        bool ok[K];
        double sd[K];
        int si[K];
        bool any = false;
#pragma unroll
        for (int w = 0; w < K; ++w) {
            const int j = i - 64 * (K - w) + lane;
            const int64_t dq = re_i - wre[w];
            const int64_t dr = fe_i - wfe[w];
            ok[w] = j >= lo && dq > 0 && dr > 0 && dq <= G && dr <= G;
            any = any || ok[w];
            const int64_t mn = dq < dr ? dq : dr;
            int64_t g = dr - dq;
            g = g < 0 ? -g : g;
            if constexpr (F64) {
                const double w_i = readlane_f64(bwd, t);
                double sv = -__builtin_inf();
                if (ok[w]) {
                    const double adv = (double)mn < w_i ? (double)mn : w_i;
                    sv = __dsub_rn(__dadd_rn(wfd[w], adv), a.lut[g]);
                    if (BONUS) sv = __dadd_rn(sv, a.bonus[(base + i) * L + (L - nvalid) + (j - lo)]);
                }
                sd[w] = sv;
            } else {
                const int w_i = __builtin_amdgcn_readlane(bwi, t);
                int sv = INT_MIN;
                if (ok[w]) {
                    const int adv = (int)(mn < (int64_t)w_i ? mn : (int64_t)w_i);
                    const int pen = g > 0 ? a.go + a.ge * (int)g : 0;
                    sv = wfi[w] + adv - pen;
                }
                si[w] = sv;
            }
        }
        const bool anyok = __builtin_amdgcn_ballot_w64(any) != 0ull;
        int64_t par = -1;
        if constexpr (F64) {
            const double w_i = readlane_f64(bwd, t);
            double m = sd[0];
#pragma unroll
            for (int w = 1; w < K; ++w) m = fmax(m, sd[w]);
            const double best = readlane_f64(wave_max_to63(m), 63);
            double fi = w_i;
            if (anyok && best > w_i) {
#pragma unroll
                for (int w = 0; w < K; ++w) {
                    const uint64_t hit = __builtin_amdgcn_ballot_w64(ok[w] && sd[w] == best);
                    if (par < 0 && hit != 0ull)
                        par = i - 64 * (K - w) + (int)__builtin_ctzll(hit);
                }
                fi = best;
            }
            if (lane == 0) { a.fd[base + i] = fi; a.parent[base + i] = par; }
#pragma unroll
            for (int w = 0; w < K; ++w) {
                const double in = w + 1 < K ? readlane_f64(wfd[w + 1 < K ? w + 1 : w], 0) : fi;
                wfd[w] = shl_f64(in, wfd[w]);
            }
        } else {
            const int w_i = __builtin_amdgcn_readlane(bwi, t);
            int m = si[0];
#pragma unroll
            for (int w = 1; w < K; ++w) m = max(m, si[w]);
            const int best = __builtin_amdgcn_readlane(seg_scan_max<64>(m), 63);
            int fi = w_i;
            if (anyok && best > w_i) {
#pragma unroll
                for (int w = 0; w < K; ++w) {
                    const uint64_t hit = __builtin_amdgcn_ballot_w64(ok[w] && si[w] == best);
                    if (par < 0 && hit != 0ull)
                        par = i - 64 * (K - w) + (int)__builtin_ctzll(hit);
                }
                fi = best;
            }
            if (lane == 0) { a.fi[base + i] = fi; a.parent[base + i] = par; }
#pragma unroll
            for (int w = 0; w < K; ++w) {
                const int in = w + 1 < K ? __builtin_amdgcn_readlane(wfi[w + 1 < K ? w + 1 : w], 0)
                                         : fi;
                wfi[w] = from_next_lane(in, wfi[w]);
            }
        }
#pragma unroll
        for (int w = 0; w < K; ++w) {
            const int64_t inr = w + 1 < K ? readlane_i64(wre[w + 1 < K ? w + 1 : w], 0) : re_i;
            const int64_t inf = w + 1 < K ? readlane_i64(wfe[w + 1 < K ? w + 1 : w], 0) : fe_i;
            wre[w] = shl_i64(inr, wre[w]);
            wfe[w] = shl_i64(inf, wfe[w]);
        }
    }
}

}  // namespace gmfa

using namespace gmfa;

constexpr int kMaxLookback = 256;  // 4 windows of 64 predecessors

// the kernel for a lookback: one window up to 64, else 2 or 4
template <bool F64, bool BONUS>
static void launch_chain_dp(const ChainArgs &a, hipStream_t s)
{
    const dim3 g((unsigned)a.n_segments), b(64);
    if (a.L <= 64)
        hipLaunchKernelGGL((chain_dp_kernel<F64, BONUS>), g, b, 0, s, a);
    else if (a.L <= 128)
        hipLaunchKernelGGL((chain_dp_wide_kernel<F64, BONUS, 2>), g, b, 0, s, a);
    else
        hipLaunchKernelGGL((chain_dp_wide_kernel<F64, BONUS, 4>), g, b, 0, s, a);
}

static int chain_common_checks(const int64_t *read_end, const int64_t *ref_end,
                               const int64_t *seg_off, int32_t n_segments, int32_t L,
                               int64_t max_gap)
{
    GMFA_REQUIRE(n_segments >= 0, "negative segment count");
    GMFA_REQUIRE(L >= 1 && L <= kMaxLookback, "max_lookback %d outside [1, %d]", L,
                 kMaxLookback);
    GMFA_REQUIRE(max_gap >= 0, "negative max_gap");
    if (n_segments == 0) return GMFA_OK;
    GMFA_REQUIRE(read_end && ref_end && seg_off, "NULL pointer");
    return GMFA_OK;
}

extern "C" int gmfa_chain_dp_f64(const int64_t *read_end, const int64_t *ref_end,
                                 const double *weight, const int64_t *seg_off,
                                 int32_t n_segments, int32_t max_lookback, int64_t max_gap,
                                 const double *pen_lut, const double *bonus, double *f,
                                 int64_t *parent, void *stream)
{
    int rc = chain_common_checks(read_end, ref_end, seg_off, n_segments, max_lookback, max_gap);
    if (rc != GMFA_OK || n_segments == 0) return rc;
    GMFA_REQUIRE(weight && pen_lut && f && parent, "NULL pointer");
    GMFA_REQUIRE(max_gap >= 1, "max_gap must be >= 1 for the penalty table");
    ChainArgs a{};
    a.read_end = read_end; a.ref_end = ref_end; a.wd = weight; a.seg_off = seg_off;
    a.n_segments = n_segments; a.L = max_lookback; a.max_gap = max_gap;
    a.lut = pen_lut; a.bonus = bonus; a.fd = f; a.parent = parent;
    if (bonus)
        launch_chain_dp<true, true>(a, (hipStream_t)stream);
    else
        launch_chain_dp<true, false>(a, (hipStream_t)stream);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
}

extern "C" int gmfa_chain_dp_i32(const int64_t *read_end, const int64_t *ref_end,
                                 const int32_t *weight, const int64_t *seg_off,
                                 int32_t n_segments, int32_t max_lookback, int64_t max_gap,
                                 int32_t gap_open, int32_t gap_extend, int32_t *f,
                                 int64_t *parent, void *stream)
{
    int rc = chain_common_checks(read_end, ref_end, seg_off, n_segments, max_lookback, max_gap);
    if (rc != GMFA_OK || n_segments == 0) return rc;
    GMFA_REQUIRE(weight && f && parent, "NULL pointer");
    GMFA_REQUIRE(gap_open >= 0 && gap_extend >= 0 && (int64_t)gap_extend * max_gap < (1 << 22),
                 "integer chaining penalties out of range");
    ChainArgs a{};
    a.read_end = read_end; a.ref_end = ref_end; a.wi = weight; a.seg_off = seg_off;
    a.n_segments = n_segments; a.L = max_lookback; a.max_gap = max_gap;
    a.go = gap_open; a.ge = gap_extend; a.fi = f; a.parent = parent;
    launch_chain_dp<false, false>(a, (hipStream_t)stream);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
}
