// Traceback-walk pieces shared by the SW kernels (sw.hip, sw_split.hip):
// the walk context, small wave helpers and the plane / sequence walk
// (extension.py:243-285 decisions replayed from the 4-bit planes).
#pragma once

#include "common.h"

#include <climits>
#include <utility>

namespace gmfa {

struct WalkCtx {
    const uint32_t *qseq, *tseq, *qmask, *tmask;
    int64_t q_words, t_words;
    int32_t nm, hb, W, Gs, ma, mi, go, ge;
};

struct WalkOut {
    const int64_t *cig_off;
    uint32_t *cigar;
    int32_t *ncig, *rstart, *tstart, *counts;
};

__device__ __forceinline__ int64_t uni64(int64_t v)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ int64_t clampi64(int64_t x, int64_t hi)
{
    return x < 0 ? 0 : (x > hi ? hi : x);
}

template <int G>
__device__ __forceinline__ int uni(int x)
{
    if constexpr (G == 64) return __builtin_amdgcn_readfirstlane(x);
    return x;
}

template <int G>
__device__ __forceinline__ uint32_t uniu(uint32_t x)
{
    if constexpr (G == 64) return (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
    return x;
}

// wave-wide "all lanes true"
__device__ __forceinline__ bool wave_all(bool p)
{
    return __builtin_amdgcn_ballot_w64(!p) == 0ull;
}

__device__ __forceinline__ int wave_max(int x)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x = max(x, __shfl_xor(x, o));
    return x;
}

// a * b + c on 24-bit unsigned operands in one VALU op.  Written as asm
// because the compiler proves a is 0/1 and would lower the product to a
// compare + select + add.
__device__ __forceinline__ int mad_u24(uint32_t a, uint32_t b, int c)
{
    int r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// the same with a wave-uniform b read from an SGPR (no VGPR copy of it)
__device__ __forceinline__ int mad_u24s(uint32_t a, uint32_t b, int c)
{
    int r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
    return r;
}

// (x << 4) | CC in one VALU op (CC an inline constant, 0 .. 15)
template <int CC>
__device__ __forceinline__ int lshl4_or(int x)
{
    static_assert(CC >= 0 && CC < 16, "4-bit constant");
    int r;
    asm("v_lshl_or_b32 %0, %1, 4, %2" : "=v"(r) : "v"(x), "i"(CC));
    return r;
}

// the median of x, lo and hi in one VALU op (x clamped to [lo, hi] when
// lo <= hi); asm, so the compiler does not merge lo into a max3 before it
__device__ __forceinline__ int med3i(int x, int lo, int hi)
{
    int r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(lo), "v"(hi));
    return r;
}

// f(std::integral_constant<int, K>{}) for K = 0 .. N-1, in order
template <typename F, int... K>
__device__ __forceinline__ void unroll_ic_seq(F &f, std::integer_sequence<int, K...>)
{
    (f(std::integral_constant<int, K>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void unroll_ic(F &&f)
{
    unroll_ic_seq(f, std::make_integer_sequence<int, N>{});
}

__device__ __forceinline__ int64_t readlane64(int64_t v, int l)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), l);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// -------------------------------------------------------------------------
// traceback walk (extension.py:249-285 decisions replayed from the planes)
// -------------------------------------------------------------------------
// Decision nibble of cell (i, d) of a pair's planes: bit 3 f_cont, bit 2
// not_diag, bit 1 not_f, bit 0 e_cont.
template <int C>
__device__ __forceinline__ int64_t plane_word(int i, int d, int Gs)
{
    const int l = d / C;
    return ((int64_t)((i - 1) >> 3) * Gs + l) * C + (d - l * C);
}

// 32 bases of a 2-bit stream starting at base e (two cached words).
struct SeqWindow {
    int64_t w = -2;
    uint32_t lo = 0u, hi = 0u;
};

template <typename Ld>
__device__ __forceinline__ uint32_t seq_bits(const uint32_t *src, int64_t words, int64_t e,
                                             SeqWindow &c, Ld ld)
{
    const int64_t w = e >> 4;
    if (w != c.w) {
        if (w == c.w - 1) {  // walking backwards: slide the window down one word
            c.hi = c.lo;
            c.lo = ld(src, w);
        } else {
            c.lo = ld(src, w);
            c.hi = (w + 1 < words) ? ld(src, w + 1) : 0u;
        }
        c.w = w;
    }
    return __builtin_amdgcn_alignbit(c.hi, c.lo, (uint32_t)(e & 15) * 2u);
}

// Windows of the wavefront walk: a plane window holds 4 blocks x 16
// diagonals (lane l: block top - (l >> 4), diagonal d0 + (l & 15)); a
// sequence window 64 words (lane x: word top - x).
struct SeqWin {
    int top = -1;  // -1: nothing loaded
    uint32_t v = 0u;
};

__device__ __forceinline__ uint32_t plane_win_load(const uint32_t *tb, int top, int d0,
                                                   int blkstride, int W)
{
    const int l = (int)(threadIdx.x & 63);
    const int b = top - (l >> 4), d = d0 + (l & 15);
    // blocks 0 .. top were written (the walk starts at or below top), and
    // diagonals 0 .. W-1 lie inside every block's blkstride >= W dwords
    return (b >= 0 && d >= 0 && d < W) ? tb[(int64_t)b * blkstride + d] : 0u;
}

// The walk (extension.py:243-285) run by a whole wavefront (every lane
// active, state wave-uniform), up to 64 cells a step.  Lane k gathers the
// decision nibble of the k-th cell ahead on the current move -- diagonal
// (i-k, j-k), vertical gap (i-k, j), horizontal gap (i, j-k) -- from the
// plane window (ds_bpermute over the window registers) and, on a diagonal
// run, its two bases from the sequence windows; ballots then give how far the
// move goes and where its mismatches lie.  A diagonal run stops before the
// first cell that is not a diagonal move, lies outside the window, or has
// k*match >= h (every cell before it keeps H > 0, so the reference's stop at
// H <= 0 is never stepped over: the next step re-checks h); a gap run ends at
// its first closing cell (inclusive) or before the window's edge.
//
// Plane window: lane l of v holds block top - (l >> 4), diagonal d0 + (l &
// 15) (`pld(top, d0)` loads it, 0 outside the planes); pv the 4 blocks above.
// Sequence windows: lane x holds word top - x of the pair's stream (`qsw`,
// `tsw` may come loaded; top -1: none).
template <int C, bool NMK, typename Put, typename PLd>
__device__ __forceinline__ void walk_steps_vec(const WalkCtx &c, int &i, int &j, int h, int off,
                                               int64_t qo, int64_t to, Put put, int &nr,
                                               int (&cnt)[4], PLd pld, SeqWin qsw, SeqWin tsw)
{
    const int k = (int)(threadIdx.x & 63);
    const bool econt_ok = c.go != c.ge;
    int state = 0;  // 0 H, 1 F (I), 2 E (D)
    int run_op = -1, run_len = 0;
    int top = INT_MIN / 2, d0 = 0;
    uint32_t v = 0u, pv = 0u;
    SeqWin qmsw, tmsw;
    // sequence positions from the pair's first words (32-bit)
    const int64_t qw0 = qo >> 4, tw0 = to >> 4;
    const int qb = (int)(qo & 15), tbs = (int)(to & 15);
    const int qwn = (int)min(c.q_words - qw0, (int64_t)INT_MAX);
    const int twn = (int)min(c.t_words - tw0, (int64_t)INT_MAX);
    auto emit = [&](int op, int n) {
        if (op == run_op) {
            run_len += n;
        } else {
            if (run_len) put(nr++, ((uint32_t)run_len << 4) | (uint32_t)run_op);
            run_op = op;
            run_len = n;
        }
    };
    // counts kept in scalars (an array indexed by the state would go to scratch)
    int n_eq = 0, n_x = 0, n_i = 0, n_d = 0;
    // a sequence window holding words lo .. hi
    auto seq_cover = [&](const uint32_t *src, int words, SeqWin s, int lo, int hi) -> SeqWin {
        if (hi > s.top || s.top - lo > 63) {
            s.top = hi;
            const int x = hi - k;
            s.v = (x >= 0 && x < words) ? src[x] : 0u;
        }
        return s;
    };
    // 2-bit field of base e (lane k's own) from a window holding its word
    auto seq_field = [&](SeqWin s, int e) -> uint32_t {
        const uint32_t w =
            (uint32_t)__builtin_amdgcn_ds_bpermute((s.top - (e >> 4)) << 2, (int)s.v);
        return (w >> ((e & 15) * 2)) & 3u;
    };
    while (i > 0 && j > 0) {
        const int d = j - i - off + c.hb;
        if (d < 0 || d >= c.W) break;  // unreachable for planes this library wrote
        // the current cell into the window: a new window, or the one above
        {
            const int rb = top - ((i - 1) >> 3), rd = d - d0;
            if ((uint32_t)rd > 15u || (uint32_t)rb > 7u) {
                top = (i - 1) >> 3;
                d0 = d - 8;
                v = pld(top, d0);
                pv = pld(top - 4, d0);
            } else if (rb >= 4) {
                v = pv;
                top -= 4;
                pv = pld(top - 4, d0);
            }
        }
        // lane k's cell on the current move: row rk, diagonal dk
        const int rk = state == 2 ? i : i - k;
        const int dk = state == 0 ? d : (state == 1 ? d + k : d - k);
        const int rb = top - ((rk - 1) >> 3), rd = dk - d0;
        const bool inwin = (uint32_t)rb <= 7u && (uint32_t)rd <= 15u;
        const int idx = (((rb & 3) << 4) | (rd & 15)) << 2;
        const uint32_t wv = (uint32_t)__builtin_amdgcn_ds_bpermute(idx, (int)v);
        const uint32_t wp = (uint32_t)__builtin_amdgcn_ds_bpermute(idx, (int)pv);
        const uint32_t nib = ((rb < 4 ? wv : wp) >> (28 - 4 * ((rk - 1) & 7))) & 15u;
        if (state == 0) {
            if (h <= 0) break;
            const int lim = min(i, j);
            // bases of lanes k < min(lim, 64): query i-k-1, target j-k-1
            const int eq = qb + i - 1, et = tbs + j - 1, kl = min(lim, 64) - 1;
            qsw = seq_cover(c.qseq + qw0, qwn, qsw, (eq - kl) >> 4, eq >> 4);
            tsw = seq_cover(c.tseq + tw0, twn, tsw, (et - kl) >> 4, et >> 4);
            uint32_t mis = seq_field(qsw, eq - k) ^ seq_field(tsw, et - k);
            if (NMK && c.nm) {
                qmsw = seq_cover(c.qmask + qw0, qwn, qmsw, (eq - kl) >> 4, eq >> 4);
                tmsw = seq_cover(c.tmask + tw0, twn, tmsw, (et - kl) >> 4, et >> 4);
                mis |= seq_field(qmsw, eq - k) | seq_field(tmsw, et - k);
            }
            const bool stop = k >= lim || !inwin || (nib & 4u) || k * c.ma >= h;
            const uint64_t sm = __builtin_amdgcn_ballot_w64(stop);
            const int m = sm ? (int)__builtin_ctzll(sm) : 64;
            if (m == 0) {
                // the current cell (in the window, H > 0) is no diagonal move
                const uint32_t nib0 = (uint32_t)__builtin_amdgcn_readlane((int)nib, 0);
                state = (nib0 & 2u) ? 2 : 1;
                continue;
            }
            uint64_t mm = __builtin_amdgcn_ballot_w64(mis != 0u);
            if (m < 64) mm &= (1ull << m) - 1ull;
            const int nmis = __builtin_popcountll(mm);
            // runs of = and X in lane (walk) order
            for (int k0 = 0; k0 < m;) {
                const bool x = (mm >> k0) & 1ull;
                const uint64_t ends = (x ? ~mm : mm) >> k0;
                const int len = min(ends ? (int)__builtin_ctzll(ends) : 64 - k0, m - k0);
                emit(x ? GMFA_OP_X : GMFA_OP_EQ, len);
                k0 += len;
            }
            h -= c.ma * (m - nmis) - c.mi * nmis;
            n_eq += m - nmis;
            n_x += nmis;
            i -= m;
            j -= m;
        } else {
            // a gap: cells 0 .. a-1 are available, the first closing one ends it
            const bool avail = state == 1 ? (k < i && dk < c.W && inwin)
                                          : (k < j && dk >= 0 && inwin);
            const bool cont = state == 1 ? (nib & 8u) != 0u : (econt_ok && (nib & 1u));
            const uint64_t na = __builtin_amdgcn_ballot_w64(!avail);
            const uint64_t cl = __builtin_amdgcn_ballot_w64(avail && !cont);
            const int a = na ? (int)__builtin_ctzll(na) : 64;  // >= 1: the current cell
            const int cc = cl ? (int)__builtin_ctzll(cl) : 64;
            int L;
            if (cc < a) {
                L = cc + 1;
                h += c.ge * (L - 1) + c.go;
            } else {
                L = a;
                h += c.ge * L;
            }
            // (selects, not branches: merged branches would address i or j
            // through a pointer and move them to scratch)
            const int Li = state == 1 ? L : 0;
            emit(state == 1 ? GMFA_OP_I : GMFA_OP_D, L);
            n_i += Li;
            n_d += L - Li;
            i -= Li;
            j -= L - Li;
            if (cc < a) state = 0;
        }
    }
    if (run_len) put(nr++, ((uint32_t)run_len << 4) | (uint32_t)run_op);
    cnt[0] += n_eq;
    cnt[1] += n_x;
    cnt[2] += n_i;
    cnt[3] += n_d;
}

// The walk (extension.py:243-285) of the lane walker (one lane per pair).
// `ld(ptr, idx)` loads a dword (plane or sequence); `put(x, run)` stores CIGAR
// run x.  Diagonal moves are taken in bulk: the not_diag bits of the rows
// below the current one in the same plane dword give the run length (up to 8
// rows), one XOR of 2-bit base windows gives its matches and mismatches, and
// the tracked H only has to stay positive, which h > (m-1)*match guarantees.
// The plane dword of the same diagonal one block up is prefetched.  NMK =
// false: the caller's batch has no N masks (c.nm is 0), so the mask windows
// and their registers go away.
template <int C, bool NMK, typename Ld, typename Put>
__device__ __forceinline__ void walk_steps_lane(const WalkCtx &c, int &i, int &j, int h, int off,
                                                int64_t qo, int64_t to, const uint32_t *tb, Ld ld,
                                                Put put, int &nr, int (&cnt)[4])
{
    const bool econt_ok = c.go != c.ge;
    const int blkstride = c.Gs * C;  // dwords between vertically adjacent blocks
    int state = 0;  // 0 H, 1 F (I), 2 E (D)
    int run_op = -1, run_len = 0;
    int64_t cur = -1, pf_addr = -1;
    uint32_t v = 0u, pf = 0u;
    SeqWindow qwin, twin, qmwin, tmwin;
    auto emit = [&](int op, int n) {
        if (op == run_op) {
            run_len += n;
        } else {
            if (run_len) put(nr++, ((uint32_t)run_len << 4) | (uint32_t)run_op);
            run_op = op;
            run_len = n;
        }
    };
    while (i > 0 && j > 0) {
        const int d = j - i - off + c.hb;
        if (d < 0 || d >= c.W) break;  // unreachable for planes this library wrote
        const int64_t addr = plane_word<C>(i, d, c.Gs);
        if (addr != cur) {
            v = (addr == pf_addr) ? pf : ld(tb, addr);
            cur = addr;
            pf_addr = (i > 8) ? addr - blkstride : -1;
            if (pf_addr >= 0) pf = ld(tb, pf_addr);
        }
        const int r0 = (i - 1) & 7;
        const uint32_t nib = (v >> (28 - 4 * r0)) & 15u;
        if (state == 0) {
            if (h <= 0) break;
            if (!(nib & 4u)) {
                // rows r0, r0-1, ... of this dword while not_diag stays 0
                const uint32_t y = (v & 0x44444444u) & (~0u << (30 - 4 * r0));
                int m = y ? r0 - ((30 - (int)__builtin_ctz(y)) >> 2) : r0 + 1;
                m = min(m, min(i, j));
                if (h <= (m - 1) * c.ma) m = 1;
                const uint32_t lenmask = (m >= 16) ? ~0u : ((1u << (2 * m)) - 1u);
                // absolute 64-bit positions on the shared stream bases
                // (pair-relative bases per lane measured slower)
                const int64_t eq = qo + i - m, et = to + j - m;
                uint32_t x = seq_bits(c.qseq, c.q_words, eq, qwin, ld) ^
                             seq_bits(c.tseq, c.t_words, et, twin, ld);
                uint32_t xm = 0u;
                if (NMK && c.nm)
                    xm = seq_bits(c.qmask, c.q_words, eq, qmwin, ld) |
                         seq_bits(c.tmask, c.t_words, et, tmwin, ld);
                x = ((x | (x >> 1)) | xm) & 0x55555555u;
                x &= lenmask;  // bit 2t: base i-m+t mismatches
                const int nmis = __builtin_popcount(x);
                h -= c.ma * (m - nmis) - c.mi * nmis;
                cnt[0] += m - nmis;
                cnt[1] += nmis;
                if (x == 0u) {
                    emit(GMFA_OP_EQ, m);
                } else {
                    for (int t = m - 1; t >= 0; --t)
                        emit(((x >> (2 * t)) & 1u) ? GMFA_OP_X : GMFA_OP_EQ, 1);
                }
                i -= m;
                j -= m;
            } else {
                state = (nib & 2u) ? 2 : 1;
            }
        } else if (state == 1) {
            emit(GMFA_OP_I, 1);
            cnt[2]++;
            --i;
            if (nib & 8u) h += c.ge;
            else { h += c.go; state = 0; }
        } else {
            emit(GMFA_OP_D, 1);
            cnt[3]++;
            --j;
            if (econt_ok && (nib & 1u)) h += c.ge;
            else { h += c.go; state = 0; }
        }
    }
    if (run_len) put(nr++, ((uint32_t)run_len << 4) | (uint32_t)run_op);
}

// The walk over planes in device memory (`tb`).  WIN = false: the lane
// walker (walk_steps_lane); WIN = true: a whole wavefront (walk_steps_vec,
// windows loaded by the walk; `ld` is not called).
template <int C, bool WIN = false, bool NMK = true, typename Ld, typename Put>
__device__ __forceinline__ void walk_steps(const WalkCtx &c, int &i, int &j, int h, int off,
                                           int64_t qo, int64_t to, const uint32_t *tb, Ld ld,
                                           Put put, int &nr, int (&cnt)[4])
{
    if constexpr (WIN) {
        const int blkstride = c.Gs * C;
        auto pld = [&](int top, int d0) { return plane_win_load(tb, top, d0, blkstride, c.W); };
        walk_steps_vec<C, NMK>(c, i, j, h, off, qo, to, put, nr, cnt, pld, SeqWin{}, SeqWin{});
    } else {
        walk_steps_lane<C, NMK>(c, i, j, h, off, qo, to, tb, ld, put, nr, cnt);
    }
}

// Fused-path walk of one pair by a whole wavefront (state wave-uniform),
// plane windows from `pld` (as in walk_steps_vec), sequence windows `qsw` /
// `tsw` if the caller holds them.  NM: the batch may hold N masks.  The
// first 64 CIGAR runs are kept in a register (run x in lane x) and stored
// once, reversed, at the end; a longer CIGAR spills them to its output and
// is reversed there.
template <int C, bool NM, typename PLd>
__device__ __forceinline__ void walk_pair_with(const WalkCtx &c, const WalkOut &o, int p, int i,
                                               int j, int h, int off, int64_t qo, int64_t to,
                                               PLd pld, SeqWin qsw, SeqWin tsw)
{
    const int lane = (int)(threadIdx.x & 63);
    uint32_t *out = o.cigar + o.cig_off[p];
    int nr = 0;
    int cnt[4] = {0, 0, 0, 0};
    uint32_t runs = 0u;
    auto put = [&](int x, uint32_t vrun) {
        if (x < 64) {
            if (lane == x) runs = vrun;
        } else {
            if (x == 64) out[lane] = runs;  // spill the first 64 runs
            if (lane == 0) out[x] = vrun;
        }
    };
    walk_steps_vec<C, NM>(c, i, j, h, off, qo, to, put, nr, cnt, pld, qsw, tsw);
    if (nr <= 64) {
        // lane x stores run nr-1-x
        const int rv = __builtin_amdgcn_ds_bpermute((nr - 1 - lane) << 2, (int)runs);
        if (lane < nr) out[lane] = (uint32_t)rv;
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (int x = lane; x < nr / 2; x += 64) {
            const int y = nr - 1 - x;
            const uint32_t va = out[x], vb = out[y];
            out[x] = vb;
            out[y] = va;
        }
    }
    if (lane == 0) {
        o.ncig[p] = nr;
        o.rstart[p] = i;
        o.tstart[p] = j;
        o.counts[4 * p + 0] = cnt[0];
        o.counts[4 * p + 1] = cnt[1];
        o.counts[4 * p + 2] = cnt[2];
        o.counts[4 * p + 3] = cnt[3];
    }
}

// walk_pair_with over planes in device memory
template <int C, bool NM>
__device__ __forceinline__ void walk_pair(const WalkCtx &c, const WalkOut &o, int p, int i, int j,
                                          int h, int off, int64_t qo, int64_t to,
                                          const uint32_t *tb)
{
    const int blkstride = c.Gs * C;
    auto pld = [&](int top, int d0) { return plane_win_load(tb, top, d0, blkstride, c.W); };
    walk_pair_with<C, NM>(c, o, p, i, j, h, off, qo, to, pld, SeqWin{}, SeqWin{});
}

}  // namespace gmfa
