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

// Windows of a walk that a whole wavefront runs with wave-uniform state: the
// lanes load many dwords at once and each step picks its dword with a
// readlane, so a dependent load is paid once per window instead of once per
// plane block or 16-base word.  Plane window: 4 blocks x 16 diagonals (lane
// l holds block top - (l >> 4), diagonal d0 + (l & 15)), plus the 4 blocks
// above it loaded one window ahead.  Sequence window: 64 words (lane l holds
// word top - l); the walk only moves towards lower words.
struct PlaneWin {
    int top = INT_MIN / 2, d0 = 0;
    uint32_t v = 0u, pv = 0u;  // this window, the 4 blocks above it
};
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

// `src`, `words` and `e` count from the pair's first sequence word, so the
// positions fit 32 bits.
__device__ __forceinline__ uint32_t seq_bits_win(const uint32_t *src, int words, int e, SeqWin &s)
{
    const int w = e >> 4;
    if (w + 1 > s.top || s.top - w > 63) {
        s.top = w + 1;
        const int x = s.top - (int)(threadIdx.x & 63);
        s.v = (x >= 0 && x < words) ? src[x] : 0u;
    }
    const int k = (int)(s.top - w);  // 1 .. 63: word w in lane k, word w+1 in lane k-1
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)s.v, k);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)s.v, k - 1);
    return __builtin_amdgcn_alignbit(hi, lo, (uint32_t)(e & 15) * 2u);
}

// The walk (extension.py:243-285) shared by both walkers.  `ld(ptr, idx)`
// loads a dword (plane or sequence); `put(x, run)` stores CIGAR run x.
// Diagonal moves are taken in bulk: the not_diag bits of the rows below the
// current one in the same plane dword give the run length (up to 8 rows),
// one XOR of 2-bit base windows gives its matches and mismatches, and the
// tracked H only has to stay positive, which h > (m-1)*match guarantees.
// WIN = false (a lane per pair): the plane dword of the same diagonal one
// block up is prefetched.  WIN = true (a wavefront per pair, every lane
// active, state wave-uniform): planes and sequences come from the windows
// above and `ld` is not called.  NMK = false: the caller's batch has no N
// masks (c.nm is 0), so the mask windows and their registers go away.
template <int C, bool WIN = false, bool NMK = true, typename Ld, typename Put>
__device__ __forceinline__ void walk_steps(const WalkCtx &c, int &i, int &j, int h, int off,
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
    PlaneWin pw;
    SeqWin qsw, tsw, qmsw, tmsw;
    // WIN: sequence positions from the pair's first words (32-bit)
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
    while (i > 0 && j > 0) {
        const int d = j - i - off + c.hb;
        if (d < 0 || d >= c.W) break;  // unreachable for planes this library wrote
        if constexpr (WIN) {
            // the walk moves up at most one block per step, so it leaves a
            // window through its top-4 edge (into the window loaded ahead)
            // or sideways
            const int blk = (i - 1) >> 3;
            int rb = pw.top - blk, rd = d - pw.d0;
            if ((uint32_t)rb > 3u || (uint32_t)rd > 15u) {
                if ((uint32_t)(rb - 4) <= 3u && (uint32_t)rd <= 15u) {
                    pw.v = pw.pv;
                    pw.top -= 4;
                } else {
                    pw.top = blk;
                    pw.d0 = d - 8;
                    pw.v = plane_win_load(tb, pw.top, pw.d0, blkstride, c.W);
                }
                pw.pv = plane_win_load(tb, pw.top - 4, pw.d0, blkstride, c.W);
                rb = pw.top - blk;
                rd = d - pw.d0;
            }
            v = (uint32_t)__builtin_amdgcn_readlane((int)pw.v, (rb << 4) | rd);
        } else {
            const int64_t addr = plane_word<C>(i, d, c.Gs);
            if (addr != cur) {
                v = (addr == pf_addr) ? pf : ld(tb, addr);
                cur = addr;
                pf_addr = (i > 8) ? addr - blkstride : -1;
                if (pf_addr >= 0) pf = ld(tb, pf_addr);
            }
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
                uint32_t x, xm = 0u;
                if constexpr (WIN) {
                    const int eq = qb + i - m, et = tbs + j - m;
                    x = seq_bits_win(c.qseq + qw0, qwn, eq, qsw) ^
                        seq_bits_win(c.tseq + tw0, twn, et, tsw);
                    if (NMK && c.nm)
                        xm = seq_bits_win(c.qmask + qw0, qwn, eq, qmsw) |
                             seq_bits_win(c.tmask + tw0, twn, et, tmsw);
                } else {
                    // a lane per pair: absolute 64-bit positions on the shared
                    // stream bases (pair-relative bases per lane measured slower)
                    const int64_t eq = qo + i - m, et = to + j - m;
                    x = seq_bits(c.qseq, c.q_words, eq, qwin, ld) ^
                        seq_bits(c.tseq, c.t_words, et, twin, ld);
                    if (NMK && c.nm)
                        xm = seq_bits(c.qmask, c.q_words, eq, qmwin, ld) |
                             seq_bits(c.tmask, c.t_words, et, tmwin, ld);
                }
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

// Fused-path walk of one pair by a whole wavefront (state wave-uniform).
// NM: the batch may hold N masks.
template <int C, bool NM>
__device__ __forceinline__ void walk_pair(const WalkCtx &c, const WalkOut &o, int p, int i, int j,
                                          int h, int off, int64_t qo, int64_t to,
                                          const uint32_t *tb)
{
    const int lane = (int)(threadIdx.x & 63);
    uint32_t *out = o.cigar + o.cig_off[p];
    int nr = 0;
    int cnt[4] = {0, 0, 0, 0};
    auto ld = [&](const uint32_t *src, int64_t idx) {
        return (uint32_t)__builtin_amdgcn_readfirstlane((int)src[idx]);
    };
    auto put = [&](int x, uint32_t vrun) {
        if (lane == 0) out[x] = vrun;
    };
    walk_steps<C, true, NM>(c, i, j, h, off, qo, to, tb, ld, put, nr, cnt);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    for (int x = lane; x < nr / 2; x += 64) {
        const int y = nr - 1 - x;
        const uint32_t va = out[x], vb = out[y];
        out[x] = vb;
        out[y] = va;
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

}  // namespace gmfa
