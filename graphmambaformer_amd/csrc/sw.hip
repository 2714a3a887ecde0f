// Banded affine-gap (Gotoh) Smith-Waterman for MI355X (gfx950).
//
// Reference semantics: graphmambaformer/alignment/extension.py:78-302
// (banded_affine_sw_batch + traceback_banded) of sarakh1999/GraphMambaFormer.
//
// Work layout
// -----------
// A pair is owned by a group of G lanes (G = 64: one pair per wavefront; G =
// 16 / 32: several short-band pairs per wavefront).  Lane l of the group owns
// C consecutive band diagonals d = l*C .. l*C+C-1 and keeps their H and F of
// the previous query row in registers.  One row of the band is one step:
//   * the vertical term F(i,d) needs row i-1 at diagonal d+1, which is the
//     lane's own next register or, for its last one, lane l+1's first
//     register -> one DPP wave_shl:1 per row;
//   * the horizontal term uses the reference's closed form (extension.py:
//     14-33): E(d) = max_{d'<d} (M(d') + d'*e) - o - (d-1)*e, an exclusive
//     max-plus prefix: serial inside the lane, then a DPP segmented max-scan
//     across lanes (row_shr 1/2/4/8, row_bcast 15/31);
//   * the target bases of a lane shift by one per row; they come from a
//     2-bit packed window loaded once per 8 rows, the query base is
//     wave-uniform (SGPR) when G = 64.
// Values are kept in the row-shifted domain X~(i) = X(i) - i*match, which turns
// the diagonal step into H~(i-1,d) - mismatch_bit*(match+mismatch) (one mad)
// and the zero clamp into a per-row scalar.  All arithmetic is exact int32;
// -inf is kNeg with enough headroom that it never wraps.
//
// Traceback planes
// ----------------
// 4 bits per cell, 8 rows per block; lane l's C dwords of a block hold, row by
// row, C "F continues" bits then (not-diagonal, not-F, E continues) for each
// cell -- exactly the decisions traceback_banded makes (extension.py:249-283):
//   not_diag = H != max(H(i-1,j-1),0) + sub        (diag has priority)
//   not_f    = H != F
//   f_cont   = F(i-1,j) - e > H(i-1,j) - o          (stay in the vertical gap)
//   e_cont   = E(i,j-1) - e > H(i,j-1) - o          (stay in the horizontal gap)
// e_cont is stored with the left neighbour's cell as "G > A" (prefix before
// d-1 beats d-1), which equals the reference's test whenever open > extend;
// for open == extend the reference never continues a horizontal gap and the
// walker ignores the bit.  Bits are shifted into the dwords with v_alignbit
// from sign bits, so a block's dword holds 32 decisions, first one in bit 31.
//
// Best cell: the reference keeps the first row with a strict improvement and
// the first column of that row (extension.py:178-182), i.e. the
// lexicographically smallest (i, j) among cells with the maximal H.  Each lane
// tracks its own first maximum; the group reduces (max H, min i, min d).
#include "common.h"

#include <algorithm>
#include <climits>

namespace gmfa {

struct SwKArgs {
    const uint32_t *qseq, *tseq, *qmask, *tmask;
    int64_t q_words, t_words;
    const int64_t *q_off, *t_off;
    const int32_t *qlen, *tlen, *boff, *order;
    int32_t n_pairs, n_items;
    int32_t hb, W, ma, mi, go, ge;
    int32_t *score, *qend, *tend;
    uint32_t *tb;
    const int64_t *tb_off;
    int32_t *rowkey;
    const int64_t *rk_off;
    uint32_t *counter;
};

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

template <int C, int G, bool TB, bool NM, bool RK>
__global__ __launch_bounds__(64) void sw_forward_kernel(SwKArgs a)
{
    constexpr int NCW = (C + 15) / 16;  // target words feeding the cells
    constexpr int NA = NCW + 1;         // aligned window words
    constexpr int NWR = NA + 1;         // raw words loaded per block
    const int lane = (int)threadIdx.x;
    const int grp = lane / G;
    const int gl = lane % G;

    const int W = a.W, hb = a.hb, ma = a.ma, mi = a.mi, go = a.go, ge = a.ge;
    const int lstar = (W - 1) / C;  // group lane holding diagonal W-1
    const int kstar = (W - 1) % C;
    const bool dead_hi = gl >= lstar;                     // registers k > kstar dead
    const bool guard = (kstar == C - 1) && (gl == lstar); // its d+1 is the band guard
    const bool lane_live = gl <= lstar;
    const int lCe = gl * C * ge;
    const int cH = go + ma, cF = ge + ma, mms = ma + mi;
    const int64_t t_hi = a.t_words - 1, q_hi = a.q_words - 1;

    for (;;) {
        int item = 0;
        if (lane == 0) item = (int)atomicAdd(a.counter, 1u);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= a.n_items) break;

        const int pidx = item * (64 / G) + grp;
        const bool has = pidx < a.n_pairs;
        const int p = has ? (a.order ? a.order[pidx] : pidx) : 0;
        const int qlen = uni<G>(has ? a.qlen[p] : 0);
        const int tlen = uni<G>(has ? a.tlen[p] : 0);
        const int off = uni<G>(has ? a.boff[p] : 0);
        const int64_t qo = has ? a.q_off[p] : 0;
        const int64_t to = has ? a.t_off[p] : 0;
        const int maxq = (G == 64) ? qlen : wave_max(qlen);

        int Hp[C], Fp[C];
        uint32_t acc[C];
#pragma unroll
        for (int k = 0; k < C; ++k) { Hp[k] = 0; Fp[k] = kNeg; acc[k] = 0u; }
        int bestK = 31, besti = 0;

        const int nblk = (maxq + 7) >> 3;
        const int nblk_mine = (qlen + 7) >> 3;
        const int64_t tb_base = (TB && has) ? a.tb_off[p] : 0;
        const int64_t rk_base = (RK && has) ? a.rk_off[p] : 0;

        // target element (0-based in the pair) of cell 0 at the block's first row
        int64_t e_t = to + (int64_t)(1 + off - hb + gl * C - 1);
        const int64_t e_q = qo;  // query element of row 1

        uint32_t wt[NWR], wn[NWR];
        auto load_t = [&](int64_t e, uint32_t (&w)[NWR], const uint32_t *src) {
            const int64_t wb = e >> 4;
#pragma unroll
            for (int x = 0; x < NWR; ++x) w[x] = src[clampi64(wb + x, t_hi)];
        };
        if (maxq > 0) {
            load_t(e_t, wt, a.tseq);
            if (NM) load_t(e_t, wn, a.tmask);
        }

        for (int b = 0; b < nblk; ++b) {
            const int i0 = 8 * b + 1;
            // align this block's windows, then prefetch the next block
            uint32_t ta[NA], na[NA];
            {
                const uint32_t sh = (uint32_t)(e_t & 15) * 2u;
#pragma unroll
                for (int x = 0; x < NA; ++x) {
                    ta[x] = __builtin_amdgcn_alignbit(wt[x + 1], wt[x], sh);
                    if (NM) na[x] = __builtin_amdgcn_alignbit(wn[x + 1], wn[x], sh);
                }
            }
            e_t += 8;
            if (b + 1 < nblk) {
                load_t(e_t, wt, a.tseq);
                if (NM) load_t(e_t, wn, a.tmask);
            }
            uint32_t qbits, qnbits = 0u;
            {
                const int64_t eq = e_q + 8 * b;
                const int64_t wq = eq >> 4;
                const uint32_t sh = (uint32_t)(eq & 15) * 2u;
                const uint32_t w0 = a.qseq[clampi64(wq, q_hi)];
                const uint32_t w1 = a.qseq[clampi64(wq + 1, q_hi)];
                qbits = uniu<G>(__builtin_amdgcn_alignbit(w1, w0, sh));
                if (NM) {
                    const uint32_t m0 = a.qmask[clampi64(wq, q_hi)];
                    const uint32_t m1 = a.qmask[clampi64(wq + 1, q_hi)];
                    qnbits = uniu<G>(__builtin_amdgcn_alignbit(m1, m0, sh));
                }
            }

            const bool fast_l = (i0 >= hb + 1 - off) && (i0 + 7 <= min(qlen, tlen - hb - off));
            const bool fast = (G == 64) ? fast_l : wave_all(fast_l || !has);

            auto row = [&](auto slow_tag, const int r) {
                constexpr bool SLOW = decltype(slow_tag)::value;
                const int i = i0 + r;
                const int zrow = -i * ma;
                const int zprev = zrow + ma;
                const uint32_t qc = (qbits >> (2 * r)) & 3u;
                const uint32_t qpat = qc * 0x55555555u;
                const bool qn = NM && ((qnbits >> (2 * r)) & 1u);
                uint32_t z[NCW];
#pragma unroll
                for (int x = 0; x < NCW; ++x) {
                    uint32_t y = __builtin_amdgcn_alignbit(ta[x + 1], ta[x], 2 * r) ^ qpat;
                    y |= y >> 1;
                    if (NM) {
                        y |= __builtin_amdgcn_alignbit(na[x + 1], na[x], 2 * r);
                        y = qn ? ~0u : y;
                    }
                    z[x] = y;
                }
                int lo = 0;
                uint32_t span = 0;
                if constexpr (SLOW) {
                    const int dlo = hb + 1 - off - i;
                    const int dhi = tlen + hb - off - i;
                    const int l0 = max(dlo, 0);
                    const int h0 = (i <= qlen) ? min(dhi, W - 1) : -1;
                    lo = l0 - gl * C;
                    span = (uint32_t)max(h0 - l0 + 1, 0);
                }
                int hN = from_next_lane(zprev, Hp[0]);
                int fN = from_next_lane(kNeg, Fp[0]);
                hN = guard ? zprev : hN;
                fN = guard ? kNeg : fN;

                int Mv[C], Fv[C], Dg[C], Av[C], Pv[C];
#pragma unroll
                for (int k = 0; k < C; ++k) {
                    const int uh = (k + 1 < C) ? Hp[(k + 1) % C] : hN;
                    const int uf = (k + 1 < C) ? Fp[(k + 1) % C] : fN;
                    const int av = uh - cH;
                    const int bv = uf - cF;
                    Fv[k] = max(av, bv);
                    if (TB) {
                        constexpr int dummy = 0;
                        (void)dummy;
                        const int n = r * 4 * C + k;
                        acc[n >> 5] = insert_sign(acc[n >> 5], av - bv);
                    }
                    const uint32_t mm = __builtin_amdgcn_ubfe(z[k >> 4], 2 * (k & 15), 1);
                    Dg[k] = Hp[k] - __mul24((int)mm, mms);
                    const int m = max(max(Dg[k], Fv[k]), zrow);
                    Mv[k] = m;
                    int A = m + k * ge;
                    if constexpr (SLOW) {
                        const bool live = (uint32_t)(k - lo) < span;
                        A = live ? A : kNeg;
                    }
                    Av[k] = A;
                    Pv[k] = k ? max(Pv[k ? k - 1 : 0], A) : A;
                }
                int X = from_prev_lane(seg_scan_max<G>(Pv[C - 1] + lCe));
                if constexpr (G < 64) X = (gl == 0) ? kNeg : X;
                const int Xl = X - lCe;
                const int Glast = (C > 1) ? max(Xl, Pv[C > 1 ? C - 2 : 0]) : Xl;
                const int dprev = from_prev_lane(Av[C - 1] - Glast);

                int rk = INT_MIN;
                int Gkm1 = Xl;
#pragma unroll
                for (int k = 0; k < C; ++k) {
                    const int Gk = k ? max(Xl, Pv[k ? k - 1 : 0]) : Xl;
                    const int E = Gk - (go + (k - 1) * ge);
                    int H = max(Mv[k], E);
                    int Fk = Fv[k];
                    if constexpr (SLOW) {
                        const bool live = (uint32_t)(k - lo) < span;
                        H = live ? H : zrow;
                        Fk = live ? Fk : kNeg;
                    } else {
                        if (k > kstar) {
                            H = dead_hi ? zrow : H;
                            Fk = dead_hi ? kNeg : Fk;
                        }
                    }
                    if (TB) {
                        const int n = r * 4 * C + C + 3 * k;
                        acc[n >> 5] = insert_sign(acc[n >> 5], Dg[k] - H);
                        acc[(n + 1) >> 5] = insert_sign(acc[(n + 1) >> 5], Fk - H);
                        acc[(n + 2) >> 5] =
                            insert_sign(acc[(n + 2) >> 5], k ? (Av[k ? k - 1 : 0] - Gkm1) : dprev);
                    }
                    rk = max(rk, (int)(((uint32_t)H << 5) | (uint32_t)(31 - k)));
                    Hp[k] = H;  // Hp[k] is not read again in this row
                    Fv[k] = Fk;
                    Gkm1 = Gk;
                }
#pragma unroll
                for (int k = 0; k < C; ++k) Fp[k] = Fv[k];
                const int tk = rk + (int)((uint32_t)(i * ma) << 5);
                if (tk > (bestK | 31)) { bestK = tk; besti = i; }
                if constexpr (RK) {
                    // lane-local (max H, first d) of this row; the host takes the
                    // max over a row's lanes (keys order by H, then by smaller d)
                    if (has && lane_live && i <= qlen) {
                        const int htrue = (rk >> 5) + i * ma;
                        const int d = gl * C + 31 - (rk & 31);
                        a.rowkey[rk_base + (int64_t)(i - 1) * (lstar + 1) + gl] =
                            (int)(((uint32_t)htrue << 11) | (uint32_t)(2047 - d));
                    }
                }
            };
            // Hp[k] is overwritten in pass 2 only after pass 1 consumed it, and
            // Hp[k+1] (read by cell k) is consumed in pass 1 as well.
            if (fast) {
#pragma unroll
                for (int r = 0; r < 8; ++r) row(std::false_type{}, r);
            } else {
#pragma unroll
                for (int r = 0; r < 8; ++r) row(std::true_type{}, r);
            }
            if (TB && has && lane_live && b < nblk_mine) {
                uint32_t *dst = a.tb + tb_base + ((int64_t)b * (lstar + 1) + gl) * C;
#pragma unroll
                for (int k = 0; k < C; ++k) dst[k] = acc[k];
            }
        }

        // group reduction of the lane-first maxima: (max H, min i, min d)
        int bH = bestK >> 5;
        const int bd = gl * C + (31 - (bestK & 31));
        if (!lane_live || besti == 0) bH = 0;
        uint64_t key = ((uint64_t)(uint32_t)bH << 40) |
                       ((uint64_t)(0xFFFFFu - (uint32_t)besti) << 12) | (uint64_t)(0xFFFu - (uint32_t)bd);
        if (bH == 0) key = 0;
#pragma unroll
        for (int o = G / 2; o >= 1; o >>= 1) {
            const uint32_t lo32 = (uint32_t)__shfl_xor((int)(uint32_t)key, o);
            const uint32_t hi32 = (uint32_t)__shfl_xor((int)(uint32_t)(key >> 32), o);
            const uint64_t other = ((uint64_t)hi32 << 32) | lo32;
            key = other > key ? other : key;
        }
        if (gl == 0 && has) {
            if (key == 0) {
                a.score[p] = 0; a.qend[p] = 0; a.tend[p] = 0;
            } else {
                const int H = (int)(key >> 40);
                const int bi = (int)(0xFFFFFu - (uint32_t)((key >> 12) & 0xFFFFFu));
                const int dd = (int)(0xFFFu - (uint32_t)(key & 0xFFFu));
                a.score[p] = H;
                a.qend[p] = bi;
                a.tend[p] = bi + off - hb + dd;
            }
        }
    }
}

__global__ void reset_counter_kernel(uint32_t *c) { *c = 0u; }

// -------------------------------------------------------------------------
// traceback walk: one thread per pair
// -------------------------------------------------------------------------
struct TbKArgs {
    const uint32_t *qseq, *tseq, *qmask, *tmask;
    const int64_t *q_off, *t_off;
    const int32_t *boff;
    int32_t n_pairs, hb, W, C, Gs, ma, mi, go, ge;
    int32_t nm;
    const int32_t *score, *qend, *tend;
    const uint32_t *tb;
    const int64_t *tb_off;
    const int64_t *cig_off;
    uint32_t *cigar;
    int32_t *ncig, *rstart, *tstart, *counts;
};

__device__ __forceinline__ uint32_t base2(const uint32_t *s, int64_t e)
{
    return (s[e >> 4] >> ((uint32_t)(e & 15) * 2u)) & 3u;
}

__global__ __launch_bounds__(64) void sw_traceback_kernel(TbKArgs a)
{
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= a.n_pairs) return;
    const int C = a.C, Gs = a.Gs, hb = a.hb, W = a.W;
    const int off = a.boff[p];
    const int64_t qo = a.q_off[p], to = a.t_off[p];
    const uint32_t *tb = a.tb + a.tb_off[p];
    uint32_t *out = a.cigar + a.cig_off[p];
    const bool econt_ok = a.go != a.ge;
    int i = a.qend[p], j = a.tend[p];
    int h = a.score[p];
    int state = 0;  // 0 H, 1 F (I), 2 E (D)
    int run_op = -1, run_len = 0, nr = 0;
    int cnt[4] = {0, 0, 0, 0};
    auto emit = [&](int op) {
        if (op == run_op) {
            ++run_len;
        } else {
            if (run_len) out[nr++] = ((uint32_t)run_len << 4) | (uint32_t)run_op;
            run_op = op;
            run_len = 1;
        }
    };
    while (i > 0 && j > 0) {
        const int d = j - i - off + hb;
        if (d < 0 || d >= W) break;  // unreachable for planes this library wrote
        const int blk = (i - 1) >> 3, r = (i - 1) & 7, l = d / C, k = d - l * C;
        const uint32_t *cell = tb + ((int64_t)blk * Gs + l) * C;
        const int nf = r * 4 * C + k;
        const int nd = r * 4 * C + C + 3 * k;
        auto bit = [&](int n) { return (cell[n >> 5] >> (31 - (n & 31))) & 1u; };
        if (state == 0) {
            if (h <= 0) break;
            if (!bit(nd)) {
                const int64_t eq = qo + i - 1, et = to + j - 1;
                bool match = base2(a.qseq, eq) == base2(a.tseq, et);
                if (a.nm) match = match && !(base2(a.qmask, eq) & 1u) && !(base2(a.tmask, et) & 1u);
                emit(match ? GMFA_OP_EQ : GMFA_OP_X);
                cnt[match ? 0 : 1]++;
                h -= match ? a.ma : -a.mi;
                --i;
                --j;
            } else if (!bit(nd + 1)) {
                state = 1;
            } else {
                state = 2;
            }
        } else if (state == 1) {
            const bool cont = bit(nf);
            emit(GMFA_OP_I);
            cnt[2]++;
            --i;
            if (cont) h += a.ge;
            else { h += a.go; state = 0; }
        } else {
            const bool cont = econt_ok && bit(nd + 2);
            emit(GMFA_OP_D);
            cnt[3]++;
            --j;
            if (cont) h += a.ge;
            else { h += a.go; state = 0; }
        }
    }
    if (run_len) out[nr++] = ((uint32_t)run_len << 4) | (uint32_t)run_op;
    for (int x = 0, y = nr - 1; x < y; ++x, --y) {
        const uint32_t t = out[x];
        out[x] = out[y];
        out[y] = t;
    }
    a.ncig[p] = nr;
    a.rstart[p] = i;
    a.tstart[p] = j;
    a.counts[4 * p + 0] = cnt[0];
    a.counts[4 * p + 1] = cnt[1];
    a.counts[4 * p + 2] = cnt[2];
    a.counts[4 * p + 3] = cnt[3];
}

// -------------------------------------------------------------------------
// int16 codes -> 2-bit stream + N mask; one thread per output word
// -------------------------------------------------------------------------
__global__ void pack_codes_kernel(const int16_t *codes, int64_t L, const int32_t *lens,
                                  const int64_t *word_off, uint32_t *seq, uint32_t *mask,
                                  int32_t *bad)
{
    const int64_t b = blockIdx.y;
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t len = lens[b];
    const int64_t nw = (len + 15) >> 4;
    if (w >= nw) return;
    uint32_t s = 0u, m = 0u;
    int badc = 0;
    const int16_t *row = codes + b * L;
#pragma unroll
    for (int x = 0; x < 16; ++x) {
        const int64_t e = w * 16 + x;
        int c = -1;
        if (e < len) c = row[e];
        if (c >= 1 && c <= 4) {
            s |= (uint32_t)(c - 1) << (2 * x);
        } else {
            m |= 1u << (2 * x);
            if (c == 0 || c > 5) badc = 1;
        }
    }
    seq[word_off[b] + w] = s;
    mask[word_off[b] + w] = m;
    if (badc) atomicOr(bad, 1);
}

// -------------------------------------------------------------------------
// host side: variant table
// -------------------------------------------------------------------------
using SwKernel = void (*)(SwKArgs);

template <int C, int G>
SwKernel pick_kernel(uint32_t flags)
{
    const bool tb = flags & GMFA_SW_TRACEBACK, nm = flags & GMFA_SW_NMASK, rk = flags & GMFA_SW_ROWKEYS;
    if (rk && tb && nm) return sw_forward_kernel<C, G, true, true, true>;
    if (rk && tb) return sw_forward_kernel<C, G, true, false, true>;
    if (rk && nm) return sw_forward_kernel<C, G, false, true, true>;
    if (rk) return sw_forward_kernel<C, G, false, false, true>;
    if (tb && nm) return sw_forward_kernel<C, G, true, true, false>;
    if (tb) return sw_forward_kernel<C, G, true, false, false>;
    if (nm) return sw_forward_kernel<C, G, false, true, false>;
    return sw_forward_kernel<C, G, false, false, false>;
}

struct TileEntry {
    int C, G;
    SwKernel (*get)(uint32_t);
};

#define GMFA_TILE(C, G) {C, G, &pick_kernel<C, G>}
static const TileEntry kTiles[] = {
    GMFA_TILE(2, 16),  GMFA_TILE(5, 16),  GMFA_TILE(8, 16),
    GMFA_TILE(4, 32),  GMFA_TILE(8, 32),
    GMFA_TILE(2, 64),  GMFA_TILE(4, 64),  GMFA_TILE(8, 64),  GMFA_TILE(9, 64),
    GMFA_TILE(12, 64), GMFA_TILE(16, 64), GMFA_TILE(24, 64), GMFA_TILE(32, 64),
};
#undef GMFA_TILE

static const TileEntry *find_tile(int C, int G)
{
    for (const auto &t : kTiles)
        if (t.C == C && t.G == G) return &t;
    return nullptr;
}

// default: the compiled tile with the least padding, fewer lanes per pair
// first (more pairs per wavefront), then more cells per lane
static const TileEntry *default_tile(int W)
{
    const TileEntry *best = nullptr;
    double best_eff = -1.0;
    for (const auto &t : kTiles) {
        if (t.C * t.G < W) continue;
        const int lanes = (W + t.C - 1) / t.C;
        const double eff = (double)W / (double)(t.G * t.C) * (lanes <= t.G ? 1.0 : 0.0);
        if (eff > best_eff + 1e-9) { best = &t; best_eff = eff; }
    }
    return best;
}

}  // namespace gmfa

using namespace gmfa;

extern "C" int gmfa_sw_resolve_tile(int32_t half_band, int32_t cells_per_lane,
                                    int32_t lanes_per_pair, gmfa_sw_tile *tile)
{
    GMFA_REQUIRE(tile != nullptr, "tile is NULL");
    GMFA_REQUIRE(half_band >= 0 && half_band <= 1023, "half_band %d outside [0, 1023]", half_band);
    const int W = 2 * half_band + 1;
    const TileEntry *t = nullptr;
    if (cells_per_lane > 0 || lanes_per_pair > 0) {
        t = find_tile(cells_per_lane, lanes_per_pair);
        if (!t) {
            set_error("no compiled SW tile with %d cells per lane and %d lanes per pair",
                      cells_per_lane, lanes_per_pair);
            return GMFA_E_UNSUPPORTED;
        }
        if (t->C * t->G < W) {
            set_error("tile %dx%d cannot hold a band of width %d", t->C, t->G, W);
            return GMFA_E_UNSUPPORTED;
        }
    } else {
        t = default_tile(W);
        if (!t) {
            set_error("no compiled SW tile holds a band of width %d", W);
            return GMFA_E_UNSUPPORTED;
        }
    }
    tile->cells_per_lane = t->C;
    tile->lanes_per_pair = t->G;
    tile->lanes_stored = (W + t->C - 1) / t->C;
    tile->half_band = half_band;
    return GMFA_OK;
}

extern "C" int gmfa_sw_list_tiles(int32_t *cells_per_lane, int32_t *lanes_per_pair, int32_t cap)
{
    const int n = (int)(sizeof(kTiles) / sizeof(kTiles[0]));
    for (int k = 0; k < n && k < cap; ++k) {
        if (cells_per_lane) cells_per_lane[k] = kTiles[k].C;
        if (lanes_per_pair) lanes_per_pair[k] = kTiles[k].G;
    }
    return n;
}

extern "C" int64_t gmfa_sw_tb_bytes(int32_t qlen, const gmfa_sw_tile *tile)
{
    if (!tile || qlen < 0) return GMFA_E_INVALID;
    const int64_t nblk = ((int64_t)qlen + 7) / 8;
    return nblk * tile->lanes_stored * tile->cells_per_lane * 4;
}

extern "C" int gmfa_pack_codes(const int16_t *codes, int64_t B, int64_t L, const int32_t *lens,
                               const int64_t *word_off, int64_t total_words, uint32_t *seq,
                               uint32_t *mask, int32_t *bad_codes, void *stream)
{
    GMFA_REQUIRE(B >= 0 && L >= 0, "negative shape");
    GMFA_REQUIRE(B <= 65535, "pack_codes handles at most 65535 rows per call (got %lld)",
                 (long long)B);
    if (B == 0 || L == 0) return GMFA_OK;
    GMFA_REQUIRE(codes && lens && word_off && seq && mask && bad_codes, "NULL pointer");
    (void)total_words;
    const int64_t nw = (L + 15) / 16;
    dim3 grid((unsigned)((nw + 255) / 256), (unsigned)B);
    hipLaunchKernelGGL(pack_codes_kernel, grid, dim3(256), 0, (hipStream_t)stream, codes, L, lens,
                       word_off, seq, mask, bad_codes);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
}

static int resident_waves(SwKernel k)
{
    int dev = 0, cus = 0, nb = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1024;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 1024;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)k, 64, 0) != hipSuccess ||
        nb <= 0)
        nb = 8;
    return cus * nb;
}

extern "C" int gmfa_sw_forward(const gmfa_sw_batch *batch, const gmfa_sw_params *params,
                               const gmfa_sw_out *out, void *stream)
{
    GMFA_REQUIRE(batch && params && out, "NULL argument");
    const gmfa_sw_params &P = *params;
    GMFA_REQUIRE(P.match >= 0 && P.mismatch >= 0 && P.gap_extend >= 0,
                 "scores must be non-negative integers");
    GMFA_REQUIRE(P.gap_open >= P.gap_extend, "gap_open (%d) must be >= gap_extend (%d)",
                 P.gap_open, P.gap_extend);
    GMFA_REQUIRE(P.match <= 1000 && P.mismatch <= 1000 && P.gap_open <= 100000,
                 "scoring magnitudes too large for the int32 DP");
    gmfa_sw_tile tile;
    int rc = gmfa_sw_resolve_tile(P.half_band, P.cells_per_lane, P.lanes_per_pair, &tile);
    if (rc != GMFA_OK) return rc;
    const int n = batch->n_pairs;
    GMFA_REQUIRE(n >= 0, "negative n_pairs");
    if (n == 0) return GMFA_OK;
    GMFA_REQUIRE(batch->qseq && batch->tseq && batch->q_off && batch->t_off && batch->qlen &&
                     batch->tlen && batch->band_offset,
                 "NULL batch pointer");
    GMFA_REQUIRE(batch->q_words > 0 && batch->t_words > 0, "empty sequence streams");
    GMFA_REQUIRE(batch->max_qlen >= 0 && batch->max_qlen <= (1 << 20),
                 "query length bound %d outside [0, 2^20]", batch->max_qlen);
    GMFA_REQUIRE((int64_t)batch->max_qlen * P.match < (1 << 24),
                 "query length x match score must stay below 2^24 for the int32 DP");
    GMFA_REQUIRE(!(P.flags & GMFA_SW_ROWKEYS) || (int64_t)batch->max_qlen * P.match < (1 << 19),
                 "per-row keys need query length x match below 2^19");
    const bool nm = P.flags & GMFA_SW_NMASK;
    GMFA_REQUIRE(!nm || (batch->qmask && batch->tmask), "GMFA_SW_NMASK needs qmask and tmask");
    GMFA_REQUIRE(out->score && out->query_end && out->target_end && out->work_counter,
                 "NULL output pointer");
    GMFA_REQUIRE(!(P.flags & GMFA_SW_TRACEBACK) || (out->tb && out->tb_off),
                 "GMFA_SW_TRACEBACK needs tb and tb_off");
    GMFA_REQUIRE(!(P.flags & GMFA_SW_ROWKEYS) || (out->rowkey && out->rowkey_off),
                 "GMFA_SW_ROWKEYS needs rowkey and rowkey_off");
    const TileEntry *t = find_tile(tile.cells_per_lane, tile.lanes_per_pair);
    SwKernel k = t->get(P.flags);

    SwKArgs a;
    a.qseq = batch->qseq; a.tseq = batch->tseq; a.qmask = batch->qmask; a.tmask = batch->tmask;
    a.q_words = batch->q_words; a.t_words = batch->t_words;
    a.q_off = batch->q_off; a.t_off = batch->t_off;
    a.qlen = batch->qlen; a.tlen = batch->tlen; a.boff = batch->band_offset; a.order = batch->order;
    a.n_pairs = n;
    const int ppw = 64 / t->G;
    a.n_items = (n + ppw - 1) / ppw;
    a.hb = P.half_band; a.W = 2 * P.half_band + 1;
    a.ma = P.match; a.mi = P.mismatch; a.go = P.gap_open; a.ge = P.gap_extend;
    a.score = out->score; a.qend = out->query_end; a.tend = out->target_end;
    a.tb = out->tb; a.tb_off = out->tb_off;
    a.rowkey = out->rowkey; a.rk_off = out->rowkey_off;
    a.counter = out->work_counter;

    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, a.counter);
    GMFA_CHECK_HIP(hipGetLastError());
    const int waves = std::max(1, std::min(a.n_items, resident_waves(k)));
    hipLaunchKernelGGL(k, dim3(waves), dim3(64), 0, s, a);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
}

extern "C" int gmfa_sw_traceback(const gmfa_sw_batch *batch, const gmfa_sw_params *params,
                                 const int32_t *score, const int32_t *query_end,
                                 const int32_t *target_end, const uint32_t *tb,
                                 const int64_t *tb_off, const int64_t *cigar_off, uint32_t *cigar,
                                 int32_t *n_cigar, int32_t *read_start, int32_t *ref_start,
                                 int32_t *counts, void *stream)
{
    GMFA_REQUIRE(batch && params, "NULL argument");
    gmfa_sw_tile tile;
    int rc = gmfa_sw_resolve_tile(params->half_band, params->cells_per_lane,
                                  params->lanes_per_pair, &tile);
    if (rc != GMFA_OK) return rc;
    const int n = batch->n_pairs;
    if (n == 0) return GMFA_OK;
    GMFA_REQUIRE(score && query_end && target_end && tb && tb_off && cigar_off && cigar &&
                     n_cigar && read_start && ref_start && counts,
                 "NULL pointer");
    const bool nm = params->flags & GMFA_SW_NMASK;
    GMFA_REQUIRE(!nm || (batch->qmask && batch->tmask), "GMFA_SW_NMASK needs qmask and tmask");
    TbKArgs a;
    a.qseq = batch->qseq; a.tseq = batch->tseq; a.qmask = batch->qmask; a.tmask = batch->tmask;
    a.q_off = batch->q_off; a.t_off = batch->t_off; a.boff = batch->band_offset;
    a.n_pairs = n; a.hb = params->half_band; a.W = 2 * params->half_band + 1;
    a.C = tile.cells_per_lane; a.Gs = tile.lanes_stored;
    a.ma = params->match; a.mi = params->mismatch; a.go = params->gap_open; a.ge = params->gap_extend;
    a.nm = nm ? 1 : 0;
    a.score = score; a.qend = query_end; a.tend = target_end;
    a.tb = tb; a.tb_off = tb_off; a.cig_off = cigar_off; a.cigar = cigar;
    a.ncig = n_cigar; a.rstart = read_start; a.tstart = ref_start; a.counts = counts;
    hipLaunchKernelGGL(sw_traceback_kernel, dim3((n + 63) / 64), dim3(64), 0, (hipStream_t)stream, a);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
}
