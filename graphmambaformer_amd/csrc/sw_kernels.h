// Device side of the banded affine-gap (Gotoh) Smith-Waterman path: the
// forward kernel, the split-pair kernel and the variant tables.  Included by
// sw.hip (host side) and by the translation units that instantiate the tiles
// (sw_tiles_*.hip, sw_split.hip), so each compiles in parallel.
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
// Values are kept in the row-shifted domain X'(i) = X(i) + i*mismatch, which
// turns the diagonal step into H'(i-1,d) + match_bit*(match+mismatch) (one
// v_mad_u32_u24) and the zero clamp into a per-row scalar.  All arithmetic is exact int32;
// -inf is kNeg with enough headroom that it never wraps.
//
// Traceback planes
// ----------------
// 4 bits per cell, 8 rows per block.  Dword k of lane l's C dwords of a block
// holds the 8 rows of diagonal d = l*C + k, row r in bits 31-4r .. 28-4r as
// (f_cont, not_diag, not_f, e_cont) -- exactly the decisions traceback_banded
// makes (extension.py:249-283):
//   not_diag = H != max(H(i-1,j-1),0) + sub        (diag has priority)
//   not_f    = H != F
//   f_cont   = F(i-1,j) - e > H(i-1,j) - o          (stay in the vertical gap)
//   e_cont   = E(i,j-1) - e > H(i,j-1) - o          (stay in the horizontal gap)
// e_cont is stored with the left neighbour's cell as "G > A" (prefix before
// d-1 beats d-1), which equals the reference's test whenever open > extend;
// for open == extend the reference never continues a horizontal gap and the
// walker ignores the bit.  Bits are shifted into the dwords with v_alignbit
// from sign bits.  A diagonal run of the walk stays inside one dword for up to
// 8 rows.
//
// Best cell: the reference keeps the first row with a strict improvement and
// the first column of that row (extension.py:178-182), i.e. the
// lexicographically smallest (i, j) among cells with the maximal H.  Each lane
// tracks its own first maximum; the group reduces (max H, min i, min d).
#pragma once

#include "common.h"
#include "sw_internal.h"
#include "sw_walk.h"

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
    int32_t walk;  // walk every pair right after its forward pass (CIGAR out)
    int32_t blkstride;   // plane dwords per 8-row block (lanes_stored * C of the tile)
    int32_t split_rows;  // split launches: pairs with at least this many rows run as slices
    int32_t prio_rows;   // offline: pairs with at least this many rows run at raised priority
    WalkCtx wc;
    WalkOut wo;
    // continuous serving (serve_tail == nullptr: one offline batch).  Request
    // k may be claimed once the host has published k < *serve_tail; the
    // flags live in host-coherent memory.
    const int64_t *serve_tail;    // requests published so far
    const int32_t *serve_closed;  // nonzero: nothing more will be published
    int32_t *serve_done;          // [n_pairs] set to 1 when request k is complete
    int32_t *serve_gave_up;       // set to 1 if a wave's wait hit deadline_ticks
    int32_t *serve_started;       // set to 1 by the first wavefront of the launch
    int64_t slab_words;           // traceback plane words per wave slot
    uint64_t deadline_ticks;      // bound of one idle wait (wall-clock ticks)
    uint64_t poll_ticks;          // least interval between two reads of the host words
    ServeDevCtl *dctl;            // device mirrors of the host words (serving)
    // online serving (ring_n > 0): requests arrive in a ring in host-coherent
    // memory while the kernel runs (fields as in ServeLaunch)
    int32_t ring_n, stage_words, cigar_cap;
    const int64_t *ring_tail;
    const ServeDesc *ring_desc;
    const uint32_t *ring_q;
    ServeRes *ring_res;
    uint32_t *ring_cigar;
    int64_t *ring_done;
    uint64_t *ring_cq;
    uint32_t *stage;
    // paged traceback planes (serving; page_shift < 0: per-slot slabs)
    int32_t page_shift, n_pages;
    int32_t *page_owner;
    uint64_t alloc_ticks;
};

// The freshest known tail of a serving queue, and whether the host closed it.
// Reads the device mirrors; when they are older than poll_ticks, the one
// wavefront that wins the compare-and-swap on last_poll reads the host words
// (relaxed, system scope: only the values themselves travel; the requests
// were written before the tail) and raises the mirrors for everyone.  The
// host publishes its final tail before closing, and the mirror tail is raised
// before the mirror closed flag (release / acquire at device scope), so a
// wavefront that sees closed also sees the final tail.
__device__ __forceinline__ int64_t doorbell(ServeDevCtl *d, const int64_t *host_tail,
                                            const int32_t *host_closed, uint64_t poll_ticks,
                                            int &closed)
{
    const bool l0 = (threadIdx.x & 63) == 0;
    closed = __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(&d->closed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    if (closed) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    int64_t tail = uni64((int64_t)__hip_atomic_load(&d->tail, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT));
    if (closed) return tail;
    const uint64_t now = (uint64_t)wall_clock64();
    const uint64_t last = (uint64_t)uni64((int64_t)__hip_atomic_load(
        &d->last_poll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    if (now - last < poll_ticks) return tail;
    unsigned long long won = 0ull;
    if (l0) won = atomicCAS(&d->last_poll, (unsigned long long)last, (unsigned long long)now);
    if ((uint64_t)uni64((int64_t)won) != last) return tail;  // another wavefront polls this time
    const int hc = __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(host_closed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM));
    if (hc) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // the final tail is older
    const int64_t ht =
        uni64(__hip_atomic_load(host_tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM));
    if (ht > tail) {
        if (l0) atomicMax(&d->tail, (unsigned long long)ht);
        tail = ht;
    }
    if (hc) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (l0) __hip_atomic_store(&d->closed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        closed = 1;
    }
    return tail;
}

// Staged serving: wait until request `item` is published or the queue is
// closed; returns the number of published requests.  Every wait ends: at the
// latest after deadline_ticks of the constant-rate wall clock (s_sleep 127 =
// 8128 cycles).  (Scalars by value: taking the kernel argument block by
// reference would move it to scratch memory.)
__device__ __forceinline__ int serve_wait(ServeDevCtl *d, const int64_t *tail_p,
                                          const int32_t *closed_p, int32_t *gave_up,
                                          uint64_t deadline_ticks, uint64_t poll_ticks, int item)
{
    const uint64_t t0 = (uint64_t)wall_clock64();
    int nap = 1;
    for (;;) {
        int closed = 0;
        const int64_t tail = doorbell(d, tail_p, closed_p, poll_ticks, closed);
        const int n = (int)(tail < (int64_t)INT_MAX ? tail : (int64_t)INT_MAX);
        if (item < n || closed) return n;
        if ((uint64_t)wall_clock64() - t0 > deadline_ticks) {
            if (threadIdx.x == 0)
                __hip_atomic_store(gave_up, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            return n;
        }
        for (int z = 0; z < nap; ++z) __builtin_amdgcn_s_sleep(127);
        nap = nap < 16 ? 2 * nap : 16;
    }
}

// Online serving: claim the next published, unclaimed ticket.  A wave claims
// only a ticket the host has already published (compare-and-swap on the
// device head below the tail), so a wave that stops waiting holds nothing and
// the host can relaunch the class later without losing a request.  Returns
// false once the queue is closed and drained, or after idle_ticks of the
// constant-rate wall clock without a claim (the wave then exits; the host
// relaunches the class when requests are pending and every wave has gone).
__device__ __forceinline__ bool serve_claim(ServeDevCtl *d, const int64_t *tail_p,
                                            const int32_t *closed_p, uint64_t idle_ticks,
                                            uint64_t poll_ticks, int64_t &ticket)
{
    const uint64_t t0 = (uint64_t)wall_clock64();
    int nap = 1;
    for (;;) {
        int closed = 0;
        const int64_t tail = doorbell(d, tail_p, closed_p, poll_ticks, closed);
        int64_t head = uni64((int64_t)__hip_atomic_load(&d->head, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT));
        while (head < tail) {
            unsigned long long got = 0ull;
            if ((threadIdx.x & 63) == 0) got = atomicCAS(&d->head, (unsigned long long)head,
                                                         (unsigned long long)head + 1ull);
            const int64_t g = uni64((int64_t)got);
            if (g == head) {
                ticket = head;
                return true;
            }
            head = g;
        }
        if (closed) return false;
        if ((uint64_t)wall_clock64() - t0 > idle_ticks) return false;
        for (int z = 0; z < nap; ++z) __builtin_amdgcn_s_sleep(127);
        nap = nap < 16 ? 2 * nap : 16;
    }
}

// Paged traceback planes (serving).  A request's planes live in pages of
// 2^shift 8-row blocks taken from its class's pool, all or nothing (a short
// pass gives every page back, so no two waves each hold part of what the
// other needs).  A pass starts at a shared hint that moves by `need` per
// request, so consecutive requests look at consecutive pages; it reads 64
// owner words at a time and tries compare-and-swap only on as many of the
// free-looking ones as it still needs (no page is won and handed back).  The
// page list goes to `pl`, the wavefront's own LDS array (one wavefront per
// workgroup).  False after `bound` ticks without enough free pages.
__device__ __forceinline__ bool take_pages(int32_t *owner, int n_pages, int need,
                                           unsigned long long *hint, int32_t *pl, uint64_t bound)
{
    const int lane = (int)(threadIdx.x & 63);
    const uint64_t t0 = (uint64_t)wall_clock64();
    int nap = 1;
    for (;;) {
        unsigned long long h0 = 0ull;
        if (lane == 0) h0 = atomicAdd(hint, (unsigned long long)need);
        const int start = (int)((uint64_t)uni64((int64_t)h0) % (uint64_t)n_pages);
        int got = 0;
        for (int scanned = 0; scanned < n_pages && got < need;) {
            const int pg = (int)((uint32_t)(start + scanned + lane) % (uint32_t)n_pages);
            const bool in = scanned + lane < n_pages;
            const bool looks_free =
                in && __hip_atomic_load(&owner[pg], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
            const uint64_t mf = __ballot(looks_free);
            const int rf = (int)__builtin_amdgcn_mbcnt_hi(
                (uint32_t)(mf >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mf, 0u));
            const bool tried = looks_free && rf < need - got;
            const bool won = tried && atomicCAS(&owner[pg], 0, 1) == 0;
            const uint64_t mw = __ballot(won);
            if (won) {
                const int rw = (int)__builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(mw >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mw, 0u));
                pl[got + rw] = pg;
            }
            got += (int)__popcll(mw);
            // past the last page tried (free-looking pages beyond it stay for
            // the next pass), or past all 64 when fewer were tried than looked free
            const uint64_t mt = __ballot(tried);
            scanned += (int)__popcll(mf) > (int)__popcll(mt) ? 64 - (int)__clzll(mt) : 64;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (got >= need) return true;
        for (int e = lane; e < got; e += 64) atomicExch(&owner[pl[e]], 0);
        if ((uint64_t)wall_clock64() - t0 > bound) return false;
        for (int z = 0; z < nap; ++z) __builtin_amdgcn_s_sleep(127);
        nap = nap < 16 ? 2 * nap : 16;
    }
}

// called by the whole wavefront
__device__ __forceinline__ void free_pages(int32_t *owner, const int32_t *pl, int need)
{
    for (int e = (int)(threadIdx.x & 63); e < need; e += 64) atomicExch(&owner[pl[e]], 0);
}

// Physical page of plane page `page` (wave-uniform).  One LDS address for
// every lane, so the value is right whichever lanes are active (a register
// table read with readlane is not: a lane outside the EXEC mask, such as a
// band lane past the stored ones during the plane store, holds no value).
__device__ __forceinline__ uint32_t page_of(const int32_t *pl, int page)
{
    return (uint32_t)__builtin_amdgcn_readfirstlane(pl[page]);
}

// Plane window of the wavefront walk (lane l: block top - (l >> 4),
// diagonal d0 + (l & 15); 0 outside the planes) from paged planes: the page
// list `pl` in LDS (each lane reads its block's page; the walk runs with
// every lane active), pages of 2^shift blocks of blkstride dwords.
__device__ __forceinline__ uint32_t plane_win_load_paged(const uint32_t *tb, const int32_t *pl,
                                                         int shift, int64_t blkstride, int top,
                                                         int d0, int W)
{
    const int l = (int)(threadIdx.x & 63);
    const int b = top - (l >> 4), d = d0 + (l & 15);
    if (b < 0 || d < 0 || d >= W) return 0u;
    const int64_t pg = (int64_t)(uint32_t)pl[b >> shift];
    return tb[((pg << shift) + (b & ((1 << shift) - 1))) * blkstride + d];
}

// A store into host-coherent memory (the online rings): system scope, so it
// writes through the L2.  A release at system scope writes the L2's dirty
// lines back, but the code generated for the completion record's release
// store has been seen to issue that write-back and then the record's store
// without waiting in between, so a host that saw the record could still read
// the entry's old contents; write-through stores, waited for by the release
// fence before the record, leave nothing to write back.
__device__ __forceinline__ void host_store(void *p, uint32_t v)
{
    __hip_atomic_store(static_cast<uint32_t *>(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Online serving walk: the runs go straight to the request's ring entry in
// host memory.  Run x (counted from the alignment's end) belongs at
// cig[cap - 1 - x], so the entry ends with the runs in forward order and no
// reversal pass reads host memory back; runs are gathered 64 at a time (run x
// in lane x & 63) and stored by the whole wavefront at once.  Runs past cap
// are counted but not stored.  The outputs record follows as one store per
// field lane.
template <int C, bool NM, typename PLd>
__device__ __forceinline__ void walk_pair_ring(const WalkCtx &c, uint32_t *cig, int cap,
                                               ServeRes *res, int i, int j, int h, int off,
                                               int64_t qo, int64_t to, PLd pld,
                                               uint64_t claim_tick)
{
    int lane = (int)(threadIdx.x & 63);
    asm volatile("" : "+v"(lane));  // keep lane-derived addresses inside the request loop
    int nr = 0;
    int cnt[4] = {0, 0, 0, 0};
    uint32_t buf = 0u;
    auto put = [&](int x, uint32_t vrun) {
        if (lane == (x & 63)) buf = vrun;
        if ((x & 63) == 63) {
            const int y = cap - 1 - (x - 63 + lane);
            if (y >= 0) host_store(cig + y, buf);
        }
    };
    const int score = h, qend = i, tend = j;
    walk_steps_vec<C, NM>(c, i, j, h, off, qo, to, put, nr, cnt, pld, SeqWin{}, SeqWin{});
    if (lane < (nr & 63)) {
        const int y = cap - 1 - ((nr & ~63) + lane);
        if (y >= 0) host_store(cig + y, buf);
    }
    const uint64_t done_tick = (uint64_t)wall_clock64();
    const int v = lane == 0 ? score : lane == 1 ? qend : lane == 2 ? tend : lane == 3 ? i
                : lane == 4 ? j : lane == 5 ? nr : lane == 6 ? cnt[0] : lane == 7 ? cnt[1]
                : lane == 8 ? cnt[2] : lane == 9 ? cnt[3]
                : lane == 10 ? (int)(uint32_t)claim_tick : lane == 11 ? (int)(claim_tick >> 32)
                : lane == 12 ? (int)(uint32_t)done_tick : (int)(done_tick >> 32);
    if (lane < 14) host_store(reinterpret_cast<uint32_t *>(res) + lane, (uint32_t)v);
}

struct TbKArgs {
    WalkCtx c;
    WalkOut o;
    const int64_t *q_off, *t_off;
    const int32_t *boff;
    int32_t n_pairs;
    int32_t per_wave;  // pairs per wavefront (lanes 0 .. per_wave-1 walk)
    const int32_t *score, *qend, *tend;
    const uint32_t *tb;
    const int64_t *tb_off;
};

// Standalone walk: one lane per pair (up to 64 pairs per wavefront), the
// current plane dword and 16-base query / target words cached in registers.
template <int C>
__global__ __launch_bounds__(64) void sw_traceback_kernel(TbKArgs a)
{
    const int p = (int)(blockIdx.x * a.per_wave + threadIdx.x);
    if ((int)threadIdx.x >= a.per_wave || p >= a.n_pairs) return;
    const WalkCtx &c = a.c;
    int i = a.qend[p], j = a.tend[p];
    const int h = a.score[p], off = a.boff[p];
    const int64_t qo = a.q_off[p], to = a.t_off[p];
    const uint32_t *tb = a.tb + a.tb_off[p];
    uint32_t *out = a.o.cigar + a.o.cig_off[p];
    int nr = 0;
    int cnt[4] = {0, 0, 0, 0};
    auto ld = [&](const uint32_t *src, int64_t idx) { return src[idx]; };
    auto put = [&](int x, uint32_t vrun) { out[x] = vrun; };
    walk_steps<C>(c, i, j, h, off, qo, to, tb, ld, put, nr, cnt);
    for (int x = 0, y = nr - 1; x < y; ++x, --y) {
        const uint32_t t = out[x];
        out[x] = out[y];
        out[y] = t;
    }
    a.o.ncig[p] = nr;
    a.o.rstart[p] = i;
    a.o.tstart[p] = j;
    a.o.counts[4 * p + 0] = cnt[0];
    a.o.counts[4 * p + 1] = cnt[1];
    a.o.counts[4 * p + 2] = cnt[2];
    a.o.counts[4 * p + 3] = cnt[3];
}

// Standalone walk of a batch whose pairs would each get a wavefront of their
// own anyway (at most 8 wavefronts per SIMD): a wavefront per pair (every
// lane active, state wave-uniform) reading planes and sequences through the
// lane windows, so each pair's dependent loads come once per window instead
// of once per block or 16-base word.
template <int C>
__global__ __launch_bounds__(64) void sw_traceback_wave_kernel(TbKArgs a)
{
    const int p = (int)blockIdx.x;
    if (p >= a.n_pairs) return;
    walk_pair<C, true>(a.c, a.o, p, a.qend[p], a.tend[p], a.score[p], a.boff[p], a.q_off[p],
                       a.t_off[p], a.tb + a.tb_off[p]);
}

// Scalar-unit operations for the scalar cell of KS == C tiles.  Compiled
// from C, its compares and max chain went to the vector unit; these small
// asm statements keep them on the scalar unit while the scheduler can still
// interleave them with the rows' vector instructions.
__device__ __forceinline__ int s_max(int a, int b)
{
    int r;
    asm("s_max_i32 %0, %1, %2" : "=s"(r) : "s"(a), "s"(b) : "scc");
    return r;
}
// c ? a : b for a nonzero test of c
__device__ __forceinline__ int s_sel_nz(int c, int a, int b)
{
    int r;
    asm("s_cmp_lg_u32 %1, 0\n\ts_cselect_b32 %0, %2, %3" : "=s"(r) : "s"(c), "s"(a), "s"(b) : "scc");
    return r;
}
// x < 0 ? a : b
__device__ __forceinline__ int s_sel_neg(int x, int a, int b)
{
    int r;
    asm("s_cmp_lt_i32 %1, 0\n\ts_cselect_b32 %0, %2, %3" : "=s"(r) : "s"(x), "s"(a), "s"(b) : "scc");
    return r;
}
// a == b ? x : 0
__device__ __forceinline__ int s_sel_eq(int a, int b, int x)
{
    int r;
    asm("s_cmp_eq_u32 %1, %2\n\ts_cselect_b32 %0, %3, 0" : "=s"(r) : "s"(a), "s"(b), "s"(x) : "scc");
    return r;
}
// (t > (bK | 31)) ? (bK, bi) = (t, i)
__device__ __forceinline__ void s_best(int t, int i, int &bK, int &bi)
{
    int u;
    asm("s_or_b32 %[u], %[bK], 31\n\t"
        "s_cmp_gt_i32 %[t], %[u]\n\t"
        "s_cselect_b32 %[bK], %[t], %[bK]\n\t"
        "s_cselect_b32 %[bi], %[i], %[bi]"
        : [bK] "+s"(bK), [bi] "+s"(bi), [u] "=&s"(u)
        : [t] "s"(t), [i] "s"(i)
        : "scc");
}

// One row of the scalar cell (wave-uniform state and inputs):
//   dg = H(i-1, d) + (match ? ma + mi : 0)      (row-shifted domain)
//   f  = max(zprev - cH, kNeg - cF)               (no band cell at d+1)
//   e  = g63 + max(d63, 0) - (go + (C-1)*ge)      (lane 63's prefix, one on)
//   h  = max(dg, f, zrow, e); out of the matrix (xl < 0): h = zrow, f = kNeg
// then the plane bits (f_cont 0: F(i-1, d+1) is out of band; not_diag,
// not_f, e_cont = sign of lane 63's A - G) and the first strict maximum.
template <bool SLOW>
__device__ __forceinline__ void scalar_cell(int &Hs, int &Fs, uint32_t &acc, int &bK, int &bi,
                                            int tf, int qc, int mf, int mms, int zprev, int zrow,
                                            int cH, int bv, int d63, int g63, int k1, int imi5,
                                            int i, int xl, int neg)
{
    // the asm takes SGPRs; values the compiler placed in VGPRs (the carried
    // state also feeds the DPP "old" operand) move over here, the others
    // cost nothing
#define GMFA_UNI(x) x = __builtin_amdgcn_readfirstlane(x)
    GMFA_UNI(Hs); GMFA_UNI(Fs); GMFA_UNI(bK); GMFA_UNI(bi);
    acc = (uint32_t)__builtin_amdgcn_readfirstlane((int)acc);
    GMFA_UNI(tf); GMFA_UNI(qc); GMFA_UNI(mf); GMFA_UNI(mms); GMFA_UNI(zprev); GMFA_UNI(zrow);
    GMFA_UNI(cH); GMFA_UNI(bv); GMFA_UNI(d63); GMFA_UNI(g63); GMFA_UNI(k1); GMFA_UNI(imi5);
    GMFA_UNI(i); GMFA_UNI(xl); GMFA_UNI(neg);
#undef GMFA_UNI
    const int sub = s_sel_nz(mf, 0, s_sel_eq(tf, qc, mms));
    const int dg = (int)((uint32_t)Hs + (uint32_t)sub);
    int fs = s_max((int)((uint32_t)zprev - (uint32_t)cH), bv);
    const int es = (int)((uint32_t)s_max(d63, 0) + (uint32_t)g63 - (uint32_t)k1);
    int hs = s_max(s_max(s_max(dg, fs), zrow), es);
    if constexpr (SLOW) {
        hs = s_sel_neg(xl, zrow, hs);
        fs = s_sel_neg(xl, neg, fs);
    }
    const uint32_t bits = ((((uint32_t)dg - (uint32_t)hs) >> 31) << 2) |
                          ((((uint32_t)fs - (uint32_t)hs) >> 31) << 1) | ((uint32_t)d63 >> 31);
    acc = (acc << 4) | bits;
    s_best((int)((((uint32_t)hs << 5) | 31u) - (uint32_t)imi5), i, bK, bi);
    Hs = hs;
    Fs = fs;
}

// -------------------------------------------------------------------------
// Split pairs: the longest pairs of a batch over several wavefronts
// -------------------------------------------------------------------------
// One pair on one wavefront cannot finish before its own row chain does: the
// longest pair of a batch of uneven lengths sets the batch's end.  Here a
// workgroup of K wavefronts takes such a pair and gives wavefront s the band
// slice of diagonals [s*64*C, (s+1)*64*C): lane l of slice s owns diagonals
// s*64*C + l*C .. +C-1, as lane s*64 + l of one K*64-lane wavefront would.
// Two values cross a slice boundary per query row:
//   * to the right, the horizontal prefix P(i) = max over the diagonals left
//     of the boundary of M(i,d) + d*e (and the e_cont difference of the last
//     diagonal), needed by the whole row i of the next slice;
//   * to the left, the vertical input H(i-1, dN), F(i-1, dN) of the slice's
//     last diagonal, where dN is the next slice's first diagonal.
// The first makes slice s depend on slice s-1 in the same row, the second on
// slice s+1 in the row above, so the two cannot be ordered row by row.  The
// slice therefore computes the cell (i-1, dN) itself, a "ghost" cell on the
// scalar unit, from its own prefix P(i-1) and slice s+1's row i-2 values at
// dN and dN+1.  Then slice s runs row i at step i + s, reading values its
// neighbours wrote in step i + s - 1 through LDS (two parity slots), with one
// workgroup barrier per step.  The ghost cell equals what slice s+1 computes
// for the same cell (exact int32), so every cell, plane nibble and best key is
// the one-wavefront result; planes keep the tile's layout (dword d of a block
// = band diagonal d), so the walkers are unchanged.
//
// A split launch is a persistent grid of K-wavefront workgroups beside the
// plain kernel: a workgroup claims the next of the split pairs (the
// longest-first prefix of the batch whose queries have at least split_rows
// rows) and runs it as K slices (the slices hand their best keys to
// wavefront 0, which then walks the CIGAR); when the split pairs run out it
// exits, and the plain kernel, launched beside it, takes the other pairs.

// LDS words per slice and parity slot: P, e_cont difference, H(dN), H(dN+1),
// F(dN+1) (slot s + 1 holds slice s's, read by slices s - 1 and s + 1)
constexpr int kXw = 8;

// Forward pass of one pair on slice s of K.  Returns the wavefront's best-cell key
// (H << 40 | (2^20-1-i) << 12 | (4095-d)), 0 if no positive cell.
// KST: compile-time register of diagonal W-1 in its lane (-1: runtime).
// pl != nullptr: the planes are paged (serving): block b of the pair lives in
// page pl[b >> a.page_shift] of the pool at a.tb, tb_base unused.
template <int C, int K, int KST>
__device__ __forceinline__ uint64_t slice_forward(const SwKArgs &a, int s, int qlen, int tlen,
                                                  int off, int64_t qo, int64_t to,
                                                  int64_t tb_base, int *xch,
                                                  const int32_t *pl = nullptr)
{
    constexpr int NCW = (C + 15) / 16;  // target words feeding the cells
    constexpr int NA = NCW + 1;         // aligned window words
    constexpr int NWR = NA + 1;         // raw words loaded per block
    static_assert(K > 1 && C >= 2 && C <= 24, "slices: K > 1, 2 <= C <= 24 (ghost bases in two window words)");
    const int gl = lane_id();  // (threadIdx.x & 63 took 256 VGPRs + AGPRs here)
    const int W = a.W, hb = a.hb, ma = a.ma, mi = a.mi, go = a.go, ge = a.ge;
    const int sbase = s * 64 * C;      // first diagonal of the slice
    const int gbase = sbase + gl * C;  // first diagonal of this lane
    const int glane = s * 64 + gl;     // lane index over the whole band
    const int lstar = (W - 1) / C;
    const int kstar = KST >= 0 ? KST : (W - 1) % C;
    const bool dead_hi = glane >= lstar;
    const bool guard = (kstar == C - 1) && (glane == lstar);
    const bool lane_live = glane <= lstar;
    const int lCe = gbase * ge;
    const int cH = go - mi, cF = ge - mi;
    const uint32_t mms = (uint32_t)(ma + mi);
    const int64_t t_hi = a.t_words - 1, q_hi = a.q_words - 1;
    const bool has_left = K > 1 && s > 0;
    const bool has_right = K > 1 && s + 1 < K;
    const int dN = sbase + 64 * C;  // the next slice's first diagonal

    int Hp[C], Fp[C];
    uint32_t acc[C];
#pragma unroll
    for (int k = 0; k < C; ++k) { Hp[k] = 0; Fp[k] = kNeg; acc[k] = 0u; }
    int bestK = 255, besti = 0;
    // split state: this slice's prefix through its last diagonal in the
    // previous row (the ghost cell's E), and the previous row's query base
    int Pprev = 0;
    uint32_t qlast = 0u;

    const int nblk = (qlen + 7) >> 3;
    int64_t e_t = to + (int64_t)(off - hb + gbase);  // target element of cell 0, first row
    uint32_t wt[NWR];
    auto load_t = [&](int64_t e) __attribute__((always_inline)) {
        const int64_t wb = e >> 4;
#pragma unroll
        for (int x = 0; x < NWR; ++x) wt[x] = a.tseq[clampi64(wb + x, t_hi)];
    };
    if (nblk > 0) load_t(e_t);
    int fb0 = ((hb + 1 - off) - 1 + 7) >> 3;
    fb0 = max(fb0, 0);
    int fb1 = min(qlen, tlen - hb - off);
    fb1 = fb1 >= 8 ? (fb1 >> 3) : 0;
    fb1 = min(max(fb1, fb0), nblk);
    fb0 = min(fb0, fb1);

    // steps before this slice's first row (the slices to its left start first)
    if constexpr (K > 1) {
        for (int x = 0; x < s; ++x) __syncthreads();
    }

    for (int b = 0; b < nblk; ++b) {
        const int i0 = 8 * b + 1;
        int bk = INT_MIN;
        uint32_t ta[NA];
        {
            const uint32_t sh = (uint32_t)(e_t & 15) * 2u;
#pragma unroll
            for (int x = 0; x < NA; ++x) ta[x] = __builtin_amdgcn_alignbit(wt[x + 1], wt[x], sh);
        }
        // lane 63's bases: the ghost cell of row i-1 sits at position C + r - 1
        uint64_t tw63 = 0u;
        if constexpr (K > 1)
            tw63 = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)ta[1], 63) << 32) |
                   (uint32_t)__builtin_amdgcn_readlane((int)ta[0], 63);
        e_t += 8;
        if (b + 1 < nblk) load_t(e_t);
        uint32_t qbits;
        {
            const int64_t eq = qo + 8 * b;
            const int64_t wq = eq >> 4;
            const uint32_t sh = (uint32_t)(eq & 15) * 2u;
            const uint32_t w0 = a.qseq[clampi64(wq, q_hi)];
            const uint32_t w1 = a.qseq[clampi64(wq + 1, q_hi)];
            qbits = (uint32_t)__builtin_amdgcn_readfirstlane((int)__builtin_amdgcn_alignbit(w1, w0, sh));
        }

        auto row = [&](auto slow_tag, const int r) __attribute__((always_inline)) {
            constexpr bool SLOW = decltype(slow_tag)::value;
            const int i = i0 + r;
            const int zrow = i * mi;  // H = 0 in this row's domain
            const int zprev = zrow - mi;
            const uint32_t qc = (qbits >> (2 * r)) & 3u;
            const uint32_t qpat = qc * 0x55555555u;
            // this step's inputs, written by the neighbours in the previous step
            int Pin = 0, dlin = 0, ghH = zprev, ghF = kNeg;
            const int par = (i + s) & 1;
            if constexpr (K > 1) {
                // slot x + 1 holds slice x's outputs; slots 0 and K + 1 stand
                // for the missing neighbours of the edge slices (P = 0), so
                // both reads are unconditional and share one wait
                const int *in = xch + (par ^ 1) * (K + 2) * kXw;
                const int vP = in[s * kXw + 0], vD = in[s * kXw + 1];
                const int vH0 = in[(s + 2) * kXw + 2], vH1 = in[(s + 2) * kXw + 3];
                const int vF1 = in[(s + 2) * kXw + 4];
                Pin = __builtin_amdgcn_readfirstlane(vP);
                dlin = __builtin_amdgcn_readfirstlane(vD);
                // the ghost cell (i-1, dN) from slice s+1's row i-2 (the band
                // guard for the last slice)
                const int rH0 = __builtin_amdgcn_readfirstlane(vH0);
                const int rH1 = __builtin_amdgcn_readfirstlane(vH1);
                const int rF1 = __builtin_amdgcn_readfirstlane(vF1);
                const int ig = i - 1;
                const int jg = ig + off - hb + dN;  // 1-based column
                const bool inside = has_right && ig >= 1 && ig <= qlen && jg >= 1 && jg <= tlen;
                const uint32_t tq = (uint32_t)(tw63 >> (2 * (C + r - 1))) & 3u;
                const uint32_t qq = r > 0 ? ((qbits >> (2 * (r - 1))) & 3u) : qlast;
                const int hd = rH0 + (tq == qq ? (int)mms : 0);
                const int fg = max(rH1 - cH, rF1 - cF);
                const int eg = Pprev - go - (dN - 1) * ge;
                const int hg = max(max(hd, fg), max(zprev, eg));
                ghH = inside ? hg : zprev;
                ghF = inside ? fg : kNeg;
            }
            uint32_t z[NCW];
#pragma unroll
            for (int x = 0; x < NCW; ++x) {
                uint32_t y = __builtin_amdgcn_alignbit(ta[x + 1], ta[x], 2 * r) ^ qpat;
                y |= y >> 1;
                z[x] = ~y & 0x55555555u;  // 2-bit field k = 1 <=> cell k matches
            }
            int lo = 0;
            uint32_t span = 0;
            if constexpr (SLOW) {
                const int dlo = hb + 1 - off - i;
                const int dhi = tlen + hb - off - i;
                const int l0 = max(dlo, 0);
                const int h0 = (i <= qlen) ? min(dhi, W - 1) : -1;
                lo = l0 - gbase;
                span = (uint32_t)max(h0 - l0 + 1, 0);
            }
            // lane 63's d+1: the ghost cell, or the band guard
            int hN = from_next_lane(ghH, Hp[0]);
            int fN = from_next_lane(ghF, Fp[0]);
            hN = guard ? zprev : hN;
            fN = guard ? kNeg : fN;

            int Mv[C], Fv[C], Dg[C], Av[C], Gv[C];
#pragma unroll
            for (int k = 0; k < C; ++k) {
                const int uh = (k + 1 < C) ? Hp[(k + 1) % C] : hN;
                const int uf = (k + 1 < C) ? Fp[(k + 1) % C] : fN;
                const int av = uh - cH;
                const int bv = uf - cF;
                Fv[k] = max(av, bv);
                acc[k] = insert_sign(acc[k], av - bv);  // f_cont
                const uint32_t mt = __builtin_amdgcn_ubfe(z[k >> 4], 2 * (k & 15), 2);
                Dg[k] = mad_u24(mt, mms, Hp[k]);
                const int m = max(max(Dg[k], Fv[k]), zrow);
                Mv[k] = m;
                int A = m + k * ge;
                if constexpr (SLOW) {
                    const bool live = (uint32_t)(k - lo) < span;
                    A = live ? A : kNeg;
                }
                Av[k] = A;
            }
            int T = Av[0];
#pragma unroll
            for (int k = 1; k < C; k += 2)
                T = (k + 1 < C) ? max(max(T, Av[k]), Av[(k + 1) % C]) : max(T, Av[k]);
            const int S = seg_scan_max<64>(T + lCe);
            int X = from_prev_lane_or0(S);
            if constexpr (K > 1) X = max(X, has_left ? Pin : INT_MIN);
            const int Xl = X - lCe;
            Gv[0] = Xl;
#pragma unroll
            for (int k = 1; k < C; ++k) Gv[k] = max(Gv[k - 1], Av[k - 1]);
            const int Glast = Gv[C - 1];
            const int dlast = Av[C - 1] - Glast;
            int dprev = from_prev_lane_or0(dlast);
            if constexpr (K > 1) dprev = (has_left && gl == 0) ? dlin : dprev;
            int rk = INT_MIN;
            int Gkm1 = Xl;
#pragma unroll
            for (int k = 0; k < C; ++k) {
                const int Gk = Gv[k];
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
                acc[k] = insert_sign(acc[k], Dg[k] - H);  // not_diag
                acc[k] = insert_sign(acc[k], Fk - H);     // not_f
                acc[k] = insert_sign(acc[k], k ? (Av[k ? k - 1 : 0] - Gkm1) : dprev);  // e_cont
                rk = max(rk, (int)(((uint32_t)H << 8) | (uint32_t)(31 - k)));
                Hp[k] = H;
                Fv[k] = Fk;
                Gkm1 = Gk;
            }
#pragma unroll
            for (int k = 0; k < C; ++k) Fp[k] = Fv[k];
            bk = max(bk, rk - (int)(((uint32_t)(i * mi) << 8) - ((uint32_t)(7 - r) << 5)));
            // finish the row's plane bits and best key here: left free, the
            // compiler defers them to the end of the block and keeps every
            // row's temporaries alive meanwhile (256 VGPRs + AGPRs)
#pragma unroll
            for (int k = 0; k < C; ++k) asm volatile("" : "+v"(acc[k]));
            asm volatile("" : "+v"(bk));
            if constexpr (K > 1) {
                // this step's outputs, then the step barrier
                int *out = xch + par * (K + 2) * kXw + (s + 1) * kXw;
                const int P = max(__builtin_amdgcn_readlane(S, 63), Pin);
                if (gl == 63) {
                    out[0] = P;
                    out[1] = dlast;
                }
                Pprev = P;
                if (gl == 0) {
                    out[2] = Hp[0];
                    out[3] = Hp[1];
                    out[4] = Fp[1];
                }
                __syncthreads();
            }
        };
        if (b >= fb0 && b < fb1) {
#pragma unroll
            for (int r = 0; r < 8; ++r) row(std::false_type{}, r);
        } else {
#pragma unroll
            for (int r = 0; r < 8; ++r) row(std::true_type{}, r);
        }
        qlast = (qbits >> 14) & 3u;
        if (bk > (bestK | 255)) {
            bestK = bk;
            besti = i0 + 7 - ((bk >> 5) & 7);
        }
        if (lane_live) {
            uint32_t *dst = a.tb + tb_base + (int64_t)b * a.blkstride + gbase;
            if (pl) {
                const int32_t pmask = (1 << a.page_shift) - 1;
                dst = a.tb + ((int64_t)page_of(pl, b >> a.page_shift) << a.page_shift) * a.blkstride +
                      (int64_t)(b & pmask) * a.blkstride + gbase;
            }
#pragma unroll
            for (int k = 0; k < C; ++k) dst[k] = acc[k];
        }
    }
    // steps after this slice's last row (the slices to its right end later)
    if constexpr (K > 1) {
        for (int x = s + 1; x < K; ++x) __syncthreads();
    }

    int bH = bestK >> 8;
    const int bd = gbase + (31 - (bestK & 31));
    if (!lane_live || besti == 0) bH = 0;
    uint64_t key = ((uint64_t)(uint32_t)bH << 40) |
                   ((uint64_t)(0xFFFFFu - (uint32_t)besti) << 12) | (uint64_t)(0xFFFu - (uint32_t)bd);
    if (bH == 0) key = 0;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const uint32_t lo32 = (uint32_t)__shfl_xor((int)(uint32_t)key, o);
        const uint32_t hi32 = (uint32_t)__shfl_xor((int)(uint32_t)(key >> 32), o);
        const uint64_t other = ((uint64_t)hi32 << 32) | lo32;
        key = other > key ? other : key;
    }
    return key;
}

// The split kernel (see the comment above slice_forward): a workgroup of KSL
// wavefronts counts itself resident (ctl[2]), then claims the next split pair
// (ctl[0], up to ctl[1] pairs, the longest-first prefix of the batch's order
// that sw_split_count_kernel found), runs it as KSL slices at raised wave priority (so the plain
// kernel's waves on the same SIMDs fill the slices' stalls instead of
// delaying them), and wavefront 0 walks it while the others wait for the next
// claim.  CW: the regular tile's cells per lane (the plane layout the walk
// reads).
template <int CS, int KSL, int KSS, int CW>
__global__ __launch_bounds__(64 * KSL) void sw_split_kernel(SwKArgs a, uint32_t *ctl)
{
    __shared__ int xch[2 * (KSL + 2) * kXw];
    __shared__ int claim[2];
    __shared__ unsigned long long keys[KSL];
    // wave-uniform (from threadIdx.x the compiler would treat the slice index
    // as divergent and mask every slice-dependent branch)
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = lane_id();
    // resident: the plain launch waits for every split workgroup to be here
    if (threadIdx.x == 0) atomicAdd(&ctl[2], 1u);
    for (;;) {
        __syncthreads();  // the previous pair (wavefront 0's walk included) is done
        if (threadIdx.x == 0) {
            claim[0] = (int)atomicAdd(&ctl[0], 1u);
            claim[1] = (int)__hip_atomic_load(&ctl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(claim[0]);
        if (item >= __builtin_amdgcn_readfirstlane(claim[1])) return;
        const int p = __builtin_amdgcn_readfirstlane(a.order[item]);
        const int qlen = __builtin_amdgcn_readfirstlane(a.qlen[p]);
        const int tlen = __builtin_amdgcn_readfirstlane(a.tlen[p]);
        const int off = __builtin_amdgcn_readfirstlane(a.boff[p]);
        const int64_t qo = uni64(a.q_off[p]), to = uni64(a.t_off[p]);
        const int64_t tb_base = uni64(a.tb_off[p]);
        // row 0 of every slice in both parity slots: H = 0, F = -inf
        for (int x = threadIdx.x; x < 2 * (KSL + 2) * kXw; x += 64 * KSL)
            xch[x] = ((x % kXw) == 4) ? kNeg : 0;
        __syncthreads();
        __builtin_amdgcn_s_setprio(3);
        const uint64_t key = slice_forward<CS, KSL, KSS>(a, wave, qlen, tlen, off, qo, to,
                                                         tb_base, xch);
        if (lane == 0) keys[wave] = key;
        __syncthreads();  // every slice's planes and key
        if (wave == 0) {
            uint64_t best = 0;
#pragma unroll
            for (int w = 0; w < KSL; ++w) best = keys[w] > best ? keys[w] : best;
            best = (uint64_t)uni64((int64_t)best);
            int res_s = 0, res_i = 0, res_j = 0;
            if (best != 0) {
                res_s = (int)(best >> 40);
                res_i = (int)(0xFFFFFu - (uint32_t)((best >> 12) & 0xFFFFFu));
                res_j = res_i + off - a.hb + (int)(0xFFFu - (uint32_t)(best & 0xFFFu));
            }
            if (lane == 0) {
                a.score[p] = res_s;
                a.qend[p] = res_i;
                a.tend[p] = res_j;
            }
            walk_pair<CW, false>(a.wc, a.wo, p, res_i, res_j, res_s, off, qo, to, a.tb + tb_base);
        }
        __builtin_amdgcn_s_setprio(0);
    }
}
// Row modes of the forward kernel (see its row lambda).
constexpr int kRowFast = 0, kRowEdge = 1, kRowMasked = 2;

// 2-bit fields 0 .. n-1 of a word set (n clamped to [0, 16]).
__device__ __forceinline__ uint32_t low_fields(int n)
{
    n = min(max(n, 0), 16);
    return n >= 16 ? ~0u : ((1u << (2 * n)) - 1u);
}

// Serving with split requests (a staged class of gmfa_serve_run whose params
// ask for a split): a persistent grid of KSL-wavefront workgroups, one
// request at a time per workgroup, every request run as KSL band slices
// (slice_forward), so a class of long requests has each one's rows spread
// over KSL wavefronts.  Wavefront 0 claims the next request of the class
// (the class's work counter), waits until the host has published it, takes
// its traceback pages (or uses the workgroup's slab), and after the slices
// walks it and raises its completion flag.  Every wait is bounded as in the
// one-wavefront serving kernel.
template <int CS, int KSL, int KSS, int CW>
__global__ __launch_bounds__(64 * KSL) void sw_split_serve_kernel(SwKArgs a)
{
    __shared__ int xch[2 * (KSL + 2) * kXw];
    __shared__ int claim[3];  // item, published requests, planes ready
    __shared__ unsigned long long keys[KSL];
    __shared__ int32_t page_list[kMaxPages];
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = lane_id();
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.serve_started)
        __hip_atomic_store(a.serve_started, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const bool paged = a.page_shift >= 0;
    for (;;) {
        __syncthreads();  // the previous request is complete
        if (wave == 0) {
            int item = 0;
            if (lane == 0) item = (int)atomicAdd(a.counter, 1u);
            item = __builtin_amdgcn_readfirstlane(item);
            int n = 0, ready = 1;
            if (item < a.n_pairs) {
                n = serve_wait(a.dctl, a.serve_tail, a.serve_closed, a.serve_gave_up,
                               a.deadline_ticks, a.poll_ticks, item);
                if (item < n && paged) {
                    const int ql = __builtin_amdgcn_readfirstlane(a.qlen[item]);
                    const int need = (((ql + 7) >> 3) + (1 << a.page_shift) - 1) >> a.page_shift;
                    if (!take_pages(a.page_owner, a.n_pages, need, &a.dctl->page_hint, page_list,
                                    a.alloc_ticks)) {
                        if (lane == 0)
                            __hip_atomic_store(a.serve_gave_up, 1, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_SYSTEM);
                        ready = 0;
                    }
                }
            }
            if (lane == 0) {
                claim[0] = item;
                claim[1] = n;
                claim[2] = ready;
            }
        }
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(claim[0]);
        if (item >= __builtin_amdgcn_readfirstlane(claim[1])) return;
        const bool ready = __builtin_amdgcn_readfirstlane(claim[2]) != 0;
        const int p = item;  // staged requests in arrival order
        const int qlen = __builtin_amdgcn_readfirstlane(a.qlen[p]);
        const int tlen = __builtin_amdgcn_readfirstlane(a.tlen[p]);
        const int off = __builtin_amdgcn_readfirstlane(a.boff[p]);
        const int64_t qo = uni64(a.q_off[p]), to = uni64(a.t_off[p]);
        const int64_t tb_base = paged ? 0 : (int64_t)blockIdx.x * a.slab_words;
        const int need = paged ? (((qlen + 7) >> 3) + (1 << a.page_shift) - 1) >> a.page_shift : 0;
        uint64_t key = 0;
        if (ready) {
            for (int x = threadIdx.x; x < 2 * (KSL + 2) * kXw; x += 64 * KSL)
                xch[x] = ((x % kXw) == 4) ? kNeg : 0;
            __syncthreads();
            key = slice_forward<CS, KSL, KSS>(a, wave, qlen, tlen, off, qo, to, tb_base, xch,
                                              paged ? page_list : nullptr);
        }
        if (lane == 0) keys[wave] = key;
        __syncthreads();  // every slice's planes and key
        if (wave == 0) {
            uint64_t best = 0;
#pragma unroll
            for (int w = 0; w < KSL; ++w) best = keys[w] > best ? keys[w] : best;
            best = (uint64_t)uni64((int64_t)best);
            int res_s = 0, res_i = 0, res_j = 0;
            if (best != 0) {
                res_s = (int)(best >> 40);
                res_i = (int)(0xFFFFFu - (uint32_t)((best >> 12) & 0xFFFFFu));
                res_j = res_i + off - a.hb + (int)(0xFFFu - (uint32_t)(best & 0xFFFu));
            }
            if (lane == 0) {
                a.score[p] = res_s;
                a.qend[p] = res_i;
                a.tend[p] = res_j;
            }
            if (paged) {
                const int sh = a.page_shift, Wb = a.W;
                const int64_t bs = a.blkstride;
                auto pld = [&](int top, int d0) {
                    return plane_win_load_paged(a.tb, page_list, sh, bs, top, d0, Wb);
                };
                walk_pair_with<CW, false>(a.wc, a.wo, p, res_i, res_j, res_s, off, qo, to, pld,
                                          SeqWin{}, SeqWin{});
            } else {
                walk_pair<CW, false>(a.wc, a.wo, p, res_i, res_j, res_s, off, qo, to,
                                     a.tb + tb_base);
            }
            if (!ready && lane == 0) a.wo.ncig[p] = -1;
        }
        __syncthreads();  // the walk's stores by every lane of wavefront 0, then the flag
        if (wave == 0) {
            if (lane == 0)
                __hip_atomic_store(a.serve_done + p, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            // after that release (dirty plane lines written back), free the pages
            if (paged && ready) free_pages(a.page_owner, page_list, need);
        }
    }
}

// SV: serving variant (persistent queue fed by the host; staged requests of
// gmfa_serve_run or the online ring of gmfa_serve_open), a separate
// instantiation so the offline kernels carry none of its state.
//
// KS == C (XS): a band one diagonal wider than the lanes hold, W = C*64 + 1
// (half_band = 32*C, e.g. 256 -> 8x64 + 1).  The 64 lanes own diagonals
// 0 .. W-2 with no dead lane; diagonal W-1 is a scalar cell (lane 64,
// register 0 of the plane layout) computed on the scalar unit from lane 63's
// prefix and window, while the vector unit, the bound of this kernel, does
// 8 cells per lane instead of 9 with 7 dead lanes.
template <int C, int G, int KS, bool TB, bool NM, bool RK, bool SV = false>
__global__ __launch_bounds__(64) void sw_forward_kernel(SwKArgs a)
{
    constexpr bool XS = KS == C;
    static_assert(!XS || (G == 64 && C + 8 <= 32), "scalar-cell tiles: 64 lanes, C <= 24");
    constexpr int NCW = (C + 15) / 16;  // target words feeding the cells
    constexpr int NA = NCW + 1;         // aligned window words
    constexpr int NWR = NA + 1;         // raw words loaded per block
    // (as lane_id() or threadIdx.x & 63 the 24 x 64 tile took 256 VGPRs +
    // AGPRs instead of 202)
    const int lane = (int)threadIdx.x;
    const int grp = lane / G;
    const int gl = lane % G;

    const int W = a.W, hb = a.hb, ma = a.ma, mi = a.mi, go = a.go, ge = a.ge;
    const int lstar = (W - 1) / C;  // group lane holding diagonal W-1
    // register of diagonal W-1 in that lane: compile time when KS >= 0 (the
    // host picks such a tile only if (W-1) % C == KS)
    const int kstar = KS >= 0 ? KS : (W - 1) % C;
    const bool dead_hi = gl >= lstar;                     // registers k > kstar dead
    const bool guard = (kstar == C - 1) && (gl == lstar); // its d+1 is the band guard
    const bool lane_live = gl <= lstar;
    const uint32_t live_mask = 0u - (uint32_t)lane_live;
    const int lCe = gl * C * ge;
    const int cH = go - mi, cF = ge - mi;
    const uint32_t mms = (uint32_t)(ma + mi);
    const int64_t t_hi = a.t_words - 1, q_hi = a.q_words - 1;

    // online form: a runtime switch of the serving instantiation (a separate
    // instantiation scheduled the DP with up to 90 more VGPRs)
    bool online = false;
    if constexpr (SV) {
        online = a.ring_n > 0;
        // tell the host this class's kernel is running (the host checks that
        // every class runs side by side)
        if (blockIdx.x == 0 && threadIdx.x == 0 && a.serve_started)
            __hip_atomic_store(a.serve_started, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    bool first = true;
    for (;;) {
        int item = 0;
        int n_items = a.n_items;
        // online serving: the request's fields come from its ring entry; its
        // query sits in this wave's stage slot (the a.qseq stream of online
        // launches is the stage array), so the forward pass is the staged one
        int64_t ticket = 0;
        int ring_r = 0;
        int r_qlen = 0, r_tlen = 0, r_off = 0;
        int64_t r_to = 0;
        uint64_t claim_tick = 0;
        if (online) {
            if (!serve_claim(a.dctl, a.ring_tail, a.serve_closed, a.deadline_ticks, a.poll_ticks,
                             ticket))
                break;
            claim_tick = (uint64_t)wall_clock64();
            // The host wrote the entry before publishing the ticket (release).
            // The entry is read with system-scope loads issued after the
            // claim's tail load has returned: an acquire would add an L2
            // invalidation per request under every working wavefront, which
            // only non-coherent loads need.
            ring_r = (int)(ticket & (int64_t)(a.ring_n - 1));  // ring_n is a power of two
            // lane index through an opaque copy: nothing derived from it is
            // hoisted out of the request loop to stay live across the DP
            int ln = lane;
            asm volatile("" : "+v"(ln));
            const uint32_t *dw = reinterpret_cast<const uint32_t *>(a.ring_desc + ring_r);
            const uint32_t dv =
                ln < 6 ? __hip_atomic_load(dw + ln, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) : 0u;
            r_to = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)dv, 1) << 32) |
                             (uint32_t)__builtin_amdgcn_readlane((int)dv, 0));
            r_qlen = __builtin_amdgcn_readlane((int)dv, 2);
            r_tlen = __builtin_amdgcn_readlane((int)dv, 3);
            r_off = __builtin_amdgcn_readlane((int)dv, 4);
            const int nw = __builtin_amdgcn_readlane((int)dv, 5);
            // the query moves once into the stage slot; the forward pass and
            // the walk then read it from device memory
            const int64_t sw2 = 2 * (int64_t)a.stage_words;
            uint32_t *stg = a.stage + (int64_t)blockIdx.x * sw2;
            const uint32_t *src = a.ring_q + (int64_t)ring_r * sw2;
            // four loads in flight per lane before a wait: each is a round
            // trip over the host link
            auto copy = [&](uint32_t *dst, const uint32_t *from) __attribute__((always_inline)) {
                for (int x = ln; x < nw; x += 256) {
                    uint32_t v[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        v[u] = x + 64 * u < nw ? __hip_atomic_load(from + x + 64 * u,
                                                                   __ATOMIC_RELAXED,
                                                                   __HIP_MEMORY_SCOPE_SYSTEM)
                                               : 0u;
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (x + 64 * u < nw) dst[x + 64 * u] = v[u];
                }
            };
            copy(stg, src);
            if (NM) copy(stg + a.stage_words, src + a.stage_words);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        } else {
            // offline launches without a split: a wave's first item is its
            // workgroup index, later ones are claimed after the grid's; a
            // grid that covers every item claims nothing (a claim round trip
            // per wave, all at once at the start, costs more than a short
            // pair's forward pass)
            const bool stat = !SV && a.split_rows == 0;
            if (stat && first) {
                item = (int)blockIdx.x;
            } else if (stat && (int)gridDim.x >= n_items) {
                item = n_items + (int)gridDim.x;  // nothing left; the counter stays untouched
            } else {
                if (lane == 0) item = (int)atomicAdd(a.counter, 1u);
                item = __builtin_amdgcn_readfirstlane(item) + (stat ? (int)gridDim.x : 0);
            }
            first = false;
            // serving: one pair per wavefront (G == 64), bounded by what the host
            // has published
            if constexpr (SV) {
                if (item >= a.n_pairs) break;  // past the class's last request: nothing to wait for
                n_items = serve_wait(a.dctl, a.serve_tail, a.serve_closed, a.serve_gave_up,
                                     a.deadline_ticks, a.poll_ticks, item);
            }
            if (item >= n_items) {
                // every wave ends on one failing claim, so the claim that
                // returns n_items + grid - 1 (after the grid's first items:
                // n_items - 1) is the launch's last: that wave leaves the
                // counter at zero for the next launch (no reset kernel per
                // launch; split launches set it up anew)
                if constexpr (!SV) {
                    if (lane == 0 && item == n_items + (int)gridDim.x - 1) *a.counter = 0u;
                }
                break;
            }
        }

        const int pidx = item * (64 / G) + grp;
        const bool has = online || pidx < (SV ? n_items : a.n_pairs);
        // a one-pair-per-wavefront grid covering the batch needs no
        // processing order (one dependent load less before the pair's fields)
        const bool use_order = a.order && (G < 64 || SV || a.split_rows > 0 ||
                                           (int)gridDim.x < n_items);
        const int p = (has && !online) ? (use_order ? a.order[pidx] : pidx) : 0;
        const int qlen = online ? r_qlen : uni<G>(has ? a.qlen[p] : 0);
        const int tlen = online ? r_tlen : uni<G>(has ? a.tlen[p] : 0);
        const int off = online ? r_off : uni<G>(has ? a.boff[p] : 0);
        const int64_t qo = online ? (int64_t)blockIdx.x * 32 * (int64_t)a.stage_words
                                  : (has ? a.q_off[p] : 0);
        const int64_t to = online ? r_to : (has ? a.t_off[p] : 0);
        const int maxq = (G == 64) ? qlen : wave_max(qlen);
        // the longest pairs set a batch's end (and a served class's tail
        // latency): at raised wave priority they keep their SIMD's issue
        // slots, the other waves there fill their stalls
        const bool prio = a.prio_rows > 0 && maxq >= a.prio_rows;
        if (prio) __builtin_amdgcn_s_setprio(2);

        // serving with paged planes: the request's pages before its forward
        // pass (admission by the planes' HBM footprint)
        int32_t *pl = nullptr;
        int need_pages = 0;
        bool paged = false, nopages = false;
        if constexpr (SV && TB) {
            __shared__ int32_t page_list[kMaxPages];
            pl = page_list;
            paged = a.page_shift >= 0;
            if (paged && has) {
                need_pages = (((qlen + 7) >> 3) + (1 << a.page_shift) - 1) >> a.page_shift;
                if (!take_pages(a.page_owner, a.n_pages, need_pages, &a.dctl->page_hint, pl,
                                a.alloc_ticks)) {
                    // no pages within the bound: the request skips the forward
                    // pass (best cell at row 0, so the walk reads no plane) and
                    // completes through the normal path with n_cigar = -1; the
                    // host is told
                    if (lane == 0)
                        __hip_atomic_store(a.serve_gave_up, 1, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_SYSTEM);
                    nopages = true;
                    need_pages = 0;
                }
            }
        }
        const int64_t page_words = paged ? ((int64_t)(lstar + 1) * C) << a.page_shift : 0;
        const int32_t page_mask = paged ? (1 << a.page_shift) - 1 : 0;

        int Hp[C], Fp[C];
        uint32_t acc[C];
#pragma unroll
        for (int k = 0; k < C; ++k) { Hp[k] = 0; Fp[k] = kNeg; acc[k] = 0u; }
        // lane's first maximum: key H << 8 | (7 - r) << 5 | (31 - k), row besti
        int bestK = 255, besti = 0;
        // XS: the scalar cell's H and F of the previous row, plane bits and
        // first maximum (wave-uniform)
        int Hs = 0, Fs = kNeg, bestKs = 31, bestis = 0;
        uint32_t accs = 0u;
        (void)Hs; (void)Fs; (void)bestKs; (void)bestis; (void)accs;

        const int nblk = nopages ? 0 : (maxq + 7) >> 3;
        const int nblk_mine = (qlen + 7) >> 3;
        // serving, and offline batches walked in the launch with slab planes
        // (slab_words > 0): each wave slot (pair group of a wave) owns one
        // plane slab, reused pair after pair
        const int64_t tb_base =
            (TB && has) ? ((SV || a.slab_words > 0)
                               ? ((int64_t)blockIdx.x * (64 / G) + grp) * a.slab_words
                               : a.tb_off[p])
                        : 0;
        const int64_t rk_base = (RK && has) ? a.rk_off[p] : 0;

        // target element (0-based in the pair) of cell 0 at the block's first row
        int64_t e_t = to + (int64_t)(1 + off - hb + gl * C - 1);
        const int64_t e_q = qo;  // query element of row 1

        uint32_t wt[NWR], wn[NWR];
        auto load_t = [&](int64_t e, uint32_t (&w)[NWR], const uint32_t *src) __attribute__((always_inline)) {
            const int64_t wb = e >> 4;
#pragma unroll
            for (int x = 0; x < NWR; ++x) w[x] = src[clampi64(wb + x, t_hi)];
        };
        if (maxq > 0) {
            load_t(e_t, wt, a.tseq);
            if (NM) load_t(e_t, wn, a.tmask);
        }

        // Blocks whose 8 rows have every band cell inside the matrix form one
        // contiguous range [fb0, fb1); the edge blocks around it are masked.
        int fb0 = ((hb + 1 - off) - 1 + 7) >> 3;  // first b with 8b+1 >= hb+1-off
        fb0 = max(fb0, 0);
        int fb1 = min(qlen, tlen - hb - off);     // rows <= fb1 are fast-capable
        fb1 = fb1 >= 8 ? (fb1 >> 3) : 0;          // blocks with 8b+8 <= that bound
        if (!has) { fb0 = 0; fb1 = nblk; }
        if constexpr (G < 64) {
            fb0 = wave_max(fb0);
            fb1 = -wave_max(-fb1);
        }
        fb1 = min(max(fb1, fb0), nblk);
        fb0 = min(fb0, fb1);
        for (int b = 0; b < nblk; ++b) {
            const int i0 = 8 * b + 1;
            int rkeys[8];
            int bk = INT_MIN;  // the block's best key (true H)
            (void)rkeys;
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
            // XS: lane 63's window also holds the scalar cell's bases (its
            // cells C .. C+7 over the block's rows)
            uint64_t tws = 0, nws = 0;
            int rks[8];
            (void)tws; (void)nws; (void)rks;
            if constexpr (XS) {
                tws = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)ta[1], G - 1) << 32) |
                      (uint32_t)__builtin_amdgcn_readlane((int)ta[0], G - 1);
                if (NM)
                    nws = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)na[1], G - 1) << 32) |
                          (uint32_t)__builtin_amdgcn_readlane((int)na[0], G - 1);
            }
            // edge blocks: the window's fields whose columns lie outside the
            // matrix (j < 1 or j > tlen), as 2-bit fields of 11 (never match)
            const bool interior = b >= fb0 && b < fb1;
            uint32_t om[NA];
            if constexpr (!RK) {
                if (!interior) {
                    const int c0 = (int)(e_t - to);  // column - 1 of the window's field 0
#pragma unroll
                    for (int x = 0; x < NA; ++x)
                        om[x] = low_fields(-c0 - 16 * x) | ~low_fields(tlen - c0 - 16 * x);
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

            // Row modes.  kRowFast: every band cell of the block is inside the
            // matrix.  kRowEdge: cells outside it (left of column 1, right of
            // column tlen, below row qlen) are computed like the others but
            // never match.  Such a cell then never exceeds a cell before it in
            // (row, diagonal) order that is inside the matrix (a left cell
            // holds H = 0 exactly; a right or lower one takes its value from a
            // cell above or to its left with a non-positive step), so the
            // first maximum, and every inside cell's value and plane bits,
            // are those of the masked form; only the outside cells' plane
            // bits, which no walk reads, differ.  kRowMasked (row keys, which
            // need each row's maximum over the inside cells alone): outside
            // cells hold H = 0 and F = -inf.
            auto row = [&](auto mode_tag, const int r) __attribute__((always_inline)) {
                constexpr int MODE = decltype(mode_tag)::value;
                constexpr bool SLOW = MODE == kRowMasked;
                constexpr bool EDGE = MODE == kRowEdge;
                const int i = i0 + r;
                const int zrow = i * mi;  // H = 0 in this row's domain
                const int zprev = zrow - mi;
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
                    if constexpr (EDGE) {
                        y |= __builtin_amdgcn_alignbit(om[x + 1], om[x], 2 * r);
                        y = (i > qlen) ? ~0u : y;
                    }
                    z[x] = ~y & 0x55555555u;  // 2-bit field k = 1 <=> cell k matches
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
                // lane 63's d+1: the band guard, or the scalar cell (XS)
                int hN = from_next_lane(XS ? Hs : zprev, Hp[0]);
                int fN = from_next_lane(XS ? Fs : kNeg, Fp[0]);
                hN = guard ? zprev : hN;
                fN = guard ? kNeg : fN;

                int Mv[C], Fv[C], Dg[C], Av[C], Gv[C];
#pragma unroll
                for (int k = 0; k < C; ++k) {
                    const int uh = (k + 1 < C) ? Hp[(k + 1) % C] : hN;
                    const int uf = (k + 1 < C) ? Fp[(k + 1) % C] : fN;
                    const int av = uh - cH;
                    const int bv = uf - cF;
                    Fv[k] = max(av, bv);
                    if (TB) acc[k] = insert_sign(acc[k], av - bv);  // f_cont
                    const uint32_t mt = __builtin_amdgcn_ubfe(z[k >> 4], 2 * (k & 15), 2);
                    Dg[k] = mad_u24(mt, mms, Hp[k]);
                    const int m = max(max(Dg[k], Fv[k]), zrow);
                    Mv[k] = m;
                    int A = m + k * ge;
                    if constexpr (SLOW) {
                        const bool live = (uint32_t)(k - lo) < span;
                        A = live ? A : kNeg;
                    }
                    Av[k] = A;
                }
                // lane total max_k A_k as a max3 tree, then the cross-lane scan
                int T = Av[0];
#pragma unroll
                for (int k = 1; k < C; k += 2) T = (k + 1 < C) ? max(max(T, Av[k]), Av[(k + 1) % C]) : max(T, Av[k]);
                // lane 0's prefix is 0, not -inf: every live A = M + k*e >= zrow
                // >= 0 (row-shifted domain), so a zero prefix never wins E (E
                // <= 0 <= zrow <= M), and a masked cell's A stays below it
                int X = from_prev_lane_or0(seg_scan_max<G>(T + lCe));
                if constexpr (G < 64) X = (gl == 0) ? kNeg : X;
                const int Xl = X - lCe;
                // exclusive prefix inside the lane: G_0 = X, G_{k+1} = max(G_k, A_k)
                Gv[0] = Xl;
#pragma unroll
                for (int k = 1; k < C; ++k) Gv[k] = max(Gv[k - 1], Av[k - 1]);
                const int Glast = Gv[C - 1];
                const int dlast = Av[C - 1] - Glast;
                // lane 0's e_cont bit at d = 0 is never read (no band cell
                // left of it)
                const int dprev = from_prev_lane_or0(dlast);
                if constexpr (XS) {
                    // the scalar cell d = W-1 as register C of lane 63: E from
                    // that lane's prefix, no band cell at d+1 (H 0 and F -inf
                    // there, as for a guard lane)
                    const int d63 = __builtin_amdgcn_readlane(dlast, G - 1);
                    const int g63 = __builtin_amdgcn_readlane(Glast, G - 1);
                    const int tf = (int)((uint32_t)(tws >> (2 * (C + r))) & 3u);
                    int mf = NM ? (int)(((uint32_t)qn) | ((uint32_t)(nws >> (2 * (C + r))) & 1u))
                                : 0;
                    // the cell is out of the matrix unless every term is >= 0
                    const int xl = (SLOW || EDGE) ? (int)((uint32_t)(qlen - i) |
                                                          (uint32_t)(tlen + hb - off - i - (W - 1)) |
                                                          (uint32_t)((W - 1) - (hb + 1 - off - i)))
                                                  : 0;
                    if constexpr (EDGE) mf |= (int)((uint32_t)xl >> 31);  // outside: never matches
                    scalar_cell<SLOW>(Hs, Fs, accs, bestKs, bestis, tf, (int)qc, mf, (int)mms, zprev,
                                zrow, cH, kNeg - cF, d63, g63, go + (C - 1) * ge,
                                (int)((uint32_t)(i * mi) << 5), i, xl, kNeg);
                    if constexpr (RK)
                        rks[r] = (int)(((uint32_t)(Hs - i * mi) << 11) |
                                       (uint32_t)(2047 - (W - 1)));
                }

                int rk = INT_MIN;
                int rkh = INT_MIN;  // registers past kstar (compile-time kstar only)
                int Gkm1 = Xl;
#pragma unroll
                for (int k = 0; k < C; ++k) {
                    const int Gk = Gv[k];
                    const int E = Gk - (go + (k - 1) * ge);
                    int H = max(Mv[k], E);
                    int Fk = Fv[k];
                    if constexpr (SLOW) {
                        const bool live = (uint32_t)(k - lo) < span;
                        H = live ? H : zrow;
                        Fk = live ? Fk : kNeg;
                    } else if constexpr (KS >= 0) {
                        // registers past diagonal W-1 (of lane lstar and the
                        // dead lanes) hold values of no cell: only the first
                        // of them, the d+1 input of diagonal W-1, must read
                        // as the band guard; the others feed dead cells alone,
                        // and the best key leaves them out below
                        if (k == kstar + 1) {
                            H = dead_hi ? zrow : H;
                            Fk = dead_hi ? kNeg : Fk;
                        }
                    } else {
                        // runtime kstar: every register past diagonal W-1
                        // reads as out of band
                        if (k > kstar) {
                            H = dead_hi ? zrow : H;
                            Fk = dead_hi ? kNeg : Fk;
                        }
                    }
                    if (TB) {
                        acc[k] = insert_sign(acc[k], Dg[k] - H);  // not_diag
                        acc[k] = insert_sign(acc[k], Fk - H);     // not_f
                        acc[k] = insert_sign(acc[k], k ? (Av[k ? k - 1 : 0] - Gkm1) : dprev);  // e_cont
                    }
                    const int key = (int)(((uint32_t)H << 8) | (uint32_t)(31 - k));
                    if (!SLOW && KS >= 0 && k > kstar)
                        rkh = max(rkh, key);
                    else
                        rk = max(rk, key);
                    Hp[k] = H;  // Hp[k] is not read again in this row
                    Fv[k] = Fk;
                    Gkm1 = Gk;
                }
#pragma unroll
                for (int k = 0; k < C; ++k) Fp[k] = Fv[k];
                if constexpr (!SLOW && KS >= 0 && KS + 1 < C) rk = max(rk, dead_hi ? INT_MIN : rkh);
                // the row's key in true H with the row in bits 5-7 (one scalar:
                // per-cell constants above 64 would each take a register), then
                // the block's running maximum (keys order by H, then by smaller
                // row, then by smaller d)
                bk = max(bk, rk - (int)(((uint32_t)(i * mi) << 8) - ((uint32_t)(7 - r) << 5)));
                if constexpr (RK) {
                    // the row's (max H, first d) as one key (keys order by H, then
                    // by smaller d): lane-local key, then a max over the group;
                    // the group's last lane ends up holding it
                    const int htrue = (rk >> 8) - i * mi;
                    const int d = gl * C + 31 - (rk & 31);
                    // live keys are >= 1 (H >= 0, 2047 - d >= 1), so a dead
                    // lane's key 0 never wins; an AND, not a select: the select
                    // form made the scheduler keep every row live (256 VGPRs +
                    // 256 AGPRs + scratch with traceback planes)
                    rkeys[r] = (int)((((uint32_t)htrue << 11) | (uint32_t)(2047 - d)) & live_mask);
                }
            };
            // Hp[k] is overwritten in pass 2 only after pass 1 consumed it, and
            // Hp[k+1] (read by cell k) is consumed in pass 1 as well.
            if (interior) {
#pragma unroll
                for (int r = 0; r < 8; ++r) row(std::integral_constant<int, kRowFast>{}, r);
            } else if constexpr (RK) {
#pragma unroll
                for (int r = 0; r < 8; ++r) row(std::integral_constant<int, kRowMasked>{}, r);
            } else {
#pragma unroll
                for (int r = 0; r < 8; ++r) row(std::integral_constant<int, kRowEdge>{}, r);
            }
            // strict improvement in H over the lane's best so far
            if (bk > (bestK | 255)) {
                bestK = bk;
                besti = i0 + 7 - ((bk >> 5) & 7);
            }
            if constexpr (RK) {
                // group max of each row's lane key, after the block: eight
                // independent scans off the rows' H dependency chain
#pragma unroll
                for (int r = 0; r < 8; ++r) rkeys[r] = seg_scan_max<G>(rkeys[r]);
                if constexpr (XS) {
#pragma unroll
                    for (int r = 0; r < 8; ++r) rkeys[r] = max(rkeys[r], rks[r]);
                }
                if (has && gl == G - 1) {
#pragma unroll
                    for (int r = 0; r < 8; ++r)
                        if (i0 + r <= qlen) a.rowkey[rk_base + (int64_t)(i0 + r - 1)] = rkeys[r];
                }
            }
            if (TB && has && lane_live && b < nblk_mine) {
                uint32_t *dst = a.tb + tb_base + ((int64_t)b * (lstar + 1) + gl) * C;
                if constexpr (SV) {
                    if (paged)
                        dst = a.tb + (int64_t)page_of(pl, b >> a.page_shift) * page_words +
                              ((int64_t)(b & page_mask) * (lstar + 1) + gl) * C;
                }
#pragma unroll
                for (int k = 0; k < C; ++k) dst[k] = acc[k];
                // XS: the scalar cell's plane dword, lane 64's register 0
                if constexpr (XS)
                    if (gl == G - 1) dst[C] = accs;
            }
        }

        // group reduction of the lane-first maxima: (max H, min i, min d)
        int bH = bestK >> 8;
        const int bd = gl * C + (31 - (bestK & 31));
        if (!lane_live || besti == 0) bH = 0;
        uint64_t key = ((uint64_t)(uint32_t)bH << 40) |
                       ((uint64_t)(0xFFFFFu - (uint32_t)besti) << 12) | (uint64_t)(0xFFFu - (uint32_t)bd);
        if (bH == 0) key = 0;
        if constexpr (XS) {
            // the scalar cell's first maximum (diagonal W-1) joins lane 0's
            const int bHs = bestKs >> 5;
            if (bestis != 0 && bHs != 0 && gl == 0) {
                const uint64_t ks = ((uint64_t)(uint32_t)bHs << 40) |
                                    ((uint64_t)(0xFFFFFu - (uint32_t)bestis) << 12) |
                                    (uint64_t)(0xFFFu - (uint32_t)(W - 1));
                key = ks > key ? ks : key;
            }
        }
#pragma unroll
        for (int o = G / 2; o >= 1; o >>= 1) {
            const uint32_t lo32 = (uint32_t)__shfl_xor((int)(uint32_t)key, o);
            const uint32_t hi32 = (uint32_t)__shfl_xor((int)(uint32_t)(key >> 32), o);
            const uint64_t other = ((uint64_t)hi32 << 32) | lo32;
            key = other > key ? other : key;
        }
        int res_s = 0, res_i = 0, res_j = 0;
        if (key != 0) {
            res_s = (int)(key >> 40);
            res_i = (int)(0xFFFFFu - (uint32_t)((key >> 12) & 0xFFFFFu));
            res_j = res_i + off - hb + (int)(0xFFFu - (uint32_t)(key & 0xFFFu));
        }
        if (gl == 0 && has && !online) {
            a.score[p] = res_s;
            a.qend[p] = res_i;
            a.tend[p] = res_j;
        }
        if constexpr (TB && !RK && SV) {
            if (online) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                if (paged) {
                    const int64_t bs = (int64_t)(lstar + 1) * C;
                    const int sh = a.page_shift;
                    auto pld = [&](int top, int d0) {
                        return plane_win_load_paged(a.tb, pl, sh, bs, top, d0, W);
                    };
                    walk_pair_ring<C, NM>(a.wc, a.ring_cigar + (int64_t)ring_r * a.cigar_cap,
                                          a.cigar_cap, a.ring_res + ring_r, res_i, res_j, res_s,
                                          off, qo, to, pld, claim_tick);
                } else {
                    const uint32_t *tbp = a.tb + tb_base;
                    const int bs = a.wc.Gs * C;
                    auto pld = [&](int top, int d0) { return plane_win_load(tbp, top, d0, bs, W); };
                    walk_pair_ring<C, NM>(a.wc, a.ring_cigar + (int64_t)ring_r * a.cigar_cap,
                                          a.cigar_cap, a.ring_res + ring_r, res_i, res_j, res_s,
                                          off, qo, to, pld, claim_tick);
                }
                if (nopages && lane == 0) host_store(&a.ring_res[ring_r].n_cigar, ~0u);
                // every output of the request (stored by the wave's lanes) before
                // its done word and its completion-queue record: the system
                // release fence waits for the write-through stores above
                __syncthreads();
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
                if (lane == 0) {
                    // the done word needs no release of its own: the record's
                    // release below orders it and the outputs before the record
                    __hip_atomic_store(a.ring_done + ring_r, ticket + 1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_SYSTEM);
                    const unsigned long long pos = atomicAdd(&a.dctl->cq_pos, 1ull);
                    __hip_atomic_store(a.ring_cq + (pos & (unsigned long long)(a.ring_n - 1)),
                                       ((uint64_t)(pos + 1) << 32) | (uint32_t)ring_r,
                                       __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
                }
                // after the release above (which wrote this L2's dirty plane
                // lines back), the pages may go to a wave on another XCD
                if (paged) free_pages(a.page_owner, pl, need_pages);
                if (prio) __builtin_amdgcn_s_setprio(0);
                continue;
            }
        }
        if constexpr (TB && !RK) {
            if (a.walk) {
                // this wave's plane stores must be visible to its own loads
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                const WalkCtx &wc = a.wc;
                const WalkOut &wo = a.wo;
#pragma unroll 1
                for (int g = 0; g < 64 / G; ++g) {
                    const int L0 = g * G;
                    if (!__builtin_amdgcn_readlane((int)has, L0)) continue;
                    const int pg = __builtin_amdgcn_readlane(p, L0);
                    const int offg = __builtin_amdgcn_readlane(off, L0);
                    const int64_t qog = readlane64(qo, L0), tog = readlane64(to, L0);
                    const int64_t tbg = readlane64(tb_base, L0);
                    if (paged) {
                        const int64_t bs = (int64_t)wc.Gs * C;
                        const int sh = a.page_shift;
                        auto pld = [&](int top, int d0) {
                            return plane_win_load_paged(a.tb, pl, sh, bs, top, d0, W);
                        };
                        walk_pair_with<C, NM>(wc, wo, pg, __builtin_amdgcn_readlane(res_i, L0),
                                              __builtin_amdgcn_readlane(res_j, L0),
                                              __builtin_amdgcn_readlane(res_s, L0), offg, qog,
                                              tog, pld, SeqWin{}, SeqWin{});
                    } else {
                        walk_pair<C, NM>(wc, wo, pg, __builtin_amdgcn_readlane(res_i, L0),
                                     __builtin_amdgcn_readlane(res_j, L0),
                                     __builtin_amdgcn_readlane(res_s, L0), offg, qog, tog,
                                     a.tb + tbg);
                    }
                }
            }
        }
        if (prio) __builtin_amdgcn_s_setprio(0);
        if constexpr (SV) {
            if (nopages && lane == 0) a.wo.ncig[p] = -1;
            __syncthreads();  // the walk's stores by every lane, then the flag
            if (has && lane == 0) {
                // every output of this pair (stored by this wave) before the flag
                __hip_atomic_store(a.serve_done + p, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            // after that release (dirty plane lines written back), free the pages
            if (paged) free_pages(a.page_owner, pl, need_pages);
        }
    }
}

// -------------------------------------------------------------------------
// Skewed tile: 2 diagonals per lane, 63 lanes per pair (bands up to 126)
// -------------------------------------------------------------------------
// The latency tile of batches of short-band pairs too small to fill the
// device (one pair per wavefront, about one wavefront per SIMD).  Lane l owns
// diagonals 2l ("even") and 2l+1 ("odd"); cell (i, d) is computed at step
// 2(i-1) + d, so in double step s lane l computes row s - l + 1 (its even
// cell, then its odd cell), lane l-1 is a row ahead and lane l+1 a row
// behind.  Every input of a cell is a register of the lane or one DPP shift
// away:
//   even (i, 2l):    left (i, 2l-1)   lane l-1's odd cell of double step s-1
//                    up   (i-1, 2l+1) own odd cell of s-1
//                    diag (i-1, 2l)   own even cell of s-1
//   odd  (i, 2l+1):  left (i, 2l)     own even cell of s
//                    up   (i-1, 2l+2) lane l+1's even cell of s
//                    diag (i-1, 2l+1) own odd cell of s-1
// E and F follow the recurrences themselves (extension.py:139-176), so a
// double step is a short dependency chain with no prefix scan across the
// lanes: the row-sweep tiles pay a 6-step DPP scan per row, which a lone
// wavefront on its SIMD cannot overlap with anything.  A pair of qlen rows
// takes 8*ceil(qlen/8) + ceil(W/2) - 1 double steps.
//
// Cells outside the matrix (rows < 1 or > qlen, columns < 1 or > tlen) are
// computed like the others with their bases forced to mismatch, as the edge
// blocks of the row-sweep kernel: a cell above row 1 or left of column 1
// then holds H = 0 (its inputs are such cells or the boundary), F <= -o and
// E <= -o, which is what the recurrences read there; a cell below row qlen or
// right of column tlen never exceeds a cell before it in (row, diagonal)
// order.  Cells past the band (d >= W) are forced to H = 0, F = -inf (the
// band guard).  Left of diagonal 0 the DPP shift brings H = 0, E = -inf.
//
// Planes: the layout of a 2-cell tile with ceil(W/2) lanes stored (dword d of
// block b at b * blkstride + d, rows in nibbles from the top), the same
// decision bits as the row-sweep kernel; a lane stores its two dwords of a
// block in the double step that completes the block's last row.  The first
// maximum in (row, diagonal) order: keys H << 4 | (15 - cell) over a window
// of 8 double steps (cells 2k + 0/1 in lane order, which is (row, diagonal)
// order), a strict improvement across windows, then the wave reduction of
// the row-sweep kernel.
template <bool TB, bool NM>
__global__ __launch_bounds__(64) void sw_skew_kernel(SwKArgs a)
{
    const int l = (int)threadIdx.x;
    const int W = a.W, hb = a.hb, ma = a.ma, mi = a.mi, go = a.go, ge = a.ge;
    const int nl = (W + 1) >> 1;  // lanes holding a band diagonal
    const bool eout = 2 * l >= W, oout = 2 * l + 1 >= W;
    const uint32_t mms = (uint32_t)(ma + mi);
    const int64_t t_hi = a.t_words - 1, q_hi = a.q_words - 1;
    // rows of this lane that a window's end leaves past its last complete block
    const int bsh = (8 - (l & 7)) & 7;
    // per-lane forms of the band guard: H capped at 0 and F - e at -inf past
    // the band
    const int hcap_e = eout ? 0 : INT_MAX, hcap_o = oout ? 0 : INT_MAX;
    const int gcap_e = eout ? (1 << 29) : ge, gcap_o = oout ? (1 << 29) : ge;
    const int ebias = (int)(0x80000000u - (uint32_t)ge);
    // loop constants held in VGPRs: the DPP forms of sub and xor read their
    // second operand from a VGPR
    int vgo = go, vzero = 0;
    uint32_t vsign = 0x80000000u;
    asm volatile("" : "+v"(vgo), "+v"(vsign), "+v"(vzero));

    for (bool first = true;; first = false) {
        // the first item is the workgroup index, later ones are claimed after
        // the grid's (a grid that covers every item claims nothing)
        int item = 0;
        if (first) {
            item = (int)blockIdx.x;
        } else if ((int)gridDim.x >= a.n_items) {
            break;
        } else {
            if (l == 0) item = (int)atomicAdd(a.counter, 1u);
            item = __builtin_amdgcn_readfirstlane(item) + (int)gridDim.x;
        }
        if (item >= a.n_items) {
            // the launch's last claim leaves the counter at zero
            if (l == 0 && item == a.n_items + (int)gridDim.x - 1) *a.counter = 0u;
            break;
        }
        // a grid with a wavefront for every pair needs no processing order
        // (one dependent load less before the pair's fields)
        const int p =
            uni<64>((a.order && (int)gridDim.x < a.n_items) ? a.order[item] : item);
        const int qlen = uni<64>(a.qlen[p]), tlen = uni<64>(a.tlen[p]), off = uni<64>(a.boff[p]);
        const int64_t qo = uni64(a.q_off[p]), to = uni64(a.t_off[p]);
        const bool prio = a.prio_rows > 0 && qlen >= a.prio_rows;
        if (prio) __builtin_amdgcn_s_setprio(2);
        const int nblk = (qlen + 7) >> 3;
        const int nwin = (8 * nblk + nl - 1 + 7) >> 3;
        const int64_t tb_base =
            TB ? (a.slab_words > 0 ? (int64_t)blockIdx.x * a.slab_words : a.tb_off[p]) : 0;
        // sequence windows for the fused walk, loaded now (each 64 words
        // ending at the pair's last word, when that covers the pair): the walk
        // then starts without two dependent loads
        const bool walk = TB && a.walk;
        SeqWin qsw, tsw;
        if (walk) {
            const int qtop = (int)(((qo & 15) + qlen - 1) >> 4) + 1;
            const int ttop = (int)(((to & 15) + tlen - 1) >> 4) + 1;
            const int64_t qw0 = qo >> 4, tw0 = to >> 4;
            // lane x holds word top - x of the pair's stream
            if (qlen > 0 && qtop <= 63) {
                const int x = qtop - l;
                qsw.top = qtop;
                qsw.v = (x >= 0 && qw0 + x < a.q_words) ? a.qseq[qw0 + x] : 0u;
            }
            if (tlen > 0 && ttop <= 63) {
                const int x = ttop - l;
                tsw.top = ttop;
                tsw.v = (x >= 0 && tw0 + x < a.t_words) ? a.tseq[tw0 + x] : 0u;
            }
        }

        // double steps s of this lane's cells inside the matrix: row s-l+1 in
        // [1, qlen], even column s+l+1+off-hb in [1, tlen] (odd: one more)
        const int c0 = hb - off - l, c1 = tlen - 1 + hb - off - l;
        const int lo_e = max(l, c0), hi_e = min(qlen - 1 + l, c1);
        const int lo_o = max(l, c0 - 1), hi_o = min(qlen - 1 + l, c1 - 1);

        // 2-bit bases of the window's 8 double steps (16 loaded): query from
        // row s0-l+1, target from the even column s0+l+1+off-hb (the odd cell
        // reads one base further); words before a stream's start are clamped,
        // their cells lie outside the matrix
        int64_t eq = qo - l, et = to + l + off - hb;
        uint32_t q0, q1, t0, t1, qm0 = 0u, qm1 = 0u, tm0 = 0u, tm1 = 0u;
        auto ld2 = [&](const uint32_t *src, int64_t e, int64_t hi, uint32_t &w0, uint32_t &w1)
            __attribute__((always_inline)) {
            const int64_t wb = e >> 4;
            w0 = src[clampi64(wb, hi)];
            w1 = src[clampi64(wb + 1, hi)];
        };
        ld2(a.qseq, eq, q_hi, q0, q1);
        ld2(a.tseq, et, t_hi, t0, t1);
        if (NM) {
            ld2(a.qmask, eq, q_hi, qm0, qm1);
            ld2(a.tmask, et, t_hi, tm0, tm1);
        }

        // Registers between double steps: the lane's cells of the previous
        // row (H for the diagonal terms, the odd cell's F - e, and its E - e
        // with the sign bit flipped: the zero that the DPP shift brings into
        // lane 0 reads as E = INT_MIN left of diagonal 0, as H reads 0).
        int He = 0, Ho = 0, Fgo = kNeg, Ebx = 0;
        uint32_t acc_e = 0u, acc_o = 0u, prev_e = 0u, prev_o = 0u;
        int bestK = 0, bestS = 0;
        for (int w = 0; w < nwin; ++w) {
            const int s0 = 8 * w;
            uint32_t ze, zo;
            {
                const uint32_t shq = (uint32_t)(eq & 15) * 2u, sht = (uint32_t)(et & 15) * 2u;
                const uint32_t qw = __builtin_amdgcn_alignbit(q1, q0, shq);
                const uint32_t tw = __builtin_amdgcn_alignbit(t1, t0, sht);
                uint32_t ye = qw ^ tw, yo = qw ^ (tw >> 2);
                ye |= ye >> 1;
                yo |= yo >> 1;
                if (NM) {
                    const uint32_t qn = __builtin_amdgcn_alignbit(qm1, qm0, shq);
                    const uint32_t tn = __builtin_amdgcn_alignbit(tm1, tm0, sht);
                    ye |= qn | tn;
                    yo |= qn | (tn >> 2);
                }
                // bit 2k: cell k of the window matches and lies inside the matrix
                ze = ~ye & (low_fields(hi_e - s0 + 1) & ~low_fields(lo_e - s0)) & 0x55555555u;
                zo = ~yo & (low_fields(hi_o - s0 + 1) & ~low_fields(lo_o - s0)) & 0x55555555u;
            }
            eq += 8;
            et += 8;
            if (w + 1 < nwin) {
                ld2(a.qseq, eq, q_hi, q0, q1);
                ld2(a.tseq, et, t_hi, t0, t1);
                if (NM) {
                    ld2(a.qmask, eq, q_hi, qm0, qm1);
                    ld2(a.tmask, et, t_hi, tm0, tm1);
                }
            }

            int bk = 0;  // the window's best key, H << 4 | (15 - cell)
            unroll_ic<8>([&](auto kc) __attribute__((always_inline)) {
                constexpr int k = decltype(kc)::value;
                // even cell (i, 2l): left from lane l-1, up and diagonal own
                const int ae = __builtin_amdgcn_update_dpp(0, Ho, kWaveShr1, 0xf, 0xf, true) - vgo;
                const int be = (int)((uint32_t)__builtin_amdgcn_update_dpp(0, Ebx, kWaveShr1, 0xf,
                                                                           0xf, true) ^ vsign);
                const int Ee = max(ae, be);
                const int Hgo = Ho - go;
                const int Fe = max(Hgo, Fgo);
                const int De = mad_u24s(__builtin_amdgcn_ubfe(ze, 2 * k, 1), mms, He - mi);
                // H >= 0, and 0 past the band
                const int He2 = med3i(max(max(De, Ee), Fe), vzero, hcap_e);
                if (TB) {
                    // "continue the gap" bits as open candidate < max: the
                    // maxima are computed anyway, and no operand of the
                    // difference comes through a DPP shift
                    acc_e = insert_sign(acc_e, Hgo - Fe);   // f_cont
                    acc_e = insert_sign(acc_e, De - He2);   // not_diag
                    acc_e = insert_sign(acc_e, Fe - He2);   // not_f
                    acc_e = insert_sign(acc_e, ae - Ee);    // e_cont
                }
                const int ke = lshl4_or<15 - 2 * k>(He2);
                // F - e past the band: -inf
                const int Hge = He2 - go, Ege = Ee - ge, Fge = Fe - gcap_e;
                // odd cell (i, 2l+1): left own even, up from lane l+1's even
                const int Eo = max(Hge, Ege);
                const int fa = __builtin_amdgcn_update_dpp(0, Hge, kWaveShl1, 0xf, 0xf, true);
                const int fb = __builtin_amdgcn_update_dpp(0, Fge, kWaveShl1, 0xf, 0xf, true);
                const int Fo = max(fa, fb);
                const int Do = mad_u24s(__builtin_amdgcn_ubfe(zo, 2 * k, 1), mms, Ho - mi);
                const int Ho2 = med3i(max(max(Do, Eo), Fo), vzero, hcap_o);
                if (TB) {
                    acc_o = insert_sign(acc_o, fa - Fo);    // f_cont
                    acc_o = insert_sign(acc_o, Do - Ho2);   // not_diag
                    acc_o = insert_sign(acc_o, Fo - Ho2);   // not_f
                    acc_o = insert_sign(acc_o, Hge - Eo);   // e_cont
                }
                bk = max(max(bk, ke), lshl4_or<14 - 2 * k>(Ho2));
                He = He2;
                Ho = Ho2;
                Ebx = Eo + ebias;
                Fgo = Fo - gcap_o;
            });
            if constexpr (TB) {
                // this lane's block completed in this window: its rows end at
                // row s0 - l + 8 - bsh (a multiple of 8), the last 16 rows are
                // this window's and the previous window's accumulators
                const int sblk = (s0 - l + 7 - bsh) >> 3;
                if (l < nl && sblk >= 0 && sblk < nblk) {
                    uint32_t *dst = a.tb + tb_base + (int64_t)sblk * a.blkstride + 2 * l;
                    dst[0] = __builtin_amdgcn_alignbit(prev_e, acc_e, 4 * bsh);
                    dst[1] = __builtin_amdgcn_alignbit(prev_o, acc_o, 4 * bsh);
                }
                prev_e = acc_e;
                prev_o = acc_o;
            }
            // strict improvement in H over the lane's best so far
            const bool up = (bk >> 4) > (bestK >> 4);
            bestK = up ? bk : bestK;
            bestS = up ? s0 : bestS;
        }
        int bestH = bestK >> 4, besti = 0, bestd = 0;
        {
            const int c = 15 - (bestK & 15);  // cell 2k + (0 even, 1 odd)
            besti = bestS + (c >> 1) - l + 1;
            bestd = 2 * l + (c & 1);
        }

        // wave reduction of the lane-first maxima: (max H, min i, min d), as
        // two DPP max-scans: the largest H, then, among the lanes holding it,
        // the smallest (i << 7 | d) (d < 127 here)
        const int bh = max(bestH, 0);
        const int hmax = __builtin_amdgcn_readlane(seg_scan_max<64>(bh), 63);
        int res_s = 0, res_i = 0, res_j = 0;
        if (hmax > 0) {
            const uint32_t pos = ((uint32_t)besti << 7) | (uint32_t)bestd;
            const int ck = bh == hmax ? (int)(0x7FFFFFFFu - pos) : INT_MIN;
            const uint32_t pmin =
                0x7FFFFFFFu - (uint32_t)__builtin_amdgcn_readlane(seg_scan_max<64>(ck), 63);
            res_s = hmax;
            res_i = (int)(pmin >> 7);
            res_j = res_i + off - hb + (int)(pmin & 127u);
        }
        if (l == 0) {
            a.score[p] = res_s;
            a.qend[p] = res_i;
            a.tend[p] = res_j;
        }
        if constexpr (TB) {
            if (walk) {
                // this wave's plane stores must be visible to its own loads
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                const uint32_t *tbp = a.tb + tb_base;
                const int bs = a.blkstride;
                auto pld = [&](int top, int d0) { return plane_win_load(tbp, top, d0, bs, W); };
                walk_pair_with<2, NM>(a.wc, a.wo, p, res_i, res_j, res_s, off, qo, to, pld, qsw,
                                      tsw);
            }
        }
        if (prio) __builtin_amdgcn_s_setprio(0);
    }
}

// -------------------------------------------------------------------------
// host side: variant table
// -------------------------------------------------------------------------
using SwKernel = void (*)(SwKArgs);

template <int C, int G, int KS>
SwKernel pick_kernel(uint32_t flags)
{
    const bool tb = flags & GMFA_SW_TRACEBACK, nm = flags & GMFA_SW_NMASK, rk = flags & GMFA_SW_ROWKEYS;
    if (rk && tb && nm) return sw_forward_kernel<C, G, KS, true, true, true>;
    if (rk && tb) return sw_forward_kernel<C, G, KS, true, false, true>;
    if (rk && nm) return sw_forward_kernel<C, G, KS, false, true, true>;
    if (rk) return sw_forward_kernel<C, G, KS, false, false, true>;
    if (tb && nm) return sw_forward_kernel<C, G, KS, true, true, false>;
    if (tb) return sw_forward_kernel<C, G, KS, true, false, false>;
    if (nm) return sw_forward_kernel<C, G, KS, false, true, false>;
    return sw_forward_kernel<C, G, KS, false, false, false>;
}

// serving variant: traceback + fused CIGAR walk, one pair per wavefront
template <int C, int G, int KS>
SwKernel pick_serve_kernel(uint32_t flags)
{
    if constexpr (G == 64) {
        if (flags & GMFA_SW_NMASK) return sw_forward_kernel<C, G, KS, true, true, false, true>;
        return sw_forward_kernel<C, G, KS, true, false, false, true>;
    }
    return nullptr;
}

// KS >= 0: the tile serves only bands with (W-1) % C == KS (compile-time dead
// registers; KS == C-1 is an exact fit with none).  KS == -1: any band.
struct TileEntry {
    int C, G, KS;
    SwKernel (*get)(uint32_t);
    void (*walk)(TbKArgs);
    SwKernel (*serve)(uint32_t);
    void (*walk_wave)(TbKArgs);
};

#define GMFA_TILE(C, G, KS)                                                                   \
    {C, G, KS, &pick_kernel<C, G, KS>, &sw_traceback_kernel<C>, &pick_serve_kernel<C, G, KS>, \
     &sw_traceback_wave_kernel<C>}

// The skewed tile (sw_skew_kernel): offline batches only (no row keys, no
// serving form); its planes are those of a 2-cell tile, walked by the C = 2
// walkers.
template <int = 0>
SwKernel pick_skew_kernel(uint32_t flags)
{
    if (flags & GMFA_SW_ROWKEYS) return nullptr;
    const bool tb = flags & GMFA_SW_TRACEBACK, nm = flags & GMFA_SW_NMASK;
    if (tb && nm) return sw_skew_kernel<true, true>;
    if (tb) return sw_skew_kernel<true, false>;
    if (nm) return sw_skew_kernel<false, true>;
    return sw_skew_kernel<false, false>;
}
template <int = 0>
SwKernel no_serve_kernel(uint32_t) { return nullptr; }

constexpr int kSkewLanes = 63;  // lanes_per_pair of the skewed tile (lane 63 idle)
#define GMFA_SKEW_TILE                                                                       \
    {2, kSkewLanes, -1, &pick_skew_kernel<>, &sw_traceback_kernel<2>, &no_serve_kernel<>, \
     &sw_traceback_wave_kernel<2>}

// Split launches: (slice cells per lane CS, slices K, compile-time register
// of diagonal W-1 in a slice tile) over a regular 64-lane tile (C, KS)
using SplitKernel = void (*)(SwKArgs, uint32_t *);

struct SplitEntry {
    int CS, K, KSS, C, KS;
    SplitKernel k;
    SwKernel serve;  // staged serving class of split requests
};

#define GMFA_SPLIT(CS, K, KSS, C, KS) \
    {CS, K, KSS, C, KS, &sw_split_kernel<CS, K, KSS, C>, &sw_split_serve_kernel<CS, K, KSS, C>}

// The compiled variants, one group per translation unit (n = entries).
const TileEntry *sw_tiles_24(int *n);      // sw_tiles_24.hip: 24 x 64
const TileEntry *sw_tiles_wide(int *n);    // sw_tiles_wide.hip: 16 and 32 x 64
const TileEntry *sw_tiles_mid(int *n);     // sw_tiles_mid.hip: 8, 9 and 12 x 64
const TileEntry *sw_tiles_small(int *n);   // sw_tiles_small.hip: 2 and 4 x 64, 16 and 32 lanes
const SplitEntry *sw_split_entries(int *n);  // sw_split.hip

}  // namespace gmfa
