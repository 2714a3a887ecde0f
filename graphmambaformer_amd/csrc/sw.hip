// Banded affine-gap (Gotoh) Smith-Waterman for MI355X (gfx950), host side:
// variant tables, launch logic and the small kernels around the forward pass
// (work-counter resets, split-launch count and wait, batch x-drop, 2-bit
// packing).  The forward and split kernels and their layout are in
// sw_kernels.h; their instantiations in sw_tiles_*.hip and sw_split.hip.
//
// Reference semantics: graphmambaformer/alignment/extension.py:78-302
// (banded_affine_sw_batch + traceback_banded) of sarakh1999/GraphMambaFormer.
#include "sw_kernels.h"

#include <algorithm>
#include <climits>
#include <mutex>
#include <vector>

namespace gmfa {


// Before a split launch: n_split = the length of the longest prefix (at most
// cap) of the longest-first order whose queries have at least split_rows
// rows.  The plain kernel's work counter starts after it (ctl[0] of the plain
// launch = counter[0]), the split kernel's at 0 with ctl[1] = n_split.  One
// workgroup of 256 threads.
__global__ __launch_bounds__(256) void sw_split_count_kernel(const int32_t *qlen,
                                                             const int32_t *order, int32_t n,
                                                             int32_t split_rows, int32_t cap,
                                                             uint32_t *counter)
{
    __shared__ int first_short;
    if (threadIdx.x == 0) first_short = cap;
    __syncthreads();
    for (int k = threadIdx.x; k < cap && k < n; k += 256)
        if (qlen[order[k]] < split_rows) atomicMin(&first_short, k);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int ns = min(first_short, n);
        counter[0] = (uint32_t)ns;  // the plain kernel's first item
        counter[1] = 0u;            // the split kernel's next item
        counter[2] = (uint32_t)ns;  // split items
        counter[3] = 0u;            // split workgroups resident
    }
}

// Between the count kernel and the plain kernel on the caller's stream: waits
// until all `groups` split workgroups (launched on the side stream) are
// resident, at most deadline_ticks of the constant-rate wall clock, so the
// plain kernel's wavefronts never take the slots the split workgroups need.
// One thread.
__global__ void sw_split_wait_kernel(const uint32_t *started, uint32_t groups,
                                     uint64_t deadline_ticks)
{
    const uint64_t t0 = wall_clock64();
    while (__hip_atomic_load(started, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < groups &&
           (uint64_t)wall_clock64() - t0 < deadline_ticks)
        __builtin_amdgcn_s_sleep(2);
}


__global__ void reset_counter_kernel(uint32_t *c) { *c = 0u; }

__global__ void reset_pages_kernel(int32_t *owner, int32_t n)
{
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += gridDim.x * blockDim.x)
        owner[i] = 0;
}

// zero a serving class's device control block (a kernel, not a memset: no
// memset nodes on a stream a caller might capture)
__global__ void reset_serve_ctl_kernel(ServeDevCtl *d)
{
    unsigned long long *w = reinterpret_cast<unsigned long long *>(d);
    if (threadIdx.x < sizeof(ServeDevCtl) / 8) w[threadIdx.x] = 0ull;
}

// The same x-drop for groups of any size: one 256-thread workgroup per group,
// each thread owning pairs p0 + tid, p0 + tid + 256, ...; per-pair state
// (best H, its row, its diagonal, best so far) in scratch (4 int32 per pair);
// rows advance together with one barrier each, which also decides the stop.
__global__ __launch_bounds__(256) void sw_xdrop_wide_kernel(
    const int32_t *rowkey, const int64_t *rk_off, const int32_t *qlen, const int32_t *boff,
    const int32_t *group_off, int32_t thr, int32_t hb, int32_t *st, int32_t *score,
    int32_t *qend, int32_t *tend)
{
    __shared__ int red[256];
    const int g = blockIdx.x, tid = threadIdx.x;
    const int p0 = group_off[g], p1 = group_off[g + 1];
    int mq = 0;
    for (int p = p0 + tid; p < p1; p += 256) {
        mq = max(mq, qlen[p]);
        st[4 * (int64_t)p + 0] = 0;  // best H
        st[4 * (int64_t)p + 1] = 0;  // its row
        st[4 * (int64_t)p + 2] = 0;  // its diagonal
        st[4 * (int64_t)p + 3] = 0;  // best so far of the rows seen (x-drop reference)
    }
    red[tid] = mq;
    __syncthreads();
    for (int o = 128; o >= 1; o >>= 1) {
        if (tid < o) red[tid] = max(red[tid], red[tid + o]);
        __syncthreads();
    }
    const int maxq = red[0];
    for (int i = 1; i <= maxq; ++i) {
        int live = 0;
        for (int p = p0 + tid; p < p1; p += 256) {
            const int key = i <= qlen[p] ? rowkey[rk_off[p] + i - 1] : 0;
            const int h = key >> 11;
            int32_t *sp = st + 4 * (int64_t)p;
            if (h > sp[0]) { sp[0] = h; sp[1] = i; sp[2] = 2047 - (key & 2047); }
            const int seen = max(sp[3], h);
            sp[3] = seen;
            if (thr < 0 || seen - h <= thr) live = 1;
        }
        if (!__syncthreads_or(live)) break;
    }
    for (int p = p0 + tid; p < p1; p += 256) {
        const int32_t *sp = st + 4 * (int64_t)p;
        score[p] = sp[0];
        qend[p] = sp[0] > 0 ? sp[1] : 0;
        tend[p] = sp[0] > 0 ? sp[1] + boff[p] - hb + sp[2] : 0;
    }
}

// The same x-drop for small groups (at most kXdropRowsPairs pairs, e.g. one
// read's chains): the 64 lanes take 64 consecutive rows, the pairs are a
// loop.  Per pair and tile: one coalesced load of 64 row keys, a wave prefix
// max for each row's best-so-far (carried across tiles), and the rows where
// that minus the row's best exceeds x_drop; the stop row is the first row
// where it holds for every pair (the first set bit of a ballot).  Each
// pair's first strict maximum over the rows up to the stop is a wave max of
// (h << 6 | 63 - lane) keys against the carry.  Per-pair state sits in lane
// k of a register.  The lane-per-pair kernel below walks the rows one at a
// time with a pair per lane, which leaves most lanes idle for such groups.
constexpr int kXdropRowsPairs = 16;

__global__ __launch_bounds__(64) void sw_xdrop_rows_kernel(
    const int32_t *rowkey, const int64_t *rk_off, const int32_t *qlen, const int32_t *boff,
    const int32_t *group_off, int32_t thr, int32_t hb, int32_t *score, int32_t *qend,
    int32_t *tend)
{
    const int g = blockIdx.x;
    const int lane = threadIdx.x;
    const int p0 = group_off[g];
    const int n = min(group_off[g + 1] - p0, kXdropRowsPairs);
    const int ql_l = lane < n ? qlen[p0 + lane] : 0;
    const int64_t base_l = lane < n ? rk_off[p0 + lane] : 0;
    const int maxq = wave_max(ql_l);
    int seen_l = 0, bh_l = 0, bi_l = 0, bd_l = 0;  // lane k: pair k's state
    // the keys of the next 64 rows are loaded while this chunk is scanned
    // (two wavefronts per SIMD cannot hide the load latency otherwise)
    int nx[kXdropRowsPairs];
    auto load_rows = [&](int t0v) {
#pragma unroll
        for (int k = 0; k < kXdropRowsPairs; ++k) {
            nx[k] = 0;
            if (k < n) {
                const int ql = __builtin_amdgcn_readlane(ql_l, k);
                const int64_t base = readlane64(base_l, k);
                const int iv = t0v + lane;
                nx[k] = iv <= ql ? rowkey[base + iv - 1] : 0;  // rows past qlen: H = 0
            }
        }
    };
    load_rows(1);
    for (int t0 = 1; t0 <= maxq; t0 += 64) {
        const int i = t0 + lane;  // this lane's row
        int kk[kXdropRowsPairs];
#pragma unroll
        for (int k = 0; k < kXdropRowsPairs; ++k) kk[k] = nx[k];
        if (t0 + 64 <= maxq) load_rows(t0 + 64);
        bool all = thr >= 0 && i <= maxq;
#pragma unroll
        for (int k = 0; k < kXdropRowsPairs; ++k) {
            if (k < n) {
                const int key = kk[k];
                const int h = key >> 11;
                const int pm = max(__builtin_amdgcn_readlane(seen_l, k), seg_scan_max<64>(h));
                all = all && pm - h > thr;
                const int carry = __builtin_amdgcn_readlane(pm, 63);
                seen_l = lane == k ? carry : seen_l;
            }
        }
        const uint64_t stop = __builtin_amdgcn_ballot_w64(all);
        const int last = stop ? (int)__builtin_ctzll(stop) : 63;  // last row lane counted
#pragma unroll
        for (int k = 0; k < kXdropRowsPairs; ++k) {
            if (k < n) {
                const int v = lane <= last
                                  ? (int)(((uint32_t)(kk[k] >> 11) << 6) | (uint32_t)(63 - lane))
                                  : -1;
                const int best = __builtin_amdgcn_readlane(seg_scan_max<64>(v), 63);
                const int h = best >> 6, at = 63 - (best & 63);
                if (h > __builtin_amdgcn_readlane(bh_l, k)) {
                    const int key = __builtin_amdgcn_readlane(kk[k], at);
                    bh_l = lane == k ? h : bh_l;
                    bi_l = lane == k ? t0 + at : bi_l;
                    bd_l = lane == k ? 2047 - (key & 2047) : bd_l;
                }
            }
        }
        if (stop) break;
    }
    if (lane < n) {
        const int p = p0 + lane;
        score[p] = bh_l;
        qend[p] = bh_l > 0 ? bi_l : 0;
        tend[p] = bh_l > 0 ? bi_l + boff[p] - hb + bd_l : 0;
    }
}

// Batch-wide x-drop (extension.py:186-193) per group of pairs: the reference
// stops a batch after the first row where every pair's best-so-far minus the
// row's best exceeds x_drop; the result is the first strict maximum over the
// rows up to and including that one.  One wavefront per group, lane k = pair
// group_off[g] + k (groups of at most 64 pairs); rows are read 16 at a time.
__global__ __launch_bounds__(64) void sw_xdrop_kernel(const int32_t *rowkey, const int64_t *rk_off,
                                                      const int32_t *qlen, const int32_t *boff,
                                                      const int32_t *group_off, int32_t thr,
                                                      int32_t hb, int32_t *score, int32_t *qend,
                                                      int32_t *tend)
{
    const int g = blockIdx.x;
    const int lane = threadIdx.x;
    const int p0 = group_off[g];
    const int n = min(group_off[g + 1] - p0, 64);
    const bool act = lane < n;
    const int p = p0 + (act ? lane : 0);
    const int ql = act ? qlen[p] : 0;
    const int64_t base = act ? rk_off[p] : 0;
    const int maxq = wave_max(ql);
    const uint64_t need = __builtin_amdgcn_ballot_w64(act);
    int seen = 0, bh = 0, bi = 0, bd = 0;
    bool stop = false;
    for (int i0 = 1; i0 <= maxq && !stop; i0 += 16) {
        int kk[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = i0 + r;
            kk[r] = (act && i <= ql) ? rowkey[base + i - 1] : 0;  // rows past qlen: H = 0
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = i0 + r;
            if (i > maxq) break;
            const int h = kk[r] >> 11;
            if (h > bh) { bh = h; bi = i; bd = 2047 - (kk[r] & 2047); }
            seen = max(seen, h);
            if (thr >= 0 && (__builtin_amdgcn_ballot_w64(act && seen - h > thr) & need) == need) {
                stop = true;
                break;
            }
        }
    }
    if (act) {
        score[p] = bh;
        qend[p] = bh > 0 ? bi : 0;
        tend[p] = bh > 0 ? bi + boff[p] - hb + bd : 0;
    }
}

// -------------------------------------------------------------------------
// int16 codes -> 2-bit stream + N mask; one thread per output word
// -------------------------------------------------------------------------
// A flat byte stream of codes -> 2-bit stream + N mask, one thread per output
// word.  two_bit: codes 0..3 (any other value is rejected); else reference codes
// (1..4 = ACGT, 5 = N -> mask, 0 or above 5 rejected).  flags bit 0: a rejected
// code, bit 1: an N.
__global__ void pack_flat_kernel(const uint8_t *codes, int64_t n, int32_t two_bit, uint32_t *seq,
                                 uint32_t *mask, int32_t *flags)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nw = (n + 15) >> 4;
    if (w >= nw) return;
    uint32_t sv = 0u, mv = 0u;
    int f = 0;
#pragma unroll
    for (int x = 0; x < 16; ++x) {
        const int64_t e = w * 16 + x;
        if (e >= n) {
            mv |= 1u << (2 * x);  // padding never matches
            continue;
        }
        const int c = codes[e];
        if (two_bit) {
            if (c > 3) f |= 1;
            sv |= (uint32_t)(c & 3) << (2 * x);
        } else if (c >= 1 && c <= 4) {
            sv |= (uint32_t)(c - 1) << (2 * x);
        } else {
            mv |= 1u << (2 * x);
            f |= (c == 5) ? 2 : 1;
        }
    }
    seq[w] = sv;
    mask[w] = mv;
    if (f) atomicOr(flags, f);
}

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
// host side: variant tables
// -------------------------------------------------------------------------
// Every compiled tile, in the order find_tile and default_tile scan them: the
// exact fits of the benchmark bands (W=65 as 5x13, W=513 as 9x57 or 8x64 + the
// scalar cell, W=1503 as 24x63 with 15 live registers in the last lane), then
// the general tiles, any band up to C*G.
static const std::vector<TileEntry> &tiles()
{
    static const std::vector<TileEntry> all = [] {
        static const int order[][3] = {
            {5, 16, 4},   {9, 64, 8},   {8, 64, 8},   {24, 64, 14}, {2, 16, -1},  {5, 16, -1},
            {8, 16, -1},  {4, 32, -1},  {8, 32, -1},  {2, 64, -1},  {4, 64, -1},  {8, 64, -1},
            {9, 64, -1},  {12, 64, -1}, {16, 64, -1}, {24, 64, -1}, {32, 64, -1},
            {1, 64, 1},   {2, kSkewLanes, -1},
        };
        std::vector<TileEntry> pool, v;
        for (auto part : {sw_tiles_24, sw_tiles_wide, sw_tiles_mid, sw_tiles_small}) {
            int n = 0;
            const TileEntry *e = part(&n);
            pool.insert(pool.end(), e, e + n);
        }
        for (const auto &o : order)
            for (const auto &t : pool)
                if (t.C == o[0] && t.G == o[1] && t.KS == o[2]) v.push_back(t);
        return v;
    }();
    return all;
}

static const std::vector<SplitEntry> &splits()
{
    static const std::vector<SplitEntry> all = [] {
        int n = 0;
        const SplitEntry *e = sw_split_entries(&n);
        return std::vector<SplitEntry>(e, e + n);
    }();
    return all;
}

// band width a tile holds: C*G, plus the scalar cell of KS == C tiles
static int tile_capacity(const TileEntry &t) { return t.C * t.G + (t.KS == t.C ? 1 : 0); }

static bool tile_serves(const TileEntry &t, int W)
{
    if (t.KS == t.C) return W == t.C * t.G + 1;  // scalar-cell tile: exactly that band
    if (t.C * t.G < W) return false;
    return t.KS < 0 || (W - 1) % t.C == t.KS;
}

// explicit (C, G): the specialised variant for this band if one exists, else
// the general one
static const TileEntry *find_tile(int C, int G, int W)
{
    const TileEntry *gen = nullptr;
    for (const auto &t : tiles()) {
        if (t.C != C || t.G != G) continue;
        if (t.KS >= 0 && tile_serves(t, W)) return &t;
        if (t.KS < 0) gen = &t;
    }
    return gen;
}

// Latency tiles, for batches too small to fill the device and never the
// default: a one-cell lane (1 x 64 + the scalar cell for W = 65) and the
// skewed tile (2 x 63).  A large batch of short pairs runs faster packed
// several pairs to a wavefront.
static bool latency_tile(const TileEntry &t) { return t.C == 1 || t.G == kSkewLanes; }

// default: the tile with the least padding (live cells / lane slots), then
// specialised before general, then more cells per lane, latency tiles aside
static const TileEntry *default_tile(int W)
{
    const TileEntry *best = nullptr;
    double best_eff = -1.0;
    for (const auto &t : tiles()) {
        if (!tile_serves(t, W) || latency_tile(t)) continue;
        const double eff = (double)W / (double)tile_capacity(t) + (t.KS >= 0 ? 1e-3 : 0.0);
        if (eff > best_eff + 1e-9) { best = &t; best_eff = eff; }
    }
    return best;
}

// split workgroups of a launch, one pair each: at most this many of the
// longest pairs run split, longer-than-split_rows pairs beyond them run plain
constexpr int kSplitGroups = 512;  // split pairs (and workgroups) of one launch, at most

// The library's side stream and fork / join events per device (created on
// first use, kept until the process ends): the split kernel runs on it beside
// the plain kernel on the caller's stream.  The mutex keeps one launch's
// record / wait / launch sequence whole when host threads launch at once.
struct SideStream {
    hipStream_t s = nullptr;  // non-blocking: the split kernel under graph capture
    // Eager launches: the split kernel on qa, the plain kernel on qb.  Both are
    // created with an all-CU mask, which gives each a hardware queue of its own
    // (queues are otherwise shared once a process holds more streams than the
    // runtime's queue pool, and two kernels on one queue run one after the
    // other).  Such streams are blocking, so with the null stream as the caller
    // neither may be the caller's stream.
    hipStream_t qa = nullptr, qb = nullptr;
    hipEvent_t fork = nullptr, join = nullptr, join2 = nullptr;
};
constexpr int kMaxSideDevices = 64;
static SideStream g_side[kMaxSideDevices];
static std::mutex g_side_mu;

// Count kernel on s; eager: split kernel on qa, wait + plain kernels on qb;
// under capture: split kernel on the side stream, wait + plain kernels on s;
// joined back into s (a caller capturing s into a HIP graph captures all of it)
static int split_launch(const SplitEntry &e, int groups, const SwKArgs &a, SwKernel plain,
                        int plain_waves, const gmfa_sw_batch *batch, int split_rows, hipStream_t s)
{
    int dev = 0;
    GMFA_CHECK_HIP(hipGetDevice(&dev));
    GMFA_REQUIRE(dev >= 0 && dev < kMaxSideDevices, "device %d beyond the %d the library tracks",
                 dev, kMaxSideDevices);
    std::lock_guard<std::mutex> lk(g_side_mu);
    SideStream &ss = g_side[dev];
    if (!ss.s) {
        int cus = 0;
        GMFA_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        GMFA_REQUIRE(cus > 0 && cus <= 4096, "device reports %d compute units", cus);
        std::vector<uint32_t> mask((cus + 31) / 32, ~0u);
        if (cus % 32) mask.back() = (1u << (cus % 32)) - 1u;
        GMFA_CHECK_HIP(hipExtStreamCreateWithCUMask(&ss.qa, (uint32_t)mask.size(), mask.data()));
        GMFA_CHECK_HIP(hipExtStreamCreateWithCUMask(&ss.qb, (uint32_t)mask.size(), mask.data()));
        GMFA_CHECK_HIP(hipStreamCreateWithFlags(&ss.s, hipStreamNonBlocking));
        GMFA_CHECK_HIP(hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming));
        GMFA_CHECK_HIP(hipEventCreateWithFlags(&ss.join, hipEventDisableTiming));
        GMFA_CHECK_HIP(hipEventCreateWithFlags(&ss.join2, hipEventDisableTiming));
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    GMFA_CHECK_HIP(hipStreamIsCapturing(s, &cap));
    const bool eager = cap == hipStreamCaptureStatusNone;
    hipStream_t sa = eager ? ss.qa : ss.s;  // the split kernel
    hipStream_t sb = eager ? ss.qb : s;     // the wait kernel and the plain kernel
    hipLaunchKernelGGL(sw_split_count_kernel, dim3(1), dim3(256), 0, s, batch->qlen, batch->order,
                       batch->n_pairs, split_rows, kSplitGroups, a.counter);
    GMFA_CHECK_HIP(hipGetLastError());
    GMFA_CHECK_HIP(hipEventRecord(ss.fork, s));
    GMFA_CHECK_HIP(hipStreamWaitEvent(sa, ss.fork, 0));
    if (eager) GMFA_CHECK_HIP(hipStreamWaitEvent(sb, ss.fork, 0));
    hipLaunchKernelGGL(e.k, dim3(groups), dim3(64 * e.K), 0, sa, a, a.counter + 1);
    GMFA_CHECK_HIP(hipGetLastError());
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
        rate_khz <= 0)
        rate_khz = 100000;
    hipLaunchKernelGGL(sw_split_wait_kernel, dim3(1), dim3(1), 0, sb, a.counter + 3,
                       (uint32_t)groups, (uint64_t)rate_khz * 5);  // 5 ms
    GMFA_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(plain, dim3(plain_waves), dim3(64), 0, sb, a);
    GMFA_CHECK_HIP(hipGetLastError());
    GMFA_CHECK_HIP(hipEventRecord(ss.join, sa));
    GMFA_CHECK_HIP(hipStreamWaitEvent(s, ss.join, 0));
    if (eager) {
        GMFA_CHECK_HIP(hipEventRecord(ss.join2, sb));
        GMFA_CHECK_HIP(hipStreamWaitEvent(s, ss.join2, 0));
    }
    return GMFA_OK;
}

static bool split_serves(const SplitEntry &e, int W, const gmfa_sw_tile &t)
{
    if (t.lanes_per_pair != 64 || t.cells_per_lane != e.C) return false;
    if (e.KS >= 0 && (W - 1) % e.C != e.KS) return false;
    if (e.KSS >= 0 && (W - 1) % e.CS != e.KSS) return false;
    // exactly K slices, each with a live lane, inside the tile's plane stride
    if (!((e.K - 1) * 64 * e.CS < W && W <= e.K * 64 * e.CS)) return false;
    return (W + e.CS - 1) / e.CS * e.CS <= t.lanes_stored * t.cells_per_lane;
}

}  // namespace gmfa

using namespace gmfa;

extern "C" int gmfa_sw_split_variants(int32_t half_band, int32_t cells_per_lane,
                                      int32_t *slice_cells, int32_t *slices, int32_t cap)
{
    GMFA_REQUIRE(half_band >= 0 && half_band <= 1023, "half_band %d outside [0, 1023]", half_band);
    gmfa_sw_tile t;
    if (gmfa_sw_resolve_tile(half_band, cells_per_lane, cells_per_lane > 0 ? 64 : 0, &t) != GMFA_OK)
        return 0;
    const int W = 2 * half_band + 1;
    int n = 0;
    for (const auto &e : splits()) {
        if (!split_serves(e, W, t)) continue;
        if (n < cap) {
            if (slice_cells) slice_cells[n] = e.CS;
            if (slices) slices[n] = e.K;
        }
        ++n;
    }
    return n;
}

extern "C" int gmfa_sw_resolve_tile(int32_t half_band, int32_t cells_per_lane,
                                    int32_t lanes_per_pair, gmfa_sw_tile *tile)
{
    GMFA_REQUIRE(tile != nullptr, "tile is NULL");
    GMFA_REQUIRE(half_band >= 0 && half_band <= 1023, "half_band %d outside [0, 1023]", half_band);
    const int W = 2 * half_band + 1;
    const TileEntry *t = nullptr;
    if (cells_per_lane > 0 || lanes_per_pair > 0) {
        t = find_tile(cells_per_lane, lanes_per_pair, W);
        if (!t) {
            set_error("no compiled SW tile with %d cells per lane and %d lanes per pair",
                      cells_per_lane, lanes_per_pair);
            return GMFA_E_UNSUPPORTED;
        }
        if (tile_capacity(*t) < W) {
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
    // each (C, G) shape once: the general tiles, then the shapes compiled only
    // as a specialised variant (which serve their band alone); a specialised
    // variant of a general shape is picked automatically for the band it serves
    int n = 0;
    auto general = [](int C, int G) {
        for (const auto &u : tiles())
            if (u.C == C && u.G == G && u.KS < 0) return true;
        return false;
    };
    for (const auto &t : tiles()) {
        if (t.KS >= 0 && general(t.C, t.G)) continue;
        if (n < cap) {
            if (cells_per_lane) cells_per_lane[n] = t.C;
            if (lanes_per_pair) lanes_per_pair[n] = t.G;
        }
        ++n;
    }
    return n;
}

extern "C" int64_t gmfa_sw_tb_bytes(int32_t qlen, const gmfa_sw_tile *tile)
{
    if (!tile || qlen < 0) return GMFA_E_INVALID;
    const int64_t nblk = ((int64_t)qlen + 7) / 8;
    return nblk * tile->lanes_stored * tile->cells_per_lane * 4;
}

extern "C" int gmfa_pack_flat(const uint8_t *codes, int64_t n, int32_t two_bit, uint32_t *seq,
                              uint32_t *mask, int32_t *flags, void *stream)
{
    GMFA_REQUIRE(n >= 0, "negative length");
    if (n == 0) return GMFA_OK;
    GMFA_REQUIRE(codes && seq && mask && flags, "NULL pointer");
    const int64_t nw = (n + 15) / 16;
    hipLaunchKernelGGL(pack_flat_kernel, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, codes, n, two_bit, seq, mask, flags);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
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

extern "C" int gmfa_sw_resident_waves(const gmfa_sw_params *params, int32_t *waves)
{
    GMFA_REQUIRE(params && waves, "NULL argument");
    gmfa_sw_tile tile;
    const int rc = gmfa_sw_resolve_tile(params->half_band, params->cells_per_lane,
                                        params->lanes_per_pair, &tile);
    if (rc != GMFA_OK) return rc;
    const TileEntry *t = find_tile(tile.cells_per_lane, tile.lanes_per_pair,
                                   2 * params->half_band + 1);
    SwKernel k = t->get(params->flags);
    GMFA_REQUIRE(k != nullptr, "the %dx%d tile has no form with per-row keys", t->C, t->G);
    *waves = resident_waves(k);
    return GMFA_OK;
}

static WalkCtx make_walk_ctx(const gmfa_sw_batch *batch, const gmfa_sw_params *params,
                             const gmfa_sw_tile &tile, bool nm)
{
    WalkCtx c;
    c.qseq = batch->qseq; c.tseq = batch->tseq;
    c.q_words = batch->q_words; c.t_words = batch->t_words;
    c.qmask = nm ? batch->qmask : nullptr; c.tmask = nm ? batch->tmask : nullptr;
    c.nm = nm ? 1 : 0;
    c.hb = params->half_band; c.W = 2 * params->half_band + 1;
    c.Gs = tile.lanes_stored;
    c.ma = params->match; c.mi = params->mismatch; c.go = params->gap_open; c.ge = params->gap_extend;
    return c;
}

namespace gmfa {

int serve_reset_pages(int32_t *owner, int32_t n_pages, hipStream_t s)
{
    if (!owner || n_pages <= 0) return GMFA_OK;
    const int blocks = std::min(1024, (n_pages + 255) / 256);
    hipLaunchKernelGGL(reset_pages_kernel, dim3(blocks), dim3(256), 0, s, owner, n_pages);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
}

int sw_launch(const gmfa_sw_batch *batch, const gmfa_sw_params *params, const gmfa_sw_out *out,
              const ServeLaunch *serve, hipStream_t s, const SlabPlanes *slabs)
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
    const bool online = serve && serve->ring_n > 0;
    // online serving: requests come from the ring; the batch gives the
    // reference (tseq / tmask / t_words) and the query length bound only
    const int n = online ? 1 : batch->n_pairs;
    GMFA_REQUIRE(n >= 0, "negative n_pairs");
    if (n == 0) return GMFA_OK;
    GMFA_REQUIRE(online || (batch->qseq && batch->tseq && batch->q_off && batch->t_off &&
                            batch->qlen && batch->tlen && batch->band_offset),
                 "NULL batch pointer");
    GMFA_REQUIRE(batch->tseq, "NULL target stream");
    GMFA_REQUIRE(batch->qseq && batch->q_words > 0 && batch->t_words > 0, "empty sequence streams");
    GMFA_REQUIRE(batch->max_qlen >= 0 && batch->max_qlen <= (1 << 20),
                 "query length bound %d outside [0, 2^20]", batch->max_qlen);
    GMFA_REQUIRE((int64_t)batch->max_qlen * (P.match + P.mismatch) < (1 << 24),
                 "query length x (match + mismatch) must stay below 2^24 for the int32 DP");
    GMFA_REQUIRE(!(P.flags & GMFA_SW_ROWKEYS) || (int64_t)batch->max_qlen * P.match < (1 << 19),
                 "per-row keys need query length x match below 2^19");
    const bool nm = P.flags & GMFA_SW_NMASK;
    GMFA_REQUIRE(!nm || (batch->qmask && batch->tmask), "GMFA_SW_NMASK needs qmask and tmask");
    GMFA_REQUIRE(online || (out->score && out->query_end && out->target_end &&
                            (out->work_counter || slabs)),
                 "NULL output pointer");
    GMFA_REQUIRE(!(P.flags & GMFA_SW_TRACEBACK) || slabs || (out->tb && (out->tb_off || serve)),
                 "GMFA_SW_TRACEBACK needs tb and tb_off");
    GMFA_REQUIRE(!slabs || (!serve && (P.flags & GMFA_SW_TRACEBACK) && (P.flags & GMFA_SW_CIGAR) &&
                            !(P.flags & GMFA_SW_ROWKEYS)),
                 "slab planes need GMFA_SW_TRACEBACK | GMFA_SW_CIGAR and no row keys");
    GMFA_REQUIRE(!(P.flags & GMFA_SW_ROWKEYS) || (out->rowkey && out->rowkey_off),
                 "GMFA_SW_ROWKEYS needs rowkey and rowkey_off");
    const TileEntry *t = find_tile(tile.cells_per_lane, tile.lanes_per_pair, 2 * P.half_band + 1);
    SwKernel k = serve ? t->serve(P.flags) : t->get(P.flags);
    GMFA_REQUIRE(serve || k != nullptr, "the %dx%d tile has no form with per-row keys", t->C, t->G);
    int block = 64;  // threads per workgroup of the launch
    const SplitEntry *split = nullptr;
    int split_groups = 0;

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
    a.tb = slabs ? slabs->tb : out->tb;
    a.tb_off = slabs ? nullptr : out->tb_off;
    a.rowkey = out->rowkey; a.rk_off = out->rowkey_off;
    a.counter = slabs ? slabs->counter : out->work_counter;
    a.walk = 0;
    a.blkstride = tile.lanes_stored * tile.cells_per_lane;
    a.split_rows = 0;
    a.prio_rows = std::max(0, P.prio_rows);
    a.wc = make_walk_ctx(batch, params, tile, nm);
    a.wo = WalkOut{};
    if (P.flags & GMFA_SW_CIGAR) {
        GMFA_REQUIRE(P.flags & GMFA_SW_TRACEBACK, "GMFA_SW_CIGAR needs GMFA_SW_TRACEBACK");
        GMFA_REQUIRE(!(P.flags & GMFA_SW_ROWKEYS),
                     "GMFA_SW_CIGAR walks the unadjusted best cell; x-drop batches use gmfa_sw_traceback");
        GMFA_REQUIRE(online || (out->cigar_off && out->cigar && out->n_cigar &&
                                out->read_start && out->ref_start && out->counts),
                     "GMFA_SW_CIGAR needs the CIGAR outputs");
        a.walk = 1;
        a.wo.cig_off = out->cigar_off; a.wo.cigar = out->cigar; a.wo.ncig = out->n_cigar;
        a.wo.rstart = out->read_start; a.wo.tstart = out->ref_start; a.wo.counts = out->counts;
    }

    a.serve_tail = nullptr; a.serve_closed = nullptr; a.serve_done = nullptr;
    a.serve_gave_up = nullptr; a.slab_words = 0; a.deadline_ticks = 0;
    a.serve_started = nullptr;
    a.poll_ticks = 0; a.dctl = nullptr;
    a.ring_n = 0; a.stage_words = 0; a.cigar_cap = 0;
    a.ring_tail = nullptr; a.ring_desc = nullptr; a.ring_q = nullptr;
    a.ring_res = nullptr; a.ring_cigar = nullptr; a.ring_done = nullptr; a.stage = nullptr;
    a.ring_cq = nullptr;
    a.page_shift = -1; a.n_pages = 0; a.page_owner = nullptr; a.alloc_ticks = 0;
    int waves = 1;
    if (!serve) {
        waves = std::max(1, std::min(a.n_items, resident_waves(k)));
        if (P.max_waves > 0) waves = std::min(waves, P.max_waves);
        // split launch: the longest pairs over several wavefronts each (needs
        // the fused walk, no N masks or row keys, a 64-lane tile and a
        // longest-first order); bands without a split variant run plainly
        if (!slabs && P.split_rows > 0 && P.flags == (GMFA_SW_TRACEBACK | GMFA_SW_CIGAR) &&
            batch->order != nullptr) {
            for (const auto &e : splits()) {
                if (!split_serves(e, 2 * P.half_band + 1, tile)) continue;
                if (P.split_slices > 0 && e.K != P.split_slices) continue;
                split = &e;
                break;
            }
        }
        if (split) {
            // the plain kernel keeps its full grid: sw_split_wait_kernel holds
            // it back until the split workgroups are resident, its wavefronts
            // beyond the free slots start as split workgroups retire (those
            // without a split pair at once)
            // at most 3/4 of the resident wavefronts in split workgroups, so
            // the plain kernel starts beside them; split pairs beyond the
            // workgroups queue behind the first ones (shorter, longest first)
            split_groups = std::min({kSplitGroups, n,
                                     std::max(1, resident_waves(k) * 3 / 4 / split->K)});
            a.split_rows = P.split_rows;
        }
        if (slabs) {
            // one slab per pair group, sized for the longest query (rounded
            // to 256 bytes); as many wavefronts as the slabs fit
            const int64_t need = (gmfa_sw_tb_bytes(batch->max_qlen, &tile) / 4 + 63) / 64 * 64;
            a.slab_words = std::max<int64_t>(need, 64);
            const int64_t fit = slabs->cap_words / (a.slab_words * ppw);
            GMFA_REQUIRE(fit >= 1,
                         "the traceback planes of %d pair(s) of query length %d need %lld bytes, "
                         "above the %lld bytes of the library's static scratch",
                         ppw, batch->max_qlen, (long long)(a.slab_words * ppw * 4),
                         (long long)(slabs->cap_words * 4));
            waves = (int)std::min<int64_t>(waves, fit);
        }
    } else {
        GMFA_REQUIRE(t->G == 64 && k != nullptr,
                     "serving needs a tile with one pair per wavefront (64 lanes per pair), got %d",
                     t->G);
        GMFA_REQUIRE((P.flags & GMFA_SW_CIGAR) && !(P.flags & GMFA_SW_ROWKEYS),
                     "serving needs GMFA_SW_TRACEBACK | GMFA_SW_CIGAR and no row keys");
        GMFA_REQUIRE(batch->order == nullptr, "serving takes requests in arrival order (order must be NULL)");
        GMFA_REQUIRE(serve->n_slots >= 1, "n_slots must be positive");
        if (serve->page_blocks > 0) {
            // paged planes: pages of page_blocks 8-row blocks from one pool
            const int pb = serve->page_blocks;
            GMFA_REQUIRE((pb & (pb - 1)) == 0, "page_blocks %d is not a power of two", pb);
            const int64_t nblk = (batch->max_qlen + 7) / 8;
            const int64_t pages = (nblk + pb - 1) / pb;
            GMFA_REQUIRE(pages <= kMaxPages,
                         "the longest query needs %lld pages of %d blocks; at most %d (use larger "
                         "pages)", (long long)pages, pb, kMaxPages);
            GMFA_REQUIRE(serve->n_pages >= std::max<int64_t>(pages, 1),
                         "a pool of %d pages cannot hold the longest query (%lld pages)",
                         serve->n_pages, (long long)pages);
            GMFA_REQUIRE(serve->page_owner, "paged planes need page_owner");
            int sh = 0;
            while ((1 << sh) < pb) ++sh;
            a.page_shift = sh;
            a.n_pages = serve->n_pages;
            a.page_owner = serve->page_owner;
            a.alloc_ticks = serve->alloc_ticks;
        } else {
            const int64_t need = gmfa_sw_tb_bytes(batch->max_qlen, &tile) / 4;
            GMFA_REQUIRE(serve->slab_words >= need,
                         "slab of %lld plane words per slot is below the %lld the longest query "
                         "needs", (long long)serve->slab_words, (long long)need);
        }
        a.serve_tail = serve->tail; a.serve_closed = serve->closed; a.serve_done = serve->done;
        a.serve_gave_up = serve->gave_up; a.slab_words = serve->slab_words;
        a.serve_started = serve->started;
        a.deadline_ticks = serve->deadline_ticks;
        a.poll_ticks = serve->poll_ticks;
        GMFA_REQUIRE(serve->dctl && ((uintptr_t)serve->dctl & 63) == 0,
                     "serving needs 64 bytes of 64-byte aligned device control words");
        a.dctl = serve->dctl;
        waves = serve->n_slots;
        if (P.split_rows > 0) {
            // a class of split requests: every request runs as band slices
            // on a workgroup of several wavefronts (n_slots workgroups)
            GMFA_REQUIRE(!online, "online sessions take no split classes (split_rows must be 0)");
            GMFA_REQUIRE(!nm, "split serving classes take no N masks");
            const SplitEntry *se = nullptr;
            for (const auto &e : splits()) {
                if (!split_serves(e, 2 * P.half_band + 1, tile)) continue;
                if (P.split_slices > 0 && e.K != P.split_slices) continue;
                se = &e;
                break;
            }
            GMFA_REQUIRE(se != nullptr,
                         "no compiled split variant for half_band %d on the %dx%d tile with %d "
                         "slices", P.half_band, tile.cells_per_lane, tile.lanes_per_pair,
                         P.split_slices);
            k = se->serve;
            block = 64 * se->K;
        }
        if (online) {
            GMFA_REQUIRE(serve->ring_tail && serve->ring_desc && serve->ring_q &&
                             serve->ring_res && serve->ring_cigar && serve->ring_done &&
                             serve->ring_cq &&
                             serve->stage && serve->closed,
                         "NULL online serving pointer");
            GMFA_REQUIRE(serve->stage_words >= serve_stage_words(batch->max_qlen),
                         "stage of %d words per slot is below the %d the longest query needs",
                         serve->stage_words, serve_stage_words(batch->max_qlen));
            GMFA_REQUIRE(serve->cigar_cap >= 1, "cigar_cap must be positive");
            GMFA_REQUIRE((serve->ring_n & (serve->ring_n - 1)) == 0, "ring must be a power of two");
            GMFA_REQUIRE(batch->qseq == serve->stage &&
                             batch->q_words >= 2 * (int64_t)serve->stage_words * serve->n_slots &&
                             (!nm || batch->qmask == serve->stage + serve->stage_words),
                         "online batch: qseq must be the stage array (qmask one stage slot on)");
            a.ring_n = serve->ring_n; a.stage_words = serve->stage_words;
            a.cigar_cap = serve->cigar_cap; a.ring_tail = serve->ring_tail;
            a.ring_desc = serve->ring_desc; a.ring_q = serve->ring_q; a.ring_res = serve->ring_res;
            a.ring_cigar = serve->ring_cigar; a.ring_done = serve->ring_done;
            a.ring_cq = serve->ring_cq;
            a.stage = serve->stage;
        }
    }
    if (serve && !online) {
        // staged: the work counter is the low word of the control block
        a.counter = reinterpret_cast<uint32_t *>(a.dctl);
        hipLaunchKernelGGL(reset_serve_ctl_kernel, dim3(1), dim3(64), 0, s, a.dctl);
        GMFA_CHECK_HIP(hipGetLastError());
    } else if (split) {
        return split_launch(*split, split_groups, a, k, waves, batch, P.split_rows, s);
    } else if (slabs) {
        // the library's static scratch: cleared before each launch (the
        // kernel also leaves it at zero)
        hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, a.counter);
        GMFA_CHECK_HIP(hipGetLastError());
    }
    // other offline launches: the kernel's last wave leaves the caller's
    // counter at zero, so it is zero here (the caller zeroes it once, before
    // its first launch)
    hipLaunchKernelGGL(k, dim3(waves), dim3(block), 0, s, a);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
}

}  // namespace gmfa

extern "C" int gmfa_sw_forward(const gmfa_sw_batch *batch, const gmfa_sw_params *params,
                               const gmfa_sw_out *out, void *stream)
{
    return sw_launch(batch, params, out, nullptr, (hipStream_t)stream);
}

extern "C" int gmfa_sw_xdrop(const int32_t *rowkey, const int64_t *rowkey_off,
                             const int32_t *qlen, const int32_t *band_offset,
                             const int32_t *group_off, int32_t n_groups, int32_t max_group,
                             int32_t x_drop, int32_t half_band, int32_t *scratch,
                             int32_t *score, int32_t *query_end, int32_t *target_end,
                             void *stream)
{
    GMFA_REQUIRE(n_groups >= 0, "negative n_groups");
    if (n_groups == 0) return GMFA_OK;
    GMFA_REQUIRE(max_group >= 1, "max_group must be positive (got %d)", max_group);
    GMFA_REQUIRE(half_band >= 0 && half_band <= 1023, "half_band %d outside [0, 1023]", half_band);
    GMFA_REQUIRE(rowkey && rowkey_off && qlen && band_offset && group_off && score && query_end &&
                     target_end,
                 "NULL pointer");
    const int thr = x_drop < 0 ? -1 : x_drop;
    if (max_group <= kXdropRowsPairs) {
        hipLaunchKernelGGL(sw_xdrop_rows_kernel, dim3(n_groups), dim3(64), 0,
                           (hipStream_t)stream, rowkey, rowkey_off, qlen, band_offset, group_off,
                           thr, half_band, score, query_end, target_end);
    } else if (max_group <= 64) {
        hipLaunchKernelGGL(sw_xdrop_kernel, dim3(n_groups), dim3(64), 0, (hipStream_t)stream,
                           rowkey, rowkey_off, qlen, band_offset, group_off, thr, half_band, score,
                           query_end, target_end);
    } else {
        GMFA_REQUIRE(scratch, "groups above 64 pairs need scratch (4 int32 per pair)");
        hipLaunchKernelGGL(sw_xdrop_wide_kernel, dim3(n_groups), dim3(256), 0,
                           (hipStream_t)stream, rowkey, rowkey_off, qlen, band_offset, group_off,
                           thr, half_band, scratch, score, query_end, target_end);
    }
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
    a.c = make_walk_ctx(batch, params, tile, nm);
    a.o.cig_off = cigar_off; a.o.cigar = cigar; a.o.ncig = n_cigar;
    a.o.rstart = read_start; a.o.tstart = ref_start; a.o.counts = counts;
    a.q_off = batch->q_off; a.t_off = batch->t_off; a.boff = batch->band_offset;
    a.n_pairs = n;
    a.score = score; a.qend = query_end; a.tend = target_end;
    a.tb = tb; a.tb_off = tb_off;
    const TileEntry *t = find_tile(tile.cells_per_lane, tile.lanes_per_pair, a.c.W);
    // A wavefront per pair (the 64-cell wavefront walk), except for batches
    // of many short pairs: pairs per wavefront of the lane walk are as few
    // as still let every wavefront be resident at once (8 per SIMD), and the
    // lane walk (each lane a serial chain of dependent loads) takes over when
    // that packs at least 8 pairs of at most 1,024 rows per wavefront.
    // Measured, standalone walk: 16,384 long-read pairs 5.97 ms on lanes (2
    // per wavefront) against 1.63 ms on wavefronts; 16,384 150 bp pairs 0.058
    // against 0.030 ms; 65,536 150 bp pairs 0.078 ms on lanes (8 per
    // wavefront) against 0.090 ms.
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    const int64_t slots = (int64_t)8 * 4 * cus;  // 8 wavefronts on each SIMD
    a.per_wave = (int)std::max<int64_t>(1, std::min<int64_t>(64, (n + slots - 1) / slots));
    const bool lanes = a.per_wave >= 8 && batch->max_qlen <= 1024;
    if (!lanes)
        hipLaunchKernelGGL(t->walk_wave, dim3(n), dim3(64), 0, (hipStream_t)stream, a);
    else
        hipLaunchKernelGGL(t->walk, dim3((n + a.per_wave - 1) / a.per_wave), dim3(64), 0,
                           (hipStream_t)stream, a);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
}
