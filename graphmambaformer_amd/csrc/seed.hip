// Stage-1 k-mer seeding on the device: minimizer selection of read k-mers and
// lookup in a sorted k-mer table, for a batch of reads at once.
//
// Reference semantics: graphmambaformer/alignment/seeding.py of
// sarakh1999/GraphMambaFormer --
//   pack_kmers      101-126  k-mer = sum digit[o] * 4^(k-1-o), digits A..T =
//                            0..3; a window with N is invalid and packs to 0
//   hash64          83-98    minimap2's 64-bit integer mix
//   minimizer_mask  172-199  position p is selected iff hash[p] equals the
//                            minimum of some window of w consecutive k-mers
//                            containing p (ties all kept); invalid k-mers take
//                            part with hash64(0)
//   GPUKmerIndex    720-798  selected = valid & minimizer; (start, count) of
//                            each selected k-mer in the CSR table; anchors
//                            expanded in position order, table order within a
//                            position (and cuda_kernels.py:44-74 kmer_lookup)
//
// Layout and work split
// ---------------------
// Reads are one 2-bit stream + N mask (16 bases per word).  A workgroup of
// 256 threads owns a tile of 256 consecutive k-mer positions of one read: it
// stages the bases of the tile plus a (w-1)-position halo on each side in
// LDS, computes packed k-mers and hashes into LDS, the window minima, the
// selection, and looks every selected k-mer up.  The table lookup is a
// binary search restricted to one bucket of a directory over the top
// dir_bits of the key (a few entries instead of ~23 dependent HBM loads).
// Then a per-read scan of the hit counts and a scan over reads give every
// anchor its output slot, and an expand pass writes (read_pos, ref_pos).
#include "common.h"

#include <algorithm>

namespace gmfa {

constexpr int kTile = 256;
constexpr int kMaxWindow = 128;
constexpr int kMaxK = 31;

__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x = ~x + (x << 21);
    x ^= x >> 24;
    x = x + (x << 3) + (x << 8);
    x ^= x >> 14;
    x = x + (x << 2) + (x << 4);
    x ^= x >> 28;
    x = x + (x << 31);
    return x;
}

struct KIndex {
    const int64_t *keys, *offsets, *positions;
    const int32_t *dir;
    int64_t n_keys;
    int32_t k, shift;
};

// lower_bound of q in its directory bucket; (start, count) or (0, 0)
__device__ __forceinline__ void table_lookup(const KIndex &ix, int64_t q, int64_t &start,
                                             int32_t &count)
{
    start = 0;
    count = 0;
    if (q < 0 || (ix.k < 32 && (uint64_t)q >= (1ull << (2 * ix.k)))) return;
    const int64_t b = (int64_t)((uint64_t)q >> ix.shift);
    int64_t lo = ix.dir[b], hi = ix.dir[b + 1];
    const int64_t end = hi;
    while (lo < hi) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        if (ix.keys[mid] < q) lo = mid + 1;
        else hi = mid;
    }
    if (lo < end && ix.keys[lo] == q) {
        start = ix.offsets[lo];
        count = (int32_t)(ix.offsets[lo + 1] - ix.offsets[lo]);
    }
}

__global__ __launch_bounds__(256) void kmer_lookup_kernel(KIndex ix, const int64_t *queries,
                                                          int64_t n, int64_t *out_start,
                                                          int32_t *out_count)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int64_t s;
    int32_t c;
    table_lookup(ix, queries[t], s, c);
    out_start[t] = s;
    out_count[t] = c;
}

struct KReads {
    const uint32_t *seq, *mask;
    int64_t n_words;
    const int64_t *read_off, *pos_off;
    const int32_t *read_len, *tile_read, *tile_start;
    int32_t window;
};

__device__ __forceinline__ int base_code(const KReads &r, int64_t e)
{
    const int64_t wi = e >> 4;
    if (wi < 0 || wi >= r.n_words) return 4;
    const int sh = (int)(e & 15) * 2;
    if (r.mask && ((r.mask[wi] >> sh) & 1u)) return 4;
    return (int)((r.seq[wi] >> sh) & 3u);
}

// Selection + lookup for one tile of one read: pos_start / pos_count per
// k-mer position (count 0 when not selected, not valid or not in the table).
__global__ __launch_bounds__(256) void kmer_select_kernel(KIndex ix, KReads rd,
                                                          int64_t *pos_start, int32_t *pos_count)
{
    __shared__ uint8_t bases[kTile + 2 * (kMaxWindow - 1) + kMaxK];
    __shared__ uint64_t hs[kTile + 2 * (kMaxWindow - 1)];
    __shared__ uint64_t pk[kTile + 2 * (kMaxWindow - 1)];
    __shared__ uint8_t okv[kTile + 2 * (kMaxWindow - 1)];
    __shared__ uint64_t wmin[kTile + kMaxWindow - 1];
    const int tid = threadIdx.x;
    const int r = rd.tile_read[blockIdx.x];
    const int t0 = rd.tile_start[blockIdx.x];
    const int k = ix.k;
    const int L = rd.read_len[r];
    const int m = L - k + 1;  // k-mer positions of the read (> 0 for a tile to exist)
    const int we = max(1, min(rd.window, m));
    const int H = we - 1;
    const int nw = m - we + 1;
    const int64_t roff = rd.read_off[r];
    // bases of positions [t0 - H, t0 + kTile + H + k - 1)
    const int nb = kTile + 2 * H + k - 1;
    for (int x = tid; x < nb; x += blockDim.x) {
        const int j = t0 - H + x;
        bases[x] = (j >= 0 && j < L) ? (uint8_t)base_code(rd, roff + j) : (uint8_t)4;
    }
    __syncthreads();
    // packed k-mers and hashes of positions [t0 - H, t0 + kTile + H)
    const int nh = kTile + 2 * H;
    for (int x = tid; x < nh; x += blockDim.x) {
        const int p = t0 - H + x;
        uint64_t h = ~0ull, v = 0;
        bool ok = false;
        if (p >= 0 && p < m) {
            ok = true;
            for (int o = 0; o < k; ++o) {
                const int c = bases[x + o];
                ok = ok && c < 4;
                v = v * 4u + (uint64_t)(c & 3);
            }
            if (!ok) v = 0;
            h = mix64(v);
        }
        hs[x] = h;
        pk[x] = v;
        okv[x] = ok ? 1 : 0;
    }
    __syncthreads();
    // window minima of windows starting at [t0 - H, t0 + kTile)
    for (int j = tid; j < kTile + H; j += blockDim.x) {
        const int s = t0 - H + j;
        uint64_t mn = ~0ull;
        if (s >= 0 && s < nw) {
            for (int o = 0; o < we; ++o) mn = min(mn, hs[j + o]);
        }
        wmin[j] = mn;
    }
    __syncthreads();
    const int p = t0 + tid;
    if (p >= m) return;
    bool sel = false;
    const uint64_t hp = hs[tid + H];
    for (int o = 0; o < we; ++o) {
        const int s = p - o;
        if (s >= 0 && s < nw && hp == wmin[tid + H - o]) sel = true;
    }
    int64_t st = 0;
    int32_t cnt = 0;
    if (sel && okv[tid + H]) table_lookup(ix, (int64_t)pk[tid + H], st, cnt);
    const int64_t g = rd.pos_off[r] + p;
    pos_start[g] = st;
    pos_count[g] = cnt;
}

// Exclusive scan of one read's hit counts (one workgroup per read); the
// read's total goes to read_total[r].
__global__ __launch_bounds__(256) void kmer_read_scan_kernel(const int64_t *pos_off,
                                                             const int32_t *pos_count,
                                                             int64_t *pos_aoff, int64_t *read_total)
{
    __shared__ int64_t part[256];
    const int r = blockIdx.x;
    const int tid = threadIdx.x;
    const int64_t a = pos_off[r], b = pos_off[r + 1];
    int64_t carry = 0;
    for (int64_t c0 = a; c0 < b; c0 += 256) {
        const int64_t x = c0 + tid;
        const int64_t v = x < b ? pos_count[x] : 0;
        part[tid] = v;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const int64_t y = tid >= o ? part[tid - o] : 0;
            __syncthreads();
            part[tid] += y;
            __syncthreads();
        }
        if (x < b) pos_aoff[x] = carry + part[tid] - v;
        carry += part[255];
        __syncthreads();
    }
    if (tid == 0) read_total[r] = carry;
}

// Exclusive scan over reads (one workgroup): read_anchor_off[0..n_reads].
__global__ __launch_bounds__(256) void kmer_total_scan_kernel(const int64_t *read_total,
                                                              int32_t n_reads,
                                                              int64_t *read_anchor_off)
{
    __shared__ int64_t part[256];
    const int tid = threadIdx.x;
    int64_t carry = 0;
    for (int c0 = 0; c0 < n_reads; c0 += 256) {
        const int x = c0 + tid;
        const int64_t v = x < n_reads ? read_total[x] : 0;
        part[tid] = v;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const int64_t y = tid >= o ? part[tid - o] : 0;
            __syncthreads();
            part[tid] += y;
            __syncthreads();
        }
        if (x < n_reads) read_anchor_off[x] = carry + part[tid] - v;
        carry += part[255];
        __syncthreads();
    }
    if (tid == 0) read_anchor_off[n_reads] = carry;
}

// Anchors of one tile: read_pos / ref_pos at read_anchor_off[r] + pos_aoff.
__global__ __launch_bounds__(256) void kmer_expand_kernel(KIndex ix, KReads rd,
                                                          const int64_t *pos_start,
                                                          const int32_t *pos_count,
                                                          const int64_t *pos_aoff,
                                                          const int64_t *read_anchor_off,
                                                          int64_t *read_pos, int64_t *ref_pos)
{
    const int r = rd.tile_read[blockIdx.x];
    const int p = rd.tile_start[blockIdx.x] + threadIdx.x;
    const int m = rd.read_len[r] - ix.k + 1;
    if (p >= m) return;
    const int64_t g = rd.pos_off[r] + p;
    const int32_t c = pos_count[g];
    if (c == 0) return;
    const int64_t base = read_anchor_off[r] + pos_aoff[g];
    const int64_t s = pos_start[g];
    for (int32_t j = 0; j < c; ++j) {
        read_pos[base + j] = p;
        ref_pos[base + j] = ix.positions[s + j];
    }
}

static int make_index(const gmfa_kmer_index *in, KIndex &ix)
{
    GMFA_REQUIRE(in, "NULL index");
    GMFA_REQUIRE(in->k >= 1 && in->k <= kMaxK, "k %d outside [1, %d]", in->k, kMaxK);
    GMFA_REQUIRE(in->dir_bits >= 0 && in->dir_bits <= 2 * in->k && in->dir_bits <= 28,
                 "dir_bits %d outside [0, min(2k, 28)]", in->dir_bits);
    GMFA_REQUIRE(in->n_keys >= 0 && in->n_keys < (int64_t)1 << 31, "n_keys out of range");
    GMFA_REQUIRE(in->dir && (in->n_keys == 0 || (in->keys && in->offsets && in->positions)),
                 "NULL index array");
    ix.keys = in->keys;
    ix.offsets = in->offsets;
    ix.positions = in->positions;
    ix.dir = in->dir;
    ix.n_keys = in->n_keys;
    ix.k = in->k;
    ix.shift = 2 * in->k - in->dir_bits;
    return GMFA_OK;
}

static int make_reads(const gmfa_kmer_reads *in, int k, KReads &rd)
{
    GMFA_REQUIRE(in, "NULL reads");
    GMFA_REQUIRE(in->window >= 1 && in->window <= kMaxWindow, "window %d outside [1, %d]",
                 in->window, kMaxWindow);
    GMFA_REQUIRE(in->n_reads >= 0 && in->n_tiles >= 0, "negative sizes");
    GMFA_REQUIRE(in->n_reads == 0 ||
                     (in->seq && in->read_off && in->read_len && in->pos_off && in->n_words > 0),
                 "NULL read array");
    GMFA_REQUIRE(in->n_tiles == 0 || (in->tile_read && in->tile_start), "NULL tile array");
    (void)k;
    rd.seq = in->seq;
    rd.mask = in->mask;
    rd.n_words = in->n_words;
    rd.read_off = in->read_off;
    rd.pos_off = in->pos_off;
    rd.read_len = in->read_len;
    rd.tile_read = in->tile_read;
    rd.tile_start = in->tile_start;
    rd.window = in->window;
    return GMFA_OK;
}

}  // namespace gmfa

using namespace gmfa;

extern "C" int gmfa_kmer_lookup(const gmfa_kmer_index *index, const int64_t *queries, int64_t n,
                                int64_t *out_start, int32_t *out_count, void *stream)
{
    KIndex ix;
    int rc = make_index(index, ix);
    if (rc != GMFA_OK) return rc;
    GMFA_REQUIRE(n >= 0, "negative n");
    if (n == 0) return GMFA_OK;
    GMFA_REQUIRE(queries && out_start && out_count, "NULL pointer");
    hipLaunchKernelGGL(kmer_lookup_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, ix, queries, n, out_start, out_count);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
}

extern "C" int gmfa_kmer_seed(const gmfa_kmer_index *index, const gmfa_kmer_reads *reads,
                              int64_t *pos_start, int32_t *pos_count, int64_t *pos_aoff,
                              int64_t *read_total, int64_t *read_anchor_off, void *stream)
{
    KIndex ix;
    KReads rd;
    int rc = make_index(index, ix);
    if (rc != GMFA_OK) return rc;
    rc = make_reads(reads, ix.k, rd);
    if (rc != GMFA_OK) return rc;
    GMFA_REQUIRE(read_total && read_anchor_off, "NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    if (reads->n_tiles > 0) {
        GMFA_REQUIRE(pos_start && pos_count && pos_aoff, "NULL pointer");
        hipLaunchKernelGGL(kmer_select_kernel, dim3(reads->n_tiles), dim3(kTile), 0, s, ix, rd,
                           pos_start, pos_count);
        GMFA_CHECK_HIP(hipGetLastError());
    }
    if (reads->n_reads > 0) {
        hipLaunchKernelGGL(kmer_read_scan_kernel, dim3(reads->n_reads), dim3(256), 0, s,
                           reads->pos_off, pos_count, pos_aoff, read_total);
        GMFA_CHECK_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(kmer_total_scan_kernel, dim3(1), dim3(256), 0, s, read_total,
                       reads->n_reads, read_anchor_off);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
}

extern "C" int gmfa_kmer_emit(const gmfa_kmer_index *index, const gmfa_kmer_reads *reads,
                              const int64_t *pos_start, const int32_t *pos_count,
                              const int64_t *pos_aoff, const int64_t *read_anchor_off,
                              int64_t *read_pos, int64_t *ref_pos, void *stream)
{
    KIndex ix;
    KReads rd;
    int rc = make_index(index, ix);
    if (rc != GMFA_OK) return rc;
    rc = make_reads(reads, ix.k, rd);
    if (rc != GMFA_OK) return rc;
    if (reads->n_tiles == 0) return GMFA_OK;
    GMFA_REQUIRE(pos_start && pos_count && pos_aoff && read_anchor_off && read_pos && ref_pos,
                 "NULL pointer");
    hipLaunchKernelGGL(kmer_expand_kernel, dim3(reads->n_tiles), dim3(kTile), 0,
                       (hipStream_t)stream, ix, rd, pos_start, pos_count, pos_aoff,
                       read_anchor_off, read_pos, ref_pos);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
}
