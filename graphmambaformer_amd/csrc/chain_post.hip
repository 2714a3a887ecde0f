// Chaining around the DP on the device: the per-segment stable sort into DP
// order, the claimed-walk backtrack and the per-read primary/secondary
// selection, so a batch of reads goes from anchors to chains without a
// host round trip of f / parent.
//
// Reference semantics (graphmambaformer/alignment of sarakh1999/GraphMambaFormer):
//   types.py:149-152      sorted_by_ref: np.lexsort((read_end, ref_end)), stable
//   chaining.py:530-576   _backtrack: visit anchors by argsort(-f, stable);
//                         follow parent until -1 or a claimed anchor (claiming
//                         as it goes); score = f[i] - f[stop] (0 if none); keep
//                         if >= min_chain_anchors members and >= min_chain_score;
//                         members reported in read order
//   chaining.py:579-617   select: stable sort by -score; stop below
//                         best * secondary_score_ratio (once one is kept); drop a
//                         chain whose read overlap with the primary exceeds
//                         secondary_overlap; first kept is primary; at most
//                         max_chains
//   chaining.py:620-642   chain: strand +1 segment first, then -1
//
// A segment = one (read, strand) anchor set, one workgroup each.  Sorting is
// a bottom-up merge sort by rank: each element finds its output slot in the
// merged run with one binary search in the other run (lower bound from the
// left run, upper bound from the right run), which is stable and needs no
// shared state besides a barrier per pass.  It runs on global scratch (two
// int32 permutation buffers per segment), so any segment size works.
#include "common.h"

namespace gmfa {

// merge sort of positions 0..n-1 by `less`; returns the buffer holding the
// sorted permutation (a or b).  All threads of the workgroup must call it.
template <class Less>
__device__ int32_t *merge_sort_perm(int n, int32_t *a, int32_t *b, Less less)
{
    const int tid = threadIdx.x, bs = blockDim.x;
    for (int i = tid; i < n; i += bs) a[i] = i;
    __syncthreads();
    int32_t *src = a, *dst = b;
    for (int run = 1; run < n; run <<= 1) {
        for (int i = tid; i < n; i += bs) {
            const int start = (i / (2 * run)) * (2 * run);
            const int mid = min(start + run, n), end = min(start + 2 * run, n);
            const int32_t x = src[i];
            int pos;
            if (i < mid) {
                int lo = mid, hi = end;  // right-run elements strictly before x
                while (lo < hi) {
                    const int m = (lo + hi) >> 1;
                    if (less(src[m], x)) lo = m + 1;
                    else hi = m;
                }
                pos = (i - start) + (lo - mid);
            } else {
                int lo = start, hi = mid;  // left-run elements not after x
                while (lo < hi) {
                    const int m = (lo + hi) >> 1;
                    if (!less(x, src[m])) lo = m + 1;
                    else hi = m;
                }
                pos = (i - mid) + (lo - start);
            }
            dst[start + pos] = x;
        }
        __syncthreads();
        int32_t *t = src;
        src = dst;
        dst = t;
    }
    return src;
}

struct SortArgs {
    const int64_t *seg_off;
    const int64_t *re, *fe, *rp, *fp, *id;
    const double *w;
    int32_t *scratch;  // 2n
    int64_t *o_re, *o_fe, *o_rp, *o_fp, *o_id;
    double *o_w;
    int64_t n;
};

// Segments of at most this many anchors sort (and backtrack) with their keys
// and permutations in LDS: the merge sort's binary searches and the claimed
// walk are chains of dependent loads, ~100 cycles each from LDS against ~1 us
// from L2/HBM.  Larger segments (any size) use global scratch.  The bound
// sets the LDS of a workgroup (backtrack: 33 B per anchor, 16.5 KB), so 9
// segments walk per CU at once (4 at a bound of 1,024); Poisson(400) anchor
// sets stay under it.
constexpr int kLdsAnchors = 512;

__global__ __launch_bounds__(256) void chain_sort_kernel(SortArgs a)
{
    __shared__ int64_t k_fe[kLdsAnchors], k_re[kLdsAnchors];
    __shared__ int32_t p_a[kLdsAnchors], p_b[kLdsAnchors];
    const int s = blockIdx.x;
    const int64_t base = a.seg_off[s];
    const int n = (int)(a.seg_off[s + 1] - base);
    if (n <= 0) return;
    const int64_t *re = a.re + base, *fe = a.fe + base;
    const int32_t *perm;
    if (n <= kLdsAnchors) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            k_fe[i] = fe[i];
            k_re[i] = re[i];
        }
        __syncthreads();
        perm = merge_sort_perm(n, p_a, p_b, [&](int32_t x, int32_t y) {
            return k_fe[x] < k_fe[y] || (k_fe[x] == k_fe[y] && k_re[x] < k_re[y]);
        });
    } else {
        perm = merge_sort_perm(n, a.scratch + base, a.scratch + a.n + base,
                               [&](int32_t x, int32_t y) {
                                   return fe[x] < fe[y] || (fe[x] == fe[y] && re[x] < re[y]);
                               });
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int64_t src = base + perm[i], dst = base + i;
        a.o_re[dst] = a.re[src];
        a.o_fe[dst] = a.fe[src];
        a.o_rp[dst] = a.rp[src];
        a.o_fp[dst] = a.fp[src];
        a.o_id[dst] = a.id[src];
        a.o_w[dst] = a.w[src];
    }
}

struct BtArgs {
    const int64_t *seg_off;
    const double *f;
    const int64_t *parent;  // segment-local, -1 = none
    const int64_t *rp, *re, *fp, *fe, *id;
    int64_t min_anchors;
    double min_score;
    int32_t *scratch;  // 2n
    uint8_t *claimed;  // n
    int32_t *n_chains;      // [n_seg]
    double *c_score;        // chain records at the segment's base (capacity n)
    int64_t *c_span;        // 4 per record: read_start, read_end, ref_start, ref_end
    int64_t *c_mem;         // 2 per record: member offset (absolute), member count
    int64_t *members;       // anchor ids, n
    int32_t *bad_parent;    // set when a parent index leaves its segment
    int64_t n;
};

// Kept chain c of a segment: spans over its members and the members as anchor
// ids in read order (the walk collected them from the chain's end back),
// by one wavefront.
__device__ __forceinline__ void chain_record(const BtArgs &a, int64_t base, int64_t rec,
                                             const int32_t *walk, int start, int cnt)
{
    const int lane = (int)(threadIdx.x & 63);
    int64_t rs = INT64_MAX, re_ = INT64_MIN, fs = INT64_MAX, fe_ = INT64_MIN;
    for (int t = lane; t < cnt; t += 64) {
        const int64_t l = base + walk[start + t];
        rs = min(rs, a.rp[l]);
        re_ = max(re_, a.re[l]);
        fs = min(fs, a.fp[l]);
        fe_ = max(fe_, a.fe[l]);
        a.members[base + start + (cnt - 1 - t)] = a.id[l];
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        rs = min(rs, (int64_t)__shfl_xor((long long)rs, o));
        re_ = max(re_, (int64_t)__shfl_xor((long long)re_, o));
        fs = min(fs, (int64_t)__shfl_xor((long long)fs, o));
        fe_ = max(fe_, (int64_t)__shfl_xor((long long)fe_, o));
    }
    if (lane == 0) {
        a.c_span[4 * rec + 0] = rs;
        a.c_span[4 * rec + 1] = re_;
        a.c_span[4 * rec + 2] = fs;
        a.c_span[4 * rec + 3] = fe_;
        a.c_mem[2 * rec + 0] = base + start;
        a.c_mem[2 * rec + 1] = cnt;
    }
}

// One workgroup per segment.  The stable argsort of -f, then one thread walks
// the claimed-walk (chaining.py:530-576): anchors by descending f; each
// unclaimed one follows its parents, claiming, until -1 or a claimed anchor;
// the chain (score f[i] - f[stop], members) is kept if it has enough members
// and score.  Segments of up to kLdsAnchors anchors keep f, parent, the
// order, the claim bits and the walked members in LDS and record the kept
// chains afterwards, a wavefront per chain; larger ones walk in global
// memory.
__global__ __launch_bounds__(256) void chain_backtrack_kernel(BtArgs a)
{
    __shared__ double l_f[kLdsAnchors];
    // l_pc[i]: parent(i) + 1 in bits 0-30, bit 31 set once i is claimed
    __shared__ int32_t l_pc[kLdsAnchors], p_a[kLdsAnchors], p_b[kLdsAnchors];
    __shared__ int32_t l_walk[kLdsAnchors], l_start[kLdsAnchors], l_cnt[kLdsAnchors];
    __shared__ int l_nc;
    const int s = blockIdx.x;
    const int64_t base = a.seg_off[s];
    const int n = (int)(a.seg_off[s + 1] - base);
    if (n <= 0) {
        if (threadIdx.x == 0) a.n_chains[s] = 0;
        return;
    }
    const double *f = a.f + base;
    const int64_t *parent = a.parent + base;
    if (n <= kLdsAnchors) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            l_f[i] = f[i];
            int64_t pi = parent[i];
            if (pi >= n) {
                atomicOr(a.bad_parent, 1);
                pi = -1;
            }
            l_pc[i] = (int32_t)(pi < 0 ? 0 : pi + 1);
        }
        __syncthreads();
        const int32_t *order =
            merge_sort_perm(n, p_a, p_b, [&](int32_t x, int32_t y) { return l_f[x] > l_f[y]; });
        if (threadIdx.x < 64) {
            // wavefront 0: the lanes look at 64 anchors of the order at once
            // for the next unclaimed one (claims only change inside a walk,
            // and the next look starts right after the anchor walked); the
            // walk itself is wave-uniform, one dependent LDS load per member
            const int lane = (int)threadIdx.x;
            int nc = 0, nm = 0, r = 0;
            while (r < n) {
                const int rr = r + lane;
                const int oi = rr < n ? order[rr] : 0;
                const uint64_t m = __builtin_amdgcn_ballot_w64(rr < n && l_pc[oi] >= 0);
                if (m == 0ull) {
                    r += 64;
                    continue;
                }
                const int k = (int)__builtin_ctzll(m);
                r += k;
                const int i = __builtin_amdgcn_readlane(oi, k);
                int node = i;
                const int start = nm;
                int w = l_pc[node];
                while (w >= 0) {  // node unclaimed
                    if (lane == 0) {
                        l_pc[node] = (int32_t)((uint32_t)w | 0x80000000u);
                        l_walk[nm] = node;
                    }
                    ++nm;
                    node = w - 1;
                    if (node < 0) break;
                    w = l_pc[node];
                }
                const double sc = l_f[i] - (node >= 0 ? l_f[node] : 0.0);
                ++r;
                if (nm - start < a.min_anchors || sc < a.min_score) {
                    nm = start;
                    continue;
                }
                if (lane == 0) {
                    a.c_score[base + nc] = sc;
                    l_start[nc] = start;
                    l_cnt[nc] = nm - start;
                }
                ++nc;
            }
            if (lane == 0) {
                l_nc = nc;
                a.n_chains[s] = nc;
            }
        }
        __syncthreads();
        const int nc = l_nc;
        for (int c = (int)(threadIdx.x >> 6); c < nc; c += (int)(blockDim.x >> 6))
            chain_record(a, base, base + c, l_walk, l_start[c], l_cnt[c]);
        return;
    }
    uint8_t *claimed = a.claimed + base;
    for (int i = threadIdx.x; i < n; i += blockDim.x) claimed[i] = 0;
    int32_t *order = merge_sort_perm(n, a.scratch + base, a.scratch + a.n + base,
                                     [&](int32_t x, int32_t y) { return f[x] > f[y]; });
    if (threadIdx.x != 0) return;
    int64_t *mem = a.members + base;
    int nc = 0, nm = 0;
    for (int r = 0; r < n; ++r) {
        const int i = order[r];
        if (claimed[i]) continue;
        int64_t node = i;
        const int start = nm;
        while (node >= 0 && !claimed[node]) {
            claimed[node] = 1;
            mem[nm++] = node;
            node = parent[node];
            if (node >= n) {
                atomicOr(a.bad_parent, 1);
                node = -1;
            }
        }
        const double sc = f[i] - (node >= 0 ? f[node] : 0.0);
        if (nm - start < a.min_anchors || sc < a.min_score) {
            nm = start;
            continue;
        }
        int64_t rs = INT64_MAX, re_ = INT64_MIN, fs = INT64_MAX, fe_ = INT64_MIN;
        for (int x = start, y = nm - 1; x <= y; ++x, --y) {  // reverse into read order
            const int64_t t = mem[x];
            mem[x] = mem[y];
            mem[y] = t;
        }
        for (int x = start; x < nm; ++x) {
            const int64_t l = base + mem[x];
            rs = min(rs, a.rp[l]);
            re_ = max(re_, a.re[l]);
            fs = min(fs, a.fp[l]);
            fe_ = max(fe_, a.fe[l]);
            mem[x] = a.id[l];
        }
        const int64_t rec = base + nc;
        a.c_score[rec] = sc;
        a.c_span[4 * rec + 0] = rs;
        a.c_span[4 * rec + 1] = re_;
        a.c_span[4 * rec + 2] = fs;
        a.c_span[4 * rec + 3] = fe_;
        a.c_mem[2 * rec + 0] = base + start;
        a.c_mem[2 * rec + 1] = nm - start;
        ++nc;
    }
    a.n_chains[s] = nc;
}

struct SelArgs {
    int32_t n_reads;
    const int32_t *read_seg;  // 2 per read: segment of strand +1, of strand -1 (-1: none)
    const int64_t *seg_off;
    const int32_t *n_chains;
    const double *c_score;
    const int64_t *c_span;
    double ratio, overlap;
    int32_t max_chains;
    int64_t *scratch;         // per read: capacity = its anchors
    const int64_t *read_base; // [n_reads] scratch offset of each read
    int32_t *n_kept;          // [n_reads]
    int64_t *kept;            // [n_reads * max_chains] chain record indices, primary first
};

__global__ __launch_bounds__(64) void chain_select_kernel(SelArgs a)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.n_reads) return;
    int64_t *cand = a.scratch + a.read_base[r];
    int nc = 0;
    for (int h = 0; h < 2; ++h) {  // strand +1 chains first, then -1 (chaining.py:632-641)
        const int s = a.read_seg[2 * r + h];
        if (s < 0) continue;
        const int64_t base = a.seg_off[s];
        for (int k = 0; k < a.n_chains[s]; ++k) cand[nc++] = base + k;
    }
    // stable sort by descending score (insertion sort keeps equal scores in order)
    for (int x = 1; x < nc; ++x) {
        const int64_t c = cand[x];
        const double sc = a.c_score[c];
        int y = x - 1;
        while (y >= 0 && a.c_score[cand[y]] < sc) {
            cand[y + 1] = cand[y];
            --y;
        }
        cand[y + 1] = c;
    }
    int kept = 0;
    int64_t *out = a.kept + (int64_t)r * a.max_chains;
    if (nc > 0) {
        const double cutoff = a.c_score[cand[0]] * a.ratio;
        for (int x = 0; x < nc; ++x) {
            const int64_t c = cand[x];
            if (a.c_score[c] < cutoff && kept) break;
            if (kept) {
                // overlap with the primary (the first kept chain) of the read spans
                const int64_t p = out[0];
                const int64_t lo = max(a.c_span[4 * c], a.c_span[4 * p]);
                const int64_t hi = min(a.c_span[4 * c + 1], a.c_span[4 * p + 1]);
                double ov = 0.0;
                if (hi > lo) {
                    const int64_t sa = a.c_span[4 * c + 1] - a.c_span[4 * c];
                    const int64_t sb = a.c_span[4 * p + 1] - a.c_span[4 * p];
                    ov = (double)(hi - lo) / (double)max(min(sa, sb), (int64_t)1);
                }
                if (ov > a.overlap) continue;
            }
            out[kept++] = c;
            if (kept >= a.max_chains) break;
        }
    }
    a.n_kept[r] = kept;
}

}  // namespace gmfa

using namespace gmfa;

extern "C" int gmfa_chain_sort(const int64_t *seg_off, int32_t n_segments, int64_t n,
                               const int64_t *read_end, const int64_t *ref_end,
                               const double *weight, const int64_t *read_pos,
                               const int64_t *ref_pos, const int64_t *anchor_id, int32_t *scratch,
                               int64_t *o_read_end, int64_t *o_ref_end, double *o_weight,
                               int64_t *o_read_pos, int64_t *o_ref_pos, int64_t *o_anchor_id,
                               void *stream)
{
    GMFA_REQUIRE(n_segments >= 0 && n >= 0 && n < ((int64_t)1 << 31), "sizes out of range");
    if (n_segments == 0 || n == 0) return GMFA_OK;
    GMFA_REQUIRE(seg_off && read_end && ref_end && weight && read_pos && ref_pos && anchor_id &&
                     scratch && o_read_end && o_ref_end && o_weight && o_read_pos && o_ref_pos &&
                     o_anchor_id,
                 "NULL pointer");
    SortArgs a{seg_off, read_end, ref_end, read_pos, ref_pos, anchor_id, weight, scratch,
               o_read_end, o_ref_end, o_read_pos, o_ref_pos, o_anchor_id, o_weight, n};
    hipLaunchKernelGGL(chain_sort_kernel, dim3(n_segments), dim3(256), 0, (hipStream_t)stream, a);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
}

extern "C" int gmfa_chain_backtrack_device(const int64_t *seg_off, int32_t n_segments, int64_t n,
                                           const double *f, const int64_t *parent,
                                           const int64_t *read_pos, const int64_t *read_end,
                                           const int64_t *ref_pos, const int64_t *ref_end,
                                           const int64_t *anchor_id, int64_t min_anchors,
                                           double min_score, int32_t *scratch, uint8_t *claimed,
                                           int32_t *n_chains, double *chain_score,
                                           int64_t *chain_span, int64_t *chain_members,
                                           int64_t *members, int32_t *bad_parent, void *stream)
{
    GMFA_REQUIRE(n_segments >= 0 && n >= 0 && n < ((int64_t)1 << 31), "sizes out of range");
    if (n_segments == 0) return GMFA_OK;
    GMFA_REQUIRE(seg_off && n_chains && bad_parent, "NULL pointer");
    GMFA_REQUIRE(n == 0 || (f && parent && read_pos && read_end && ref_pos && ref_end &&
                            anchor_id && scratch && claimed && chain_score && chain_span &&
                            chain_members && members),
                 "NULL pointer");
    BtArgs a{seg_off, f, parent, read_pos, read_end, ref_pos, ref_end, anchor_id, min_anchors,
             min_score, scratch, claimed, n_chains, chain_score, chain_span, chain_members,
             members, bad_parent, n};
    hipLaunchKernelGGL(chain_backtrack_kernel, dim3(n_segments), dim3(256), 0,
                       (hipStream_t)stream, a);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
}

extern "C" int gmfa_chain_select(int32_t n_reads, const int32_t *read_seg, const int64_t *seg_off,
                                 const int32_t *n_chains, const double *chain_score,
                                 const int64_t *chain_span, double score_ratio,
                                 double max_overlap, int32_t max_chains, int64_t *scratch,
                                 const int64_t *read_base, int32_t *n_kept, int64_t *kept,
                                 void *stream)
{
    GMFA_REQUIRE(n_reads >= 0 && max_chains >= 1, "n_reads >= 0 and max_chains >= 1 required");
    if (n_reads == 0) return GMFA_OK;
    GMFA_REQUIRE(read_seg && seg_off && n_chains && chain_score && chain_span && scratch &&
                     read_base && n_kept && kept,
                 "NULL pointer");
    SelArgs a{n_reads, read_seg, seg_off, n_chains, chain_score, chain_span, score_ratio,
              max_overlap, max_chains, scratch, read_base, n_kept, kept};
    hipLaunchKernelGGL(chain_select_kernel, dim3((n_reads + 63) / 64), dim3(64), 0,
                       (hipStream_t)stream, a);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
}
