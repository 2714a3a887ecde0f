// Library plumbing of libgmfa: thread-local error string, version and device
// queries, and the host-side chain backtrack.
#include "common.h"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

namespace gmfa {

static thread_local std::string g_error;

void set_error(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
}

}  // namespace gmfa

using namespace gmfa;

extern "C" const char *gmfa_last_error(void) { return g_error.c_str(); }

extern "C" int gmfa_abi_version(void) { return GMFA_ABI_VERSION; }

extern "C" int gmfa_device_info(int32_t *cu_count, int32_t *warp_size, char *arch_name,
                                int32_t arch_len)
{
    int dev = 0;
    GMFA_CHECK_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    GMFA_CHECK_HIP(hipGetDeviceProperties(&prop, dev));
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (warp_size) *warp_size = prop.warpSize;
    if (arch_name && arch_len > 0) {
        std::strncpy(arch_name, prop.gcnArchName, (size_t)arch_len - 1);
        arch_name[arch_len - 1] = '\0';
    }
    return GMFA_OK;
}

// AffineChainer._backtrack (chaining.py:530-576): visit anchors by descending
// f with ties in index order (np.argsort(-f, kind="stable")), follow parents
// until -1 or an anchor claimed by a better chain, score = f[i] - f[stop].
extern "C" int64_t gmfa_chain_backtrack(int64_t n, const double *f, const int64_t *parent,
                                        int64_t min_anchors, double min_score,
                                        int64_t *members, int64_t *chain_start,
                                        double *chain_score)
{
    GMFA_REQUIRE(n >= 0, "negative anchor count");
    GMFA_REQUIRE(chain_start != nullptr, "NULL chain_start");
    chain_start[0] = 0;
    if (n == 0) return 0;
    GMFA_REQUIRE(f && parent && members && chain_score, "NULL pointer");
    std::vector<int64_t> order((size_t)n);
    std::iota(order.begin(), order.end(), (int64_t)0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t x, int64_t y) { return -f[x] < -f[y]; });
    std::vector<uint8_t> claimed((size_t)n, 0);
    int64_t nc = 0, nm = 0;
    for (int64_t r = 0; r < n; ++r) {
        const int64_t i = order[(size_t)r];
        if (claimed[(size_t)i]) continue;
        int64_t node = i;
        const int64_t start = nm;
        while (node >= 0 && !claimed[(size_t)node]) {
            if (node >= n) {
                set_error("parent index %lld out of range", (long long)node);
                return GMFA_E_INVALID;
            }
            claimed[(size_t)node] = 1;
            members[nm++] = node;
            node = parent[(size_t)node];
        }
        const double s = f[i] - (node >= 0 ? f[node] : 0.0);
        if (nm - start < min_anchors || s < min_score) {
            nm = start;
            continue;
        }
        std::reverse(members + start, members + nm);
        chain_score[nc] = s;
        chain_start[++nc] = nm;
    }
    return nc;
}

// ---------------------------------------------------------------------------
// BAM CIGAR records (graphmambaformer/data/formats.py:52, 343-382 _make_segment,
// with ExtensionEngine's soft clips, extension.py:628-640)
// ---------------------------------------------------------------------------
namespace gmfa {

__device__ __forceinline__ bool query_op(uint32_t op)
{
    return op == 0 || op == GMFA_OP_I || op == GMFA_OP_S || op == GMFA_OP_EQ || op == GMFA_OP_X;
}

// One record per thread: S(read_start) + runs + S(read_len - read_end), zero
// runs dropped and equal neighbours merged; a record is mapped with a
// non-empty CIGAR, a positive reference span and a sequence; reverse-strand
// mapped records are reversed to the forward strand and flagged 0x10; a
// query-consuming length other than the read length unmaps (0x4, no CIGAR).
__global__ void cigar_bam_kernel(int64_t n, const uint32_t *cigar, const int64_t *cig_off,
                                 const int32_t *n_cigar, const int32_t *read_start,
                                 const int32_t *read_end, const int32_t *read_len,
                                 const int64_t *ref_start, const int64_t *ref_end,
                                 const int32_t *strand, uint32_t *out, const int64_t *out_off,
                                 int32_t *out_n, int32_t *flag)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint32_t *in = cigar + cig_off[r];
    uint32_t *o = out + out_off[r];
    int nr = 0;
    auto put = [&](uint32_t op, uint32_t len) {
        if (len == 0) return;
        if (nr > 0 && (o[nr - 1] & 15u) == op) o[nr - 1] += len << 4;
        else o[nr++] = (len << 4) | op;
    };
    put(GMFA_OP_S, (uint32_t)max(read_start[r], 0));
    for (int x = 0; x < n_cigar[r]; ++x) put(in[x] & 15u, in[x] >> 4);
    put(GMFA_OP_S, (uint32_t)max(read_len[r] - read_end[r], 0));
    int64_t qlen = 0;
    for (int x = 0; x < nr; ++x)
        if (query_op(o[x] & 15u)) qlen += o[x] >> 4;
    // the reference sets the reverse flag (and reverses) before the SAM length
    // check, so an unmapped reverse-strand record keeps 0x10
    bool mapped = nr > 0 && ref_end[r] > ref_start[r] && read_len[r] > 0;
    int f = 0;
    if (mapped && strand[r] < 0) {
        f |= 0x10;
        for (int x = 0, y = nr - 1; x < y; ++x, --y) {
            const uint32_t t = o[x];
            o[x] = o[y];
            o[y] = t;
        }
    }
    if (mapped && qlen != (int64_t)read_len[r]) mapped = false;
    if (!mapped) {
        f |= 0x4;
        nr = 0;
    }
    out_n[r] = nr;
    flag[r] = f;
}

}  // namespace gmfa

extern "C" int gmfa_cigar_bam(int64_t n, const uint32_t *cigar, const int64_t *cigar_off,
                              const int32_t *n_cigar, const int32_t *read_start,
                              const int32_t *read_end, const int32_t *read_len,
                              const int64_t *ref_start, const int64_t *ref_end,
                              const int32_t *strand, uint32_t *out, const int64_t *out_off,
                              int32_t *out_n, int32_t *flag, void *stream)
{
    GMFA_REQUIRE(n >= 0, "negative n");
    if (n == 0) return GMFA_OK;
    GMFA_REQUIRE(cigar && cigar_off && n_cigar && read_start && read_end && read_len && ref_start &&
                     ref_end && strand && out && out_off && out_n && flag,
                 "NULL pointer");
    hipLaunchKernelGGL(gmfa::cigar_bam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, n, cigar, cigar_off, n_cigar, read_start, read_end,
                       read_len, ref_start, ref_end, strand, out, out_off, out_n, flag);
    GMFA_CHECK_HIP(hipGetLastError());
    return GMFA_OK;
}
