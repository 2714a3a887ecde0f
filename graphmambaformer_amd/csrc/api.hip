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
