// Continuous serving of banded SW + CIGAR requests on one device.
//
// The reference aligns one read after another (pipeline.py:355-363 fans
// reads over host threads, each calling ExtensionEngine.extend_chains,
// extension.py:511-641, with its own small batch).  Here a request stream is
// served by persistent forward kernels, one per band class, each on its own
// stream: a wavefront claims the next request index from the class's work
// counter, waits until the host has published it, aligns it, walks its CIGAR
// in the same launch and raises the request's completion flag.  No launch per
// request, no batch boundary: a finished wavefront immediately takes the next
// request (continuous batching), and classes run side by side on the CUs.
//
// Control words live in host-coherent pinned memory: the queue tails and
// "closed" flags are written by this host thread and read by the device with
// system-scope acquire loads; completion flags are written by the device
// with system-scope release stores and polled here.  Every device wait is
// bounded (deadline in wall-clock ticks), and the host closes every queue
// before it returns, so every wavefront exits.
#include "common.h"
#include "sw_internal.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace gmfa {
namespace {

// one 128-byte line per control word
struct alignas(128) CtlWord {
    int32_t v;
    int32_t pad[31];
};

struct ClassCtl {
    CtlWord tail, closed, gave_up;
};

}  // namespace
}  // namespace gmfa

using namespace gmfa;

extern "C" int gmfa_serve_run(const gmfa_serve_class *classes, int32_t n_classes, double timeout_s,
                              gmfa_serve_report *report)
{
    GMFA_REQUIRE(classes && report, "NULL argument");
    GMFA_REQUIRE(n_classes >= 1 && n_classes <= 8, "n_classes %d outside [1, 8]", n_classes);
    GMFA_REQUIRE(timeout_s > 0.0 && timeout_s <= 3600.0, "timeout_s must be in (0, 3600]");
    std::memset(report, 0, sizeof(*report));
    int64_t total = 0;
    for (int c = 0; c < n_classes; ++c) {
        const gmfa_serve_class &k = classes[c];
        const int np = k.batch.n_pairs;
        GMFA_REQUIRE(np >= 0, "class %d: negative n_pairs", c);
        if (np == 0) continue;
        GMFA_REQUIRE(k.arrival_s && k.complete_s, "class %d: NULL arrival_s / complete_s", c);
        for (int i = 0; i < np; ++i) {
            const double t = k.arrival_s[i];
            GMFA_REQUIRE(std::isfinite(t) && t >= 0.0 && (i == 0 || t >= k.arrival_s[i - 1]),
                         "class %d: arrival_s must be finite, >= 0 and nondecreasing (index %d)",
                         c, i);
        }
        total += np;
    }

    int dev = 0, khz = 0;
    GMFA_CHECK_HIP(hipGetDevice(&dev));
    GMFA_CHECK_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    GMFA_REQUIRE(khz > 0, "device reports no wall-clock rate");
    // inputs written by the caller on other streams must be complete
    GMFA_CHECK_HIP(hipDeviceSynchronize());

    // control block: per-class words, then the completion flags
    std::vector<int64_t> flag_off(n_classes + 1, 0);
    for (int c = 0; c < n_classes; ++c) flag_off[c + 1] = flag_off[c] + classes[c].batch.n_pairs;
    const size_t ctl_bytes = sizeof(ClassCtl) * (size_t)n_classes;
    const size_t bytes = ctl_bytes + sizeof(int32_t) * (size_t)std::max<int64_t>(total, 1);
    void *host = nullptr;
    GMFA_CHECK_HIP(hipHostMalloc(&host, bytes, hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(host, 0, bytes);
    void *devp = nullptr;
    if (hipHostGetDevicePointer(&devp, host, 0) != hipSuccess) {
        (void)hipHostFree(host);
        set_error("hipHostGetDevicePointer failed for the serving control block");
        return GMFA_E_RUNTIME;
    }
    ClassCtl *ctl = (ClassCtl *)host;
    ClassCtl *ctl_d = (ClassCtl *)devp;
    int32_t *done = (int32_t *)((char *)host + ctl_bytes);
    int32_t *done_d = (int32_t *)((char *)devp + ctl_bytes);

    std::vector<hipStream_t> streams(n_classes, nullptr);
    std::vector<char> launched(n_classes, 0);
    auto close_all = [&]() {
        for (int c = 0; c < n_classes; ++c) __atomic_store_n(&ctl[c].closed.v, 1, __ATOMIC_RELEASE);
    };
    auto finish = [&](int rc) {
        close_all();
        for (int c = 0; c < n_classes; ++c) {
            if (streams[c]) {
                (void)hipStreamSynchronize(streams[c]);
                (void)hipStreamDestroy(streams[c]);
            }
        }
        for (int c = 0; c < n_classes; ++c)
            if (__atomic_load_n(&ctl[c].gave_up.v, __ATOMIC_ACQUIRE)) report->waves_gave_up = 1;
        (void)hipHostFree(host);
        return rc;
    };

    for (int c = 0; c < n_classes; ++c) {
        const gmfa_serve_class &k = classes[c];
        if (k.batch.n_pairs == 0) continue;
        for (int i = 0; i < k.batch.n_pairs; ++i) k.complete_s[i] = -1.0;
        if (hipStreamCreateWithFlags(&streams[c], hipStreamNonBlocking) != hipSuccess) {
            streams[c] = nullptr;
            set_error("hipStreamCreateWithFlags failed");
            return finish(GMFA_E_RUNTIME);
        }
        ServeLaunch sl;
        sl.tail = &ctl_d[c].tail.v;
        sl.closed = &ctl_d[c].closed.v;
        sl.done = done_d + flag_off[c];
        sl.gave_up = &ctl_d[c].gave_up.v;
        sl.slab_words = k.slab_words;
        sl.deadline_ticks = (uint64_t)((timeout_s + 5.0) * (double)khz * 1000.0);
        sl.n_slots = k.n_slots;
        const int rc = sw_launch(&k.batch, &k.params, &k.out, &sl, streams[c]);
        if (rc != GMFA_OK) return finish(rc);
        launched[c] = 1;
    }

    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    auto now_s = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };
    std::vector<int64_t> next(n_classes, 0), lo(n_classes, 0);
    int64_t remaining = total;
    double last = 0.0;
    while (remaining > 0) {
        const double now = now_s();
        if (now > timeout_s) {
            report->timed_out = 1;
            break;
        }
        for (int c = 0; c < n_classes; ++c) {
            const gmfa_serve_class &k = classes[c];
            int64_t x = next[c];
            while (x < k.batch.n_pairs && k.arrival_s[x] <= now) ++x;
            if (x != next[c]) {
                next[c] = x;
                __atomic_store_n(&ctl[c].tail.v, (int32_t)x, __ATOMIC_RELEASE);
            }
        }
        for (int c = 0; c < n_classes; ++c) {
            const gmfa_serve_class &k = classes[c];
            const int32_t *dn = done + flag_off[c];
            for (int64_t x = lo[c]; x < next[c]; ++x) {
                if (k.complete_s[x] < 0.0 && __atomic_load_n(&dn[x], __ATOMIC_ACQUIRE)) {
                    last = now_s();
                    k.complete_s[x] = last;
                    --remaining;
                }
            }
            while (lo[c] < next[c] && k.complete_s[lo[c]] >= 0.0) ++lo[c];
        }
    }
    report->completed = total - remaining;
    report->wall_s = remaining == 0 ? last : now_s();
    (void)launched;
    const int rc = finish(GMFA_OK);
    GMFA_CHECK_HIP(hipGetLastError());
    if (report->timed_out) {
        set_error("serving run timed out after %.1f s with %lld of %lld requests completed",
                  timeout_s, (long long)report->completed, (long long)total);
        return GMFA_E_RUNTIME;
    }
    return rc;
}
