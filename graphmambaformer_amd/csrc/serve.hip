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
// "closed" flags are written by this host thread; on the device, idle
// wavefronts poll device-memory mirrors of them, which one wavefront at a time
// (at most once per microsecond) refreshes with relaxed system-scope loads.
// Completion flags (and, for online sessions, completion-queue records) are
// written by the device with system-scope release stores and polled here.
// Every device wait is bounded (deadline in wall-clock ticks), and the host
// closes every queue before it returns, so every wavefront exits.
#include "common.h"
#include "sw_internal.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace gmfa {
namespace {

// one 128-byte line per control word
struct alignas(128) CtlWord {
    int32_t v;
    int32_t pad[31];
};

struct alignas(128) CtlWord64 {
    int64_t v;
    int64_t pad[15];
};

struct ClassCtl {
    CtlWord64 tail;
    CtlWord closed, gave_up, started;
};

// least interval between two reads of a class's host words by its wavefronts
constexpr double kPollSeconds = 1e-6;

// The class kernels are persistent and must run side by side.  HIP maps
// streams onto a per-process pool of hardware queues (GPU_MAX_HW_QUEUES, 4
// by default) and dispatches in order within a queue, so a class kernel whose
// stream shares a queue with another class's running kernel waits until that
// one exits.  Wait (bounded) until every class kernel has signalled that it
// runs; report the queue sharing instead of serving at a fraction of the
// device.
constexpr double kStartSeconds = 5.0;

int wait_started(const std::vector<int32_t *> &flags)
{
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        int missing = -1;
        for (size_t c = 0; c < flags.size(); ++c)
            if (flags[c] && !__atomic_load_n(flags[c], __ATOMIC_ACQUIRE)) missing = (int)c;
        if (missing < 0) return GMFA_OK;
        if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >
            kStartSeconds) {
            set_error("serving class %d's kernel did not start within %.0f s while the other "
                      "classes run: its stream shares a HIP hardware queue with another class's "
                      "persistent kernel (raise GPU_MAX_HW_QUEUES above the number of streams "
                      "the process uses)",
                      missing, kStartSeconds);
            return GMFA_E_RUNTIME;
        }
        std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

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
        GMFA_REQUIRE(std::isfinite(k.page_wait_s) && k.page_wait_s >= 0.0 &&
                         k.page_wait_s <= 3600.0,
                     "class %d: page_wait_s must be in [0, 3600]", c);
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
        ServeLaunch sl{};
        sl.tail = &ctl_d[c].tail.v;
        sl.closed = &ctl_d[c].closed.v;
        sl.done = done_d + flag_off[c];
        sl.gave_up = &ctl_d[c].gave_up.v;
        sl.started = &ctl_d[c].started.v;
        sl.slab_words = k.slab_words;
        sl.page_blocks = k.page_blocks;
        sl.n_pages = k.n_pages;
        sl.page_owner = k.page_owner;
        sl.alloc_ticks = (uint64_t)((k.page_wait_s > 0.0 ? k.page_wait_s : timeout_s + 5.0) *
                                    (double)khz * 1000.0);
        if (k.page_blocks > 0) {
            const int rc = serve_reset_pages(k.page_owner, k.n_pages, streams[c]);
            if (rc != GMFA_OK) return finish(rc);
        }
        sl.deadline_ticks = (uint64_t)((timeout_s + 5.0) * (double)khz * 1000.0);
        sl.poll_ticks = (uint64_t)(kPollSeconds * (double)khz * 1000.0) + 1;
        sl.dctl = reinterpret_cast<ServeDevCtl *>(k.out.work_counter);
        sl.n_slots = k.n_slots;
        const int rc = sw_launch(&k.batch, &k.params, &k.out, &sl, streams[c]);
        if (rc != GMFA_OK) return finish(rc);
        launched[c] = 1;
    }
    {
        std::vector<int32_t *> flags(n_classes, nullptr);
        for (int c = 0; c < n_classes; ++c)
            if (launched[c]) flags[c] = &ctl[c].started.v;
        const int rc = wait_started(flags);
        if (rc != GMFA_OK) return finish(rc);
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
                __atomic_store_n(&ctl[c].tail.v, x, __ATOMIC_RELEASE);
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

// ---------------------------------------------------------------------------
// Online sessions (gmfa_serve_open / submit / poll / close)
// ---------------------------------------------------------------------------
namespace gmfa {
namespace {

struct OnlineCtl {
    CtlWord64 tail;
    CtlWord closed, gave_up, started;
};

struct OnlineClass {
    gmfa_serve_class_cfg cfg{};
    int32_t stage_words = 0;
    bool nm = false;
    hipStream_t stream = nullptr;
    void *host = nullptr;  // one host-coherent block per class
    OnlineCtl *ctl = nullptr;
    ServeDesc *desc = nullptr;
    uint32_t *q = nullptr;
    ServeRes *res = nullptr;
    uint32_t *cig = nullptr;
    int64_t *done = nullptr;
    uint64_t *cq = nullptr;
    int64_t cq_head = 0;             // next completion-queue position to read
    ServeLaunch sl{};
    gmfa_sw_batch batch{};
    gmfa_sw_out out{};  // only tb: the outputs go to the ring
    int64_t next = 0;                // next ticket
    int64_t lo = 0;                  // oldest ticket not yet returned
    std::vector<int64_t> occupant;   // ticket held by each ring entry, -1 when free
    std::chrono::steady_clock::time_point last_query{};
};

template <class T>
T *dev_ptr(void *host_base, void *dev_base, T *p)
{
    return (T *)((char *)dev_base + ((char *)p - (char *)host_base));
}

}  // namespace
}  // namespace gmfa

struct gmfa_serve_session {
    std::vector<gmfa::OnlineClass> cls;
    const uint32_t *ref_seq = nullptr, *ref_mask = nullptr;
    int64_t ref_bases = 0;
    uint64_t idle_ticks = 0;
    int device = 0;
    int64_t submitted = 0, returned = 0, completed = 0, launches = 0;
    double ticks_per_s = 0.0;
};

namespace gmfa {
namespace {

__global__ void zero_ctl_kernel(ServeDevCtl *d)
{
    unsigned long long *w = reinterpret_cast<unsigned long long *>(d);
    if (threadIdx.x < sizeof(ServeDevCtl) / 8) w[threadIdx.x] = 0ull;
}

int launch_class(gmfa_serve_session *s, OnlineClass &c)
{
    const int rc = sw_launch(&c.batch, &c.cfg.params, &c.out, &c.sl, c.stream);
    if (rc == GMFA_OK) ++s->launches;
    return rc;
}

// Relaunch a class whose wavefronts have all exited (idle) while requests of
// it are still pending.  The stream query is rate-limited (it takes the
// runtime's stream lock) unless forced: a class that was empty until now
// (its waves may have idled out) or a closing session.
int ensure_running(gmfa_serve_session *s, OnlineClass &c, bool force)
{
    if (c.lo == c.next) return GMFA_OK;
    const auto now = std::chrono::steady_clock::now();
    if (!force && now - c.last_query < std::chrono::microseconds(200)) return GMFA_OK;
    c.last_query = now;
    const hipError_t q = hipStreamQuery(c.stream);
    if (q == hipErrorNotReady) return GMFA_OK;
    if (q != hipSuccess) {
        set_error("serving class kernel failed: %s", hipGetErrorString(q));
        return GMFA_E_RUNTIME;
    }
    const int32_t ring = c.cfg.ring;
    bool pending = false;
    for (int64_t t = c.lo; t < c.next && !pending; ++t) {
        const int r = (int)(t % ring);
        if (c.occupant[r] == t && __atomic_load_n(&c.done[r], __ATOMIC_ACQUIRE) != t + 1)
            pending = true;
    }
    return pending ? launch_class(s, c) : GMFA_OK;
}

void free_session(gmfa_serve_session *s)
{
    for (auto &c : s->cls) {
        if (c.ctl) __atomic_store_n(&c.ctl->closed.v, 1, __ATOMIC_RELEASE);
    }
    for (auto &c : s->cls) {
        if (c.stream) {
            (void)hipStreamSynchronize(c.stream);
            (void)hipStreamDestroy(c.stream);
            c.stream = nullptr;
        }
        if (c.host) {
            (void)hipHostFree(c.host);
            c.host = nullptr;
        }
    }
    delete s;
}

}  // namespace
}  // namespace gmfa

extern "C" int32_t gmfa_serve_stage_words(int32_t max_qlen)
{
    return max_qlen < 0 ? 0 : serve_stage_words(max_qlen);
}

extern "C" int gmfa_serve_open(const uint32_t *ref_seq, const uint32_t *ref_mask,
                               int64_t ref_bases, const gmfa_serve_class_cfg *classes,
                               int32_t n_classes, double idle_timeout_s,
                               gmfa_serve_session **session)
{
    GMFA_REQUIRE(session, "NULL session pointer");
    *session = nullptr;
    GMFA_REQUIRE(ref_seq && classes, "NULL argument");
    GMFA_REQUIRE(ref_bases >= 1, "empty reference");
    GMFA_REQUIRE(n_classes >= 1 && n_classes <= 8, "n_classes %d outside [1, 8]", n_classes);
    GMFA_REQUIRE(idle_timeout_s > 0.0 && idle_timeout_s <= 3600.0,
                 "idle_timeout_s must be in (0, 3600]");
    for (int c = 0; c < n_classes; ++c) {
        const gmfa_serve_class_cfg &k = classes[c];
        GMFA_REQUIRE(k.max_qlen >= 0 && k.max_qlen <= (1 << 20),
                     "class %d: max_qlen %d outside [0, 2^20]", c, k.max_qlen);
        GMFA_REQUIRE(k.ring >= 1 && k.ring <= (1 << 20) && (k.ring & (k.ring - 1)) == 0,
                     "class %d: ring %d is not a power of two in [1, 2^20]", c, k.ring);
        GMFA_REQUIRE(k.n_slots >= 1, "class %d: n_slots must be positive", c);
        GMFA_REQUIRE(k.cigar_cap >= 1 && k.cigar_cap <= (1 << 22),
                     "class %d: cigar_cap %d outside [1, 2^22]", c, k.cigar_cap);
        GMFA_REQUIRE(k.tb && k.stage && k.head, "class %d: NULL tb / stage / head", c);
        GMFA_REQUIRE(std::isfinite(k.page_wait_s) && k.page_wait_s >= 0.0 &&
                         k.page_wait_s <= 3600.0,
                     "class %d: page_wait_s must be in [0, 3600]", c);
        GMFA_REQUIRE(!(k.params.flags & GMFA_SW_NMASK) || ref_mask,
                     "class %d: GMFA_SW_NMASK needs ref_mask", c);
        GMFA_REQUIRE(k.params.split_rows == 0,
                     "class %d: online sessions take no split classes (split_rows must be 0; "
                     "gmfa_serve_run serves them)", c);
    }
    int dev = 0, khz = 0;
    GMFA_CHECK_HIP(hipGetDevice(&dev));
    GMFA_CHECK_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    GMFA_REQUIRE(khz > 0, "device reports no wall-clock rate");
    // the caller's uploads (reference, zeroed buffers) must be complete
    GMFA_CHECK_HIP(hipDeviceSynchronize());

    gmfa_serve_session *s = new gmfa_serve_session();
    s->ref_seq = ref_seq;
    s->ref_mask = ref_mask;
    s->ref_bases = ref_bases;
    s->idle_ticks = (uint64_t)(idle_timeout_s * (double)khz * 1000.0);
    s->ticks_per_s = (double)khz * 1000.0;
    s->device = dev;
    s->cls.resize(n_classes);
    for (int ci = 0; ci < n_classes; ++ci) {
        OnlineClass &c = s->cls[ci];
        const gmfa_serve_class_cfg &k = classes[ci];
        c.cfg = k;
        c.nm = (k.params.flags & GMFA_SW_NMASK) != 0;
        c.stage_words = serve_stage_words(k.max_qlen);
        c.occupant.assign(k.ring, -1);
        const size_t ring = (size_t)k.ring;
        const size_t b_ctl = sizeof(OnlineCtl);
        const size_t b_desc = sizeof(ServeDesc) * ring;
        const size_t b_q = sizeof(uint32_t) * 2 * (size_t)c.stage_words * ring;
        const size_t b_res = sizeof(ServeRes) * ring;
        const size_t b_cig = (sizeof(uint32_t) * (size_t)k.cigar_cap * ring + 127) & ~(size_t)127;
        const size_t b_done = sizeof(int64_t) * ring;
        const size_t b_cq = sizeof(uint64_t) * ring;
        const size_t bytes = b_ctl + b_desc + b_q + b_res + b_cig + b_done + b_cq;
        void *host = nullptr, *devp = nullptr;
        if (hipHostMalloc(&host, bytes, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
            gmfa::free_session(s);
            set_error("hipHostMalloc of %zu bytes for the serving rings failed", bytes);
            return GMFA_E_RUNTIME;
        }
        c.host = host;
        std::memset(host, 0, bytes);
        if (hipHostGetDevicePointer(&devp, host, 0) != hipSuccess) {
            gmfa::free_session(s);
            set_error("hipHostGetDevicePointer failed for the serving rings");
            return GMFA_E_RUNTIME;
        }
        char *h = (char *)host;
        c.ctl = (OnlineCtl *)h;
        c.desc = (ServeDesc *)(h + b_ctl);
        c.q = (uint32_t *)(h + b_ctl + b_desc);
        c.res = (ServeRes *)(h + b_ctl + b_desc + b_q);
        c.cig = (uint32_t *)(h + b_ctl + b_desc + b_q + b_res);
        c.done = (int64_t *)(h + b_ctl + b_desc + b_q + b_res + b_cig);
        c.cq = (uint64_t *)(h + b_ctl + b_desc + b_q + b_res + b_cig + b_done);

        c.batch = gmfa_sw_batch{};
        c.batch.max_qlen = k.max_qlen;
        c.batch.qseq = k.stage;
        c.batch.qmask = c.nm ? k.stage + c.stage_words : nullptr;
        c.batch.q_words = 2 * (int64_t)c.stage_words * k.n_slots;
        c.batch.tseq = ref_seq;
        c.batch.tmask = c.nm ? ref_mask : nullptr;
        c.batch.t_words = (ref_bases + 15) / 16;

        c.out = gmfa_sw_out{};
        c.out.tb = k.tb;

        ServeLaunch &sl = c.sl;
        sl.tail = nullptr;
        sl.closed = &dev_ptr(host, devp, c.ctl)->closed.v;
        sl.done = nullptr;
        sl.gave_up = &dev_ptr(host, devp, c.ctl)->gave_up.v;
        sl.started = &dev_ptr(host, devp, c.ctl)->started.v;
        sl.slab_words = k.slab_words;
        sl.deadline_ticks = s->idle_ticks;
        sl.poll_ticks = (uint64_t)(kPollSeconds * (double)khz * 1000.0) + 1;
        sl.dctl = reinterpret_cast<ServeDevCtl *>(k.head);
        sl.n_slots = k.n_slots;
        sl.ring_n = k.ring;
        sl.ring_tail = &dev_ptr(host, devp, c.ctl)->tail.v;
        sl.ring_desc = dev_ptr(host, devp, c.desc);
        sl.ring_q = dev_ptr(host, devp, c.q);
        sl.ring_res = dev_ptr(host, devp, c.res);
        sl.ring_cigar = dev_ptr(host, devp, c.cig);
        sl.ring_done = dev_ptr(host, devp, c.done);
        sl.ring_cq = dev_ptr(host, devp, c.cq);
        sl.stage = k.stage;
        sl.stage_words = c.stage_words;
        sl.cigar_cap = k.cigar_cap;
        sl.page_blocks = k.page_blocks;
        sl.n_pages = k.n_pages;
        sl.page_owner = k.page_owner;
        // a wait for pages is bounded by page_wait_s, by default the longer of
        // the idle bound and 60 s
        sl.alloc_ticks = (uint64_t)((k.page_wait_s > 0.0 ? k.page_wait_s
                                                         : std::max(idle_timeout_s, 60.0)) *
                                    (double)khz * 1000.0);

        if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) {
            c.stream = nullptr;
            gmfa::free_session(s);
            set_error("hipStreamCreateWithFlags failed");
            return GMFA_E_RUNTIME;
        }
        // the head starts at ticket 0, the mirrors at 0 (a kernel, not a
        // memset: no memset nodes on a stream a caller might capture)
        hipLaunchKernelGGL(zero_ctl_kernel, dim3(1), dim3(64), 0, c.stream, sl.dctl);
        if (k.page_blocks > 0 && serve_reset_pages(k.page_owner, k.n_pages, c.stream) != GMFA_OK) {
            gmfa::free_session(s);
            set_error("page-owner reset launch failed");
            return GMFA_E_RUNTIME;
        }
        if (hipGetLastError() != hipSuccess) {
            gmfa::free_session(s);
            set_error("head reset launch failed");
            return GMFA_E_RUNTIME;
        }
        const int rc = gmfa::launch_class(s, c);
        if (rc != GMFA_OK) {
            const std::string msg = gmfa_last_error();
            gmfa::free_session(s);
            set_error("%s", msg.c_str());
            return rc;
        }
    }
    std::vector<int32_t *> flags;
    for (auto &c : s->cls) flags.push_back(&c.ctl->started.v);
    const int src = wait_started(flags);
    if (src != GMFA_OK) {
        const std::string msg = gmfa_last_error();
        gmfa::free_session(s);
        set_error("%s", msg.c_str());
        return src;
    }
    *session = s;
    return GMFA_OK;
}

namespace gmfa {
namespace {

// 16 codes 0..3 (one per byte) -> one 2-bit word, base k at bits 2k..2k+1,
// eight bytes at a time; bad != 0 if a byte is above 3.
inline uint32_t pack16_2bit(const uint8_t *p, uint32_t &bad)
{
    uint64_t w[2];
    std::memcpy(w, p, 16);
    uint32_t out = 0u;
    uint64_t hi = 0u;
    for (int h = 0; h < 2; ++h) {
        uint64_t x = w[h];
        hi |= x & 0xFCFCFCFCFCFCFCFCull;
        x = (x | (x >> 6)) & 0x000F000F000F000Full;   // byte pairs -> 4-bit nibbles
        x = (x | (x >> 12)) & 0x000000FF000000FFull;  // nibble pairs -> bytes
        x = (x | (x >> 24)) & 0xFFFFull;              // two bytes -> 16 bits
        out |= (uint32_t)x << (16 * h);
    }
    bad = hi != 0u;
    return out;
}

// Write one request into the next ring entry of its class (not yet published:
// the caller advances the class's tail afterwards).
int stage_request(gmfa_serve_session *s, int32_t cls, const uint8_t *codes, int32_t qlen,
                  int32_t two_bit, int64_t t_off, int32_t tlen, int32_t band_offset,
                  int64_t *ticket)
{
    GMFA_REQUIRE(cls >= 0 && cls < (int)s->cls.size(), "class %d outside [0, %d)", cls,
                 (int)s->cls.size());
    OnlineClass &c = s->cls[cls];
    GMFA_REQUIRE(qlen >= 0 && qlen <= c.cfg.max_qlen, "qlen %d outside [0, %d]", qlen,
                 c.cfg.max_qlen);
    GMFA_REQUIRE(qlen == 0 || codes, "NULL codes");
    GMFA_REQUIRE(tlen >= 0 && t_off >= 0 && t_off + (int64_t)tlen <= s->ref_bases,
                 "window [%lld, %lld) leaves the reference of %lld bases", (long long)t_off,
                 (long long)(t_off + tlen), (long long)s->ref_bases);
    GMFA_REQUIRE(band_offset >= -(1 << 30) && band_offset <= (1 << 30),
                 "band_offset %d outside [-2^30, 2^30]", band_offset);
    const int64_t t = c.next;
    const int r = (int)(t & (c.cfg.ring - 1));
    if (c.occupant[r] >= 0) {
        set_error("class %d: ring of %d requests is full (poll, then submit again)", cls,
                  c.cfg.ring);
        return GMFA_E_AGAIN;
    }
    // pack into the entry: 2-bit words, then N-mask words
    const int nw = (qlen + 15) / 16;
    uint32_t *qw = c.q + (size_t)r * 2 * c.stage_words;
    uint32_t *mw = qw + c.stage_words;
    for (int w = 0; w < nw; ++w) {
        uint32_t v = 0u, m = 0u, bad = 0u;
        const int b0 = 16 * w, b1 = std::min(qlen, b0 + 16);
        if (two_bit && b1 - b0 == 16) {
            v = pack16_2bit(codes + b0, bad);
            GMFA_REQUIRE(bad == 0u, "a code outside 0..3 in bases %d..%d", b0, b1 - 1);
        } else if (two_bit) {
            for (int b = b0; b < b1; ++b) {
                const uint32_t x = codes[b];
                bad |= x & ~3u;
                v |= (x & 3u) << (2 * (b - b0));
            }
            GMFA_REQUIRE(bad == 0u, "a code outside 0..3 in bases %d..%d", b0, b1 - 1);
        } else {
            for (int b = b0; b < b1; ++b) {
                uint32_t x = codes[b];
                GMFA_REQUIRE(x >= 1u && x <= 5u, "code %u at %d outside 1..5", x, b);
                uint32_t n = 0u;
                if (x == 5u) {
                    GMFA_REQUIRE(c.nm, "an N (code 5 at %d) needs a GMFA_SW_NMASK class", b);
                    x = 1u;
                    n = 1u;
                }
                v |= (x - 1u) << (2 * (b - b0));
                m |= n << (2 * (b - b0));
            }
        }
        qw[w] = v;
        if (c.nm) mw[w] = m;
    }
    ServeDesc &d = c.desc[r];
    d.t_off = t_off;
    d.qlen = qlen;
    d.tlen = tlen;
    d.band_offset = band_offset;
    d.q_words = nw;
    c.occupant[r] = t;
    c.next = t + 1;
    ++s->submitted;
    *ticket = t;
    return GMFA_OK;
}

int publish(gmfa_serve_session *s, OnlineClass &c, bool was_empty)
{
    __atomic_store_n(&c.ctl->tail.v, c.next, __ATOMIC_RELEASE);
    return ensure_running(s, c, was_empty);
}

}  // namespace
}  // namespace gmfa

extern "C" int gmfa_serve_submit(gmfa_serve_session *s, int32_t cls, const uint8_t *codes,
                                 int32_t qlen, int32_t two_bit, int64_t t_off, int32_t tlen,
                                 int32_t band_offset, int64_t *ticket)
{
    GMFA_REQUIRE(s && ticket, "NULL argument");
    GMFA_REQUIRE(cls >= 0 && cls < (int)s->cls.size(), "class %d outside [0, %d)", cls,
                 (int)s->cls.size());
    OnlineClass &c = s->cls[cls];
    const bool was_empty = c.lo == c.next;
    const int rc = stage_request(s, cls, codes, qlen, two_bit, t_off, tlen, band_offset, ticket);
    if (rc != GMFA_OK) return rc;
    return publish(s, c, was_empty);
}

extern "C" int gmfa_serve_submit_many(gmfa_serve_session *s, int32_t n, const int32_t *cls,
                                      const uint8_t *codes, const int64_t *code_off,
                                      const int32_t *qlen, int32_t two_bit, const int64_t *t_off,
                                      const int32_t *tlen, const int32_t *band_offset,
                                      int64_t *tickets, int32_t *n_submitted)
{
    GMFA_REQUIRE(s && n_submitted, "NULL argument");
    *n_submitted = 0;
    GMFA_REQUIRE(n >= 0, "negative n");
    if (n == 0) return GMFA_OK;
    GMFA_REQUIRE(cls && code_off && qlen && t_off && tlen && band_offset && tickets,
                 "NULL request array");
    const int nc = (int)s->cls.size();
    std::vector<char> touched(nc, 0), was_empty(nc, 0);
    for (int c = 0; c < nc; ++c) was_empty[c] = s->cls[c].lo == s->cls[c].next;
    int rc = GMFA_OK, k = 0;
    for (; k < n; ++k) {
        if (code_off[k] < 0) {
            set_error("request %d: negative code offset", k);
            rc = GMFA_E_INVALID;
            break;
        }
        rc = stage_request(s, cls[k], codes ? codes + code_off[k] : nullptr, qlen[k], two_bit,
                           t_off[k], tlen[k], band_offset[k], &tickets[k]);
        if (rc != GMFA_OK) break;
        touched[cls[k]] = 1;
    }
    *n_submitted = k;
    // one tail store per class for the whole burst
    for (int c = 0; c < nc; ++c) {
        if (!touched[c]) continue;
        const int prc = publish(s, s->cls[c], was_empty[c] != 0);
        if (prc != GMFA_OK && rc == GMFA_OK) rc = prc;
    }
    return rc;
}

extern "C" int gmfa_serve_poll(gmfa_serve_session *s, gmfa_serve_result *results,
                               int32_t max_results, uint32_t *cigar, int64_t cigar_cap,
                               int32_t *n_results)
{
    GMFA_REQUIRE(s && n_results, "NULL argument");
    GMFA_REQUIRE(max_results >= 0 && (max_results == 0 || results), "bad results buffer");
    GMFA_REQUIRE(cigar_cap >= 0 && (cigar_cap == 0 || cigar), "bad cigar buffer");
    *n_results = 0;
    int n = 0;
    int64_t coff = 0;
    for (int ci = 0; ci < (int)s->cls.size(); ++ci) {
        OnlineClass &c = s->cls[ci];
        const int32_t ring = c.cfg.ring, cap = c.cfg.cigar_cap;
        // completions in the order the device recorded them; a record not yet
        // written (its stamp is not the expected position) ends the scan
        while (n < max_results) {
            const uint64_t v = __atomic_load_n(&c.cq[c.cq_head & (ring - 1)], __ATOMIC_ACQUIRE);
            if ((uint32_t)(v >> 32) != (uint32_t)(c.cq_head + 1)) break;
            const uint32_t r = (uint32_t)v;
            const int64_t t = r < (uint32_t)ring ? c.occupant[r] : -1;
            if (t < 0) {
                set_error("class %d: corrupt completion record %llx", ci, (unsigned long long)v);
                *n_results = n;
                return GMFA_E_RUNTIME;
            }
            const ServeRes &e = c.res[r];
            const int nc = e.n_cigar;
            const bool nopages = nc < 0;
            const bool trunc = nc > cap;
            const int keep = (trunc || nopages) ? 0 : nc;
            if (coff + keep > cigar_cap) break;
            ++c.cq_head;
            gmfa_serve_result &o = results[n++];
            o.ticket = t;
            o.cls = ci;
            o.status = nopages ? GMFA_SERVE_NO_PAGES : trunc ? GMFA_SERVE_TRUNCATED : 0;
            o.score = e.score;
            o.query_end = e.query_end;
            o.target_end = e.target_end;
            o.read_start = e.read_start;
            o.ref_start = e.ref_start;
            o.n_cigar = nc;
            for (int k = 0; k < 4; ++k) o.counts[k] = e.counts[k];
            o.cigar_off = coff;
            o.claim_tick = e.claim_tick;
            o.done_tick = e.done_tick;
            if (keep)
                std::memcpy(cigar + coff, c.cig + (size_t)r * cap + (cap - keep),
                            sizeof(uint32_t) * (size_t)keep);
            coff += keep;
            c.occupant[r] = -1;
            ++s->returned;
            ++s->completed;
        }
        while (c.lo < c.next && c.occupant[(int)(c.lo % ring)] != c.lo) ++c.lo;
        const int rc = ensure_running(s, c, false);
        if (rc != GMFA_OK) {
            *n_results = n;
            return rc;
        }
    }
    *n_results = n;
    return GMFA_OK;
}

extern "C" int gmfa_serve_session_stats(const gmfa_serve_session *s, gmfa_serve_stats *st)
{
    GMFA_REQUIRE(s && st, "NULL argument");
    st->submitted = s->submitted;
    st->returned = s->returned;
    st->in_flight = s->submitted - s->returned;
    st->launches = s->launches;
    st->ticks_per_s = s->ticks_per_s;
    return GMFA_OK;
}

extern "C" int gmfa_serve_close(gmfa_serve_session *s, gmfa_serve_report *report)
{
    GMFA_REQUIRE(s, "NULL session");
    // publish nothing more, let the waves drain what is published, then count
    for (auto &c : s->cls) __atomic_store_n(&c.ctl->closed.v, 1, __ATOMIC_RELEASE);
    int rc = GMFA_OK;
    for (auto &c : s->cls) {
        // a class whose waves all left on idle with requests pending runs once
        // more to drain them (it sees closed and exits when done)
        if (rc == GMFA_OK) rc = ensure_running(s, c, true);
        if (hipStreamSynchronize(c.stream) != hipSuccess && rc == GMFA_OK) {
            set_error("serving class kernel failed while draining");
            rc = GMFA_E_RUNTIME;
        }
    }
    int64_t done_unpolled = 0;
    for (auto &c : s->cls) {
        for (int64_t t = c.lo; t < c.next; ++t) {
            const int r = (int)(t % c.cfg.ring);
            if (c.occupant[r] == t && __atomic_load_n(&c.done[r], __ATOMIC_ACQUIRE) == t + 1)
                ++done_unpolled;
        }
    }
    if (report) {
        std::memset(report, 0, sizeof(*report));
        report->completed = s->completed + done_unpolled;
    }
    gmfa::free_session(s);
    return rc;
}
