// gmfa_sw_forward_noalloc: the forward pass with the CIGAR walk fused
// (gmfa_sw_forward with GMFA_SW_TRACEBACK | GMFA_SW_CIGAR) for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops).  Its scratch -- the work counter and the traceback planes, one slab
// per wavefront reused pair after pair -- lives in `scratch`, a static device
// array of GMFA_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds
// until the process ends.  Calls on one GPU share it: calls that overlap on
// one GPU (two streams, two threads) race.
#include "common.h"
#include "sw_internal.h"

#include <algorithm>
#include <mutex>

namespace gmfa {

static __device__ __align__(256) uint32_t scratch[GMFA_NOALLOC_SCRATCH_BYTES / 4];

// Device address of `scratch` per device.  In host code the array's name is
// a host-side stand-in, not a device address, so kernels get these instead.
constexpr int kMaxDevices = 64;
static uint32_t *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_ptr(uint32_t **p)
{
    int dev = 0;
    GMFA_CHECK_HIP(hipGetDevice(&dev));
    GMFA_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d the library tracks", dev,
                 kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *q = nullptr;
        GMFA_CHECK_HIP(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
        GMFA_REQUIRE(q != nullptr, "no device address for the static scratch");
        g_scratch_ptr[dev] = static_cast<uint32_t *>(q);
    }
    *p = g_scratch_ptr[dev];
    return GMFA_OK;
}

}  // namespace gmfa

using namespace gmfa;

extern "C" int gmfa_sw_forward_noalloc(const gmfa_sw_batch *batch, const gmfa_sw_params *params,
                                       const gmfa_sw_out *out, void *stream)
{
    GMFA_REQUIRE(batch && params && out, "NULL argument");
    GMFA_REQUIRE((params->flags & GMFA_SW_TRACEBACK) && (params->flags & GMFA_SW_CIGAR) &&
                     !(params->flags & GMFA_SW_ROWKEYS),
                 "gmfa_sw_forward_noalloc needs GMFA_SW_TRACEBACK | GMFA_SW_CIGAR and no row keys");
    GMFA_REQUIRE(!out->tb && !out->tb_off && !out->work_counter && !out->rowkey,
                 "gmfa_sw_forward_noalloc keeps its planes and work counter in the library's "
                 "scratch: out->tb, tb_off, work_counter and rowkey must be NULL");
    GMFA_REQUIRE(batch->n_pairs >= 0, "negative n_pairs");
    if (batch->n_pairs == 0) return GMFA_OK;
    // the size limit before any device call: one wavefront's slabs must fit
    gmfa_sw_tile tile;
    int rc = gmfa_sw_resolve_tile(params->half_band, params->cells_per_lane,
                                  params->lanes_per_pair, &tile);
    if (rc != GMFA_OK) return rc;
    GMFA_REQUIRE(batch->max_qlen >= 0 && batch->max_qlen <= (1 << 20),
                 "query length bound %d outside [0, 2^20]", batch->max_qlen);
    const int64_t cap_words = GMFA_NOALLOC_SCRATCH_BYTES / 4 - 64;  // word 0: the work counter
    const int64_t slab = (gmfa_sw_tb_bytes(batch->max_qlen, &tile) / 4 + 63) / 64 * 64;
    const int64_t per_wave = std::max<int64_t>(slab, 64) * (64 / tile.lanes_per_pair);
    GMFA_REQUIRE(per_wave <= cap_words,
                 "the traceback planes of one wavefront (query length %d, half_band %d, tile "
                 "%dx%d) need %lld bytes, above the %lld bytes of the library's static scratch",
                 batch->max_qlen, params->half_band, tile.cells_per_lane, tile.lanes_per_pair,
                 (long long)(per_wave * 4), (long long)(cap_words * 4));
    uint32_t *base = nullptr;
    rc = scratch_ptr(&base);
    if (rc != GMFA_OK) return rc;
    SlabPlanes sp;
    sp.counter = base;
    sp.tb = base + 64;
    sp.cap_words = cap_words;
    return sw_launch(batch, params, out, nullptr, (hipStream_t)stream, &sp);
}
