// Internal interface between the SW kernels (sw.hip) and the serving runtime
// (serve.hip).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gmfa.h"

namespace gmfa {

// One submitted request of an online session, written by the host into
// host-coherent memory before the class's tail is advanced past it.
struct alignas(32) ServeDesc {
    int64_t t_off;        // first target base in the resident reference
    int32_t qlen, tlen;
    int32_t band_offset;
    int32_t q_words;      // 2-bit words of the query in its ring entry
    int32_t pad[2];
};

// The outputs of one request of an online session (host-coherent memory),
// complete once the entry's done word holds ticket + 1.
struct alignas(64) ServeRes {
    int32_t score, query_end, target_end, read_start, ref_start, n_cigar;
    int32_t counts[4];
    int64_t claim_tick, done_tick;  // device wall clock at the claim / the record
    int32_t pad[2];
};
static_assert(sizeof(ServeRes) == 64 && sizeof(ServeDesc) == 32, "ring entry layouts");

// Device-memory control words of one serving class (64 bytes of the caller's
// device memory: gmfa_serve_class.out.work_counter / gmfa_serve_class_cfg.head).
// Idle wavefronts poll these mirrors in device memory; at most one wavefront
// per poll interval reads the host-coherent words over the host link and
// raises the mirrors (a read of host memory by every idle wavefront clogs the
// memory queues that the working wavefronts share).
struct alignas(64) ServeDevCtl {
    unsigned long long head;       // online: next ticket; staged: uint32 work counter (low word)
    unsigned long long tail;       // mirror of the host tail
    int32_t closed;                // mirror of the host closed flag
    int32_t pad0;
    unsigned long long last_poll;  // wall-clock tick of the last read of the host words
    unsigned long long cq_pos;     // online: completions so far (completion queue position)
    unsigned long long page_hint;  // paged planes: where the next page search starts
    unsigned long long pad[2];
};
static_assert(sizeof(ServeDevCtl) == 64, "serving device control block");

// Persistent serving launch.  Staged form (ring_n == 0): the forward kernel
// claims request k only once k < *tail, walks it into a CIGAR in the same
// launch, then sets done[k] = 1.  Online form (ring_n > 0): a wavefront claims
// the next published ticket from *head, copies that ring entry's query into
// its slot of `stage`, aligns it and writes its outputs into the host ring.
// Pointers named "host-coherent" are device-visible pointers into pinned
// memory that the host thread reads and writes while the kernel runs.
struct ServeLaunch {
    const int64_t *tail;      // staged: requests published (host-coherent)
    const int32_t *closed;    // host-coherent: nothing more will be published
    int32_t *done;            // staged: [n_pairs] completion flags (host-coherent)
    int32_t *gave_up;         // host-coherent: a wave's idle wait hit its bound
    int32_t *started;         // host-coherent: set by the launch's first wavefront
    int64_t slab_words;       // traceback plane words per wave slot (out->tb holds n_slots slabs)
    uint64_t deadline_ticks;  // bound of one idle wait of a wave, wall-clock ticks
    uint64_t poll_ticks;      // least interval between two reads of the host words
    int32_t n_slots;          // persistent wavefronts
    // online form
    int32_t ring_n;                  // ring entries (0: staged form)
    const int64_t *ring_tail;        // host-coherent: tickets published
    ServeDevCtl *dctl;               // device control words (staged: out->work_counter)
    const ServeDesc *ring_desc;      // host-coherent [ring_n]
    const uint32_t *ring_q;          // host-coherent [ring_n][2 * stage_words]
    ServeRes *ring_res;              // host-coherent [ring_n]
    uint32_t *ring_cigar;            // host-coherent [ring_n][cigar_cap]
    int64_t *ring_done;              // host-coherent [ring_n]: ticket + 1 when complete
    uint64_t *ring_cq;               // host-coherent [ring_n] completion queue: completion k
                                     // writes (k + 1) << 32 | entry at k mod ring_n
    uint32_t *stage;                 // device [n_slots][2 * stage_words]
    int32_t stage_words;             // query words per slot (mask words follow)
    int32_t cigar_cap;               // CIGAR runs kept per request
    // paged traceback planes (page_blocks > 0): out->tb is a pool of n_pages
    // pages of page_blocks 8-row plane blocks; a wavefront takes the pages of
    // its request (all or nothing) before the forward pass and frees them after
    // the walk, so requests are admitted by the planes' HBM footprint.
    int32_t page_blocks;             // power of two
    int32_t n_pages;
    int32_t *page_owner;             // device [n_pages], 0 = free
    uint64_t alloc_ticks;            // bound of one wait for pages
};

// pages one request may hold (the wavefront's page list is an LDS array)
constexpr int kMaxPages = 256;

// Zero a class's page-owner words (a kernel: no memset nodes on a stream a
// caller might capture).
int serve_reset_pages(int32_t *owner, int32_t n_pages, hipStream_t s);

// Offline batch whose traceback planes live in per-wave slabs instead of one
// region per pair: the CIGAR walk runs in the same launch, so a slab is free
// again once its pair is walked.  Pair group g of wavefront w writes its
// planes at tb + (w * pairs_per_wave + g) * slab_words; the launch takes as
// many wavefronts as the slabs fit in cap_words (gmfa_sw_forward_noalloc).
struct SlabPlanes {
    uint32_t *tb;        // cap_words words of device memory
    uint32_t *counter;   // work counter (one word, reset by the launch)
    int64_t cap_words;   // words at tb
};

// Validates and launches the forward kernel (work-counter reset + the kernel)
// on stream s; serve == nullptr and slabs == nullptr is the offline batch of
// gmfa_sw_forward.  The online form does not reset its head (relaunches
// continue the ticket count).
int sw_launch(const gmfa_sw_batch *batch, const gmfa_sw_params *params, const gmfa_sw_out *out,
              const ServeLaunch *serve, hipStream_t s, const SlabPlanes *slabs = nullptr);

// Query words a stage slot holds for queries up to max_qlen bases.
inline int32_t serve_stage_words(int32_t max_qlen) { return (max_qlen + 15) / 16 + 2; }

}  // namespace gmfa
