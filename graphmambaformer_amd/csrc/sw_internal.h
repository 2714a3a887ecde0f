// Internal interface between the SW kernels (sw.hip) and the serving runtime
// (serve.hip).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gmfa.h"

namespace gmfa {

// Persistent serving launch: the forward kernel claims request k only once
// k < *tail, walks it into a CIGAR in the same launch, then sets done[k] = 1.
// tail / closed / done / gave_up are device-visible pointers into
// host-coherent memory.
struct ServeLaunch {
    const int32_t *tail;
    const int32_t *closed;
    int32_t *done;
    int32_t *gave_up;
    int64_t slab_words;       // traceback plane words per wave slot (out->tb holds n_slots slabs)
    uint64_t deadline_ticks;  // bound of one idle wait of a wave, wall-clock ticks
    int32_t n_slots;          // persistent wavefronts
};

// Validates and launches the forward kernel (work-counter reset + the kernel)
// on stream s; serve == nullptr is the offline batch of gmfa_sw_forward.
int sw_launch(const gmfa_sw_batch *batch, const gmfa_sw_params *params, const gmfa_sw_out *out,
              const ServeLaunch *serve, hipStream_t s);

}  // namespace gmfa
