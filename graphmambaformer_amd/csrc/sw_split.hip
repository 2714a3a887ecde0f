// Split-pair kernels: W = 1503 (half_band 751) over the 24 x 64 tile as two
// slices of 12 x 64, three of 8 x 64 or four of 6 x 64
// (one translation unit per group of variants, so the library builds in
// parallel; the host side that picks among them is sw.hip).
#include "sw_kernels.h"

namespace gmfa {

const SplitEntry *sw_split_entries(int *n)
{
    static const SplitEntry entries[] = {
        GMFA_SPLIT(12, 2, 2, 24, 14),
        GMFA_SPLIT(8, 3, 6, 24, 14),
        GMFA_SPLIT(6, 4, 2, 24, 14),
    };
    *n = (int)(sizeof(entries) / sizeof(entries[0]));
    return entries;
}

}  // namespace gmfa
