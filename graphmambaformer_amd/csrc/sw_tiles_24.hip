// Forward-kernel variants of the 24 x 64 tile: the exact fit of W = 1503
// (half_band 751, 15 live registers in the last lane) and the general one
// (one translation unit per group of variants, so the library builds in
// parallel; the host side that picks among them is sw.hip).
#include "sw_kernels.h"

namespace gmfa {

const TileEntry *sw_tiles_24(int *n)
{
    static const TileEntry entries[] = {
        GMFA_TILE(24, 64, 14),
        GMFA_TILE(24, 64, -1),
    };
    *n = (int)(sizeof(entries) / sizeof(entries[0]));
    return entries;
}

}  // namespace gmfa
