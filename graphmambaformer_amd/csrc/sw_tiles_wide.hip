// Forward-kernel variants of the general 16 x 64 and 32 x 64 tiles
// (one translation unit per group of variants, so the library builds in
// parallel; the host side that picks among them is sw.hip).
#include "sw_kernels.h"

namespace gmfa {

const TileEntry *sw_tiles_wide(int *n)
{
    static const TileEntry entries[] = {
        GMFA_TILE(16, 64, -1),
        GMFA_TILE(32, 64, -1),
    };
    *n = (int)(sizeof(entries) / sizeof(entries[0]));
    return entries;
}

}  // namespace gmfa
