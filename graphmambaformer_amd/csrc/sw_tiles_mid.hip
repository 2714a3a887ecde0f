// Forward-kernel variants of the 8, 9 and 12 x 64 tiles: W = 513 as 9 x 57
// or 8 x 64 plus the scalar cell, and the general ones
// (one translation unit per group of variants, so the library builds in
// parallel; the host side that picks among them is sw.hip).
#include "sw_kernels.h"

namespace gmfa {

const TileEntry *sw_tiles_mid(int *n)
{
    static const TileEntry entries[] = {
        GMFA_TILE(9, 64, 8),
        GMFA_TILE(8, 64, 8),
        GMFA_TILE(8, 64, -1),
        GMFA_TILE(9, 64, -1),
        GMFA_TILE(12, 64, -1),
    };
    *n = (int)(sizeof(entries) / sizeof(entries[0]));
    return entries;
}

}  // namespace gmfa
