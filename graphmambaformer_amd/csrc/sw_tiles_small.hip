// Forward-kernel variants of the narrow tiles: 2 and 4 x 64, the 16- and
// 32-lane tiles that pack several short-band pairs into one wavefront (W = 65
// as 5 x 13 among them), W = 65 as one cell per lane plus the scalar cell
// (the latency tile of small serving classes) and the skewed tile (two
// diagonals per lane on 63 lanes: the latency tile of small offline batches
// of bands up to 126)
// (one translation unit per group of variants, so the library builds in
// parallel; the host side that picks among them is sw.hip).
#include "sw_kernels.h"

namespace gmfa {

const TileEntry *sw_tiles_small(int *n)
{
    static const TileEntry entries[] = {
        GMFA_TILE(5, 16, 4),
        GMFA_TILE(2, 16, -1),
        GMFA_TILE(5, 16, -1),
        GMFA_TILE(8, 16, -1),
        GMFA_TILE(4, 32, -1),
        GMFA_TILE(8, 32, -1),
        GMFA_TILE(2, 64, -1),
        GMFA_TILE(4, 64, -1),
        GMFA_TILE(1, 64, 1),
        GMFA_SKEW_TILE,
    };
    *n = (int)(sizeof(entries) / sizeof(entries[0]));
    return entries;
}

}  // namespace gmfa
