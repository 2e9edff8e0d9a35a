// Workgroup -> output-tile mapping shared by the GEMM kernels.
#pragma once
#include <hip/hip_runtime.h>

// The grid is padded to a multiple of 8 workgroups.  Hardware dispatches
// workgroup w to XCD w % 8, so first renumber ids so that each XCD owns one
// contiguous range, then walk that range in groups of GROUP_M tile rows
// (column-major inside a group): the tiles resident on one XCD at a time then
// share a few A row-blocks and B column-blocks in that XCD's L2.
// Ids past the last tile come back as (tiles_m, tiles_n).
__device__ __forceinline__ void scirs_tile_of(int wg, int nwg_pad, int tiles_m, int tiles_n, int &tm, int &tn)
{
    constexpr int NXCD = 8;
    constexpr int GROUP_M = 8;
    const int id = (wg % NXCD) * (nwg_pad / NXCD) + wg / NXCD;
    if (id >= tiles_m * tiles_n) { tm = tiles_m; tn = tiles_n; return; }
    const int group = id / (GROUP_M * tiles_n);
    const int first_m = group * GROUP_M;
    const int gm = min(tiles_m - first_m, GROUP_M);
    const int in_group = id % (GROUP_M * tiles_n);
    tm = first_m + in_group % gm;
    tn = in_group / gm;
}
