// Output tile of the 32 -> 32 convolution kernel and its launch grid (shared by
// the kernel and the host code that launches it).
#pragma once
#include <hip/hip_runtime.h>

namespace ctlab {
constexpr int kTX = 8, kTZ = 32;                    // 8 x-columns x 32 z of one y row

// 32 -> 32 layers: the tile for nyb consecutive y rows per block
__host__ __device__ inline long conv32_tiles(int XN, int YN, int ZN, int nyb)
{
    return (long)((YN + nyb - 1) / nyb) * ((XN + kTX - 1) / kTX) * ((ZN + kTZ - 1) / kTZ);
}

// rows per block: the largest even count up to 32 that still gives one block
// per CU (256 on MI355X); a block's set-up (all weights, two bricks) is then
// spread over many rows
inline int conv32_rows(int XN, int YN, int ZN)
{
    int nyb = 32;
    while (nyb > 2 && conv32_tiles(XN, YN, ZN, nyb) < 256) nyb >>= 1;
    return nyb;
}

// tiles rounded up to a multiple of the 8 XCDs (the kernel deals contiguous
// tile ranges to the XCDs and drops the blocks past the last tile)
inline unsigned conv32_grid(int XN, int YN, int ZN, int nyb)
{
    return (unsigned)((conv32_tiles(XN, YN, ZN, nyb) + 7) / 8 * 8);
}
// first layer (1 -> 32): the same tile for kFY consecutive y rows per block
constexpr int kFY = 8;

__host__ __device__ inline long conv_first_tiles(int XN, int YN, int ZN)
{
    return (long)((YN + kFY - 1) / kFY) * ((XN + kTX - 1) / kTX) * ((ZN + kTZ - 1) / kTZ);
}
inline unsigned conv_first_grid(int XN, int YN, int ZN)
{
    return (unsigned)((conv_first_tiles(XN, YN, ZN) + 7) / 8 * 8);
}
// output layer (32 -> 1): 8 x-columns x 32 z, for kLY consecutive y rows per block
constexpr int kLX = 8, kLY = 16;

__host__ __device__ inline long conv_last_tiles(int XN, int YN, int ZN)
{
    return (long)((YN + kLY - 1) / kLY) * ((XN + kLX - 1) / kLX) * ((ZN + kTZ - 1) / kTZ);
}
inline unsigned conv_last_grid(int XN, int YN, int ZN)
{
    return (unsigned)((conv_last_tiles(XN, YN, ZN) + 7) / 8 * 8);
}
}  // namespace ctlab
