// Launch geometry shared by the stand-alone FFT and the fused kernels.
#pragma once

namespace fsim {

// Column tile width: 16 columns (128-byte row segments) while an L x CW
// complex tile fits 128 KB of LDS (gfx950 allows 160 KB per workgroup).
template <int L>
struct ColW {
    static constexpr int v0 = 131072 / (8 * L);
    static constexpr int value = v0 >= 16 ? 16 : (v0 < 1 ? 1 : v0);
};

// Tile width for a column kernel whose transform uses T threads and E elements
// per thread: as ColW, capped so that the block stays within 1024 threads
// (512 when E >= 32, which needs more than 128 VGPRs per thread).
template <int L, int T, int E>
struct ColTile {
    static constexpr int maxthreads = E >= 32 ? 512 : 1024;
    static constexpr int v0 = ColW<L>::value;
    static constexpr int v1 = maxthreads / T;
    static constexpr int value = v0 < v1 ? v0 : v1;
};

// Work index of this block such that each of the 8 XCDs (workgroups are
// dealt to them round robin by linear id) walks one contiguous range of
// [0, total): neighbouring tiles of a component run on one XCD.  The grid
// has xcd_grid(total) blocks; indices >= total have no work.
__device__ __forceinline__ int xcd_linear_at(unsigned b, int total) {
    const int per = (total + 7) / 8;
    return (int)(b % 8) * per + (int)(b / 8);
}
__device__ __forceinline__ int xcd_linear(int total) { return xcd_linear_at(blockIdx.x, total); }
static inline unsigned xcd_grid(int total) { return (unsigned)((total + 7) / 8) * 8; }

}  // namespace fsim
