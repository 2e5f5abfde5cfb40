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

}  // namespace fsim
