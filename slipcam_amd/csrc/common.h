// common.h -- shared types and helpers of the framecor HIP kernels (gfx950 / CDNA4)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

namespace framecor {

constexpr int kWave = 64;

// The frames, the window grid and the correlation sizes (cxx/framecor.h:22-35 of the reference),
// plus the derived sizes the kernels use. Passed by value to every kernel.
struct Geometry {
    int h, w;           // frame size
    int W, R, L, M;     // window, half search range, search window W + 2R, surface 2R + 1
    int ny, nx, nwin;   // window grid
    int stat;           // SNR box
    int bh, bw;         // the box-sum maps: (h - W + 1) x (w - W + 1)
    int64_t pixels;     // h * w
    int64_t satSize;    // (h + 1) * (w + 1)
    int64_t boxSize;    // bh * bw
};

// Per-window constants on the device, written by set_reference
struct WindowData {
    const int *down, *across;   // first pixels of the rows / columns of windows
    const float *norm;          // FP32 norm of the zero-mean reference window (0: flat, not valid)
    const int64_t *tsum;        // sum(T) over the W x W reference window (exact)
    const int64_t *tvar;        // W^2 sum(T^2) - sum(T)^2 (exact)
};

// One group of frames on the device: inputs, tables and outputs (outputs already offset to the
// group's first frame; q = f * nwin + k indexes them)
struct FrameGroup {
    const uint8_t *frames;      // n x h x w
    const int *gross;           // n x nwin x 2, or null
    const int64_t *sat;         // n x (h + 1) x (w + 1), the summed-area table of the pixels
    const int32_t *box, *box2;  // n x bh x bw: W x W box sums of the pixels and of their squares
    const int2 *screen;         // n x bh x bw: {box, 1 / sqrt(W^2 box2 - box^2) float bits} (int8 form)
    int n;
    float *offset, *peak, *snr, *cov;
};

// Blocks of a grid are dealt round-robin to the 8 XCDs of an MI355X, each with its own L2. Map the
// hardware block index to a logical index such that each XCD runs a contiguous range of logical
// blocks (neighbouring windows share frame rows and box-sum rows in that XCD's L2).
__device__ inline int xcdRemap(int b, int nblocks)
{
    const int xcd = b & 7, slot = b >> 3, per = nblocks >> 3, extra = nblocks & 7;
    return xcd * per + min(xcd, extra) + slot;
}

// the outputs of a window that could not be correlated (cxx/surface.h:48-54)
__device__ inline void invalidateOut(int64_t q, float *offset, float *peak, float *snr, float *cov)
{
    offset[2 * q] = __builtin_nanf("");
    offset[2 * q + 1] = __builtin_nanf("");
    peak[q] = 0.0f;
    snr[q] = 0.0f;
    cov[3 * q] = 99.0f;
    cov[3 * q + 1] = 99.0f;
    cov[3 * q + 2] = 0.0f;
}

// the search window of window k with gross offset (gy, gx); whether it is in the frame
// (cxx/surface.h:39-45)
__device__ inline bool searchWindow(const Geometry &g, int y0, int x0, int gy, int gx, int &sy, int &sx)
{
    sy = y0 + gy - g.R;
    sx = x0 + gx - g.R;
    return sy >= 0 && sx >= 0 && sy + g.L <= g.h && sx + g.L <= g.w;
}

// wave-wide reductions (64 lanes)
__device__ inline double waveSumD(double x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
    return x;
}

__device__ inline int waveSumI(int x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
    return x;
}

// the larger (value, index) pair; equal values: the lower index (the first in row-major order)
template <class T>
__device__ inline void bestOf(T &v, int &i, T v2, int i2)
{
    if (v2 > v || (v2 == v && i2 < i)) {
        v = v2;
        i = i2;
    }
}

template <class T>
__device__ inline void waveBest(T &v, int &i)
{
    for (int o = 32; o > 0; o >>= 1) {
        const T v2 = __shfl_xor(v, o, kWave);
        const int i2 = __shfl_xor(i, o, kWave);
        bestOf(v, i, v2, i2);
    }
}

} // namespace framecor

// host-side error helpers
#define FC_HIP_CHECK(call)                                                                          \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            throw std::runtime_error(std::string("framecor: ") + #call + ": " + hipGetErrorString(e_)); \
    } while (0)
