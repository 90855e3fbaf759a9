// common.h -- shared types and helpers of the framecor HIP kernels (gfx950 / CDNA4)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_set>

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
    const int32_t *box2;        // n x bh x bw: W x W box sums of the squares of the pixels (int8 form)
    const int2 *screen;         // n x bh x bw: {box, 1 / sqrt(W^2 box2 - box^2) float bits} (int8 form)
    const float *sigma;         // n x bh x bw: sqrtf(float(W^2 box2 - box^2) / float(W^2)), 0 if flat (FFT form)
    int n;
    float *offset, *peak, *snr, *cov;
};

constexpr int kXcds = 8;          // XCDs of an MI355X, each with its own L2
constexpr int kQueueClips = 48;   // clips per launch of a work queue

// The work queue of one tile shape over clips of a batch (a kernel argument). Items are (clip, window,
// run of frames), numbered clip by clip and, within a clip, window-major (consecutive items share a
// template; locate) or run of frames major (locateRunMajor): segment s covers items first[s] .. first[s + 1] - 1 of clip clip[s], nwin x nfb items of
// fpb frames (the last run of a window may be shorter). The items are dealt into kXcds contiguous
// ranges, one per XCD, each drained through its own counter (counters: kXcds ints, zero at launch);
// a block takes items from its own XCD's range (blocks are dealt to the XCDs round-robin), then steals
// from the others; -1 when every range is drained.
struct WorkQueue {
    int *counters;
    int nitems, nseg;
    int clip[kQueueClips], first[kQueueClips + 1], nfb[kQueueClips], fpb[kQueueClips], frames[kQueueClips];

    __device__ int next(int block) const
    {
        const int home = block & (kXcds - 1);
        for (int t = 0; t < kXcds; t++) {
            const int x = (home + t) & (kXcds - 1);
            const int lo = int(int64_t(nitems) * x / kXcds), hi = int(int64_t(nitems) * (x + 1) / kXcds);
            if (lo >= hi || __atomic_load_n(&counters[x], __ATOMIC_RELAXED) >= hi - lo) continue;
            const int i = atomicAdd(&counters[x], 1);
            if (i < hi - lo) return lo + i;
        }
        return -1;
    }

    __device__ void locate(int it, int &c, int &k, int &fBegin, int &fEnd) const
    {
        int s = 0;
        while (s + 1 < nseg && it >= first[s + 1]) s++;
        c = clip[s];
        const int local = it - first[s];
        k = local / nfb[s];
        fBegin = (local - k * nfb[s]) * fpb[s];
        fEnd = min(fBegin + fpb[s], frames[s]);
    }

    // the same with the items of each clip numbered run of frames major: consecutive items are one run
    // of frames over consecutive windows, so an XCD's L2 holds rows of that run's frames (the int8
    // kernels: 3 % less time on a 16-clip batch than window major)
    __device__ void locateRunMajor(int it, int &c, int &k, int &fBegin, int &fEnd) const
    {
        int s = 0;
        while (s + 1 < nseg && it >= first[s + 1]) s++;
        c = clip[s];
        const int local = it - first[s], nw = (first[s + 1] - first[s]) / nfb[s];
        const int run = local / nw;
        k = local - run * nw;
        fBegin = run * fpb[s];
        fEnd = min(fBegin + fpb[s], frames[s]);
    }
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
    // in int64: gross offsets are any int32
    const int64_t ly = int64_t(y0) + gy - g.R, lx = int64_t(x0) + gx - g.R;
    const bool in = ly >= 0 && lx >= 0 && ly + g.L <= g.h && lx + g.L <= g.w;
    sy = in ? int(ly) : 0;
    sx = in ? int(lx) : 0;
    return in;
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

namespace framecor {

// lets a kernel use up to the whole 160 KB of LDS of a CU (once per kernel and process)
inline void allowLargeLds(const void *kernel)
{
    static std::mutex lock;
    static std::unordered_set<const void *> done;
    std::lock_guard<std::mutex> guard(lock);
    if (!done.insert(kernel).second) return;
    (void)hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipGetLastError();
}

} // namespace framecor
