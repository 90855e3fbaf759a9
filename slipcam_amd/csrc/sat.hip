// sat.hip -- summed-area tables of uint8 frames (exact int64) and the W x W box-sum maps
//
// The tables are those of the reference (cxx/cpu.cpp:24-42, cxx/cuda.cu:72-110): (h + 1) x (w + 1)
// int64 tables of the pixels and of their squares, row and column 0 zero. Here:
//   satRows  -- one wavefront per frame row: an inclusive prefix scan of 64 pixels per step with a
//               running carry (the scan is exact in int32: a row of squares sums to at most
//               w * 65025 < 2^31 for w < 33000), written as int64 row prefixes;
//   satCols  -- one thread per table column: the running column sums of the row prefixes, loads
//               batched eight rows ahead;
//   boxSums  -- the W x W box sums of the pixels and of their squares at every window position,
//               (h - W + 1) x (w - W + 1) int32 per frame, from four table reads each (exact: a box of
//               squares sums to at most W^2 * 65025 < 2^31 for W <= 181), and optionally a packed map of
//               {sum F, 1 / sqrt(W^2 sum F^2 - (sum F)^2) as FP32 (0 for a flat box)} that every window
//               touching that position shares (one 8-byte load per surface point in the int8 form's
//               screen). The correlation kernels read these maps instead of eight int64 table entries.

#include "common.h"
#include "kernels.h"

namespace framecor {

namespace {

constexpr int kRowsPerBlock = 4;

__device__ inline int waveInclusiveScan(int x, int lane)
{
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    return x;
}

__global__ void __launch_bounds__(kWave * kRowsPerBlock)
satRows(const uint8_t *frames, int h, int w, int64_t satSize, int64_t *sat, int64_t *sat2)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int y = blockIdx.x * kRowsPerBlock + (threadIdx.x >> 6);
    const int f = blockIdx.y;
    if (y >= h) return;
    const int64_t stride = w + 1;
    const uint8_t *row = frames + (size_t(f) * h + y) * w;
    int64_t *s = sat + f * satSize + (y + 1) * stride;
    int64_t *s2 = sat2 + f * satSize + (y + 1) * stride;
    if (y == 0) {   // the zero first row
        int64_t *z = sat + f * satSize, *z2 = sat2 + f * satSize;
        for (int x = lane; x <= w; x += kWave) {
            z[x] = 0;
            z2[x] = 0;
        }
    }
    if (lane == 0) {
        s[0] = 0;
        s2[0] = 0;
    }
    int carry = 0, carry2 = 0;
    for (int x0 = 0; x0 < w; x0 += kWave) {
        const int x = x0 + lane;
        const int px = x < w ? int(row[x]) : 0;
        const int a = waveInclusiveScan(px, lane) + carry;
        const int b = waveInclusiveScan(px * px, lane) + carry2;
        if (x < w) {
            s[x + 1] = a;
            s2[x + 1] = b;
        }
        carry = __shfl(a, kWave - 1, kWave);
        carry2 = __shfl(b, kWave - 1, kWave);
    }
}

__global__ void __launch_bounds__(256)
satCols(int h, int w, int64_t satSize, int64_t *sat, int64_t *sat2)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int f = blockIdx.y;
    if (x > w) return;
    const int64_t stride = w + 1;
    int64_t *s = sat + f * satSize + x, *s2 = sat2 + f * satSize + x;
    int64_t a = 0, b = 0;
    constexpr int kAhead = 8;
    int y = 1;
    for (; y + kAhead - 1 <= h; y += kAhead) {
        int64_t ra[kAhead], rb[kAhead];
#pragma unroll
        for (int i = 0; i < kAhead; i++) {
            ra[i] = s[(y + i) * stride];
            rb[i] = s2[(y + i) * stride];
        }
#pragma unroll
        for (int i = 0; i < kAhead; i++) {
            a += ra[i];
            b += rb[i];
            s[(y + i) * stride] = a;
            s2[(y + i) * stride] = b;
        }
    }
    for (; y <= h; y++) {
        a += s[y * stride];
        b += s2[y * stride];
        s[y * stride] = a;
        s2[y * stride] = b;
    }
}

__global__ void __launch_bounds__(256)
boxSums(const int64_t *sat, const int64_t *sat2, int h, int w, int W, int64_t satSize, int64_t boxSize,
        int32_t *box, int32_t *box2, int2 *screen)
{
    const int bw = w - W + 1;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int f = blockIdx.z;
    if (x >= bw) return;
    const int64_t stride = w + 1;
    const int64_t *s = sat + f * satSize, *s2 = sat2 + f * satSize;
    const int64_t t0 = y * stride + x, t1 = (y + W) * stride + x;
    const int64_t b = s[t1 + W] - s[t0 + W] - s[t1] + s[t0];
    const int64_t b2 = s2[t1 + W] - s2[t0 + W] - s2[t1] + s2[t0];
    const int64_t o = f * boxSize + int64_t(y) * bw + x;
    box[o] = int32_t(b);
    box2[o] = int32_t(b2);
    if (screen) {
        const int64_t v = int64_t(W) * W * b2 - b * b;
        screen[o] = make_int2(int32_t(b), __float_as_int(v > 0 ? float(1.0 / sqrt(double(v))) : 0.0f));
    }
}

} // namespace

void launchSummedArea(const uint8_t *frames, int n, int h, int w, int64_t *sat, int64_t *sat2,
                      hipStream_t stream)
{
    if (n <= 0) return;
    const int64_t satSize = int64_t(h + 1) * (w + 1);
    hipLaunchKernelGGL(satRows, dim3((h + kRowsPerBlock - 1) / kRowsPerBlock, n), dim3(kWave * kRowsPerBlock),
                       0, stream, frames, h, w, satSize, sat, sat2);
    hipLaunchKernelGGL(satCols, dim3((w + 255) / 256, n), dim3(256), 0, stream, h, w, satSize, sat, sat2);
    FC_HIP_CHECK(hipGetLastError());
}

void launchBoxSums(const int64_t *sat, const int64_t *sat2, int n, int h, int w, int W, int32_t *box,
                   int32_t *box2, int2 *screen, hipStream_t stream)
{
    if (n <= 0) return;
    const int64_t satSize = int64_t(h + 1) * (w + 1);
    const int bh = h - W + 1, bw = w - W + 1;
    const int64_t boxSize = int64_t(bh) * bw;
    hipLaunchKernelGGL(boxSums, dim3((bw + 255) / 256, bh, n), dim3(256), 0, stream, sat, sat2, h, w, W,
                       satSize, boxSize, box, box2, screen);
    FC_HIP_CHECK(hipGetLastError());
}

} // namespace framecor
