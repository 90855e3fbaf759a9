// sat.hip -- summed-area tables of uint8 frames (exact int64) and the W x W box-sum maps
//
// The tables are those of the reference (cxx/cpu.cpp:24-42, cxx/cuda.cu:72-110): (h + 1) x (w + 1)
// int64 tables of the pixels and of their squares, row and column 0 zero. Here:
//   satRows  -- one wavefront per frame row: an inclusive prefix scan of 64 pixels per step with a
//               running carry (the scan is exact in int32: a row of squares sums to at most
//               w * 65025 < 2^31 for w < 33000), written as int64 row prefixes;
//   satCols  -- one thread per table column: the running column sums of the row prefixes, loads
//               batched eight rows ahead;
//   boxMaps  -- the W x W box sums of the pixels and of their squares at every window position,
//               (h - W + 1) x (w - W + 1) int32 per frame, straight from the frames with running
//               vertical sums and prefix scans across columns (no table: the correlator never
//               materialises the int64 tables; exact: a box of squares sums to at most
//               W^2 * 65025 < 2^31 for W <= 181), and optionally a packed map of
//               {sum F, 1 / sqrt(W^2 sum F^2 - (sum F)^2) to FP32 accuracy (0 for a flat box)} that every window
//               touching that position shares (one 8-byte load per surface point in the int8 form's
//               screen), or the FFT form's map of sqrtf(float(W^2 sum F^2 - (sum F)^2) / float(W^2))
//               (0 for a flat box): the FP32 deviation that cxx/surface.h:30-35 evaluates per surface
//               point, here once per position and frame. The correlation kernels read these maps instead of eight int64 table entries;
//               the maps equal the four-term differences of the tables (cxx/surface.h:21-25) exactly.

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

// The W x W box sums straight from the frames (no table): one block of kBoxThreads threads holds
// kBoxThreads consecutive input columns x0 .. (one per thread) and kBoxThreads - W + 1 output columns;
// it walks kBoxRows output rows down the frame, kBoxRound rows per round. Per row each thread moves the
// vertical W-row sums of its column one row down (one pixel in, one out; the pixels of a round are
// loaded a round ahead), every wave scans them into prefix sums across its 64 columns with DPP row
// shifts and row broadcasts, the waves' totals go through LDS (one barrier per round), and each output
// is the difference of two block prefixes W apart (a second barrier per round). The arithmetic is
// modulo 2^32 (uint32): the prefixes of squares may wrap, but every box sum is below 2^31, so the
// differences are exact. W^2 sum F^2 - (sum F)^2 is an integer below 2^53, exact in FP64.
constexpr int kBoxThreads = 512, kBoxRows = 64, kBoxRound = 4;

// inclusive prefix sum over the 64 lanes of a wave: shifts by 1, 2, 4, 8 within each row of 16 lanes
// (lanes shifted in from outside the row add 0), then lane 15 of each row into the next row and lane
// 31 into the upper two
__device__ inline uint32_t waveScanDpp(uint32_t v)
{
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x111, 0xf, 0xf, true));    // row_shr:1
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x112, 0xf, 0xf, true));    // row_shr:2
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x114, 0xf, 0xf, true));    // row_shr:4
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x118, 0xf, 0xf, true));    // row_shr:8
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x142, 0xa, 0xf, false));   // row_bcast:15
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x143, 0xc, 0xf, false));   // row_bcast:31
    return v;
}

// block (bx, by) of frame f: output columns bx (kBoxThreads - W + 1) .., rows by kBoxRows ..
__device__ __attribute__((always_inline)) inline void boxMapsBlock(const uint8_t *frames, int h, int w, int W,
                                                                   int64_t boxSize, int32_t *box, int32_t *box2,
                                                                   int2 *screen, float *sigma, int bx, int by, int f)
{
    constexpr int R = kBoxRound, NW = kBoxThreads / kWave;
    __shared__ uint32_t pre[R][2][kBoxThreads];
    __shared__ uint32_t tot[R][2][NW];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid >> 6;
    const int bh = h - W + 1, bw = w - W + 1, tx = kBoxThreads - W + 1;
    const int x0 = bx * tx, y0 = by * kBoxRows;
    const int x = x0 + tid, yEnd = min(y0 + kBoxRows, bh);
    const bool mine = x < w;
    // every pixel load is unconditional, at a clamped address inside the frame (column w - 1, row h - 1)
    // where its value is not used: no branch around the loads, so each wait counts exactly the loads
    // still in flight and the next round's loads stay in flight through this round
    const uint8_t *F = frames + size_t(f) * h * w + min(x, w - 1);
    // the vertical sums of rows y0 .. y0 + W - 1, eight loads in flight
    uint32_t vs = 0, vs2 = 0;
    {
        int i = 0;
        for (; i + 8 <= W; i += 8) {
            uint32_t p[8];
#pragma unroll
            for (int t = 0; t < 8; t++) p[t] = F[size_t(y0 + i + t) * w];
#pragma unroll
            for (int t = 0; t < 8; t++) {
                vs += p[t];
                vs2 += p[t] * p[t];
            }
        }
        for (; i < W; i++) {
            const uint32_t p = F[size_t(y0 + i) * w];
            vs += p;
            vs2 += p * p;
        }
        if (!mine) vs = vs2 = 0;
    }
    const double nd = double(W) * W;
    // the pixels entering (row y + W) and leaving (row y) the vertical sums after output row y, for the
    // rows y = yr .. yr + R - 1 of a round (used only where a later row is output: then y + W < h)
    auto loadRound = [&](int yr, uint32_t (&in)[R], uint32_t (&out)[R]) {
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int y = yr + k;
            in[k] = F[size_t(min(y + W, h - 1)) * w];
            out[k] = F[size_t(min(y, h - 1)) * w];
        }
    };
    uint32_t cin[R], cout[R];
    loadRound(y0, cin, cout);
    for (int yr = y0; yr < yEnd; yr += R) {
        uint32_t nin[R], nout[R];
        loadRound(yr + R, nin, nout);
        uint32_t s[R], s2[R];
#pragma unroll
        for (int k = 0; k < R; k++) {
            s[k] = vs;
            s2[k] = vs2;
            if (mine && yr + k + 1 < yEnd) {
                vs += cin[k] - cout[k];
                vs2 += cin[k] * cin[k] - cout[k] * cout[k];
            }
        }
#pragma unroll
        for (int k = 0; k < R; k++) {
            s[k] = waveScanDpp(s[k]);
            s2[k] = waveScanDpp(s2[k]);
        }
        if (lane == kWave - 1)
#pragma unroll
            for (int k = 0; k < R; k++) {
                tot[k][0][wave] = s[k];
                tot[k][1][wave] = s2[k];
            }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < R; k++) {
#pragma unroll
            for (int j = 0; j < NW - 1; j++)
                if (j < wave) {
                    s[k] += tot[k][0][j];
                    s2[k] += tot[k][1][j];
                }
            pre[k][0][tid] = s[k];
            pre[k][1][tid] = s2[k];
        }
        __syncthreads();
        if (tid < tx && x < bw)
#pragma unroll
            for (int k = 0; k < R; k++) {
                const int y = yr + k;
                if (y < yEnd) {
                    const uint32_t lo = tid ? pre[k][0][tid - 1] : 0u, lo2 = tid ? pre[k][1][tid - 1] : 0u;
                    const int32_t bs = int32_t(pre[k][0][tid + W - 1] - lo), bs2 = int32_t(pre[k][1][tid + W - 1] - lo2);
                    const int64_t o = f * boxSize + int64_t(y) * bw + x;
                    if (box) box[o] = bs;
                    if (box2) box2[o] = bs2;
                    // W^2 sum F^2 - (sum F)^2, exact: every term an integer below 2^53
                    const double v = fma(nd, double(bs2), -(double(bs) * double(bs)));
                    if (screen) screen[o] = make_int2(bs, __float_as_int(v > 0 ? rsqrtf(float(v)) : 0.0f));
                    if (sigma) sigma[o] = v > 0 ? sqrtf(float(v) / float(nd)) : 0.0f;
                }
            }
#pragma unroll
        for (int k = 0; k < R; k++) {
            cin[k] = nin[k];
            cout[k] = nout[k];
        }
    }
}

__global__ void __launch_bounds__(kBoxThreads)
boxMaps(const uint8_t *frames, int h, int w, int W, int64_t boxSize, int32_t *box, int32_t *box2, int2 *screen,
        float *sigma)
{
    boxMapsBlock(frames, h, w, W, boxSize, box, box2, screen, sigma, blockIdx.x, blockIdx.y, blockIdx.z);
}

// the box maps of several clips in one launch: block b belongs to the job j with first[j] <= b < first[j + 1]
__global__ void __launch_bounds__(kBoxThreads) boxMapsJobs(const BoxJob *__restrict__ jobs, int njobs)
{
    const int b = blockIdx.x;
    int j = 0;
    while (j + 1 < njobs && b >= jobs[j + 1].first) j++;
    const BoxJob &jb = jobs[j];
    const int bw = jb.w - jb.W + 1, bh = jb.h - jb.W + 1, tx = kBoxThreads - jb.W + 1;
    const int gx = (bw + tx - 1) / tx, gy = (bh + kBoxRows - 1) / kBoxRows;
    const int local = b - jb.first, bx = local % gx, rest = local / gx, by = rest % gy, f = rest / gy;
    boxMapsBlock(jb.frames, jb.h, jb.w, jb.W, int64_t(bh) * bw, jb.box, jb.box2, jb.screen, jb.sigma, bx, by, f);
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

int boxJobBlocks(int n, int h, int w, int W)
{
    const int bh = h - W + 1, bw = w - W + 1, tx = kBoxThreads - W + 1;
    if (n <= 0 || bh <= 0 || bw <= 0) return 0;
    return ((bw + tx - 1) / tx) * ((bh + kBoxRows - 1) / kBoxRows) * n;
}

void launchBoxJobs(const BoxJob *jobs, int njobs, int blocks, hipStream_t stream)
{
    if (njobs <= 0 || blocks <= 0) return;
    hipLaunchKernelGGL(boxMapsJobs, dim3(blocks), dim3(kBoxThreads), 0, stream, jobs, njobs);
    FC_HIP_CHECK(hipGetLastError());
}

void launchBoxSums(const uint8_t *frames, int n, int h, int w, int W, int32_t *box, int32_t *box2, int2 *screen,
                   float *sigma, hipStream_t stream)
{
    if (n <= 0) return;
    const int bh = h - W + 1, bw = w - W + 1, tx = kBoxThreads - W + 1;
    if (bh <= 0 || bw <= 0 || W > 181) throw std::invalid_argument("framecor: box sums need 1 <= W <= 181 and W <= h, w");
    const int64_t boxSize = int64_t(bh) * bw;
    hipLaunchKernelGGL(boxMaps, dim3((bw + tx - 1) / tx, (bh + kBoxRows - 1) / kBoxRows, n), dim3(kBoxThreads), 0,
                       stream, frames, h, w, W, boxSize, box, box2, screen, sigma);
    FC_HIP_CHECK(hipGetLastError());
}

} // namespace framecor
