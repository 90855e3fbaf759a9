// kernels.h -- host-side launchers of the framecor kernels
#pragma once

#include "common.h"

namespace framecor {

// ---- summed-area tables and box sums (sat.hip)

// the (h + 1) x (w + 1) int64 tables of n frames (n x h x w uint8), pixels and squares
void launchSummedArea(const uint8_t *frames, int n, int h, int w, int64_t *sat, int64_t *sat2, hipStream_t stream);
// the W x W box maps of n frames (n x h x w uint8), n x (h - W + 1) x (w - W + 1) each, straight from the
// pixels; each output may be null: box and box2, the int32 sums of the pixels and of their squares;
// screen, {box, 1 / sqrt(W^2 box2 - box^2) as float bits (0 where that is not positive)}; sigma, the
// float sqrtf(float(W^2 box2 - box^2) / float(W^2)) (0 where that is not positive)
void launchBoxSums(const uint8_t *frames, int n, int h, int w, int W, int32_t *box, int32_t *box2, int2 *screen,
                   float *sigma, hipStream_t stream);

// the box maps of n frames of one clip (as launchBoxSums) as one job of a launch over several clips;
// first: the job's first block in the launch (the jobs in order, each boxJobBlocks(n, h, w, W) blocks)
struct BoxJob {
    const uint8_t *frames;
    int32_t *box, *box2;
    int2 *screen;
    float *sigma;
    int n, h, w, W;
    int first;
};
int boxJobBlocks(int n, int h, int w, int W);
// jobs: a device array of njobs jobs; blocks: the sum of their blocks
void launchBoxJobs(const BoxJob *jobs, int njobs, int blocks, hipStream_t stream);

// ---- the int8-MFMA form (int8_corr.hip)

struct Int8Args {
    Geometry g;
    WindowData wd;
    FrameGroup fg;
    const uint8_t *tmpl;   // nwin x W x W, the reference windows
    int nut;               // 16-row bands of the surface: the wavefronts of a block
    int ni4;               // template row groups of 4
    int tstride;           // LDS bytes per template row
    int cstride;           // LDS bytes per chip row
    int crows;             // LDS chip rows
    int fpb, nfb;          // frames per block, frame blocks
    int alias;             // the one-frame kernel: the raw surface shares the chip's LDS
};

// The LDS of the one-frame-at-a-time int8 kernel (byte offsets): the template rows, a zero block of
// 16 njb bytes, the chip, the SNR box (min(stat, M)^2 int32) and one result per wave
struct Int8Lds {
    int zero, chip, box, parts, total;
};

__host__ __device__ inline Int8Lds int8Lds(int ni4, int tstride, int crows, int cstride, int njb, int M, int stat, int nw)
{
    Int8Lds l;
    l.zero = 4 * ni4 * tstride;
    l.chip = l.zero + ((16 * njb + 15) & ~15);
    l.box = l.chip + ((crows * cstride + 15) & ~15);
    const int nn = stat < M ? stat : M;
    l.parts = l.box + ((4 * nn * nn + 15) & ~15);
    l.total = l.parts + 16 * nw;
    return l;
}

struct Int8Layout {
    int W;                // the template side
    int njb, nvt, nut, ni4, tstride, cstride, crows;
    size_t shmem;         // the one-frame-at-a-time kernel
    bool alias;           // (unused: the one-frame kernel keeps no full surface)
    bool supported;
    bool edge;            // M = 16 c + 1: c core bands plus a wave for the edge row and one for the edge column
    int edgeTstride, edgeCrows;   // the one-frame kernel with the edge waves (variant simple): its LDS
    size_t edgeShmem;
    size_t pipedShmem;    // the pipelined producer/helper kernel
    bool pipedOk;
    // the 32 x 32 kernel (int8_wide.hip) where M = 32 nb + 1: nb bands, nj template blocks, its LDS
    bool wideOk;
    int nb, nj, wideTstride, wideCstride, wideCrows;
    size_t wideShmem;
    // the four-wave pipelined kernel (int8_quad.hip) where M = 16 c + 1, c = 4 or 6, and W = 32 or 64: the
    // LDS of its block
    bool quadOk;
    int quadTstride, quadCrows;
    size_t quadShmem;
    // one wave per window where W <= 64 and M = 16 c + 1 <= 65 (c bands, the edge row and column) or
    // M <= 48 (full bands): its chip rows and the LDS of its block
    bool soloOk;
    int soloCrows;
    size_t soloShmem;
};

// Auto: Solo where it applies and W <= 33, else Quad where it applies, else Simple; Simple: the one-frame-at-a-time kernel (two blocks per CU),
// with the edge row and the edge column in a wave each where M = 16 c + 1; Piped: the producer/helper pipeline where its buffers fit
// in LDS; Padded: the one-frame kernel with ceil(M / 16) full bands (no edge waves); Wide: the
// one-frame kernel on 32 x 32 tiles (v_mfma_i32_32x32x32_i8) where M = 32 c + 1, c <= 4 (else as Auto);
// Quad: blocks of four waves with equal MFMA shares, software-pipelined over frames, where M = 16 c + 1,
// c = 4 or 6, and W = 32 or 64 (else as Auto); Solo: one wave per window and frame where W <= 64 and
// M = 16 c + 1 <= 65 (c bands, the edge row and column) or M <= 48 (full bands) (else as Simple)
enum class Int8Variant { Auto = 0, Simple = 1, Piped = 2, Padded = 3, Wide = 4, Quad = 5, Solo = 6 };

Int8Layout int8Layout(const Geometry &g);
void launchInt8(Int8Args a, const Int8Layout &l, Int8Variant variant, hipStream_t stream);

// one int8 kernel: its entry, block size, dynamic LDS and the variant it is (Auto resolved)
struct Int8Kernel {
    const void *func;
    int threads;
    size_t shmem;
    Int8Variant variant;
};

// the kernel of a tile shape (l) and variant; queue: the persistent work-queue form
Int8Kernel int8Kernel(const Int8Layout &l, Int8Variant variant, bool queue);
// the layout fields of the arguments (nut, ni4, strides, alias) for the kernel variant (resolved)
void fillInt8Args(Int8Args &a, const Int8Layout &l, Int8Variant variant);
// the 32 x 32 kernel of nb bands and nj (njb) template blocks; null where none is instantiated
const void *int8WideKernel(int nb, int nj, int njb, bool queue);
// the four-wave kernel of c core bands, a template of W columns in njb blocks; null where none is
// instantiated; its block size and LDS
const void *int8QuadKernel(int c, int W, int njb, bool queue);
int int8QuadThreads();
size_t int8QuadShmem(int ni4, int tstride, int crows, int cstride, int njb, const Geometry &g);
// the persistent form over a work queue; clips: device array of the clips' arguments (layout filled)
void launchInt8Queue(const Int8Kernel &k, const WorkQueue &q, const Int8Args *clips, int blocks, hipStream_t stream);

// ---- exact NCC surfaces and the second peak (int8_corr.hip)

struct SurfaceArgs {
    Geometry g;
    WindowData wd;
    const uint8_t *frame;        // h x w: each window is correlated with its search window here
    const int *gross;            // nwin x 2 integer offsets (down, across) of the search windows, or null
    const uint8_t *tmpl;         // nwin x W x W, the reference windows
    const int32_t *box, *box2;   // bh x bw box maps of the frame
    double *surf;                // count x M x M, the exact FP64 NCC surfaces
    int k0, count;               // windows k0 .. k0 + count - 1
    // set by the launcher: template column blocks and row groups, LDS strides and rows, tiles per side
    int njb, ni4, tstride, cstride, crows, nt;
};

// the LDS of the surface kernel (template rows and the chip); above 160 KB it is not supported
size_t surfaceShmem(const Geometry &g);
void launchNccSurfaces(SurfaceArgs s, hipStream_t stream);
// For i < count: out[k0 + i] = the largest value of surf[i] (M x M) at a local maximum (a point equal to
// the largest value of its 5 x 5 neighbourhood clipped to the surface) outside the square of rows and
// columns lo .. hi - 1; 0 where there is none; 1 where window k0 + i's template is flat (tvar 0)
void launchSecondPeak(const double *surf, int count, int M, int lo, int hi, const int64_t *tvar, int k0, double *out,
                      hipStream_t stream);
// For i < count, window k = k0 + i: the first maximum (row-major) (u, v) of surf[i], each coordinate
// refined by the parabola through it and its two neighbours (0 at the border or where the parabola is
// flat, |den| < 1e-12), relative to the centre: offsets[2k] = u + du - (M - 1) / 2, offsets[2k + 1] =
// v + dv - (M - 1) / 2, peak[k] = the maximum; NaN, NaN, 0 for a flat template or a surface of NaN
void launchParabolaPeaks(const double *surf, int count, int M, const int64_t *tvar, int k0, double *offsets,
                         double *peak, hipStream_t stream);

// ---- the FP32 FFT form (fft_corr.hip)

constexpr int kMaxFactors = 12;

struct FftPlan {
    int nf;                  // number of radix passes
    int fac[kMaxFactors];    // radices, in the order of the forward (decimation in frequency) passes
};

struct FftArgs {
    Geometry g;
    WindowData wd;
    FrameGroup fg;
    const float2 *spectra;   // nwin x np x rs: the spectra of the reference windows (kernel layout)
    const float2 *twiddle;   // L: exp(-2 pi i k / L)
    const float *templ;      // nwin x W x W zero-mean reference windows (set_reference only)
    float2 *spectraOut;      // set_reference only
    FftPlan plan;
    int np;                  // row pairs: (L + 1) / 2
    int H;                   // L / 2 + 1, the half-spectrum width
    int rs;                  // complex slots per row pair: 2 H
    int fpb, nfb;            // frames per block, frame blocks
    // search windows whose half spectrum does not fit in LDS: each block's spectrum in a workspace in
    // global memory (np x rs complex per block, block b at work + b * workStride), and a persistent grid
    // of workBlocks blocks; null when the spectrum is held in LDS
    float2 *work;
    size_t workStride;
    int workBlocks;
};

// the largest search window of the FFT form (a line of a transform pass fits in one round of a block)
constexpr int kFftMaxL = 2048;
// the FFT form's plan of L (radices 8, 4, 2, 5, 3 and then any prime; L <= kFftMaxL), false if none
bool fftPlanFor(int L, FftPlan &plan);
// whether the half spectrum of L is held in LDS (else in a global workspace)
bool fftInLds(int L);
// the LDS of a block: the half spectrum when it fits, plus the twiddle and position tables
size_t fftShmem(int L);
// the workspace of one block when the spectrum is not in LDS (bytes; 0 if it is), and the persistent grid
// on a device of cus compute units
size_t fftWorkPerBlock(int L);
int fftWorkBlocks(int L, int cus);
// the spectra of the reference windows (args: g, wd, twiddle, templ, spectraOut, plan, np, H, rs)
void launchFftReference(const FftArgs &a, hipStream_t stream);
// correlate a group of frames
void launchFftCorrelate(FftArgs a, hipStream_t stream);

// the register-resident form of the FFT correlation for L = 64, 96, 128, 160, 192 (fft_fast.hip); its
// reference spectra have their own layout (the same size, nwin x (L / 2 + 1) x L complex)
bool fftFastFor(int L);
size_t fftFastShmem(int L);
void launchFftFastReference(const FftArgs &a, hipStream_t stream);
void launchFftFastCorrelate(const FftArgs &a, hipStream_t stream);

// the persistent work-queue kernel of the FFT form (search-window size L: its LDS)
struct FftKernel {
    const void *func;
    int threads;
    size_t shmem;
};
FftKernel fftQueueKernel(int L);
FftKernel fftFastQueueKernel(int L);
// clips: device array of the clips' arguments (np, H, rs, plan, spectra, twiddle filled)
void launchFftQueue(const FftKernel &k, const WorkQueue &q, const FftArgs *clips, int blocks, hipStream_t stream);

// ---- the differentiable form (diff.hip)

struct DiffArgs {
    Geometry g;
    WindowData wd;
    const uint8_t *tmpl;      // nwin x W x W, the reference windows
    const float *frames;      // n x h x w float32
    int n;
    const int *gross;         // n x nwin x 2, or null
    const float *shifts;      // n x nwin x 2: (sy, sx), continuous, relative to the window + gross
    float *ncc;               // n x nwin
    const float *gradNcc;     // backward: n x nwin
    float *gradFrames;        // backward: n x h x w, accumulated
    float *gradShifts;        // backward: n x nwin x 2
    float *shiftsOut;         // refine: n x nwin x 2
    int steps;                // refine: Gauss-Newton steps
};

enum class DiffOp { Forward, Backward, Refine };

void launchDiff(DiffArgs d, DiffOp op, hipStream_t stream);

} // namespace framecor
