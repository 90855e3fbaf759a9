// fft_corr.hip -- the FP32 FFT form: one block per (window, run of frames), the half spectrum in LDS
//
// Per search window (cxx/cpu.cpp:117-167 of the reference, as one fused kernel): the L x L chip minus
// its mean (its exact pixel sum, taken while it is loaded) is loaded into LDS as packed row pairs; forward row passes,
// the split into half spectra, forward column passes; the product with the conjugate of the cached
// spectrum of the reference window (read per frame, coalesced: from HBM for the first frame of a
// run, from L2 for the others); inverse column passes, the combine, inverse row passes over the M
// surface rows only; then, in LDS, the
// normalization of the M x M surface with the box sums (cxx/surface.h:30-35: c / L^2 over
// norm * sqrt(v / n)), the first maximum, and the reference's FP32 quadratic fit, FP64 SNR and
// covariance. HBM traffic per window and frame: the chip, the reference spectrum, the box sums of the
// surface and the outputs; nothing else leaves LDS.

#include <algorithm>

#include "common.h"
#include "epilogue.h"
#include "fft_phases.h"
#include "kernels.h"

namespace framecor {

namespace {

constexpr int kFftThreads = 512;
constexpr int kIpt = 4;   // items per thread and round of the split and combine steps

struct BlockSync {
    __device__ void operator()() const { __syncthreads(); }
};

struct FftShared {
    float2 *Z;      // np x rs
    float2 *tw;     // L
    int *pos;       // L
    float *redv;    // waves
    int *redi;      // waves
    double *redd;   // waves
    int *sums;      // waves: the chip sum
};

// the block's buffers: the half spectrum Z in LDS, or (zg not null) in the block's global workspace,
// then the tables and reductions in LDS
__device__ inline FftShared carve(unsigned char *lds, const FftArgs &a, float2 *zg)
{
    FftShared s;
    float2 *z = reinterpret_cast<float2 *>(lds);
    s.Z = zg ? zg : z;
    s.tw = zg ? z : z + size_t(a.np) * a.rs;
    s.redd = reinterpret_cast<double *>(s.tw + a.g.L);
    s.pos = reinterpret_cast<int *>(s.redd + kFftThreads / kWave);
    s.redv = reinterpret_cast<float *>(s.pos + a.g.L);
    s.redi = reinterpret_cast<int *>(s.redv + kFftThreads / kWave);
    s.sums = s.redi + kFftThreads / kWave;
    return s;
}

__device__ inline void loadTables(const FftShared &s, const FftArgs &a, int tid)
{
    const int L = a.g.L;
    for (int k = tid; k < L; k += kFftThreads) {
        s.tw[k] = a.twiddle[k];
        s.pos[k] = digitReversed(k, a.plan.fac, a.plan.nf, L);
    }
}

// forward row passes, split, forward column passes of the packed data in s.Z
__device__ inline void forward2d(const FftShared &s, const FftArgs &a, int tid)
{
    const int L = a.g.L;
    BlockSync sync;
    transformLines<false>(s.Z, s.tw, L, a.plan.fac, a.plan.nf, a.np, RowAddr{a.rs}, tid, kFftThreads, sync);
    splitPairs<kIpt>(s.Z, s.pos, L, a.H, a.rs, a.np, tid, kFftThreads, sync);
    transformLines<false>(s.Z, s.tw, L, a.plan.fac, a.plan.nf, a.H, ColAddr{a.rs, a.H}, tid, kFftThreads, sync);
}

// the spectra of the reference windows: one block per window (a spectrum that does not fit in LDS is
// transformed in place in the output)
__global__ void __launch_bounds__(kFftThreads) fftReference(FftArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const int tid = threadIdx.x, k = blockIdx.x;
    float2 *out = a.spectraOut + size_t(k) * a.np * a.rs;
    const FftShared s = carve(lds, a, a.work ? out : nullptr);
    const int W = a.g.W;
    loadTables(s, a, tid);
    // zero-padded template, packed: row y -> half (y & 1) of pair y >> 1
    float *zf = reinterpret_cast<float *>(s.Z);
    for (int i = tid; i < a.np * a.rs * 2; i += kFftThreads) zf[i] = 0.0f;
    __syncthreads();
    const float *T = a.templ + size_t(k) * W * W;
    for (int idx = tid; idx < W * W; idx += kFftThreads) {
        const int y = idx / W, x = idx - y * W;
        zf[2 * ((y >> 1) * a.rs + x) + (y & 1)] = T[idx];
    }
    __syncthreads();
    forward2d(s, a, tid);
    if (!a.work)
        for (int i = tid; i < a.np * a.rs; i += kFftThreads) out[i] = s.Z[i];
}

// window k against frames fBegin .. fEnd - 1, by the whole block; zg: the block's global workspace for
// the half spectrum, or null (in LDS)
__device__ __attribute__((always_inline)) inline void fftWindow(const FftArgs &a, int k, int fBegin, int fEnd,
                                                                unsigned char *lds, float2 *zg)
{
    const FftShared s = carve(lds, a, zg);
    const Geometry &g = a.g;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid >> 6;
    const int L = g.L, M = g.M, W = g.W;
    BlockSync sync;

    loadTables(s, a, tid);
    const float normk = a.wd.norm[k];
    const int y0 = a.wd.down[k / g.nx], x0 = a.wd.across[k % g.nx];
    const float2 *ref = a.spectra + size_t(k) * a.np * a.rs;
    float *zf = reinterpret_cast<float *>(s.Z);

    for (int f = fBegin; f < fEnd; f++) {
        const int64_t q = int64_t(f) * g.nwin + k;
        int gy = 0, gx = 0;
        if (a.fg.gross) {
            gy = a.fg.gross[2 * q];
            gx = a.fg.gross[2 * q + 1];
        }
        int sy, sx;
        if (!(normk > 0) || !searchWindow(g, y0, x0, gy, gx, sy, sx)) {
            if (tid == 0) invalidateOut(q, a.fg.offset, a.fg.peak, a.fg.snr, a.fg.cov);
            continue;
        }
        // the chip minus its mean (cxx/cpu.cpp:130-137: mean = float(box sum over L x L) / float(L^2),
        // buf = px - mean). The pixels go to LDS as floats (exact) while the block sums them exactly in
        // int32; then each becomes px - mean, the one FP32 rounding of the reference
        __syncthreads();   // the previous frame's epilogue read the surface
        const uint8_t *src = a.fg.frames + size_t(f) * g.pixels + size_t(sy) * g.w + sx;
        int part = 0;
        for (int idx = tid; idx < L * L; idx += kFftThreads) {
            const int y = idx / L, x = idx - y * L;
            const int px = src[size_t(y) * g.w + x];
            part += px;
            zf[2 * ((y >> 1) * a.rs + x) + (y & 1)] = float(px);
        }
        if (L & 1)   // the imaginary half of the last pair: zero
            for (int x = tid; x < L; x += kFftThreads) zf[2 * ((a.np - 1) * a.rs + x) + 1] = 0.0f;
        part = waveSumI(part);
        if (lane == 0) s.sums[wave] = part;
        __syncthreads();
        int boxL = 0;
        for (int t = 0; t < kFftThreads / kWave; t++) boxL += s.sums[t];
        const float mean = float(boxL) / (float(L) * float(L));
        for (int idx = tid; idx < L * L; idx += kFftThreads) {
            const int y = idx / L, x = idx - y * L;
            float &z = zf[2 * ((y >> 1) * a.rs + x) + (y & 1)];
            z = z - mean;
        }
        __syncthreads();

        forward2d(s, a, tid);
        // times the conjugate spectrum of the reference window
        for (int i = tid; i < a.np * a.rs; i += kFftThreads) s.Z[i] = cmulc(s.Z[i], ref[i]);
        __syncthreads();
        transformLines<true>(s.Z, s.tw, L, a.plan.fac, a.plan.nf, a.H, ColAddr{a.rs, a.H}, tid, kFftThreads, sync);
        const int npOut = (M + 1) / 2;
        combinePairs<kIpt>(s.Z, s.pos, L, a.H, a.rs, npOut, tid, kFftThreads, sync);
        transformLines<true>(s.Z, s.tw, L, a.plan.fac, a.plan.nf, npOut, RowAddr{a.rs}, tid, kFftThreads, sync);

        // normalize in place and find the first maximum: c / L^2 over norm * sqrtf(float(var) / float(n)),
        // the second factor from the frame's sigma map (0 where var <= 0)
        const float scale = 1.0f / (float(L) * float(L));
        const int n = W * W;
        const float *sig = a.fg.sigma + f * g.boxSize + int64_t(sy) * g.bw + sx;
        float best = -2.0f;
        int ib = 0x7fffffff;
        for (int idx = tid; idx < M * M; idx += kFftThreads) {
            const int u = idx / M, v = idx - u * M;
            float &c = zf[2 * ((u >> 1) * a.rs + v) + (u & 1)];
            const float sg = sig[int64_t(u) * g.bw + v];
            float val = 0.0f;
            {
#pragma clang fp contract(off)
                const float cc = c * scale;
                if (sg > 0 && normk > 0) val = cc / (normk * sg);
            }
            c = val;
            bestOf(best, ib, val, idx);
        }
        waveBest(best, ib);
        if (lane == 0) {
            s.redv[wave] = best;
            s.redi[wave] = ib;
        }
        __syncthreads();
        best = -2.0f;
        ib = 0x7fffffff;
        for (int t = 0; t < kFftThreads / kWave; t++) bestOf(best, ib, s.redv[t], s.redi[t]);
        if (!(best > 0)) {   // uniform
            if (tid == 0) invalidateOut(q, a.fg.offset, a.fg.peak, a.fg.snr, a.fg.cov);
            continue;
        }
        const int ub = ib / M, vb = ib - ub * M;
        auto at = [&](int u, int v) { return zf[2 * ((u >> 1) * a.rs + v) + (u & 1)]; };
        // the stat-box sum of squares (FP64)
        const int nn = g.stat < M ? g.stat : M;
        const int ulo = max(ub - nn / 2, 0), uhi = min(ub - nn / 2 + nn, M);
        const int vlo = max(vb - nn / 2, 0), vhi = min(vb - nn / 2 + nn, M);
        const int bwid = vhi - vlo, items = (uhi - ulo) * bwid;
        double sum = 0.0;
        for (int it = tid; it < items; it += kFftThreads) {
            const int u = ulo + it / bwid, v = vlo + it % bwid;
            const double c = at(u, v);
            sum += c * c;
        }
        sum = waveSumD(sum);
        if (lane == 0) s.redd[wave] = sum;
        __syncthreads();
        if (tid == 0) {
            double total = 0.0;
            for (int t = 0; t < kFftThreads / kWave; t++) total += s.redd[t];
            const bool inside = ub > 0 && ub < M - 1 && vb > 0 && vb < M - 1;
            float fs[3][3];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) fs[i][j] = inside ? at(ub + i - 1, vb + j - 1) : 0.0f;
            fs[1][1] = at(ub, vb);
            writePeakF32(fs, inside, ub, vb, g.R, gy, gx, total, items, double(n), q, a.fg.offset, a.fg.peak,
                         a.fg.snr, a.fg.cov);
        }
    }
}

__global__ void __launch_bounds__(kFftThreads) fftCorrelate(FftArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const int logical = xcdRemap(blockIdx.x, gridDim.x);
    const int k = logical / a.nfb, fblock = logical - k * a.nfb;
    const int fBegin = fblock * a.fpb, fEnd = min(fBegin + a.fpb, a.fg.n);
    fftWindow(a, k, fBegin, fEnd, lds, nullptr);
}

// the same for search windows whose spectrum is in the global workspace: a persistent grid (one
// workspace per block) over the (window, run of frames) items, in the order of the blocks above
__global__ void __launch_bounds__(kFftThreads) fftCorrelateGlobal(FftArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    float2 *zg = a.work + size_t(blockIdx.x) * a.workStride;
    const int items = a.g.nwin * a.nfb;
    for (int it = blockIdx.x; it < items; it += gridDim.x) {
        const int k = it / a.nfb, fblock = it - k * a.nfb;
        const int fBegin = fblock * a.fpb, fEnd = min(fBegin + a.fpb, a.fg.n);
        fftWindow(a, k, fBegin, fEnd, lds, zg);
        __syncthreads();   // the item's epilogue read the workspace
    }
}

// the persistent form over the work queue of one search-window size (every clip of a batch with this
// L; clips[c]: clip c's arguments), as int8Queue
__global__ void __launch_bounds__(kFftThreads) fftQueue(WorkQueue q, const FftArgs *__restrict__ clips)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int item[2];
    for (int round = 0;; round++) {
        if (threadIdx.x == 0) item[round & 1] = q.next(blockIdx.x);
        __syncthreads();   // also: the previous item is over (its LDS may be rewritten)
        const int it = __builtin_amdgcn_readfirstlane(item[round & 1]);
        if (it < 0) break;
        int c, k, fBegin, fEnd;
        q.locate(it, c, k, fBegin, fEnd);
        const FftArgs &a = clips[c];
        fftWindow(a, k, fBegin, fEnd, lds, a.work ? a.work + size_t(blockIdx.x) * a.workStride : nullptr);
    }
}

void configure(const void *kern)
{
    (void)hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipGetLastError();
}

} // namespace

bool fftPlanFor(int L, FftPlan &plan)
{
    plan.nf = 0;
    if (L < 2 || L > kFftMaxL) return false;
    int m = L;
    // radix 8 first (fewest passes), then 4, 2, 5, 3, then the remaining primes (direct passes)
    static const int first[5] = {8, 4, 2, 5, 3};
    for (int r : first)
        while (m % r == 0) {
            if (plan.nf == kMaxFactors) return false;
            plan.fac[plan.nf++] = r;
            m /= r;
        }
    for (int r = 7; m > 1; r += 2)
        while (m % r == 0) {
            if (plan.nf == kMaxFactors) return false;
            plan.fac[plan.nf++] = r;
            m /= r;
        }
    return true;
}

namespace {

size_t fftTables(int L)
{
    return sizeof(float2) * L + sizeof(double) * (kFftThreads / kWave) + sizeof(int) * L
        + (sizeof(float) + 2 * sizeof(int)) * (kFftThreads / kWave);
}

size_t fftSpectrum(int L)
{
    const size_t np = (L + 1) / 2, rs = 2 * (L / 2 + 1);
    return sizeof(float2) * np * rs;
}

} // namespace

bool fftInLds(int L) { return fftSpectrum(L) + fftTables(L) <= 160 * 1024; }

size_t fftShmem(int L) { return fftTables(L) + (fftInLds(L) ? fftSpectrum(L) : 0); }

size_t fftWorkPerBlock(int L) { return fftInLds(L) ? 0 : fftSpectrum(L); }

int fftWorkBlocks(int L, int cus)
{
    // two blocks per CU, within 512 MiB of workspace
    const size_t per = std::max<size_t>(fftSpectrum(L), 1);
    return int(std::max<size_t>(1, std::min<size_t>(size_t(2) * cus, (size_t(512) << 20) / per)));
}

void launchFftReference(const FftArgs &a, hipStream_t stream)
{
    static bool configured = false;
    if (!configured) {
        configure(reinterpret_cast<const void *>(fftReference));
        configured = true;
    }
    hipLaunchKernelGGL(fftReference, dim3(a.g.nwin), dim3(kFftThreads), fftShmem(a.g.L), stream, a);
    FC_HIP_CHECK(hipGetLastError());
}

void launchFftCorrelate(FftArgs a, hipStream_t stream)
{
    if (a.fg.n <= 0) return;
    static bool configured = false;
    if (!configured) {
        configure(reinterpret_cast<const void *>(fftCorrelate));
        configured = true;
    }
    // runs of at most 16 frames per block (at 720p, window 64, search 48: 6 % faster than one frame
    // per block), and enough blocks to fill the GPU several times over; consecutive logical blocks are
    // the runs of one window, so an XCD's blocks share that window's spectrum in its L2
    const int target = 2048, maxFpb = 16;
    int nfb = std::max((a.fg.n + maxFpb - 1) / maxFpb, (target + a.g.nwin - 1) / a.g.nwin);
    if (nfb > a.fg.n) nfb = a.fg.n;
    a.fpb = (a.fg.n + nfb - 1) / nfb;
    a.nfb = (a.fg.n + a.fpb - 1) / a.fpb;
    const int64_t blocks = int64_t(a.g.nwin) * a.nfb;
    if (blocks > 0x7fffffff) throw std::invalid_argument("framecor: too many windows x frames in one group");
    if (a.work) {
        static bool configuredGlobal = false;
        if (!configuredGlobal) {
            configure(reinterpret_cast<const void *>(fftCorrelateGlobal));
            configuredGlobal = true;
        }
        const int grid = int(std::min<int64_t>(blocks, a.workBlocks));
        hipLaunchKernelGGL(fftCorrelateGlobal, dim3(unsigned(grid)), dim3(kFftThreads), fftShmem(a.g.L), stream, a);
    }
    else {
        hipLaunchKernelGGL(fftCorrelate, dim3(unsigned(blocks)), dim3(kFftThreads), fftShmem(a.g.L), stream, a);
    }
    FC_HIP_CHECK(hipGetLastError());
}

FftKernel fftQueueKernel(int L)
{
    FftKernel k{};
    k.func = reinterpret_cast<const void *>(fftQueue);
    k.threads = kFftThreads;
    k.shmem = fftShmem(L);
    allowLargeLds(k.func);
    return k;
}

void launchFftQueue(const FftKernel &k, const WorkQueue &q, const FftArgs *clips, int blocks, hipStream_t stream)
{
    WorkQueue qa = q;
    const FftArgs *c = clips;
    void *args[] = {&qa, &c};
    FC_HIP_CHECK(hipLaunchKernel(k.func, dim3(blocks), dim3(k.threads), args, k.shmem, stream));
}

} // namespace framecor
