// int8_corr.hip -- the int8-MFMA form: exact integer correlation on CDNA4 matrix cores, fused with
// the normalization, an exact argmax and the FP64 peak fit
//
// For window k and frame f the raw surface is r[u][v] = sum_{i,j<W} (T[i][j]-128)(F[u+i][v+j]-128),
// with T the reference window and F the search window (chip) of the frame, pixels offset by -128 to
// int8 and summed in int32 (|r| <= 16384 W^2 < 2^31 for W <= 128). The offset is undone exactly with
// the box sums of the frame and the template sums:
//   sum(FT) = r + 128 (sum F + sum T) - 16384 W^2,
//   ncc     = (W^2 sum(FT) - sum T sum F) / sqrt((W^2 sum T^2 - (sum T)^2) (W^2 sum F^2 - (sum F)^2)),
// where the numerator and both variances are exact integers (and exact in FP64).
//
// The contraction as a GEMM. For a block of 16 surface rows (u = 16 t + m) and 16 columns
// (v = 16 v1 + n), K runs over template rows i and 16-byte column blocks jb of a Toeplitz expansion
// of the template row (the template shifted by n):
//   A[m][(i, jb, kk)] = F[16 t + m + i][16 (v1 + jb) + kk]      (16 contiguous chip bytes)
//   B[(i, jb, kk)][n] = T[i][16 jb + kk - n]                     (zero outside 0..W-1)
//   C[m][n] = sum over i, jb, kk  -> r[16 t + m][16 v1 + n].
// One v_mfma_i32_16x16x64_i8 covers four template rows (one per 16-lane group of the wavefront) and
// one jb. A chip tile (16 rows x 16 bytes at column block cb) serves every (v1, jb) with v1 + jb = cb,
// so each LDS read of A feeds up to NJB MFMAs; the B fragments of four template rows are built once
// per row group from the LDS template with byte funnel shifts and reused across the v tiles.
//
// One block per (window, run of frames); one wavefront per 16-row band of the surface (nut waves),
// each over all v tiles (NVT) and all template rows. Per frame: the chip is staged in LDS as int8;
// each wave accumulates its band in registers; the epilogue screens every point with an FP32 NCC,
// takes the exact FP64 NCC of the points within FP32 error of the block's maximum when there are
// several (so the argmax is that of the exact FP64 surface, first maximum in row-major order on ties),
// and evaluates the SNR, the 3 x 3 FP64 fit and the covariance on exact FP64 values
// (cxx/surface.h:71-146 formulas).

#include "common.h"
#include "int8_device.h"

namespace framecor {

namespace {

// Window k against frames fBegin .. fEnd - 1, by the whole block (NVT + EW waves).
// EW = 0: ceil(M / 16) = NVT full bands; EW = 2 (M = 16 NVT + 1): NVT core bands plus a wave for the edge
// row and one for the edge column.
//
// Per frame: every wave runs its MFMAs (a band, the edge row or the edge column) into registers and
// screens its own points there (FP32 screen from the frame's screen map) into parts[wave]; barrier B2;
// every wave takes the block's peak from parts (blockPeak: exact FP64 values only at near-ties) and
// writes its raw values inside the SNR box around it to LDS, and the chip of the next valid frame is
// staged (every MFMA of this frame is over); barrier B3; then the
// finisher wave (the edge-row wave, whose MFMA share is the smallest, or wave 0) writes the outputs
// while the other waves start the next frame's MFMAs. LDS: the template, a zero block, the chip, the
// SNR box (min(stat, M)^2 int32) and the waves' results -- no full surface.
template <int NVT, int NJB, int EW>
__device__ __attribute__((always_inline)) inline void int8Window(const Int8Args &a, int k, int fBegin, int fEnd,
                                                                 unsigned char *lds)
{
    const Geometry &g = a.g;
    const int tid = threadIdx.x, lane0 = tid & (kWave - 1), wave = tid >> 6, nthreads = blockDim.x;
    const int nw = nthreads / kWave;
    const int W = g.W, M = g.M, L = g.L;
    const Int8Lds sh = int8Lds(a.ni4, a.tstride, a.crows, a.cstride, NJB, M, g.stat, nw);
    int8_t *tmpl = reinterpret_cast<int8_t *>(lds);
    const int8_t *zero = tmpl + sh.zero;
    int8_t *chip = tmpl + sh.chip;
    int *boxbuf = reinterpret_cast<int *>(lds + sh.box);
    WaveBest *parts = reinterpret_cast<WaveBest *>(lds + sh.parts);

    loadTemplate(lds, sh.box, tmpl, a.tmpl + size_t(k) * W * W, W, a.tstride, tid, nthreads);
    const float normk = a.wd.norm[k];
    const int y0 = a.wd.down[k / g.nx], x0 = a.wd.across[k % g.nx];
    Moments mom;
    mom.n = int64_t(W) * W;
    mom.tvar = a.wd.tvar[k];
    mom.nd = double(mom.n);
    mom.tau = double(a.wd.tsum[k] - 128 * mom.n);
    mom.rst = mom.tvar > 0 ? float(1.0 / sqrt(double(mom.tvar))) : 0.0f;
    const bool band = EW == 0 || wave < NVT;
    const int finisher = EW == 0 ? 0 : NVT;

    // the first valid frame (the invalid ones before it written as such), its chip
    int f = fBegin;
    FrameWindow fw;
    for (;; f++) {
        if (f >= fEnd) break;
        fw = frameWindow(a, k, f, y0, x0, normk);
        if (fw.valid) break;
        if (tid == 0) invalidateOut(fw.q, a.fg.offset, a.fg.peak, a.fg.snr, a.fg.cov);
    }
    __syncthreads();   // the template (and the zeroed areas)
    if (f < fEnd)
        loadChip(chip, a.fg.frames + size_t(f) * g.pixels + size_t(fw.sy) * g.w + fw.sx, g.w, L, a.cstride, tid,
                 nthreads);
    __syncthreads();   // B1

    while (f < fEnd) {
        // the lane's coordinates anew each frame: the per-point indices derived from them are recomputed
        // here (a few VALU) rather than hoisted out of the loop and kept live (spilled) across the MFMAs
        int lane = lane0;
        asm volatile("" : "+v"(lane));
        const int g4 = lane >> 4, m16 = lane & 15;
        // the MFMAs of this wave's share into registers
        v4i acc[NVT];
        v4i es = v4i{0, 0, 0, 0};
        if (band)
            bandMfma<NVT, NJB>(acc, chip, tmpl, wave, lane, a.ni4, a.tstride, a.cstride);
        else if (wave == NVT)
            es = edgeRowMfma<NJB>(chip, tmpl, NVT, lane, a.ni4, a.tstride, a.cstride);
        else
            es = edgeColMfma<NJB>(chip, tmpl, zero, NVT, W, lane, a.tstride, a.cstride);

        // this lane's points: fn(raw, u, v)
        auto visit = [&](auto &&fn) {
            if (band) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int u = 16 * wave + 4 * g4 + r;
#pragma unroll
                    for (int v1 = 0; v1 < NVT; v1++) {
                        const int v = 16 * v1 + m16;
                        if (EW != 0 || (u < M && v < M)) fn(acc[v1][r], u, v);
                    }
                }
            }
            else if (wave == NVT) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int blk = 4 * g4 + r, v = 16 * blk + m16;
                    if (blk <= NVT && v < M) fn(es[r], 16 * NVT, v);
                }
            }
            else {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int blk = 4 * g4 + r;
                    if (blk < NVT) fn(es[r], 16 * blk + m16, 16 * NVT);
                }
            }
        };

        // the screen of this wave's points: its largest screened value and whether it stands alone
        const int64_t boff = f * g.boxSize + int64_t(fw.sy) * g.bw + fw.sx;
        const int2 *scr = a.fg.screen + boff;
        const int32_t *box2 = a.fg.box2 + boff;
        Top2 t = top2Init();
        if (band) {
            // a row of NVT points per r: one row pointer, the columns at immediate offsets; up to three
            // tiles a band (search 16), the loads of all four rows issued first (at four tiles, the
            // registers cost more than the waits save: 3-4 % slower at window 32 / search 32)
            constexpr int RL = NVT <= 3 ? 4 : 1;   // rows whose loads are in flight together
#pragma unroll
            for (int r0 = 0; r0 < 4; r0 += RL) {
                int2 p[RL][NVT];
#pragma unroll
                for (int r = 0; r < RL; r++) {
                    const int u = 16 * wave + 4 * g4 + r0 + r;
                    const int2 *row = scr + u * g.bw + m16;
#pragma unroll
                    for (int v1 = 0; v1 < NVT; v1++)
                        p[r][v1] = (EW != 0 || (u < M && 16 * v1 + m16 < M)) ? row[16 * v1] : make_int2(0, 0);
                }
#pragma unroll
                for (int r = 0; r < RL; r++) {
                    const int u = 16 * wave + 4 * g4 + r0 + r;
#pragma unroll
                    for (int v1 = 0; v1 < NVT; v1++)
                        if (EW != 0 || (u < M && 16 * v1 + m16 < M))
                            top2Add(t, screenNcc(acc[v1][r0 + r], p[r][v1].x, __int_as_float(p[r][v1].y), mom),
                                    u * M + 16 * v1 + m16);
                }
            }
        }
        else
            visit([&](int raw, int u, int v) {
                const int2 p = scr[u * g.bw + v];
                top2Add(t, screenNcc(raw, p.x, __int_as_float(p.y), mom), u * M + v);
            });
        screenPart(t, parts + wave, lane);
        __syncthreads();   // B2: every MFMA of this frame is over; parts are in

        // the block's peak (exact FP64 values only at near-ties); this wave's raw values inside its SNR box
        int ib;
        const bool valid = blockPeak(parts, nw, wave, lane, mom, scr, box2, g.bw, M, visit, ib);
        if (valid) {
            const SnrBox b = snrBox(ib, M, g.stat);
            if (!band || (16 * wave < b.uhi && 16 * wave + 16 > b.ulo))
                visit([&](int raw, int u, int v) {
                    if (u >= b.ulo && u < b.uhi && v >= b.vlo && v < b.vhi) boxbuf[(u - b.ulo) * b.bwid + v - b.vlo] = raw;
                });
        }
        // the next valid frame and its chip
        const FrameWindow cur = fw;
        const int fc = f;
        for (f++;; f++) {
            if (f >= fEnd) break;
            fw = frameWindow(a, k, f, y0, x0, normk);
            if (fw.valid) break;
            if (tid == 0) invalidateOut(fw.q, a.fg.offset, a.fg.peak, a.fg.snr, a.fg.cov);
        }
        if (f < fEnd)
            loadChip(chip, a.fg.frames + size_t(f) * g.pixels + size_t(fw.sy) * g.w + fw.sx, g.w, L, a.cstride, tid,
                     nthreads);
        __syncthreads();   // B3: the SNR box and the next chip are in
        if (wave == finisher) finishPeak(a, mom, boxbuf, true, cur, fc, valid, ib, lane);
    }
}
// The register budget of a block of NVT + EW waves: blocks of more than 8 waves run one per CU (no
// limit below the 512 registers of a lone wave); a template of 112 or more columns (NJB >= 8) leaves
// LDS for at most three blocks per CU, so its waves take up to 168 registers (3 per SIMD) rather than
// spill at 128. Blocks of up to 6 waves with smaller templates take 6 waves per SIMD (80 registers:
// four blocks of 6, six of 4 per CU): their per-frame screen and finishing are latency-bound, and the
// extra resident blocks hide more of it than the spills of the epilogue cost (-12 % at window 64 and
// search 32, -20 % at 32 / 32 against 4 waves per SIMD; 7 or 8 per SIMD are slower). Blocks of 7 or 8
// waves keep 4 per SIMD (two blocks per CU, no spills): at 6 per SIMD they gain at most 4 % while their
// spills move 5x the kernel's HBM bytes through scratch. The persistent work-queue form keeps 4 per
// SIMD throughout: its loop over items holds more state across the frames, and there the spills cost
// more than the extra blocks save (a 16-clip batch: 190 ms against 193 ms).
constexpr int int8WavesPerEu(int nvt, int njb, int ew)
{
    return nvt + ew > 8 ? 1 : njb >= 8 ? 3 : nvt + ew <= 6 ? 6 : 4;
}
constexpr int int8QueueWavesPerEu(int nvt, int njb, int ew) { return nvt + ew > 8 ? 1 : njb >= 8 ? 3 : 4; }

template <int NVT, int NJB, int EW>
__global__ void __launch_bounds__(64 * (NVT + EW))
__attribute__((amdgpu_waves_per_eu(int8WavesPerEu(NVT, NJB, EW)))) int8Correlate(Int8Args a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    // run of frames major: an XCD takes one run of frames over a contiguous range of windows, so that
    // its L2 holds rows of the frames of that run only (3 % faster at 720p than window major)
    const int logical = xcdRemap(blockIdx.x, gridDim.x);
    const int fblock = logical / a.g.nwin, k = logical - fblock * a.g.nwin;
    const int fBegin = fblock * a.fpb, fEnd = min(fBegin + a.fpb, a.fg.n);
    int8Window<NVT, NJB, EW>(a, k, fBegin, fEnd, lds);
}

// The persistent form: a grid of resident blocks drains the work queue of a tile shape (every clip
// of a batch with this window and search range; clips[c] holds clip c's arguments, fpb and nfb
// included). Each block takes the next item of its XCD's queue, then steals from the other XCDs'.
template <int NVT, int NJB, int EW>
__global__ void __launch_bounds__(64 * (NVT + EW))
__attribute__((amdgpu_waves_per_eu(int8QueueWavesPerEu(NVT, NJB, EW)))) int8Queue(WorkQueue q, const Int8Args *__restrict__ clips)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int item[2];
    for (int round = 0;; round++) {
        if (threadIdx.x == 0) item[round & 1] = q.next(blockIdx.x);
        __syncthreads();   // also: the previous item is over (its LDS may be rewritten)
        const int it = __builtin_amdgcn_readfirstlane(item[round & 1]);
        if (it < 0) break;
        int c, k, fBegin, fEnd;
        q.locateRunMajor(it, c, k, fBegin, fEnd);
        int8Window<NVT, NJB, EW>(clips[c], k, fBegin, fEnd, lds);
    }
}

// ---- the pipelined kernel: producer and helper wavefronts
//
// Waves 0 .. P-1 (P = NVT bands) only run MFMAs; waves P .. are helpers. Frames are pipelined over two
// chip buffers and two raw-surface buffers in LDS; iteration j has two phases, each ended by a block
// barrier:
//   phase A  producers: the MFMA bands of frame j from chip[j % 2], kept in registers
//            helpers:   stage the chip of frame j + 1 into chip[(j + 1) % 2]; screen frame j - 1 from
//                       surf[(j - 1) % 2] (each helper its rows: FP32 screen, exact FP64 best of its
//                       candidates into part[(j - 1) % 2][helper]); helper 0 also finishes frame j - 2
//                       from part[j % 2] and surf[j % 2] (peak, SNR, 3 x 3 FP64 fit, covariance)
//   phase B  producers: write frame j's raw surface into surf[j % 2]
// Every buffer a phase reads was written before the previous barrier and is written again only after
// the next one, so one frame's MFMAs overlap the staging, screening and peak fitting of its neighbours.

template <int NVT, int NJB>
__device__ __attribute__((always_inline)) inline void int8PipedWindow(const Int8Args &a, int k, int fBegin, int fEnd,
                                                                      unsigned char *lds)
{
    const Geometry &g = a.g;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid >> 6, nthreads = blockDim.x;
    const int P = a.nut, nh = nthreads / kWave - P;
    const bool producer = wave < P;
    const int h = wave - P, htid = tid - kWave * P, hthreads = kWave * nh;
    const int F = fEnd - fBegin;
    const int W = g.W, M = g.M, L = g.L;

    const int tbytes = 4 * a.ni4 * a.tstride, cbytes = a.crows * a.cstride, sints = (M * M + 3) & ~3;
    int8_t *tmpl = reinterpret_cast<int8_t *>(lds);
    int8_t *chips[2] = {tmpl + tbytes, tmpl + tbytes + cbytes};
    int *surfs[2] = {reinterpret_cast<int *>(tmpl + tbytes + 2 * cbytes),
                     reinterpret_cast<int *>(tmpl + tbytes + 2 * cbytes) + sints};
    WaveBest *parts = reinterpret_cast<WaveBest *>(surfs[1] + sints);   // 2 x nh

    loadTemplate(lds, tbytes + 2 * cbytes, tmpl, a.tmpl + size_t(k) * W * W, W, a.tstride, tid, nthreads);
    const float normk = a.wd.norm[k];
    const int y0 = a.wd.down[k / g.nx], x0 = a.wd.across[k % g.nx];
    Moments mom;
    mom.n = int64_t(W) * W;
    mom.tvar = a.wd.tvar[k];
    mom.nd = double(mom.n);
    mom.tau = double(a.wd.tsum[k] - 128 * mom.n);
    mom.rst = mom.tvar > 0 ? float(1.0 / sqrt(double(mom.tvar))) : 0.0f;
    if (!producer && F > 0) {
        const FrameWindow fw = frameWindow(a, k, fBegin, y0, x0, normk);
        if (fw.valid)
            loadChip(chips[0], a.fg.frames + size_t(fBegin) * g.pixels + size_t(fw.sy) * g.w + fw.sx, g.w, L,
                     a.cstride, htid, hthreads);
    }
    __syncthreads();

    const int g4 = lane >> 4, m16 = lane & 15;
    for (int j = 0; j < F + 2; j++) {
        v4i acc[NVT];
        bool wrote = false;
        // phase A
        if (producer) {
            if (j < F) {
                const FrameWindow fw = frameWindow(a, k, fBegin + j, y0, x0, normk);
                if (fw.valid) {
                    bandMfma<NVT, NJB>(acc, chips[j & 1], tmpl, wave, lane, a.ni4, a.tstride, a.cstride);
                    wrote = true;
                }
            }
        }
        else {
            if (j + 1 < F) {
                const int f = fBegin + j + 1;
                const FrameWindow fw = frameWindow(a, k, f, y0, x0, normk);
                if (fw.valid)
                    loadChip(chips[(j + 1) & 1], a.fg.frames + size_t(f) * g.pixels + size_t(fw.sy) * g.w + fw.sx,
                             g.w, L, a.cstride, htid, hthreads);
            }
            if (j >= 1 && j <= F) {
                const int f = fBegin + j - 1;
                const FrameWindow fw = frameWindow(a, k, f, y0, x0, normk);
                if (fw.valid) helperScreen(a, mom, surfs[(j - 1) & 1], fw, f, h, nh, lane, parts + ((j - 1) & 1) * nh + h);
            }
            if (h == 0 && j >= 2) {
                const int f = fBegin + j - 2;
                const FrameWindow fw = frameWindow(a, k, f, y0, x0, normk);
                helperFinish(a, mom, surfs[j & 1], fw, f, parts + (j & 1) * nh, nh, lane);
            }
        }
        __syncthreads();
        // phase B
        if (wrote) {
            int *S = surfs[j & 1];
#pragma unroll
            for (int v1 = 0; v1 < NVT; v1++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int u = 16 * wave + 4 * g4 + r, v = 16 * v1 + m16;
                    if (u < M && v < M) S[u * M + v] = acc[v1][r];
                }
        }
        __syncthreads();
    }
}

template <int NVT, int NJB>
__global__ void __launch_bounds__(64 * kPipedWaves) int8Piped(Int8Args a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const int logical = xcdRemap(blockIdx.x, gridDim.x);
    const int k = logical / a.nfb, fblock = logical - k * a.nfb;
    const int fBegin = fblock * a.fpb, fEnd = min(fBegin + a.fpb, a.fg.n);
    int8PipedWindow<NVT, NJB>(a, k, fBegin, fEnd, lds);
}

template <int NVT, int NJB>
__global__ void __launch_bounds__(64 * kPipedWaves) int8PipedQueue(WorkQueue q, const Int8Args *__restrict__ clips)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int item[2];
    for (int round = 0;; round++) {
        if (threadIdx.x == 0) item[round & 1] = q.next(blockIdx.x);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(item[round & 1]);
        if (it < 0) break;
        int c, k, fBegin, fEnd;
        q.locate(it, c, k, fBegin, fEnd);
        int8PipedWindow<NVT, NJB>(clips[c], k, fBegin, fEnd, lds);
    }
}

// ---- one wave per window (small surfaces)
//
// A surface of at most four 16-row bands and an edge row and column (M = 16 c + 1 <= 65) or at most three
// full bands (M <= 48), against a template of at most 64 columns, gives a block of several waves little
// MFMA work per frame against its fixed costs (barriers among its waves, the rounds of map loads, the
// finisher's tail), with only a few windows per CU in flight. Here one wave does all of a window's
// frame: its bands (each template fragment built once for all of them), where M = 16 c + 1 the edge row
// and column as the edge waves compute them, the screen of every point, the peak (blockPeak with one
// part), the SNR box in LDS and the fit (finishPeak), waiting for no other wave. The template rows are
// the four-wave kernel's (the narrowest conflict-free stride), the chip's rows its L rows (edge form) or
// the full bands' reach: 10 KB a block at window 32, search 16 (sixteen windows per CU), 16 KB at
// window 32, search 32 (ten). The same screen, peak and finishing as the other kernels: the same bits.

// the chip as loadChip stages it, by the one wave: up to eight rows' loads a lane in flight before any
// is waited for (a row past the chip's end loads its last row and is not stored)
__device__ inline void soloChip(int8_t *chip, const uint8_t *src, int w, int L, int cstride, int lane)
{
    const int cpr = L >> 4;
    if (cpr > 0) {
        const int rowsPer = kWave / cpr, r0 = lane / cpr, cc = lane - r0 * cpr;
        if (r0 < rowsPer)
            for (int rb = r0; rb < L; rb += 8 * rowsPer) {
                uint4 v[8];
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const int r = min(rb + i * rowsPer, L - 1);
                    __builtin_memcpy(&v[i], src + size_t(r) * w + 16 * cc, 16);
                }
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const int r = rb + i * rowsPer;
                    if (r < L) {
                        v[i].x ^= 0x80808080u;
                        v[i].y ^= 0x80808080u;
                        v[i].z ^= 0x80808080u;
                        v[i].w ^= 0x80808080u;
                        *reinterpret_cast<uint4 *>(chip + r * cstride + 16 * cc) = v[i];
                    }
                }
            }
    }
    const int tail = L & 15;
    if (tail)
        for (int c = lane; c < L * tail; c += kWave) {
            const int r = c / tail, j = 16 * cpr + (c - r * tail);
            chip[r * cstride + j] = int8_t(src[size_t(r) * w + j] ^ 0x80);
        }
}

// every band of the surface into acc[band][v tile], as bandMfma band by band, but with the template's B
// fragments of a row group built once for all the bands
template <int NVT, int NJB>
__device__ inline void soloBands(v4i (&acc)[NVT][NVT], const int8_t *chip, const int8_t *tmpl, int lane, int ni4,
                                 int tstride, int cstride)
{
    const int g4 = lane >> 4, m16 = lane & 15;
    const int bbyte = 16 - m16, bdw = bbyte >> 2, bsh = bbyte & 3;
#pragma unroll
    for (int b = 0; b < NVT; b++)
#pragma unroll
        for (int v1 = 0; v1 < NVT; v1++) acc[b][v1] = v4i{0, 0, 0, 0};
    const int8_t *crow = chip + (g4 + m16) * cstride;
    const uint32_t *trow = reinterpret_cast<const uint32_t *>(tmpl + g4 * tstride) + bdw;
    for (int i4 = 0; i4 < ni4; i4++) {
        v4i bf[NJB];
#pragma unroll
        for (int jb = 0; jb < NJB; jb++) {
            const uint32_t *t = trow + 4 * jb;
            const uint32_t w0 = t[0], w1 = t[1], w2 = t[2], w3 = t[3], w4 = t[4];
            bf[jb] = v4i{int(__builtin_amdgcn_alignbyte(w1, w0, bsh)), int(__builtin_amdgcn_alignbyte(w2, w1, bsh)),
                         int(__builtin_amdgcn_alignbyte(w3, w2, bsh)), int(__builtin_amdgcn_alignbyte(w4, w3, bsh))};
        }
#pragma unroll
        for (int b = 0; b < NVT; b++)
#pragma unroll
            for (int cb = 0; cb < NVT + NJB - 1; cb++) {
                const v4i av = *reinterpret_cast<const v4i *>(crow + 16 * b * cstride + 16 * cb);
#pragma unroll
                for (int jb = 0; jb < NJB; jb++) {
                    const int v1 = cb - jb;
                    if (v1 >= 0 && v1 < NVT) acc[b][v1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, bf[jb], acc[b][v1], 0, 0, 0);
                }
            }
        crow += 4 * cstride;
        trow += tstride;
    }
}

// NVT: the full bands (EW = 0) or the core bands c (EW = 2: with the edge row u = 16 c and column v = 16 c)
template <int NVT, int NJB, int EW>
__device__ __attribute__((always_inline)) inline void int8SoloWindow(const Int8Args &a, int k, int fBegin, int fEnd,
                                                                     unsigned char *lds)
{
    const Geometry &g = a.g;
    const int lane0 = threadIdx.x & (kWave - 1);
    const int W = g.W, M = g.M, L = g.L;
    const Int8Lds sh = int8Lds(a.ni4, a.tstride, a.crows, a.cstride, NJB, M, g.stat, 1);
    int8_t *tmpl = reinterpret_cast<int8_t *>(lds);
    const int8_t *zero = tmpl + sh.zero;
    int8_t *chip = tmpl + sh.chip;
    int *boxbuf = reinterpret_cast<int *>(lds + sh.box);
    WaveBest *parts = reinterpret_cast<WaveBest *>(lds + sh.parts);

    loadTemplate(lds, sh.box, tmpl, a.tmpl + size_t(k) * W * W, W, a.tstride, lane0, kWave);
    const float normk = a.wd.norm[k];
    const int y0 = a.wd.down[k / g.nx], x0 = a.wd.across[k % g.nx];
    Moments mom;
    mom.n = int64_t(W) * W;
    mom.tvar = a.wd.tvar[k];
    mom.nd = double(mom.n);
    mom.tau = double(a.wd.tsum[k] - 128 * mom.n);
    mom.rst = mom.tvar > 0 ? float(1.0 / sqrt(double(mom.tvar))) : 0.0f;
    __syncthreads();   // the template (and the zeroed areas)

    for (int f = fBegin; f < fEnd; f++) {
        const FrameWindow fw = frameWindow(a, k, f, y0, x0, normk);
        if (!fw.valid) {
            if (lane0 == 0) invalidateOut(fw.q, a.fg.offset, a.fg.peak, a.fg.snr, a.fg.cov);
            continue;
        }
        soloChip(chip, a.fg.frames + size_t(f) * g.pixels + size_t(fw.sy) * g.w + fw.sx, g.w, L, a.cstride, lane0);
        __syncthreads();   // the chip
        int lane = lane0;
        asm volatile("" : "+v"(lane));
        const int g4 = lane >> 4, m16 = lane & 15;
        v4i acc[NVT][NVT];
        soloBands<NVT, NJB>(acc, chip, tmpl, lane, a.ni4, a.tstride, a.cstride);
        v4i er = v4i{0, 0, 0, 0}, ec = v4i{0, 0, 0, 0};
        if constexpr (EW != 0) {
            er = edgeRowMfma<NJB>(chip, tmpl, NVT, lane, a.ni4, a.tstride, a.cstride);
            ec = edgeColMfma<NJB>(chip, tmpl, zero, NVT, W, lane, a.tstride, a.cstride);
        }

        // this wave's points: fn(raw, u, v); the edge row's and column's (as their waves hold them) after
        // the bands'
        auto visitEdges = [&](auto &&fn) {
            if constexpr (EW != 0) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int blk = 4 * g4 + r, v = 16 * blk + m16;
                    if (blk <= NVT && v < M) fn(er[r], 16 * NVT, v);
                }
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int blk = 4 * g4 + r;
                    if (blk < NVT) fn(ec[r], 16 * blk + m16, 16 * NVT);
                }
            }
        };
        auto visit = [&](auto &&fn) {
#pragma unroll
            for (int b = 0; b < NVT; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int u = 16 * b + 4 * g4 + r;
#pragma unroll
                    for (int v1 = 0; v1 < NVT; v1++) {
                        const int v = 16 * v1 + m16;
                        if (EW != 0 || (u < M && v < M)) fn(acc[b][v1][r], u, v);
                    }
                }
            visitEdges(fn);
        };
        const int64_t boff = f * g.boxSize + int64_t(fw.sy) * g.bw + fw.sx;
        const int2 *scr = a.fg.screen + boff;
        const int32_t *box2 = a.fg.box2 + boff;
        // the screen, band by band, the map loads of a band's four rows in flight together
        Top2 t = top2Init();
#pragma unroll
        for (int b = 0; b < NVT; b++) {
            int2 p[4][NVT];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int u = 16 * b + 4 * g4 + r;
                const int2 *row = scr + u * g.bw + m16;
#pragma unroll
                for (int v1 = 0; v1 < NVT; v1++)
                    p[r][v1] = (EW != 0 || (u < M && 16 * v1 + m16 < M)) ? row[16 * v1] : make_int2(0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int u = 16 * b + 4 * g4 + r;
#pragma unroll
                for (int v1 = 0; v1 < NVT; v1++)
                    if (EW != 0 || (u < M && 16 * v1 + m16 < M))
                        top2Add(t, screenNcc(acc[b][v1][r], p[r][v1].x, __int_as_float(p[r][v1].y), mom),
                                u * M + 16 * v1 + m16);
            }
        }
        visitEdges([&](int raw, int u, int v) {
            const int2 p = scr[u * g.bw + v];
            top2Add(t, screenNcc(raw, p.x, __int_as_float(p.y), mom), u * M + v);
        });
        screenPart(t, parts, lane);
        __syncthreads();
        int ib;
        const bool valid = blockPeak(parts, 1, 0, lane, mom, scr, box2, g.bw, M, visit, ib);
        if (valid) {
            const SnrBox b = snrBox(ib, M, g.stat);
            visit([&](int raw, int u, int v) {
                if (u >= b.ulo && u < b.uhi && v >= b.vlo && v < b.vhi) boxbuf[(u - b.ulo) * b.bwid + v - b.vlo] = raw;
            });
        }
        __syncthreads();   // the SNR box
        finishPeak(a, mom, boxbuf, true, fw, f, valid, ib, lane);
        __syncthreads();   // the box and the chip read: the next frame may rewrite them
    }
}

// The register budget: up to two full bands or the edge form's 2 x 2 core tiles four waves per SIMD (128
// registers; 10 KB of LDS at window 32 leaves room for sixteen blocks per CU); three full bands or 3 x 3
// core tiles three (168); 4 x 4 core tiles two (about 200 registers; at 168 they spill)
constexpr int soloWavesPerEu(int nvt, int ew) { return ew != 0 && nvt >= 4 ? 2 : nvt >= 3 ? 3 : 4; }

template <int NVT, int NJB, int EW>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(soloWavesPerEu(NVT, EW)))) int8Solo(Int8Args a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const int logical = xcdRemap(blockIdx.x, gridDim.x);
    const int fblock = logical / a.g.nwin, k = logical - fblock * a.g.nwin;
    const int fBegin = fblock * a.fpb, fEnd = min(fBegin + a.fpb, a.fg.n);
    int8SoloWindow<NVT, NJB, EW>(a, k, fBegin, fEnd, lds);
}

template <int NVT, int NJB, int EW>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(soloWavesPerEu(NVT, EW))))
int8SoloQueue(WorkQueue q, const Int8Args *__restrict__ clips)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int item[2];
    for (int round = 0;; round++) {
        if (threadIdx.x == 0) item[round & 1] = q.next(blockIdx.x);
        __syncthreads();   // also: the previous item is over (its LDS may be rewritten)
        const int it = __builtin_amdgcn_readfirstlane(item[round & 1]);
        if (it < 0) break;
        int c, k, fBegin, fEnd;
        q.locateRunMajor(it, c, k, fBegin, fEnd);
        int8SoloWindow<NVT, NJB, EW>(clips[c], k, fBegin, fEnd, lds);
    }
}

template <int NVT, int NJB, int EW>
const void *soloKernel(bool queue)
{
    return queue ? reinterpret_cast<const void *>(int8SoloQueue<NVT, NJB, EW>)
                 : reinterpret_cast<const void *>(int8Solo<NVT, NJB, EW>);
}

// nvt: the full bands (edge false: 1..3) or the core bands (edge: 1..4)
template <int NJB>
const void *soloOf(int nvt, bool edge, bool queue)
{
    if (edge) switch (nvt) {
        case 1: return soloKernel<1, NJB, 2>(queue);
        case 2: return soloKernel<2, NJB, 2>(queue);
        case 3: return soloKernel<3, NJB, 2>(queue);
        case 4: return soloKernel<4, NJB, 2>(queue);
        default: return nullptr;
        }
    switch (nvt) {
    case 1: return soloKernel<1, NJB, 0>(queue);
    case 2: return soloKernel<2, NJB, 0>(queue);
    case 3: return soloKernel<3, NJB, 0>(queue);
    default: return nullptr;
    }
}

// the kernels of one tile shape: kind 0 the one-frame kernel with full bands, 1 the same with the edge
// row and column waves, 2 pipelined; direct (one block per window and run of frames) or queue
template <int NVT, int NJB>
const void *kernelOf(int kind, bool queue)
{
    switch (kind) {
    case 0: return queue ? reinterpret_cast<const void *>(int8Queue<NVT, NJB, 0>)
                         : reinterpret_cast<const void *>(int8Correlate<NVT, NJB, 0>);
    case 1: return queue ? reinterpret_cast<const void *>(int8Queue<NVT, NJB, 2>)
                         : reinterpret_cast<const void *>(int8Correlate<NVT, NJB, 2>);
    default: return queue ? reinterpret_cast<const void *>(int8PipedQueue<NVT, NJB>)
                          : reinterpret_cast<const void *>(int8Piped<NVT, NJB>);
    }
}

template <int NJB>
const void *kernelNvt(int nvt, int kind, bool queue)
{
    switch (nvt) {
    case 1: return kernelOf<1, NJB>(kind, queue);
    case 2: return kernelOf<2, NJB>(kind, queue);
    case 3: return kernelOf<3, NJB>(kind, queue);
    case 4: return kernelOf<4, NJB>(kind, queue);
    case 5: return kernelOf<5, NJB>(kind, queue);
    case 6: return kernelOf<6, NJB>(kind, queue);
    case 7: return kernelOf<7, NJB>(kind, queue);
    case 8: return kernelOf<8, NJB>(kind, queue);
    case 9: return kernelOf<9, NJB>(kind, queue);
    default: return nullptr;
    }
}

// ---- exact NCC surfaces: the whole M x M FP64 surface of each window against one frame
//
// One block of 256 threads per window. The template and the chip are staged in LDS as in the bands;
// the ceil(M / 16)^2 tiles of 16 x 16 surface points are dealt to the four waves, each tile the K loop
// of the bands (template row groups i4, column blocks jb) with its B fragments built per tile. Every
// point is normalized with the exact FP64 NCC of the frame's box maps and written as a double.

constexpr int kSurfThreads = 256;

__global__ void __launch_bounds__(kSurfThreads) nccSurfaces(SurfaceArgs s)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const Geometry &g = s.g;
    const int k = s.k0 + blockIdx.x;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    int gy = 0, gx = 0;
    if (s.gross) {
        gy = s.gross[2 * k];
        gx = s.gross[2 * k + 1];
    }
    int sy, sx;
    if (!searchWindow(g, s.wd.down[k / g.nx], s.wd.across[k % g.nx], gy, gx, sy, sx)) {
        // the search window leaves the frame (uniform over the block): a surface of NaN
        double *bad = s.surf + size_t(blockIdx.x) * g.M * g.M;
        for (int p = tid; p < g.M * g.M; p += kSurfThreads) bad[p] = __builtin_nan("");
        return;
    }
    int8_t *tmpl = reinterpret_cast<int8_t *>(lds);
    const int tbytes = 4 * s.ni4 * s.tstride;
    int8_t *chip = tmpl + tbytes;
    loadTemplate(lds, tbytes + s.crows * s.cstride, tmpl, s.tmpl + size_t(k) * g.W * g.W, g.W, s.tstride, tid,
                 kSurfThreads);
    loadChip(chip, s.frame + size_t(sy) * g.w + sx, g.w, g.L, s.cstride, tid, kSurfThreads);
    __syncthreads();

    Moments mom;
    mom.n = int64_t(g.W) * g.W;
    mom.tvar = s.wd.tvar[k];
    mom.nd = double(mom.n);
    mom.tau = double(s.wd.tsum[k] - 128 * mom.n);
    mom.rst = 0.0f;
    const int M = g.M;
    double *out = s.surf + size_t(blockIdx.x) * M * M;
    const int g4 = lane >> 4, m16 = lane & 15;
    const int bbyte = 16 - m16, bdw = bbyte >> 2, bsh = bbyte & 3;
    const int nt = s.nt;
    for (int tile = wave; tile < nt * nt; tile += kSurfThreads / kWave) {
        const int t = tile / nt, v1 = tile - t * nt;
        v4i acc = v4i{0, 0, 0, 0};
        const int8_t *crow = chip + (16 * t + g4 + m16) * s.cstride + 16 * v1;
        const uint32_t *trow = reinterpret_cast<const uint32_t *>(tmpl + g4 * s.tstride) + bdw;
        for (int i4 = 0; i4 < s.ni4; i4++) {
            for (int jb = 0; jb < s.njb; jb++) {
                const uint32_t *tp = trow + 4 * jb;
                const uint32_t w0 = tp[0], w1 = tp[1], w2 = tp[2], w3 = tp[3], w4 = tp[4];
                const v4i b = v4i{int(__builtin_amdgcn_alignbyte(w1, w0, bsh)), int(__builtin_amdgcn_alignbyte(w2, w1, bsh)),
                                  int(__builtin_amdgcn_alignbyte(w3, w2, bsh)), int(__builtin_amdgcn_alignbyte(w4, w3, bsh))};
                const v4i av = *reinterpret_cast<const v4i *>(crow + 16 * jb);
                acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, b, acc, 0, 0, 0);
            }
            crow += 4 * s.cstride;
            trow += s.tstride;
        }
        const int v = 16 * v1 + m16;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int u = 16 * t + 4 * g4 + r;
            if (u < M && v < M) {
                const int64_t b = int64_t(sy + u) * g.bw + sx + v;
                out[u * M + v] = exactNcc(acc[r], s.box[b], s.box2[b], mom);
            }
        }
    }
}

// One block per surface: the first maximum and its parabolic refinement along each axis (the peak of
// python/slipcam/stabilize.py:22-45 ncc_peak)
__device__ inline double parabola(double c, double m, double p)
{
    const double den = (m - 2.0 * c) + p;
    return fabs(den) < 1e-12 ? 0.0 : 0.5 * (m - p) / den;
}

__global__ void __launch_bounds__(256) parabolaPeaks(const double *surf, int M, const int64_t *tvar, int k0,
                                                     double *offsets, double *peak)
{
    __shared__ double pv[256 / kWave];
    __shared__ int pi[256 / kWave];
    const double *S = surf + size_t(blockIdx.x) * M * M;
    const int k = k0 + blockIdx.x;
    double best = -3.0;
    int ib = 0x7fffffff;
    for (int p = threadIdx.x; p < M * M; p += 256) {
        const double x = S[p];
        if (x > best) {   // ascending p: the first of equal values stays
            best = x;
            ib = p;
        }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const double b2 = __shfl_xor(best, o, kWave);
        const int i2 = __shfl_xor(ib, o, kWave);
        if (b2 > best || (b2 == best && i2 < ib)) {
            best = b2;
            ib = i2;
        }
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        pv[threadIdx.x / kWave] = best;
        pi[threadIdx.x / kWave] = ib;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int w = 1; w < 256 / kWave; w++)
        if (pv[w] > best || (pv[w] == best && pi[w] < ib)) {
            best = pv[w];
            ib = pi[w];
        }
    if (tvar[k] == 0 || isnan(S[0]) || ib >= M * M) {
        offsets[2 * k] = __builtin_nan("");
        offsets[2 * k + 1] = __builtin_nan("");
        peak[k] = 0.0;
        return;
    }
    const int u = ib / M, v = ib - u * M;
    const double du = u > 0 && u < M - 1 ? parabola(best, S[ib - M], S[ib + M]) : 0.0;
    const double dv = v > 0 && v < M - 1 ? parabola(best, S[ib - 1], S[ib + 1]) : 0.0;
    const double c0 = double(M - 1) / 2.0;
    offsets[2 * k] = (double(u) + du) - c0;
    offsets[2 * k + 1] = (double(v) + dv) - c0;
    peak[k] = best;
}

// One block per surface: the local maxima outside the excluded square, their largest value
constexpr int kPeakThreads = 256;

__global__ void __launch_bounds__(kPeakThreads) secondPeak(const double *surf, int M, int lo, int hi,
                                                           const int64_t *tvar, int k0, double *out)
{
    __shared__ double part[kPeakThreads / kWave];
    const double *S = surf + size_t(blockIdx.x) * M * M;
    double best = -2.0;   // below every NCC value: no local maximum yet
    for (int p = threadIdx.x; p < M * M; p += kPeakThreads) {
        const int u = p / M, v = p - u * M;
        if (u >= lo && u < hi && v >= lo && v < hi) continue;
        const double x = S[p];
        if (!(x > best)) continue;   // cannot raise this thread's best
        bool peak = true;
        for (int uu = max(u - 2, 0); uu <= min(u + 2, M - 1) && peak; uu++)
            for (int vv = max(v - 2, 0); vv <= min(v + 2, M - 1); vv++)
                if (S[uu * M + vv] > x) {
                    peak = false;
                    break;
                }
        if (peak) best = x;
    }
    for (int o = kWave / 2; o > 0; o >>= 1) best = fmax(best, __shfl_xor(best, o, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kPeakThreads / kWave; w++) best = fmax(best, part[w]);
        const int k = k0 + blockIdx.x;
        out[k] = tvar[k] == 0 ? 1.0 : best > -2.0 ? best : 0.0;
    }
}

} // namespace

// the instantiated template widths: NJB 16-byte blocks cover templates of W <= 16 NJB - 15
int int8JBlocks(int W)
{
    const int need = (W + 15 + 15) / 16;
    if (need <= 2) return 2;
    if (need <= 3) return 3;
    if (need <= 5) return 5;
    if (need <= 9) return 9;
    return 0;
}

Int8Layout int8Layout(const Geometry &g)
{
    Int8Layout l{};
    l.W = g.W;
    l.njb = int8JBlocks(g.W);
    l.nvt = (g.M + 15) / 16;
    l.nut = l.nvt;
    l.ni4 = (g.W + 3) / 4;
    // template rows: at least 16 (NJB + 2) bytes, and 32 + 128 k bytes so that the four rows a
    // fragment build reads at once (the four lane groups) fall on disjoint LDS banks
    l.tstride = 32;
    while (l.tstride < 16 * (l.njb + 2)) l.tstride += 128;
    int cs = 16 * (l.nvt + l.njb - 1);
    if (cs < g.L) cs = g.L;
    cs = (cs + 15) / 16;
    if (cs % 2 == 0) cs++;   // an odd number of 16-byte units: conflict-free 16-byte row reads
    l.cstride = 16 * cs;
    l.crows = 16 * l.nut + 4 * l.ni4;
    // (one result per wave: nut, or nut + 1 with the edge row and column in two waves)
    l.shmem = size_t(int8Lds(l.ni4, l.tstride, l.crows, l.cstride, l.njb, g.M, g.stat, l.nut + 1).total);
    l.alias = false;
    l.supported = l.njb > 0 && l.nvt <= 9 && l.shmem <= 160 * 1024;
    l.edge = g.M % 16 == 1 && g.M > 1;
    // the edge form reads no chip row below the L of the search window (the padded bands do), and for
    // templates of 113-128 columns it takes template rows of 176 bytes, the narrowest multiple of 16 that
    // holds a row (the four rows of a fragment build then share one bank of their five): 52 KB at
    // window 128, search 16, three blocks per CU against two
    l.edgeTstride = l.tstride;
    l.edgeCrows = l.crows;
    l.edgeShmem = l.shmem;
    if (l.edge && l.njb >= 8) {
        l.edgeTstride = 16 * ((16 * l.njb + 20 + 15) / 16);
        l.edgeCrows = g.L;
        l.edgeShmem = size_t(int8Lds(l.ni4, l.edgeTstride, l.edgeCrows, l.cstride, l.njb, g.M, g.stat, l.nut + 1).total);
    }
    // the pipelined kernel: two chips, two raw surfaces, two sets of helper results
    const int nh = kPipedWaves - l.nut;
    l.pipedShmem = size_t(4 * l.ni4) * l.tstride + 2 * size_t(l.crows) * l.cstride
        + 2 * sizeof(int) * size_t((g.M * g.M + 3) & ~3) + 2 * sizeof(WaveBest) * size_t(nh > 0 ? nh : 0);
    l.pipedOk = l.supported && nh >= 2 && l.pipedShmem <= 160 * 1024;
    // the 32 x 32 kernel: template rows with 32 zero bytes in front (the 32 x 32 fragments read bytes
    // 32 - n .. 32 + 16 nj + 3, the 16 x 16 edge fragments 16 bytes in), at a stride whose two rows of a
    // fragment build (nine dwords each) fall on disjoint banks; chip rows of an odd number of 16-byte units
    l.nb = (g.M - 1) / 32;
    l.nj = (g.W + 46) / 16;
    l.wideOk = l.supported && g.M % 32 == 1 && l.nb >= 1 && l.nb <= 4 && int8WideKernel(l.nb, l.nj, l.njb, false);
    if (l.wideOk) {
        int ts = std::max(32 + 16 * l.nj + 4, 36 + 16 * l.njb);
        ts = (ts + 15) & ~15;
        while ((ts / 4) % 64 < 12 || (ts / 4) % 64 > 52) ts += 16;
        l.wideTstride = ts;
        int wcs = std::max(std::max(16 * (2 * l.nb + l.nj - 2), 16 * (2 * l.nb + l.njb)), g.L);
        wcs = (wcs + 15) / 16;
        if (wcs % 2 == 0) wcs++;
        l.wideCstride = 16 * wcs;
        l.wideCrows = 32 * l.nb + 4 * l.ni4;
        l.wideShmem = size_t(int8Lds(l.ni4, l.wideTstride, l.wideCrows, l.wideCstride, l.njb, g.M, g.stat, l.nb + 2).total);
        l.wideOk = l.wideShmem <= 160 * 1024;
    }
    // the four-wave kernel: the chip's L rows only, and the narrowest template rows (a multiple of 16
    // bytes, for the edge column's 16-byte reads) whose four rows of a fragment build (five dwords each
    // at the same offset) fall on disjoint banks; four blocks per CU at window 64, search 48
    {
        int ts = (16 * l.njb + 20 + 15) & ~15;
        for (;; ts += 16) {
            bool ok = true;
            for (int r = 1; r < 4 && ok; r++) {
                const int d = (r * ts / 4) % 64;
                ok = d >= 5 && d <= 59;
            }
            for (int r1 = 1; r1 < 4 && ok; r1++)
                for (int r2 = r1 + 1; r2 < 4 && ok; r2++) {
                    const int d = ((r2 - r1) * ts / 4) % 64;
                    ok = d >= 5 && d <= 59;
                }
            if (ok) break;
        }
        l.quadTstride = ts;
    }
    l.quadCrows = g.L;
    l.quadShmem = int8QuadShmem(l.ni4, l.quadTstride, l.quadCrows, l.cstride, l.njb, g);
    l.quadOk = l.supported && l.edge && l.quadShmem <= 80 * 1024 && int8QuadKernel(l.nvt - 1, g.W, l.njb, false);
    // one wave per window: the four-wave kernel's template rows, the chip's L rows (edge form) or the full
    // bands' reach, one screen part
    l.soloCrows = l.edge ? g.L : l.crows;
    l.soloShmem = size_t(int8Lds(l.ni4, l.quadTstride, l.soloCrows, l.cstride, l.njb, g.M, g.stat, 1).total);
    l.soloOk = l.supported && l.njb <= 5 && (l.edge ? l.nvt - 1 <= 4 : l.nvt <= 3);
    return l;
}

Int8Kernel int8Kernel(const Int8Layout &l, Int8Variant variant, bool queue)
{
    if (!l.supported) throw std::invalid_argument("framecor: the int8 form supports windows up to 129 pixels and search ranges up to 71 pixels");
    if (l.wideOk && variant == Int8Variant::Wide) {
        Int8Kernel k{};
        k.func = int8WideKernel(l.nb, l.nj, l.njb, queue);
        k.threads = kWave * (l.nb + 2);
        k.shmem = l.wideShmem;
        k.variant = Int8Variant::Wide;
        allowLargeLds(k.func);
        return k;
    }
    // automatic: the one-wave kernel for templates of up to 33 columns (one clip at 720p, 64 frames: -22 %
    // at window 32 / search 16, -3 % at search 32 against the four-wave kernel; +5 % at window 64), else
    // the four-wave kernel where it applies
    const bool autoSolo = variant == Int8Variant::Auto && l.soloOk && l.njb <= 3;
    if (l.quadOk && !autoSolo && (variant == Int8Variant::Quad || variant == Int8Variant::Auto)) {
        Int8Kernel k{};
        k.func = int8QuadKernel(l.nvt - 1, l.W, l.njb, queue);
        k.threads = int8QuadThreads();
        k.shmem = l.quadShmem;
        k.variant = Int8Variant::Quad;
        allowLargeLds(k.func);
        return k;
    }
    if (l.soloOk && (variant == Int8Variant::Solo || autoSolo)) {
        Int8Kernel k{};
        switch (l.njb) {
        case 2: k.func = soloOf<2>(l.edge ? l.nvt - 1 : l.nvt, l.edge, queue); break;
        case 3: k.func = soloOf<3>(l.edge ? l.nvt - 1 : l.nvt, l.edge, queue); break;
        case 5: k.func = soloOf<5>(l.edge ? l.nvt - 1 : l.nvt, l.edge, queue); break;
        default: break;
        }
        if (!k.func) throw std::invalid_argument("framecor: no one-wave int8 kernel for this window and search range");
        k.threads = kWave;
        k.shmem = l.soloShmem;
        k.variant = Int8Variant::Solo;
        allowLargeLds(k.func);
        return k;
    }
    const bool piped = variant == Int8Variant::Piped && l.pipedOk;
    const bool edge = !piped && variant != Int8Variant::Padded && l.edge;
    const int kind = piped ? 2 : edge ? 1 : 0, nvt = edge ? l.nvt - 1 : l.nvt;
    Int8Kernel k{};
    switch (l.njb) {
    case 2: k.func = kernelNvt<2>(nvt, kind, queue); break;
    case 3: k.func = kernelNvt<3>(nvt, kind, queue); break;
    case 5: k.func = kernelNvt<5>(nvt, kind, queue); break;
    case 9: k.func = kernelNvt<9>(nvt, kind, queue); break;
    default: break;
    }
    if (!k.func) throw std::invalid_argument("framecor: no int8 kernel for this window and search range");
    k.threads = kWave * (piped ? kPipedWaves : nvt + (edge ? 2 : 0));
    k.shmem = piped ? l.pipedShmem : edge ? l.edgeShmem : l.shmem;
    k.variant = piped ? Int8Variant::Piped : edge ? Int8Variant::Simple : Int8Variant::Padded;
    allowLargeLds(k.func);
    return k;
}

void fillInt8Args(Int8Args &a, const Int8Layout &l, Int8Variant variant)
{
    const bool wide = variant == Int8Variant::Wide && l.wideOk, quad = variant == Int8Variant::Quad && l.quadOk;
    const bool edge = variant == Int8Variant::Simple && l.edge, solo = variant == Int8Variant::Solo && l.soloOk;
    a.nut = wide ? l.nb : l.nut;
    a.ni4 = l.ni4;
    a.tstride = wide ? l.wideTstride : quad || solo ? l.quadTstride : edge ? l.edgeTstride : l.tstride;
    a.cstride = wide ? l.wideCstride : l.cstride;
    a.crows = wide ? l.wideCrows : quad ? l.quadCrows : solo ? l.soloCrows : edge ? l.edgeCrows : l.crows;
    a.alias = l.alias ? 1 : 0;
}

void launchInt8(Int8Args a, const Int8Layout &l, Int8Variant variant, hipStream_t stream)
{
    if (a.fg.n <= 0) return;
    const Int8Kernel k = int8Kernel(l, variant, false);
    fillInt8Args(a, l, k.variant);
    // frames per block: at most 8 (at 720p, window 64, search 48, 64 frames in blocks of 8 run 2 % faster
    // than in blocks of 16 with the four-wave kernel, blocks of 4, 6, 11 or 13 slower, blocks of 64 slower
    // still), and enough blocks to fill the GPU several times over
    const int target = 2048, maxFpb = 8;
    int nfb = std::max((a.fg.n + maxFpb - 1) / maxFpb, (target + a.g.nwin - 1) / a.g.nwin);
    if (nfb > a.fg.n) nfb = a.fg.n;
    if (nfb < 1) nfb = 1;
    a.fpb = (a.fg.n + nfb - 1) / nfb;
    a.nfb = (a.fg.n + a.fpb - 1) / a.fpb;
    const int blocks = a.g.nwin * a.nfb;
    void *args[] = {&a};
    FC_HIP_CHECK(hipLaunchKernel(k.func, dim3(blocks), dim3(k.threads), args, k.shmem, stream));
}

void launchInt8Queue(const Int8Kernel &k, const WorkQueue &q, const Int8Args *clips, int blocks, hipStream_t stream)
{
    WorkQueue qa = q;
    const Int8Args *c = clips;
    void *args[] = {&qa, &c};
    FC_HIP_CHECK(hipLaunchKernel(k.func, dim3(blocks), dim3(k.threads), args, k.shmem, stream));
}

// ---- the surface kernel's layout and launchers

namespace {

struct SurfaceLayout {
    int njb, ni4, tstride, cstride, crows, nt;
    size_t shmem;
};

SurfaceLayout surfaceLayout(const Geometry &g)
{
    SurfaceLayout l;
    l.njb = (g.W + 30) / 16;   // the 16-byte column blocks that meet the template at some shift n < 16
    l.ni4 = (g.W + 3) / 4;
    l.nt = (g.M + 15) / 16;
    l.tstride = 32;
    while (l.tstride < 16 * (l.njb + 2)) l.tstride += 128;
    int cs = 16 * (l.nt + l.njb - 1);
    if (cs < g.L) cs = g.L;
    cs = (cs + 15) / 16;
    if (cs % 2 == 0) cs++;
    l.cstride = 16 * cs;
    l.crows = 16 * l.nt + 4 * l.ni4;
    l.shmem = size_t(4 * l.ni4) * l.tstride + size_t(l.crows) * l.cstride;
    return l;
}

} // namespace

size_t surfaceShmem(const Geometry &g) { return surfaceLayout(g).shmem; }

void launchNccSurfaces(SurfaceArgs s, hipStream_t stream)
{
    if (s.count <= 0) return;
    const SurfaceLayout l = surfaceLayout(s.g);
    if (l.shmem > 160 * 1024) throw std::invalid_argument("framecor: the search window is too large for the NCC surfaces");
    s.njb = l.njb;
    s.ni4 = l.ni4;
    s.tstride = l.tstride;
    s.cstride = l.cstride;
    s.crows = l.crows;
    s.nt = l.nt;
    allowLargeLds(reinterpret_cast<const void *>(nccSurfaces));
    void *args[] = {&s};
    FC_HIP_CHECK(hipLaunchKernel(reinterpret_cast<const void *>(nccSurfaces), dim3(s.count), dim3(kSurfThreads), args,
                                 l.shmem, stream));
}

void launchParabolaPeaks(const double *surf, int count, int M, const int64_t *tvar, int k0, double *offsets,
                         double *peak, hipStream_t stream)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(parabolaPeaks, dim3(count), dim3(256), 0, stream, surf, M, tvar, k0, offsets, peak);
    FC_HIP_CHECK(hipGetLastError());
}

void launchSecondPeak(const double *surf, int count, int M, int lo, int hi, const int64_t *tvar, int k0, double *out,
                      hipStream_t stream)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(secondPeak, dim3(count), dim3(kPeakThreads), 0, stream, surf, M, lo, hi, tvar, k0, out);
    FC_HIP_CHECK(hipGetLastError());
}

} // namespace framecor
