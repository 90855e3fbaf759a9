// int8_wide.hip -- the int8-MFMA form on 32 x 32 tiles (v_mfma_i32_32x32x32_i8) for surfaces of
// M = 32 c + 1 points (search ranges R = 16 c, c = 1..4)
//
// On gfx950 v_mfma_i32_16x16x64_i8 (32768 ops) issues about every 28 cycles per SIMD in a register-only
// loop, v_mfma_i32_32x32x32_i8 (65536 ops) every 32 cycles: the 32 x 32 instruction does 1.75x the
// work per cycle. The same Toeplitz contraction as the 16 x 16 bands, on 32-row bands of 32-column
// tiles:
//   A[m][(i, jb, kk)] = F[32 t + m + i][32 v1 + 16 jb + kk]      (16 contiguous chip bytes)
//   B[(i, jb, kk)][n] = T[i][16 jb + kk - n]                      (zero outside 0..W-1; n < 32)
//   C[m][n] = sum over i, jb, kk  -> r[32 t + m][32 v1 + n].
// Lane (h, m) of the A operand (h = lane / 32, m = lane % 32) holds 16 bytes of row m and lane (h, n)
// of B 16 bytes of column n, both at the K bytes of half h: one MFMA covers two template rows
// (i = 2 i2 + h) and one 16-byte column block jb. Register q of lane (h, n) of the 32 x 32 result
// holds C[8 (q / 4) + 4 h + q % 4][n]. jb runs over NJ = ceil((W + 31) / 16) blocks (the 32 shifts
// of a tile widen the Toeplitz band by 16 bytes over the 16 x 16 form: 1.2x the MFMA work at
// W = 64). A chip tile at column block cb = 2 v1 + jb serves every (v1, jb) on that diagonal, so each
// LDS read of A feeds up to NB MFMAs; the B fragments of two template rows are built once per row pair
// from the LDS template (32 zero bytes in front of each row: shifts n < 32) with byte funnel shifts.
//
// One block per (window, run of frames): NB band waves (32 rows, all NB tiles each), a wave for the
// edge row u = 32 c and one for the edge column v = 32 c (the 16 x 16 fragments of int8_device.h,
// which read the template 16 bytes into each row). The frame loop, the screen, the exact candidates,
// the SNR box and the FP64 fit are those of the 16 x 16 kernel (int8_corr.hip), so the outputs are the
// same bits.

#include "common.h"
#include "int8_device.h"

namespace framecor {

namespace {

typedef int v16i __attribute__((ext_vector_type(16)));

// the raw int32 surface of one 32-row band (rows 32 band .. + 31, all NB tiles) of a staged chip
template <int NB, int NJ>
__device__ inline void wideBandMfma(v16i (&acc)[NB], const int8_t *chip, const int8_t *tmpl, int band, int lane,
                                    int ni2, int tstride, int cstride)
{
    const int h = lane >> 5, n = lane & 31;
    // B fragment geometry: lane (h, n) takes template row 2 i2 + h, bytes 32 + 16 jb - n .. + 15 of the
    // zero-padded LDS row, i.e. dwords (32 - n) / 4 + 4 jb .. + 4, funnel-shifted
    const int bbyte = 32 - n, bdw = bbyte >> 2, bsh = bbyte & 3;
#pragma unroll
    for (int v1 = 0; v1 < NB; v1++)
#pragma unroll
        for (int q = 0; q < 16; q++) acc[v1][q] = 0;
    const int8_t *crow = chip + (32 * band + h + n) * cstride;
    const uint32_t *trow = reinterpret_cast<const uint32_t *>(tmpl + h * tstride) + bdw;
    const int cstep = 2 * cstride, tstep = tstride >> 1;   // 2 rows: bytes; dwords per 2 rows
    for (int i2 = 0; i2 < ni2; i2++) {
        v4i b[NJ];
#pragma unroll
        for (int jb = 0; jb < NJ; jb++) {
            const uint32_t *t = trow + 4 * jb;
            const uint32_t w0 = t[0], w1 = t[1], w2 = t[2], w3 = t[3], w4 = t[4];
            b[jb] = v4i{int(__builtin_amdgcn_alignbyte(w1, w0, bsh)), int(__builtin_amdgcn_alignbyte(w2, w1, bsh)),
                        int(__builtin_amdgcn_alignbyte(w3, w2, bsh)), int(__builtin_amdgcn_alignbyte(w4, w3, bsh))};
        }
#pragma unroll
        for (int cb = 0; cb < 2 * (NB - 1) + NJ; cb++) {
            const v4i av = *reinterpret_cast<const v4i *>(crow + 16 * cb);
#pragma unroll
            for (int v1 = 0; v1 < NB; v1++) {
                const int jb = cb - 2 * v1;
                if (jb >= 0 && jb < NJ) acc[v1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, b[jb], acc[v1], 0, 0, 0);
            }
        }
        crow += cstep;
        trow += tstep;
    }
}

// Window k against frames fBegin .. fEnd - 1, by the whole block (NB + 2 waves): the frame loop of the
// 16 x 16 kernel with 32-row bands. NJB: the 16-byte template blocks of the edge fragments.
template <int NB, int NJ, int NJB>
__device__ __attribute__((always_inline)) inline void int8WideWindow(const Int8Args &a, int k, int fBegin, int fEnd,
                                                                     unsigned char *lds)
{
    const Geometry &g = a.g;
    const int tid = threadIdx.x, lane0 = tid & (kWave - 1), wave = tid >> 6, nthreads = blockDim.x;
    const int nw = nthreads / kWave;
    const int W = g.W, M = g.M, L = g.L;
    constexpr int c = 2 * NB;   // the core in 16-point units: the edge row and column are u, v = 16 c
    const Int8Lds sh = int8Lds(a.ni4, a.tstride, a.crows, a.cstride, NJB, M, g.stat, nw);
    int8_t *tmpl = reinterpret_cast<int8_t *>(lds);
    const int8_t *tmpl16 = tmpl + 16;   // the rows as the 16 x 16 fragments read them (data at byte 16)
    const int8_t *zero = tmpl + sh.zero;
    int8_t *chip = tmpl + sh.chip;
    int *boxbuf = reinterpret_cast<int *>(lds + sh.box);
    WaveBest *parts = reinterpret_cast<WaveBest *>(lds + sh.parts);

    loadTemplate(lds, sh.box, tmpl, a.tmpl + size_t(k) * W * W, W, a.tstride, tid, nthreads, 32);
    const float normk = a.wd.norm[k];
    const int y0 = a.wd.down[k / g.nx], x0 = a.wd.across[k % g.nx];
    Moments mom;
    mom.n = int64_t(W) * W;
    mom.tvar = a.wd.tvar[k];
    mom.nd = double(mom.n);
    mom.tau = double(a.wd.tsum[k] - 128 * mom.n);
    mom.rst = mom.tvar > 0 ? float(1.0 / sqrt(double(mom.tvar))) : 0.0f;
    const bool band = wave < NB;
    const int finisher = NB;
    const int ni2 = (W + 1) >> 1;

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
        // the lane's coordinates anew each frame (recomputed rather than kept live across the MFMAs)
        int lane = lane0;
        asm volatile("" : "+v"(lane));
        const int g4 = lane >> 4, m16 = lane & 15, h = lane >> 5, n32 = lane & 31;
        v16i acc[NB];
        v4i es = v4i{0, 0, 0, 0};
        if (band)
            wideBandMfma<NB, NJ>(acc, chip, tmpl, wave, lane, ni2, a.tstride, a.cstride);
        else if (wave == NB)
            es = edgeRowMfma<NJB>(chip, tmpl16, c, lane, a.ni4, a.tstride, a.cstride);
        else
            es = edgeColMfma<NJB>(chip, tmpl16, zero, c, W, lane, a.tstride, a.cstride);

        // this lane's points in ascending index order: fn(raw, u, v)
        auto visit = [&](auto &&fn) {
            if (band) {
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const int u = 32 * wave + 8 * (q >> 2) + 4 * h + (q & 3);
#pragma unroll
                    for (int v1 = 0; v1 < NB; v1++) fn(acc[v1][q], u, 32 * v1 + n32);
                }
            }
            else if (wave == NB) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int blk = 4 * g4 + r, v = 16 * blk + m16;
                    if (blk <= c && v < M) fn(es[r], 16 * c, v);
                }
            }
            else {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int blk = 4 * g4 + r;
                    if (blk < c) fn(es[r], 16 * blk + m16, 16 * c);
                }
            }
        };

        // the screen of this wave's points: its largest screened value and whether it stands alone
        const int64_t boff = f * g.boxSize + int64_t(fw.sy) * g.bw + fw.sx;
        const int2 *scr = a.fg.screen + boff;
        const int32_t *box2 = a.fg.box2 + boff;
        Top2 t = top2Init();
        if (band) {
            // a row of NB points per q: one row pointer, the columns at immediate offsets
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int u = 32 * wave + 8 * (q >> 2) + 4 * h + (q & 3);
                const int2 *row = scr + u * g.bw + n32;
                int2 p[NB];
#pragma unroll
                for (int v1 = 0; v1 < NB; v1++) p[v1] = row[32 * v1];
#pragma unroll
                for (int v1 = 0; v1 < NB; v1++)
                    top2Add(t, screenNcc(acc[v1][q], p[v1].x, __int_as_float(p[v1].y), mom), u * M + 32 * v1 + n32);
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
            if (!band || (32 * wave < b.uhi && 32 * wave + 32 > b.ulo))
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

// blocks of NB + 2 <= 6 waves, three or four per CU by LDS: up to 128 registers a lane (4 waves a SIMD)
template <int NB, int NJ, int NJB>
__global__ void __launch_bounds__(64 * (NB + 2)) __attribute__((amdgpu_waves_per_eu(4))) int8Wide(Int8Args a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    // run of frames major, as the 16 x 16 kernel
    const int logical = xcdRemap(blockIdx.x, gridDim.x);
    const int fblock = logical / a.g.nwin, k = logical - fblock * a.g.nwin;
    const int fBegin = fblock * a.fpb, fEnd = min(fBegin + a.fpb, a.fg.n);
    int8WideWindow<NB, NJ, NJB>(a, k, fBegin, fEnd, lds);
}

// the persistent form over the work queue of a tile shape (clips[c]: clip c's arguments)
template <int NB, int NJ, int NJB>
__global__ void __launch_bounds__(64 * (NB + 2)) __attribute__((amdgpu_waves_per_eu(4)))
int8WideQueue(WorkQueue q, const Int8Args *__restrict__ clips)
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
        int8WideWindow<NB, NJ, NJB>(clips[c], k, fBegin, fEnd, lds);
    }
}

template <int NJ, int NJB>
const void *wideOf(int nb, bool queue)
{
    switch (nb) {
    case 1: return queue ? reinterpret_cast<const void *>(int8WideQueue<1, NJ, NJB>)
                         : reinterpret_cast<const void *>(int8Wide<1, NJ, NJB>);
    case 2: return queue ? reinterpret_cast<const void *>(int8WideQueue<2, NJ, NJB>)
                         : reinterpret_cast<const void *>(int8Wide<2, NJ, NJB>);
    case 3: return queue ? reinterpret_cast<const void *>(int8WideQueue<3, NJ, NJB>)
                         : reinterpret_cast<const void *>(int8Wide<3, NJ, NJB>);
    case 4: return queue ? reinterpret_cast<const void *>(int8WideQueue<4, NJ, NJB>)
                         : reinterpret_cast<const void *>(int8Wide<4, NJ, NJB>);
    default: return nullptr;
    }
}

} // namespace

// the (NJ, NJB) pairs of templates W = 2 .. 129: NJ = ceil((W + 31) / 16), NJB = int8JBlocks(W)
const void *int8WideKernel(int nb, int nj, int njb, bool queue)
{
    switch (nj * 16 + njb) {
    case 3 * 16 + 2: return wideOf<3, 2>(nb, queue);     // W 2..17
    case 4 * 16 + 3: return wideOf<4, 3>(nb, queue);     // W 18..33
    case 5 * 16 + 5: return wideOf<5, 5>(nb, queue);     // W 34..49
    case 6 * 16 + 5: return wideOf<6, 5>(nb, queue);     // W 50..65
    case 7 * 16 + 9: return wideOf<7, 9>(nb, queue);     // W 66..81
    case 8 * 16 + 9: return wideOf<8, 9>(nb, queue);     // W 82..97
    case 9 * 16 + 9: return wideOf<9, 9>(nb, queue);     // W 98..113
    case 10 * 16 + 9: return wideOf<10, 9>(nb, queue);   // W 114..129
    default: return nullptr;
    }
}

} // namespace framecor
