// int8_quad.hip -- the int8-MFMA form in blocks of four wavefronts with equal MFMA shares
//
// The same exact integer contraction, screen, candidates, SNR box and FP64 fit as the one-frame kernel
// of int8_corr.hip (so the outputs are the same bits), for surfaces of M = 16 c + 1 points (search
// ranges R = 8 c) with c = 4 or 6 and templates of 32 or 64 columns. There a block has c band waves and
// two light edge waves, so two of the four SIMDs of a CU carry twice the MFMA work of the other two.
// Here a block has four waves, one per SIMD, and the c x c core tiles plus the edge row and column are
// dealt so that every wave carries the same MFMA work to within one tile:
//   c = 4: wave w the band w; the edge row to wave 0, the edge column to wave 1
//   c = 6: wave w the band w and tiles 3 (w & 1) .. + 2 of band 4 + (w >> 1); edges to waves 0 and 1
// A wave's two band segments and the edge row share the template's B fragments, built once per row
// group of four; the edge column runs on independent MFMA chains with its extent known at compile time.
// The LDS holds the chip's L rows only and template rows of the narrowest conflict-free stride (112
// bytes at window 64), 37 KB at window 64, search 48: four blocks run per CU, a SIMD holds one wave of
// each, and the blocks are independent, so the per-frame screen and peak fit of one overlap the MFMAs
// of the others. The peak fit of a frame alternates between waves 2 and 3, which carry no edge.

#include "common.h"
#include "int8_device.h"

namespace framecor {

namespace {

constexpr int kQuadWaves = 4;

// the second band segment of wave w (c = 6: three tiles; c = 4: none)
template <int C>
struct QuadShare {
    static constexpr int T2 = C == 6 ? 3 : 0;
    __device__ static int band2(int w) { return 4 + (w >> 1); }
    __device__ static int tile2(int w) { return 3 * (w & 1); }
};

// The raw int32 surfaces of a wave's segments: the NT1 tiles of band b1, the NT2 tiles of band b2 from
// tile t2, and (EROW) the edge row u = 16 c, all over every template row group with one set of B
// fragments per row group (the bands' A and B as bandMfma, the edge row's A as edgeRowMfma)
template <int NT1, int NT2, int NJB, bool EROW>
__device__ inline void quadMfma(v4i (&acc1)[NT1], v4i (&acc2)[NT2 > 0 ? NT2 : 1], v4i &er, const int8_t *chip,
                                const int8_t *tmpl, int b1, int b2, int t2, int c, int lane, int ni4, int tstride,
                                int cstride)
{
    const int g4 = lane >> 4, m16 = lane & 15;
    const int bbyte = 16 - m16, bdw = bbyte >> 2, bsh = bbyte & 3;
#pragma unroll
    for (int v1 = 0; v1 < NT1; v1++) acc1[v1] = v4i{0, 0, 0, 0};
#pragma unroll
    for (int v1 = 0; v1 < (NT2 > 0 ? NT2 : 1); v1++) acc2[v1] = v4i{0, 0, 0, 0};
    er = v4i{0, 0, 0, 0};
    const int8_t *crow1 = chip + (16 * b1 + g4 + m16) * cstride;
    const int8_t *crow2 = chip + (16 * b2 + g4 + m16) * cstride + 16 * t2;
    const int8_t *crowE = chip + (16 * c + g4) * cstride + 16 * min(m16, c);
    const uint32_t *trow = reinterpret_cast<const uint32_t *>(tmpl + g4 * tstride) + bdw;
    const int cstep = 4 * cstride, tstep = tstride;   // 4 rows: bytes; dwords per 4 rows
    for (int i4 = 0; i4 < ni4; i4++) {
        v4i b[NJB];
#pragma unroll
        for (int jb = 0; jb < NJB; jb++) {
            const uint32_t *t = trow + 4 * jb;
            const uint32_t w0 = t[0], w1 = t[1], w2 = t[2], w3 = t[3], w4 = t[4];
            b[jb] = v4i{int(__builtin_amdgcn_alignbyte(w1, w0, bsh)), int(__builtin_amdgcn_alignbyte(w2, w1, bsh)),
                        int(__builtin_amdgcn_alignbyte(w3, w2, bsh)), int(__builtin_amdgcn_alignbyte(w4, w3, bsh))};
        }
#pragma unroll
        for (int cb = 0; cb < NT1 + NJB - 1; cb++) {
            const v4i av = *reinterpret_cast<const v4i *>(crow1 + 16 * cb);
#pragma unroll
            for (int jb = 0; jb < NJB; jb++) {
                const int v1 = cb - jb;
                if (v1 >= 0 && v1 < NT1) acc1[v1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, b[jb], acc1[v1], 0, 0, 0);
            }
        }
        if constexpr (NT2 > 0) {
#pragma unroll
            for (int cb = 0; cb < NT2 + NJB - 1; cb++) {
                const v4i av = *reinterpret_cast<const v4i *>(crow2 + 16 * cb);
#pragma unroll
                for (int jb = 0; jb < NJB; jb++) {
                    const int v1 = cb - jb;
                    if (v1 >= 0 && v1 < NT2)
                        acc2[v1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, b[jb], acc2[v1], 0, 0, 0);
                }
            }
        }
        if constexpr (EROW) {
#pragma unroll
            for (int jb = 0; jb < NJB; jb++) {
                const v4i av = *reinterpret_cast<const v4i *>(crowE + 16 * jb);
                er = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, b[jb], er, 0, 0, 0);
            }
        }
        crow1 += cstep;
        crow2 += cstep;
        crowE += cstep;
        trow += tstep;
    }
}

// the FP32 screen of NT tiles of band `band` from tile t0: rows in ascending order, columns ascending
// within a row (top2Add's order); up to four tiles with the loads of all four rows issued first
template <int NT>
__device__ inline void screenTiles(Top2 &t, const v4i (&acc)[NT], const int2 *scr, int bw, int M, int band, int t0,
                                   int g4, int m16, const Moments &mom)
{
    if constexpr (NT <= 4) {
        // every row's map loads in flight together (at most 16 a lane), then the points in order
        int2 p[4][NT];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int2 *row = scr + (16 * band + 4 * g4 + r) * bw + 16 * t0 + m16;
#pragma unroll
            for (int v1 = 0; v1 < NT; v1++) p[r][v1] = row[16 * v1];
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int u = 16 * band + 4 * g4 + r;
#pragma unroll
            for (int v1 = 0; v1 < NT; v1++)
                top2Add(t, screenNcc(acc[v1][r], p[r][v1].x, __int_as_float(p[r][v1].y), mom),
                        u * M + 16 * (t0 + v1) + m16);
        }
    }
    else {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int u = 16 * band + 4 * g4 + r;
            const int2 *row = scr + u * bw + 16 * t0 + m16;
            int2 p[NT];
#pragma unroll
            for (int v1 = 0; v1 < NT; v1++) p[v1] = row[16 * v1];
#pragma unroll
            for (int v1 = 0; v1 < NT; v1++)
                top2Add(t, screenNcc(acc[v1][r], p[v1].x, __int_as_float(p[v1].y), mom), u * M + 16 * (t0 + v1) + m16);
        }
    }
}

// The edge column v = 16 c (rows 0 .. 16 c - 1) as edgeColMfma (int8_device.h) computes it, for a template
// of W = 4 NI4 columns: the column blocks and row groups known at compile time, one accumulator per
// column block (independent MFMA chains, summed exactly in int32 at the end) and every LDS read of a
// row group issued ahead of its MFMAs
template <int NI4>
__device__ inline v4i edgeColQuad(const int8_t *chip, const int8_t *tmpl, const int8_t *zero, int c, int lane,
                                  int tstride, int cstride)
{
    constexpr int W = 4 * NI4, NJ = (W + 15) >> 4, NQ = (W + 18) >> 2;
    const int g4 = lane >> 4, m16 = lane & 15;
    v4i acc[NJ];
#pragma unroll
    for (int jb = 0; jb < NJ; jb++) acc[jb] = v4i{0, 0, 0, 0};
    const int8_t *abase = chip + 16 * min(m16, c - 1) * cstride + 16 * c;
#pragma unroll 2
    for (int q = 0; q < NQ; q++) {
        const int ip = 4 * q + g4, ti = ip - m16;
        const bool inside = ip < W + 15 && ti >= 0 && ti < W;
        const int8_t *arow = abase + (ip < W + 15 ? ip : 0) * cstride;
        const int8_t *brow = inside ? tmpl + ti * tstride + 16 : zero;
        v4i av[NJ], bv[NJ];
#pragma unroll
        for (int jb = 0; jb < NJ; jb++) {
            av[jb] = *reinterpret_cast<const v4i *>(arow + 16 * jb);
            bv[jb] = *reinterpret_cast<const v4i *>(brow + 16 * jb);
        }
#pragma unroll
        for (int jb = 0; jb < NJ; jb++) acc[jb] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[jb], bv[jb], acc[jb], 0, 0, 0);
    }
#pragma unroll
    for (int jb = 1; jb < NJ; jb++) acc[0] += acc[jb];
    return acc[0];
}

// top2Add for points in any order: of equal values the lower index is the largest, and an equal value
// of either index counts as a second candidate
__device__ inline void top2AddAny(Top2 &t, float x, int idx)
{
    const bool win = x > t.f1 || (x == t.f1 && idx < t.i1);
    t.f2 = win ? t.f1 : fmaxf(t.f2, x);
    t.f1 = win ? x : t.f1;
    t.i1 = win ? idx : t.i1;
}

// Window k against frames fBegin .. fEnd - 1, by a block of four waves (C core bands, M = 16 C + 1).
// Per frame: every wave runs its MFMAs into registers and screens its own points there (FP32 screen)
// into parts[wave]; barrier B2; every wave takes the block's peak from parts (blockPeak: exact FP64
// values only at near-ties) and writes its raw values inside the SNR box around it to LDS, and the chip
// of the next valid frame is staged; barrier B3; then the finisher (wave 2 or 3 on alternate frames)
// writes the outputs while the other waves start the next frame's MFMAs.
template <int C, int NJB, int NI4>
__device__ __attribute__((always_inline)) inline void quadWindow(const Int8Args &a, int k, int fBegin, int fEnd,
                                                                 unsigned char *lds)
{
    using Share = QuadShare<C>;
    constexpr int NT2 = Share::T2;
    const Geometry &g = a.g;
    const int tid = threadIdx.x, lane0 = tid & (kWave - 1), wave = tid >> 6;
    constexpr int nthreads = kWave * kQuadWaves;
    const int W = g.W, M = g.M, L = g.L;
    const Int8Lds sh = int8Lds(a.ni4, a.tstride, a.crows, a.cstride, NJB, M, g.stat, kQuadWaves);
    int8_t *tmpl = reinterpret_cast<int8_t *>(lds);
    const int8_t *zero = tmpl + sh.zero;
    int8_t *chip = tmpl + sh.chip;
    int *boxbuf = reinterpret_cast<int *>(lds + sh.box);
    WaveBest *parts = reinterpret_cast<WaveBest *>(lds + sh.parts);
    // the raw values of the edge row (v = 0 .. M - 1) and the edge column (u = 0 .. M - 2) of the frame,
    // kept in LDS by the wave that computes them rather than in registers of every wave
    int *eRow = reinterpret_cast<int *>(lds + ((sh.total + 15) & ~15)), *eCol = eRow + M;

    loadTemplate(lds, sh.box, tmpl, a.tmpl + size_t(k) * W * W, W, a.tstride, tid, nthreads);
    const float normk = a.wd.norm[k];
    const int y0 = a.wd.down[k / g.nx], x0 = a.wd.across[k % g.nx];
    Moments mom;
    mom.n = int64_t(W) * W;
    mom.tvar = a.wd.tvar[k];
    mom.nd = double(mom.n);
    mom.tau = double(a.wd.tsum[k] - 128 * mom.n);
    mom.rst = mom.tvar > 0 ? float(1.0 / sqrt(double(mom.tvar))) : 0.0f;
    const bool erow = wave == 0, ecol = wave == 1;
    const int b1 = wave, b2 = Share::band2(wave), t2 = Share::tile2(wave);

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

    int finisher = 2;
    while (f < fEnd) {
        // the lane's coordinates anew each frame (recomputed rather than kept live across the MFMAs)
        int lane = lane0;
        asm volatile("" : "+v"(lane));
        const int g4 = lane >> 4, m16 = lane & 15;
        v4i acc1[C], acc2[NT2 > 0 ? NT2 : 1];
        if (erow) {
            v4i er;
            quadMfma<C, NT2, NJB, true>(acc1, acc2, er, chip, tmpl, b1, b2, t2, C, lane, a.ni4, a.tstride, a.cstride);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int blk = 4 * g4 + r, v = 16 * blk + m16;
                if (blk <= C && v < M) eRow[v] = er[r];
            }
        }
        else {
            v4i er;
            quadMfma<C, NT2, NJB, false>(acc1, acc2, er, chip, tmpl, b1, b2, t2, C, lane, a.ni4, a.tstride, a.cstride);
        }
        if (ecol) {
            const v4i ec = edgeColQuad<NI4>(chip, tmpl, zero, C, lane, a.tstride, a.cstride);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int blk = 4 * g4 + r;
                if (blk < C) eCol[16 * blk + m16] = ec[r];
            }
        }

        // this lane's points: fn(raw, u, v)
        auto visit = [&](auto &&fn) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int u = 16 * b1 + 4 * g4 + r;
#pragma unroll
                for (int v1 = 0; v1 < C; v1++) fn(acc1[v1][r], u, 16 * v1 + m16);
            }
            if constexpr (NT2 > 0) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int u = 16 * b2 + 4 * g4 + r;
#pragma unroll
                    for (int v1 = 0; v1 < NT2; v1++) fn(acc2[v1][r], u, 16 * (t2 + v1) + m16);
                }
            }
            if (erow) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int blk = 4 * g4 + r, v = 16 * blk + m16;
                    if (blk <= C && v < M) fn(eRow[v], 16 * C, v);
                }
            }
            if (ecol) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int blk = 4 * g4 + r, u = 16 * blk + m16;
                    if (blk < C) fn(eCol[u], u, 16 * C);
                }
            }
        };

        // the screen of this wave's points (ascending index within each list): its largest screened value
        // and whether it stands alone
        const int64_t boff = f * g.boxSize + int64_t(fw.sy) * g.bw + fw.sx;
        const int2 *scr = a.fg.screen + boff;
        const int32_t *box2 = a.fg.box2 + boff;
        Top2 t = top2Init();
        screenTiles<C>(t, acc1, scr, g.bw, M, b1, 0, g4, m16, mom);
        if constexpr (NT2 > 0) screenTiles<NT2>(t, acc2, scr, g.bw, M, b2, t2, g4, m16, mom);
        if (erow) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int blk = 4 * g4 + r, v = 16 * blk + m16;
                if (blk <= C && v < M) {
                    const int2 p = scr[16 * C * g.bw + v];
                    const int raw = eRow[v];
                    top2Add(t, screenNcc(raw, p.x, __int_as_float(p.y), mom), 16 * C * M + v);
                }
            }
        }
        if (ecol) {
            // the edge column's rows interleave with the bands' (not in ascending index order)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int blk = 4 * g4 + r, u = 16 * blk + m16;
                if (blk < C) {
                    const int2 p = scr[u * g.bw + 16 * C];
                    const int raw = eCol[u];
                    top2AddAny(t, screenNcc(raw, p.x, __int_as_float(p.y), mom), u * M + 16 * C);
                }
            }
        }
        screenPart(t, parts + wave, lane);
        __syncthreads();   // B2: every MFMA of this frame is over; parts are in

        // the block's peak (exact FP64 values only at near-ties); this wave's raw values inside its SNR box
        int ib;
        const bool valid = blockPeak(parts, kQuadWaves, wave, lane, mom, scr, box2, g.bw, M, visit, ib);
        if (valid) {
            const SnrBox bx = snrBox(ib, M, g.stat);
            auto put = [&](int raw, int u, int v) {
                if (u >= bx.ulo && u < bx.uhi && v >= bx.vlo && v < bx.vhi) boxbuf[(u - bx.ulo) * bx.bwid + v - bx.vlo] = raw;
            };
            // only the 16 x 16 tiles that meet the box (wave-uniform tests)
            if (16 * b1 < bx.uhi && 16 * b1 + 16 > bx.ulo) {
#pragma unroll
                for (int v1 = 0; v1 < C; v1++)
                    if (16 * v1 < bx.vhi && 16 * v1 + 16 > bx.vlo)
#pragma unroll
                        for (int r = 0; r < 4; r++) put(acc1[v1][r], 16 * b1 + 4 * g4 + r, 16 * v1 + m16);
            }
            if constexpr (NT2 > 0) {
                if (16 * b2 < bx.uhi && 16 * b2 + 16 > bx.ulo) {
#pragma unroll
                    for (int v1 = 0; v1 < NT2; v1++)
                        if (16 * (t2 + v1) < bx.vhi && 16 * (t2 + v1) + 16 > bx.vlo)
#pragma unroll
                            for (int r = 0; r < 4; r++) put(acc2[v1][r], 16 * b2 + 4 * g4 + r, 16 * (t2 + v1) + m16);
                }
            }
            if (erow && bx.uhi > 16 * C) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int blk = 4 * g4 + r, v = 16 * blk + m16;
                    if (blk <= C && v < M) put(eRow[v], 16 * C, v);
                }
            }
            if (ecol && bx.vhi > 16 * C) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int blk = 4 * g4 + r;
                    if (blk < C) put(eCol[16 * blk + m16], 16 * blk + m16, 16 * C);
                }
            }
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
        finisher ^= 1;     // 2, 3, 2, ...
    }
}

// four blocks of four waves per CU (the LDS of the 720p shape): up to 128 registers a lane
template <int C, int NJB, int NI4>
__global__ void __launch_bounds__(kWave * kQuadWaves) __attribute__((amdgpu_waves_per_eu(4))) int8Quad(Int8Args a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    // run of frames major, as the one-frame kernel
    const int logical = xcdRemap(blockIdx.x, gridDim.x);
    const int fblock = logical / a.g.nwin, k = logical - fblock * a.g.nwin;
    const int fBegin = fblock * a.fpb, fEnd = min(fBegin + a.fpb, a.fg.n);
    quadWindow<C, NJB, NI4>(a, k, fBegin, fEnd, lds);
}

// One work item of the persistent form of c = 6, as a function of its own. Inlined into the item loop
// of the persistent kernel at 128 registers, the code of c = 6 came out wrong (every window of a
// window-64, search-48 batch differed from the one-clip kernel, with no fault), while the same source
// compiled with this call boundary, or with 168 registers, gives the one-clip kernel's bits; c = 4 is
// inlined. The batching tests cover both.
template <int C, int NJB, int NI4>
__device__ __attribute__((noinline)) void quadItem(const Int8Args &a, int k, int fBegin, int fEnd, unsigned char *lds)
{
    quadWindow<C, NJB, NI4>(a, k, fBegin, fEnd, lds);
}

// the persistent form over the work queue of a tile shape (clips[c]: clip c's arguments)
template <int C, int NJB, int NI4>
__global__ void __launch_bounds__(kWave * kQuadWaves) __attribute__((amdgpu_waves_per_eu(4)))
int8QuadQueue(WorkQueue q, const Int8Args *__restrict__ clips)
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
        if constexpr (C == 6)
            quadItem<C, NJB, NI4>(clips[c], k, fBegin, fEnd, lds);
        else
            quadWindow<C, NJB, NI4>(clips[c], k, fBegin, fEnd, lds);
    }
}

template <int C, int NJB, int NI4>
const void *quadOf(bool queue)
{
    return queue ? reinterpret_cast<const void *>(int8QuadQueue<C, NJB, NI4>)
                 : reinterpret_cast<const void *>(int8Quad<C, NJB, NI4>);
}

} // namespace

int int8QuadThreads() { return kWave * kQuadWaves; }

size_t int8QuadShmem(int ni4, int tstride, int crows, int cstride, int njb, const Geometry &g)
{
    // the one-frame layout with four results, then the edge row and column values
    const int total = int8Lds(ni4, tstride, crows, cstride, njb, g.M, g.stat, kQuadWaves).total;
    return size_t(((total + 15) & ~15) + 8 * g.M);
}

// c core bands (M = 16 c + 1) and a template of W = 32 or 64 columns (njb blocks); null where none is
// instantiated
const void *int8QuadKernel(int c, int W, int njb, bool queue)
{
    if (W == 64 && njb == 5) {
        if (c == 4) return quadOf<4, 5, 16>(queue);
        if (c == 6) return quadOf<6, 5, 16>(queue);
    }
    if (W == 32 && njb == 3) {
        if (c == 4) return quadOf<4, 3, 8>(queue);
        if (c == 6) return quadOf<6, 3, 8>(queue);
    }
    return nullptr;
}

} // namespace framecor
