// int8_device.h -- device helpers shared by the int8-MFMA kernels (int8_corr.hip, int8_wide.hip):
// the exact and screened NCC of a raw int32 correlation, the two-largest screen, the 16 x 16 MFMA
// fragments of a band and of the edge row and column, the staging of the template and the chip in LDS,
// the SNR box and the FP64 peak fit of a frame
#pragma once

#include "common.h"
#include "epilogue.h"
#include "kernels.h"

namespace framecor {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kPipedWaves = 12;   // producers + helpers per block of the pipelined kernel
constexpr int kEdgeChains = 3;    // independent accumulators of the edge row and column waves

struct WaveBest {
    double value;
    int index;
    int pad;
};

struct Moments {
    int64_t n;      // W^2
    int64_t tvar;   // W^2 sum T^2 - (sum T)^2
    double nd;      // W^2
    double tau;     // sum (T - 128) = sum T - 128 W^2
    float rst;      // 1 / sqrt(tvar)
};

// the exact numerator W^2 sum(FT) - sum T sum F = W^2 r - tau phi, phi = sum F - 128 W^2: an integer
// below 2^53, so the FP64 product tau phi and the fused multiply-add are exact
__device__ inline double numerator(int raw, int fs, const Moments &m)
{
    const double phi = double(fs - 128 * int(m.n));
    return fma(m.nd, double(raw), -(m.tau * phi));
}

// the exact NCC in FP64 (the same value as the direct oracle's); 0 where the search box (or the
// template) is flat, as cxx/surface.h:30-35
__device__ inline double exactNcc(int raw, int fs, int fs2, const Moments &m)
{
    const int64_t vf = m.n * int64_t(fs2) - int64_t(fs) * int64_t(fs);
    if (vf <= 0 || m.tvar <= 0) return 0.0;
    return numerator(raw, fs, m) / sqrt(double(m.tvar) * double(vf));
}

// the same to FP32 accuracy (a relative error of a few 1e-7), with the frame's 1 / sqrt(vf) map
__device__ inline float screenNcc(int raw, int fs, float rsig, const Moments &m)
{
    return float(numerator(raw, fs, m)) * rsig * m.rst;
}

// a lane's two largest screened values and the index (u M + v) of the largest; equal values: the lower
// index
struct Top2 {
    float f1, f2;
    int i1;
};

__device__ inline Top2 top2Init() { return Top2{-3.0f, -3.0f, 0x7fffffff}; }

// A lane adds its points in ascending index order, so of equal values the first stays the largest and
// only a strictly larger value replaces it. Branch-free: the second largest is the larger of the old
// second and the smaller of (x, the old largest).
__device__ inline void top2Add(Top2 &t, float x, int idx)
{
    const bool win = x > t.f1;
    t.f2 = fmaxf(t.f2, fminf(x, t.f1));
    t.f1 = fmaxf(t.f1, x);
    t.i1 = win ? idx : t.i1;
}

// The screen's bound: the FP32 screen value of a point is within a few 1e-7 (relative) of its exact NCC,
// so a point screened below screenThreshold(the largest screened value) is not the exact maximum.
__device__ inline float screenThreshold(float fmax) { return fmax - 1e-5f * fabsf(fmax) - 1e-30f; }

// A wave's screen result for its block, in *part: its largest screened value and that point's index, and
// (pad = 1) whether no other point of the wave is screened within the bound of it.
__device__ inline void screenPart(const Top2 &t, WaveBest *part, int lane)
{
    float fmax = t.f1;
    int imax = t.i1;
    waveBest(fmax, imax);
    const float thr = screenThreshold(fmax);
    const bool single = __popcll(__ballot(t.f1 >= thr)) == 1 && !__any(t.f2 >= thr);
    if (lane == 0) {
        part->value = fmax;
        part->index = imax;
        part->pad = single;
    }
}

// The block's peak from its nw waves' screen parts (read behind the barrier that follows them). When one
// point is screened above the bound of every other point of the block, it is the maximum of the exact
// FP64 surface and no FP64 value is needed here. Otherwise (near-ties) every wave takes the exact NCC of
// each of its points screened within the bound (visit(fn) calls fn(raw, u, v) for every point of the
// wave) and the block the largest of them, the first in row-major order on equal values, behind two more
// barriers: every wave decides from the same parts, so all of them take these or none. Returns whether
// the peak is positive (else the frame's outputs are invalid); *ib is its index u M + v.
template <typename Visit>
__device__ inline bool blockPeak(WaveBest *parts, int nw, int wave, int lane, const Moments &mom, const int2 *scr,
                                 const int32_t *box2, int bw, int M, Visit &&visit, int &ib)
{
    double top = -3.0;
    int itop = 0x7fffffff, wtop = 0;
    for (int w = 0; w < nw; w++) {
        const double v = parts[w].value;
        const int i = parts[w].index;
        if (v > top || (v == top && i < itop)) {
            top = v;
            itop = i;
            wtop = w;
        }
    }
    const float thr = screenThreshold(float(top));
    bool clear = parts[wtop].pad != 0;
    for (int w = 0; w < nw; w++) clear = clear && (w == wtop || parts[w].value < thr);
    if (clear) {
        ib = itop;
        return top > 0;
    }
    double dbest = -3.0;
    int dib = 0x7fffffff;
    visit([&](int raw, int u, int v) {
        // each point's FP64 numerator recomputed here from a raw value the compiler cannot match to the
        // screen's: else it keeps every point's numerator live from the screen to this rarely taken path
        // (two registers a point: the four-wave kernel's c = 4 form spilled 72 bytes a lane, none now)
        asm volatile("" : "+v"(raw));
        const int o = u * bw + v;
        const int2 p = scr[o];
        if (screenNcc(raw, p.x, __int_as_float(p.y), mom) >= thr)
            bestOf(dbest, dib, exactNcc(raw, p.x, box2[o], mom), u * M + v);
    });
    waveBest(dbest, dib);
    __syncthreads();   // every wave has read the screen parts
    if (lane == 0) {
        parts[wave].value = dbest;
        parts[wave].index = dib;
    }
    __syncthreads();
    double best = -3.0;
    ib = 0x7fffffff;
    for (int w = 0; w < nw; w++) bestOf(best, ib, parts[w].value, parts[w].index);
    return best > 0;
}

// the raw int32 surface of one 16-row band (rows 16 band .. + 15, all NVT v tiles) of a staged chip
template <int NVT, int NJB>
__device__ inline void bandMfma(v4i (&acc)[NVT], const int8_t *chip, const int8_t *tmpl, int band, int lane, int ni4,
                                int tstride, int cstride)
{
    const int g4 = lane >> 4, m16 = lane & 15;
    // B fragment geometry: lane (g4, n = m16) takes template row 4 i4 + g4, bytes 16 + 16 jb - n .. + 15
    // of the zero-padded LDS row, i.e. dwords (16 - n) / 4 + 4 jb .. + 4, funnel-shifted
    const int bbyte = 16 - m16, bdw = bbyte >> 2, bsh = bbyte & 3;
#pragma unroll
    for (int v1 = 0; v1 < NVT; v1++) acc[v1] = v4i{0, 0, 0, 0};
    const int8_t *crow = chip + (16 * band + g4 + m16) * cstride;
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
        for (int cb = 0; cb < NVT + NJB - 1; cb++) {
            const v4i av = *reinterpret_cast<const v4i *>(crow + 16 * cb);
#pragma unroll
            for (int jb = 0; jb < NJB; jb++) {
                const int v1 = cb - jb;
                if (v1 >= 0 && v1 < NVT) acc[v1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, b[jb], acc[v1], 0, 0, 0);
            }
        }
        crow += cstep;
        trow += tstep;
    }
}

// The edge row and column of a surface of M = 16 c + 1 points (every search range R = 8 c):
// the c x c core tiles come from c bands of c tiles; the row u = 16 c and the column v = 16 c are
// computed here instead of as a 17th band and a 17th tile column (16 rows or columns computed to keep
// one). Every LDS read is unconditional (lanes whose outputs are discarded read clamped addresses), so
// the loads of an iteration are issued together ahead of its MFMAs.
//
// The edge row: the v blocks (0..c) on the MFMA's M axis, the template Toeplitz on N as in the bands:
//   A[m][(i, jb, kk)] = F[16 c + i][16 (m + jb) + kk],  B = the bands' fragments,  C[m][n] = r[16 c][16 m + n]
// (rows m > c of C are discarded; their lanes read the chip at m = c).
template <int NJB>
__device__ inline v4i edgeRowMfma(const int8_t *chip, const int8_t *tmpl, int c, int lane, int ni4, int tstride,
                                  int cstride)
{
    // kEdgeChains independent MFMA chains (a chain issues an MFMA only once the previous one has
    // finished), column block jb into chain jb % kEdgeChains, summed exactly in int32 at the end
    v4i acc[kEdgeChains];
#pragma unroll
    for (int k = 0; k < kEdgeChains; k++) acc[k] = v4i{0, 0, 0, 0};
    const int g4 = lane >> 4, m16 = lane & 15;
    const int bbyte = 16 - m16, bdw = bbyte >> 2, bsh = bbyte & 3;
    const int8_t *crow = chip + (16 * c + g4) * cstride + 16 * min(m16, c);
    const uint32_t *trow = reinterpret_cast<const uint32_t *>(tmpl + g4 * tstride) + bdw;
    for (int i4 = 0; i4 < ni4; i4++) {
        v4i av[NJB], b[NJB];
#pragma unroll
        for (int jb = 0; jb < NJB; jb++) {
            const uint32_t *t = trow + 4 * jb;
            const uint32_t w0 = t[0], w1 = t[1], w2 = t[2], w3 = t[3], w4 = t[4];
            b[jb] = v4i{int(__builtin_amdgcn_alignbyte(w1, w0, bsh)), int(__builtin_amdgcn_alignbyte(w2, w1, bsh)),
                        int(__builtin_amdgcn_alignbyte(w3, w2, bsh)), int(__builtin_amdgcn_alignbyte(w4, w3, bsh))};
            av[jb] = *reinterpret_cast<const v4i *>(crow + 16 * jb);
        }
#pragma unroll
        for (int jb = 0; jb < NJB; jb++)
            acc[jb % kEdgeChains] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[jb], b[jb], acc[jb % kEdgeChains], 0, 0, 0);
        crow += 4 * cstride;
        trow += tstride;
    }
#pragma unroll
    for (int k = 1; k < kEdgeChains; k++) acc[0] += acc[k];
    return acc[0];
}

// The edge column: a Toeplitz expansion of the template along its columns' transpose, the u blocks
// (0..c-1) on M and the 16 rows of a block on N:
//   A[m][(i', jb, kk)] = F[16 m + i'][16 c + 16 jb + kk],  B[(i', jb, kk)][n] = T[i' - n][16 jb + kk],
//   C[m][n] = sum over i' in [0, W + 15), jb in [0, ceil(W / 16))  -> r[16 m + n][16 c]
// i' is taken four at a time (i' = 4 q + g4, one per 16-lane group), jb inner (immediate LDS offsets).
// Rows m >= c of C are discarded (their lanes read the chip at m = c - 1); a B fragment outside the
// template (i' - n outside 0..W-1) reads the zero block (16 NJB zero bytes).
template <int NJB>
__device__ inline v4i edgeColMfma(const int8_t *chip, const int8_t *tmpl, const int8_t *zero, int c, int W, int lane,
                                  int tstride, int cstride)
{
    v4i acc[kEdgeChains];   // independent chains, as the edge row
#pragma unroll
    for (int k = 0; k < kEdgeChains; k++) acc[k] = v4i{0, 0, 0, 0};
    const int g4 = lane >> 4, m16 = lane & 15;
    const int nJ = (W + 15) >> 4, nq = (W + 18) >> 2;
    const int8_t *abase = chip + 16 * min(m16, c - 1) * cstride + 16 * c;
    for (int q = 0; q < nq; q++) {
        const int ip = 4 * q + g4, ti = ip - m16;
        const bool inside = ip < W + 15 && ti >= 0 && ti < W;
        const int8_t *arow = abase + (ip < W + 15 ? ip : 0) * cstride;
        const int8_t *brow = inside ? tmpl + ti * tstride + 16 : zero;
        v4i av[NJB], bv[NJB];
#pragma unroll
        for (int jb = 0; jb < NJB; jb++)
            if (jb < nJ) {
                av[jb] = *reinterpret_cast<const v4i *>(arow + 16 * jb);
                bv[jb] = *reinterpret_cast<const v4i *>(brow + 16 * jb);
            }
#pragma unroll
        for (int jb = 0; jb < NJB; jb++)
            if (jb < nJ)
                acc[jb % kEdgeChains] =
                    __builtin_amdgcn_mfma_i32_16x16x64_i8(av[jb], bv[jb], acc[jb % kEdgeChains], 0, 0, 0);
    }
#pragma unroll
    for (int k = 1; k < kEdgeChains; k++) acc[0] += acc[k];
    return acc[0];
}

// the chip of a frame (L x L at (sy, sx)) as int8 (pixels ^ 0x80 = pixels - 128) into LDS, by threads
// t0 .. t0 + nt - 1 of the block (t: this thread's index among them)
__device__ inline void loadChip(int8_t *chip, const uint8_t *src, int w, int L, int cstride, int t, int nt)
{
    // thread t takes the 16-byte column chunk t % cpr of rows t / cpr, t / cpr + nt / cpr, ... (cpr <= nt)
    const int cpr = L >> 4;
    if (cpr > 0) {
        const int rowsPer = nt / cpr, r0 = t / cpr, cc = t - r0 * cpr;
        if (r0 < rowsPer) {
            const uint8_t *s = src + size_t(r0) * w + 16 * cc;
            int8_t *d = chip + r0 * cstride + 16 * cc;
            const size_t sstep = size_t(rowsPer) * w;
            const int dstep = rowsPer * cstride;
            for (int r = r0; r < L; r += rowsPer) {
                uint4 v;
                __builtin_memcpy(&v, s, 16);
                v.x ^= 0x80808080u;
                v.y ^= 0x80808080u;
                v.z ^= 0x80808080u;
                v.w ^= 0x80808080u;
                *reinterpret_cast<uint4 *>(d) = v;
                s += sstep;
                d += dstep;
            }
        }
    }
    const int tail = L & 15;
    if (tail)
        for (int c = t; c < L * tail; c += nt) {
            const int r = c / tail, j = 16 * cpr + (c - r * tail);
            chip[r * cstride + j] = int8_t(src[size_t(r) * w + j] ^ 0x80);
        }
}

// zero the template and chip areas (the padding stays zero), then the template as int8 rows at byte pad
// of each row (16: the 16 x 16 fragments, shifts n < 16; 32: the 32 x 32 fragments)
__device__ inline void loadTemplate(unsigned char *lds, int zeroBytes, int8_t *tmpl, const uint8_t *T, int W,
                                    int tstride, int tid, int nthreads, int pad = 16)
{
    for (int i = tid * 16; i < zeroBytes; i += nthreads * 16) *reinterpret_cast<v4i *>(lds + i) = v4i{0, 0, 0, 0};
    __syncthreads();
    for (int idx = tid; idx < W * W; idx += nthreads) {
        const int i = idx / W, j = idx - i * W;
        tmpl[i * tstride + pad + j] = int8_t(T[idx] ^ 0x80);
    }
}

// the window k of frame f: its gross offset, search window and validity (uniform over the block)
struct FrameWindow {
    int64_t q;
    int gy, gx, sy, sx;
    bool valid;
};

__device__ inline FrameWindow frameWindow(const Int8Args &a, int k, int f, int y0, int x0, float normk)
{
    FrameWindow fw;
    fw.q = int64_t(f) * a.g.nwin + k;
    fw.gy = 0;
    fw.gx = 0;
    if (a.fg.gross) {
        fw.gy = a.fg.gross[2 * fw.q];
        fw.gx = a.fg.gross[2 * fw.q + 1];
    }
    fw.valid = normk > 0 && searchWindow(a.g, y0, x0, fw.gy, fw.gx, fw.sy, fw.sx);
    return fw;
}

// floor(x / d) for 0 <= x < 2^20, 1 <= d < 2^12 by a multiply-high with m = ceil(2^32 / d) (d > 1)
__device__ inline int quot(int x, unsigned m, int d) { return d == 1 ? x : int(__umulhi(unsigned(x), m)); }
__device__ inline unsigned quotMagic(int d) { return unsigned((0x100000000ull + d - 1) / unsigned(d)); }

constexpr int kBatch = 8;   // surface points per lane whose loads are in flight together

// helper wave h of nh: the rows h, h + nh, ... of the raw surface S of frame f: the FP32 screen, then
// the exact FP64 best among the points within the screen's error of the wave's maximum, into *out
__device__ inline void helperScreen(const Int8Args &a, const Moments &mom, const int *S, const FrameWindow &fw, int f,
                                    int h, int nh, int lane, WaveBest *out)
{
    const Geometry &g = a.g;
    const int M = g.M;
    const int64_t boff = f * g.boxSize + int64_t(fw.sy) * g.bw + fw.sx;
    const int2 *scr = a.fg.screen + boff;
    const int32_t *box2 = a.fg.box2 + boff;
    const int nrows = (M - h + nh - 1) / nh, total = nrows * M;
    const unsigned mM = quotMagic(M);
    // pass 1: the two largest screened values of each lane, with the index of the largest
    float f1 = -3.0f, f2 = -3.0f;
    int i1 = 0x7fffffff;
    for (int e0 = lane; e0 < total; e0 += kBatch * kWave) {
        int raw[kBatch], fs[kBatch], at[kBatch];
        float rs[kBatch];
#pragma unroll
        for (int t = 0; t < kBatch; t++) {
            const int e = e0 + t * kWave;
            at[t] = -1;
            if (e < total) {
                const int r = quot(e, mM, M), v = e - r * M, u = h + r * nh;
                at[t] = u * M + v;
                raw[t] = S[at[t]];
                const int2 p = scr[u * g.bw + v];
                fs[t] = p.x;
                rs[t] = __int_as_float(p.y);
            }
        }
#pragma unroll
        for (int t = 0; t < kBatch; t++)
            if (at[t] >= 0) {
                const float x = screenNcc(raw[t], fs[t], rs[t], mom);
                if (x > f1 || (x == f1 && at[t] < i1)) {
                    f2 = f1;
                    f1 = x;
                    i1 = at[t];
                }
                else if (x > f2)
                    f2 = x;
            }
    }
    float fmax = f1;
    int imax = i1;
    waveBest(fmax, imax);
    const float thr = fmax - 1e-5f * fabsf(fmax) - 1e-30f;
    double dbest = -3.0;
    int dibest = 0x7fffffff;
    if (__any(f2 >= thr)) {
        // a lane holds two candidates: every point within the screen's error of the maximum
        for (int e = lane; e < total; e += kWave) {
            const int r = quot(e, mM, M), v = e - r * M, u = h + r * nh;
            const int2 p = scr[u * g.bw + v];
            const int raw = S[u * M + v];
            if (screenNcc(raw, p.x, __int_as_float(p.y), mom) >= thr)
                bestOf(dbest, dibest, exactNcc(raw, p.x, box2[u * g.bw + v], mom), u * M + v);
        }
    }
    else if (f1 >= thr) {
        const int u = i1 / M, v = i1 - u * M, o = u * g.bw + v;
        dbest = exactNcc(S[i1], scr[o].x, box2[o], mom);
        dibest = i1;
    }
    waveBest(dbest, dibest);
    if (lane == 0) {
        out->value = dbest;
        out->index = dibest;
    }
}

// The SNR box around the peak (ub, vb) of an M x M surface: n = min(stat, M) points a side, from
// (peak - n / 2), clipped to the surface (cxx/surface.h:108-124)
struct SnrBox {
    int ub, vb, ulo, uhi, vlo, vhi, bwid, items;
    bool inside;   // the peak has its 3 x 3 neighbours
};

__device__ inline SnrBox snrBox(int ib, int M, int stat)
{
    SnrBox b;
    b.ub = ib / M;
    b.vb = ib - b.ub * M;
    const int nn = stat < M ? stat : M;
    b.ulo = max(b.ub - nn / 2, 0);
    b.uhi = min(b.ub - nn / 2 + nn, M);
    b.vlo = max(b.vb - nn / 2, 0);
    b.vhi = min(b.vb - nn / 2 + nn, M);
    b.bwid = b.vhi - b.vlo;
    b.items = (b.uhi - b.ulo) * b.bwid;
    b.inside = b.ub > 0 && b.ub < M - 1 && b.vb > 0 && b.vb < M - 1;
    return b;
}

// one wave: the outputs of frame f from its peak ib (u M + v; valid: the peak is positive): the SNR over
// the stat box and the 3 x 3 FP64 fit and covariance on exact values (cxx/surface.h:71-146 formulas). The
// raw surface S is the whole M x M surface, or (boxed) only the SNR box, row-major from its corner.
__device__ inline void finishPeak(const Int8Args &a, const Moments &mom, const int *S, bool boxed, const FrameWindow &fw,
                                  int f, bool valid, int ib, int lane)
{
    const Geometry &g = a.g;
    const int M = g.M;
    if (!fw.valid || !valid) {
        if (lane == 0) invalidateOut(fw.q, a.fg.offset, a.fg.peak, a.fg.snr, a.fg.cov);
        return;
    }
    const int64_t boff = f * g.boxSize + int64_t(fw.sy) * g.bw + fw.sx;
    const int2 *scr = a.fg.screen + boff;
    const int32_t *box2 = a.fg.box2 + boff;
    const SnrBox b = snrBox(ib, M, g.stat);
    const int sStride = boxed ? b.bwid : M, su = boxed ? b.ulo : 0, sv = boxed ? b.vlo : 0;
    double sum = 0.0;
    // the stat box, kBatch points per lane in flight; lanes 0..8 also take the 3 x 3 neighbourhood (lane 4
    // the peak itself, the only one when the peak is on the surface's border)
    double mine = 0.0;
    for (int it0 = lane; it0 < b.items; it0 += kBatch * kWave) {
        int raw[kBatch], fs[kBatch], f2v[kBatch];
        bool ok[kBatch];
#pragma unroll
        for (int t = 0; t < kBatch; t++) {
            const int it = it0 + t * kWave;
            ok[t] = it < b.items;
            if (ok[t]) {
                const int u = b.ulo + it / b.bwid, v = b.vlo + it % b.bwid, o = u * g.bw + v;
                raw[t] = S[(u - su) * sStride + v - sv];
                fs[t] = scr[o].x;
                f2v[t] = box2[o];
            }
        }
#pragma unroll
        for (int t = 0; t < kBatch; t++)
            if (ok[t]) {
                const double d = exactNcc(raw[t], fs[t], f2v[t], mom);
                sum += d * d;
            }
    }
    if (lane < 9 && (b.inside || lane == 4)) {
        const int u = b.ub + lane / 3 - 1, v = b.vb + lane % 3 - 1, o = u * g.bw + v;
        mine = exactNcc(S[(u - su) * sStride + v - sv], scr[o].x, box2[o], mom);
    }
    sum = waveSumD(sum);
    double fsq[3][3];
    for (int i = 0; i < 9; i++) fsq[i / 3][i % 3] = __shfl(mine, i, kWave);
    if (lane == 0)
        writePeakF64(fsq, b.inside, b.ub, b.vb, g.R, fw.gy, fw.gx, sum, b.items, mom.nd, fw.q, a.fg.offset, a.fg.peak,
                     a.fg.snr, a.fg.cov);
}

// one wave: the peak of frame f from the helpers' results over the whole raw surface S
__device__ inline void helperFinish(const Int8Args &a, const Moments &mom, const int *S, const FrameWindow &fw, int f,
                                    const WaveBest *part, int nh, int lane)
{
    double best = -3.0;
    int ib = 0x7fffffff;
    for (int t = 0; t < nh; t++) bestOf(best, ib, part[t].value, part[t].index);
    finishPeak(a, mom, S, false, fw, f, best > 0, ib, lane);
}

} // namespace

} // namespace framecor
