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
// takes the exact FP64 NCC of the points within FP32 error of the wave's maximum (so the argmax is
// that of the exact FP64 surface, first maximum in row-major order on ties), and evaluates the SNR,
// the 3 x 3 FP64 fit and the covariance on exact FP64 values (cxx/surface.h:71-146 formulas).

#include "common.h"
#include "epilogue.h"
#include "kernels.h"

namespace framecor {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

struct WaveBest {
    double value;
    int index;
    int pad;
};

struct Moments {
    int64_t n;      // W^2
    int64_t tsum;   // sum T
    int64_t tvar;   // W^2 sum T^2 - (sum T)^2
};

// the exact NCC in FP64 from the raw int8 sum and the box sums of the frame; 0 where the search box
// (or the template) is flat, as cxx/surface.h:30-35
__device__ inline double exactNcc(int raw, int fs, int fs2, const Moments &m)
{
    const int64_t sft = int64_t(raw) + 128 * (int64_t(fs) + m.tsum) - 16384 * m.n;
    const int64_t vf = m.n * int64_t(fs2) - int64_t(fs) * int64_t(fs);
    if (vf <= 0 || m.tvar <= 0) return 0.0;
    const double num = double(m.n) * double(sft) - double(m.tsum) * double(fs);
    return num / sqrt(double(m.tvar) * double(vf));
}

// the same to FP32 accuracy (a relative error of a few 1e-7), to screen the argmax candidates
__device__ inline float screenNcc(int raw, int fs, int fs2, const Moments &m)
{
    const int64_t sft = int64_t(raw) + 128 * (int64_t(fs) + m.tsum) - 16384 * m.n;
    const int64_t vf = m.n * int64_t(fs2) - int64_t(fs) * int64_t(fs);
    if (vf <= 0 || m.tvar <= 0) return 0.0f;
    const int64_t num = m.n * sft - m.tsum * int64_t(fs);
    return float(num) * rsqrtf(float(m.tvar) * float(vf));
}

template <int NVT, int NJB>
__global__ void __launch_bounds__(64 * NVT) int8Correlate(Int8Args a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const Geometry &g = a.g;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid >> 6, nthreads = blockDim.x;
    const int logical = xcdRemap(blockIdx.x, gridDim.x);
    const int k = logical / a.nfb, fblock = logical - k * a.nfb;
    const int fBegin = fblock * a.fpb, fEnd = min(fBegin + a.fpb, a.fg.n);
    const int W = g.W, M = g.M, L = g.L;

    int8_t *tmpl = reinterpret_cast<int8_t *>(lds);
    int8_t *chip = tmpl + 4 * a.ni4 * a.tstride;
    WaveBest *wbest = reinterpret_cast<WaveBest *>(chip + a.crows * a.cstride);
    double *psum = reinterpret_cast<double *>(wbest + a.nut);
    double *f9 = psum + a.nut;
    int *pcnt = reinterpret_cast<int *>(f9 + 9);

    // zero the template and chip areas (their padding stays zero), then the template as int8
    {
        const int bytes = 4 * a.ni4 * a.tstride + a.crows * a.cstride;
        for (int i = tid * 16; i < bytes; i += nthreads * 16) *reinterpret_cast<v4i *>(lds + i) = v4i{0, 0, 0, 0};
    }
    __syncthreads();
    {
        const uint8_t *T = a.tmpl + size_t(k) * W * W;
        for (int idx = tid; idx < W * W; idx += nthreads) {
            const int i = idx / W, j = idx - i * W;
            tmpl[i * a.tstride + 16 + j] = int8_t(T[idx] ^ 0x80);
        }
    }
    __syncthreads();

    const float normk = a.wd.norm[k];
    const int y0 = a.wd.down[k / g.nx], x0 = a.wd.across[k % g.nx];
    Moments mom;
    mom.n = int64_t(W) * W;
    mom.tsum = a.wd.tsum[k];
    mom.tvar = a.wd.tvar[k];

    const int g4 = lane >> 4, m16 = lane & 15;
    // B fragment geometry: lane (g4, n = m16) takes template row 4 i4 + g4, bytes 16 + 16 jb - n ..
    // + 15 of the zero-padded LDS row, i.e. dwords (16 - n) / 4 + 4 jb .. + 4, funnel-shifted
    const int bbyte = 16 - m16, bdw = bbyte >> 2, bsh = bbyte & 3;

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
            continue;   // uniform over the block
        }

        // the chip, pixels ^ 0x80 = pixels - 128 as int8 (the previous frame's reads ended at B3)
        {
            const uint8_t *src = a.fg.frames + size_t(f) * g.pixels + size_t(sy) * g.w + sx;
            const int cpr = L >> 4;
            for (int c = tid; c < L * cpr; c += nthreads) {
                const int r = c / cpr, cc = c - r * cpr;
                uint4 v;
                __builtin_memcpy(&v, src + size_t(r) * g.w + 16 * cc, 16);
                v.x ^= 0x80808080u;
                v.y ^= 0x80808080u;
                v.z ^= 0x80808080u;
                v.w ^= 0x80808080u;
                *reinterpret_cast<uint4 *>(chip + r * a.cstride + 16 * cc) = v;
            }
            const int tail = L & 15;
            if (tail)
                for (int c = tid; c < L * tail; c += nthreads) {
                    const int r = c / tail, j = 16 * cpr + (c - r * tail);
                    chip[r * a.cstride + j] = int8_t(src[size_t(r) * g.w + j] ^ 0x80);
                }
        }
        __syncthreads();   // B1

        // the band of this wave: rows 16 wave .. + 15 of the surface, all v tiles
        v4i acc[NVT];
#pragma unroll
        for (int v1 = 0; v1 < NVT; v1++) acc[v1] = v4i{0, 0, 0, 0};
        {
            const int8_t *crow = chip + (16 * wave + g4 + m16) * a.cstride;
            const uint32_t *trow = reinterpret_cast<const uint32_t *>(tmpl + g4 * a.tstride) + bdw;
            const int cstep = 4 * a.cstride, tstep = a.tstride;   // 4 rows: bytes; dwords per 4 rows
            for (int i4 = 0; i4 < a.ni4; i4++) {
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

        // epilogue 1: screen the band in FP32, then the exact FP64 maximum of the wave
        const int32_t *box = a.fg.box + f * g.boxSize + int64_t(sy) * g.bw + sx;
        const int32_t *box2 = a.fg.box2 + f * g.boxSize + int64_t(sy) * g.bw + sx;
        float fv[NVT][4];
        float fbest = -3.0f;
        int ibest = 0x7fffffff;
#pragma unroll
        for (int v1 = 0; v1 < NVT; v1++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int u = 16 * wave + 4 * g4 + r, v = 16 * v1 + m16;
                fv[v1][r] = -3.0f;
                if (u < M && v < M) {
                    const int64_t o = int64_t(u) * g.bw + v;
                    fv[v1][r] = screenNcc(acc[v1][r], box[o], box2[o], mom);
                    bestOf(fbest, ibest, fv[v1][r], u * M + v);
                }
            }
        waveBest(fbest, ibest);
        const float thr = fbest - 1e-5f * fabsf(fbest) - 1e-30f;
        double dbest = -3.0;
        int dibest = 0x7fffffff;
#pragma unroll
        for (int v1 = 0; v1 < NVT; v1++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (fv[v1][r] >= thr) {
                    const int u = 16 * wave + 4 * g4 + r, v = 16 * v1 + m16;
                    const int64_t o = int64_t(u) * g.bw + v;
                    bestOf(dbest, dibest, exactNcc(acc[v1][r], box[o], box2[o], mom), u * M + v);
                }
        waveBest(dbest, dibest);
        if (lane == 0) {
            wbest[wave].value = dbest;
            wbest[wave].index = dibest;
        }
        __syncthreads();   // B2

        // epilogue 2: every wave finds the block's peak; the bands that meet the stat box add their
        // exact squared surface, and the 3 x 3 around the peak
        double best = -3.0;
        int ib = 0x7fffffff;
        for (int t = 0; t < a.nut; t++) bestOf(best, ib, wbest[t].value, wbest[t].index);
        if (!(best > 0)) {   // uniform: every thread read the same values
            if (tid == 0) invalidateOut(q, a.fg.offset, a.fg.peak, a.fg.snr, a.fg.cov);
            continue;
        }
        const int ub = ib / M, vb = ib - ub * M;
        const int nn = g.stat < M ? g.stat : M;
        const int ulo = max(ub - nn / 2, 0), uhi = min(ub - nn / 2 + nn, M);
        const int vlo = max(vb - nn / 2, 0), vhi = min(vb - nn / 2 + nn, M);
        double sum = 0.0;
        int count = 0;
        if (16 * wave + 15 >= ulo && 16 * wave < uhi) {
#pragma unroll
            for (int v1 = 0; v1 < NVT; v1++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int u = 16 * wave + 4 * g4 + r, v = 16 * v1 + m16;
                    if (u >= ulo && u < uhi && v >= vlo && v < vhi) {
                        const int64_t o = int64_t(u) * g.bw + v;
                        const double d = exactNcc(acc[v1][r], box[o], box2[o], mom);
                        sum += d * d;
                        count++;
                        if (u >= ub - 1 && u <= ub + 1 && v >= vb - 1 && v <= vb + 1)
                            f9[(u - ub + 1) * 3 + (v - vb + 1)] = d;
                    }
                }
            sum = waveSumD(sum);
            count = waveSumI(count);
        }
        if (lane == 0) {
            psum[wave] = sum;
            pcnt[wave] = count;
        }
        __syncthreads();   // B3

        if (tid == 0) {
            double total = 0.0;
            int cnt = 0;
            for (int t = 0; t < a.nut; t++) {
                total += psum[t];
                cnt += pcnt[t];
            }
            const bool inside = ub > 0 && ub < M - 1 && vb > 0 && vb < M - 1;
            double fs[3][3];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) fs[i][j] = inside ? f9[i * 3 + j] : 0.0;
            fs[1][1] = best;
            writePeakF64(fs, inside, ub, vb, g.R, gy, gx, total, cnt, double(mom.n), q, a.fg.offset, a.fg.peak,
                         a.fg.snr, a.fg.cov);
        }
    }
}

template <int NVT, int NJB>
void launchOne(const Int8Args &a, int blocks, size_t shmem, hipStream_t stream)
{
    auto kern = int8Correlate<NVT, NJB>;
    static bool configured = false;
    if (!configured) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipGetLastError();
        configured = true;
    }
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(kWave * a.nut), shmem, stream, a);
}

template <int NJB>
bool dispatchNvt(int nvt, const Int8Args &a, int blocks, size_t shmem, hipStream_t stream)
{
    switch (nvt) {
    case 1: launchOne<1, NJB>(a, blocks, shmem, stream); return true;
    case 2: launchOne<2, NJB>(a, blocks, shmem, stream); return true;
    case 3: launchOne<3, NJB>(a, blocks, shmem, stream); return true;
    case 4: launchOne<4, NJB>(a, blocks, shmem, stream); return true;
    case 5: launchOne<5, NJB>(a, blocks, shmem, stream); return true;
    case 6: launchOne<6, NJB>(a, blocks, shmem, stream); return true;
    case 7: launchOne<7, NJB>(a, blocks, shmem, stream); return true;
    case 8: launchOne<8, NJB>(a, blocks, shmem, stream); return true;
    case 9: launchOne<9, NJB>(a, blocks, shmem, stream); return true;
    default: return false;
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
    l.njb = int8JBlocks(g.W);
    l.nvt = (g.M + 15) / 16;
    l.nut = l.nvt;
    l.ni4 = (g.W + 3) / 4;
    l.tstride = 16 * (l.njb + 2);
    int cs = 16 * (l.nvt + l.njb - 1);
    if (cs < g.L) cs = g.L;
    cs = (cs + 15) / 16;
    if (cs % 2 == 0) cs++;   // an odd number of 16-byte units: conflict-free 16-byte row reads
    l.cstride = 16 * cs;
    l.crows = 16 * l.nut + 4 * l.ni4;
    l.shmem = size_t(4 * l.ni4) * l.tstride + size_t(l.crows) * l.cstride + sizeof(WaveBest) * l.nut
        + sizeof(double) * (l.nut + 9) + sizeof(int) * l.nut;
    l.supported = l.njb > 0 && l.nvt <= 9 && l.shmem <= 160 * 1024;
    return l;
}

void launchInt8(Int8Args a, const Int8Layout &l, hipStream_t stream)
{
    if (a.fg.n <= 0) return;
    if (!l.supported) throw std::invalid_argument("framecor: the int8 form supports windows up to 129 pixels and search ranges up to 71 pixels");
    a.nut = l.nut;
    a.ni4 = l.ni4;
    a.tstride = l.tstride;
    a.cstride = l.cstride;
    a.crows = l.crows;
    // frames per block: enough blocks to fill the GPU several times over
    const int target = 2048;
    int nfb = (target + a.g.nwin - 1) / a.g.nwin;
    if (nfb > a.fg.n) nfb = a.fg.n;
    if (nfb < 1) nfb = 1;
    a.fpb = (a.fg.n + nfb - 1) / nfb;
    a.nfb = (a.fg.n + a.fpb - 1) / a.fpb;
    const int blocks = a.g.nwin * a.nfb;
    bool ok = false;
    switch (l.njb) {
    case 2: ok = dispatchNvt<2>(l.nvt, a, blocks, l.shmem, stream); break;
    case 3: ok = dispatchNvt<3>(l.nvt, a, blocks, l.shmem, stream); break;
    case 5: ok = dispatchNvt<5>(l.nvt, a, blocks, l.shmem, stream); break;
    case 9: ok = dispatchNvt<9>(l.nvt, a, blocks, l.shmem, stream); break;
    default: break;
    }
    if (!ok) throw std::invalid_argument("framecor: no int8 kernel for this window and search range");
    FC_HIP_CHECK(hipGetLastError());
}

} // namespace framecor
