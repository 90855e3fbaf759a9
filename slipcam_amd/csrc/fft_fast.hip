// fft_fast.hip -- the FP32 FFT form for search windows of L = 16 N2 pixels (N2 = 4, 6, 8, 10, 12:
// L = 64, 96, 128, 160, 192): register-resident transforms by lane groups, four block barriers a frame
//
// The correlation of one search window is the reference's (cxx/cpu.cpp:117-167): the L x L chip minus
// its mean, its 2-D real FFT, the product with the conjugate spectrum of the zero-padded reference
// window, the inverse transform, normalization (cxx/surface.h:30-35), first maximum and writePeak.
// One block of NW waves (8 to 16) takes one window and a run of frames; per frame, in LDS:
//
//   S      L rows of L / 2 complex (row stride P, plus one slot per 16 rows: bank-conflict free column
//          reads), the half spectrum with the DC and Nyquist columns packed into slot 0 (both are real
//          along x, so X(0) + i X(L/2) is one complex column: L / 2 columns to transform, not L / 2 + 1)
//
//   chip   the L x L uint8 chip staged into the S rows (row y at the start of S row y) while the block
//          sums its pixels exactly; barrier; mean = float(sum) / float(L^2) as the reference
//   rows   per row pair, z = (row 2p - mean) + i (row 2p + 1 - mean), its L-point FFT, and the split
//          of Z(f) and conj Z(L - f) into the two rows' half spectra; barrier
//   cols   per column, the L-point FFT along y, the product with the cached conjugate reference spectrum
//          (column 0: the DC and Nyquist columns unpacked, multiplied and repacked in one step), and the
//          inverse FFT along y, keeping the M output rows; barrier
//   irows  per pair of output rows, the two inverse real transforms as one complex one, normalization
//          with the box sums, the surface row written back as floats, the first maximum per lane;
//          barrier; the block's first maximum; wave 0: the FP64 SNR box, FP32 fit and covariance
//          (writePeakF32) while the other waves stage the next frame's chip (after one more barrier)
//
// Every 1-D transform of L = 16 N2 points is two register-resident steps done by one lane group of 16
// lanes inside one wave: step 1, 16 tasks j: the DFT-N2 of elements j + 16 m (m < N2), times
// exp(-2 pi i j k2 / L); step 2, N2 tasks k2: the DFT-16 of elements (j, k2), j < 16, giving frequency
// k2 + N2 k1 at element k1 + 16 k2. The two steps exchange through the line's own LDS storage, so only
// the wave synchronizes between them (the lanes of a wave run in lockstep; a fence keeps the compiler's
// order); the inverse runs the steps in reverse, so frequencies stay in that order and no reordering
// pass exists. Forward column DFT, spectrum product and inverse column DFT are one step in registers.
// A wave takes four lines at a time (a lane group each). The spectra of the reference windows are made
// by the same row and column steps and stored per (column, k2) as the 16 values a lane multiplies by.

#include "common.h"
#include "epilogue.h"
#include "fft_phases.h"
#include "kernels.h"

namespace framecor {

namespace {

// ---- compile-time roots of unity and register DFTs

constexpr double kPi = 3.14159265358979323846;

constexpr double taylorSin(double x)
{
    double term = x, sum = 0.0;
    for (int k = 1; k < 40; k++) {
        sum += term;
        term *= -x * x / double((2 * k) * (2 * k + 1));
    }
    return sum;
}

// exp(-2 pi i e / N), e < N: the Taylor series in FP64 of the angle reduced to (-pi, pi], rounded to
// FP32; the roots on the axes exactly
template <int N>
struct Roots {
    float re[N], im[N];
};

template <int N>
constexpr Roots<N> makeRoots()
{
    Roots<N> r{};
    for (int e = 0; e < N; e++) {
        double a = -2.0 * kPi * e / N;
        if (a <= -kPi) a += 2.0 * kPi;
        double c = taylorSin(a + kPi / 2), s = taylorSin(a);
        if ((4 * e) % N == 0) {
            const int q = (4 * e) / N;   // 0, 1, 2, 3: angles 0, -pi/2, -pi, -3pi/2
            c = q == 0 ? 1.0 : q == 2 ? -1.0 : 0.0;
            s = q == 1 ? -1.0 : q == 3 ? 1.0 : 0.0;
        }
        r.re[e] = float(c);
        r.im[e] = float(s);
    }
    return r;
}

template <int N>
constexpr Roots<N> kRoots = makeRoots<N>();

// exp(-+2 pi i e / N) (INV: the conjugate); e is a constant once the loops around are unrolled
template <int N, bool INV>
__device__ inline float2 root(int e)
{
    const float re = kRoots<N>.re[e % N], im = kRoots<N>.im[e % N];
    return make_float2(re, INV ? -im : im);
}

// the DFT of N values in registers, natural order in and out (forward: exp(-2 pi i n k / N); INV: the
// unnormalized inverse). N = A B: the DFT-B of the A subsequences of stride A, twiddles, the DFT-A.
template <int N, bool INV>
__device__ inline void dftReg(float2 *v)
{
    if constexpr (N == 1) {
    }
    else if constexpr (N == 2) dft2<INV>(v);
    else if constexpr (N == 3) dft3<INV>(v);
    else if constexpr (N == 4) dft4<INV>(v);
    else if constexpr (N == 5) dft5<INV>(v);
    else if constexpr (N == 8) dft8<INV>(v);
    else {
        constexpr int A = N % 4 == 0 ? 4 : N % 2 == 0 ? 2 : N % 3 == 0 ? 3 : 5;
        constexpr int B = N / A;
        static_assert(A * B == N && B > 1, "dftReg: no split");
        float2 y[A][B];
#pragma unroll
        for (int n1 = 0; n1 < A; n1++) {
#pragma unroll
            for (int n2 = 0; n2 < B; n2++) y[n1][n2] = v[n1 + A * n2];
            dftReg<B, INV>(y[n1]);
#pragma unroll
            for (int k2 = 1; k2 < B; k2++)
                if (n1 > 0) y[n1][k2] = cmul(y[n1][k2], root<N, INV>(n1 * k2));
        }
#pragma unroll
        for (int k2 = 0; k2 < B; k2++) {
            float2 z[A];
#pragma unroll
            for (int n1 = 0; n1 < A; n1++) z[n1] = y[n1][k2];
            dftReg<A, INV>(z);
#pragma unroll
            for (int k1 = 0; k1 < A; k1++) v[B * k1 + k2] = z[k1];
        }
    }
}

// the lanes of a wave exchanging values through LDS: they run in lockstep and LDS operations of one wave
// complete in order, so only the compiler's order of the accesses has to be kept
__device__ inline void waveFence()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- the layout of one search-window size

template <int L_, int NW_>
struct Fast {
    static constexpr int L = L_, NW = NW_, N2 = L / 16, H = L / 2;
    static constexpr int threads = NW * kWave;
    // row stride of S (complex): odd (columns read 16 consecutive rows conflict-free) and room for a
    // line's intermediate elements (j, k2) at 17 k2 + j in the two rows of a pair
    static constexpr int P = (L / 2 + N2 / 2) | 1;
    static constexpr int slots = L * P + (L >> 4);
    static_assert(L % 16 == 0 && N2 >= 4 && N2 <= 12, "Fast: L = 16 N2, N2 = 4 .. 12");
    static_assert(17 * (N2 - 1) + 15 < 2 * P && 9 * (N2 - 1) + 8 <= 2 * P, "Fast: pair storage");
    __host__ __device__ static constexpr int idx(int y, int k) { return y * P + (y >> 4) + k; }
    // LDS: S, the twiddle table (L complex), per-wave partial sums and maxima
    static constexpr size_t shmem = sizeof(float2) * (size_t(slots) + L) + 16 * NW + 16;
};

struct Parts {
    int *sums;      // NW: the chip sum
    float *bestv;   // NW
    int *besti;     // NW
};

template <class F>
__device__ inline Parts carveParts(float2 *tw)
{
    Parts p;
    p.sums = reinterpret_cast<int *>(tw + F::L);
    p.bestv = reinterpret_cast<float *>(p.sums + F::NW);
    p.besti = reinterpret_cast<int *>(p.bestv + F::NW);
    return p;
}

// the forward row transform of pair p (rows 2p, 2p + 1 of S hold the pair's input; after it, their half
// spectra): task l of lane group g; in(x, a, b): the two rows' values at column x
template <class F, class In>
__device__ inline void rowForward(float2 *S, const float2 *tw, int p, int l, In in)
{
    constexpr int N2 = F::N2, P = F::P;
    float2 *pair = S + F::idx(2 * p, 0);
    {   // step 1: task j = l
        const int j = l;
        float2 v[N2];
#pragma unroll
        for (int m = 0; m < N2; m++) v[m] = in(pair, j + 16 * m);
        dftReg<N2, false>(v);
#pragma unroll
        for (int k2 = 1; k2 < N2; k2++) v[k2] = cmul(v[k2], tw[j * k2]);
#pragma unroll
        for (int k2 = 0; k2 < N2; k2++) pair[17 * k2 + j] = v[k2];
    }
    waveFence();
    const int k2 = l;
    float2 u[16];
    if (k2 < N2) {   // step 2: task k2 = l, Z(k2 + N2 k1) in u[k1]; the upper halves out for the partners
#pragma unroll
        for (int j = 0; j < 16; j++) u[j] = pair[17 * k2 + j];
        dftReg<16, false>(u);
#pragma unroll
        for (int k1 = 8; k1 < 16; k1++) pair[9 * k2 + k1 - 8] = u[k1];
    }
    waveFence();
    float2 A[8], B[8];
    if (k2 < N2) {
        // the split: A(f) = (Z(f) + conj Z(L - f)) / 2, B(f) = (Z(f) - conj Z(L - f)) / 2i for f = k2 + N2 k1,
        // k1 < 8; Z(L - f) is element 15 - k1 of task N2 - k2, or 16 - k1 of task 0 when k2 = 0
        const int k2p = k2 == 0 ? 0 : N2 - k2;
#pragma unroll
        for (int k1 = 0; k1 < 8; k1++) {
            const int k1p = k2 == 0 ? 16 - k1 : 15 - k1;
            const float2 zk = u[k1];
            const float2 zc = (k2 == 0 && k1 == 0) ? u[0] : pair[9 * k2p + k1p - 8];
            A[k1] = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y - zc.y));
            B[k1] = make_float2(0.5f * (zk.y + zc.y), -0.5f * (zk.x - zc.x));
        }
    }
    waveFence();
    if (k2 < N2) {
#pragma unroll
        for (int k1 = 0; k1 < 8; k1++) {
            const int f = k2 + N2 * k1;
            if (f == 0) {   // the real DC and Nyquist values of both rows, packed: (X(0), X(L / 2))
                pair[0] = make_float2(u[0].x, u[8].x);
                pair[P] = make_float2(u[0].y, u[8].y);
            }
            else {
                pair[f] = A[k1];
                pair[P + f] = B[k1];
            }
        }
    }
}

// the forward column transform of column k: step 1 (task j = l; the column's elements stay in S,
// element (j, k2) at row j + 16 k2), then step 2 (colStep2: the spectrum C(k2 + N2 k1) in u[k1] of task
// k2 = l < N2)
template <class F>
__device__ inline void colStep1(float2 *S, const float2 *tw, int k, int l)
{
    constexpr int N2 = F::N2;
    float2 *col = S + k + F::idx(l, 0);   // rows l + 16 m at m (16 P + 1)
    const int j = l;
    float2 v[N2];
#pragma unroll
    for (int m = 0; m < N2; m++) v[m] = col[m * (16 * F::P + 1)];
    dftReg<N2, false>(v);
#pragma unroll
    for (int k2 = 1; k2 < N2; k2++) v[k2] = cmul(v[k2], tw[j * k2]);
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) col[k2 * (16 * F::P + 1)] = v[k2];
}

template <class F>
__device__ inline void colStep2(const float2 *S, int k, int k2, float2 (&u)[16])
{
    const float2 *col = S + k + k2 * (16 * F::P + 1);   // rows 16 k2 + j at j P
#pragma unroll
    for (int j = 0; j < 16; j++) u[j] = col[j * F::P];
    dftReg<16, false>(u);
}

// the partner of frequency f = k2 + N2 k1 in the packed column 0: L - f, as (task, element)
template <class F>
__device__ inline int partnerRow(int k2, int k1)
{
    const int k2p = k2 == 0 ? 0 : F::N2 - k2;
    const int k1p = k2 == 0 ? (16 - k1) & 15 : 15 - k1;
    return 16 * k2p + k1p;
}

// inverse step 1 of column k: task j = l, the output rows j + 16 m < M
template <class F>
__device__ inline void colInverseStep1(float2 *S, int k, int l, int M)
{
    constexpr int N2 = F::N2;
    const int j = l;
    float2 *col = S + k + F::idx(j, 0);
    float2 v[N2];
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) v[k2] = col[k2 * (16 * F::P + 1)];
    dftReg<N2, true>(v);
#pragma unroll
    for (int m = 0; m < N2; m++)
        if (j + 16 * m < M) col[m * (16 * F::P + 1)] = v[m];
}

// inverse step 2 of a column (task k2, the product in u): element (j, k2) at row 16 k2 + j
template <class F>
__device__ inline void colInverseStep2(float2 *S, const float2 *tw, int k, int k2, float2 (&u)[16])
{
    dftReg<16, true>(u);
#pragma unroll
    for (int j = 1; j < 16; j++) u[j] = cmulc(u[j], tw[j * k2]);
    float2 *cq = S + k + k2 * (16 * F::P + 1);
#pragma unroll
    for (int j = 0; j < 16; j++) cq[j * F::P] = u[j];
}

// column k >= 1 of the correlation: forward transform, times the conjugate reference spectrum (16 values
// per task, read from global memory after the forward transform), inverse transform, the M output rows
template <class F>
__device__ inline void columnTask(float2 *S, const float2 *tw, const float2 *spec, int kc, int l, int M)
{
    constexpr int N2 = F::N2;
    colStep1<F>(S, tw, kc, l);
    waveFence();
    const int k2 = l;
    if (k2 < N2) {
        float2 u[16];
        colStep2<F>(S, kc, k2, u);
        const float4 *s4 = reinterpret_cast<const float4 *>(spec + (size_t(kc) * N2 + k2) * 16);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float4 sp = s4[i];
            u[2 * i] = cmul(u[2 * i], make_float2(sp.x, sp.y));
            u[2 * i + 1] = cmul(u[2 * i + 1], make_float2(sp.z, sp.w));
        }
        colInverseStep2<F>(S, tw, kc, k2, u);
    }
    waveFence();
    colInverseStep1<F>(S, kc, l, M);
}

// column 0, the DC and Nyquist columns packed as D + i N: C'(f) = C(f) P(f) + conj C(L - f) Q(f) with
// P, Q from the reference spectra, which is D(f) conj TD(f) + i N(f) conj TN(f); C(L - f) is read back
// from the lanes that hold it
template <class F>
__device__ inline void packedColumnTask(float2 *S, const float2 *tw, const float2 *spec, int l, int M)
{
    constexpr int N2 = F::N2, H = F::H, P = F::P;
    colStep1<F>(S, tw, 0, l);
    waveFence();
    const int k2 = l;
    float2 u[16];
    if (k2 < N2) {
        colStep2<F>(S, 0, k2, u);
        float2 *cq = S + k2 * (16 * P + 1);
#pragma unroll
        for (int k1 = 0; k1 < 16; k1++) cq[k1 * P] = u[k1];
    }
    waveFence();
    if (k2 < N2) {
        const float2 *pv = spec + size_t(k2) * 16, *qv = spec + (size_t(H) * N2 + k2) * 16;
        // fully unrolled: u stays in registers (a partly unrolled loop indexes it at run time, in scratch)
#pragma unroll
        for (int k1 = 0; k1 < 16; k1++) {
            const float2 zc = S[F::idx(partnerRow<F>(k2, k1), 0)];
            u[k1] = cadd(cmul(u[k1], pv[k1]), cmul(make_float2(zc.x, -zc.y), qv[k1]));
        }
    }
    waveFence();
    if (k2 < N2) colInverseStep2<F>(S, tw, 0, k2, u);
    waveFence();
    colInverseStep1<F>(S, 0, l, M);
}

// ---- the correlation of window k with frames fBegin .. fEnd - 1 (the whole block)

template <class F>
__device__ __attribute__((always_inline)) inline void fastWindow(const FftArgs &a, int k, int fBegin, int fEnd,
                                                                 unsigned char *lds)
{
    constexpr int L = F::L, N2 = F::N2, H = F::H, P = F::P, NW = F::NW;
    const Geometry &g = a.g;
    float2 *S = reinterpret_cast<float2 *>(lds);
    float2 *tw = S + F::slots;
    const Parts parts = carveParts<F>(tw);
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid >> 6, grp = lane >> 4, l = lane & 15;
    const int M = g.M, W = g.W;

    for (int t = tid; t < L; t += F::threads) tw[t] = a.twiddle[t];
    const float normk = a.wd.norm[k];
    const int y0 = a.wd.down[k / g.nx], x0 = a.wd.across[k % g.nx];
    const float2 *spec = a.spectra + size_t(k) * (H + 1) * L;

    for (int f = fBegin; f < fEnd; f++) {
        const int64_t q = int64_t(f) * g.nwin + k;
        int gy = 0, gx = 0;
        if (a.fg.gross) {
            gy = a.fg.gross[2 * q];
            gx = a.fg.gross[2 * q + 1];
        }
        int sy, sx;
        __syncthreads();   // the previous frame's epilogue read the surface in S (and the twiddle table is in)
        if (!(normk > 0) || !searchWindow(g, y0, x0, gy, gx, sy, sx)) {   // uniform
            if (tid == 0) invalidateOut(q, a.fg.offset, a.fg.peak, a.fg.snr, a.fg.cov);
            continue;
        }

        // ---- the chip into the S rows (row y at the start of S row y) and its exact pixel sum
        const uint8_t *src = a.fg.frames + size_t(f) * g.pixels + size_t(sy) * g.w + sx;
        unsigned part = 0;
        for (int c = tid; c < L * (L / 16); c += F::threads) {
            const int y = c / (L / 16), cc = c - y * (L / 16);
            uint4 v;
            __builtin_memcpy(&v, src + size_t(y) * g.w + 16 * cc, 16);
            part = __builtin_amdgcn_sad_u8(v.x, 0u, part);
            part = __builtin_amdgcn_sad_u8(v.y, 0u, part);
            part = __builtin_amdgcn_sad_u8(v.z, 0u, part);
            part = __builtin_amdgcn_sad_u8(v.w, 0u, part);
            uint2 *d = reinterpret_cast<uint2 *>(S + F::idx(y, 0)) + 2 * cc;
            d[0] = make_uint2(v.x, v.y);
            d[1] = make_uint2(v.z, v.w);
        }
        const int wsum = waveSumI(int(part));
        if (lane == 0) parts.sums[wave] = wsum;
        __syncthreads();
        int boxL = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) boxL += parts.sums[w];
        // cxx/cpu.cpp:130-137: mean = float(box sum over L x L) / float(L^2), buf = px - mean
        const float mean = float(boxL) / (float(L) * float(L));

        // ---- rows (each stage takes the lane index anew: its lane-dependent addresses are made there, not
        // kept in registers across the frame loop)
        int lr = l;
        asm volatile("" : "+v"(lr));
        lr &= 15;
        for (int b = wave; 4 * b < H; b += NW) {
            const int p = 4 * b + grp;
            if (p < H)
                rowForward<F>(S, tw, p, lr, [mean](const float2 *pair, int x) {
                    const uint8_t *ra = reinterpret_cast<const uint8_t *>(pair);
                    const uint8_t *rb = reinterpret_cast<const uint8_t *>(pair + P);
                    return make_float2(float(ra[x]) - mean, float(rb[x]) - mean);
                });
            waveFence();
        }
        __syncthreads();

        // ---- columns: forward, times the conjugate reference spectrum, inverse; the M output rows.
        // Columns 1 .. H - 1 four to a wave at a time; the packed column 0 by the first lane group of the
        // last wave (which has the fewest of the others) in a pass of its own
        int lc = l;
        asm volatile("" : "+v"(lc));
        lc &= 15;
        for (int b = wave; 4 * b < H - 1; b += NW) {
            const int kc = 1 + 4 * b + grp;
            if (kc < H) columnTask<F>(S, tw, spec, kc, lc, M);
            waveFence();
        }
        if (wave == NW - 1 && grp == 0) packedColumnTask<F>(S, tw, spec, lc, M);
        __syncthreads();

        // ---- output row pairs: inverse real transforms, normalization, first maximum
        const float scale = 1.0f / (float(L) * float(L));
        const int n = W * W;
        // the sigma map sqrtf(float(var) / float(n)) of the surface points (u, x) at u bw + x (32-bit offsets
        // from a uniform base; 0 where var <= 0)
        const float *sig = a.fg.sigma + int64_t(f) * g.boxSize + int64_t(sy) * g.bw + sx;
        const int bw = g.bw;
        float best = -2.0f;
        int ib = 0x7fffffff;
        const int nq = (M + 1) / 2;
        int li = l;
        asm volatile("" : "+v"(li));
        li &= 15;
        for (int b = wave; 4 * b < nq; b += NW) {
            const int qp = 4 * b + grp;
            if (qp < nq) {
                float2 *pr = S + F::idx(2 * qp, 0);
                const bool hasB = 2 * qp + 1 < M;
                const int k2 = li;
                if (k2 < N2) {   // inverse step 2 with the combine: z(f) = R_a(f) + i R_b(f), Hermitian for f > L / 2
                    float2 u[16];
#pragma unroll
                    for (int k1 = 0; k1 < 16; k1++) {
                        const int fq = k2 + N2 * k1;
                        float2 z;
                        if (k1 < 8) {
                            const float2 sa = pr[fq], sb = hasB ? pr[P + fq] : make_float2(0.0f, 0.0f);
                            z = fq == 0 ? make_float2(sa.x, sb.x) : make_float2(sa.x - sb.y, sa.y + sb.x);
                        }
                        else if (fq == H) {
                            z = make_float2(pr[0].y, hasB ? pr[P].y : 0.0f);
                        }
                        else {
                            const float2 sa = pr[L - fq], sb = hasB ? pr[P + L - fq] : make_float2(0.0f, 0.0f);
                            z = make_float2(sa.x + sb.y, sb.x - sa.y);
                        }
                        u[k1] = z;
                        // the loads of eight elements at a time (fewer live registers)
                        if (k1 == 7) __builtin_amdgcn_sched_barrier(0);
                    }
                    dftReg<16, true>(u);
#pragma unroll
                    for (int j = 1; j < 16; j++) u[j] = cmulc(u[j], tw[j * k2]);
#pragma unroll
                    for (int j = 0; j < 16; j++) pr[17 * k2 + j] = u[j];
                }
                waveFence();
                {   // inverse step 1: outputs x = j + 16 m of rows 2q (real part) and 2q + 1 (imaginary)
                    const int j = li;
                    float2 v[N2];
#pragma unroll
                    for (int k2b = 0; k2b < N2; k2b++) v[k2b] = pr[17 * k2b + j];
                    dftReg<N2, true>(v);
                    float *surfA = reinterpret_cast<float *>(pr), *surfB = reinterpret_cast<float *>(pr + P);
#pragma unroll
                    for (int m = 0; m < N2; m++) {
                        const int x = j + 16 * m;
                        if (x < M) {
#pragma unroll
                            for (int r = 0; r < 2; r++) {
                                const int u = 2 * qp + r;
                                if (r == 1 && !hasB) continue;
                                const float c = r == 0 ? v[m].x : v[m].y;
                                const float sg = sig[u * bw + x];
                                float val = 0.0f;
                                {
#pragma clang fp contract(off)
                                    const float cc = c * scale;
                                    if (sg > 0 && normk > 0) val = cc / (normk * sg);
                                }
                                (r == 0 ? surfA : surfB)[x] = val;
                                bestOf(best, ib, val, u * M + x);
                            }
                        }
                    }
                }
            }
            waveFence();
        }
        waveBest(best, ib);
        if (lane == 0) {
            parts.bestv[wave] = best;
            parts.besti[wave] = ib;
        }
        __syncthreads();
        best = -2.0f;
        ib = 0x7fffffff;
#pragma unroll
        for (int w = 0; w < NW; w++) bestOf(best, ib, parts.bestv[w], parts.besti[w]);
        if (!(best > 0)) {   // uniform
            if (tid == 0) invalidateOut(q, a.fg.offset, a.fg.peak, a.fg.snr, a.fg.cov);
            continue;
        }
        if (wave == 0) {   // the SNR box (FP64), the fit and the covariance; the others go on to the next chip
            const int ub = ib / M, vb = ib - ub * M;
            auto at = [&](int uu, int vv) { return reinterpret_cast<const float *>(S + F::idx(uu, 0))[vv]; };
            const int nn = g.stat < M ? g.stat : M;
            const int ulo = max(ub - nn / 2, 0), uhi = min(ub - nn / 2 + nn, M);
            const int vlo = max(vb - nn / 2, 0), vhi = min(vb - nn / 2 + nn, M);
            const int bwid = vhi - vlo, items = (uhi - ulo) * bwid;
            double sum = 0.0;
            for (int it = lane; it < items; it += kWave) {
                const int uu = ulo + it / bwid, vv = vlo + it % bwid;
                const double c = at(uu, vv);
                sum += c * c;
            }
            sum = waveSumD(sum);
            if (lane == 0) {
                const bool inside = ub > 0 && ub < M - 1 && vb > 0 && vb < M - 1;
                float fs[3][3];
                for (int i = 0; i < 3; i++)
                    for (int jj = 0; jj < 3; jj++) fs[i][jj] = inside ? at(ub + i - 1, vb + jj - 1) : 0.0f;
                fs[1][1] = at(ub, vb);
                writePeakF32(fs, inside, ub, vb, g.R, gy, gx, sum, items, double(n), q, a.fg.offset, a.fg.peak,
                             a.fg.snr, a.fg.cov);
            }
        }
    }
}

template <class F>
__global__ void __launch_bounds__(F::threads) fastCorrelate(FftArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const int logical = xcdRemap(blockIdx.x, gridDim.x);
    const int k = logical / a.nfb, fblock = logical - k * a.nfb;
    const int fBegin = fblock * a.fpb, fEnd = min(fBegin + a.fpb, a.fg.n);
    fastWindow<F>(a, k, fBegin, fEnd, lds);
}

// the persistent form over the work queue of a batch (every clip of this L)
template <class F>
__global__ void __launch_bounds__(F::threads) fastQueue(WorkQueue q, const FftArgs *__restrict__ clips)
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
        fastWindow<F>(clips[c], k, fBegin, fEnd, lds);
    }
}

// the spectra of the reference windows (one block per window) in the layout the column step reads:
// per column kc < L / 2 and task k2 < N2, the 16 values conj T(k2 + N2 k1, kc), k1 < 16; column 0 (the
// packed DC and Nyquist columns) holds P = (conj TD + conj TN) / 2 there and Q = (conj TD - conj TN) / 2
// at column L / 2, where TD and TN are the spectra of the two packed columns
template <class F>
__global__ void __launch_bounds__(F::threads) fastReference(FftArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    constexpr int L = F::L, N2 = F::N2, H = F::H, NW = F::NW;
    float2 *S = reinterpret_cast<float2 *>(lds);
    float2 *tw = S + F::slots;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid >> 6, grp = lane >> 4, l = lane & 15;
    const int k = blockIdx.x, W = a.g.W;
    for (int t = tid; t < L; t += F::threads) tw[t] = a.twiddle[t];
    const float *T = a.templ + size_t(k) * W * W;
    __syncthreads();
    for (int b = wave; 4 * b < H; b += NW) {
        const int p = 4 * b + grp;
        if (p < H)
            rowForward<F>(S, tw, p, l, [T, W, p](const float2 *, int x) {
                const float va = (x < W && 2 * p < W) ? T[2 * p * W + x] : 0.0f;
                const float vb = (x < W && 2 * p + 1 < W) ? T[(2 * p + 1) * W + x] : 0.0f;
                return make_float2(va, vb);
            });
        waveFence();
    }
    __syncthreads();
    float2 *out = a.spectraOut + size_t(k) * (H + 1) * L;
    for (int b = wave; 4 * b < H; b += NW) {
        const int kc = 4 * b + grp;
        if (kc < H) {
            colStep1<F>(S, tw, kc, l);
            waveFence();
            const int k2 = l;
            if (k2 < N2) {
                float2 u[16];
                colStep2<F>(S, kc, k2, u);
                float2 *o = out + (size_t(kc) * N2 + k2) * 16;
                if (kc != 0) {
#pragma unroll
                    for (int k1 = 0; k1 < 16; k1++) o[k1] = make_float2(u[k1].x, -u[k1].y);
                }
                else {
                    float2 *cq = S + kc + k2 * (16 * F::P + 1);
#pragma unroll
                    for (int k1 = 0; k1 < 16; k1++) cq[k1 * F::P] = u[k1];
                    waveFence();
                    float2 *oq = out + (size_t(H) * N2 + k2) * 16;
#pragma unroll
                    for (int k1 = 0; k1 < 16; k1++) {
                        const float2 c = u[k1], cr = S[F::idx(partnerRow<F>(k2, k1), 0)];
                        // TD = (C + conj Cr) / 2, TN = (C - conj Cr) / 2i
                        const float2 td = make_float2(0.5f * (c.x + cr.x), 0.5f * (c.y - cr.y));
                        const float2 tn = make_float2(0.5f * (c.y + cr.y), -0.5f * (c.x - cr.x));
                        o[k1] = make_float2(0.5f * (td.x + tn.x), -0.5f * (td.y + tn.y));
                        oq[k1] = make_float2(0.5f * (td.x - tn.x), -0.5f * (td.y - tn.y));
                    }
                }
            }
        }
        waveFence();
    }
}

// waves per block: the L / 8 batches of four lines (rows or columns) spread evenly over the four SIMDs
// (L = 64: 8 waves, one batch each; 96: 12, one each; 128 - 192: 16, one or two each, the same per SIMD)
constexpr int fastWaves(int L) { return L <= 64 ? 8 : L <= 96 ? 12 : 16; }

template <int L>
using FastL = Fast<L, fastWaves(L)>;

template <class F>
void configureOnce()
{
    static bool done = false;
    if (done) return;
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(fastCorrelate<F>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(fastReference<F>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipGetLastError();
    done = true;
}

template <class F>
void launchCorrelate(FftArgs a, hipStream_t stream)
{
    configureOnce<F>();
    const int target = 2048, maxFpb = 16;
    int nfb = std::max((a.fg.n + maxFpb - 1) / maxFpb, (target + a.g.nwin - 1) / a.g.nwin);
    if (nfb > a.fg.n) nfb = a.fg.n;
    a.fpb = (a.fg.n + nfb - 1) / nfb;
    a.nfb = (a.fg.n + a.fpb - 1) / a.fpb;
    const int64_t blocks = int64_t(a.g.nwin) * a.nfb;
    if (blocks > 0x7fffffff) throw std::invalid_argument("framecor: too many windows x frames in one group");
    hipLaunchKernelGGL(fastCorrelate<F>, dim3(unsigned(blocks)), dim3(F::threads), F::shmem, stream, a);
    FC_HIP_CHECK(hipGetLastError());
}

template <class F>
void launchReference(const FftArgs &a, hipStream_t stream)
{
    configureOnce<F>();
    hipLaunchKernelGGL(fastReference<F>, dim3(a.g.nwin), dim3(F::threads), F::shmem, stream, a);
    FC_HIP_CHECK(hipGetLastError());
}

template <class F>
FftKernel queueKernel()
{
    FftKernel k{};
    k.func = reinterpret_cast<const void *>(fastQueue<F>);
    k.threads = F::threads;
    k.shmem = F::shmem;
    allowLargeLds(k.func);
    return k;
}

} // namespace

bool fftFastFor(int L) { return L == 64 || L == 96 || L == 128 || L == 160 || L == 192; }

#define FC_FAST_SWITCH(L, CALL)                                       \
    switch (L) {                                                      \
    case 64: CALL(FastL<64>); break;                                  \
    case 96: CALL(FastL<96>); break;                                  \
    case 128: CALL(FastL<128>); break;                                \
    case 160: CALL(FastL<160>); break;                                \
    case 192: CALL(FastL<192>); break;                                \
    default: throw std::invalid_argument("framecor: no fast FFT kernel for this search window"); \
    }

void launchFftFastReference(const FftArgs &a, hipStream_t stream)
{
#define FC_CALL(F) launchReference<F>(a, stream)
    FC_FAST_SWITCH(a.g.L, FC_CALL)
#undef FC_CALL
}

void launchFftFastCorrelate(const FftArgs &a, hipStream_t stream)
{
    if (a.fg.n <= 0) return;
#define FC_CALL(F) launchCorrelate<F>(a, stream)
    FC_FAST_SWITCH(a.g.L, FC_CALL)
#undef FC_CALL
}

FftKernel fftFastQueueKernel(int L)
{
    FftKernel k{};
#define FC_CALL(F) k = queueKernel<F>()
    FC_FAST_SWITCH(L, FC_CALL)
#undef FC_CALL
    return k;
}

size_t fftFastShmem(int L)
{
    size_t s = 0;
#define FC_CALL(F) s = F::shmem
    FC_FAST_SWITCH(L, FC_CALL)
#undef FC_CALL
    return s;
}

} // namespace framecor
