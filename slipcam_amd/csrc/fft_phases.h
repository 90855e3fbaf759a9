// fft_phases.h -- the phases of the LDS-resident 2-D real FFT correlation (FP32 FFT form)
//
// One block holds the half spectrum of one L x L search window in LDS as np = (L + 1) / 2 row pairs of
// rs = 2 H complex slots (H = L / 2 + 1): rows 2p and 2p + 1 are first packed as one complex row
// z = a + i b (the two-for-one real transform), and after the row transforms they are split into their
// half spectra A (slots 0..H-1) and B (slots H..2H-1) of the pair's region.
//
// Transforms are in place, mixed radix: forward passes are decimation in frequency (natural order
// in, digit-reversed order out); inverse passes are the exact reverse (digit-reversed in, natural out),
// so no reordering pass is ever made. Along x the split and the combine steps read the digit-reversed
// row spectra through a position table; along y the spectrum stays digit-reversed, and the cached
// spectra of the reference windows (made by the same phases) share that order, so the product with
// their conjugates is pointwise.
//
// Each phase is a loop of independent items over (tid, nt); the caller separates phases with a block
// barrier. The functions are __host__ __device__ so that the same code can be stepped on a CPU.
#pragma once

#include <hip/hip_runtime.h>

namespace framecor {

// floor(x / d) for 0 <= x < 2^20 and 1 <= d < 2^12 as a multiply-high by m = ceil(2^32 / d): the error
// of x m / 2^32 against x / d is below x / 2^32 < 1 / d, which cannot carry it past the next integer
__host__ __device__ inline unsigned magicOf(unsigned d) { return unsigned((0x100000000ull + d - 1) / d); }

__host__ __device__ inline int divMagic(int x, unsigned m, int d)
{
    // d == 1: m would be 2^32
    if (d == 1) return x;
#ifdef __HIP_DEVICE_COMPILE__
    return int(__umulhi(unsigned(x), m));
#else
    return int((unsigned long long)unsigned(x) * m >> 32);
#endif
}

__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a times conj(b)
__host__ __device__ inline float2 cmulc(float2 a, float2 b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// a times -i (forward) or +i (inverse)
template <bool INV>
__host__ __device__ inline float2 rotq(float2 a)
{
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// in-place DFTs of small radix: forward sums x[m] exp(-2 pi i m c / R); inverse exp(+...)
template <bool INV>
__host__ __device__ inline void dft2(float2 *v)
{
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <bool INV>
__host__ __device__ inline void dft3(float2 *v)
{
    const float h = 0.86602540378443864676f;   // sin(2 pi / 3)
    const float2 s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    const float2 m = make_float2(v[0].x - 0.5f * s.x, v[0].y - 0.5f * s.y);
    const float2 r = rotq<INV>(make_float2(h * d.x, h * d.y));
    v[0] = cadd(v[0], s);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
}

template <bool INV>
__host__ __device__ inline void dft4(float2 *v)
{
    const float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const float2 t2 = cadd(v[1], v[3]), t3 = rotq<INV>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
}

template <bool INV>
__host__ __device__ inline void dft5(float2 *v)
{
    const float c1 = 0.30901699437494742410f;    // cos(2 pi / 5)
    const float c2 = -0.80901699437494742410f;   // cos(4 pi / 5)
    const float s1 = 0.95105651629515357212f;    // sin(2 pi / 5)
    const float s2 = 0.58778525229247312917f;    // sin(4 pi / 5)
    const float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    const float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    const float2 x0 = v[0];
    const float2 p1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
    const float2 p2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
    // forward: X1 = p1 - i q1, X4 = p1 + i q1, X2 = p2 - i q2, X3 = p2 + i q2 (inverse: signs swap)
    const float2 q1 = rotq<INV>(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
    const float2 q2 = rotq<INV>(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
    v[0] = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
    v[1] = cadd(p1, q1);
    v[4] = csub(p1, q1);
    v[2] = cadd(p2, q2);
    v[3] = csub(p2, q2);
}

template <bool INV>
__host__ __device__ inline void dft8(float2 *v)
{
    float2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft4<INV>(e);
    dft4<INV>(o);
    const float h = 0.70710678118654752440f;
    // W8^c o[c], W8 = exp(-i pi / 4) (forward) or exp(+i pi / 4) (inverse)
    const float2 w1 = INV ? make_float2(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
                          : make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    const float2 w2 = rotq<INV>(o[2]);
    const float2 w3 = INV ? make_float2(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
                          : make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], w1);
    v[5] = csub(e[1], w1);
    v[2] = cadd(e[2], w2);
    v[6] = csub(e[2], w2);
    v[3] = cadd(e[3], w3);
    v[7] = csub(e[3], w3);
}

template <int R, bool INV>
__host__ __device__ inline void dftR(float2 *v)
{
    if constexpr (R == 2) dft2<INV>(v);
    else if constexpr (R == 3) dft3<INV>(v);
    else if constexpr (R == 4) dft4<INV>(v);
    else if constexpr (R == 5) dft5<INV>(v);
    else if constexpr (R == 8) dft8<INV>(v);
}

// element e of line l of a row-pair pass: the packed row of pair l
struct RowAddr {
    int rs;
    __host__ __device__ int operator()(int l, int e) const { return l * rs + e; }
};

// element y of column k of the half spectrum: row y is half (y & 1) of pair y >> 1
struct ColAddr {
    int rs, H;
    __host__ __device__ int operator()(int l, int e) const { return (e >> 1) * rs + (e & 1) * H + l; }
};

// one radix-R pass over nlines lines of length L; Lj: the length of the sub-transforms of this pass;
// tw: exp(-2 pi i t / L), t < L. Forward (INV false): DFT then twiddle; inverse: untwiddle then DFT.
template <int R, bool INV, class Addr>
__host__ __device__ inline void radixPass(float2 *Z, const float2 *tw, int L, int nlines, int Lj, Addr addr,
                                          int tid, int nt)
{
    const int Sj = Lj / R, per = L / R, tstep = L / Lj;
    const int items = nlines * per;
    const unsigned mPer = magicOf(per), mSj = magicOf(Sj);
    for (int it = tid; it < items; it += nt) {
        const int line = divMagic(it, mPer, per), w = it - line * per;
        const int b = divMagic(w, mSj, Sj), n1 = w - b * Sj;
        const int e0 = b * Lj + n1;
        float2 v[R];
#pragma unroll
        for (int m = 0; m < R; m++) v[m] = Z[addr(line, e0 + Sj * m)];
        if (!INV) {
            dftR<R, false>(v);
#pragma unroll
            for (int c = 1; c < R; c++) v[c] = cmul(v[c], tw[n1 * c * tstep]);
        }
        else {
#pragma unroll
            for (int c = 1; c < R; c++) v[c] = cmulc(v[c], tw[n1 * c * tstep]);
            dftR<R, true>(v);
        }
#pragma unroll
        for (int m = 0; m < R; m++) Z[addr(line, e0 + Sj * m)] = v[m];
    }
}

// a pass of any radix R by direct sums with the twiddle table (the radices without a register butterfly:
// primes from 7 up, of any size). An item is one output element; the items are taken in rounds of whole
// lines, at most IPT per thread, so that every input of a round is read (into registers) before any
// output of it is written: the pass is in place. Needs IPT nt >= L (a line in one round).
template <int IPT, bool INV, class Addr, class Sync>
__host__ __device__ inline void directPass(int R, float2 *Z, const float2 *tw, int L, int nlines, int Lj, Addr addr,
                                           int tid, int nt, Sync sync)
{
    const int Sj = Lj / R, tstep = L / Lj, rstep = L / R;
    const int linesPerRound = (IPT * nt) / L > 0 ? (IPT * nt) / L : 1;
    const unsigned mL = magicOf(L), mLj = magicOf(Lj), mSj = magicOf(Sj);
    for (int l0 = 0; l0 < nlines; l0 += linesPerRound) {
        const int l1 = l0 + linesPerRound < nlines ? l0 + linesPerRound : nlines;
        const int items = (l1 - l0) * L;
        float2 o[IPT];
#pragma unroll
        for (int i = 0; i < IPT; i++) {
            const int it = tid + i * nt;
            if (it < items) {
                const int ll = divMagic(it, mL, L), line = l0 + ll, e = it - ll * L;
                const int b = divMagic(e, mLj, Lj), r = e - b * Lj;
                const int c = divMagic(r, mSj, Sj), n1 = r - c * Sj;
                const int e0 = b * Lj + n1;
                float2 acc = make_float2(0.0f, 0.0f);
                for (int m = 0; m < R; m++) {
                    float2 v = Z[addr(line, e0 + Sj * m)];
                    if (INV && m > 0) v = cmulc(v, tw[n1 * m * tstep]);
                    const float2 w = tw[((m * c) % R) * rstep];
                    acc = cadd(acc, INV ? cmulc(v, w) : cmul(v, w));
                }
                o[i] = (!INV && c > 0) ? cmul(acc, tw[n1 * c * tstep]) : acc;
            }
        }
        sync();
#pragma unroll
        for (int i = 0; i < IPT; i++) {
            const int it = tid + i * nt;
            if (it < items) {
                const int ll = divMagic(it, mL, L);
                Z[addr(l0 + ll, it - ll * L)] = o[i];
            }
        }
        sync();
    }
}

template <bool INV, class Addr, class Sync>
__host__ __device__ inline void anyPass(int R, float2 *Z, const float2 *tw, int L, int nlines, int Lj, Addr addr,
                                        int tid, int nt, Sync sync)
{
    switch (R) {
    case 2: radixPass<2, INV>(Z, tw, L, nlines, Lj, addr, tid, nt); break;
    case 3: radixPass<3, INV>(Z, tw, L, nlines, Lj, addr, tid, nt); break;
    case 4: radixPass<4, INV>(Z, tw, L, nlines, Lj, addr, tid, nt); break;
    case 5: radixPass<5, INV>(Z, tw, L, nlines, Lj, addr, tid, nt); break;
    case 8: radixPass<8, INV>(Z, tw, L, nlines, Lj, addr, tid, nt); break;
    default: directPass<4, INV>(R, Z, tw, L, nlines, Lj, addr, tid, nt, sync); break;
    }
}

// the transforms of nlines lines: forward (decimation in frequency) or inverse (the reverse passes);
// sync() separates the passes
template <bool INV, class Addr, class Sync>
__host__ __device__ inline void transformLines(float2 *Z, const float2 *tw, int L, const int *fac, int nf, int nlines,
                                               Addr addr, int tid, int nt, Sync sync)
{
    if (!INV) {
        int Lj = L;
        for (int j = 0; j < nf; j++) {
            anyPass<false>(fac[j], Z, tw, L, nlines, Lj, addr, tid, nt, sync);
            Lj /= fac[j];
            sync();
        }
    }
    else {
        for (int j = nf - 1; j >= 0; j--) {
            int Lj = L;
            for (int i = 0; i < j; i++) Lj /= fac[i];
            anyPass<true>(fac[j], Z, tw, L, nlines, Lj, addr, tid, nt, sync);
            sync();
        }
    }
}

// the position of frequency k in the digit-reversed output of the forward passes
__host__ __device__ inline int digitReversed(int k, const int *fac, int nf, int L)
{
    int pos = 0, S = L;
    for (int j = 0; j < nf; j++) {
        S /= fac[j];
        pos += (k % fac[j]) * S;
        k /= fac[j];
    }
    return pos;
}

// the split of the packed row spectra of pairs [p0, p1) into their half spectra: item (p, k), k < H:
// A = (Z(k) + conj Z(L-k)) / 2 into slot k, B = (Z(k) - conj Z(L-k)) / 2i into slot H + k. Items read
// slots that other items write, so values go through registers: a round reads at most IPT items per
// thread, syncs, then writes them.
template <int IPT, class Sync>
__host__ __device__ inline void splitPairs(float2 *Z, const int *pos, int L, int H, int rs, int np, int tid, int nt,
                                           Sync sync)
{
    const int pairsPerRound = (IPT * nt) / H > 0 ? (IPT * nt) / H : 1;
    for (int p0 = 0; p0 < np; p0 += pairsPerRound) {
        const int p1 = p0 + pairsPerRound < np ? p0 + pairsPerRound : np;
        const int items = (p1 - p0) * H;
        const unsigned mH = magicOf(H);
        float2 A[IPT], B[IPT];
#pragma unroll
        for (int i = 0; i < IPT; i++) {
            const int it = tid + i * nt;
            if (it < items) {
                const int pp = divMagic(it, mH, H), p = p0 + pp, k = it - pp * H;
                const float2 zk = Z[p * rs + pos[k]];
                const float2 zc = Z[p * rs + pos[(L - k) % L]];
                // conj(Z(L-k)) = (zc.x, -zc.y)
                A[i] = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y - zc.y));
                B[i] = make_float2(0.5f * (zk.y + zc.y), -0.5f * (zk.x - zc.x));
            }
        }
        sync();
#pragma unroll
        for (int i = 0; i < IPT; i++) {
            const int it = tid + i * nt;
            if (it < items) {
                const int pp = divMagic(it, mH, H), p = p0 + pp, k = it - pp * H;
                Z[p * rs + k] = A[i];
                Z[p * rs + H + k] = B[i];
            }
        }
        sync();
    }
}

// the inverse of the split for pairs [0, np): z(k) = a(k) + i b(k) over k < L from the half spectra
// of rows 2p (slots 0..H-1) and 2p+1 (slots H..), extended by Hermitian symmetry, written at the
// digit-reversed positions for the inverse row passes. The imaginary parts of the columns 0 and L/2
// (L even) are dropped, as a complex-to-real transform does; rows at or beyond L (odd L) are zero.
template <int IPT, class Sync>
__host__ __device__ inline void combinePairs(float2 *Z, const int *pos, int L, int H, int rs, int np, int tid, int nt,
                                             Sync sync)
{
    const int pairsPerRound = (IPT * nt) / L > 0 ? (IPT * nt) / L : 1;
    for (int p0 = 0; p0 < np; p0 += pairsPerRound) {
        const int p1 = p0 + pairsPerRound < np ? p0 + pairsPerRound : np;
        const int items = (p1 - p0) * L;
        const unsigned mL = magicOf(L);
        float2 z[IPT];
#pragma unroll
        for (int i = 0; i < IPT; i++) {
            const int it = tid + i * nt;
            if (it < items) {
                const int pp = divMagic(it, mL, L), p = p0 + pp, k = it - pp * L;
                const bool hasB = 2 * p + 1 < L;
                float2 a, b = make_float2(0.0f, 0.0f);
                if (k < H) {
                    a = Z[p * rs + k];
                    if (hasB) b = Z[p * rs + H + k];
                    if (k == 0 || 2 * k == L) {
                        a.y = 0.0f;
                        b.y = 0.0f;
                    }
                }
                else {
                    const float2 ac = Z[p * rs + (L - k)];
                    a = make_float2(ac.x, -ac.y);
                    if (hasB) {
                        const float2 bc = Z[p * rs + H + (L - k)];
                        b = make_float2(bc.x, -bc.y);
                    }
                }
                z[i] = make_float2(a.x - b.y, a.y + b.x);
            }
        }
        sync();
#pragma unroll
        for (int i = 0; i < IPT; i++) {
            const int it = tid + i * nt;
            if (it < items) {
                const int pp = divMagic(it, mL, L), p = p0 + pp, k = it - pp * L;
                Z[p * rs + pos[k]] = z[i];
            }
        }
        sync();
    }
}

} // namespace framecor
