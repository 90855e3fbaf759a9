// epilogue.h -- the peak of a correlation surface: sub-pixel fit, SNR and covariance (device code)
//
// The formulas are those of the reference's writePeak (cxx/surface.h:71-146): the quadratic fitted to
// the 3 x 3 surface around the integer peak (coefficients c1..c5, accepted when c3 < 0, det > 0 and
// |du|, |dv| <= 1, else one parabola per axis), the SNR as the squared peak over the mean squared
// surface of the stat box without the peak, and the pycuampcor covariance from the curvature. They are
// evaluated here in the same operation order, with FP contraction off, so that the same surface gives
// the same bits as a CPU evaluation of those formulas. fitF32 is the reference's FP32 fit (the FFT
// form); fitF64 evaluates the same fit in FP64 (the int8 form, on the exact FP64 surface).
#pragma once

#include "common.h"

namespace framecor {

__device__ inline float parabolaF32(float m, float c, float p)
{
#pragma clang fp contract(off)
    const float den = m - 2 * c + p;
    return den < 0 ? 0.5f * (m - p) / den : 0.0f;
}

__device__ inline double parabolaF64(double m, double c, double p)
{
#pragma clang fp contract(off)
    const double den = m - 2 * c + p;
    return den < 0 ? 0.5 * (m - p) / den : 0.0;
}

// f[i][j]: the surface at (u + i - 1, v + j - 1)
__device__ inline void fitF32(const float f[3][3], float &du, float &dv)
{
#pragma clang fp contract(off)
    const float c1 = (f[2][0] + f[2][1] + f[2][2] - f[0][0] - f[0][1] - f[0][2]) / 6;
    const float c2 = (f[0][2] + f[1][2] + f[2][2] - f[0][0] - f[1][0] - f[2][0]) / 6;
    const float c3 = (f[0][0] + f[0][1] + f[0][2] + f[2][0] + f[2][1] + f[2][2]
                      - 2 * (f[1][0] + f[1][1] + f[1][2])) / 6;
    const float c5 = (f[0][0] + f[1][0] + f[2][0] + f[0][2] + f[1][2] + f[2][2]
                      - 2 * (f[0][1] + f[1][1] + f[2][1])) / 6;
    const float c4 = (f[2][2] - f[2][0] - f[0][2] + f[0][0]) / 4;
    const float det = 4 * c3 * c5 - c4 * c4;
    bool fitted = false;
    du = 0.0f;
    dv = 0.0f;
    if (c3 < 0 && det > 0) {
        du = (c4 * c2 - 2 * c5 * c1) / det;
        dv = (c4 * c1 - 2 * c3 * c2) / det;
        fitted = fabsf(du) <= 1 && fabsf(dv) <= 1;
    }
    if (!fitted) {
        du = parabolaF32(f[0][1], f[1][1], f[2][1]);
        dv = parabolaF32(f[1][0], f[1][1], f[1][2]);
    }
}

__device__ inline void fitF64(const double f[3][3], double &du, double &dv)
{
#pragma clang fp contract(off)
    const double c1 = (f[2][0] + f[2][1] + f[2][2] - f[0][0] - f[0][1] - f[0][2]) / 6;
    const double c2 = (f[0][2] + f[1][2] + f[2][2] - f[0][0] - f[1][0] - f[2][0]) / 6;
    const double c3 = (f[0][0] + f[0][1] + f[0][2] + f[2][0] + f[2][1] + f[2][2]
                       - 2 * (f[1][0] + f[1][1] + f[1][2])) / 6;
    const double c5 = (f[0][0] + f[1][0] + f[2][0] + f[0][2] + f[1][2] + f[2][2]
                       - 2 * (f[0][1] + f[1][1] + f[2][1])) / 6;
    const double c4 = (f[2][2] - f[2][0] - f[0][2] + f[0][0]) / 4;
    const double det = 4 * c3 * c5 - c4 * c4;
    bool fitted = false;
    du = 0.0;
    dv = 0.0;
    if (c3 < 0 && det > 0) {
        du = (c4 * c2 - 2 * c5 * c1) / det;
        dv = (c4 * c1 - 2 * c3 * c2) / det;
        fitted = fabs(du) <= 1 && fabs(dv) <= 1;
    }
    if (!fitted) {
        du = parabolaF64(f[0][1], f[1][1], f[2][1]);
        dv = parabolaF64(f[1][0], f[1][1], f[1][2]);
    }
}

// the SNR from the sum of the squared surface over the stat box (count points, the peak included)
__device__ inline float snrOf(double f0, double sum, int count)
{
#pragma clang fp contract(off)
    const double mean = count > 1 ? (sum - f0 * f0) / (count - 1) : 0.0;
    return mean > 0 ? float(f0 * f0 / mean) : 0.0f;
}

// the covariance of the offsets from the curvature at an inside peak; f: the 3 x 3 surface in its own
// precision S (float for the FFT form, double for the int8 form): the sums of neighbours are taken
// in S, then promoted, as in cxx/surface.h:128-131; T: the pixels of the reference window
template <class S>
__device__ inline void covarianceOf(const S f[3][3], double T, float *cov)
{
#pragma clang fp contract(off)
    const double f0 = f[1][1];
    const S sxx = f[2][1] + f[0][1];
    const S syy = f[1][2] + f[1][0];
    const S sxy = f[2][2] + f[0][0] - f[2][0] - f[0][2];
    const double dxx = -(sxx - 2.0 * f0) * T;
    const double dyy = -(syy - 2.0 * f0) * T;
    const double dxy = sxy * 0.25 * T;
    double n2 = 1.0 - f0 > 0 ? 1.0 - f0 : 0.0;
    const double n4 = n2 * n2 * 0.5 * T;
    n2 *= 2;
    const double w = dxy * dxy - dxx * dyy;
    cov[0] = 99.0f;
    cov[1] = 99.0f;
    cov[2] = 0.0f;
    if (fabs(w) >= 1e-2) {
        const double w2 = w * w;
        cov[0] = float((-n2 * w * dyy + n4 * (dyy * dyy + dxy * dxy)) / w2);
        cov[1] = float((-n2 * w * dxx + n4 * (dxx * dxx + dxy * dxy)) / w2);
        cov[2] = float((n2 * w * dxy - n4 * (dxx + dyy) * dxy) / w2);
    }
}

// all the outputs of window q from its integer peak (u, v), the 3 x 3 surface around it (when
// inside) and the stat-box sum of squares; the FP32 (reference) fit
__device__ inline void writePeakF32(const float f[3][3], bool inside, int u, int v, int R, int gy, int gx,
                                    double sum, int count, double T, int64_t q, float *offset,
                                    float *peak, float *snr, float *cov)
{
#pragma clang fp contract(off)
    float du = 0.0f, dv = 0.0f;
    if (inside) fitF32(f, du, dv);
    offset[2 * q] = float(gy + u - R) + du;
    offset[2 * q + 1] = float(gx + v - R) + dv;
    const float f0 = f[1][1];
    peak[q] = f0;
    snr[q] = snrOf(double(f0), sum, count);
    float c[3] = {99.0f, 99.0f, 0.0f};
    if (inside) covarianceOf<float>(f, T, c);
    cov[3 * q] = c[0];
    cov[3 * q + 1] = c[1];
    cov[3 * q + 2] = c[2];
}

// the same with the FP64 fit on an FP64 surface; the offsets are rounded to float once
__device__ inline void writePeakF64(const double f[3][3], bool inside, int u, int v, int R, int gy, int gx,
                                    double sum, int count, double T, int64_t q, float *offset,
                                    float *peak, float *snr, float *cov)
{
#pragma clang fp contract(off)
    double du = 0.0, dv = 0.0;
    if (inside) fitF64(f, du, dv);
    offset[2 * q] = float(double(gy + u - R) + du);
    offset[2 * q + 1] = float(double(gx + v - R) + dv);
    const double f0 = f[1][1];
    peak[q] = float(f0);
    snr[q] = snrOf(f0, sum, count);
    float c[3] = {99.0f, 99.0f, 0.0f};
    if (inside) covarianceOf<double>(f, T, c);
    cov[3 * q] = c[0];
    cov[3 * q + 1] = c[1];
    cov[3 * q + 2] = c[2];
}

} // namespace framecor
