// diff.hip -- the differentiable form: the NCC of each window with a continuously shifted, bilinearly
// sampled search chip, its hand-written backward pass, and a Gauss-Newton refinement of the shift
//
// For frame f and window k (reference window T at (y0, x0), integer gross offset (gy, gx)) and a
// continuous shift s = (sy, sx), the chip is C[i][j] = bilinear(F, y0 + gy + sy + i, x0 + gx + sx + j),
// i, j < W. With a = frac(sy), b = frac(sx) every pixel of the chip has the same bilinear weights, so
// C is a 2 x 2 filter of the (W + 1) x (W + 1) frame region at (y0 + gy + floor(sy), x0 + gx + floor(sx)).
//   ncc   = sum t^ (C - mean C) / sigma,   t^ = (T - mean T) / |T - mean T|,   sigma = |C - mean C|
//   d ncc / d C[i][j] = (t^[i][j] - ncc c^[i][j]) / sigma,   c^ = (C - mean C) / sigma
//   d C / d sy = (1 - b)(F[y+1][x] - F[y][x]) + b (F[y+1][x+1] - F[y][x+1])       (y, x: the top-left tap)
//   d C / d sx = (1 - a)(F[y][x+1] - F[y][x]) + a (F[y+1][x+1] - F[y+1][x])
// The gradient of the frame is the transposed 2 x 2 filter of d ncc / d C over the region, added with
// one atomic per region pixel (windows overlap). A window whose region leaves the frame, or whose
// template or chip is flat, gives ncc 0 and zero gradients.
//
// Gauss-Newton on r = t^ - c^(s) (|r|^2 = 2 - 2 ncc): with D = dC/ds, D~ = D - mean D and
// alpha = c^ . D~, the Jacobian J = dc^/ds = (D~ - c^ alpha) / sigma gives
//   J^T J = (D~ . D~ - alpha alpha^T) / sigma^2,   J^T r = (t^ . D~ - alpha ncc) / sigma,
// so each step needs only nine block sums; s += (J^T J)^-1 J^T r, each step bounded to 1 px.
//
// One block of 256 threads per (frame, window); the chip and the gradient live in LDS.

#include "common.h"
#include "kernels.h"

namespace framecor {

namespace {

constexpr int kDiffThreads = 256;
constexpr int kDiffWaves = kDiffThreads / kWave;

struct Tap {
    int y, x;     // the top-left frame pixel of chip pixel (0, 0)
    float a, b;   // bilinear weights along y and x
    bool valid;
};

__device__ inline Tap tapOf(const DiffArgs &d, int k, int64_t q, float sy, float sx)
{
    const Geometry &g = d.g;
    int gy = 0, gx = 0;
    if (d.gross) {
        gy = d.gross[2 * q];
        gx = d.gross[2 * q + 1];
    }
    const float fy = floorf(sy), fx = floorf(sx);
    Tap t;
    t.a = sy - fy;
    t.b = sx - fx;
    t.y = t.x = 0;
    // the shift's magnitude is checked as a float before any conversion (a shift beyond the frame,
    // however large, is not valid), and the tap's bounds in int64 (gross offsets are any int32)
    t.valid = isfinite(sy) && isfinite(sx) && fabsf(fy) <= float(g.h) && fabsf(fx) <= float(g.w);
    if (!t.valid) return t;
    const int64_t ty = int64_t(d.wd.down[k / g.nx]) + gy + int64_t(fy);
    const int64_t tx = int64_t(d.wd.across[k % g.nx]) + gx + int64_t(fx);
    t.valid = ty >= 0 && tx >= 0 && ty + g.W + 1 <= g.h && tx + g.W + 1 <= g.w;
    if (t.valid) {
        t.y = int(ty);
        t.x = int(tx);
    }
    return t;
}

// block-wide sums of NV doubles (every thread gets the totals)
template <int NV>
__device__ inline void blockSum(double (&v)[NV], double *red)
{
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) v[i] = waveSumD(v[i]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; i++) red[wave * NV + i] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; i++) {
        double s = 0.0;
        for (int w = 0; w < kDiffWaves; w++) s += red[w * NV + i];
        v[i] = s;
    }
}

struct TemplateNorm {
    float mean, inv;   // t^ = (T - mean) * inv
    bool valid;
};

__device__ inline TemplateNorm templateNorm(const DiffArgs &d, int k)
{
    const double n = double(d.g.W) * d.g.W;
    TemplateNorm t;
    const int64_t var = d.wd.tvar[k];
    t.valid = var > 0;
    t.mean = float(double(d.wd.tsum[k]) / n);
    t.inv = t.valid ? float(sqrt(n / double(var))) : 0.0f;
    return t;
}

// the chip value and its derivatives at (i, j)
struct Sample {
    float c, dy, dx;
};

__device__ inline Sample sampleAt(const float *F, int w, const Tap &t, int i, int j)
{
    const float *p = F + int64_t(t.y + i) * w + t.x + j;
    const float f00 = p[0], f01 = p[1], f10 = p[w], f11 = p[w + 1];
    Sample s;
    s.c = (1.0f - t.a) * ((1.0f - t.b) * f00 + t.b * f01) + t.a * ((1.0f - t.b) * f10 + t.b * f11);
    s.dy = (1.0f - t.b) * (f10 - f00) + t.b * (f11 - f01);
    s.dx = (1.0f - t.a) * (f01 - f00) + t.a * (f11 - f10);
    return s;
}

// the moments of the chip: mean and ncc terms; returns false for a flat chip
struct ChipStats {
    float mean, sigma, ncc;
};

__device__ inline bool chipStats(const DiffArgs &d, const float *F, const uint8_t *T, const Tap &tap,
                                 const TemplateNorm &tn, float *chip, double *red, ChipStats &st)
{
    const int W = d.g.W, n = W * W, tid = threadIdx.x;
    double v1[1] = {0.0};
    for (int idx = tid; idx < n; idx += kDiffThreads) {
        const int i = idx / W, j = idx - i * W;
        const float c = sampleAt(F, d.g.w, tap, i, j).c;
        chip[idx] = c;
        v1[0] += c;
    }
    blockSum<1>(v1, red);
    const float mean = float(v1[0] / n);
    double v2[2] = {0.0, 0.0};
    for (int idx = tid; idx < n; idx += kDiffThreads) {
        const float dc = chip[idx] - mean;
        const float th = (float(T[idx]) - tn.mean) * tn.inv;
        v2[0] += double(dc) * dc;
        v2[1] += double(th) * dc;
    }
    blockSum<2>(v2, red);
    st.mean = mean;
    st.sigma = float(sqrt(v2[0]));
    st.ncc = st.sigma > 0 ? float(v2[1] / sqrt(v2[0])) : 0.0f;
    return st.sigma > 0;
}

__global__ void __launch_bounds__(kDiffThreads) diffForward(DiffArgs d)
{
    extern __shared__ __align__(16) unsigned char lds[];
    double *red = reinterpret_cast<double *>(lds);
    float *chip = reinterpret_cast<float *>(red + 3 * kDiffWaves * 4);
    const int64_t q = blockIdx.x;
    const int f = int(q / d.g.nwin), k = int(q - int64_t(f) * d.g.nwin);
    const Tap tap = tapOf(d, k, q, d.shifts[2 * q], d.shifts[2 * q + 1]);
    const TemplateNorm tn = templateNorm(d, k);
    if (!tap.valid || !tn.valid) {   // uniform
        if (threadIdx.x == 0) d.ncc[q] = 0.0f;
        return;
    }
    const float *F = d.frames + int64_t(f) * d.g.pixels;
    const uint8_t *T = d.tmpl + size_t(k) * d.g.W * d.g.W;
    ChipStats st;
    chipStats(d, F, T, tap, tn, chip, red, st);
    if (threadIdx.x == 0) d.ncc[q] = st.ncc;
}

__global__ void __launch_bounds__(kDiffThreads) diffBackward(DiffArgs d)
{
    extern __shared__ __align__(16) unsigned char lds[];
    double *red = reinterpret_cast<double *>(lds);
    float *chip = reinterpret_cast<float *>(red + 3 * kDiffWaves * 4);
    const int64_t q = blockIdx.x;
    const int f = int(q / d.g.nwin), k = int(q - int64_t(f) * d.g.nwin);
    const int W = d.g.W, n = W * W, tid = threadIdx.x;
    const Tap tap = tapOf(d, k, q, d.shifts[2 * q], d.shifts[2 * q + 1]);
    const TemplateNorm tn = templateNorm(d, k);
    const float gin = d.gradNcc[q];
    if (!tap.valid || !tn.valid || gin == 0.0f) {
        if (tid == 0) {
            d.gradShifts[2 * q] = 0.0f;
            d.gradShifts[2 * q + 1] = 0.0f;
        }
        return;
    }
    const float *F = d.frames + int64_t(f) * d.g.pixels;
    const uint8_t *T = d.tmpl + size_t(k) * W * W;
    ChipStats st;
    if (!chipStats(d, F, T, tap, tn, chip, red, st)) {
        if (tid == 0) {
            d.gradShifts[2 * q] = 0.0f;
            d.gradShifts[2 * q + 1] = 0.0f;
        }
        return;
    }
    // dL/dC into the chip array, and the shift gradient
    const float inv = 1.0f / st.sigma;
    double gs[2] = {0.0, 0.0};
    for (int idx = tid; idx < n; idx += kDiffThreads) {
        const int i = idx / W, j = idx - i * W;
        const Sample s = sampleAt(F, d.g.w, tap, i, j);
        const float th = (float(T[idx]) - tn.mean) * tn.inv;
        const float ch = (s.c - st.mean) * inv;
        const float dc = gin * (th - st.ncc * ch) * inv;
        gs[0] += double(dc) * s.dy;
        gs[1] += double(dc) * s.dx;
        chip[idx] = dc;
    }
    blockSum<2>(gs, red);   // its barriers also publish chip[] to the block
    if (tid == 0) {
        d.gradShifts[2 * q] = float(gs[0]);
        d.gradShifts[2 * q + 1] = float(gs[1]);
    }
    // the region gradient: the transposed bilinear filter of dL/dC
    const float w00 = (1.0f - tap.a) * (1.0f - tap.b), w01 = (1.0f - tap.a) * tap.b;
    const float w10 = tap.a * (1.0f - tap.b), w11 = tap.a * tap.b;
    float *G = d.gradFrames + int64_t(f) * d.g.pixels;
    const int R1 = W + 1;
    for (int idx = tid; idx < R1 * R1; idx += kDiffThreads) {
        const int p = idx / R1, r = idx - p * R1;
        float v = 0.0f;
        if (p < W && r < W) v += w00 * chip[p * W + r];
        if (p < W && r > 0) v += w01 * chip[p * W + r - 1];
        if (p > 0 && r < W) v += w10 * chip[(p - 1) * W + r];
        if (p > 0 && r > 0) v += w11 * chip[(p - 1) * W + r - 1];
        if (v != 0.0f) unsafeAtomicAdd(G + int64_t(tap.y + p) * d.g.w + tap.x + r, v);
    }
}

__global__ void __launch_bounds__(kDiffThreads) diffRefine(DiffArgs d)
{
    extern __shared__ __align__(16) unsigned char lds[];
    double *red = reinterpret_cast<double *>(lds);
    const int64_t q = blockIdx.x;
    const int f = int(q / d.g.nwin), k = int(q - int64_t(f) * d.g.nwin);
    const int W = d.g.W, n = W * W, tid = threadIdx.x;
    const TemplateNorm tn = templateNorm(d, k);
    const float *F = d.frames + int64_t(f) * d.g.pixels;
    const uint8_t *T = d.tmpl + size_t(k) * W * W;
    float sy = d.shifts[2 * q], sx = d.shifts[2 * q + 1];
    float ncc = 0.0f;
    bool ok = tn.valid;
    for (int step = 0; step <= d.steps && ok; step++) {
        const Tap tap = tapOf(d, k, q, sy, sx);
        if (!tap.valid) {
            ok = false;
            break;
        }
        double m[3] = {0.0, 0.0, 0.0};
        for (int idx = tid; idx < n; idx += kDiffThreads) {
            const Sample s = sampleAt(F, d.g.w, tap, idx / W, idx % W);
            m[0] += s.c;
            m[1] += s.dy;
            m[2] += s.dx;
        }
        blockSum<3>(m, red);
        const float mc = float(m[0] / n), my = float(m[1] / n), mx = float(m[2] / n);
        // cc, cy, cx, tc, ty, tx, yy, xx, xy
        double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int idx = tid; idx < n; idx += kDiffThreads) {
            const Sample s = sampleAt(F, d.g.w, tap, idx / W, idx % W);
            const double c = s.c - mc, y = s.dy - my, x = s.dx - mx;
            const double t = (float(T[idx]) - tn.mean) * tn.inv;
            v[0] += c * c;
            v[1] += c * y;
            v[2] += c * x;
            v[3] += t * c;
            v[4] += t * y;
            v[5] += t * x;
            v[6] += y * y;
            v[7] += x * x;
            v[8] += x * y;
        }
        blockSum<9>(v, red);
        if (!(v[0] > 0)) {
            ok = false;
            break;
        }
        const double sigma = sqrt(v[0]);
        const double rho = v[3] / sigma, ay = v[1] / sigma, ax = v[2] / sigma;
        ncc = float(rho);
        if (step == d.steps) break;
        const double s2 = v[0];
        const double hyy = (v[6] - ay * ay) / s2, hxx = (v[7] - ax * ax) / s2, hxy = (v[8] - ay * ax) / s2;
        const double by = (v[4] - ay * rho) / sigma, bx = (v[5] - ax * rho) / sigma;
        const double det = hyy * hxx - hxy * hxy;
        if (!(det > 0)) break;
        double dy = (hxx * by - hxy * bx) / det, dx = (hyy * bx - hxy * by) / det;
        const double big = fmax(fabs(dy), fabs(dx));
        if (big > 1.0) {
            dy /= big;
            dx /= big;
        }
        sy = float(double(sy) + dy);
        sx = float(double(sx) + dx);
    }
    if (tid == 0) {
        d.shiftsOut[2 * q] = ok ? sy : __builtin_nanf("");
        d.shiftsOut[2 * q + 1] = ok ? sx : __builtin_nanf("");
        d.ncc[q] = ok ? ncc : 0.0f;
    }
}

size_t diffShmem(const Geometry &g)
{
    return sizeof(double) * 3 * kDiffWaves * 4 + sizeof(float) * size_t(g.W) * g.W;
}

void configure(const void *kern)
{
    (void)hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipGetLastError();
}

} // namespace

void launchDiff(DiffArgs d, DiffOp op, hipStream_t stream)
{
    const int64_t blocks = int64_t(d.n) * d.g.nwin;
    if (blocks <= 0) return;
    if (blocks > 0x7fffffff) throw std::invalid_argument("framecor: too many frames x windows");
    static bool configured = false;
    if (!configured) {
        configure(reinterpret_cast<const void *>(diffForward));
        configure(reinterpret_cast<const void *>(diffBackward));
        configure(reinterpret_cast<const void *>(diffRefine));
        configured = true;
    }
    const size_t shmem = diffShmem(d.g);
    switch (op) {
    case DiffOp::Forward:
        hipLaunchKernelGGL(diffForward, dim3(unsigned(blocks)), dim3(kDiffThreads), shmem, stream, d);
        break;
    case DiffOp::Backward:
        hipLaunchKernelGGL(diffBackward, dim3(unsigned(blocks)), dim3(kDiffThreads), shmem, stream, d);
        break;
    case DiffOp::Refine:
        hipLaunchKernelGGL(diffRefine, dim3(unsigned(blocks)), dim3(kDiffThreads), sizeof(double) * 3 * kDiffWaves * 4,
                           stream, d);
        break;
    }
    FC_HIP_CHECK(hipGetLastError());
}

} // namespace framecor
