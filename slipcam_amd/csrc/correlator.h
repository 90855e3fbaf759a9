// correlator.h -- the correlator object behind the C ABI (capi.hip) and the batches of correlators
// (batch.hip): one clip's grid, reference windows, device buffers and launch sequence, plus the error
// plumbing of the C entry points (statuses and the per-thread message)
//
// Per group of frames: the W x W box-sum maps straight from the frames (boxMaps: the sums that the
// reference reads from its summed-area tables), then one correlation launch (int8Correlate or
// fftCorrelate) that writes the outputs. The reference windows are prepared on the host in
// set_reference exactly as cxx/common.cpp:30-48 (mean in FP64, d = float(px - mean),
// norm = float(sqrt(sum d^2)) in row-major order) so that the validity of a window (norm > 0) and the
// FFT form's normalization use the reference's norm bits.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "common.h"
#include "framecor_hip.h"
#include "kernels.h"

namespace framecor {

// the message of the last failing call of this thread
inline std::string &lastError()
{
    static thread_local std::string msg;
    return msg;
}

inline int fail(int status, const std::string &msg)
{
    lastError() = msg;
    return status;
}

template <class F>
int guarded(F &&f)
{
    try {
        f();
        lastError().clear();
        return FRAMECOR_OK;
    }
    catch (const std::invalid_argument &e) {
        return fail(FRAMECOR_EINVAL, e.what());
    }
    catch (const std::bad_alloc &) {
        return fail(FRAMECOR_ENOMEM, "framecor: out of memory");
    }
    catch (const std::exception &e) {
        return fail(FRAMECOR_ERUNTIME, e.what());
    }
    catch (...) {
        return fail(FRAMECOR_ERUNTIME, "framecor: unknown error");
    }
}

[[noreturn]] inline void invalid(const std::string &msg) { throw std::invalid_argument("framecor: " + msg); }

// device memory owned by a correlator
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer() { release(); }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    void ensure(size_t n)
    {
        if (n <= bytes && ptr) return;
        release();
        const hipError_t e = hipMalloc(&ptr, std::max<size_t>(n, 16));
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            ptr = nullptr;
            throw std::bad_alloc();
        }
        FC_HIP_CHECK(e);
        bytes = n;
    }
    template <class T>
    T *as() const { return static_cast<T *>(ptr); }
};

template <class F>
void parallelFor(int n, F &&body)
{
    const int nt = std::max(1, std::min<int>(16, int(std::thread::hardware_concurrency())));
    if (n < 64 || nt == 1) {
        for (int i = 0; i < n; i++) body(i);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
        pool.emplace_back([&, t] {
            for (int i = t; i < n; i += nt) body(i);
        });
    for (auto &th : pool) th.join();
}

} // namespace framecor

namespace framecor {

// the reference's parameter checks, with its messages (cxx/common.cpp:15-28)
inline void checkParams(int height, int width, const int *down, int ndown, const int *across, int nacross, int window,
                 int search, int batch, int stat)
{
    if (height <= 0 || width <= 0) invalid("the frame size must be positive");
    if (window < 3) invalid("the window must be at least 3 pixels");
    if (search < 1) invalid("the half search range must be at least 1 pixel");
    if (ndown <= 0 || nacross <= 0 || !down || !across) invalid("no windows");
    for (int i = 0; i < ndown; i++)
        if (down[i] < 0 || down[i] + window > height)
            invalid("a window is outside the frame (down " + std::to_string(down[i]) + ")");
    for (int i = 0; i < nacross; i++)
        if (across[i] < 0 || across[i] + window > width)
            invalid("a window is outside the frame (across " + std::to_string(across[i]) + ")");
    if (stat < 3) invalid("the SNR window must be at least 3 pixels");
    if (batch < 1) invalid("the batch must be at least 1 window");
}

// the reference's checks, then what the method supports
inline void checkAll(int height, int width, const int *down, int ndown, const int *across, int nacross, int window,
              int search, int method, int batch, int stat)
{
    checkParams(height, width, down, ndown, across, nacross, window, search, batch, stat);
    if (method != FRAMECOR_METHOD_FFT && method != FRAMECOR_METHOD_INT8) invalid("the method must be fft or int8");
    if (window > 181) invalid("the window must be at most 181 pixels (exact int32 box sums)");
    Geometry g{};
    g.W = window;
    g.R = search;
    g.L = window + 2 * search;
    g.M = 2 * search + 1;
    if (method == FRAMECOR_METHOD_INT8) {
        if (!int8Layout(g).supported)
            invalid("the int8 method supports windows up to 129 pixels and search ranges up to 71 pixels");
    }
    else {
        FftPlan plan;
        if (!fftPlanFor(g.L, plan))
            invalid("the fft method needs a search window (window + 2 search = " + std::to_string(g.L)
                    + ") of at most " + std::to_string(kFftMaxL) + " pixels");
    }
}

// the device address of the no-allocation entry point's scratch array on device (the current
// device), resolved once per device (noalloc.hip); called when a correlator is created
void scratchSetup(int device);

} // namespace framecor

using namespace framecor;

// the correlator behind the C ABI (one clip: its grid, its reference windows, its buffers)
struct framecor_correlator {
    Geometry g{};
    std::vector<int> down, across;
    int method = FRAMECOR_METHOD_FFT;
    int batch = 4096;
    int device = 0;
    hipStream_t stream = nullptr;
    bool ready = false;
    int maxGroup = 1;
    Int8Layout i8{};
    Int8Variant variant = Int8Variant::Auto;
    FftPlan plan{};

    // per window, written by set_reference
    DeviceBuffer dDown, dAcross, dNorm, dTsum, dTvar, dTmpl, dSpectra, dTwiddle;
    // the FFT form's per-block spectra when they do not fit in LDS (workBlocks blocks)
    DeviceBuffer dWork;
    int workBlocks = 0;
    // per group of frames
    DeviceBuffer dBox2, dScreen, dSigma;
    int groupCap = 0;
    // host-call staging
    DeviceBuffer dFrames, dGross, dOffset, dPeak, dSnr, dCov;
    // NCC surfaces and periodicity: the box maps of one frame, surfaces of a run of windows, results
    DeviceBuffer dSurfBox, dSurfBox2, dSurf, dPer;
    // per-stage timing: three events per group of the last call
    bool timing = false;
    std::vector<hipEvent_t> events;
    int eventsUsed = 0;
    // The order of the correlator's work across streams. Every call that enqueues work on the
    // correlator's buffers (templates, norms, spectra, box maps, staging) first makes its stream wait
    // for the last such work of any earlier call (after(s): the event lastUse), and records lastUse
    // behind its own work (done(s)). So set_reference on the correlator's own stream waits for a device
    // correlate still running on the caller's stream, and device calls on two streams run in call order.
    // Not under stream capture (a captured call is ordered by its graph).
    hipEvent_t lastUse = nullptr;
    bool used = false;

    ~framecor_correlator()
    {
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        if (stream || lastUse) (void)hipSetDevice(device);
        if (used) (void)hipEventSynchronize(lastUse);
        if (lastUse) (void)hipEventDestroy(lastUse);
        if (stream) (void)hipStreamDestroy(stream);
    }

    static bool capturing(hipStream_t s)
    {
        hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &st) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        return st != hipStreamCaptureStatusNone;
    }

    void after(hipStream_t s)
    {
        if (used && !capturing(s)) FC_HIP_CHECK(hipStreamWaitEvent(s, lastUse, 0));
    }

    void done(hipStream_t s)
    {
        if (capturing(s)) return;
        FC_HIP_CHECK(hipEventRecord(lastUse, s));
        used = true;
    }

    // checks the parameters and sets up the correlator on its device (before set_reference)
    void init(int height, int width, const int *down_, int ndown, const int *across_, int nacross, int window,
              int search, int method_, int batch_, int stat_size, int device_)
    {
        checkAll(height, width, down_, ndown, across_, nacross, window, search, method_, batch_, stat_size);
        g.h = height;
        g.w = width;
        g.W = window;
        g.R = search;
        g.L = window + 2 * search;
        g.M = 2 * search + 1;
        g.ny = ndown;
        g.nx = nacross;
        g.nwin = ndown * nacross;
        g.stat = stat_size;
        g.bh = height - window + 1;
        g.bw = width - window + 1;
        g.pixels = int64_t(height) * width;
        g.satSize = int64_t(height + 1) * (width + 1);
        g.boxSize = int64_t(g.bh) * g.bw;
        down.assign(down_, down_ + ndown);
        across.assign(across_, across_ + nacross);
        method = method_;
        batch = batch_;
        device = device_;
        if (method == FRAMECOR_METHOD_INT8)
            i8 = int8Layout(g);
        else
            fftPlanFor(g.L, plan);
        // groups of frames: at most 2 GiB of tables
        const size_t perFrame = bytesPerFrame();
        maxGroup = int(std::max<size_t>(1, std::min<size_t>(256, (size_t(2) << 30) / perFrame)));

        // the device
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
            (void)hipGetLastError();
            throw std::runtime_error("framecor: no GPU is available");
        }
        if (device < 0 || device >= count) invalid("no GPU " + std::to_string(device));
        FC_HIP_CHECK(hipSetDevice(device));
        FC_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        FC_HIP_CHECK(hipEventCreateWithFlags(&lastUse, hipEventDisableTiming));
        dDown.ensure(sizeof(int) * ndown);
        dAcross.ensure(sizeof(int) * nacross);
        FC_HIP_CHECK(hipMemcpy(dDown.ptr, down_, sizeof(int) * ndown, hipMemcpyHostToDevice));
        FC_HIP_CHECK(hipMemcpy(dAcross.ptr, across_, sizeof(int) * nacross, hipMemcpyHostToDevice));
        if (method == FRAMECOR_METHOD_FFT) {
            const int L = g.L;
            std::vector<float2> tw(L);
            for (int k = 0; k < L; k++) {
                const double ang = -2.0 * M_PI * double(k) / double(L);
                tw[k] = make_float2(float(std::cos(ang)), float(std::sin(ang)));
            }
            dTwiddle.ensure(sizeof(float2) * L);
            FC_HIP_CHECK(hipMemcpy(dTwiddle.ptr, tw.data(), sizeof(float2) * L, hipMemcpyHostToDevice));
            const size_t np = (L + 1) / 2, rs = 2 * (L / 2 + 1);
            dSpectra.ensure(sizeof(float2) * np * rs * size_t(g.nwin));
            if (!fftInLds(L)) {
                int cus = 0;
                FC_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
                workBlocks = fftWorkBlocks(L, cus);
                dWork.ensure(fftWorkPerBlock(L) * size_t(workBlocks));
            }
        }
    }

    // the global workspace of the FFT form's blocks, when the spectrum does not fit in LDS
    void setWork(FftArgs &a) const
    {
        a.work = dWork.as<float2>();
        a.workStride = fftWorkPerBlock(g.L) / sizeof(float2);
        a.workBlocks = workBlocks;
    }

    WindowData windowData() const
    {
        WindowData wd;
        wd.down = dDown.as<int>();
        wd.across = dAcross.as<int>();
        wd.norm = dNorm.as<float>();
        wd.tsum = dTsum.as<int64_t>();
        wd.tvar = dTvar.as<int64_t>();
        return wd;
    }

    size_t bytesPerFrame() const
    {
        // int8: the screen map {box, 1 / sigma} and box2; fft: sigma
        return size_t(g.boxSize) * (method == FRAMECOR_METHOD_INT8 ? 12 : 4);
    }

    void ensureGroup(int n)
    {
        const int want = std::min(n, maxGroup);
        if (want <= groupCap) return;
        const bool isInt8 = method == FRAMECOR_METHOD_INT8;
        if (isInt8) dBox2.ensure(size_t(want) * g.boxSize * sizeof(int32_t));
        else dSigma.ensure(size_t(want) * g.boxSize * sizeof(float));
        if (isInt8) dScreen.ensure(size_t(want) * g.boxSize * sizeof(int2));
        groupCap = want;
    }

    void mark(hipStream_t s)
    {
        if (!timing) return;
        if (eventsUsed == int(events.size())) {
            hipEvent_t e;
            FC_HIP_CHECK(hipEventCreate(&e));
            events.push_back(e);
        }
        FC_HIP_CHECK(hipEventRecord(events[eventsUsed++], s));
    }

    // one group of frames, all pointers on the device, with the correlator's box-map buffers
    void runGroup(const uint8_t *frames, int n, const int32_t *gross, float *offset, float *peak, float *snr,
                  float *cov, hipStream_t s)
    {
        runGroupWith(frames, n, gross, offset, peak, snr, cov, s, dBox2.as<int32_t>(), dScreen.as<int2>(),
                     dSigma.as<float>(), true);
    }

    // the same with given box-map buffers (int8: box2 and screen; FFT: sigma, n x boxSize each);
    // timed: the per-stage events of set_timing
    void runGroupWith(const uint8_t *frames, int n, const int32_t *gross, float *offset, float *peak, float *snr,
                      float *cov, hipStream_t s, int32_t *box2, int2 *screen, float *sigma, bool timed)
    {
        const bool isInt8 = method == FRAMECOR_METHOD_INT8;
        if (!isInt8) box2 = nullptr, screen = nullptr;
        else sigma = nullptr;
        if (timed) mark(s);
        launchBoxSums(frames, n, g.h, g.w, g.W, nullptr, box2, screen, sigma, s);
        if (timed) mark(s);
        FrameGroup fg;
        fg.frames = frames;
        fg.gross = gross;
        fg.box2 = box2;
        fg.screen = screen;
        fg.sigma = sigma;
        fg.n = n;
        fg.offset = offset;
        fg.peak = peak;
        fg.snr = snr;
        fg.cov = cov;
        if (isInt8) {
            Int8Args a{};
            a.g = g;
            a.wd = windowData();
            a.fg = fg;
            a.tmpl = dTmpl.as<uint8_t>();
            launchInt8(a, i8, variant, s);
        }
        else {
            FftArgs a{};
            a.g = g;
            a.wd = windowData();
            a.fg = fg;
            a.spectra = dSpectra.as<float2>();
            a.twiddle = dTwiddle.as<float2>();
            setWork(a);
            if (fftFastFor(g.L)) {
                launchFftFastCorrelate(a, s);
                if (timed) mark(s);
                return;
            }
            a.plan = plan;
            a.np = (g.L + 1) / 2;
            a.H = g.L / 2 + 1;
            a.rs = 2 * a.H;
            launchFftCorrelate(a, s);
        }
        if (timed) mark(s);
    }

    void correlateDevice(const uint8_t *frames, int n, const int32_t *gross, float *offset, float *peak,
                         float *snr, float *cov, hipStream_t s)
    {
        if (!ready) throw std::runtime_error("framecor: set_reference first");
        if (n < 0) invalid("the number of frames must not be negative");
        eventsUsed = 0;
        if (n == 0) return;
        FC_HIP_CHECK(hipSetDevice(device));
        ensureGroup(n);
        after(s);
        for (int f0 = 0; f0 < n; f0 += groupCap) {
            const int nf = std::min(groupCap, n - f0);
            const size_t q0 = size_t(f0) * g.nwin;
            runGroup(frames + size_t(f0) * g.pixels, nf, gross ? gross + 2 * q0 : nullptr, offset + 2 * q0,
                     peak + q0, snr + q0, cov + 3 * q0, s);
        }
        done(s);
    }

    // correlateDevice with the box maps of each group in a caller's device buffer (scratch, bytes long,
    // 256-byte aligned) instead of the correlator's own: no device allocation and no timing events;
    // groups of as many frames as the buffer holds. A frame whose maps exceed the buffer: EINVAL.
    void correlateScratch(const uint8_t *frames, int n, const int32_t *gross, float *offset, float *peak, float *snr,
                          float *cov, hipStream_t s, void *scratch, size_t bytes)
    {
        if (!ready) throw std::runtime_error("framecor: set_reference first");
        if (n < 0) invalid("the number of frames must not be negative");
        if (method == FRAMECOR_METHOD_FFT && !fftInLds(g.L))
            invalid("the no-allocation entry point does not take FFT search windows whose spectrum does not fit in "
                    "LDS (window + 2 search = " + std::to_string(g.L) + "): they need the correlator's workspace");
        const size_t perFrame = bytesPerFrame();
        if (perFrame > bytes)
            invalid("the box-sum maps of one frame (" + std::to_string(perFrame) + " bytes) exceed the "
                    + std::to_string(bytes >> 20) + " MiB scratch of the no-allocation entry point");
        if (n == 0) return;
        FC_HIP_CHECK(hipSetDevice(device));
        const int G = int(std::min<size_t>(size_t(maxGroup), bytes / perFrame));
        const size_t plane = size_t(G) * g.boxSize;
        unsigned char *base = static_cast<unsigned char *>(scratch);
        int32_t *box2 = nullptr;
        int2 *screen = nullptr;
        float *sigma = nullptr;
        if (method == FRAMECOR_METHOD_INT8) {
            screen = reinterpret_cast<int2 *>(base);
            box2 = reinterpret_cast<int32_t *>(base + plane * sizeof(int2));
        }
        else {
            sigma = reinterpret_cast<float *>(base);
        }
        after(s);
        for (int f0 = 0; f0 < n; f0 += G) {
            const int nf = std::min(G, n - f0);
            const size_t q0 = size_t(f0) * g.nwin;
            runGroupWith(frames + size_t(f0) * g.pixels, nf, gross ? gross + 2 * q0 : nullptr, offset + 2 * q0,
                         peak + q0, snr + q0, cov + 3 * q0, s, box2, screen, sigma, false);
        }
        done(s);
    }

    void correlateHost(const uint8_t *frames, int n, const int32_t *gross, float *offset, float *peak, float *snr,
                       float *cov)
    {
        if (!ready) throw std::runtime_error("framecor: set_reference first");
        if (n < 0) invalid("the number of frames must not be negative");
        eventsUsed = 0;
        if (n == 0) return;
        FC_HIP_CHECK(hipSetDevice(device));
        ensureGroup(n);
        const int G = groupCap;
        dFrames.ensure(size_t(G) * g.pixels);
        dOffset.ensure(size_t(G) * g.nwin * 2 * sizeof(float));
        dPeak.ensure(size_t(G) * g.nwin * sizeof(float));
        dSnr.ensure(size_t(G) * g.nwin * sizeof(float));
        dCov.ensure(size_t(G) * g.nwin * 3 * sizeof(float));
        if (gross) dGross.ensure(size_t(G) * g.nwin * 2 * sizeof(int32_t));
        after(stream);
        for (int f0 = 0; f0 < n; f0 += G) {
            const int nf = std::min(G, n - f0);
            const size_t q0 = size_t(f0) * g.nwin, pairs = size_t(nf) * g.nwin;
            FC_HIP_CHECK(hipMemcpyAsync(dFrames.ptr, frames + size_t(f0) * g.pixels, size_t(nf) * g.pixels,
                                        hipMemcpyHostToDevice, stream));
            if (gross)
                FC_HIP_CHECK(hipMemcpyAsync(dGross.ptr, gross + 2 * q0, pairs * 2 * sizeof(int32_t),
                                            hipMemcpyHostToDevice, stream));
            runGroup(dFrames.as<uint8_t>(), nf, gross ? dGross.as<int32_t>() : nullptr, dOffset.as<float>(),
                     dPeak.as<float>(), dSnr.as<float>(), dCov.as<float>(), stream);
            FC_HIP_CHECK(hipMemcpyAsync(offset + 2 * q0, dOffset.ptr, pairs * 2 * sizeof(float), hipMemcpyDeviceToHost, stream));
            FC_HIP_CHECK(hipMemcpyAsync(peak + q0, dPeak.ptr, pairs * sizeof(float), hipMemcpyDeviceToHost, stream));
            FC_HIP_CHECK(hipMemcpyAsync(snr + q0, dSnr.ptr, pairs * sizeof(float), hipMemcpyDeviceToHost, stream));
            FC_HIP_CHECK(hipMemcpyAsync(cov + 3 * q0, dCov.ptr, pairs * 3 * sizeof(float), hipMemcpyDeviceToHost, stream));
            FC_HIP_CHECK(hipStreamSynchronize(stream));
        }
        done(stream);
    }

    DiffArgs diffArgs(const float *frames, int n, const int32_t *gross, const float *shifts, float *ncc) const
    {
        if (!ready) throw std::runtime_error("framecor: set_reference first");
        if (n < 0) invalid("the number of frames must not be negative");
        DiffArgs d{};
        d.g = g;
        d.wd = windowData();
        d.tmpl = dTmpl.as<uint8_t>();
        d.frames = frames;
        d.n = n;
        d.gross = gross;
        d.shifts = shifts;
        d.ncc = ncc;
        return d;
    }

    // every window's search window (no gross offset) must lie in the frame
    void requireSearchInFrame(const char *what) const
    {
        for (int y : down)
            if (y - g.R < 0 || y - g.R + g.L > g.h)
                invalid(std::string(what) + ": the search window of the windows at row " + std::to_string(y)
                        + " leaves the frame");
        for (int x : across)
            if (x - g.R < 0 || x - g.R + g.L > g.w)
                invalid(std::string(what) + ": the search window of the windows at column " + std::to_string(x)
                        + " leaves the frame");
    }

    // the box maps of one frame on the device, for the surfaces
    // (dSurfBox, dSurfBox2: ensured by the caller, before it orders its stream)
    void surfaceBoxes(const uint8_t *frame, hipStream_t s)
    {
        launchBoxSums(frame, 1, g.h, g.w, g.W, dSurfBox.as<int32_t>(), dSurfBox2.as<int32_t>(), nullptr, nullptr, s);
    }

    void launchSurfaces(const uint8_t *frame, const int32_t *gross, int k0, int count, double *surf, hipStream_t s)
    {
        SurfaceArgs a{};
        a.g = g;
        a.wd = windowData();
        a.frame = frame;
        a.gross = gross;
        a.tmpl = dTmpl.as<uint8_t>();
        a.box = dSurfBox.as<int32_t>();
        a.box2 = dSurfBox2.as<int32_t>();
        a.surf = surf;
        a.k0 = k0;
        a.count = count;
        launchNccSurfaces(a, s);
    }

    // inFrame: every search window must lie in the frame (no gross offsets)
    void checkSurfaces(const char *what, bool inFrame) const
    {
        if (!ready) throw std::runtime_error("framecor: set_reference first");
        if (inFrame) requireSearchInFrame(what);
        if (surfaceShmem(g) > 160 * 1024)
            invalid(std::string(what) + ": the search window (window + 2 search = " + std::to_string(g.L)
                    + ") is too large for the NCC surfaces");
    }

    // the exact FP64 NCC surfaces (count x M x M) of windows k0 .. k0 + count - 1 against a frame on the
    // device, each search window at its window's position plus its gross offset (nwin x 2 on the device,
    // or null); NaN surfaces where a search window leaves the frame
    void surfaces(const uint8_t *frame, const int32_t *gross, int k0, int count, double *surf, hipStream_t s)
    {
        checkSurfaces("surfaces", false);
        if (k0 < 0 || count < 0 || count > g.nwin - k0) invalid("the windows must be windows of the grid");
        if (count == 0) return;
        FC_HIP_CHECK(hipSetDevice(device));
        dSurfBox.ensure(size_t(g.boxSize) * sizeof(int32_t));
        dSurfBox2.ensure(size_t(g.boxSize) * sizeof(int32_t));
        after(s);
        surfaceBoxes(frame, s);
        launchSurfaces(frame, gross, k0, count, surf, s);
        done(s);
    }

    int surfaceChunk() const
    {
        const size_t perWindow = size_t(g.M) * g.M * sizeof(double);
        return int(std::max<size_t>(1, std::min<size_t>(g.nwin, (size_t(256) << 20) / perWindow)));
    }

    // the peak of every window's NCC surface against a frame on the device, as
    // python/slipcam/stabilize.py:22-45 ncc_peak: offsets (nwin x 2 doubles: down, across, relative to
    // the window's position plus its gross offset) and peak (nwin doubles), on the device
    void peaksDevice(const uint8_t *frame, const int32_t *gross, double *offsets, double *peak, hipStream_t s)
    {
        checkSurfaces("ncc peaks", false);
        FC_HIP_CHECK(hipSetDevice(device));
        const int chunk = surfaceChunk();
        dSurf.ensure(size_t(chunk) * g.M * g.M * sizeof(double));
        dSurfBox.ensure(size_t(g.boxSize) * sizeof(int32_t));
        dSurfBox2.ensure(size_t(g.boxSize) * sizeof(int32_t));
        after(s);
        surfaceBoxes(frame, s);
        for (int k0 = 0; k0 < g.nwin; k0 += chunk) {
            const int cnt = std::min(chunk, g.nwin - k0);
            launchSurfaces(frame, gross, k0, cnt, dSurf.as<double>(), s);
            launchParabolaPeaks(dSurf.as<double>(), cnt, g.M, dTvar.as<int64_t>(), k0, offsets, peak, s);
        }
        done(s);
    }

    // the same with host arrays (synchronous)
    void peaksHost(const uint8_t *frame, const int32_t *gross, double *offsets, double *peak)
    {
        checkSurfaces("ncc peaks", false);
        FC_HIP_CHECK(hipSetDevice(device));
        dFrames.ensure(size_t(g.pixels));
        dPer.ensure(sizeof(double) * 3 * size_t(g.nwin));
        if (gross) dGross.ensure(sizeof(int32_t) * 2 * size_t(g.nwin));
        after(stream);
        FC_HIP_CHECK(hipMemcpyAsync(dFrames.ptr, frame, size_t(g.pixels), hipMemcpyHostToDevice, stream));
        if (gross) {
            FC_HIP_CHECK(hipMemcpyAsync(dGross.ptr, gross, sizeof(int32_t) * 2 * size_t(g.nwin), hipMemcpyHostToDevice,
                                        stream));
        }
        double *dOff = dPer.as<double>(), *dPk = dOff + 2 * size_t(g.nwin);
        peaksDevice(dFrames.as<uint8_t>(), gross ? dGross.as<int32_t>() : nullptr, dOff, dPk, stream);
        FC_HIP_CHECK(hipMemcpyAsync(offsets, dOff, sizeof(double) * 2 * size_t(g.nwin), hipMemcpyDeviceToHost, stream));
        FC_HIP_CHECK(hipMemcpyAsync(peak, dPk, sizeof(double) * size_t(g.nwin), hipMemcpyDeviceToHost, stream));
        FC_HIP_CHECK(hipStreamSynchronize(stream));
        done(stream);
    }

    // the second peak of every window's NCC surface against a frame on the host (the periodicity of
    // python/slipcam/auto_slip.py:105-129 when the frame is the reference): the largest local maximum
    // outside the square of rows and columns search - exclude .. search + exclude, bounds as a Python
    // slice of the surface
    void periodicity(const uint8_t *frame, int exclude, double *out)
    {
        checkSurfaces("periodicity", true);
        FC_HIP_CHECK(hipSetDevice(device));
        const int M = g.M;
        auto bound = [M](int64_t i) {
            if (i < 0) i = std::max<int64_t>(i + M, 0);
            return int(std::min<int64_t>(i, M));
        };
        const int lo = bound(int64_t(g.R) - exclude), hi = bound(int64_t(g.R) + exclude + 1);
        const int chunk = surfaceChunk();
        dFrames.ensure(size_t(g.pixels));
        dSurf.ensure(size_t(chunk) * M * M * sizeof(double));
        dPer.ensure(sizeof(double) * g.nwin);
        dSurfBox.ensure(size_t(g.boxSize) * sizeof(int32_t));
        dSurfBox2.ensure(size_t(g.boxSize) * sizeof(int32_t));
        after(stream);
        FC_HIP_CHECK(hipMemcpyAsync(dFrames.ptr, frame, size_t(g.pixels), hipMemcpyHostToDevice, stream));
        surfaceBoxes(dFrames.as<uint8_t>(), stream);
        for (int k0 = 0; k0 < g.nwin; k0 += chunk) {
            const int cnt = std::min(chunk, g.nwin - k0);
            launchSurfaces(dFrames.as<uint8_t>(), nullptr, k0, cnt, dSurf.as<double>(), stream);
            launchSecondPeak(dSurf.as<double>(), cnt, M, lo, hi, dTvar.as<int64_t>(), k0, dPer.as<double>(), stream);
        }
        FC_HIP_CHECK(hipMemcpyAsync(out, dPer.ptr, sizeof(double) * g.nwin, hipMemcpyDeviceToHost, stream));
        FC_HIP_CHECK(hipStreamSynchronize(stream));
        done(stream);
    }

    void setReference(const uint8_t *frame)
    {
        FC_HIP_CHECK(hipSetDevice(device));
        const int W = g.W, nwin = g.nwin, n = W * W;
        std::vector<uint8_t> tmpl(size_t(nwin) * n);
        std::vector<float> norm(nwin), templ;
        std::vector<int64_t> tsum(nwin), tvar(nwin);
        const bool fft = method == FRAMECOR_METHOD_FFT;
        if (fft) templ.resize(size_t(nwin) * n);
        parallelFor(nwin, [&](int k) {
            const int y0 = down[k / g.nx], x0 = across[k % g.nx];
            uint8_t *t = tmpl.data() + size_t(k) * n;
            int64_t s = 0, s2 = 0;
            for (int i = 0; i < W; i++)
                for (int j = 0; j < W; j++) {
                    const int64_t px = frame[size_t(y0 + i) * g.w + x0 + j];
                    t[i * W + j] = uint8_t(px);
                    s += px;
                    s2 += px * px;
                }
            tsum[k] = s;
            tvar[k] = int64_t(n) * s2 - s * s;
            // the zero-mean window and its norm, as cxx/common.cpp:30-48
            const double mean = double(s) / (double(W) * W);
            double sq = 0.0;
            for (int i = 0; i < n; i++) {
                const float d = float(t[i] - mean);
                if (fft) templ[size_t(k) * n + i] = d;
                sq += double(d) * d;
            }
            norm[k] = float(std::sqrt(sq));
        });
        dTmpl.ensure(tmpl.size());
        dNorm.ensure(sizeof(float) * nwin);
        dTsum.ensure(sizeof(int64_t) * nwin);
        dTvar.ensure(sizeof(int64_t) * nwin);
        // the templates may still be read by an earlier call's kernels on another stream
        after(stream);
        FC_HIP_CHECK(hipMemcpyAsync(dTmpl.ptr, tmpl.data(), tmpl.size(), hipMemcpyHostToDevice, stream));
        FC_HIP_CHECK(hipMemcpyAsync(dNorm.ptr, norm.data(), sizeof(float) * nwin, hipMemcpyHostToDevice, stream));
        FC_HIP_CHECK(hipMemcpyAsync(dTsum.ptr, tsum.data(), sizeof(int64_t) * nwin, hipMemcpyHostToDevice, stream));
        FC_HIP_CHECK(hipMemcpyAsync(dTvar.ptr, tvar.data(), sizeof(int64_t) * nwin, hipMemcpyHostToDevice, stream));
        if (fft) {
            DeviceBuffer dTempl;
            dTempl.ensure(sizeof(float) * templ.size());
            FC_HIP_CHECK(hipMemcpyAsync(dTempl.ptr, templ.data(), sizeof(float) * templ.size(), hipMemcpyHostToDevice, stream));
            FftArgs a{};
            a.g = g;
            a.wd = windowData();
            a.twiddle = dTwiddle.as<float2>();
            a.templ = dTempl.as<float>();
            a.spectraOut = dSpectra.as<float2>();
            // a spectrum that does not fit in LDS is transformed in place in the output
            a.work = fftInLds(g.L) ? nullptr : dSpectra.as<float2>();
            a.plan = plan;
            a.np = (g.L + 1) / 2;
            a.H = g.L / 2 + 1;
            a.rs = 2 * a.H;
            if (fftFastFor(g.L))
                launchFftFastReference(a, stream);
            else
                launchFftReference(a, stream);
            FC_HIP_CHECK(hipStreamSynchronize(stream));
        }
        else {
            FC_HIP_CHECK(hipStreamSynchronize(stream));
        }
        done(stream);
        ready = true;
    }
};

