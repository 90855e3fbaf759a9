// capi.hip -- the C ABI of libframecor_hip.so (include/framecor_hip.h) over the correlator object
// of correlator.h

#include "correlator.h"

#ifndef FRAMECOR_VERSION
#define FRAMECOR_VERSION "0.1.0"
#endif

using namespace framecor;

extern "C" {

int framecor_check_params(int height, int width, const int *down, int ndown, const int *across, int nacross,
                          int window, int search, int method, int batch, int stat_size)
{
    return guarded([&] { checkAll(height, width, down, ndown, across, nacross, window, search, method, batch, stat_size); });
}


const char *framecor_version(void) { return FRAMECOR_VERSION; }

int framecor_has_gpu(void)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return count > 0;
}

const char *framecor_last_error(void) { return lastError().c_str(); }

int framecor_create(int height, int width, const int *down, int ndown, const int *across, int nacross, int window,
                    int search, int method, int batch, int stat_size, int device, framecor_correlator **out)
{
    if (!out) return fail(FRAMECOR_EINVAL, "framecor: no output handle");
    *out = nullptr;
    return guarded([&] {
        auto c = std::make_unique<framecor_correlator>();
        c->init(height, width, down, ndown, across, nacross, window, search, method, batch, stat_size, device);
        scratchSetup(device);
        *out = c.release();
    });
}

void framecor_destroy(framecor_correlator *c) { delete c; }

const char *framecor_backend(const framecor_correlator *c) { return c ? "hip" : ""; }

int framecor_shape(const framecor_correlator *c, int *ndown, int *nacross)
{
    if (!c || !ndown || !nacross) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    *ndown = c->g.ny;
    *nacross = c->g.nx;
    return FRAMECOR_OK;
}

int framecor_set_reference(framecor_correlator *c, const uint8_t *frame)
{
    if (!c || !frame) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    return guarded([&] { c->setReference(frame); });
}

int framecor_correlate(framecor_correlator *c, const uint8_t *frames, int n, const int32_t *gross, float *offset,
                       float *peak, float *snr, float *cov)
{
    if (!c) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    if (n > 0 && (!frames || !offset || !peak || !snr || !cov)) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    return guarded([&] { c->correlateHost(frames, n, gross, offset, peak, snr, cov); });
}

int framecor_correlate_device(framecor_correlator *c, const uint8_t *frames, int n, const int32_t *gross,
                              float *offset, float *peak, float *snr, float *cov, void *stream)
{
    if (!c) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    if (n > 0 && (!frames || !offset || !peak || !snr || !cov)) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    return guarded([&] {
        c->correlateDevice(frames, n, gross, offset, peak, snr, cov,
                           static_cast<hipStream_t>(stream));
    });
}

int framecor_ncc_forward(framecor_correlator *c, const float *frames, int n, const int32_t *gross,
                         const float *shifts, float *ncc, void *stream)
{
    if (!c) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    if (n > 0 && (!frames || !shifts || !ncc)) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    return guarded([&] {
        DiffArgs d = c->diffArgs(frames, n, gross, shifts, ncc);
        FC_HIP_CHECK(hipSetDevice(c->device));
        const hipStream_t s = static_cast<hipStream_t>(stream);
        c->after(s);
        launchDiff(d, DiffOp::Forward, s);
        c->done(s);
    });
}

int framecor_ncc_backward(framecor_correlator *c, const float *frames, int n, const int32_t *gross,
                          const float *shifts, const float *grad_ncc, float *grad_frames, float *grad_shifts,
                          void *stream)
{
    if (!c) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    if (n > 0 && (!frames || !shifts || !grad_ncc || !grad_frames || !grad_shifts))
        return fail(FRAMECOR_EINVAL, "framecor: null argument");
    return guarded([&] {
        DiffArgs d = c->diffArgs(frames, n, gross, shifts, nullptr);
        d.gradNcc = grad_ncc;
        d.gradFrames = grad_frames;
        d.gradShifts = grad_shifts;
        FC_HIP_CHECK(hipSetDevice(c->device));
        const hipStream_t s = static_cast<hipStream_t>(stream);
        c->after(s);
        launchDiff(d, DiffOp::Backward, s);
        c->done(s);
    });
}

int framecor_ncc_refine(framecor_correlator *c, const float *frames, int n, const int32_t *gross,
                        const float *shifts_in, int steps, float *shifts_out, float *ncc, void *stream)
{
    if (!c) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    if (n > 0 && (!frames || !shifts_in || !shifts_out || !ncc)) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    if (steps < 0 || steps > 100) return fail(FRAMECOR_EINVAL, "framecor: steps must be in 0..100");
    return guarded([&] {
        DiffArgs d = c->diffArgs(frames, n, gross, shifts_in, ncc);
        d.shiftsOut = shifts_out;
        d.steps = steps;
        FC_HIP_CHECK(hipSetDevice(c->device));
        const hipStream_t s = static_cast<hipStream_t>(stream);
        c->after(s);
        launchDiff(d, DiffOp::Refine, s);
        c->done(s);
    });
}

int framecor_set_variant(framecor_correlator *c, int variant)
{
    if (!c) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    if (variant < 0 || variant > 6) return fail(FRAMECOR_EINVAL, "framecor: the kernel variant must be 0 to 6");
    c->variant = static_cast<Int8Variant>(variant);
    return FRAMECOR_OK;
}

int framecor_set_timing(framecor_correlator *c, int enable)
{
    if (!c) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    c->timing = enable != 0;
    c->eventsUsed = 0;
    return FRAMECOR_OK;
}

int framecor_get_timing(framecor_correlator *c, float *ms, int *launches)
{
    if (!c || !ms || !launches) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    return guarded([&] {
        ms[0] = ms[1] = 0.0f;
        *launches = c->eventsUsed / 3;
        if (c->eventsUsed == 0) return;
        FC_HIP_CHECK(hipSetDevice(c->device));
        FC_HIP_CHECK(hipEventSynchronize(c->events[c->eventsUsed - 1]));
        for (int i = 0; i + 2 < c->eventsUsed; i += 3)
            for (int s = 0; s < 2; s++) {
                float t = 0.0f;
                FC_HIP_CHECK(hipEventElapsedTime(&t, c->events[i + s], c->events[i + s + 1]));
                ms[s] += t;
            }
    });
}

int framecor_summed_area_device(const uint8_t *frames, int n, int height, int width, int64_t *sat, int64_t *sat2,
                                void *stream)
{
    return guarded([&] {
        if (n < 0 || height <= 0 || width <= 0) invalid("the frame size must be positive");
        if (n > 65535) invalid("at most 65535 frames per call");
        if (n > 0 && (!frames || !sat || !sat2)) invalid("null argument");
        launchSummedArea(frames, n, height, width, sat, sat2, static_cast<hipStream_t>(stream));
    });
}

int framecor_ncc_surfaces_device(framecor_correlator *c, const uint8_t *frame, const int32_t *gross, int k0, int count,
                                 double *surfaces, void *stream)
{
    if (!c) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    if (count > 0 && (!frame || !surfaces)) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    return guarded([&] { c->surfaces(frame, gross, k0, count, surfaces, static_cast<hipStream_t>(stream)); });
}

int framecor_ncc_peaks(framecor_correlator *c, const uint8_t *frame, const int32_t *gross, double *offsets,
                       double *peak)
{
    if (!c || !frame || !offsets || !peak) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    return guarded([&] { c->peaksHost(frame, gross, offsets, peak); });
}

int framecor_ncc_peaks_device(framecor_correlator *c, const uint8_t *frame, const int32_t *gross, double *offsets,
                              double *peak, void *stream)
{
    if (!c || !frame || !offsets || !peak) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    return guarded([&] { c->peaksDevice(frame, gross, offsets, peak, static_cast<hipStream_t>(stream)); });
}

int framecor_periodicity(framecor_correlator *c, const uint8_t *frame, int exclude, double *out)
{
    if (!c || !frame || !out) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    return guarded([&] { c->periodicity(frame, exclude, out); });
}

int framecor_box_sums_device(const uint8_t *frames, int n, int height, int width, int window, int32_t *box,
                             int32_t *box2, void *stream)
{
    return guarded([&] {
        if (n < 0 || height <= 0 || width <= 0) invalid("the frame size must be positive");
        if (window < 1 || window > 181 || window > height || window > width)
            invalid("the window must be 1 .. 181 pixels and fit in the frame");
        if (n > 65535) invalid("at most 65535 frames per call");
        if (n > 0 && (!frames || !box || !box2)) invalid("null argument");
        launchBoxSums(frames, n, height, width, window, box, box2, nullptr, nullptr, static_cast<hipStream_t>(stream));
    });
}

} // extern "C"
