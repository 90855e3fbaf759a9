/*
 * framecor_hip.h -- the C ABI of the MI355X framecor correlator (libframecor_hip.so)
 *
 * Dense normalized cross-correlation of 8-bit frames with a reference frame on a grid of square
 * windows: the drop-in for the reference's native module slipcam.framecor._framecor, whose pybind11
 * class the Python package slipcam_amd.framecor rebuilds on top of these entry points (ctypes).
 * Plain C types only: pointers, sizes and status codes.
 *
 * Every entry point returns FRAMECOR_OK (0) or an error status; framecor_last_error() then holds the
 * message (per thread). The statuses map to the exceptions of the reference module:
 *   FRAMECOR_EINVAL   -> ValueError    (std::invalid_argument: cxx/common.cpp:17, cxx/bindings.cpp:44,50,61,79)
 *   FRAMECOR_ERUNTIME -> RuntimeError  (std::runtime_error: cxx/bindings.cpp:59, cxx/cuda.cu:28-38)
 *   FRAMECOR_ENOMEM   -> MemoryError   (std::bad_alloc: cxx/cpu.cpp:67)
 *
 * Outputs, for n frames and nwin = ndown * nacross windows, q = frame * nwin + window (windows of a
 * grid row, then the next row), as cxx/framecor.h:37-43:
 *   offset[2 q + 0..1]  (down, across) motion of the frame, NaN if not valid
 *   peak[q]             normalized correlation at the peak (0 if not valid)
 *   snr[q]              squared peak over the mean squared surface around it (0 if not valid)
 *   cov[3 q + 0..2]     covariance of the offsets (down, across, cross); (99, 99, 0) if unknown
 *
 * A correlator is not re-entrant (it owns its tables and stream), as the reference's.
 */
#ifndef FRAMECOR_HIP_H
#define FRAMECOR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FRAMECOR_OK = 0,
    FRAMECOR_EINVAL = 1,
    FRAMECOR_ERUNTIME = 2,
    FRAMECOR_ENOMEM = 3
};

/* the correlation methods (a keyword the reference does not have; it only has the FFT) */
enum {
    FRAMECOR_METHOD_FFT = 0,    /* FP32 FFT, the reference algorithm */
    FRAMECOR_METHOD_INT8 = 1    /* exact int8-MFMA correlation, FP64 peak fit */
};

typedef struct framecor_correlator framecor_correlator;

/* the version string; replaces _framecor.__version__ (cxx/bindings.cpp:113, CMakeLists.txt:12) */
const char *framecor_version(void);

/* 1 if a GPU is visible; replaces has_cuda() (cxx/framecor.h:63, cxx/cuda.cu:383-387, cxx/bindings.cpp:115) */
int framecor_has_gpu(void);

/* the message of the last failing call on this thread ("" if none) */
const char *framecor_last_error(void);

/* checks the parameters of a correlator without touching a GPU: FRAMECOR_EINVAL with the message
 * of the reference's checkParams (cxx/common.cpp:15-28) if they are not valid, or if the method does
 * not support them */
int framecor_check_params(int height, int width, const int *down, int ndown, const int *across, int nacross,
                          int window, int search, int method, int batch, int stat_size);

/* a correlator of height x width frames on the grid down[ndown] x across[nacross] of windows;
 * replaces PyCorrelator::PyCorrelator (cxx/bindings.cpp:26-45) with checkParams
 * (cxx/common.cpp:15-28) and makeCuda (cxx/cuda.cu:389-393). The parameters are checked before
 * any GPU call (FRAMECOR_EINVAL with the reference's messages). batch (>= 1) and stat_size (>= 3)
 * are the reference's keywords; method is FRAMECOR_METHOD_*. */
int framecor_create(int height, int width, const int *down, int ndown, const int *across, int nacross,
                    int window, int search, int method, int batch, int stat_size, int device,
                    framecor_correlator **out);

/* frees the correlator and its device memory (NULL is ignored) */
void framecor_destroy(framecor_correlator *c);

/* "hip"; replaces Correlator::backend (cxx/framecor.h:55, cxx/cuda.cu:299, cxx/bindings.cpp:99) */
const char *framecor_backend(const framecor_correlator *c);

/* the window grid (ndown, nacross); replaces the shape property (cxx/bindings.cpp:106, 139) */
int framecor_shape(const framecor_correlator *c, int *ndown, int *nacross);

/* sets the reference frame (height x width uint8, host memory) and caches what its windows need;
 * may be called again to replace it. Replaces PyCorrelator::setReference (cxx/bindings.cpp:47-55)
 * and Cuda::setReference (cxx/cuda.cu:301-315). Synchronous. */
int framecor_set_reference(framecor_correlator *c, const uint8_t *frame);

/* correlates n frames (n x height x width uint8, host memory) with the reference frame; gross:
 * n x nwin x 2 int32 (down, across) integer offsets of the search windows, or NULL; outputs in host
 * memory. Replaces PyCorrelator::correlate's call of Correlator::correlate (cxx/bindings.cpp:86-90,
 * cxx/framecor.h:52-53, cxx/cuda.cu:317-360). Synchronous. FRAMECOR_ERUNTIME before a reference. */
int framecor_correlate(framecor_correlator *c, const uint8_t *frames, int n, const int32_t *gross,
                       float *offset, float *peak, float *snr, float *cov);

/* the same with every array in device memory, enqueued on stream (a hipStream_t; NULL: the null
 * stream) and not synchronized: the outputs are complete when the stream is. The
 * reference has no device-resident entry point; this is the one its CUDA backend's batch loop
 * (cxx/cuda.cu:329-359) runs between its copies. */
int framecor_correlate_device(framecor_correlator *c, const uint8_t *frames, int n, const int32_t *gross,
                              float *offset, float *peak, float *snr, float *cov, void *stream);

/* framecor_correlate_device with no device allocation. No counterpart in the reference: for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The box-sum maps of each group of frames go to a static 64 MiB device array per GPU,
 * held until the process ends, instead of buffers the correlator allocates; frames are correlated in
 * groups whose maps fit in it (12 bytes per window position and frame for the int8 method, 4 for the
 * FFT: 64 MiB holds 6 or 21 frames at 720 x 1280, window 64). FRAMECOR_EINVAL if one frame's maps
 * exceed 64 MiB, and for the FFT method when its spectra do not fit in LDS (window + 2 search above
 * 197: the correlator's own workspace is needed). Every call on a GPU shares that GPU's array: calls are ordered on the GPU (each waits for the
 * previous one's work on it), except under stream capture, where captured work that overlaps other
 * calls on the same GPU races. The correlator must have been created and given its reference (those
 * calls allocate). Arguments and results as framecor_correlate_device. */
int framecor_correlate_device_noalloc(framecor_correlator *c, const uint8_t *frames, int n, const int32_t *gross,
                                      float *offset, float *peak, float *snr, float *cov, void *stream);

/* The differentiable form (no counterpart in the reference). For n float32 frames (device memory),
 * the windows of the correlator's reference and a continuous shift per window, shifts n x nwin x 2
 * (down, across, relative to the window plus its gross offset; gross as in framecor_correlate, or
 * NULL), the chip of each window is sampled bilinearly and compared with the reference window:
 *   framecor_ncc_forward   ncc[q], the NCC (0 where the chip leaves the frame or is flat)
 *   framecor_ncc_backward  given grad_ncc[q]: adds dL/dframes into grad_frames (n x h x w) and writes
 *                          dL/dshifts into grad_shifts (n x nwin x 2)
 *   framecor_ncc_refine    steps Gauss-Newton steps from shifts_in: shifts_out and their ncc (NaN / 0
 *                          where the chip leaves the frame)
 * All arrays in device memory, enqueued on stream (NULL: the null stream), not synchronized. */
int framecor_ncc_forward(framecor_correlator *c, const float *frames, int n, const int32_t *gross,
                         const float *shifts, float *ncc, void *stream);
int framecor_ncc_backward(framecor_correlator *c, const float *frames, int n, const int32_t *gross,
                          const float *shifts, const float *grad_ncc, float *grad_frames, float *grad_shifts,
                          void *stream);
int framecor_ncc_refine(framecor_correlator *c, const float *frames, int n, const int32_t *gross,
                        const float *shifts_in, int steps, float *shifts_out, float *ncc, void *stream);

/* the kernel of the int8 method (no counterpart in the reference; a tuning knob): 0 automatic (as 6
 * where it applies and the window is at most 33, else as 5 where it applies, else as 1), 1 one frame
 * at a time on 16 x 16 tiles, 2 the pipelined producer/helper kernel (falls back to 1 where its LDS
 * buffers do not fit), 3 one frame at a time with ceil(M / 16) full 16-row bands (where M = 2 search + 1 = 16 c + 1, variant 1 computes c x c tiles plus
 * the edge row and the edge column in a wave each), 4 one frame at a time on 32 x 32 tiles
 * (v_mfma_i32_32x32x32_i8) where M = 32 c + 1, c <= 4 (search 16, 32, 48, 64; else as 1), 5 one frame
 * at a time in blocks of four waves with equal MFMA shares where M = 16 c + 1, c = 4 or 6 (search 32 or
 * 48) and the window is 32 or 64 (else as 1), 6 one wave per window and frame where the window is at
 * most 64 and M = 16 c + 1 <= 65 (search 8, 16, 24 or 32: c bands, the edge row and column) or M <= 48
 * (search up to 23: full bands) (else as 1); the offsets, peaks and covariances are the same, the SNR
 * to rounding */
int framecor_set_variant(framecor_correlator *c, int variant);

/* per-stage GPU timing (no counterpart in the reference; for benchmarks): with enable 1, the
 * correlate calls record hipEvents on their stream around each stage; framecor_get_timing waits for
 * the last call's events and writes the milliseconds of its stages, summed over its groups of frames:
 * ms[0] the box-sum maps, ms[1] the correlation kernel; and *launches, the number of
 * correlation-kernel launches */
int framecor_set_timing(framecor_correlator *c, int enable);
int framecor_get_timing(framecor_correlator *c, float *ms, int *launches);

/* the (h + 1) x (w + 1) int64 summed-area tables of n frames (device memory) and of their squares,
 * row and column 0 zero; replaces satRows/satColumns (cxx/cuda.cu:72-110, cxx/cpu.cpp:24-42).
 * Enqueued on stream (NULL: the null stream), not synchronized. */
int framecor_summed_area_device(const uint8_t *frames, int n, int height, int width, int64_t *sat,
                                int64_t *sat2, void *stream);

/* ---- batches: many clips of uneven sizes in one hipGraph launch (no counterpart in the reference,
 * whose correlator takes one clip; each clip here is what one reference Correlator(...) would be:
 * cxx/bindings.cpp:26-45, with a fixed number of frames per call) */

typedef struct framecor_batch framecor_batch;

typedef struct framecor_clip {
    int height, width;          /* frame size */
    const int *down;            /* first pixels of the rows of windows (ndown) */
    int ndown;
    const int *across;          /* first pixels of the columns of windows (nacross) */
    int nacross;
    int window, search;         /* window size and half search range */
    int stat_size;              /* the SNR window (>= 3; the reference's default is 21) */
    int frames;                 /* frames of this clip per call */
} framecor_clip;

/* a batch of nclips clips correlated with one method on one GPU; each clip's parameters are checked
 * as framecor_create's (FRAMECOR_EINVAL, the message naming the clip). Clips of the same tile shape
 * (int8: window, search range and SNR box side min(stat_size, 2 search + 1); fft: window + 2 search)
 * share one persistent work-queue kernel. */
int framecor_batch_create(const framecor_clip *clips, int nclips, int method, int device, framecor_batch **out);

/* frees the batch and its device memory (NULL is ignored) */
void framecor_batch_destroy(framecor_batch *b);

/* sets the reference frame of clip (its height x width uint8, host memory); as framecor_set_reference */
int framecor_batch_set_reference(framecor_batch *b, int clip, const uint8_t *frame);

/* enable 1: the next correlate call first times the candidate plans of every tile shape that has none
 * (on a sample of its frames) and keeps the fastest; 0 (the default): shapes without a plan use the
 * automatic one. Every plan gives the same offsets, peaks and covariances (the SNR to rounding). */
int framecor_batch_set_tuning(framecor_batch *b, int enable);

/* the plan of clip's tile shape: plan[0] the int8 kernel variant (as framecor_set_variant; resolved:
 * 1 edge waves, 2 pipelined, 3 full bands, 4 32 x 32 tiles, 5 four-wave blocks, 6 one wave per window),
 * plan[1] frames per work item (0: automatic), plan[2] resident blocks per CU (0: as many as fit),
 * plan[3] 1 if tuned or set */
int framecor_batch_plan(const framecor_batch *b, int clip, int *plan);

/* sets the plan of clip's tile shape (plan[0..2] as above) */
int framecor_batch_set_plan(framecor_batch *b, int clip, const int *plan);

/* at most frames (>= 1) frames of each clip per stage (the default: every frame in one stage where the
 * box-sum maps of all clips fit in a quarter of the free device memory, at most 64 GiB; else what 2 GiB
 * of box-sum maps per clip hold, at most 256); a call runs the stages one after the other, each one
 * launch of the box sums of every clip, then the work-queue kernels */
int framecor_batch_set_stage_frames(framecor_batch *b, int frames);

/* info[0] tile shapes, info[1] stages of frames, info[2] nodes of the captured graph (0 before the first
 * call), info[3] work queues per call */
int framecor_batch_info(const framecor_batch *b, int *info);

/* correlates every clip: frames[c] (frames_c x height_c x width_c uint8), gross[c] (frames_c x nwin_c x 2
 * int32, or NULL; gross itself may be NULL) and the outputs offset[c], peak[c], snr[c], cov[c] of clip c
 * as framecor_correlate's, all in device memory. One graph launch enqueued on stream (NULL: the null
 * stream), not synchronized; the graph is captured on the first call and again when a pointer changes.
 * FRAMECOR_ERUNTIME before every clip has a reference. Not re-entrant. */
int framecor_batch_correlate_device(framecor_batch *b, const uint8_t *const *frames, const int32_t *const *gross,
                                    float *const *offset, float *const *peak, float *const *snr, float *const *cov,
                                    void *stream);

/* the W x W box sums of n frames (device memory) at every window position: box and box2,
 * n x (height - window + 1) x (width - window + 1) int32, the sums of the pixels and of their squares;
 * what boxSum (cxx/surface.h:21-25) reads from the summed-area tables, computed here straight from
 * the pixels (the correlator's first stage). window 1 .. 181. Enqueued on stream (NULL: the null
 * stream), not synchronized. */
int framecor_box_sums_device(const uint8_t *frames, int n, int height, int width, int window, int32_t *box,
                             int32_t *box2, void *stream);

/* The periodicity of every window (python/slipcam/auto_slip.py:105-129 `periodicity`): the NCC of the
 * window's reference template (set_reference first) with its own search window in frame (h x w uint8,
 * host; the reference frame for the periodicity), each window at its own position; out[k] (nwin
 * doubles, host, k = row x nx + column) = the largest value of the surface at a local maximum (equal to
 * the largest value of its 5 x 5 neighbourhood clipped to the surface, as scipy's maximum_filter with
 * size 5) outside the rows and columns search - exclude .. search + exclude (bounds as a Python slice
 * of the surface), 0 where no such point exists, 1 where the template is flat. The NCC is exact (integer
 * sums, FP64 normalization), which the reference evaluates with FP64 FFTs. Every search window must lie
 * in the frame (else FRAMECOR_EINVAL). Synchronous. */
int framecor_periodicity(framecor_correlator *c, const uint8_t *frame, int exclude, double *out);

/* The exact FP64 NCC surfaces of windows k0 .. k0 + count - 1 against one frame (h x w uint8, device),
 * each search window at its window's position plus its gross offset (gross: nwin x 2 int32 (down,
 * across) on the device, indexed by window, or NULL): surfaces (count x M x M doubles, device)
 * [i][u][v] = the NCC of window k0 + i with the frame at shift (gross + u - search, gross + v - search),
 * exact integer sums normalized in FP64 as (W^2 sum FT - sum T sum F) / sqrt((W^2 sum T^2 - (sum T)^2)
 * (W^2 sum F^2 - (sum F)^2)), 0 where either variance is 0 (cxx/surface.h:30-35); NaN throughout where
 * the search window leaves the frame. The surface that cxx/cpu.cpp:138-154 evaluates in FP32 per window
 * before its argmax. Enqueued on stream, not synchronized; one call at a time per correlator (it owns
 * the frame's box maps). */
int framecor_ncc_surfaces_device(framecor_correlator *c, const uint8_t *frame, const int32_t *gross, int k0, int count,
                                 double *surfaces, void *stream);

/* The peak of every window's NCC with one frame (h x w uint8) as python/slipcam/stabilize.py:22-45
 * `ncc_peak`, for every window of the grid at once: on the exact FP64 surface (above; gross: nwin x 2
 * int32 or NULL), the first maximum (u, v) in row-major order, each coordinate refined by the parabola
 * through the maximum and its two neighbours (0 at the surface's border or where the parabola's
 * curvature is below 1e-12): offsets[2k], offsets[2k + 1] = u + du - search, v + dv - search (doubles,
 * relative to the window's position plus its gross offset); peak[k] = the maximum. NaN, NaN and 0 for a
 * flat template or a search window leaving the frame. Host arrays, synchronous. */
int framecor_ncc_peaks(framecor_correlator *c, const uint8_t *frame, const int32_t *gross, double *offsets,
                       double *peak);

/* The same with device arrays, enqueued on stream, not synchronized. */
int framecor_ncc_peaks_device(framecor_correlator *c, const uint8_t *frame, const int32_t *gross, double *offsets,
                              double *peak, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FRAMECOR_HIP_H */
