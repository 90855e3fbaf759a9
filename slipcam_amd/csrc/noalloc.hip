// noalloc.hip -- framecor_correlate_device_noalloc: the device-resident correlation with no device
// allocation, for callers that cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc)
//
// The one buffer a correlate call needs beyond the caller's inputs and outputs and the correlator's own
// tables is the box-sum maps of a group of frames (int8: 12 bytes, FFT: 4 bytes per window position;
// the FFT form's workspace for spectra beyond LDS is not taken here: such correlators are refused).
// Here they live in `scratch`, a file-scope device array of kScratchBytes (64 MiB) that every GPU the
// library runs on holds until the process ends. Its device address is resolved once per device when a
// correlator is created on it (framecor_create calls scratchSetup with that device current) and kept
// per device; host code uses that address, never the array's name. Frames are correlated in groups whose
// maps fit in the array: a frame whose maps alone exceed it is refused (FRAMECOR_EINVAL).
//
// Every call on a GPU shares that GPU's array, which keeps the last call's contents (nothing in it has
// to start at zero: each group's maps are written before they are read). Calls are enqueued under a lock,
// and each makes its stream wait for the previous call's work on the array (an event recorded behind it),
// so calls from any host threads on any streams are ordered on the GPU. Under stream capture no event
// is waited for or recorded: work captured into graphs shares the array too, so replays (or captured
// and direct calls) that overlap on one GPU race.

#include <mutex>
#include <string>
#include <unordered_map>

#include "correlator.h"

using namespace framecor;

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct ScratchSlot {
    void *ptr = nullptr;        // the device address of this GPU's copy of scratch
    hipEvent_t last = nullptr;  // behind the last call's work on it
    bool used = false;
    std::string error;          // why the set-up failed (the entry point reports it)
};

std::mutex scratchLock;
std::unordered_map<int, ScratchSlot> slots;   // by device

} // namespace

namespace framecor {

// once per device, with that device current: the address of its copy of scratch and the event that
// orders the calls; a failure is kept for the entry point to report (the other entry points do not
// need the array)
void scratchSetup(int device)
{
    std::lock_guard<std::mutex> guard(scratchLock);
    ScratchSlot &slot = slots[device];
    if (slot.ptr) return;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && !p) e = hipErrorInvalidSymbol;
    if (e == hipSuccess && !slot.last) e = hipEventCreateWithFlags(&slot.last, hipEventDisableTiming);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        slot.error = std::string("framecor: the no-allocation scratch of GPU ") + std::to_string(device)
            + " could not be set up: " + hipGetErrorString(e);
        return;
    }
    slot.ptr = p;
    slot.error.clear();
}

} // namespace framecor

extern "C" int framecor_correlate_device_noalloc(framecor_correlator *c, const uint8_t *frames, int n,
                                                 const int32_t *gross, float *offset, float *peak, float *snr,
                                                 float *cov, void *stream)
{
    if (!c) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    if (n > 0 && (!frames || !offset || !peak || !snr || !cov)) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    return guarded([&] {
        std::lock_guard<std::mutex> guard(scratchLock);
        auto it = slots.find(c->device);
        if (it == slots.end()) throw std::runtime_error("framecor: the no-allocation scratch is not set up");
        ScratchSlot &slot = it->second;
        if (!slot.ptr) throw std::runtime_error(slot.error);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        FC_HIP_CHECK(hipSetDevice(c->device));
        const bool captured = framecor_correlator::capturing(s);
        if (slot.used && !captured) FC_HIP_CHECK(hipStreamWaitEvent(s, slot.last, 0));
        c->correlateScratch(frames, n, gross, offset, peak, snr, cov, s, slot.ptr, kScratchBytes);
        if (n > 0 && !captured) {
            FC_HIP_CHECK(hipEventRecord(slot.last, s));
            slot.used = true;
        }
    });
}
