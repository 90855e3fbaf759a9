// framecor_module.cpp -- the pybind11 module slipcam_amd.framecor._framecor: the reference's Python
// extension framecor._framecor (cxx/bindings.cpp:111-142), with the same names, keywords, defaults,
// argument checks and exceptions, built on the C ABI of libframecor_hip.so (include/framecor_hip.h).
// Host C++ only (g++): no HIP or torch types here; every GPU call goes through the C entry points,
// with the GIL released while they run.
//
//   from slipcam_amd.framecor._framecor import Correlator, has_cuda, __version__
//   cor = Correlator(720, 1280, down, across, window=64, search=48)      # backend='auto'
//   cor.set_reference(ref)
//   r = cor.correlate(frames, gross=None)    # dict: offsets, peak, snr, cov (float32 numpy arrays)
//
// The keyword method='fft' | 'int8' (the reference has only the FFT) comes after the reference's.

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "framecor_hip.h"

namespace py = pybind11;

namespace {

using U8 = py::array_t<uint8_t, py::array::c_style | py::array::forcecast>;
using I32 = py::array_t<int32_t, py::array::c_style | py::array::forcecast>;

// a C status as the reference module's exception: ValueError, MemoryError or RuntimeError
void raise(int status)
{
    if (status == FRAMECOR_OK) return;
    const std::string msg = framecor_last_error();
    if (status == FRAMECOR_EINVAL) throw py::value_error(msg);
    if (status == FRAMECOR_ENOMEM) throw std::bad_alloc();
    throw std::runtime_error(msg);
}

class Correlator {
public:
    Correlator(int height, int width, const std::vector<int> &down, const std::vector<int> &across, int window,
               int search, const std::string &backend, int batch, int statSize, int device, int threads,
               const std::string &method)
        : h_(height), w_(width), ny_(int(down.size())), nx_(int(across.size())), window_(window), search_(search)
    {
        (void)threads;   // the CPU threads of the reference's CPU backend; there is no CPU path here
        const int m = method == "int8" ? FRAMECOR_METHOD_INT8 : FRAMECOR_METHOD_FFT;
        // the parameters first, then the backend and the method (cxx/bindings.cpp:37-44)
        raise(framecor_check_params(height, width, down.data(), ny_, across.data(), nx_, window, search, m, batch,
                                    statSize));
        if (backend != "auto" && backend != "hip" && backend != "cuda")
            throw py::value_error("framecor: the backend must be auto, hip or cuda, not " + backend);
        if (method != "fft" && method != "int8")
            throw py::value_error("framecor: the method must be fft or int8, not " + method);
        method_ = method;
        int status;
        {
            py::gil_scoped_release release;
            status = framecor_create(height, width, down.data(), ny_, across.data(), nx_, window, search, m, batch,
                                     statSize, device, &c_);
        }
        raise(status);
        backend_ = backend == "cuda" ? "cuda" : framecor_backend(c_);
    }

    ~Correlator() { framecor_destroy(c_); }
    Correlator(const Correlator &) = delete;
    Correlator &operator=(const Correlator &) = delete;

    void setReference(const U8 &frame)
    {
        if (frame.ndim() != 2 || frame.shape(0) != h_ || frame.shape(1) != w_)
            throw py::value_error("framecor: the reference frame must be height x width");
        const uint8_t *p = frame.data();
        int status;
        {
            py::gil_scoped_release release;
            status = framecor_set_reference(c_, p);
        }
        raise(status);
        ready_ = true;
    }

    py::dict correlate(const U8 &frames, std::optional<I32> gross)
    {
        if (!ready_) throw std::runtime_error("framecor: set_reference first");
        if (frames.ndim() != 3 || frames.shape(1) != h_ || frames.shape(2) != w_)
            throw py::value_error("framecor: the frames must be frames x height x width");
        const py::ssize_t n = frames.shape(0), nwin = py::ssize_t(ny_) * nx_;
        // per-window offsets: n x 2 broadcast to every window, or n x ny x nx x 2 as given
        std::vector<int32_t> spread;
        const int32_t *g = nullptr;
        if (gross) {
            const I32 &a = *gross;
            if (a.ndim() == 2 && a.shape(0) == n && a.shape(1) == 2) {
                spread.resize(size_t(n) * nwin * 2);
                const int32_t *src = a.data();
                for (py::ssize_t f = 0; f < n; f++)
                    for (py::ssize_t k = 0; k < nwin; k++) {
                        spread[2 * (f * nwin + k)] = src[2 * f];
                        spread[2 * (f * nwin + k) + 1] = src[2 * f + 1];
                    }
                g = spread.data();
            }
            else if (a.ndim() == 4 && a.shape(0) == n && a.shape(1) == ny_ && a.shape(2) == nx_ && a.shape(3) == 2)
                g = a.data();
            else
                throw py::value_error("framecor: gross must be frames x 2 or frames x down x across x 2");
        }
        const py::ssize_t ny = ny_, nx = nx_;
        py::array_t<float> offsets({n, ny, nx, py::ssize_t(2)}), peak({n, ny, nx}), snr({n, ny, nx}),
            cov({n, ny, nx, py::ssize_t(3)});
        if (n > 0) {
            float *o = offsets.mutable_data(), *p = peak.mutable_data(), *s = snr.mutable_data(),
                  *cv = cov.mutable_data();
            const uint8_t *fr = frames.data();
            int status;
            {
                py::gil_scoped_release release;
                status = framecor_correlate(c_, fr, int(n), g, o, p, s, cv);
            }
            raise(status);
        }
        py::dict out;
        out["offsets"] = offsets;
        out["peak"] = peak;
        out["snr"] = snr;
        out["cov"] = cov;
        return out;
    }

    const std::string &backend() const { return backend_; }
    py::tuple shape() const { return py::make_tuple(ny_, nx_); }
    int window() const { return window_; }
    int search() const { return search_; }
    const std::string &method() const { return method_; }
    bool ready() const { return ready_; }
    // the C handle (framecor_correlator *) as an integer, for the package's other C entry points
    uintptr_t handle() const { return reinterpret_cast<uintptr_t>(c_); }

private:
    framecor_correlator *c_ = nullptr;
    int h_, w_, ny_, nx_, window_, search_;
    std::string backend_, method_;
    bool ready_ = false;
};

} // namespace

PYBIND11_MODULE(_framecor, m)
{
    m.doc() = "Dense normalized cross-correlation of 8-bit frames with a reference frame, on an MI355X";
    m.attr("__version__") = framecor_version();
    m.def("has_cuda", [] { return framecor_has_gpu() != 0; },
          "Whether the GPU backend can run (the reference's name; a HIP device is visible)");
    m.def("has_gpu", [] { return framecor_has_gpu() != 0; }, "Whether a HIP device is visible");

    py::class_<Correlator>(m, "Correlator",
        "The correlator of 8-bit frames with a reference frame, on a grid of square windows\n\n"
        "height, width: the size of the frames\n"
        "down, across: the first pixels of the rows and of the columns of windows\n"
        "window: the size of the windows; search: the half search range, pixels\n"
        "backend: 'auto', 'hip' or 'cuda' (the reference's name): the GPU (there is no CPU backend)\n"
        "batch: kept for the reference's API; stat_size: the SNR window\n"
        "device: the GPU; threads: kept for the reference's API\n"
        "method: 'fft' (the reference's FP32 FFT algorithm) or 'int8' (exact int8 matrix-core correlation)")
        .def(py::init<int, int, const std::vector<int> &, const std::vector<int> &, int, int, const std::string &, int,
                      int, int, int, const std::string &>(),
             py::arg("height"), py::arg("width"), py::arg("down"), py::arg("across"), py::kw_only(),
             py::arg("window") = 64, py::arg("search") = 48, py::arg("backend") = "auto", py::arg("batch") = 4096,
             py::arg("stat_size") = 21, py::arg("device") = 0, py::arg("threads") = 0, py::arg("method") = "fft")
        .def("set_reference", &Correlator::setReference, py::arg("frame"),
             "Set the reference frame (height x width, uint8)")
        .def("correlate", &Correlator::correlate, py::arg("frames"), py::arg("gross") = py::none(),
             "Correlate the frames (frames x height x width, uint8) with the reference frame\n\n"
             "gross: integer offsets (down, across) of the search windows, frames x 2 or\n"
             "frames x down x across x 2 (default: none)\n\n"
             "Returns a dict of float32 arrays: offsets (frames x down x across x 2: down, across,\n"
             "the motion of the frame; NaN if not valid), peak (the normalized correlation),\n"
             "snr, cov (the covariance of the offsets: down, across, cross)")
        .def_property_readonly("backend", &Correlator::backend)
        .def_property_readonly("shape", &Correlator::shape, "the number of windows (down, across)")
        .def_property_readonly("window", &Correlator::window)
        .def_property_readonly("search", &Correlator::search)
        .def_property_readonly("method", &Correlator::method)
        .def_property_readonly("_ready", &Correlator::ready)
        .def_property_readonly("_handle", &Correlator::handle);
}
