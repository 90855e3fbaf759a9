"""
The CPU oracle of the framecor correlator (test infrastructure).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py import this package, as
the checker of the HIP path or as the timed CPU baseline. The product (slipcam_amd) never does.

Two correlators, both in oracle/framecor_oracle.c (built by oracle/Makefile):

* ``OracleCorrelator.correlate_fft`` -- a C restatement of the reference CPU backend
  (cxx/cpu.cpp:24-167, cxx/surface.h:21-146, cxx/common.cpp:30-56) with this repository's own FP32
  mixed-radix FFT in place of FFTW;
* ``OracleCorrelator.correlate_direct`` -- the FP64 direct-NCC oracle: exact integer correlation
  sums, the NCC in FP64, first-maximum argmax and the peak fit, SNR and covariance in FP64.

Each function's pin is in tests/test_oracle.py.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, 'liboracle_framecor.so')
_lib = None


def build():
    """Compile the oracle library (gcc, OpenMP)"""
    subprocess.run(['make', '-s', '-C', _HERE, 'liboracle_framecor.so'], check=True)


def lib():
    """The loaded oracle library, built first if missing"""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    P, I, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    L.oracle_create.restype = P
    L.oracle_create.argtypes = [I, I, P, I, P, I, I, I, I, I]
    L.oracle_destroy.argtypes = [P]
    L.oracle_set_reference.restype = I
    L.oracle_set_reference.argtypes = [P, P]
    L.oracle_correlate_fft.restype = I
    L.oracle_correlate_fft.argtypes = [P, P, I, P, P, P, P, P, P]
    L.oracle_correlate_direct.restype = I
    L.oracle_correlate_direct.argtypes = [P, P, I, P, P, P, P, P, P, P]
    L.oracle_ncc_peaks.restype = I
    L.oracle_ncc_peaks.argtypes = [P, P, P, P, P]
    L.oracle_periodicity.restype = I
    L.oracle_periodicity.argtypes = [P, P, I, P]
    L.oracle_reference_stats.restype = I
    L.oracle_reference_stats.argtypes = [P, P, P, P]
    L.oracle_reference_spectra.restype = P
    L.oracle_reference_spectra.argtypes = [P]
    L.oracle_summed_area.argtypes = [P, I, I, P, P]
    L.oracle_raw_direct.argtypes = [P, I, I, P, I, I, P]
    L.oracle_write_peak_f32.argtypes = [P, I, I, I, I, I, I, I, I, I, I64, P, P, P, P]
    L.oracle_write_peak_f64.argtypes = [P, I, I, I, I, I, I, I, I, I, I64, P, P, P, P]
    L.oracle_cfft.restype = I
    L.oracle_cfft.argtypes = [I, P, I]
    L.oracle_rfft2.restype = I
    L.oracle_rfft2.argtypes = [I, P, P]
    L.oracle_irfft2.restype = I
    L.oracle_irfft2.argtypes = [I, P, P]
    L.oracle_threads.restype = I
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def threads():
    """The OpenMP threads the oracle uses by default"""
    return lib().oracle_threads()


def summed_area(frame):
    """The (h+1) x (w+1) int64 summed-area tables of a uint8 frame and of its squares"""
    frame = np.ascontiguousarray(frame, np.uint8)
    h, w = frame.shape
    sat = np.empty((h + 1, w + 1), np.int64)
    sat2 = np.empty((h + 1, w + 1), np.int64)
    lib().oracle_summed_area(_ptr(frame), h, w, _ptr(sat), _ptr(sat2))
    return sat, sat2


def raw_direct(chip, template):
    """The exact correlation sums sum_ij t[i,j] chip[u+i, v+j], (L-W+1) x (L-W+1) int64"""
    chip = np.ascontiguousarray(chip, np.uint8)
    t = np.ascontiguousarray(template, np.uint8)
    L, W = chip.shape[0], t.shape[0]
    assert chip.shape == (L, L) and t.shape == (W, W) and W <= L
    out = np.empty((L - W + 1, L - W + 1), np.int64)
    lib().oracle_raw_direct(_ptr(chip), L, L, _ptr(t), W, W, _ptr(out))
    return out


def cfft(x, sign=-1):
    """The oracle's complex FFT (FP32), unnormalized in both directions"""
    x = np.asarray(x, np.complex64)
    buf = np.ascontiguousarray(x).view(np.float32).copy()
    if lib().oracle_cfft(len(x), _ptr(buf), sign):
        raise ValueError('cfft: size not supported')
    return buf.view(np.complex64)


def rfft2(x):
    """The oracle's 2-D real FFT of an L x L float32 array: L x (L/2+1) complex64"""
    x = np.ascontiguousarray(x, np.float32)
    L = x.shape[0]
    spec = np.empty((L, L // 2 + 1), np.complex64)
    lib().oracle_rfft2(L, _ptr(x), _ptr(spec))
    return spec


def irfft2(spec):
    """The oracle's unnormalized inverse of rfft2 (L^2 times the inverse transform)"""
    spec = np.ascontiguousarray(spec, np.complex64).copy()
    L = spec.shape[0]
    out = np.empty((L, L), np.float32)
    lib().oracle_irfft2(L, _ptr(spec), _ptr(out))
    return out


def write_peak(surface, u, v, search, gross=(0, 0), stat_size=21, template_size=64 * 64, fp64=False):
    """The outputs (offset down, offset across, peak, snr, cov x 3) of a surface and its integer peak:
    the reference's FP32 fit (fp64=False) or the FP64 fit"""
    s = np.ascontiguousarray(surface, np.float64 if fp64 else np.float32)
    M = s.shape[0]
    off = np.empty(2, np.float32)
    peak = np.empty(1, np.float32)
    snr = np.empty(1, np.float32)
    cov = np.empty(3, np.float32)
    fn = lib().oracle_write_peak_f64 if fp64 else lib().oracle_write_peak_f32
    fn(_ptr(s), s.shape[1], M, u, v, search, int(gross[0]), int(gross[1]), stat_size, template_size, 0,
       _ptr(off), _ptr(peak), _ptr(snr), _ptr(cov))
    return off, peak[0], snr[0], cov


class OracleCorrelator:
    """The oracle correlators over the same window grid and outputs as framecor.Correlator"""

    def __init__(self, height, width, down, across, *, window=64, search=48, stat_size=21, threads=0):
        self.h, self.w = int(height), int(width)
        self.down = np.ascontiguousarray(down, np.int32)
        self.across = np.ascontiguousarray(across, np.int32)
        self.window, self.search, self.stat_size = int(window), int(search), int(stat_size)
        self.shape = (len(self.down), len(self.across))
        self.nwin = self.shape[0] * self.shape[1]
        self._c = lib().oracle_create(self.h, self.w, _ptr(self.down), len(self.down), _ptr(self.across),
                                      len(self.across), self.window, self.search, self.stat_size, int(threads))
        if not self._c:
            raise MemoryError('oracle_create')

    def __del__(self):
        c = getattr(self, '_c', None)
        if c:
            lib().oracle_destroy(c)
            self._c = None

    def set_reference(self, frame):
        frame = np.ascontiguousarray(frame, np.uint8)
        assert frame.shape == (self.h, self.w)
        if lib().oracle_set_reference(self._c, _ptr(frame)):
            raise MemoryError('oracle_set_reference')

    def reference_stats(self):
        """(norm float32, sum(T) int64, sum(T^2) int64) per window"""
        norm = np.empty(self.nwin, np.float32)
        s = np.empty(self.nwin, np.int64)
        s2 = np.empty(self.nwin, np.int64)
        if lib().oracle_reference_stats(self._c, _ptr(norm), _ptr(s), _ptr(s2)):
            raise RuntimeError('set_reference first')
        return norm, s, s2

    def reference_spectra(self):
        L = self.window + 2 * self.search
        p = lib().oracle_reference_spectra(self._c)
        n = self.nwin * L * (L // 2 + 1)
        arr = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_float)), shape=(2 * n,))
        return arr.view(np.complex64).reshape(self.nwin, L, L // 2 + 1).copy()

    def _prep(self, frames, gross, windows):
        frames = np.ascontiguousarray(frames, np.uint8)
        assert frames.ndim == 3 and frames.shape[1:] == (self.h, self.w)
        n = frames.shape[0]
        ny, nx = self.shape
        g = None
        if gross is not None:
            a = np.asarray(gross)
            if a.shape == (n, 2):
                g = np.ascontiguousarray(np.broadcast_to(a.astype(np.int32)[:, None, None, :], (n, ny, nx, 2)))
            else:
                assert a.shape == (n, ny, nx, 2)
                g = np.ascontiguousarray(a, np.int32)
        mask = None
        if windows is not None:
            mask = np.zeros(self.nwin, np.uint8)
            mask[np.asarray(windows)] = 1
        out = dict(offsets=np.full((n, ny, nx, 2), np.nan, np.float32), peak=np.zeros((n, ny, nx), np.float32),
                   snr=np.zeros((n, ny, nx), np.float32), cov=np.zeros((n, ny, nx, 3), np.float32))
        return frames, n, g, mask, out

    def correlate_fft(self, frames, gross=None, windows=None):
        """The reference algorithm (FP32 FFT); windows: the flat indices to correlate (default all)"""
        frames, n, g, mask, out = self._prep(frames, gross, windows)
        if n:
            rc = lib().oracle_correlate_fft(self._c, _ptr(frames), n, _ptr(g), _ptr(mask), _ptr(out['offsets']),
                                            _ptr(out['peak']), _ptr(out['snr']), _ptr(out['cov']))
            if rc:
                raise RuntimeError(f'oracle_correlate_fft: {rc}')
        return out

    def ncc_peaks(self, frame, gross=None):
        """The ncc_peak of slipcam's stabilize.py for every window against frame: (offsets down x across
        x 2, peak down x across), float64; gross: down x across x 2 integer offsets or None"""
        frame = np.ascontiguousarray(frame, np.uint8)
        assert frame.shape == (self.h, self.w)
        g = None if gross is None else np.ascontiguousarray(gross, np.int32)
        assert g is None or g.shape == self.shape + (2,)
        off = np.empty(self.shape + (2,), np.float64)
        peak = np.empty(self.shape, np.float64)
        rc = lib().oracle_ncc_peaks(self._c, _ptr(frame), _ptr(g), _ptr(off), _ptr(peak))
        if rc:
            raise RuntimeError(f'oracle_ncc_peaks: {rc}')
        return off, peak

    def periodicity(self, frame, exclude=5):
        """The second peak of each window's exact NCC with its own surroundings in frame (down x across
        float64), the periodicity of slipcam-auto"""
        frame = np.ascontiguousarray(frame, np.uint8)
        assert frame.shape == (self.h, self.w)
        out = np.empty(self.shape, np.float64)
        rc = lib().oracle_periodicity(self._c, _ptr(frame), int(exclude), _ptr(out))
        if rc:
            raise RuntimeError(f'oracle_periodicity: {rc}')
        return out

    def correlate_direct(self, frames, gross=None, windows=None, surfaces=False):
        """The FP64 direct-NCC oracle; surfaces=True adds 'surface' (n x ny x nx x M x M float64)"""
        frames, n, g, mask, out = self._prep(frames, gross, windows)
        M = 2 * self.search + 1
        surf = np.full((n,) + self.shape + (M, M), np.nan) if surfaces else None
        if n:
            rc = lib().oracle_correlate_direct(self._c, _ptr(frames), n, _ptr(g), _ptr(mask), _ptr(out['offsets']),
                                               _ptr(out['peak']), _ptr(out['snr']), _ptr(out['cov']), _ptr(surf))
            if rc:
                raise RuntimeError(f'oracle_correlate_direct: {rc}')
        if surfaces:
            out['surface'] = surf
        return out
