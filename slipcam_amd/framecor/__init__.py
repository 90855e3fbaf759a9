"""
framecor: dense normalized cross-correlation of 8-bit video frames with a reference frame, on an
MI355X (hand-written HIP kernels for gfx950, libframecor_hip.so)

The API of the reference's slipcam.framecor (python/slipcam/framecor/__init__.py, cxx/bindings.cpp):

  grid = framecor.window_grid(720, 1280, window=64, search=48, skip=16)
  cor = framecor.Correlator(720, 1280, *grid, window=64, search=48)
  cor.set_reference(frames[ref])
  result = cor.correlate(frames[first:last])     # offsets, peak, snr, cov

plus the keyword method='fft' (the reference's FP32 FFT algorithm, the default) or 'int8' (exact
integer correlation on the int8 matrix cores, FP64 peak fit), and device-resident inputs: correlate()
also takes a torch uint8 tensor on the GPU and then returns torch tensors without a host copy.

Two bindings of one library (libframecor_hip.so, its C ABI in include/framecor_hip.h): the pybind11
module _framecor, the reference's extension module (Correlator, has_cuda, __version__), which
Correlator here extends; and ctypes for the entry points beyond the reference's API (device-resident
calls, NCC surfaces, batches, the differentiable form).
"""

import ctypes
import os

import numpy as np

__all__ = ['Batch', 'Correlator', 'has_cuda', 'has_gpu', 'window_grid', '__version__']

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB_PATH = os.path.join(_PKG, 'libframecor_hip.so')

_OK, _EINVAL, _ERUNTIME, _ENOMEM = 0, 1, 2, 3
_METHODS = {'fft': 0, 'int8': 1}


def _load():
    # One HIP runtime per process: torch wheels bundle their own libamdhip64 (soname libamdhip64.so.7),
    # which this library's dependency resolves to when torch is loaded first; loaded the other way
    # round the process would map two runtimes, and torch tensors and streams handed to the library
    # would belong to another runtime. So torch, when installed, is loaded first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(_LIB_PATH):
        raise ImportError(f'framecor: {_LIB_PATH} is not built (python -m slipcam_amd.build)')
    lib = ctypes.CDLL(_LIB_PATH)
    P, I = ctypes.c_void_p, ctypes.c_int
    lib.framecor_version.restype = ctypes.c_char_p
    lib.framecor_has_gpu.restype = I
    lib.framecor_last_error.restype = ctypes.c_char_p
    lib.framecor_check_params.restype = I
    lib.framecor_check_params.argtypes = [I, I, P, I, P, I, I, I, I, I, I]
    lib.framecor_create.restype = I
    lib.framecor_create.argtypes = [I, I, P, I, P, I, I, I, I, I, I, I, ctypes.POINTER(P)]
    lib.framecor_destroy.restype = None
    lib.framecor_destroy.argtypes = [P]
    lib.framecor_backend.restype = ctypes.c_char_p
    lib.framecor_backend.argtypes = [P]
    lib.framecor_set_reference.restype = I
    lib.framecor_set_reference.argtypes = [P, P]
    lib.framecor_correlate.restype = I
    lib.framecor_correlate.argtypes = [P, P, I, P, P, P, P, P]
    lib.framecor_correlate_device.restype = I
    lib.framecor_correlate_device.argtypes = [P, P, I, P, P, P, P, P, P]
    lib.framecor_correlate_device_noalloc.restype = I
    lib.framecor_correlate_device_noalloc.argtypes = [P, P, I, P, P, P, P, P, P]
    lib.framecor_ncc_forward.restype = I
    lib.framecor_ncc_forward.argtypes = [P, P, I, P, P, P, P]
    lib.framecor_ncc_backward.restype = I
    lib.framecor_ncc_backward.argtypes = [P, P, I, P, P, P, P, P, P]
    lib.framecor_ncc_refine.restype = I
    lib.framecor_ncc_refine.argtypes = [P, P, I, P, P, I, P, P, P]
    lib.framecor_set_variant.restype = I
    lib.framecor_set_variant.argtypes = [P, I]
    lib.framecor_set_timing.restype = I
    lib.framecor_set_timing.argtypes = [P, I]
    lib.framecor_get_timing.restype = I
    lib.framecor_get_timing.argtypes = [P, P, P]
    lib.framecor_summed_area_device.restype = I
    lib.framecor_summed_area_device.argtypes = [P, I, I, I, P, P, P]
    lib.framecor_box_sums_device.restype = I
    lib.framecor_box_sums_device.argtypes = [P, I, I, I, I, P, P, P]
    lib.framecor_ncc_surfaces_device.restype = I
    lib.framecor_ncc_surfaces_device.argtypes = [P, P, P, I, I, P, P]
    lib.framecor_ncc_peaks.restype = I
    lib.framecor_ncc_peaks.argtypes = [P, P, P, P, P]
    lib.framecor_ncc_peaks_device.restype = I
    lib.framecor_ncc_peaks_device.argtypes = [P, P, P, P, P, P]
    lib.framecor_periodicity.restype = I
    lib.framecor_periodicity.argtypes = [P, P, I, P]
    lib.framecor_batch_create.restype = I
    lib.framecor_batch_create.argtypes = [P, I, I, I, ctypes.POINTER(P)]
    lib.framecor_batch_destroy.restype = None
    lib.framecor_batch_destroy.argtypes = [P]
    lib.framecor_batch_set_reference.restype = I
    lib.framecor_batch_set_reference.argtypes = [P, I, P]
    lib.framecor_batch_set_tuning.restype = I
    lib.framecor_batch_set_tuning.argtypes = [P, I]
    lib.framecor_batch_plan.restype = I
    lib.framecor_batch_plan.argtypes = [P, I, P]
    lib.framecor_batch_set_plan.restype = I
    lib.framecor_batch_set_plan.argtypes = [P, I, P]
    lib.framecor_batch_set_stage_frames.restype = I
    lib.framecor_batch_set_stage_frames.argtypes = [P, I]
    lib.framecor_batch_info.restype = I
    lib.framecor_batch_info.argtypes = [P, P]
    lib.framecor_batch_correlate_device.restype = I
    lib.framecor_batch_correlate_device.argtypes = [P, P, P, P, P, P, P, P]
    return lib


_lib = _load()
# the reference's module API (cxx/bindings.cpp:111-142): pybind11 on the same library
from ._framecor import __version__, has_cuda  # noqa: E402
from . import _framecor  # noqa: E402


def _raise(status):
    msg = _lib.framecor_last_error().decode()
    if status == _EINVAL:
        raise ValueError(msg)
    if status == _ENOMEM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def _check(status):
    if status != _OK:
        _raise(status)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def has_gpu():
    """Whether a GPU is visible to the HIP runtime (the reference's has_cuda)"""
    return bool(_lib.framecor_has_gpu())


def window_grid(height, width, window, search, skip, margin=0):
    """
    The first pixels (down, across) of the grid of windows whose search windows are in the frame:
    from margin + search, (size - 2 (margin + search) - window) // skip windows in each direction;
    a ValueError if no window fits (python/slipcam/framecor/__init__.py:18-32 of the reference)
    """
    axes = []
    for direction, size in (('down', height), ('across', width)):
        first = margin + search
        count = (size - margin - search - first - window) // skip
        if count <= 0:
            raise ValueError(f'window_grid: no window fits {direction}: frame size {size}, window {window}, '
                             f'search range {search}, margin {margin}')
        axes.append(first + skip * np.arange(count))
    return tuple(axes)


def _is_device_tensor(x):
    t = type(x)
    return t.__module__.startswith('torch') and t.__name__ == 'Tensor' and x.is_cuda


class Correlator(_framecor.Correlator):
    """
    The correlator of 8-bit frames with a reference frame, on a grid of square windows: the pybind11
    class of the reference's API (_framecor.Correlator: set_reference, correlate of numpy frames with
    the GIL released, backend, shape, window, search) plus the device-resident and extra entry points

    height, width: the size of the frames
    down, across: the first pixels of the rows and of the columns of windows
    window: the size of the windows; search: the half search range, pixels
    backend: 'hip', 'auto' or 'cuda' (the reference's name of its GPU backend): the GPU; the
    reference's 'cpu' is not a backend here (no CPU path: ValueError)
    method: 'fft' (the reference's FP32 FFT algorithm) or 'int8' (exact int8 matrix-core correlation)
    batch: at least 1 (kept for the reference's API); stat_size: the SNR window
    device: the GPU; threads: kept for the reference's API (the CPU threads of its CPU backend)
    """

    def __init__(self, height, width, down, across, *, window=64, search=48, backend='auto', batch=4096,
                 stat_size=21, device=0, threads=0, method='fft'):
        down = np.asarray(down)
        across = np.asarray(across)
        if down.ndim != 1 or across.ndim != 1:
            raise ValueError('framecor: down and across must be sequences of integers')
        super().__init__(int(height), int(width), [int(v) for v in down], [int(v) for v in across],
                         window=int(window), search=int(search), backend=backend, batch=int(batch),
                         stat_size=int(stat_size), device=int(device), threads=int(threads), method=method)
        # the C handle for the entry points beyond the reference's API (owned by the base class)
        self._c = ctypes.c_void_p(self._handle)
        self._h, self._w = int(height), int(width)
        self._device = int(device)

    def set_reference(self, frame):
        """Set the reference frame (height x width, uint8; numpy, or a torch tensor)"""
        if _is_device_tensor(frame):
            frame = frame.cpu().numpy()
        super().set_reference(frame)

    def set_variant(self, variant):
        """The int8 kernel: 'auto' (as 'solo' where it applies with windows up to 33, else as 'quad' where
        it applies, else 'simple'), 'simple' (one frame at a
        time, 16 x 16 tiles), 'piped' (producer/helper pipeline), 'padded' (one frame at a time, full
        16-row bands even where only an edge row and column remain), 'wide' (one frame at a time, 32 x 32
        tiles, search ranges 16, 32, 48, 64; else as 'simple'), 'quad' (one frame at a time in blocks of
        four waves with equal MFMA shares, search ranges 32 and 48 with windows of 32 or 64; else as
        'simple') or 'solo' (one wave per window and frame, windows up to 64 with search ranges up to 23
        or of 24 and 32; else as 'simple'); every variant gives the same offsets, peaks and covariances,
        the SNR to rounding"""
        _check(_lib.framecor_set_variant(self._c, _VARIANTS[variant]))

    def set_timing(self, enable=True):
        """Record the GPU time of each stage of the next correlate calls (hipEvents on their stream)"""
        _check(_lib.framecor_set_timing(self._c, int(bool(enable))))

    def timing(self):
        """The GPU milliseconds of the stages of the last correlate call: a dict with 'box_sums',
        'correlate' and 'launches' (waits for that call to finish)"""
        ms = (ctypes.c_float * 2)()
        launches = ctypes.c_int()
        _check(_lib.framecor_get_timing(self._c, ms, ctypes.byref(launches)))
        return dict(box_sums=ms[0], correlate=ms[1], launches=launches.value)

    def _gross(self, gross, n, module):
        ny, nx = self.shape
        g = gross
        if g.ndim == 2 and g.shape[0] == n and g.shape[1] == 2:
            return module.broadcast(g, n, ny, nx)
        if g.ndim == 4 and tuple(g.shape) == (n, ny, nx, 2):
            return g
        raise ValueError('framecor: gross must be frames x 2 or frames x down x across x 2')

    def correlate(self, frames, gross=None):
        """
        Correlate the frames (frames x height x width, uint8) with the reference frame

        gross: integer offsets (down, across) of the search windows, frames x 2 or
        frames x down x across x 2 (default: none)

        Returns a dict of float32 arrays: offsets (frames x down x across x 2: down, across, the motion
        of the frame; NaN if not valid), peak (the normalized correlation), snr, cov (the covariance of
        the offsets: down, across, cross). With a torch tensor on the GPU, torch tensors on that GPU,
        computed on the current stream.
        """
        if not self._ready:
            raise RuntimeError('framecor: set_reference first')
        if _is_device_tensor(frames):
            return self._correlate_device(frames, gross)
        return super().correlate(frames, gross)

    def _correlate_device(self, frames, gross):
        import torch
        if frames.dtype != torch.uint8 or frames.dim() != 3 or tuple(frames.shape[1:]) != (self._h, self._w):
            raise ValueError('framecor: the frames must be frames x height x width (uint8)')
        if frames.device.index != self._device:
            raise ValueError(f'framecor: the frames are on {frames.device}, the correlator on GPU {self._device}')
        frames = frames.contiguous()
        n = frames.shape[0]
        ny, nx = self.shape
        g = None
        if gross is not None:
            if not isinstance(gross, torch.Tensor):
                gross = torch.as_tensor(np.asarray(gross), device=frames.device)
            g = self._gross(gross.to(device=frames.device, dtype=torch.int32), n, _TorchOps).contiguous()
        kw = dict(dtype=torch.float32, device=frames.device)
        out = dict(offsets=torch.empty((n, ny, nx, 2), **kw), peak=torch.empty((n, ny, nx), **kw),
                   snr=torch.empty((n, ny, nx), **kw), cov=torch.empty((n, ny, nx, 3), **kw))
        if n > 0:
            stream = torch.cuda.current_stream(frames.device).cuda_stream
            _check(_lib.framecor_correlate_device(
                self._c, frames.data_ptr(), n, None if g is None else g.data_ptr(), out['offsets'].data_ptr(),
                out['peak'].data_ptr(), out['snr'].data_ptr(), out['cov'].data_ptr(), stream))
        return out


    def surfaces(self, frame):
        """
        The exact NCC surfaces of every window against a frame (height x width uint8 torch tensor on
        the GPU), each window at its own position: a float64 tensor down x across x M x M (M = 2 search
        + 1), [.., u, v] the NCC at shift (u - search, v - search); exact integer sums normalized in
        FP64, 0 where the search box or the template is flat. Computed on the current stream.
        """
        import torch
        if not self._ready:
            raise RuntimeError('framecor: set_reference first')
        if not (_is_device_tensor(frame) and frame.dtype == torch.uint8 and tuple(frame.shape) == (self._h, self._w)):
            raise ValueError('framecor: the frame must be a height x width uint8 tensor on the GPU')
        if frame.device.index != self._device:
            raise ValueError(f'framecor: the frame is on {frame.device}, the correlator on GPU {self._device}')
        frame = frame.contiguous()
        ny, nx = self.shape
        M = 2 * self.search + 1
        out = torch.empty((ny, nx, M, M), dtype=torch.float64, device=frame.device)
        stream = torch.cuda.current_stream(frame.device).cuda_stream
        _check(_lib.framecor_ncc_surfaces_device(self._c, frame.data_ptr(), None, 0, ny * nx, out.data_ptr(), stream))
        return out

    def ncc_peaks(self, frame, gross=None):
        """
        The peak of each window's exact NCC with its search window in frame (height x width uint8,
        numpy), as python/slipcam/stabilize.py:22-45 ncc_peak for every window at once: the first
        maximum, refined along each axis by the parabola through its neighbours. gross: down x across x 2
        integer offsets of the search windows (default none). Returns (offsets down x across x 2 float64:
        down, across relative to the window's position plus its gross offset; peak down x across
        float64); NaN offsets and peak 0 for a flat window or a search window leaving the frame.
        """
        if not self._ready:
            raise RuntimeError('framecor: set_reference first')
        frame = np.ascontiguousarray(frame, np.uint8)
        if frame.ndim != 2 or frame.shape != (self._h, self._w):
            raise ValueError('framecor: the frame must be height x width')
        ny, nx = self.shape
        g = None
        if gross is not None:
            g = np.asarray(gross)
            if g.shape != (ny, nx, 2):
                raise ValueError('framecor: gross must be down x across x 2')
            g = np.ascontiguousarray(g, np.int32)
        off = np.empty((ny, nx, 2), np.float64)
        peak = np.empty((ny, nx), np.float64)
        _check(_lib.framecor_ncc_peaks(self._c, _ptr(frame), _ptr(g), _ptr(off), _ptr(peak)))
        return off, peak

    def periodicity(self, frame, exclude=5):
        """The second peak of each window's NCC with its own surroundings in frame (height x width
        uint8; the reference frame for slipcam-auto's periodicity): down x across float64, 1 for a flat
        window, 0 where the surface has no local maximum outside the excluded square"""
        if not self._ready:
            raise RuntimeError('framecor: set_reference first')
        if _is_device_tensor(frame):
            frame = frame.cpu().numpy()
        frame = np.ascontiguousarray(frame, np.uint8)
        if frame.ndim != 2 or frame.shape != (self._h, self._w):
            raise ValueError('framecor: the frame must be height x width')
        out = np.empty(self.shape, np.float64)
        _check(_lib.framecor_periodicity(self._c, _ptr(frame), int(exclude), _ptr(out)))
        return out


class _Clip(ctypes.Structure):
    _fields_ = [('height', ctypes.c_int), ('width', ctypes.c_int), ('down', ctypes.c_void_p), ('ndown', ctypes.c_int),
                ('across', ctypes.c_void_p), ('nacross', ctypes.c_int), ('window', ctypes.c_int),
                ('search', ctypes.c_int), ('stat_size', ctypes.c_int), ('frames', ctypes.c_int)]


_VARIANTS = {'auto': 0, 'simple': 1, 'piped': 2, 'padded': 3, 'wide': 4, 'quad': 5, 'solo': 6}

# the tuner's set of candidate plans, in the plan cache's key: plans tuned among an earlier set (before
# the one-wave kernel and up to 16 blocks per CU) are tuned again rather than reused
_PLAN_CANDIDATES = 2


class Batch:
    """
    Many clips of uneven sizes correlated in one hipGraph launch (no counterpart in the reference: each
    clip is what one Correlator would be, with a fixed number of frames per call)

    clips: a sequence of dicts with height, width, down, across, window, search, frames (per call) and
    optionally stat_size (21). method: 'int8' or 'fft'. tune: time the candidate plans of every tile
    shape on the first call and keep the fastest (otherwise the automatic plan). plan_cache: a JSON file
    of tuned plans by (method, GPU, tile shape), read here and extended after a call that tuned (default:
    $FRAMECOR_PLAN_CACHE or ~/.cache/slipcam_amd/framecor_plans.json; None: no cache), so a later batch
    with the same tile shapes on the same kind of GPU starts with their plans and does not tune them.

      batch = framecor.Batch(clips, method='int8')
      for c, ref in enumerate(refs):
          batch.set_reference(c, ref)
      results = batch.correlate(frames)      # frames[c]: frames x height x width uint8 per clip
    """

    def __init__(self, clips, *, method='int8', device=0, tune=True, plan_cache=''):
        self._b = None
        m = _METHODS.get(method)
        if m is None:
            raise ValueError(f"framecor: the method must be fft or int8, not {method}")
        self._keep = []
        arr = (_Clip * len(clips))()
        self._clips = []
        for i, c in enumerate(clips):
            down = np.ascontiguousarray(c['down'], np.int32)
            across = np.ascontiguousarray(c['across'], np.int32)
            self._keep += [down, across]
            arr[i] = _Clip(int(c['height']), int(c['width']), down.ctypes.data, len(down), across.ctypes.data,
                           len(across), int(c['window']), int(c['search']), int(c.get('stat_size', 21)),
                           int(c['frames']))
            W, R = int(c['window']), int(c['search'])
            key = (f'W{W}R{R}S{min(int(c.get("stat_size", 21)), 2 * R + 1)}' if method == 'int8' else f'L{W + 2 * R}')
            self._clips.append(dict(height=int(c['height']), width=int(c['width']), shape=(len(down), len(across)),
                                    frames=int(c['frames']), key=key))
        handle = ctypes.c_void_p()
        _check(_lib.framecor_batch_create(ctypes.cast(arr, ctypes.c_void_p), len(clips), m, int(device),
                                          ctypes.byref(handle)))
        self._b = handle
        self._destroy = _lib.framecor_batch_destroy
        self._device = int(device)
        self._method = method
        _check(_lib.framecor_batch_set_tuning(self._b, int(bool(tune))))
        self._cache = None
        self.plans_from_cache = 0
        if tune and plan_cache is not None:
            self._cache = plan_cache or os.environ.get('FRAMECOR_PLAN_CACHE') or os.path.join(
                os.path.expanduser('~'), '.cache', 'slipcam_amd', 'framecor_plans.json')
            self._read_plans()

    def _device_key(self):
        import torch
        props = torch.cuda.get_device_properties(self._device)
        return (f'{self._method}|{getattr(props, "gcnArchName", props.name)}|{props.multi_processor_count}'
                f'|candidates {_PLAN_CANDIDATES}')

    def _read_plans(self):
        """The cached plans of this batch's tile shapes (an unreadable cache is ignored)"""
        import json
        try:
            with open(self._cache) as fh:
                plans = json.load(fh).get(self._device_key(), {})
        except (OSError, ValueError, AttributeError):
            return
        done = set()
        for c, info in enumerate(self._clips):
            p = plans.get(info['key'])
            if p is None or info['key'] in done:
                continue
            try:
                self.set_plan(c, p['variant'], int(p['frames_per_item']), int(p['blocks_per_cu']))
            except (KeyError, TypeError, ValueError):
                continue
            done.add(info['key'])
        self.plans_from_cache = len(done)

    def _write_plans(self):
        """Add this batch's tuned plans to the cache (written whole to a temporary file, then renamed)"""
        import json
        try:
            with open(self._cache) as fh:
                cache = json.load(fh)
        except (OSError, ValueError):
            cache = {}
        mine = cache.setdefault(self._device_key(), {})
        changed = False
        for c, info in enumerate(self._clips):
            p = self.plan(c)
            if p['tuned'] and mine.get(info['key']) != {k: p[k] for k in ('variant', 'frames_per_item', 'blocks_per_cu')}:
                mine[info['key']] = {k: p[k] for k in ('variant', 'frames_per_item', 'blocks_per_cu')}
                changed = True
        if not changed:
            return
        try:
            os.makedirs(os.path.dirname(os.path.abspath(self._cache)), exist_ok=True)
            tmp = f'{self._cache}.{os.getpid()}.tmp'
            with open(tmp, 'w') as fh:
                json.dump(cache, fh, indent=1, sort_keys=True)
            os.replace(tmp, self._cache)
        except OSError:
            pass

    def __del__(self):
        b, destroy = getattr(self, '_b', None), getattr(self, '_destroy', None)
        if b and destroy is not None:
            destroy(b)
            self._b = None

    def __len__(self):
        return len(self._clips)

    @property
    def method(self):
        return self._method

    def set_reference(self, clip, frame):
        """Set the reference frame of one clip (its height x width, uint8)"""
        c = self._clips[clip]
        if _is_device_tensor(frame):
            frame = frame.cpu().numpy()
        frame = np.ascontiguousarray(frame, np.uint8)
        if frame.shape != (c['height'], c['width']):
            raise ValueError(f'framecor: the reference frame of clip {clip} must be height x width')
        _check(_lib.framecor_batch_set_reference(self._b, int(clip), _ptr(frame)))

    def plan(self, clip):
        """The plan of the clip's tile shape: kernel variant, frames per work item (0: automatic),
        resident blocks per CU (0: as many as fit), and whether it was tuned or set"""
        p = (ctypes.c_int * 4)()
        _check(_lib.framecor_batch_plan(self._b, int(clip), p))
        names = {v: k for k, v in _VARIANTS.items()}
        return dict(variant=names[p[0]], frames_per_item=p[1], blocks_per_cu=p[2], tuned=bool(p[3]))

    def set_plan(self, clip, variant='auto', frames_per_item=0, blocks_per_cu=0):
        """Set the plan of the clip's tile shape (every clip of that shape)"""
        p = (ctypes.c_int * 3)(_VARIANTS[variant], int(frames_per_item), int(blocks_per_cu))
        _check(_lib.framecor_batch_set_plan(self._b, int(clip), p))

    def set_stage_frames(self, frames):
        """At most this many frames of each clip per stage (default: every frame in one stage where the
        box-sum maps of all clips fit in a quarter of the free device memory, at most 64 GiB; else what
        2 GiB of a clip's maps hold)"""
        _check(_lib.framecor_batch_set_stage_frames(self._b, int(frames)))

    def info(self):
        """tile shapes, stages of frames, nodes of the captured graph, work queues per call"""
        p = (ctypes.c_int * 4)()
        _check(_lib.framecor_batch_info(self._b, p))
        return dict(shapes=p[0], stages=p[1], graph_nodes=p[2], queues=p[3])

    def _staging(self, key, c, shape, dtype):
        """A device buffer of the batch kept across calls (stable pointers: the captured graph is reused)"""
        import torch
        buf = self._bufs.get((key, c))
        if buf is None or tuple(buf.shape) != tuple(shape) or buf.dtype != dtype:
            buf = torch.empty(shape, dtype=dtype, device=torch.device('cuda', self._device))
            self._bufs[(key, c)] = buf
        return buf

    def correlate(self, frames, gross=None):
        """
        Correlate every clip: frames[c] is clip c's frames x height x width uint8 (torch tensors on the
        GPU, or numpy arrays), gross[c] None, frames x 2 or frames x down x across x 2. Returns one dict
        per clip as Correlator.correlate, computed by one graph launch on the current stream: numpy arrays
        for host inputs; for device inputs, torch tensors on the GPU that belong to the batch and are
        overwritten by its next call (clone them to keep them). The graph is captured on the first call
        and again whenever the frames or gross tensors given are other tensors than last time.
        """
        import torch
        if not hasattr(self, '_bufs'):
            self._bufs = {}
        if len(frames) != len(self._clips):
            raise ValueError(f'framecor: {len(self._clips)} clips, {len(frames)} frame arrays')
        host = not all(_is_device_tensor(f) for f in frames)
        dev = torch.device('cuda', self._device)
        outs, fr, gr = [], [], []
        for c, (info, f) in enumerate(zip(self._clips, frames)):
            n, h, w = info['frames'], info['height'], info['width']
            ny, nx = info['shape']
            if not _is_device_tensor(f):
                f = np.ascontiguousarray(f, np.uint8)
                if f.shape != (n, h, w):
                    raise ValueError(f'framecor: the frames of clip {c} must be {n} x {h} x {w} uint8')
                f = self._staging('frames', c, (n, h, w), torch.uint8).copy_(torch.from_numpy(f))
            if f.dtype != torch.uint8 or tuple(f.shape) != (n, h, w) or f.device != dev:
                raise ValueError(f'framecor: the frames of clip {c} must be {n} x {h} x {w} uint8 on {dev}')
            if not f.is_contiguous():
                f = self._staging('frames', c, (n, h, w), torch.uint8).copy_(f)
            g = None if gross is None else gross[c]
            if g is not None:
                g = torch.as_tensor(np.asarray(g) if not isinstance(g, torch.Tensor) else g)
                if tuple(g.shape) == (n, 2):
                    g = g[:, None, None, :].expand(n, ny, nx, 2)
                elif tuple(g.shape) != (n, ny, nx, 2):
                    raise ValueError(f'framecor: gross of clip {c} must be frames x 2 or frames x down x across x 2')
                g = self._staging('gross', c, (n, ny, nx, 2), torch.int32).copy_(g)
            outs.append({k: self._staging(k, c, s, torch.float32)
                         for k, s in (('offsets', (n, ny, nx, 2)), ('peak', (n, ny, nx)), ('snr', (n, ny, nx)),
                                      ('cov', (n, ny, nx, 3)))})
            fr.append(f.data_ptr())
            gr.append(0 if g is None else g.data_ptr())
        P = ctypes.c_void_p * len(self._clips)
        arrays = [P(*fr), P(*gr) if gross is not None else None] + [P(*[o[k].data_ptr() for o in outs])
                                                                      for k in ('offsets', 'peak', 'snr', 'cov')]
        stream = torch.cuda.current_stream(dev).cuda_stream
        untuned = self._cache is not None and not all(self.plan(c)['tuned'] for c in range(len(self._clips)))
        _check(_lib.framecor_batch_correlate_device(self._b, *[None if a is None else ctypes.cast(a, ctypes.c_void_p)
                                                               for a in arrays], stream))
        if untuned:   # this call tuned the shapes without a plan
            self._write_plans()
        self._inflight = frames   # the caller's tensors stay alive while the launch may read them
        if host:
            torch.cuda.current_stream(dev).synchronize()
            return [{k: v.cpu().numpy() for k, v in o.items()} for o in outs]
        return outs


class _TorchOps:
    @staticmethod
    def broadcast(g, n, ny, nx):
        return g[:, None, None, :].expand(n, ny, nx, 2)
