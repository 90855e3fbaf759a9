"""
framecor_correlate_device_noalloc (the entry point that allocates nothing: the box-sum maps in a static
64 MiB device array per GPU) through the C ABI: the same bits as framecor_correlate_device, on two calls
in a row, the second with fewer frames, for both methods, with groups of frames split by the array's
size; and a frame too large for the array refused with a clear error.
"""

import ctypes

import numpy as np
import pytest

from slipcam_amd import framecor, synthetic

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

KEYS = ('offsets', 'peak', 'snr', 'cov')


def call(entry, cor, frames, gross):
    n = frames.shape[0]
    ny, nx = cor.shape
    kw = dict(dtype=torch.float32, device=frames.device)
    out = dict(offsets=torch.full((n, ny, nx, 2), -7.0, **kw), peak=torch.full((n, ny, nx), -7.0, **kw),
               snr=torch.full((n, ny, nx), -7.0, **kw), cov=torch.full((n, ny, nx, 3), -7.0, **kw))
    stream = torch.cuda.current_stream(frames.device).cuda_stream
    status = entry(cor._c, frames.data_ptr(), n, None if gross is None else gross.data_ptr(),
                   out['offsets'].data_ptr(), out['peak'].data_ptr(), out['snr'].data_ptr(), out['cov'].data_ptr(),
                   ctypes.c_void_p(stream))
    assert status == 0, framecor._lib.framecor_last_error().decode()
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


@pytest.fixture(scope='module')
def video():
    d = synthetic.fault_video(n=24)
    return d, torch.from_numpy(d['frames']).cuda()


@pytest.mark.parametrize('method', ['int8', 'fft'])
def test_noalloc_equals_the_device_entry_point(gpu, video, method):
    """720 x 1280, window 64, search 48: 24 then 5 frames (int8: groups of 6 frames in the 64 MiB array;
    fft: groups of 21), with per-window gross offsets on the first call"""
    d, dev = video
    down, across = framecor.window_grid(720, 1280, 64, 48, 16)
    cor = framecor.Correlator(720, 1280, down, across, window=64, search=48, method=method)
    cor.set_reference(d['ref'])
    ny, nx = cor.shape
    gross = torch.tensor(synthetic.gross_offsets(24, ny, nx, -8, 8), device='cuda')
    lib = framecor._lib
    for frames, g in ((dev, gross), (dev[3:8].contiguous(), None)):
        want = call(lib.framecor_correlate_device, cor, frames, g)
        got = call(lib.framecor_correlate_device_noalloc, cor, frames, g)
        assert np.isfinite(want['offsets'][..., 0]).mean() > 0.8
        for k in KEYS:
            assert np.array_equal(got[k], want[k], equal_nan=True), (method, frames.shape[0], k)


def test_noalloc_refuses_a_frame_larger_than_its_array(gpu):
    """2400 x 2400 frames, window 3, int8: one frame's maps are 2398^2 x 12 bytes > 64 MiB"""
    cor = framecor.Correlator(2400, 2400, [10], [10], window=3, search=2, method='int8')
    cor.set_reference(np.random.default_rng(0).integers(0, 256, (2400, 2400)).astype(np.uint8))
    frames = torch.zeros((1, 2400, 2400), dtype=torch.uint8, device='cuda')
    out = torch.empty((1, 1, 1, 4), dtype=torch.float32, device='cuda')
    status = framecor._lib.framecor_correlate_device_noalloc(cor._c, frames.data_ptr(), 1, None, out.data_ptr(),
                                                             out.data_ptr(), out.data_ptr(), out.data_ptr(), None)
    assert status == 1
    assert 'exceed the 64 MiB scratch' in framecor._lib.framecor_last_error().decode()


def test_noalloc_refuses_fft_spectra_beyond_lds(gpu):
    """The FFT method with a search window whose spectrum does not fit in LDS (window 128, search 64: 256)
    needs the correlator's workspace: refused with a clear error"""
    cor = framecor.Correlator(600, 600, [100], [100], window=128, search=64, method='fft')
    cor.set_reference(np.random.default_rng(1).integers(0, 256, (600, 600)).astype(np.uint8))
    frames = torch.zeros((1, 600, 600), dtype=torch.uint8, device='cuda')
    out = torch.empty((1, 1, 1, 4), dtype=torch.float32, device='cuda')
    status = framecor._lib.framecor_correlate_device_noalloc(cor._c, frames.data_ptr(), 1, None, out.data_ptr(),
                                                             out.data_ptr(), out.data_ptr(), out.data_ptr(), None)
    assert status == 1 and 'workspace' in framecor._lib.framecor_last_error().decode()
