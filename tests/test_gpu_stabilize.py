"""
The coarse tracking of slipcam-auto on the GPU (slipcam_amd/stabilize.py; python/slipcam/stabilize.py:
22-96): the peaks of the exact NCC surfaces with per-window gross offsets equal the oracle's bit for
bit, so coarse_track equals the same frame loop driven by the oracle, and recovers the camera motion.
"""

import numpy as np
import pytest
import torch

from oracle import OracleCorrelator
from slipcam_amd import framecor, stabilize, synthetic
from slipcam_amd.framecor import window_grid

pytestmark = pytest.mark.gpu


def textured(h=240, w=320, seed=9):
    ref = synthetic.to_uint8(synthetic.texture(h, w, sigma=2.0, seed=seed))
    ref[:80, :80] = 77
    return ref


@pytest.mark.parametrize('W,R,skip', [(32, 16, 16), (96, 32, 48)])
def test_ncc_peaks_bit_exact(gpu, W, R, skip):
    ref = textured()
    frame = np.roll(ref, (3, -2), axis=(0, 1))
    frame[100:140, 200:260] = 128
    down, across = window_grid(240, 320, W, R, skip)
    gross = np.random.default_rng(4).integers(-40, 41, (len(down), len(across), 2)).astype(np.int32)
    gross[0, 0, 0] = R - down[0] - 1          # one search window starts a row above the frame
    cor = framecor.Correlator(240, 320, down, across, window=W, search=R, method='int8')
    cor.set_reference(ref)
    ora = OracleCorrelator(240, 320, down, across, window=W, search=R)
    ora.set_reference(ref)
    for g in (None, gross):
        off, peak = cor.ncc_peaks(frame, g)
        o_off, o_peak = ora.ncc_peaks(frame, g)
        assert np.array_equal(off, o_off, equal_nan=True)
        assert np.array_equal(peak, o_peak)
    assert np.isnan(off).any() and np.isfinite(off).any()


def test_surfaces_with_gross_bit_exact(gpu):
    ref = textured()
    frame = np.roll(ref, (1, 4), axis=(0, 1))
    W, R = 32, 16
    down, across = window_grid(240, 320, W, R, 16)
    gross = np.random.default_rng(5).integers(-20, 21, (len(down), len(across), 2)).astype(np.int32)
    cor = framecor.Correlator(240, 320, down, across, window=W, search=R, method='int8')
    cor.set_reference(ref)
    M = 2 * R + 1
    out = torch.empty((len(down) * len(across), M, M), dtype=torch.float64, device='cuda')
    f = torch.as_tensor(frame, device='cuda')
    g = torch.as_tensor(gross, device='cuda')
    framecor._check(framecor._lib.framecor_ncc_surfaces_device(cor._c, f.data_ptr(), g.data_ptr(), 0, out.shape[0],
                                                               out.data_ptr(), None))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ora = OracleCorrelator(240, 320, down, across, window=W, search=R)
    ora.set_reference(ref)
    want = ora.correlate_direct(frame[None], gross=gross[None], surfaces=True)['surface'][0].reshape(-1, M, M)
    valid = np.isfinite(want[:, 0, 0])
    assert valid.any() and not valid.all()
    assert np.array_equal(got[valid], want[valid])
    # the invalid ones: NaN where the search window leaves the frame (flat templates give 0 here)
    sy = np.asarray(down)[:, None] + gross[..., 0] - R
    sx = np.asarray(across)[None, :] + gross[..., 1] - R
    outside = ((sy < 0) | (sx < 0) | (sy + W + 2 * R > 240) | (sx + W + 2 * R > 320)).ravel()
    assert np.isnan(got[outside]).all() and not np.isnan(got[~outside]).any()


def test_coarse_track_equals_the_oracle_loop(gpu):
    v = synthetic.similarity_video(n=7, ref=3)
    frames, ref = v['frames'], v['ref']
    params, used, ngrid = stabilize.coarse_track(frames, ref, v['text'])
    g = stabilize.Grid(frames[ref], v['text'], 96, 96, 32)
    ora = OracleCorrelator(480, 640, g.yc - g.half, g.xc - g.half, window=96, search=32)
    ora.set_reference(frames[ref])
    o_params, o_used = stabilize.track_frames(frames, ref, g, ora.ncc_peaks)
    assert ngrid == len(g.Xg)
    assert np.array_equal(params, o_params) and np.array_equal(used, o_used)
    err = params - v['params']
    assert np.abs(err[:, :2]).max() < 0.1 and np.abs(err[:, 2:]).max() < 3e-4


def test_coarse_track_start_and_empty_grid(gpu):
    v = synthetic.similarity_video(n=5, ref=2)
    params, used, _ = stabilize.coarse_track(v['frames'], 2, v['text'], start=1)
    assert np.isnan(params[0]).all() and used[0] == 0 and np.isfinite(params[1:]).all()
    # every window under the text mask: no window is used, every frame keeps the previous transform
    params, used, ngrid = stabilize.coarse_track(v['frames'], 2, np.ones((480, 640), bool))
    assert ngrid == 0 and (used == 0).all() and np.array_equal(params, np.zeros((5, 4)))


def test_ncc_peak_single_pair(gpu):
    ref = textured()
    t = ref[60:92, 100:132]
    search = np.roll(ref, (4, -3), axis=(0, 1))[44:108, 84:148]
    dy, dx, pk = stabilize.ncc_peak(t, search)
    assert abs(pk - 1.0) < 1e-12 and round(dy) == 4 and round(dx) == -3
    with pytest.raises(ValueError):
        stabilize.ncc_peak(t, search[:-1])
