"""
The exact NCC surfaces and the periodicity of slipcam-auto (python/slipcam/auto_slip.py:105-129) on the
GPU against the oracle: the FP64 surfaces are the oracle's bit for bit (both are the exact integer sums
normalized by the same FP64 expression), and so is the second peak taken from them.
"""

import numpy as np
import pytest

from oracle import OracleCorrelator
from slipcam_amd import auto_slip, framecor, synthetic
from slipcam_amd.framecor import window_grid

pytestmark = pytest.mark.gpu


def mixed_texture(h=240, w=320, seed=7):
    """Smoothed noise, a 12 x 12 tile repeated over the lower right quarter and a flat block"""
    img = synthetic.to_uint8(synthetic.texture(h, w, seed=seed))
    tile = np.random.default_rng(seed).integers(0, 256, (12, 12), dtype=np.uint8)
    img[h // 2:, w // 2:] = np.tile(tile, (h // 24 + 1, w // 24 + 1))[:h - h // 2, :w - w // 2]
    img[:h // 3, :w // 4] = 77
    return img


def surfaces_of(cor, frame):
    import torch
    s = cor.surfaces(torch.as_tensor(frame, device='cuda'))
    torch.cuda.synchronize()
    return s.cpu().numpy()


@pytest.mark.parametrize('W,R,skip', [(32, 16, 32), (24, 9, 20), (64, 24, 40)])
def test_surfaces_bit_exact(gpu, W, R, skip):
    d = synthetic.small_shifts(n=2)
    down, across = window_grid(256, 320, W, R, skip)
    cor = framecor.Correlator(256, 320, down, across, window=W, search=R, method='int8')
    cor.set_reference(d['ref'])
    ora = OracleCorrelator(256, 320, down, across, window=W, search=R)
    ora.set_reference(d['ref'])
    for frame in (d['ref'], d['frames'][1]):
        got = surfaces_of(cor, frame)
        want = ora.correlate_direct(frame[None], surfaces=True)['surface'][0]
        valid = np.isfinite(want[..., 0, 0])   # the oracle leaves flat windows NaN
        assert valid.any()
        assert np.array_equal(got[valid], want[valid])


def test_surfaces_720p_sampled_windows(gpu):
    d = synthetic.fault_video(n=2)
    W, R = 64, 48
    down, across = window_grid(720, 1280, W, R, 16)
    cor = framecor.Correlator(720, 1280, down, across, window=W, search=R, method='int8')
    cor.set_reference(d['ref'])
    got = surfaces_of(cor, d['frames'][1])
    ora = OracleCorrelator(720, 1280, down, across, window=W, search=R)
    ora.set_reference(d['ref'])
    sample = np.random.default_rng(5).choice(len(down) * len(across), 24, replace=False)
    want = ora.correlate_direct(d['frames'][1][None], windows=sample, surfaces=True)['surface'][0]
    got = got.reshape(-1, 2 * R + 1, 2 * R + 1)[sample]
    want = want.reshape(-1, 2 * R + 1, 2 * R + 1)[sample]
    valid = np.isfinite(want[:, 0, 0])
    assert valid.sum() > 12
    assert np.array_equal(got[valid], want[valid])


@pytest.mark.parametrize('exclude', [5, 0, 16, 20, -3])
def test_periodicity_matches_oracle(gpu, exclude):
    ref = mixed_texture()
    W, R = 32, 16
    down, across = window_grid(240, 320, W, R, 16)
    cor = framecor.Correlator(240, 320, down, across, window=W, search=R, method='int8')
    cor.set_reference(ref)
    ora = OracleCorrelator(240, 320, down, across, window=W, search=R)
    ora.set_reference(ref)
    got = cor.periodicity(ref, exclude=exclude)
    assert np.array_equal(got, ora.periodicity(ref, exclude=exclude))
    # flat windows (the block of 77) are periodic
    assert (got == 1.0).any()


def test_periodicity_720p_subgrid(gpu):
    """The default slipcam-auto geometry (window 64, search 48, skip 16) on the fault video's reference,
    against the oracle on every 6th row and 9th column of windows"""
    d = synthetic.fault_video(n=1)
    W, R = 64, 48
    down, across = window_grid(720, 1280, W, R, 16)
    got = auto_slip.periodicity(d['ref'], down, across, W, R)
    assert got.shape == (len(down), len(across))
    sub_d, sub_a = down[::6], across[::9]
    ora = OracleCorrelator(720, 1280, sub_d, sub_a, window=W, search=R)
    ora.set_reference(d['ref'])
    assert np.array_equal(got[::6, ::9], ora.periodicity(d['ref']))
    assert np.all((got >= -1.0) & (got <= 1.0))


def test_periodicity_wrapper_and_errors(gpu):
    ref = mixed_texture()
    W, R = 32, 16
    down, across = window_grid(240, 320, W, R, 16)
    cor = framecor.Correlator(240, 320, down, across, window=W, search=R, method='fft')
    cor.set_reference(ref)
    # the surfaces do not depend on the method of the correlator
    assert np.array_equal(auto_slip.periodicity(ref, down, across, W, R), cor.periodicity(ref))
    with pytest.raises(ValueError, match='leaves the frame'):
        auto_slip.periodicity(ref, [4], [40], W, R)
    with pytest.raises(ValueError, match='uint8'):
        auto_slip.periodicity(ref.astype(np.float64), down, across, W, R)
    fresh = framecor.Correlator(240, 320, down, across, window=W, search=R, method='int8')
    with pytest.raises(RuntimeError, match='set_reference'):
        fresh.periodicity(ref)


@pytest.mark.parametrize('W,R', [(144, 20), (131, 8)])
def test_periodicity_windows_beyond_the_int8_tiles(gpu, W, R):
    """Windows wider than the int8 form's tiles (129): the correlator is the FFT form's, the surfaces
    come from the same exact integer kernel, against the oracle"""
    ref = mixed_texture(256, 320, seed=11)
    down, across = window_grid(256, 320, W, R, 40)
    got = auto_slip.periodicity(ref, down, across, W, R)
    ora = OracleCorrelator(256, 320, down, across, window=W, search=R)
    ora.set_reference(ref)
    assert got.shape == (len(down), len(across)) and got.size >= 3
    assert np.array_equal(got, ora.periodicity(ref))
