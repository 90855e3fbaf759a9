"""
The reference's tests of framecor (tests/test_framecor.py of the reference), on the HIP path, for both
methods: known shifts of a smoothed texture, the two methods agreeing (the reference's cross-backend
bar), gross offsets, invalid windows, argument errors.
"""

import numpy as np
import pytest
from scipy import ndimage

from slipcam_amd import framecor

pytestmark = pytest.mark.gpu

METHODS = ['fft', 'int8']


def texture(h=240, w=320, seed=0):
    img = ndimage.gaussian_filter(np.random.default_rng(seed).normal(size=(h, w)), 2.0)
    return (img - img.min()) / (img.max() - img.min()) * 200 + 20


def frames(shifts, h=240, w=320):
    base = texture(h, w)
    to8 = lambda img: np.clip(np.rint(img), 0, 255).astype(np.uint8)  # noqa: E731
    return to8(base), np.stack([to8(ndimage.shift(base, s, order=3, mode='reflect')) for s in shifts])


def correlator(method, window=32, search=24, skip=24, h=240, w=320, margin=0):
    down, across = framecor.window_grid(h, w, window, search, skip, margin=margin)
    return framecor.Correlator(h, w, down, across, window=window, search=search, method=method)


@pytest.mark.parametrize('method', METHODS)
def test_shifts(gpu, method):
    shifts = [(0, 0), (0.3, -0.7), (5.25, -12.5), (-20.4, 15.1)]
    ref, sec = frames(shifts)
    cor = correlator(method)
    assert cor.backend == 'hip' and cor.method == method
    cor.set_reference(ref)
    r = cor.correlate(sec)
    assert r['offsets'].shape == (len(shifts),) + cor.shape + (2,)
    assert all(r[k].dtype == np.float32 for k in r)
    for k, s in enumerate(shifts):
        err = r['offsets'][k] - np.array(s, np.float32)
        assert np.all(np.isfinite(err))
        assert np.median(np.abs(err)) < 0.05, (s, np.median(np.abs(err), axis=(0, 1)))
        assert np.abs(err).max() < 0.2, (s, np.abs(err).max())
        assert np.median(r['snr'][k]) > 10
        assert np.median(r['peak'][k]) > 0.9


def test_methods_agree(gpu):
    ref, sec = frames([(0.3, -0.7), (5.25, -12.5), (-20.4, 15.1)])
    out = {}
    for method in METHODS:
        cor = correlator(method)
        cor.set_reference(ref)
        out[method] = cor.correlate(sec)
    for key, tol in (('offsets', 1e-3), ('peak', 1e-4)):
        assert np.abs(out['fft'][key] - out['int8'][key]).max() < tol, key
    assert np.abs(out['fft']['snr'] / out['int8']['snr'] - 1).max() < 1e-2


@pytest.mark.parametrize('method', METHODS)
def test_gross(gpu, method):
    shift = (30.4, -27.6)
    ref, sec = frames([shift])
    cor = correlator(method, search=6, skip=16, margin=32)
    cor.set_reference(ref)
    found = cor.correlate(sec, gross=np.array([[30, -28]]))['offsets'][0]
    assert np.median(np.abs(found - np.array(shift, np.float32))) < 0.05
    g = np.broadcast_to(np.array([30, -28], np.int32), (1,) + cor.shape + (2,))
    assert np.array_equal(cor.correlate(sec, gross=g)['offsets'][0], found, equal_nan=True)
    assert np.median(np.abs(cor.correlate(sec)['offsets'][0] - np.array(shift, np.float32))) > 1


@pytest.mark.parametrize('method', METHODS)
def test_invalid(gpu, method):
    ref, sec = frames([(0, 0)])
    cor = correlator(method)
    cor.set_reference(ref)
    r = cor.correlate(sec, gross=np.array([[0, 1000]]))
    assert np.all(np.isnan(r['offsets'])) and np.all(r['snr'] == 0)
    assert np.all(r['peak'] == 0) and np.all(r['cov'][..., :2] == 99) and np.all(r['cov'][..., 2] == 0)
    flat = ref.copy()
    flat[:120] = 100
    cor.set_reference(flat)
    r = cor.correlate(sec)
    assert np.all(np.isnan(r['offsets'][0, 0]))
    assert np.all(np.isfinite(r['offsets'][0, -1]))


@pytest.mark.parametrize('method', METHODS)
def test_arguments(gpu, method):
    cor = correlator(method)
    with pytest.raises(RuntimeError):
        cor.correlate(np.zeros((1, 240, 320), np.uint8))
    cor.set_reference(np.zeros((240, 320), np.uint8))
    with pytest.raises(ValueError):
        cor.correlate(np.zeros((1, 240, 321), np.uint8))
    with pytest.raises(ValueError):
        cor.set_reference(np.zeros((240, 321), np.uint8))
    with pytest.raises(ValueError):
        cor.correlate(np.zeros((2, 240, 320), np.uint8), gross=np.zeros((1, 2), np.int32))
    r = cor.correlate(np.zeros((0, 240, 320), np.uint8))
    assert r['offsets'].shape == (0,) + cor.shape + (2,) and r['cov'].shape == (0,) + cor.shape + (3,)
    # a flat reference: every window is invalid
    r = cor.correlate(np.zeros((1, 240, 320), np.uint8))
    assert np.all(np.isnan(r['offsets'])) and np.all(r['snr'] == 0)
