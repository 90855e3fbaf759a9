"""
The reference's five known-answer tests (tests/test_framecor.py of the reference) through the pybind11
module itself, called as the reference's tests call framecor._framecor: Correlator(..., backend='cuda')
(the reference's name of its GPU backend) and the default method, plus the int8 method.
"""

import numpy as np
import pytest

from slipcam_amd import framecor
from slipcam_amd.framecor import _framecor
from test_gpu_framecor import frames

pytestmark = pytest.mark.gpu

METHODS = ['fft', 'int8']


def correlator(method, window=32, search=24, skip=24, h=240, w=320, margin=0):
    down, across = framecor.window_grid(h, w, window, search, skip, margin=margin)
    return _framecor.Correlator(h, w, down, across, window=window, search=search, backend='cuda', method=method)


@pytest.mark.parametrize('method', METHODS)
def test_shifts(gpu, method):
    shifts = [(0, 0), (0.3, -0.7), (5.25, -12.5), (-20.4, 15.1)]
    ref, sec = frames(shifts)
    cor = correlator(method)
    assert cor.backend == 'cuda' and _framecor.has_cuda()
    cor.set_reference(ref)
    r = cor.correlate(sec)
    assert r['offsets'].shape == (len(shifts),) + cor.shape + (2,)
    for k, s in enumerate(shifts):
        err = r['offsets'][k] - np.array(s, np.float32)
        assert np.all(np.isfinite(err))
        assert np.median(np.abs(err)) < 0.05
        assert np.abs(err).max() < 0.2
        assert np.median(r['snr'][k]) > 10
        assert np.median(r['peak'][k]) > 0.9


def test_methods_agree(gpu):
    """The reference's cross-backend bar, between the FFT and int8 methods"""
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
    flat = ref.copy()
    flat[:120] = 100
    cor.set_reference(flat)
    r = cor.correlate(sec)
    assert np.all(np.isnan(r['offsets'][0, 0]))
    assert np.all(np.isfinite(r['offsets'][0, -1]))


def test_arguments(gpu):
    with pytest.raises(ValueError):
        _framecor.Correlator(240, 320, [0], [300], window=32, search=8, backend='cuda')    # outside
    with pytest.raises(ValueError):
        _framecor.Correlator(240, 320, [0], [0], window=32, search=8, backend='gpu')
    cor = correlator('fft')
    with pytest.raises(RuntimeError, match='set_reference first'):
        cor.correlate(np.zeros((1, 240, 320), np.uint8))
    cor.set_reference(np.zeros((240, 320), np.uint8))
    with pytest.raises(ValueError, match='frames x height x width'):
        cor.correlate(np.zeros((1, 240, 321), np.uint8))
    # forcecast: float frames are cast to uint8, as the reference's module does
    r = cor.correlate(np.zeros((1, 240, 320), np.float64))
    assert r['offsets'].dtype == np.float32 and np.all(np.isnan(r['offsets']))


def test_same_bits_as_the_package_class(gpu):
    """The package's Correlator is this class extended: the same native object and results"""
    ref, sec = frames([(0.3, -0.7), (5.25, -12.5)])
    a = correlator('int8')
    b = framecor.Correlator(240, 320, *framecor.window_grid(240, 320, 32, 24, 24), window=32, search=24,
                            method='int8')
    assert isinstance(b, _framecor.Correlator)
    for c in (a, b):
        c.set_reference(ref)
    ra, rb = a.correlate(sec), b.correlate(sec)
    for k in ra:
        assert np.array_equal(ra[k], rb[k], equal_nan=True), k
