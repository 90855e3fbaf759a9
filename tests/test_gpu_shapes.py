"""
Parity of both GPU forms over the range of geometries they support, against the oracle:

  int8   every tile shape family (template column blocks 2, 3, 5, 9; 1 to 9 bands; edge waves or
         padded bands; the largest window 129 and search 71) and every kernel variant of each, against
         the FP64 direct-NCC oracle, with gross offsets that push some search windows out of the frame
  fft    search windows whose size has factors the radix-2/4/8 passes do not cover (3, 5, 7, 11, the
         primes 29, 37, 67 and 197: direct passes), the largest ones held in LDS (192, 197) and larger ones
         whose spectra go to the global workspace (201, 256), every size of the register-resident
         kernel (64, 96, 128, 160, 192), against the restatement of the reference algorithm; a size
         above 2048 is refused
  both   every window invalid; a flat reference
"""

import numpy as np
import pytest

from oracle import OracleCorrelator
from slipcam_amd import framecor, synthetic
from parity import assert_fft_parity, assert_int8_parity

pytestmark = pytest.mark.gpu


def workload(W, R, n=2, seed=0):
    """A few windows (2 x 3) of a smoothed texture, n frames shifted by up to 2 px with noise, and
    gross offsets up to R / 2 + 2 (some search windows leave the frame)"""
    skip = 12
    h, w = W + 2 * R + 2 * skip, W + 2 * R + 3 * skip + 4
    base = synthetic.texture(h, w, sigma=1.5, seed=seed)
    rng = np.random.default_rng(seed + 1)
    frames = np.stack([synthetic.to_uint8(synthetic.warp(base, *rng.uniform(-2, 2, 2)) + rng.normal(0, 2, base.shape))
                       for _ in range(n)])
    down, across = framecor.window_grid(h, w, W, R, skip)
    gross = rng.integers(-(R // 2 + 2), R // 2 + 3, size=(n, len(down), len(across), 2)).astype(np.int32)
    return dict(ref=synthetic.to_uint8(base), frames=frames, down=down, across=across, gross=gross, h=h, w=w)


# the 32 x 32 kernel (search 16, 32, 48, 64) at every template width class it instantiates:
# (3, 16), (20, 16), (40, 48), (64, 16/32), (80, 32), (90, 16), (100, 48), (125, 64), (128, 64);
# the four-wave kernel at each of its shapes: (32, 32), (32, 48), (64, 32), (64, 48); the one-frame kernel
# with the edge waves and 113-128 template columns (the chip's L rows, 176-byte template rows): (100, 48),
# (125, 64), (128, 16), (128, 32), (128, 64)
INT8_SHAPES = [(3, 1), (16, 8), (20, 5), (33, 17), (48, 24), (40, 56), (64, 16), (64, 32), (96, 40), (128, 64),
               (129, 71), (3, 16), (20, 16), (40, 48), (80, 32), (90, 16), (100, 48), (125, 64), (32, 32), (32, 48),
               (64, 48), (128, 16), (128, 32)]


@pytest.mark.parametrize('W,R', INT8_SHAPES)
def test_int8_shapes_vs_fp64_oracle(gpu, W, R):
    d = workload(W, R, seed=W + R)
    o = OracleCorrelator(d['h'], d['w'], d['down'], d['across'], window=W, search=R)
    o.set_reference(d['ref'])
    ref = o.correlate_direct(d['frames'], gross=d['gross'])
    valid = np.isfinite(ref['offsets'][..., 0])
    assert valid.any()
    cor = framecor.Correlator(d['h'], d['w'], d['down'], d['across'], window=W, search=R, method='int8')
    cor.set_reference(d['ref'])
    for variant in ('simple', 'padded', 'piped', 'wide', 'quad', 'solo'):
        cor.set_variant(variant)
        got = cor.correlate(d['frames'], gross=d['gross'])
        assert_int8_parity(got, ref)


# L = 29, 35, 44, 77, 96, 192 (the largest power-of-two-and-3 size in LDS), 37 and 197 (primes: direct
# passes; 197 the largest odd size in LDS), 256 (window 128 and 160: spectra in the global workspace),
# 201 = 3 x 67 (global workspace and a direct pass of 67); 64, 96, 128, 160 and 192 are the sizes of the
# register-resident kernel
FFT_SHAPES = [(15, 7), (21, 7), (24, 10), (33, 22), (16, 40), (64, 64), (25, 6), (97, 50), (128, 64), (160, 48),
              (101, 50), (32, 16), (64, 32), (80, 40)]


@pytest.mark.parametrize('W,R', FFT_SHAPES)
def test_fft_shapes_vs_reference_algorithm(gpu, W, R):
    d = workload(W, R, seed=3 * W + R)
    o = OracleCorrelator(d['h'], d['w'], d['down'], d['across'], window=W, search=R)
    o.set_reference(d['ref'])
    ref = o.correlate_fft(d['frames'], gross=d['gross'])
    surf = o.correlate_direct(d['frames'], gross=d['gross'], surfaces=True)['surface']
    got = framecor.Correlator(d['h'], d['w'], d['down'], d['across'], window=W, search=R, method='fft')
    got.set_reference(d['ref'])
    out = got.correlate(d['frames'], gross=d['gross'])
    assert np.isfinite(ref['offsets'][..., 0]).any()
    assert assert_fft_parity(out, ref, R, surf=surf, gross=d['gross']) <= 1


def test_fft_refuses_a_search_window_beyond_2048(gpu):
    with pytest.raises(ValueError, match='at most 2048'):
        framecor.Correlator(2400, 2400, [1100], [1100], window=49, search=1000, method='fft')


@pytest.mark.parametrize('method', ['fft', 'int8'])
def test_every_window_invalid_and_flat_reference(gpu, method):
    d = workload(32, 16)
    cor = framecor.Correlator(d['h'], d['w'], d['down'], d['across'], window=32, search=16, method=method)
    cor.set_reference(d['ref'])
    far = np.full_like(d['gross'], 10 ** 4)
    r = cor.correlate(d['frames'], gross=far)
    assert np.isnan(r['offsets']).all() and (r['peak'] == 0).all() and (r['snr'] == 0).all()
    assert (r['cov'][..., :2] == 99).all() and (r['cov'][..., 2] == 0).all()
    cor.set_reference(np.full((d['h'], d['w']), 128, np.uint8))
    r = cor.correlate(d['frames'])
    assert np.isnan(r['offsets']).all() and (r['peak'] == 0).all()
