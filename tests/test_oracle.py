"""
Pins of the CPU oracle (oracle/framecor_oracle.c) before it is trusted as the checker of the HIP path.

Each function of the oracle is checked against something that fixes it independently: a library
routine (numpy.fft, numpy.cumsum, scipy.signal.correlate2d), brute force on tiny inputs, closed forms
of the peak fit, and the known-answer tests of the reference (tests/test_framecor.py:16-95 of the
reference: smoothed texture, known shifts, gross offsets, invalid windows).
"""

import os

import numpy as np
import pytest
from scipy import ndimage, signal

import oracle
from oracle import OracleCorrelator
from slipcam_amd import synthetic
from slipcam_amd.framecor import window_grid

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'tiny.npz')


# ---- the FFT (the reference links FFTW; the oracle has its own)

@pytest.mark.parametrize('n', [1, 2, 3, 4, 5, 6, 8, 12, 16, 27, 44, 64, 80, 96, 97, 128, 160, 192])
def test_cfft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
    ref = np.fft.fft(x.astype(np.complex128))
    got = oracle.cfft(x, -1)
    assert np.abs(got - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())
    inv = oracle.cfft(x, +1)
    assert np.abs(inv - np.fft.ifft(x.astype(np.complex128)) * n).max() <= 1e-5 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize('L', [27, 44, 64, 80, 160])
def test_rfft2_matches_numpy(L):
    x = np.random.default_rng(L).normal(size=(L, L)).astype(np.float32)
    spec = oracle.rfft2(x)
    ref = np.fft.rfft2(x.astype(np.float64))
    assert spec.shape == ref.shape
    assert np.abs(spec - ref).max() <= 2e-5 * np.abs(ref).max()
    back = oracle.irfft2(spec)
    assert np.abs(back / (L * L) - x).max() < 1e-4


# ---- summed-area tables (bit-exact)

@pytest.mark.parametrize('shape', [(1, 1), (3, 7), (64, 65), (240, 320)])
def test_summed_area_is_cumsum(shape):
    frame = np.random.default_rng(0).integers(0, 256, size=shape).astype(np.uint8)
    sat, sat2 = oracle.summed_area(frame)
    ref = np.zeros((shape[0] + 1, shape[1] + 1), np.int64)
    ref[1:, 1:] = frame.astype(np.int64).cumsum(0).cumsum(1)
    ref2 = np.zeros_like(ref)
    ref2[1:, 1:] = (frame.astype(np.int64) ** 2).cumsum(0).cumsum(1)
    assert np.array_equal(sat, ref) and np.array_equal(sat2, ref2)


def test_summed_area_extremes():
    frame = np.full((720, 1280), 255, np.uint8)
    sat, sat2 = oracle.summed_area(frame)
    assert sat[-1, -1] == 255 * 720 * 1280 and sat2[-1, -1] == 255 * 255 * 720 * 1280


# ---- exact correlation sums

def test_raw_direct_brute_force():
    rng = np.random.default_rng(1)
    chip = rng.integers(0, 256, size=(9, 9)).astype(np.uint8)
    t = rng.integers(0, 256, size=(4, 4)).astype(np.uint8)
    got = oracle.raw_direct(chip, t)
    M = 6
    ref = np.zeros((M, M), np.int64)
    for u in range(M):
        for v in range(M):
            for i in range(4):
                for j in range(4):
                    ref[u, v] += int(t[i, j]) * int(chip[u + i, v + j])
    assert np.array_equal(got, ref)


def test_raw_direct_is_correlate2d():
    rng = np.random.default_rng(2)
    chip = rng.integers(0, 256, size=(64, 64)).astype(np.uint8)
    t = rng.integers(0, 256, size=(32, 32)).astype(np.uint8)
    ref = signal.correlate2d(chip.astype(np.int64), t.astype(np.int64), mode='valid')
    assert np.array_equal(oracle.raw_direct(chip, t), ref)


def test_raw_direct_extreme_values():
    chip = np.full((192, 192), 255, np.uint8)
    t = np.full((128, 128), 255, np.uint8)
    assert np.all(oracle.raw_direct(chip, t) == 255 * 255 * 128 * 128)


# ---- the peak fit, SNR and covariance (cxx/surface.h:71-146), closed forms

def paraboloid(u0, v0, a=0.05, b=0.03, c=0.01, top=0.95, M=9, u=4, v=4):
    """A concave quadratic surface with its vertex at (u + u0, v + v0)"""
    uu, vv = np.mgrid[0:M, 0:M].astype(np.float64)
    du, dv = uu - (u + u0), vv - (v + v0)
    return top - a * du ** 2 - b * dv ** 2 - c * du * dv


@pytest.mark.parametrize('fp64', [False, True])
@pytest.mark.parametrize('u0,v0', [(0.0, 0.0), (0.3, -0.2), (-0.45, 0.4), (0.1, 0.49)])
def test_fit_finds_the_vertex(fp64, u0, v0):
    s = paraboloid(u0, v0)
    off, peak, snr, cov = oracle.write_peak(s, 4, 4, search=4, gross=(0, 0), fp64=fp64)
    tol = 1e-12 if fp64 else 2e-5
    assert abs(off[0] - u0) < max(tol, 1e-7) and abs(off[1] - v0) < max(tol, 1e-7)
    assert peak == np.float32(s[4, 4])


def test_fit_offsets_include_the_integer_peak_and_gross():
    s = paraboloid(0.25, -0.125, M=33, u=20, v=7)
    off, _, _, _ = oracle.write_peak(s, 20, 7, search=16, gross=(5, -3), fp64=True)
    assert np.allclose(off, [5 + 20 - 16 + 0.25, -3 + 7 - 16 - 0.125], atol=1e-6)


def test_fit_falls_back_to_parabolas():
    # convex along u (c3 > 0): the fitted quadratic is rejected; along v a parabola with its vertex
    # at +0.3; along u the parabola is not concave, so du = 0
    uu, vv = np.mgrid[0:5, 0:5].astype(np.float64)
    s = 0.5 + 0.01 * (uu - 2) ** 2 - 0.04 * (vv - 2.3) ** 2
    for fp64 in (False, True):
        off, _, _, _ = oracle.write_peak(s, 2, 2, search=2, fp64=fp64)
        assert abs(off[0]) < 1e-6 and abs(off[1] - 0.3) < 1e-5


def test_boundary_peak_has_no_fit_and_no_covariance():
    s = paraboloid(0.3, 0.3, M=9, u=0, v=4)
    off, peak, snr, cov = oracle.write_peak(s, 0, 4, search=4, fp64=True)
    assert off[0] == -4 and off[1] == 0
    assert np.array_equal(cov, np.array([99, 99, 0], np.float32))


def test_snr_of_constant_and_delta_surfaces():
    const = np.full((33, 33), 0.5)
    _, _, snr, _ = oracle.write_peak(const, 16, 16, search=16, fp64=True)
    assert abs(snr - 1.0) < 1e-7
    delta = np.zeros((33, 33))
    delta[16, 16] = 0.8
    _, _, snr, _ = oracle.write_peak(delta, 16, 16, search=16, fp64=True)
    assert snr == 0.0
    # a peak of 0.9 over a box of 0.3: 0.81 / 0.09 = 9; the box is clipped at the surface border
    s = np.full((33, 33), 0.3)
    s[1, 1] = 0.9
    _, _, snr, _ = oracle.write_peak(s, 1, 1, search=16, stat_size=21, fp64=True)
    assert abs(snr - 9.0) < 1e-6


def test_covariance_closed_form():
    # f0 = 0.9, axis neighbours 0.8, diagonals 0.7, T = 1024: dxx = dyy = 0.2 T, dxy = 0,
    # n2 = 0.2, n4 = 0.005 T: cov = n2 / d + n4 / d^2 on both axes, 0 across
    s = np.full((5, 5), 0.1)
    s[1:4, 1:4] = [[0.7, 0.8, 0.7], [0.8, 0.9, 0.8], [0.7, 0.8, 0.7]]
    T = 1024.0
    d = 0.2 * T
    expect = 0.2 / d + 0.005 * T / d ** 2
    for fp64 in (False, True):
        _, _, _, cov = oracle.write_peak(s, 2, 2, search=2, template_size=1024, fp64=fp64)
        assert abs(cov[0] - expect) < 1e-6 * expect * (1 if fp64 else 100)
        assert abs(cov[1] - expect) < 1e-6 * expect * (1 if fp64 else 100)
        assert abs(cov[2]) < 1e-9


def test_covariance_cross_term_closed_form():
    # s = 0.9 - 0.06 u^2 - 0.06 v^2 + 0.08 u v: the finite differences of a quadratic are its exact
    # second derivatives, so dxx = dyy = a = 0.12 T and dxy = b = 0.08 T; with n2 = 2 (1 - f0) and
    # n4 = (1 - f0)^2 T / 2 the pycuampcor expressions reduce to
    #   cov_uu = cov_vv = (-n2 w a + n4 (a^2 + b^2)) / w^2,  cov_uv = b (n2 w - 2 a n4) / w^2,  w = b^2 - a^2
    uu, vv = np.mgrid[0:5, 0:5].astype(np.float64) - 2
    s = 0.9 - 0.06 * uu ** 2 - 0.06 * vv ** 2 + 0.08 * uu * vv
    T = 1024.0
    a, b = 0.12 * T, 0.08 * T
    n2, n4 = 2 * 0.1, 0.1 ** 2 * T / 2
    w = b * b - a * a
    _, _, _, cov = oracle.write_peak(s, 2, 2, search=2, template_size=1024, fp64=True)
    assert abs(cov[0] - (-n2 * w * a + n4 * (a * a + b * b)) / w ** 2) < 1e-9
    assert abs(cov[1] - cov[0]) < 1e-12
    assert abs(cov[2] - b * (n2 * w - 2 * a * n4) / w ** 2) < 1e-9
    assert cov[2] < 0   # the convention of the reference: positive mixed curvature, negative cross term


# ---- the correlators: the reference's known-answer tests (tests/test_framecor.py:16-95)

def ref_texture(h=240, w=320, seed=0):
    img = ndimage.gaussian_filter(np.random.default_rng(seed).normal(size=(h, w)), 2.0)
    return (img - img.min()) / (img.max() - img.min()) * 200 + 20


def ref_frames(shifts, h=240, w=320):
    base = ref_texture(h, w)
    to8 = lambda img: np.clip(np.rint(img), 0, 255).astype(np.uint8)  # noqa: E731
    return to8(base), np.stack([to8(ndimage.shift(base, s, order=3, mode='reflect')) for s in shifts])


def ref_oracle(window=32, search=24, skip=24, h=240, w=320, margin=0):
    down, across = window_grid(h, w, window, search, skip, margin=margin)
    return OracleCorrelator(h, w, down, across, window=window, search=search)


@pytest.mark.parametrize('method', ['correlate_fft', 'correlate_direct'])
def test_known_shifts(method):
    shifts = [(0, 0), (0.3, -0.7), (5.25, -12.5), (-20.4, 15.1)]
    ref, sec = ref_frames(shifts)
    cor = ref_oracle()
    cor.set_reference(ref)
    r = getattr(cor, method)(sec)
    assert r['offsets'].shape == (len(shifts),) + cor.shape + (2,)
    for k, s in enumerate(shifts):
        err = r['offsets'][k] - np.array(s, np.float32)
        assert np.all(np.isfinite(err))
        assert np.median(np.abs(err)) < 0.05
        assert np.abs(err).max() < 0.2
        assert np.median(r['snr'][k]) > 10
        assert np.median(r['peak'][k]) > 0.9


@pytest.mark.parametrize('method', ['correlate_fft', 'correlate_direct'])
def test_gross(method):
    shift = (30.4, -27.6)
    ref, sec = ref_frames([shift])
    cor = ref_oracle(search=6, skip=16, margin=32)
    cor.set_reference(ref)
    fn = getattr(cor, method)
    found = fn(sec, gross=np.array([[30, -28]]))['offsets'][0]
    assert np.median(np.abs(found - np.array(shift, np.float32))) < 0.05
    g = np.broadcast_to(np.array([30, -28], np.int32), (1,) + cor.shape + (2,))
    assert np.array_equal(fn(sec, gross=g)['offsets'][0], found, equal_nan=True)
    assert np.median(np.abs(fn(sec)['offsets'][0] - np.array(shift, np.float32))) > 1


@pytest.mark.parametrize('method', ['correlate_fft', 'correlate_direct'])
def test_invalid(method):
    ref, sec = ref_frames([(0, 0)])
    cor = ref_oracle()
    cor.set_reference(ref)
    r = getattr(cor, method)(sec, gross=np.array([[0, 1000]]))
    assert np.all(np.isnan(r['offsets'])) and np.all(r['snr'] == 0)
    assert np.all(r['cov'][..., :2] == 99) and np.all(r['peak'] == 0)
    flat = ref.copy()
    flat[:120] = 100
    cor.set_reference(flat)
    r = getattr(cor, method)(sec)
    assert np.all(np.isnan(r['offsets'][0, 0]))
    assert np.all(np.isfinite(r['offsets'][0, -1]))


def test_fft_and_direct_agree():
    """The FP32 FFT restatement and the FP64 direct oracle: the reference's cross-backend bar
    (tests/test_framecor.py:52-62) wherever the FP64 surface has no near-tie at its peak"""
    d = synthetic.small_shifts()
    down, across = window_grid(256, 320, d['window'], d['search'], d['skip'])
    cor = OracleCorrelator(256, 320, down, across, window=d['window'], search=d['search'])
    cor.set_reference(d['ref'])
    a = cor.correlate_fft(d['frames'])
    b = cor.correlate_direct(d['frames'])
    assert np.abs(a['offsets'] - b['offsets']).max() < 1e-3
    assert np.abs(a['peak'] - b['peak']).max() < 1e-4
    assert np.abs(a['snr'] / b['snr'] - 1).max() < 1e-4
    err = b['offsets'] - d['shifts'][:, None, None, :]
    assert np.median(np.abs(err)) < 0.05


def test_direct_surface_is_textbook_ncc():
    """The FP64 surface of the direct oracle is the textbook NCC of the two windows"""
    d = synthetic.small_shifts(n=1)
    W, R = 32, 16
    down, across = window_grid(256, 320, W, R, 32)
    cor = OracleCorrelator(256, 320, down, across, window=W, search=R)
    cor.set_reference(d['ref'])
    r = cor.correlate_direct(d['frames'], windows=[9], surfaces=True)
    s = r['surface'][0].reshape(-1, 2 * R + 1, 2 * R + 1)[9]
    y0, x0 = down[9 // len(across)], across[9 % len(across)]
    t = d['ref'][y0:y0 + W, x0:x0 + W].astype(np.float64)
    t -= t.mean()
    for u, v in [(0, 0), (16, 16), (5, 29), (32, 32)]:
        f = d['frames'][0, y0 - R + u:y0 - R + u + W, x0 - R + v:x0 - R + v + W].astype(np.float64)
        f -= f.mean()
        assert abs(s[u, v] - (t * f).sum() / np.sqrt((t * t).sum() * (f * f).sum())) < 1e-12


def test_golden_fixture_reproduced():
    g = np.load(GOLDEN)
    cor = OracleCorrelator(96, 128, g['down'], g['across'], window=int(g['window']), search=int(g['search']))
    cor.set_reference(g['ref'])
    for tag, gross in (('', None), ('gross_', g['gross'])):
        r = cor.correlate_direct(g['frames'], gross=gross)
        for key in ('offsets', 'peak', 'snr', 'cov'):
            assert np.array_equal(r[key], g[f'direct_{tag}{key}'], equal_nan=True), key
        r = cor.correlate_fft(g['frames'], gross=gross)
        for key in ('offsets', 'peak', 'snr', 'cov'):
            assert np.array_equal(r[key], g[f'fft_{tag}{key}'], equal_nan=True), key
    # the known shifts are recovered by the fixture's outputs
    err = g['direct_offsets'] - g['shifts'][:, None, None, :]
    assert np.median(np.abs(err)) < 0.1


# ---- the periodicity of slipcam-auto (python/slipcam/auto_slip.py:105-129)

def periodicity_scipy(ref, y0, x0, window, search, exclude=5):
    """The reference's periodicity in another formulation of the same mathematics: the correlation by
    scipy.signal.correlate (FFT), the box sums by cumulative sums, the variance clamped at 1e-6 as the
    reference does, local maxima by maximum_filter in 'nearest' mode (the same as the reference's
    'reflect' for a 5-point filter: both extend the surface with values inside it), the excluded square
    as a numpy slice"""
    ref = ref.astype(np.float64)
    n = window * window
    out = np.ones((len(y0), len(x0)))

    def box(c):
        return c[window:, window:] - c[:-window, window:] - c[window:, :-window] + c[:-window, :-window]

    for i, y in enumerate(y0):
        for j, x in enumerate(x0):
            t = ref[y:y + window, x:x + window]
            t = t - t.mean()
            tn = np.sqrt(np.sum(t * t))
            if tn < 1e-6:
                continue
            s = ref[y - search:y + window + search, x - search:x + window + search]
            num = signal.correlate(s, t, mode='valid', method='fft')
            c1 = np.pad(s.cumsum(0).cumsum(1), ((1, 0), (1, 0)))
            c2 = np.pad((s * s).cumsum(0).cumsum(1), ((1, 0), (1, 0)))
            var = np.maximum(box(c2) - box(c1) ** 2 / n, 1e-6)
            ncc = num / (np.sqrt(var) * tn)
            peak = ncc == ndimage.maximum_filter(ncc, size=5, mode='nearest')
            peak[search - exclude:search + exclude + 1, search - exclude:search + exclude + 1] = False
            out[i, j] = ncc[peak].max() if peak.any() else 0.0
    return out


def mixed_texture(h=240, w=320, seed=7):
    """Smoothed noise, a 12 x 12 tile repeated over the lower right quarter and a flat block"""
    img = synthetic.to_uint8(ref_texture(h, w, seed))
    tile = np.random.default_rng(seed).integers(0, 256, (12, 12), dtype=np.uint8)
    img[h // 2:, w // 2:] = np.tile(tile, (h // 24 + 1, w // 24 + 1))[:h - h // 2, :w - w // 2]
    img[:h // 3, :w // 4] = 77
    return img


@pytest.mark.parametrize('exclude', [5, 0, 16, 20])
def test_periodicity_matches_scipy(exclude):
    ref = mixed_texture()
    W, R = 32, 16
    down, across = window_grid(240, 320, W, R, 16)
    cor = OracleCorrelator(240, 320, down, across, window=W, search=R)
    cor.set_reference(ref)
    got = cor.periodicity(ref, exclude=exclude)
    want = periodicity_scipy(ref, down, across, W, R, exclude)
    assert got.shape == want.shape
    assert np.abs(got - want).max() < 1e-9


def test_periodicity_closed_forms():
    W, R = 32, 16
    tile = np.random.default_rng(1).integers(0, 256, (12, 12), dtype=np.uint8)
    periodic = np.tile(tile, (20, 27))[:240, :320]
    down, across = window_grid(240, 320, W, R, 24)
    cor = OracleCorrelator(240, 320, down, across, window=W, search=R)
    cor.set_reference(periodic)
    # the window repeats at a shift of 12: its second peak is the NCC of a window with itself
    assert np.abs(cor.periodicity(periodic) - 1.0).max() < 1e-12
    # an excluded square covering the surface leaves no local maximum
    assert np.array_equal(cor.periodicity(periodic, exclude=R), np.zeros(cor.shape))
    # flat windows are periodic
    flat = np.full((240, 320), 200, np.uint8)
    cor.set_reference(flat)
    assert np.array_equal(cor.periodicity(flat), np.ones(cor.shape))
    # white noise has no second peak near 1
    noise = np.random.default_rng(2).integers(0, 256, (240, 320), dtype=np.uint8)
    cor.set_reference(noise)
    p = cor.periodicity(noise)
    assert p.max() < 0.3 and p.min() > 0.0


# ---- the peak of slipcam-auto's coarse tracking (python/slipcam/stabilize.py:22-45 ncc_peak)

def ncc_peak_scipy(template, search):
    """ncc_peak in another formulation of the same mathematics: scipy.signal.correlate (FFT), box sums
    by cumulative sums, the 1e-6 variance clamp of the reference, np.argmax, the axis parabolas"""
    t = template.astype(np.float64)
    t = t - t.mean()
    tn = np.sqrt(np.sum(t * t))
    if tn < 1e-6:
        return np.nan, np.nan, 0.0
    s = search.astype(np.float64)
    W = t.shape[0]
    c1 = np.pad(s.cumsum(0).cumsum(1), ((1, 0), (1, 0)))
    c2 = np.pad((s * s).cumsum(0).cumsum(1), ((1, 0), (1, 0)))
    box = lambda c: c[W:, W:] - c[:-W, W:] - c[W:, :-W] + c[:-W, :-W]  # noqa: E731
    ncc = signal.correlate(s, t, mode='valid', method='fft') / (
        np.sqrt(np.maximum(box(c2) - box(c1) ** 2 / (W * W), 1e-6)) * tn)
    i, j = np.unravel_index(np.argmax(ncc), ncc.shape)
    M = ncc.shape[0]

    def para(c, m, p):
        return 0.0 if abs(m - 2 * c + p) < 1e-12 else 0.5 * (m - p) / (m - 2 * c + p)
    di = para(ncc[i, j], ncc[i - 1, j], ncc[i + 1, j]) if 0 < i < M - 1 else 0.0
    dj = para(ncc[i, j], ncc[i, j - 1], ncc[i, j + 1]) if 0 < j < M - 1 else 0.0
    return i + di - (M - 1) / 2, j + dj - (M - 1) / 2, ncc[i, j]


def test_ncc_peaks_match_scipy():
    # no periodic part: a repeated texture has exact ties between equal peaks, which the exact oracle
    # breaks by the first index and FFT rounding breaks either way
    ref = synthetic.to_uint8(ref_texture(240, 320, seed=9))
    ref[:80, :80] = 77
    frame = synthetic.to_uint8(ndimage.shift(ref.astype(np.float64), (2.3, -1.6), order=3, mode='nearest'))
    W, R = 32, 16
    down, across = window_grid(240, 320, W, R, 16)
    cor = OracleCorrelator(240, 320, down, across, window=W, search=R)
    cor.set_reference(ref)
    gross = np.random.default_rng(3).integers(-25, 26, (len(down), len(across), 2))
    off, peak = cor.ncc_peaks(frame, gross)
    nan = 0
    for i, y in enumerate(down):
        for j, x in enumerate(across):
            sy, sx = y + gross[i, j, 0] - R, x + gross[i, j, 1] - R
            if sy < 0 or sx < 0 or sy + W + 2 * R > 240 or sx + W + 2 * R > 320:
                assert np.isnan(off[i, j]).all() and peak[i, j] == 0.0
                nan += 1
                continue
            dy, dx, pk = ncc_peak_scipy(ref[y:y + W, x:x + W], frame[sy:sy + W + 2 * R, sx:sx + W + 2 * R])
            if np.isnan(dy):
                assert np.isnan(off[i, j]).all() and peak[i, j] == 0.0
                continue
            assert abs(off[i, j, 0] - dy) < 1e-7 and abs(off[i, j, 1] - dx) < 1e-7, (i, j)
            assert abs(peak[i, j] - pk) < 1e-9
    assert 0 < nan < off.shape[0] * off.shape[1]


def test_ncc_peaks_known_shift():
    """A frame that is the reference moved by whole pixels: the peak is 1 at the shift, and the
    parabolas through a symmetric peak add nothing"""
    ref = synthetic.to_uint8(ref_texture(240, 320, seed=3))
    frame = np.roll(ref, (5, -7), axis=(0, 1))
    W, R = 32, 16
    down, across = window_grid(240, 320, W, R, 40)
    cor = OracleCorrelator(240, 320, down, across, window=W, search=R)
    cor.set_reference(ref)
    off, peak = cor.ncc_peaks(frame)
    assert np.abs(peak - 1.0).max() < 1e-12
    assert np.abs(np.rint(off) - np.array([5, -7])).max() == 0
    assert np.abs(off - np.array([5, -7])).max() < 0.25
