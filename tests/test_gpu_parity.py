"""
Parity of the HIP path with the CPU oracle, through the C ABI.

  summed-area tables          bit-exact against numpy cumsum (int64)
  int8 form                   against the FP64 direct-NCC oracle: the same integer peak everywhere
                              (the int8 surface is exact, and its argmax is taken on exact FP64 values),
                              offsets, peak and covariance to FP32 rounding, SNR to 1e-9 relative
  FFT form                    against the restatement of the reference CPU algorithm (FP32 FFT):
                              offsets within 1e-3 px, peak 1e-4, SNR 1e-4 relative (the reference's
                              cross-backend bar, tests/test_framecor.py:52-62, and 1e-4 on SNR), except
                              where the integer peak differs at a near-tie of the exact FP64 surface
                              (a top-2 gap below 1e-5), which the FP32 noise of two FFTs may flip
  both forms, full size       size-independent properties: results independent of how frames are
                              grouped; device-resident calls equal host calls
"""

import os

import numpy as np
import pytest

from oracle import OracleCorrelator
from slipcam_amd import framecor, synthetic

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'tiny.npz')


def _torch():
    import torch
    return torch


# ---- summed-area tables

@pytest.mark.parametrize('n,h,w', [(3, 1, 1), (2, 7, 3), (8, 256, 320), (4, 720, 1280), (1, 1080, 1920)])
def test_summed_area_bit_exact(gpu, n, h, w):
    torch = _torch()
    rng = np.random.default_rng(n * h + w)
    frames = rng.integers(0, 256, size=(n, h, w)).astype(np.uint8)
    frames[0, : min(h, 5)] = 255
    d = torch.from_numpy(frames).cuda()
    sat = torch.empty((n, h + 1, w + 1), dtype=torch.int64, device='cuda')
    sat2 = torch.empty_like(sat)
    assert framecor._lib.framecor_summed_area_device(d.data_ptr(), n, h, w, sat.data_ptr(), sat2.data_ptr(), None) == 0
    torch.cuda.synchronize()
    ref = np.zeros((n, h + 1, w + 1), np.int64)
    ref[:, 1:, 1:] = frames.astype(np.int64).cumsum(1).cumsum(2)
    ref2 = np.zeros_like(ref)
    ref2[:, 1:, 1:] = (frames.astype(np.int64) ** 2).cumsum(1).cumsum(2)
    assert np.array_equal(sat.cpu().numpy(), ref)
    assert np.array_equal(sat2.cpu().numpy(), ref2)


@pytest.mark.parametrize('n,h,w,W', [(2, 5, 7, 1), (3, 40, 33, 33), (2, 256, 320, 32), (3, 720, 1280, 64),
                                     (1, 300, 1100, 181), (2, 1080, 1920, 128), (1, 181, 600, 181)])
def test_box_sums_bit_exact(gpu, n, h, w, W):
    """The W x W box sums straight from the pixels equal the four-term differences of the int64
    summed-area tables (cxx/surface.h:21-25), bit for bit; all-255 rows reach the largest sums"""
    torch = _torch()
    rng = np.random.default_rng(n * h + w + W)
    frames = rng.integers(0, 256, size=(n, h, w)).astype(np.uint8)
    frames[0, : min(h, W + 3)] = 255
    d = torch.from_numpy(frames).cuda()
    bh, bw = h - W + 1, w - W + 1
    box = torch.empty((n, bh, bw), dtype=torch.int32, device='cuda')
    box2 = torch.empty_like(box)
    assert framecor._lib.framecor_box_sums_device(d.data_ptr(), n, h, w, W, box.data_ptr(), box2.data_ptr(), None) == 0
    torch.cuda.synchronize()
    for p, got in ((1, box), (2, box2)):
        sat = np.zeros((n, h + 1, w + 1), np.int64)
        sat[:, 1:, 1:] = (frames.astype(np.int64) ** p).cumsum(1).cumsum(2)
        ref = sat[:, W:, W:] - sat[:, :-W, W:] - sat[:, W:, :-W] + sat[:, :-W, :-W]
        assert np.array_equal(got.cpu().numpy().astype(np.int64), ref), p


# ---- helpers

def run(method, h, w, down, across, W, R, ref, frames, gross=None):
    cor = framecor.Correlator(h, w, down, across, window=W, search=R, method=method)
    cor.set_reference(ref)
    return cor.correlate(frames, gross=gross)


def near_peak(out, R, gross=None):
    """The integer peak of each window to within one pixel: round(offset - gross + R) (|du| <= 1)"""
    off = out['offsets'].astype(np.float64)
    g = 0 if gross is None else np.asarray(gross, np.float64)
    return np.rint(off - g + R)


def local_max(s, p):
    """The maximum of the surface s over the 3 x 3 neighbourhood of p (clipped)"""
    u, v = int(p[0]), int(p[1])
    return s[max(u - 1, 0):u + 2, max(v - 1, 0):v + 2].max()


def assert_int8_parity(got, ref):
    """int8 form vs the FP64 direct oracle"""
    valid = np.isfinite(ref['offsets'][..., 0])
    assert np.array_equal(np.isfinite(got['offsets'][..., 0]), valid)
    assert np.array_equal(got['snr'] == 0, ref['snr'] == 0)
    d = np.abs(got['offsets'] - ref['offsets'])[valid]
    assert d.max() <= 2e-6, d.max()
    assert np.abs(got['peak'] - ref['peak']).max() <= 1e-7
    rel = np.abs(got['snr'][valid] / ref['snr'][valid] - 1)
    assert rel.max() <= 1e-6, rel.max()
    cv = np.abs(got['cov'] - ref['cov']) / np.maximum(np.abs(ref['cov']), 1e-30)
    assert cv.max() <= 1e-5, cv.max()
    return float((got['offsets'][valid] == ref['offsets'][valid]).mean())


def assert_fft_parity(got, ref, R, surf=None, gross=None, tie=1e-5):
    """FFT form vs the FP32 FFT restatement; where the offsets differ by more than 1e-3 the two forms
    must have picked peaks whose exact FP64 values are a near-tie (surf: n x ny x nx x M x M, or None
    to require no such window)"""
    valid = np.isfinite(ref['offsets'][..., 0])
    assert np.array_equal(np.isfinite(got['offsets'][..., 0]), valid)
    differ = valid & np.any(np.abs(got['offsets'] - ref['offsets']) > 1e-3, axis=-1)
    pg, pr = near_peak(got, R, gross), near_peak(ref, R, gross)
    for idx in zip(*np.nonzero(differ)):
        assert surf is not None, f'the offsets differ at {idx}'
        a, b = local_max(surf[idx], pg[idx]), local_max(surf[idx], pr[idx])
        assert abs(a - b) < tie, (idx, a, b)
    same = valid & ~differ
    assert np.abs(got['peak'] - ref['peak'])[same].max() < 1e-4
    assert np.abs(got['snr'][same] / ref['snr'][same] - 1).max() < 1e-4
    return int(differ.sum())


# ---- golden fixture

def test_golden_fixture(gpu):
    g = np.load(GOLDEN)
    args = (96, 128, g['down'], g['across'], int(g['window']), int(g['search']), g['ref'], g['frames'])
    for tag, gross in (('', None), ('gross_', g['gross'])):
        ref_direct = {k: g[f'direct_{tag}{k}'] for k in ('offsets', 'peak', 'snr', 'cov')}
        ref_fft = {k: g[f'fft_{tag}{k}'] for k in ('offsets', 'peak', 'snr', 'cov')}
        assert_int8_parity(run('int8', *args, gross=gross), ref_direct)
        assert_fft_parity(run('fft', *args, gross=gross), ref_fft, int(g['search']), gross=gross)


# ---- the small workload (256 x 320, 8 frames, window 32, search 16), with and without gross offsets

@pytest.fixture(scope='module')
def small():
    d = synthetic.small_shifts()
    down, across = framecor.window_grid(256, 320, d['window'], d['search'], d['skip'])
    o = OracleCorrelator(256, 320, down, across, window=d['window'], search=d['search'])
    o.set_reference(d['ref'])
    gross = synthetic.gross_offsets(len(d['frames']), len(down), len(across))
    return d, down, across, o, gross


@pytest.mark.parametrize('with_gross', [False, True])
def test_small_int8_vs_fp64_oracle(gpu, small, with_gross):
    d, down, across, o, gross = small
    g = gross if with_gross else None
    got = run('int8', 256, 320, down, across, d['window'], d['search'], d['ref'], d['frames'], gross=g)
    ref = o.correlate_direct(d['frames'], gross=g)
    exact = assert_int8_parity(got, ref)
    assert exact > 0.5
    if not with_gross:
        err = got['offsets'] - d['shifts'][:, None, None, :]
        assert np.median(np.abs(err)) < 0.05


@pytest.mark.parametrize('with_gross', [False, True])
def test_small_fft_vs_reference_algorithm(gpu, small, with_gross):
    d, down, across, o, gross = small
    g = gross if with_gross else None
    got = run('fft', 256, 320, down, across, d['window'], d['search'], d['ref'], d['frames'], gross=g)
    ref = o.correlate_fft(d['frames'], gross=g)
    surf = o.correlate_direct(d['frames'], gross=g, surfaces=True)['surface']
    assert assert_fft_parity(got, ref, d['search'], surf=surf, gross=g) <= 2


# ---- the README-shaped workload (720 x 1280, window 64, search 48, skip 16: 2450 windows)

@pytest.fixture(scope='module')
def fault():
    d = synthetic.fault_video(n=6)
    down, across = framecor.window_grid(720, 1280, d['window'], d['search'], d['skip'])
    o = OracleCorrelator(720, 1280, down, across, window=d['window'], search=d['search'])
    o.set_reference(d['ref'])
    return d, down, across, o


def test_fault_int8_vs_fp64_oracle(gpu, fault):
    d, down, across, o = fault
    got = run('int8', 720, 1280, down, across, 64, 48, d['ref'], d['frames'])
    sample = np.random.default_rng(5).choice(len(down) * len(across), 48, replace=False)
    ref = o.correlate_direct(d['frames'], windows=sample)
    sel = lambda r: {k: r[k].reshape((r[k].shape[0], -1) + r[k].shape[3:])[:, sample] for k in r}  # noqa: E731
    assert_int8_parity(sel(got), sel(ref))


def test_fault_fft_vs_reference_algorithm(gpu, fault):
    d, down, across, o = fault
    got = run('fft', 720, 1280, down, across, 64, 48, d['ref'], d['frames'])
    ref = o.correlate_fft(d['frames'])
    valid = np.isfinite(ref['offsets'][..., 0])
    differ = valid & np.any(np.abs(got['offsets'] - ref['offsets']) > 1e-3, axis=-1)
    # peak flips of two FP32 FFTs: rare, and only at near-ties (checked window by window on the
    # small workload against the FP64 surfaces)
    assert differ.mean() < 2e-3, differ.mean()
    same = valid & ~differ
    assert np.median(np.abs(got['offsets'] - ref['offsets'])[same]) < 1e-5
    assert np.abs(got['peak'] - ref['peak'])[same].max() < 1e-4
    assert np.abs(got['snr'][same] / ref['snr'][same] - 1).max() < 1e-4


@pytest.mark.parametrize('method', ['fft', 'int8'])
def test_grouping_and_device_calls_agree(gpu, fault, method):
    """Frame by frame, all at once, and device-resident: the same bits"""
    torch = _torch()
    d, down, across, _ = fault
    cor = framecor.Correlator(720, 1280, down, across, window=64, search=48, method=method)
    cor.set_reference(d['ref'])
    whole = cor.correlate(d['frames'])
    one = [cor.correlate(d['frames'][f:f + 1]) for f in range(len(d['frames']))]
    dev = cor.correlate(torch.from_numpy(d['frames']).cuda())
    torch.cuda.synchronize()
    for key in whole:
        assert np.array_equal(whole[key], np.concatenate([o[key] for o in one]), equal_nan=True), key
        assert np.array_equal(whole[key], dev[key].cpu().numpy(), equal_nan=True), key


@pytest.mark.parametrize('workload', ['small', 'fault'])
def test_int8_kernel_variants_agree(gpu, small, fault, workload):
    """The one-frame-at-a-time kernel with the edge row and column waves (search 48 and 16: M = 16 c + 1),
    the same with full 16-row bands, and the pipelined producer/helper kernel: the same offsets, peaks and
    covariances, the SNR to 1e-12 (its sums are split over the helper waves)"""
    if workload == 'small':
        d, down, across, _, gross = small
        h, w, W, R, frames, g = 256, 320, d['window'], d['search'], d['frames'], gross
    else:
        d, down, across, _ = fault
        h, w, W, R, frames, g = 720, 1280, 64, 48, d['frames'], None
    out = {}
    for variant in ('simple', 'piped', 'padded'):
        cor = framecor.Correlator(h, w, down, across, window=W, search=R, method='int8')
        cor.set_variant(variant)
        cor.set_reference(d['ref'])
        out[variant] = cor.correlate(frames, gross=g)
    a = out['simple']
    ok = a['snr'] > 0
    for other in ('padded', 'piped'):   # the same surface, its SNR sums split over other waves
        b = out[other]
        for key in ('offsets', 'peak', 'cov'):
            assert np.array_equal(a[key], b[key], equal_nan=True), (other, key)
        assert np.array_equal(ok, b['snr'] > 0), other
        assert np.abs(a['snr'][ok] / b['snr'][ok] - 1).max() < 1e-12, other


@pytest.mark.parametrize('n', [37, 48])
def test_int8_runs_of_frames_agree(gpu, n):
    """A call of more than 16 frames is dealt to the blocks in runs of at most 16 frames (37: runs of
    13, 13 and 11; 48: three of 16), run of frames major: the same bits as frame by frame"""
    h, w, W, R = 480, 640, 32, 16
    base = synthetic.texture(h, w, sigma=1.5, seed=n)
    rng = np.random.default_rng(n)
    frames = np.stack([synthetic.to_uint8(synthetic.warp(base, *rng.uniform(-3, 3, 2)) + rng.normal(0, 2, base.shape))
                       for _ in range(n)])
    down, across = framecor.window_grid(h, w, W, R, 16)
    cor = framecor.Correlator(h, w, down, across, window=W, search=R, method='int8')
    cor.set_reference(synthetic.to_uint8(base))
    whole = cor.correlate(frames)
    one = [cor.correlate(frames[f:f + 1]) for f in range(n)]
    assert np.isfinite(whole['offsets'][..., 0]).mean() > 0.9
    for key in whole:
        assert np.array_equal(whole[key], np.concatenate([o[key] for o in one]), equal_nan=True), key
