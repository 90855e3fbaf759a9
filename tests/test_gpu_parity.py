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
from parity import assert_fft_parity, assert_int8_parity, flat_select, run
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


# ---- the README-shaped workload (720 x 1280, window 64, search 48, skip 16: 2450 windows): frames of the
# 64-frame fault video before, during and after the 40 px slip ramp of frames 20-40 (centre 30)

FAULT_FRAMES = [0, 22, 28, 30, 34, 40, 63]


def fault_line_windows(down, across, W, h=720, w=1280, angle=30.0, rows=2, skip=16):
    """The flat indices of the windows whose centre lies within `rows` grid rows of the fault line
    (the straight line at `angle` degrees through the image centre of synthetic.fault_video)"""
    a = np.deg2rad(angle)
    dist = (down[:, None] + W / 2 - h / 2) * np.cos(a) + (across[None, :] + W / 2 - w / 2) * np.sin(a)
    return np.flatnonzero(np.abs(dist).ravel() <= rows * skip)


@pytest.fixture(scope='module')
def fault():
    d = synthetic.fault_video(n=64)
    d['frames'] = np.ascontiguousarray(d['frames'][FAULT_FRAMES])
    down, across = framecor.window_grid(720, 1280, d['window'], d['search'], d['skip'])
    o = OracleCorrelator(720, 1280, down, across, window=d['window'], search=d['search'])
    o.set_reference(d['ref'])
    return d, down, across, o


def test_fault_int8_vs_fp64_oracle(gpu, fault):
    """Every window within two grid rows of the fault line (cut by it, or next to it: offsets of the two
    sides) and 96 others, on frames before, during and after the slip: the int8 form equals the FP64
    direct-NCC oracle (same validity and integer peaks, offsets to FP32 rounding)"""
    d, down, across, o = fault
    got = run('int8', 720, 1280, down, across, 64, 48, d['ref'], d['frames'])
    near = fault_line_windows(down, across, 64)
    others = np.setdiff1d(np.arange(len(down) * len(across)), near)
    sample = np.union1d(near, np.random.default_rng(5).choice(others, 96, replace=False))
    assert len(near) > 200
    ref = o.correlate_direct(d['frames'], windows=sample)
    assert_int8_parity(flat_select(got, sample), flat_select(ref, sample))
    # the slip is seen: after the ramp, windows on the far side move by about 40 px more than before
    off = flat_select(got, sample)['offsets']
    side = np.asarray([d['side'][int(down[k // len(across)] + 32), int(across[k % len(across)] + 32)] for k in sample])
    moved = np.nanmedian(np.linalg.norm(off[-1] - off[0], axis=-1)[side])
    assert 30 < moved < 50, moved


def test_fault_fft_vs_reference_algorithm(gpu, fault):
    """Every window of the slip frames: the FFT form against the restatement of the reference algorithm;
    every window whose offsets differ by more than 1e-3 px is checked on its exact FP64 surface to be a
    near-tie of the two peaks"""
    d, down, across, o = fault
    got = run('fft', 720, 1280, down, across, 64, 48, d['ref'], d['frames'])
    ref = o.correlate_fft(d['frames'])
    valid = np.isfinite(ref['offsets'][..., 0])
    differ = valid & np.any(np.abs(got['offsets'] - ref['offsets']) > 1e-3, axis=-1)
    ny, nx = len(down), len(across)
    wins = np.unique(np.nonzero(differ)[1] * nx + np.nonzero(differ)[2])
    surf = o.correlate_direct(d['frames'], windows=wins, surfaces=True)['surface'] if len(wins) else None
    print(f'{int(differ.sum())} of {int(valid.sum())} window-frames differ by more than 1e-3 px')
    assert assert_fft_parity(got, ref, 48, surf=surf) == int(differ.sum())
    same = valid & ~differ
    assert np.median(np.abs(got['offsets'] - ref['offsets'])[same]) < 1e-5


@pytest.mark.parametrize('variant', ['simple', 'padded', 'piped', 'wide', 'quad'])
def test_edge_peaks_w64_r48(gpu, variant):
    """Window 64, search 48 (M = 97: the edge-row and edge-column waves of the 'simple' kernel hold
    surface row and column 96): gross offsets put the true integer peak of each window on u or v in
    {0, 1, 48, 95, 96} and past the border (-1, 97, where the maximum is clipped to the border), so the
    3 x 3 fit falls back and the 21 x 21 SNR box is clipped; every kernel variant against the FP64
    direct oracle"""
    h = w = 512
    W, R = 64, 48
    down = across = np.array([112, 208, 304])
    base = synthetic.texture(h, w, sigma=1.5, seed=21)
    rng = np.random.default_rng(21)
    moves = rng.integers(-6, 7, size=(4, 2))
    frames = np.stack([synthetic.to_uint8(synthetic.warp(base, *m) + rng.normal(0, 2, base.shape)) for m in moves])
    targets = [-1, 0, 1, 48, 95, 96, 97]
    # peak at u = true move - gross + R: gross = move + R - target
    tu = rng.choice(targets, size=(4, 3, 3))
    tv = rng.choice(targets, size=(4, 3, 3))
    tu[0], tv[0] = [[0, 96, 0], [96, 48, 1], [95, 97, -1]], [[0, 0, 96], [96, 96, 48], [-1, 95, 97]]
    gross = np.stack([moves[:, None, None, 0] + R - tu, moves[:, None, None, 1] + R - tv], -1).astype(np.int32)
    ref8 = synthetic.to_uint8(base)
    cor = framecor.Correlator(h, w, down, across, window=W, search=R, method='int8')
    cor.set_variant(variant)
    cor.set_reference(ref8)
    got = cor.correlate(frames, gross=gross)
    o = OracleCorrelator(h, w, down, across, window=W, search=R)
    o.set_reference(ref8)
    want = o.correlate_direct(frames, gross=gross)
    assert_int8_parity(got, want)
    # the peaks are where they were put (clipped to the border), in every window
    u = np.rint(got['offsets'][..., 0] - gross[..., 0] + R)
    v = np.rint(got['offsets'][..., 1] - gross[..., 1] + R)
    assert np.array_equal(u, np.clip(tu, 0, 2 * R)) and np.array_equal(v, np.clip(tv, 0, 2 * R))
    border = (np.clip(tu, 0, 96) % 96 == 0) | (np.clip(tv, 0, 96) % 96 == 0)
    assert np.all(got['cov'][border] == np.array([99, 99, 0], np.float32))


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
    the same with full 16-row bands, the pipelined producer/helper kernel, the 32 x 32 kernel (M =
    32 c + 1), the four-wave kernel and the one-wave kernel (search up to 23): the same offsets, peaks and
    covariances, the SNR to 1e-12 (the pipelined kernel splits its sums over the helper waves); the
    32 x 32, four-wave and one-wave kernels give the same bits as the 16 x 16 one throughout"""
    if workload == 'small':
        d, down, across, _, gross = small
        h, w, W, R, frames, g = 256, 320, d['window'], d['search'], d['frames'], gross
    else:
        d, down, across, _ = fault
        h, w, W, R, frames, g = 720, 1280, 64, 48, d['frames'], None
    out = {}
    for variant in ('simple', 'piped', 'padded', 'wide', 'quad', 'solo'):
        cor = framecor.Correlator(h, w, down, across, window=W, search=R, method='int8')
        cor.set_variant(variant)
        cor.set_reference(d['ref'])
        out[variant] = cor.correlate(frames, gross=g)
    a = out['simple']
    for other in ('wide', 'quad', 'solo'):
        for key in a:
            assert np.array_equal(a[key], out[other][key], equal_nan=True), (other, key)
    ok = a['snr'] > 0
    for other in ('padded', 'piped'):   # the same surface, its SNR sums split over other waves
        b = out[other]
        for key in ('offsets', 'peak', 'cov'):
            assert np.array_equal(a[key], b[key], equal_nan=True), (other, key)
        assert np.array_equal(ok, b['snr'] > 0), other
        assert np.abs(a['snr'][ok] / b['snr'][ok] - 1).max() < 1e-12, other


@pytest.mark.parametrize('n', [37, 48])
def test_int8_runs_of_frames_agree(gpu, n):
    """A call of more than 8 frames is dealt to the blocks in runs of at most 8 frames (37: runs of 8, 8,
    8, 8 and 5; 48: six of 8), run of frames major: the same bits as frame by frame"""
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


@pytest.mark.parametrize('method', ['int8', 'fft'])
def test_set_reference_waits_for_device_calls(gpu, fault, method):
    """Device calls queued on another stream, then set_reference of another frame at once: the queued
    calls still see the first reference (the correlator orders its own stream after them)"""
    torch = _torch()
    d, down, across, _ = fault
    cor = framecor.Correlator(720, 1280, down, across, window=64, search=48, method=method)
    cor.set_reference(d['ref'])
    dev = torch.from_numpy(d['frames']).cuda()
    expect = {k: v.cpu().numpy() for k, v in cor.correlate(dev).items()}
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        outs = [cor.correlate(dev) for _ in range(8)]
        cor.set_reference(np.roll(d['ref'], 7, axis=1))
    torch.cuda.synchronize()
    for out in outs:
        for k in expect:
            assert np.array_equal(out[k].cpu().numpy(), expect[k], equal_nan=True), k
    moved = cor.correlate(dev)
    assert not np.array_equal(moved['offsets'].cpu().numpy(), expect['offsets'], equal_nan=True)
