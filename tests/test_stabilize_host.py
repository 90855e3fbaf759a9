"""
The host side of the coarse tracking (slipcam_amd/stabilize.py, the drop-in for python/slipcam/
stabilize.py:48-96 and python/slipcam/slip_series.py:49-72): the robust similarity fit against exact
transforms, outliers and too few points; the grid; and the frame loop driven by the oracle's peaks
recovering the known camera motion of a synthetic video.
"""

import numpy as np

from oracle import OracleCorrelator
from slipcam_amd import stabilize, synthetic


def similarity_points(params, n=30, seed=0):
    rng = np.random.default_rng(seed)
    x, y = rng.uniform(-300, 300, n), rng.uniform(-200, 200, n)
    x2, y2 = stabilize.apply(params, x, y)
    return x2 - x, y2 - y, x, y


def test_apply_is_the_similarity():
    p, q, a, s = 3.0, -2.0, 0.01, 0.02
    x, y = np.array([10.0, -5.0]), np.array([4.0, 7.0])
    x2, y2 = stabilize.apply((p, q, a, s), x, y)
    rot = np.array([[1 + s, -a], [a, 1 + s]])
    assert np.allclose(np.stack([x2, y2]), rot @ np.stack([x, y]) + np.array([[p], [q]]), atol=1e-12)


def test_fit_camera_recovers_an_exact_transform():
    truth = np.array([2.5, -1.25, 1.5e-3, -0.8e-3])
    dx, dy, x, y = similarity_points(truth)
    fit, used = stabilize.fit_camera(dx, dy, x, y)
    assert np.abs(fit - truth).max() < 1e-9 and used == 30


def test_fit_camera_drops_outliers_and_gaps():
    truth = np.array([-4.0, 3.0, -1e-3, 2e-3])
    dx, dy, x, y = similarity_points(truth, n=40, seed=1)
    dx = dx + np.random.default_rng(2).normal(0, 0.05, 40)
    dx[[3, 17]] += 6.0
    dy[25] -= 9.0
    dx[30] = np.nan
    fit, used = stabilize.fit_camera(dx, dy, x, y)
    assert used == 36
    assert np.abs(fit[:2] - truth[:2]).max() < 0.05 and np.abs(fit[2:] - truth[2:]).max() < 2e-4


def test_fit_camera_shift_model_and_too_few_points():
    dx, dy, x, y = similarity_points(np.array([1.0, 2.0, 3e-3, 0.0]), n=12)
    fit, _ = stabilize.fit_camera(dx, dy, x, y, model='shift')
    assert fit[2] == 0.0 and fit[3] == 0.0 and abs(fit[0] - (dx.mean())) < 1e-9
    dx[3:] = np.nan                       # 3 points: 6 residuals, fewer than 8
    fit, used = stabilize.fit_camera(dx, dy, x, y)
    assert np.isnan(fit).all() and used == 0


def test_grid_skips_text_and_flat_windows():
    ref = synthetic.to_uint8(synthetic.texture(480, 640, seed=1))
    ref[200:320, 300:420] = 90
    text = np.zeros((480, 640), bool)
    text[50:60, 60:70] = True
    g = stabilize.Grid(ref, text, 96, 96, 32)
    assert list(g.yc) == [80, 176, 272, 368] and list(g.xc) == [80, 176, 272, 368, 464]
    keep = g.keep.reshape(4, 5)
    assert not keep[0, 0]                 # the text
    assert not keep[2, 3]                 # the flat block (centre 272, 368)
    assert keep.sum() == 18


def test_tracking_loop_recovers_the_camera_motion():
    v = synthetic.similarity_video(n=6, ref=2)
    frames, ref = v['frames'], v['ref']
    g = stabilize.Grid(frames[ref], v['text'], 96, 96, 32)
    ora = OracleCorrelator(480, 640, g.yc - g.half, g.xc - g.half, window=96, search=32)
    ora.set_reference(frames[ref])
    params, used = stabilize.track_frames(frames, ref, g, ora.ncc_peaks)
    assert np.array_equal(params[ref], np.zeros(4)) and used[ref] == len(g.Xg)
    err = params - v['params']
    assert np.abs(err[:, :2]).max() < 0.1 and np.abs(err[:, 2:]).max() < 3e-4
    assert (used >= 10).all()
