"""
The track driver (slipcam_amd/track.py, the drop-in for the reference's slipcam-track,
python/slipcam/track_offsets.py): chunked, pipelined correlation of a memory-mapped video gives the
same bits as one correlate call over the same frames, with the reference's output keys; the command
line writes the .npz the reference writes.
"""

import os
import subprocess
import sys

import numpy as np
import pytest

from slipcam_amd import framecor, synthetic, track

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KEYS = ('offsets', 'snr', 'cov', 'peak')


@pytest.fixture(scope='module')
def video(tmp_path_factory):
    d = synthetic.fault_video(n=11)
    path = tmp_path_factory.mktemp('video') / 'frames.npy'
    np.save(path, d['frames'])
    return d, path


@pytest.mark.parametrize('method', ['fft', 'int8'])
def test_chunks_equal_one_call(gpu, video, method):
    d, path = video
    frames = np.load(path, mmap_mode='r')
    index = np.arange(2, 11)
    z = track.track(frames, 1, index, window=64, skip=16, search=48, method=method, chunk=4)
    down, across = framecor.window_grid(720, 1280, 64, 48, 16)
    cor = framecor.Correlator(720, 1280, down, across, window=64, search=48, method=method)
    cor.set_reference(d['frames'][1])
    r = cor.correlate(d['frames'][2:11])
    for key in KEYS:
        assert z[key].dtype == np.float32
        assert np.array_equal(z[key], r[key], equal_nan=True), key
    assert z['engine'] == 'framecor' and z['ref'] == 1 and np.array_equal(z['index'], index)
    assert np.array_equal(z['yc'], down + 32) and np.array_equal(z['xc'], across + 32)


def test_gathered_frames(gpu, video):
    """An index that is not a run of consecutive frames (gathered from the memmap)"""
    d, path = video
    frames = np.load(path, mmap_mode='r')
    index = np.array([10, 3, 4, 0, 7])
    z = track.track(frames, 5, index, method='int8', chunk=2)
    down, across = framecor.window_grid(720, 1280, 64, 48, 16)
    cor = framecor.Correlator(720, 1280, down, across, window=64, search=48, method='int8')
    cor.set_reference(d['frames'][5])
    r = cor.correlate(d['frames'][index])
    for key in KEYS:
        assert np.array_equal(z[key], r[key], equal_nan=True), key


def test_command_line(gpu, video, tmp_path):
    d, path = video
    out = tmp_path / 'offsets.npz'
    cmd = [sys.executable, '-m', 'slipcam_amd.track', str(path), '--ref', '0', '--first', '6', '--out', str(out),
           '--mask', '870', '25', '1240', '68', '--mask', '40', '650', '630', '700', '--method', 'int8']
    subprocess.run(cmd, check=True, cwd=ROOT, timeout=300)
    z = np.load(out)
    assert sorted(z.files) == sorted(['index', 'ref', 'masked', 'yc', 'xc', 'window', 'skip', 'search', 'engine',
                                      'offsets', 'snr', 'cov', 'peak'])
    assert str(z['engine']) == 'framecor' and int(z['ref']) == 0
    assert np.array_equal(z['index'], np.arange(6, 11))
    assert z['offsets'].shape == (5, 35, 70, 2)
    down, across = framecor.window_grid(720, 1280, 64, 48, 16)
    assert np.array_equal(z['masked'], track.boxes_mask(down, across, 64, [(870, 25, 1240, 68), (40, 650, 630, 700)]))
    assert z['masked'].sum() > 0
