"""
Writes tests/golden/tiny.npz: a small seeded workload and the outputs of the two CPU oracles on it.

  python tests/golden/make_fixtures.py

Inputs: a 96 x 128 smoothed-noise reference (sigma 1.5), 4 frames shifted by known sub-pixel
offsets with noise sigma 2 DN, window 16, search 8, skip 16, and per-window integer gross offsets
in -2..2 (a few push search windows out of the frame). Outputs: the FP64 direct-NCC oracle
('direct_*') and the restatement of the reference CPU algorithm with FP32 FFT ('fft_*'), with and
without the gross offsets. The fixture pins the oracle against regressions and lets the GPU tests
compare against fixed numbers.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import OracleCorrelator  # noqa: E402
from slipcam_amd import synthetic  # noqa: E402


def window_grid(height, width, window, search, skip, margin=0):
    axes = []
    for size in (height, width):
        first = margin + search
        count = (size - margin - search - first - window) // skip
        axes.append(first + skip * np.arange(count))
    return tuple(axes)


def make():
    h, w, W, R, skip = 96, 128, 16, 8, 16
    ref = synthetic.to_uint8(synthetic.texture(h, w, seed=10))
    shifts = np.array([[0.0, 0.0], [0.3, -0.7], [1.25, 2.5], [-2.4, 1.1]])
    noise = np.random.default_rng(11)
    base = ref.astype(np.float64)
    frames = np.stack([synthetic.to_uint8(synthetic.warp(base, s[0], s[1]) + noise.normal(0, 2.0, base.shape))
                       for s in shifts])
    down, across = window_grid(h, w, W, R, skip)
    gross = np.random.default_rng(12).integers(-2, 3, size=(len(frames), len(down), len(across), 2)).astype(np.int32)
    o = OracleCorrelator(h, w, down, across, window=W, search=R, stat_size=21)
    o.set_reference(ref)
    out = dict(ref=ref, frames=frames, shifts=shifts, down=down.astype(np.int32), across=across.astype(np.int32),
               gross=gross, window=W, search=R)
    for tag, g in (('', None), ('gross_', gross)):
        for name, fn in (('direct', o.correlate_direct), ('fft', o.correlate_fft)):
            r = fn(frames, gross=g)
            for key in ('offsets', 'peak', 'snr', 'cov'):
                out[f'{name}_{tag}{key}'] = r[key]
    np.savez_compressed(os.path.join(HERE, 'tiny.npz'), **out)


if __name__ == '__main__':
    make()
