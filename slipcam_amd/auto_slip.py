"""
The periodicity step of slipcam-auto on the GPU (python/slipcam/auto_slip.py:105-129 `periodicity`).

  per = periodicity(frames[ref], y0, x0, window, search)        # as auto_slip.py:308
  periodic = per > max_periodicity

For each window of the grid (y0 rows x x0 columns of top-left corners) the NCC of the reference
window with its own surroundings (the (window + 2 search)^2 region around it) is computed exactly
(integer sums on the int8 matrix cores, FP64 normalization), and the result is the largest value of
that surface at a local maximum (a point equal to the largest value of its 5 x 5 neighbourhood, as
scipy.ndimage.maximum_filter with size 5) outside the square of rows and columns
search - exclude .. search + exclude around the zero shift: 1 for a periodic or flat texture, small
for a unique one; 0 where no local maximum lies outside the square; 1 for a flat window.

The reference evaluates the same NCC with FP64 FFTs (scipy.signal.fftconvolve) and clamps the
variance of the search box at 1e-6; the exact values differ from it by FFT rounding (about 1e-12),
and a flat search box gives 0 here where the reference divides its FFT rounding by 1e-3. Every search
window must lie in the frame (the reference's slices silently shrink otherwise): ValueError. The
correlator holding the templates is the int8 form's up to window 129 and search 71, else the FFT
form's, whose search window (window + 2 search) must be at most 192 with no prime factor above 31:
other geometries raise ValueError.
"""

import numpy as np

from . import framecor


def periodicity(ref, y0, x0, window, search, exclude=5):
    """The second peak of the normalized correlation of each window with its own surroundings in the
    reference frame: len(y0) x len(x0) float64 (1 for a periodic or flat texture, small for a unique one)"""
    ref = np.asarray(ref)
    if ref.dtype != np.uint8 or ref.ndim != 2:
        raise ValueError('periodicity: the reference frame must be a height x width uint8 array')
    h, w = ref.shape
    y0 = np.asarray(y0, np.int64).ravel()
    x0 = np.asarray(x0, np.int64).ravel()
    window, search = int(window), int(search)
    # the int8 correlator holds the exact integer moments of the windows for any size it supports;
    # the surfaces themselves do not depend on the method
    method = 'int8' if window <= 129 and search <= 71 else 'fft'
    cor = framecor.Correlator(h, w, y0, x0, window=window, search=search, method=method)
    cor.set_reference(ref)
    return cor.periodicity(ref, exclude=exclude)
