"""
Parity checks shared by the GPU tests: the HIP path's result dicts against the oracle's.

  int8 form vs the FP64 direct-NCC oracle: the same validity and integer peaks, offsets, peak and
  covariance to FP32 rounding of the same FP64 values, SNR to 1e-6 relative
  FFT form vs the restatement of the reference CPU algorithm (FP32 FFT): offsets within 1e-3 px, peak
  1e-4, SNR 1e-4 relative (the reference's cross-backend bar, tests/test_framecor.py:52-62), except where
  the two FP32 FFTs pick peaks whose exact FP64 values are a near-tie (checked window by window on the
  FP64 surfaces)
"""

import numpy as np

from slipcam_amd import framecor


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
    if same.any():
        assert np.abs(got['peak'] - ref['peak'])[same].max() < 1e-4
        assert np.abs(got['snr'][same] / ref['snr'][same] - 1).max() < 1e-4
    return int(differ.sum())


def flat_select(r, sample):
    """The windows sample (flat indices) of every frame of a result dict: frames x len(sample) x ..."""
    return {k: r[k].reshape((r[k].shape[0], -1) + r[k].shape[3:])[:, sample] for k in r}
