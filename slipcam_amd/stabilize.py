"""
Coarse tracking of slipcam-auto on the GPU (python/slipcam/stabilize.py:22-96: ncc_peak, apply,
coarse_track, with slip_series.fit_camera, python/slipcam/slip_series.py:49-72).

A coarse grid of large windows of the reference frame (away from the burned-in text, not flat) is
tracked in each frame around the positions predicted by the similarity transform of the previous frame
(forward from the reference, then backward), and a robust similarity fit of the tracked positions gives
each frame's transform. Per frame, every window's NCC surface and its peak (first maximum, parabolic
refinement along each axis, as ncc_peak) come from one call of the correlator with the predicted
integer offsets as gross offsets (exact NCC: integer sums on the int8 matrix cores, FP64
normalization; the reference evaluates the same quotient with FP64 FFTs); the prediction and the
fit, a few dozen windows, stay on the host as in the reference.

  params, used, ngrid = coarse_track(frames, ref, text)        # as auto_slip.py:272
"""

import numpy as np

from . import framecor


def apply(params, x, y):
    """The positions (x, y) (relative to the image centre) moved by the similarity transform
    params = (p, q, a, s): x' = x + p - a y + s x, y' = y + q + a x + s y (stabilize.py:48-52)"""
    p, q, a, s = params
    return x + p - a * y + s * x, y + q + a * x + s * y


def fit_camera(dx, dy, x, y, model='similarity', iterations=3):
    """
    The robust least-squares fit of the similarity transform (p, q, a, s) to the offsets dx (across)
    and dy (down) of the points (x, y) (slip_series.py:49-72): each of the iterations solves the
    least-squares system of the points still in use, then drops the points whose residual (in either
    component) is at least max(3 x the scaled median absolute deviation of the residuals, 0.3 px);
    with model 'shift' only (p, q) (a = s = 0). Returns (params, the number of points in use), or
    (NaN x 4, 0) when fewer than 8 residuals remain.
    """
    n = len(x)
    design = np.zeros((2 * n, 4))
    design[:n, 0] = 1.0
    design[n:, 1] = 1.0
    if model != 'shift':
        design[:n, 2], design[:n, 3] = -y, x
        design[n:, 2], design[n:, 3] = x, y
    d = np.concatenate([dx, dy])
    finite = np.isfinite(dx) & np.isfinite(dy)
    use = np.concatenate([finite, finite])
    m = np.full(4, np.nan)
    for _ in range(iterations):
        if use.sum() < 8:
            return np.full(4, np.nan), 0
        m = np.linalg.lstsq(design[use], d[use], rcond=None)[0]
        res = np.where(use, d - design @ m, np.nan)
        spread = 1.4826 * np.nanmedian(np.abs(res - np.nanmedian(res)))
        both = use & (np.abs(res) < max(3 * spread, 0.3))
        both = both[:n] & both[n:]
        use = np.concatenate([both, both])
    return m, use[:n].sum()


class Grid:
    """The coarse grid of coarse_track (stabilize.py:61-69): window centres yc x xc, step apart, at least
    search + window / 2 from the borders; keep: the windows (row-major over the grid) that miss the text
    mask and whose reference pixels have a standard deviation above 2"""

    def __init__(self, refimg, text, window, step, search):
        h, w = refimg.shape
        self.h, self.w, self.window, self.search = h, w, window, search
        self.half = half = window // 2
        self.yc = np.arange(search + half, h - search - half, step)
        self.xc = np.arange(search + half, w - search - half, step)
        yy, xx = (v.ravel() for v in np.meshgrid(self.yc, self.xc, indexing='ij'))
        ref64 = refimg.astype(np.float64)
        self.keep = np.array([not text[y - half:y + half, x - half:x + half].any()
                              and ref64[y - half:y + half, x - half:x + half].std() > 2
                              for y, x in zip(yy, xx)], bool)
        self.Yg, self.Xg = yy[self.keep], xx[self.keep]
        self.X0, self.Y0 = self.Xg - w / 2, self.Yg - h / 2


def track_frames(frames, ref, grid, peaks, min_peak=0.5, start=0):
    """
    The frame loop of coarse_track (stabilize.py:72-96) over any peak finder: peaks(frame, gross)
    returns the (offsets, peak) of every window of the grid (down x across x 2 and down x across,
    offsets relative to the window's position plus gross); gross is down x across x 2 integers.
    Returns (params frames x 4, used frames).
    """
    n = frames.shape[0]
    h, w, half, search, L = grid.h, grid.w, grid.half, grid.search, grid.window + 2 * grid.search
    ny, nx = len(grid.yc), len(grid.xc)
    params = np.full((n, 4), np.nan)
    used = np.zeros(n, int)
    params[ref] = 0
    used[ref] = len(grid.Xg)
    for order in (range(ref + 1, n), range(ref - 1, start - 1, -1)):
        prev = np.zeros(4)
        for k in order:
            px, py = apply(prev, grid.X0, grid.Y0)
            ry = np.round(py + h / 2).astype(np.int64)      # round half to even, as Python's round
            rx = np.round(px + w / 2).astype(np.int64)
            y0, x0 = ry - half - search, rx - half - search
            inside = (y0 >= 0) & (x0 >= 0) & (y0 + L <= h) & (x0 + L <= w)
            gross = np.zeros((ny * nx, 2), np.int32)
            gross[grid.keep, 0] = np.where(inside, ry - grid.Yg, 0)
            gross[grid.keep, 1] = np.where(inside, rx - grid.Xg, 0)
            off, peak = peaks(frames[k], gross.reshape(ny, nx, 2))
            off = off.reshape(-1, 2)[grid.keep]
            peak = peak.ravel()[grid.keep]
            ok = inside & (peak >= min_peak)
            dx = np.where(ok, (x0 + search + half + off[:, 1]) - w / 2 - grid.X0, np.nan)
            dy = np.where(ok, (y0 + search + half + off[:, 0]) - h / 2 - grid.Y0, np.nan)
            fit, nuse = fit_camera(dx, dy, grid.X0, grid.Y0)
            if np.all(np.isfinite(fit)):
                params[k], used[k], prev = fit, nuse, fit
            else:
                params[k], used[k] = prev, 0
    return params, used


def _correlator(h, w, down, across, window, search):
    # the surfaces do not depend on the method: the int8 correlator where it supports the sizes
    method = 'int8' if window <= 129 and search <= 71 else 'fft'
    return framecor.Correlator(h, w, down, across, window=window, search=search, method=method)


def coarse_track(frames, ref, text, window=96, step=96, search=32, min_peak=0.5, start=0):
    """The similarity transform of each frame relative to the frame ref (frames x 4: p, q, a, s about the
    image centre), the number of windows used per frame, and the number of windows of the grid
    (stabilize.py:55-96); frames: frames x height x width uint8, text: height x width bool"""
    frames = np.asarray(frames)
    if frames.dtype != np.uint8 or frames.ndim != 3:
        raise ValueError('coarse_track: the frames must be a frames x height x width uint8 array')
    n, h, w = frames.shape
    text = np.asarray(text, bool)
    grid = Grid(frames[ref], text, window, step, search)
    if len(grid.Xg) == 0:
        # no window to track: every fit fails and each frame keeps the previous transform, as in the
        # reference; there is nothing to correlate
        def peaks(frame, gross):
            return np.full(gross.shape, np.nan), np.zeros(gross.shape[:2])
    else:
        cor = _correlator(h, w, grid.yc - grid.half, grid.xc - grid.half, window, search)
        cor.set_reference(frames[ref])
        peaks = cor.ncc_peaks
    params, used = track_frames(frames, ref, grid, peaks, min_peak=min_peak, start=start)
    return params, used, len(grid.Xg)


def ncc_peak(template, search):
    """The position (dy, dx) of the peak of the NCC of template in search, relative to the centre of
    search, refined by a parabola along each axis, and the peak value (stabilize.py:22-45), for a
    square template of uint8 values in a square search image larger by an even number of pixels"""
    template = np.asarray(template)
    search = np.asarray(search)
    W, Ls = template.shape[0], search.shape[0]
    if (template.ndim != 2 or search.ndim != 2 or template.shape != (W, W) or search.shape != (Ls, Ls)
            or Ls <= W or (Ls - W) % 2):
        raise ValueError('ncc_peak: a square template and a square search image larger by an even number of pixels')
    if np.any(template != np.round(template)) or template.min() < 0 or template.max() > 255 \
            or np.any(search != np.round(search)) or search.min() < 0 or search.max() > 255:
        raise ValueError('ncc_peak: the pixels must be integers 0 .. 255')
    R = (Ls - W) // 2
    ref = np.zeros((Ls, Ls), np.uint8)
    ref[R:R + W, R:R + W] = template
    cor = _correlator(Ls, Ls, [R], [R], W, R)
    cor.set_reference(ref)
    off, peak = cor.ncc_peaks(search.astype(np.uint8))
    return off[0, 0, 0], off[0, 0, 1], peak[0, 0]
