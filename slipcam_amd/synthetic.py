"""
Seeded synthetic frames for the framecor workloads (numpy/scipy; no downloads)

  texture      white noise N(0, 1), Gaussian-filtered (sigma 1.5 px), rescaled linearly to 0-255
  small_shifts 256 x 320 reference, 8 frames: the reference bilinearly shifted by a per-frame
               (down, across) offset uniform in +-3 px, plus Gaussian noise sigma 2 DN, rounded and
               clipped to uint8; window 32, search 16, skip 32 (48 windows)
  fault_video  720 x 1280, 64 frames: a straight fault at 30 degrees through the image centre; the
               whole frame shakes by a random walk (step sigma 0.5 px, bounded +-5 px) and the far side
               of the fault adds a sigmoid slip ramp to 40 px along the fault over frames 20-40
               (centre 30, scale 2.5 frames); two fixed text-like boxes of 0/255 blocks (370 x 45 at
               x 870-1240, y 25-70 and 590 x 50 at x 40-630, y 650-700); noise sigma 2 DN; window 64,
               search 48, skip 16 (2450 windows)
  gross_offsets per-window integer gross offsets uniform in -20..20
  smooth_shifts 1080 x 1920, 128 frames, window 64, search 32, skip 32: each frame moved by an integer
               offset uniform in -20..20 (the gross offset of all its windows) plus a smooth sub-pixel
               displacement field within +-0.5 px (long-wavelength products of sines); windows overlap
               (skip 32 < window 64), so the truth of a window is the mean of the field over its pixels

Seeds: texture 0, shifts 1, noise 2, gross and clip parameters 3, boxes 4.
"""

import numpy as np
from scipy import ndimage


def texture(height, width, sigma=1.5, seed=0):
    """Smoothed white noise rescaled to 0-255 (float64)"""
    img = ndimage.gaussian_filter(np.random.default_rng(seed).normal(size=(height, width)), sigma)
    return (img - img.min()) / (img.max() - img.min()) * 255.0


def to_uint8(img):
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def warp(image, dy, dx):
    """The image moved by the displacement (dy, dx) (scalars or per-pixel arrays): out(y, x) =
    image(y - dy, x - dx), bilinear, reflected at the borders"""
    h, w = image.shape
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    return ndimage.map_coordinates(image, [yy - dy, xx - dx], order=1, mode='reflect')


def small_shifts(n=8, height=256, width=320):
    """The CPU-sized workload: a dict with ref, frames, shifts (n x 2) and the grid parameters"""
    ref = to_uint8(texture(height, width))
    shifts = np.random.default_rng(1).uniform(-3, 3, size=(n, 2))
    noise = np.random.default_rng(2)
    base = ref.astype(np.float64)
    frames = np.stack([to_uint8(warp(base, s[0], s[1]) + noise.normal(0.0, 2.0, base.shape)) for s in shifts])
    return dict(ref=ref, frames=frames, shifts=shifts, window=32, search=16, skip=32)


def text_boxes(height, width, seed=4):
    """The mask and the pixel values of the two fixed text-like boxes (x0, y0, x1, y1)"""
    boxes = [(870, 25, 1240, 70), (40, 650, 630, 700)]
    rng = np.random.default_rng(seed)
    mask = np.zeros((height, width), bool)
    values = np.zeros((height, width), np.uint8)
    for x0, y0, x1, y1 in boxes:
        x1, y1 = min(x1, width), min(y1, height)
        if x0 >= x1 or y0 >= y1:
            continue
        cells = rng.integers(0, 2, size=((y1 - y0 + 6) // 7, (x1 - x0 + 4) // 5)).astype(np.uint8) * 255
        block = np.kron(cells, np.ones((7, 5), np.uint8))[:y1 - y0, :x1 - x0]
        mask[y0:y1, x0:x1] = True
        values[y0:y1, x0:x1] = block
    return mask, values, boxes


def fault_video(n=64, height=720, width=1280, angle=30.0, slip=40.0, centre=30.0, scale=2.5):
    """The README-shaped workload: a dict with ref, frames, shake (n x 2), slip (n), side (h x w bool),
    boxes and the grid parameters"""
    base = texture(height, width)
    mask, values, boxes = text_boxes(height, width)
    ref = to_uint8(base)
    ref[mask] = values[mask]
    rng = np.random.default_rng(3)
    shake = np.zeros((n, 2))
    for f in range(1, n):
        shake[f] = np.clip(shake[f - 1] + rng.normal(0.0, 0.5, 2), -5.0, 5.0)
    a = np.deg2rad(angle)
    along = np.array([-np.sin(a), np.cos(a)])       # (down, across) direction of the fault
    normal = np.array([np.cos(a), np.sin(a)])
    yy, xx = np.mgrid[0:height, 0:width]
    side = (yy - height / 2) * normal[0] + (xx - width / 2) * normal[1] > 0
    ramp = slip / (1.0 + np.exp(-(np.arange(n) - centre) / scale))
    noise = np.random.default_rng(2)
    src = base.copy()
    frames = np.empty((n, height, width), np.uint8)
    for f in range(n):
        dy = shake[f, 0] + side * ramp[f] * along[0]
        dx = shake[f, 1] + side * ramp[f] * along[1]
        frame = to_uint8(warp(src, dy, dx) + noise.normal(0.0, 2.0, src.shape))
        frame[mask] = values[mask]
        frames[f] = frame
    return dict(ref=ref, frames=frames, shake=shake, slip=ramp, along=along, side=side, boxes=boxes,
                window=64, search=48, skip=16)


RESOLUTIONS = ((480, 640), (720, 1280), (1080, 1920))


def clip_specs(n=16, seed=3, frames=(8, 256), resolutions=RESOLUTIONS, windows=(32, 64, 128), searches=(16, 32, 64)):
    """The batching workload: n clips, each with a resolution, window and search range drawn from the
    given sets, a frame count uniform in frames[0]..frames[1], skip = window / 4 and a shift sigma
    uniform in 1..5 px (a list of dicts; the grid from the reference's window_grid rule)"""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        h, w = resolutions[rng.integers(len(resolutions))]
        W = int(windows[rng.integers(len(windows))])
        R = int(searches[rng.integers(len(searches))])
        nf = int(rng.integers(frames[0], frames[1] + 1))
        sigma = float(rng.uniform(1.0, 5.0))
        skip = W // 4
        down = R + skip * np.arange((h - 2 * R - W) // skip)
        across = R + skip * np.arange((w - 2 * R - W) // skip)
        out.append(dict(height=h, width=w, window=W, search=R, skip=skip, frames=nf, sigma=sigma, down=down,
                        across=across, seed=seed * 1000 + i))
    return out


def clip_video(spec, device):
    """The reference (numpy uint8) and frames (torch uint8 on device) of one clip: the texture of its
    resolution moved by a per-frame shift N(0, sigma^2) per axis (clipped to +-(search - 1) px),
    bilinearly (borders reflected), plus noise sigma 2 DN, rounded and clipped; and the shifts (n x 2)"""
    import torch
    h, w, n = spec['height'], spec['width'], spec['frames']
    base = texture(h, w)
    ref = to_uint8(base)
    rng = np.random.default_rng(spec['seed'])
    lim = spec['search'] - 1
    shifts = np.clip(rng.normal(0.0, spec['sigma'], size=(n, 2)), -lim, lim)
    pad = int(np.ceil(np.abs(shifts).max())) + 2 if n else 2
    img = torch.tensor(np.pad(base, pad, mode='reflect'), dtype=torch.float32, device=device)
    gen = torch.Generator(device=device)
    gen.manual_seed(spec['seed'])
    frames = torch.empty((n, h, w), dtype=torch.uint8, device=device)
    for f in range(n):
        # out(y, x) = base(y - dy, x - dx): the padded image at (pad - dy + y, pad - dx + x)
        sy, sx = pad - shifts[f, 0], pad - shifts[f, 1]
        iy, ix = int(np.floor(sy)), int(np.floor(sx))
        fy, fx = float(sy - iy), float(sx - ix)
        a = img[iy:iy + h, ix:ix + w]
        b = img[iy:iy + h, ix + 1:ix + 1 + w]
        c = img[iy + 1:iy + 1 + h, ix:ix + w]
        d = img[iy + 1:iy + 1 + h, ix + 1:ix + 1 + w]
        moved = (1 - fy) * ((1 - fx) * a + fx * b) + fy * ((1 - fx) * c + fx * d)
        noise = torch.randn((h, w), generator=gen, device=device) * 2.0
        frames[f] = torch.clamp(torch.round(moved + noise), 0, 255).to(torch.uint8)
    return ref, frames, shifts


def gross_offsets(n, ny, nx, low=-20, high=20, seed=3):
    """Per-window integer gross offsets, n x ny x nx x 2 int32, uniform in low..high"""
    return np.random.default_rng(seed).integers(low, high + 1, size=(n, ny, nx, 2)).astype(np.int32)


def smooth_shifts(n=128, height=1080, width=1920, window=64, search=32, skip=32, seed=6):
    """The differentiation workload: a dict with ref, frames, gross (n x 2 int32), truth (n x ny x nx x 2:
    the gross offset plus the mean of the field over each window) and the grid parameters"""
    base = texture(height, width)
    ref = to_uint8(base)
    rng = np.random.default_rng(seed)
    gross = rng.integers(-20, 21, size=(n, 2)).astype(np.int32)
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    noise = np.random.default_rng(2)
    first = search
    down = first + skip * np.arange((height - 2 * search - window) // skip)
    across = first + skip * np.arange((width - 2 * search - window) // skip)
    frames = np.empty((n, height, width), np.uint8)
    truth = np.empty((n, len(down), len(across), 2))
    for f in range(n):
        field = []
        for axis in range(2):
            ly, lx = rng.uniform(500, 900, 2)
            py, px = rng.uniform(0, 2 * np.pi, 2)
            field.append(0.5 * np.sin(2 * np.pi * yy / ly + py) * np.sin(2 * np.pi * xx / lx + px))
        dy, dx = gross[f, 0] + field[0], gross[f, 1] + field[1]
        frames[f] = to_uint8(warp(base, dy, dx) + noise.normal(0.0, 2.0, base.shape))
        for a in range(2):
            c = np.cumsum(np.cumsum(np.pad(field[a], ((1, 0), (1, 0))), 0), 1)
            box = c[window:, window:] - c[:-window, window:] - c[window:, :-window] + c[:-window, :-window]
            truth[f, :, :, a] = gross[f, a] + box[np.ix_(down, across)] / (window * window)
    return dict(ref=ref, frames=frames, gross=gross, truth=truth, window=window, search=search, skip=skip)


def smooth_shifts_device(n=128, height=1080, width=1920, window=64, search=32, skip=32, seed=6, device='cuda'):
    """smooth_shifts made on a torch device (the full 128-frame 1080p workload in seconds): the same
    texture and reference, integer offsets and sine-product fields drawn the same way, the frames warped
    bilinearly with reflected borders and noise sigma 2 DN in float64 on the device (so not the same
    noise bits as smooth_shifts); frames is a torch uint8 tensor on the device, the rest numpy"""
    import torch
    base = texture(height, width)
    ref = to_uint8(base)
    rng = np.random.default_rng(seed)
    gross = rng.integers(-20, 21, size=(n, 2)).astype(np.int32)
    dev = torch.device(device)
    img = torch.tensor(base, dtype=torch.float64, device=dev)
    yy = torch.arange(height, dtype=torch.float64, device=dev)[:, None].expand(height, width)
    xx = torch.arange(width, dtype=torch.float64, device=dev)[None, :].expand(height, width)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)

    def reflect(i, size):
        # scipy's 'reflect' (half-sample symmetric) boundary for integer taps
        i = torch.remainder(i, 2 * size)
        return torch.where(i >= size, 2 * size - 1 - i, i)

    first = search
    down = first + skip * np.arange((height - 2 * search - window) // skip)
    across = first + skip * np.arange((width - 2 * search - window) // skip)
    frames = torch.empty((n, height, width), dtype=torch.uint8, device=dev)
    truth = np.empty((n, len(down), len(across), 2))
    for f in range(n):
        field = []
        for axis in range(2):
            ly, lx = rng.uniform(500, 900, 2)
            py, px = rng.uniform(0, 2 * np.pi, 2)
            field.append(0.5 * torch.sin(2 * np.pi * yy / ly + py) * torch.sin(2 * np.pi * xx / lx + px))
        cy, cx = yy - (gross[f, 0] + field[0]), xx - (gross[f, 1] + field[1])
        y0, x0 = torch.floor(cy), torch.floor(cx)
        ay, ax = cy - y0, cx - x0
        r0, r1 = reflect(y0.long(), height), reflect(y0.long() + 1, height)
        c0, c1 = reflect(x0.long(), width), reflect(x0.long() + 1, width)
        moved = ((1 - ay) * ((1 - ax) * img[r0, c0] + ax * img[r0, c1])
                 + ay * ((1 - ax) * img[r1, c0] + ax * img[r1, c1]))
        noise = torch.randn((height, width), generator=gen, dtype=torch.float64, device=dev) * 2.0
        frames[f] = torch.clamp(torch.round(moved + noise), 0, 255).to(torch.uint8)
        for a in range(2):
            c = torch.nn.functional.pad(field[a], (1, 0, 1, 0)).cumsum(0).cumsum(1)
            box = c[window:, window:] - c[:-window, window:] - c[window:, :-window] + c[:-window, :-window]
            truth[f, :, :, a] = gross[f, a] + box.cpu().numpy()[np.ix_(down, across)] / (window * window)
    return dict(ref=ref, frames=frames, gross=gross, truth=truth, window=window, search=search, skip=skip)


def similarity_video(n=8, height=480, width=640, ref=3, seed=8):
    """The coarse-tracking workload: a texture seen by a camera that moves by a similarity transform per
    frame (shift up to 8 px, rotation and scale up to 2e-3 about the image centre; frame ref is the
    reference, params 0), with two burned-in text boxes fixed in the frames. frame_k sampled at
    apply(params_k, x, y) (centred coordinates) shows the reference at (x, y). Returns a dict with
    frames (n x h x w uint8), params (n x 4: p, q, a, s), text (h x w bool) and ref."""
    base = texture(height, width, sigma=2.5, seed=seed)
    rng = np.random.default_rng(seed)
    params = np.column_stack([rng.uniform(-8, 8, n), rng.uniform(-8, 8, n), rng.uniform(-2e-3, 2e-3, n),
                              rng.uniform(-2e-3, 2e-3, n)])
    params[ref] = 0
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    xc, yc = xx - width / 2, yy - height / 2
    text = np.zeros((height, width), bool)
    text[20:60, 30:250] = True
    text[height - 70:height - 30, width - 300:width - 40] = True
    noise = np.random.default_rng(2)
    frames = np.empty((n, height, width), np.uint8)
    for k, (p, q, a, s) in enumerate(params):
        # the reference point (x, y) appears at x' = (1 + s) x - a y + p, y' = a x + (1 + s) y + q
        det = (1 + s) ** 2 + a * a
        x = ((1 + s) * (xc - p) + a * (yc - q)) / det
        y = (-a * (xc - p) + (1 + s) * (yc - q)) / det
        img = ndimage.map_coordinates(base, [y + height / 2, x + width / 2], order=3, mode='reflect')
        img = img + noise.normal(0.0, 1.0, img.shape)
        img[text] = np.where((xx[text] // 8 + yy[text] // 8) % 2 == 0, 0.0, 255.0)
        frames[k] = to_uint8(img)
    return dict(frames=frames, params=params, text=text, ref=ref)
