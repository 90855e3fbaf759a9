#!/usr/bin/env python3
"""
slipcam-track on an MI355X: per-window motion of every frame of a video against one reference frame

The drop-in for the reference's slipcam-track (python/slipcam/track_offsets.py): each frame from
--first to --last is correlated with the reference frame (--ref) on a grid of windows; the offsets are
in pixels, (down, across), the motion of the frame relative to the reference; the windows that overlap
the burned-in text boxes (--mask) are flagged. The output .npz has the reference's keys (index, ref,
masked, yc, xc, window, skip, search, engine='framecor', offsets, snr, cov, peak), so slipcam-auto's
cache (auto_slip.py:292-296) takes it as its own.

The frames stream through the GPU in chunks (64 frames): while the GPU correlates a chunk, the host
copies the next one from the (memory-mapped) frames into pinned memory and a copy stream moves it to
the device, so disk, PCIe and the correlation overlap; the outputs stay on the device until the end.

usage:
  python -m slipcam_amd.track run/frames.npy --ref 285 --first 140 --out run/offsets.npz \\
      --mask 870 25 1240 68 --mask 40 650 630 700 [--method int8]
"""

import argparse
import time

import numpy as np

from . import framecor


def boxes_mask(y0, x0, window, boxes):
    """The windows that overlap any of the boxes (x0, y0, x1, y1) (track_offsets.py:27-32)"""
    y0, x0 = np.asarray(y0), np.asarray(x0)
    masked = np.zeros((len(y0), len(x0)), bool)
    for bx0, by0, bx1, by1 in boxes:
        masked |= ((y0[:, None] < by1) & (y0[:, None] + window > by0) & (x0[None, :] < bx1)
                   & (x0[None, :] + window > bx0))
    return masked


def track(frames, ref, index, masked=None, window=64, skip=16, search=48, method='fft', chunk=64, device=0,
          verbose=True):
    """
    The offsets of the frames index relative to the frame ref, chunk frames per correlation call; the
    dict that main saves (track_offsets.py:35-60). frames: frames x height x width uint8 (a numpy array
    or memmap); method: 'fft' (the reference's algorithm) or 'int8'.
    """
    import torch
    n, h, w = frames.shape
    y0, x0 = framecor.window_grid(h, w, window, search, skip)
    ny, nx = len(y0), len(x0)
    if masked is None:
        masked = np.zeros((ny, nx), bool)
    cor = framecor.Correlator(h, w, y0, x0, window=window, search=search, method=method, device=device)
    index = np.asarray(index)
    dev = torch.device('cuda', device)
    out = {key: torch.empty((len(index), ny, nx) + tail, dtype=torch.float32, device=dev)
           for key, tail in (('offsets', (2,)), ('snr', ()), ('cov', (3,)), ('peak', ()))}
    tic = time.perf_counter()
    cor.set_reference(np.ascontiguousarray(frames[ref]))
    chunk = max(1, min(chunk, len(index)))
    # two pinned host buffers and two device buffers: chunk k + 1 is copied while chunk k is correlated
    pinned = [torch.empty((chunk, h, w), dtype=torch.uint8, pin_memory=True) for _ in range(2)]
    staged = [torch.empty((chunk, h, w), dtype=torch.uint8, device=dev) for _ in range(2)]
    copies = torch.cuda.Stream(device=dev)
    compute = torch.cuda.current_stream(dev)
    copied = [torch.cuda.Event() for _ in range(2)]
    done = [torch.cuda.Event() for _ in range(2)]

    def load(k0, slot):
        sel = index[k0:k0 + chunk]
        # a run of consecutive frames: a slice of the (memory-mapped) frames, without a gather
        run = np.array_equal(sel, np.arange(sel[0], sel[0] + len(sel)))
        src = frames[sel[0]:sel[0] + len(sel)] if run else frames[sel]
        done[slot].synchronize()                  # the pinned buffer's previous copy has been used
        np.copyto(pinned[slot][:len(sel)].numpy(), src)
        with torch.cuda.stream(copies):
            copies.wait_event(done[slot])         # the device buffer's previous chunk is correlated
            staged[slot][:len(sel)].copy_(pinned[slot][:len(sel)], non_blocking=True)
            copied[slot].record(copies)
        return len(sel)

    starts = list(range(0, len(index), chunk))
    for e in done:
        e.record(compute)
    count = load(starts[0], 0) if starts else 0
    for i, k0 in enumerate(starts):
        slot = i % 2
        compute.wait_event(copied[slot])
        r = cor.correlate(staged[slot][:count])
        for key in out:
            out[key][k0:k0 + count].copy_(r[key])
        done[slot].record(compute)
        if i + 1 < len(starts):
            count = load(starts[i + 1], 1 - slot)
    result = {key: v.cpu().numpy() for key, v in out.items()}
    elapsed = time.perf_counter() - tic
    if verbose:
        print(f'{len(index)} frames of {ny} x {nx} windows in {elapsed:.1f} s '
              f'({elapsed / max(len(index), 1) * 1e3:.1f} ms per frame, framecor {cor.backend} {method}), '
              f'{masked.sum()} windows masked')
    return dict(index=index, ref=ref, masked=masked, yc=y0 + window / 2, xc=x0 + window / 2, window=window,
                skip=skip, search=search, engine='framecor', **result)


def main(argv=None):
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    parser.add_argument('frames', help='video as a .npy array, n x height x width uint8 (memory-mapped)')
    parser.add_argument('--ref', type=int, required=True, help='index of the frame every other frame is matched to')
    parser.add_argument('--first', type=int, default=0, help='index of the first frame tracked (default: %(default)s)')
    parser.add_argument('--last', type=int, help='index of the last frame tracked (default: the final frame)')
    parser.add_argument('--window', type=int, default=64, help='side of each square window in pixels (default: %(default)s)')
    parser.add_argument('--skip', type=int, default=16, help='grid spacing of the windows in pixels (default: %(default)s)')
    parser.add_argument('--search', type=int, default=48,
                        help='largest offset searched along each axis in pixels (default: %(default)s)')
    parser.add_argument('--mask', type=int, nargs=4, action='append', default=[], metavar=('X0', 'Y0', 'X1', 'Y1'),
                        help='pixel box of on-screen text whose windows are flagged; may be given several times')
    parser.add_argument('--method', choices=('fft', 'int8'), default='fft',
                        help='fft: the reference algorithm; int8: exact int8 matrix-core correlation (default: %(default)s)')
    parser.add_argument('--chunk', type=int, default=64, help='frames per correlation call (default: %(default)s)')
    parser.add_argument('--out', required=True, help='path of the .npz written')
    args = parser.parse_args(argv)

    frames = np.load(args.frames, mmap_mode='r')
    n, h, w = frames.shape
    last = n - 1 if args.last is None else args.last
    y0, x0 = framecor.window_grid(h, w, args.window, args.search, args.skip)
    masked = boxes_mask(y0, x0, args.window, args.mask)
    result = track(frames, args.ref, np.arange(args.first, last + 1), masked=masked, window=args.window,
                   skip=args.skip, search=args.search, method=args.method, chunk=args.chunk)
    np.savez(args.out, **result)
    print(f'-> {args.out}')


if __name__ == '__main__':
    main()
