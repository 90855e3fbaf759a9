"""
The differentiable form of framecor: the NCC of each window of the reference with a continuously
shifted, bilinearly sampled chip of a frame, as a torch.autograd.Function over hand-written HIP
forward and backward kernels, and a Gauss-Newton refinement of the shifts.

  cor = framecor.Correlator(h, w, down, across, window=64, search=32, method='int8')
  cor.set_reference(ref)
  r = cor.correlate(frames, gross=g)                       # integer peak + quadratic fit
  shifts = torch.tensor(r['offsets'] - g, device='cuda')   # continuous part, per window
  ncc = ncc_shift(cor, frames_f32, shifts, gross=g)        # differentiable in frames and shifts
  refined, ncc = refine(cor, frames_f32, shifts, gross=g, steps=3)

frames: n x h x w float32 on the GPU; shifts: n x ny x nx x 2 (down, across) float32 on the GPU,
relative to each window's position plus its integer gross offset; gross: None, n x 2 or
n x ny x nx x 2 integers. The reference has no such form; the kernels are in csrc/diff.hip.
"""

import numpy as np
import torch

from . import _check, _lib


def _prepare(cor, frames, shifts, gross):
    if not (isinstance(frames, torch.Tensor) and frames.is_cuda and frames.dtype == torch.float32):
        raise ValueError('framecor: the frames must be a float32 tensor on the GPU')
    h, w = cor._h, cor._w
    ny, nx = cor.shape
    if frames.dim() != 3 or tuple(frames.shape[1:]) != (h, w):
        raise ValueError('framecor: the frames must be frames x height x width')
    n = frames.shape[0]
    if tuple(shifts.shape) != (n, ny, nx, 2) or shifts.dtype != torch.float32 or shifts.device != frames.device:
        raise ValueError('framecor: shifts must be a float32 tensor frames x down x across x 2 on the frames\' GPU')
    g = None
    if gross is not None:
        if isinstance(gross, torch.Tensor):
            g = gross.to(device=frames.device, dtype=torch.int32)
        else:
            # host offsets: staged in pinned memory and copied without blocking (a pageable copy would
            # wait for the work already queued on the stream)
            g = torch.from_numpy(np.ascontiguousarray(gross, np.int32)).pin_memory()
            g = g.to(device=frames.device, non_blocking=True)
        if tuple(g.shape) == (n, 2):
            g = g[:, None, None, :].expand(n, ny, nx, 2)
        elif tuple(g.shape) != (n, ny, nx, 2):
            raise ValueError('framecor: gross must be frames x 2 or frames x down x across x 2')
        g = g.contiguous()
    return frames.contiguous(), shifts.contiguous(), g, n


def _stream(t):
    return torch.cuda.current_stream(t.device).cuda_stream


class NccShift(torch.autograd.Function):
    """ncc[f, i, j] of window (i, j) against frame f sampled at its position + gross + shift"""

    @staticmethod
    def forward(ctx, frames, shifts, cor, gross):
        frames, shifts, g, n = _prepare(cor, frames, shifts, gross)
        ncc = torch.empty(shifts.shape[:-1], dtype=torch.float32, device=frames.device)
        if n:
            _check(_lib.framecor_ncc_forward(cor._c, frames.data_ptr(), n, None if g is None else g.data_ptr(),
                                             shifts.data_ptr(), ncc.data_ptr(), _stream(frames)))
        ctx.save_for_backward(frames, shifts)
        ctx.cor, ctx.gross = cor, g
        return ncc

    @staticmethod
    def backward(ctx, grad):
        frames, shifts = ctx.saved_tensors
        cor, g = ctx.cor, ctx.gross
        n = frames.shape[0]
        grad = grad.contiguous().to(torch.float32)
        gf = torch.zeros_like(frames)
        gs = torch.empty_like(shifts)
        if n:
            _check(_lib.framecor_ncc_backward(cor._c, frames.data_ptr(), n, None if g is None else g.data_ptr(),
                                              shifts.data_ptr(), grad.data_ptr(), gf.data_ptr(), gs.data_ptr(),
                                              _stream(frames)))
        return (gf if ctx.needs_input_grad[0] else None), (gs if ctx.needs_input_grad[1] else None), None, None


def ncc_shift(cor, frames, shifts, gross=None):
    """The differentiable NCC of every window at a continuous shift (n x ny x nx)"""
    return NccShift.apply(frames, shifts, cor, gross)


def refine(cor, frames, shifts, gross=None, steps=3):
    """Gauss-Newton refinement of the shifts (steps steps, each bounded to 1 px); returns
    (shifts, ncc): NaN shifts and 0 ncc where the chip leaves the frame"""
    with torch.no_grad():
        frames, shifts, g, n = _prepare(cor, frames, shifts.detach(), gross)
        out = torch.empty_like(shifts)
        ncc = torch.empty(shifts.shape[:-1], dtype=torch.float32, device=frames.device)
        if n:
            _check(_lib.framecor_ncc_refine(cor._c, frames.data_ptr(), n, None if g is None else g.data_ptr(),
                                            shifts.data_ptr(), int(steps), out.data_ptr(), ncc.data_ptr(),
                                            _stream(frames)))
        return out, ncc
