"""
The differentiable form (csrc/diff.hip, slipcam_amd/framecor/autograd.py): the forward NCC of a
bilinearly shifted chip and its hand-written backward pass, checked against an FP64 torch autograd
restatement of the same function (on the CPU) and against central finite differences (h = 1e-3); and
the Gauss-Newton refinement against the known shifts of the smooth-field workload, beside the
quadratic fit of the correlator.
"""

import numpy as np
import pytest

from slipcam_amd import framecor, synthetic
from slipcam_amd.framecor import autograd

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')


def ncc_reference(frames, ref, down, across, W, gross, shifts):
    """FP64 torch restatement: frames (n, h, w) float64, shifts (n, ny, nx, 2) float64 (both may require
    grad); gross (n, ny, nx, 2) int; returns ncc (n, ny, nx)"""
    n = frames.shape[0]
    ny, nx = len(down), len(across)
    ii = torch.arange(W, dtype=torch.float64)
    out = torch.zeros((n, ny, nx), dtype=torch.float64)
    for f in range(n):
        for a in range(ny):
            for b in range(nx):
                t = torch.as_tensor(ref[down[a]:down[a] + W, across[b]:across[b] + W], dtype=torch.float64)
                t = t - t.mean()
                t = t / t.norm()
                y = down[a] + int(gross[f, a, b, 0]) + shifts[f, a, b, 0]
                x = across[b] + int(gross[f, a, b, 1]) + shifts[f, a, b, 1]
                fy, fx = torch.floor(y.detach()), torch.floor(x.detach())
                ay, ax = y - fy, x - fx
                rows = (fy + ii).long()
                cols = (fx + ii).long()
                F = frames[f]
                f00 = F[rows][:, cols]
                f01 = F[rows][:, cols + 1]
                f10 = F[rows + 1][:, cols]
                f11 = F[rows + 1][:, cols + 1]
                C = (1 - ay) * ((1 - ax) * f00 + ax * f01) + ay * ((1 - ax) * f10 + ax * f11)
                c = C - C.mean()
                out[f, a, b] = (t * c).sum() / c.norm()
    return out


@pytest.fixture(scope='module')
def small():
    d = synthetic.smooth_shifts(n=2, height=200, width=280, window=32, search=16, skip=24)
    down, across = framecor.window_grid(200, 280, 32, 16, 24)
    rng = np.random.default_rng(7)
    shifts = rng.uniform(-0.9, 0.9, size=(2, len(down), len(across), 2))
    shifts = shifts + 0.013 * np.sign(shifts)       # away from integer kinks of the bilinear sampling
    gross = np.broadcast_to(d['gross'][:, None, None, :], shifts.shape).astype(np.int32)
    cor = framecor.Correlator(200, 280, down, across, window=32, search=16, method='int8')
    cor.set_reference(d['ref'])
    return d, down, across, shifts, gross, cor


def test_forward_matches_fp64(gpu, small):
    d, down, across, shifts, gross, cor = small
    frames = torch.tensor(d['frames'], dtype=torch.float32, device='cuda')
    s = torch.tensor(shifts, dtype=torch.float32, device='cuda')
    got = autograd.ncc_shift(cor, frames, s, gross=gross).cpu().numpy()
    ref = ncc_reference(torch.tensor(d['frames'], dtype=torch.float64), d['ref'], down, across, 32, gross,
                        torch.tensor(s.cpu().numpy(), dtype=torch.float64)).numpy()
    assert np.abs(got - ref).max() < 2e-6


def test_backward_matches_fp64_autograd(gpu, small):
    d, down, across, shifts, gross, cor = small
    frames = torch.tensor(d['frames'], dtype=torch.float32, device='cuda', requires_grad=True)
    s = torch.tensor(shifts, dtype=torch.float32, device='cuda', requires_grad=True)
    weights = torch.tensor(np.random.default_rng(8).normal(size=shifts.shape[:-1]), dtype=torch.float32)
    ncc = autograd.ncc_shift(cor, frames, s, gross=gross)
    (ncc * weights.cuda()).sum().backward()
    f64 = torch.tensor(d['frames'], dtype=torch.float64, requires_grad=True)
    s64 = torch.tensor(s.detach().cpu().numpy(), dtype=torch.float64, requires_grad=True)
    (ncc_reference(f64, d['ref'], down, across, 32, gross, s64) * weights.double()).sum().backward()
    gs, gs64 = s.grad.cpu().numpy(), s64.grad.numpy()
    assert np.abs(gs - gs64).max() < 1e-4 * max(1.0, np.abs(gs64).max())
    gf, gf64 = frames.grad.cpu().numpy(), f64.grad.numpy()
    assert np.abs(gf - gf64).max() < 1e-4 * np.abs(gf64).max()
    assert np.count_nonzero(gf64) > 0 and np.array_equal(gf64 == 0, np.abs(gf) < 1e-12 * np.abs(gf64).max())


def test_shift_gradient_central_differences(gpu, small):
    """d ncc / d shift against central differences (h = 1e-3) of the FP64 restatement and of the GPU
    forward itself (FP32: a tolerance of FP32 rounding over 2h)"""
    d, down, across, shifts, gross, cor = small
    h = 1e-3
    frames = torch.tensor(d['frames'], dtype=torch.float32, device='cuda')
    s = torch.tensor(shifts, dtype=torch.float32, device='cuda', requires_grad=True)
    autograd.ncc_shift(cor, frames, s, gross=gross).sum().backward()
    analytic = s.grad.cpu().numpy()
    f64 = torch.tensor(d['frames'], dtype=torch.float64)
    base = torch.tensor(s.detach().cpu().numpy(), dtype=torch.float64)
    for axis in range(2):
        e = torch.zeros_like(base)
        e[..., axis] = h
        fd64 = ((ncc_reference(f64, d['ref'], down, across, 32, gross, base + e)
                 - ncc_reference(f64, d['ref'], down, across, 32, gross, base - e)) / (2 * h)).numpy()
        assert np.abs(fd64 - analytic[..., axis]).max() < 2e-3 * max(1.0, np.abs(fd64).max())
        sp = (s.detach() + torch.tensor(e.numpy(), dtype=torch.float32, device='cuda'))
        sm = (s.detach() - torch.tensor(e.numpy(), dtype=torch.float32, device='cuda'))
        fd32 = ((autograd.ncc_shift(cor, frames, sp, gross=gross)
                 - autograd.ncc_shift(cor, frames, sm, gross=gross)) / (2 * h)).cpu().numpy()
        assert np.abs(fd32 - analytic[..., axis]).max() < 5e-3


def test_refine_without_steps_is_the_forward(gpu, small):
    """Zero Gauss-Newton steps return the start shifts and the forward NCC at them"""
    d, down, across, shifts, gross, cor = small
    frames = torch.tensor(d['frames'], dtype=torch.float32, device='cuda')
    s = torch.tensor(shifts, dtype=torch.float32, device='cuda')
    out, ncc = autograd.refine(cor, frames, s, gross=gross, steps=0)
    assert torch.equal(out, s)
    assert torch.allclose(ncc, autograd.ncc_shift(cor, frames, s, gross=gross), rtol=0, atol=1e-6)


def test_gauss_newton_beside_quadratic_fit(gpu):
    """Three Gauss-Newton steps from the correlator's quadratic-fit offsets, against the known shifts
    of the smooth-field workload (a reduced 1080p shape: 270 x 480, window 64, search 32, skip 32)"""
    d = synthetic.smooth_shifts(n=6, height=270, width=480)
    down, across = framecor.window_grid(270, 480, 64, 32, 32)
    cor = framecor.Correlator(270, 480, down, across, window=64, search=32, method='int8')
    cor.set_reference(d['ref'])
    gross = d['gross']
    r = cor.correlate(d['frames'], gross=gross)
    quad = r['offsets'].astype(np.float64)
    valid = np.isfinite(quad[..., 0])
    assert valid.mean() > 0.5
    g4 = np.broadcast_to(gross[:, None, None, :], quad.shape)
    start = np.where(np.isfinite(quad), quad - g4, 0.0)
    frames = torch.tensor(d['frames'], dtype=torch.float32, device='cuda')
    refined, ncc = autograd.refine(cor, frames, torch.tensor(start, dtype=torch.float32, device='cuda'),
                                            gross=gross, steps=3)
    gn = refined.cpu().numpy() + g4
    ok = valid & np.isfinite(gn[..., 0])
    err_quad = np.abs(quad - d['truth'])[ok]
    err_gn = np.abs(gn - d['truth'])[ok]
    print(f'median |error|: quadratic fit {np.median(err_quad):.4f} px, Gauss-Newton {np.median(err_gn):.4f} px')
    assert np.median(err_gn) < 0.05
    assert np.median(err_gn) <= np.median(err_quad) + 0.005
    assert np.all(ncc.cpu().numpy()[ok] > 0.5)


# ---- the differentiation workload at its own size: 1080 x 1920, window 64, search 32, skip 32 (1624
# windows), 128 frames made on the GPU

@pytest.fixture(scope='module')
def full():
    d = synthetic.smooth_shifts_device(n=128, device='cuda')
    down, across = framecor.window_grid(1080, 1920, 64, 32, 32)
    cor = framecor.Correlator(1080, 1920, down, across, window=64, search=32, method='int8')
    cor.set_reference(d['ref'])
    return d, down, across, cor


def ncc_reference_windows(frames, ref, down, across, W, gross, shifts, wins):
    """ncc_reference restricted to the windows wins ((row, column) pairs): frames (n, h, w) float64,
    shifts (n, len(wins), 2) float64, gross (n, ny, nx, 2); returns (n, len(wins))"""
    out = []
    for i, (a, b) in enumerate(wins):
        g = gross[:, a:a + 1, b:b + 1]
        out.append(ncc_reference(frames, ref, down[a:a + 1], across[b:b + 1], W, g, shifts[:, i:i + 1, None])[:, 0, 0])
    return torch.stack(out, 1)


def test_full_size_gradients_on_sampled_windows(gpu, full):
    """The 1080p geometry: forward, backward (frame and shift gradients) and central differences on 16
    sampled windows of 2 frames, against the FP64 torch restatement; the loss weights only those windows,
    so the frame gradient of the GPU (every window's contribution, atomically added) must equal theirs"""
    d, down, across, cor = full
    ny, nx = len(down), len(across)
    rng = np.random.default_rng(12)
    fsel = [5, 77]
    frames = d['frames'][fsel].to(torch.float32).requires_grad_(True)
    flat = rng.choice(ny * nx, 16, replace=False)
    wins = [(int(k // nx), int(k % nx)) for k in flat]
    # away from the integer kinks of the bilinear sampling (central differences across one fail)
    shifts = rng.choice([-1.0, 1.0], size=(2, ny, nx, 2)) * rng.uniform(0.05, 0.9, size=(2, ny, nx, 2))
    gross = np.broadcast_to(d['gross'][fsel][:, None, None, :], shifts.shape).astype(np.int32)
    weights = np.zeros((2, ny, nx))
    for a, b in wins:
        weights[:, a, b] = rng.normal(size=2)
    s = torch.tensor(shifts, dtype=torch.float32, device='cuda', requires_grad=True)
    ncc = autograd.ncc_shift(cor, frames, s, gross=gross)
    (ncc * torch.tensor(weights, dtype=torch.float32, device='cuda')).sum().backward()
    # the FP64 restatement on the sampled windows only
    f64 = frames.detach().cpu().double().requires_grad_(True)
    s64 = torch.tensor(s.detach().cpu().numpy()[:, [a for a, _ in wins], [b for _, b in wins]], dtype=torch.float64,
                       requires_grad=True)
    w_sel = torch.tensor(np.stack([weights[:, a, b] for a, b in wins], 1))
    ref_ncc = ncc_reference_windows(f64, d['ref'], down, across, 64, gross, s64, wins)
    (ref_ncc * w_sel).sum().backward()
    got = ncc.detach().cpu().numpy()[:, [a for a, _ in wins], [b for _, b in wins]]
    assert np.abs(got - ref_ncc.detach().numpy()).max() < 2e-6
    gs = s.grad.cpu().numpy()[:, [a for a, _ in wins], [b for _, b in wins]]
    assert np.abs(gs - s64.grad.numpy()).max() < 1e-4 * max(1.0, np.abs(s64.grad.numpy()).max())
    gf, gf64 = frames.grad.cpu().numpy(), f64.grad.numpy()
    assert np.abs(gf - gf64).max() < 1e-4 * np.abs(gf64).max()
    # windows with zero weight have zero shift gradients
    rest = np.ones((ny, nx), bool)
    for a, b in wins:
        rest[a, b] = False
    assert np.all(s.grad.cpu().numpy()[:, rest] == 0)
    # central differences (h = 1e-3) of the FP64 restatement
    h = 1e-3
    base = torch.tensor(s.detach().cpu().numpy()[:, [a for a, _ in wins], [b for _, b in wins]], dtype=torch.float64)
    f64d = frames.detach().cpu().double()
    for axis in range(2):
        e = torch.zeros_like(base)
        e[..., axis] = h
        fd = ((ncc_reference_windows(f64d, d['ref'], down, across, 64, gross, base + e, wins)
               - ncc_reference_windows(f64d, d['ref'], down, across, 64, gross, base - e, wins)) / (2 * h)).numpy()
        # the GPU's gradient of sum(w ncc) is w d ncc / d shift
        fdw = fd * w_sel.numpy()
        assert np.abs(fdw - gs[..., axis]).max() < 2e-3 * max(1.0, np.abs(fdw).max())


def test_full_size_gauss_newton_beside_quadratic_fit(gpu, full):
    """All 128 frames of the 1080p workload: the integer peak and quadratic fit of the int8 correlator,
    then three Gauss-Newton steps, against the known sub-pixel shifts (the mean of the displacement field
    over each window)"""
    d, down, across, cor = full
    gross = d['gross']
    r = cor.correlate(d['frames'], gross=gross)
    quad = r['offsets'].cpu().numpy().astype(np.float64)
    valid = np.isfinite(quad[..., 0])
    assert valid.mean() > 0.9
    g4 = np.broadcast_to(gross[:, None, None, :], quad.shape)
    start = np.where(np.isfinite(quad), quad - g4, 0.0)
    frames = d['frames'].to(torch.float32)
    refined, ncc = autograd.refine(cor, frames, torch.tensor(start, dtype=torch.float32, device='cuda'),
                                   gross=gross, steps=3)
    gn = refined.cpu().numpy() + g4
    ok = valid & np.isfinite(gn[..., 0])
    assert ok.mean() > 0.9
    err_quad = np.abs(quad - d['truth'])[ok]
    err_gn = np.abs(gn - d['truth'])[ok]
    print(f'1080p, 128 frames, {int(ok.sum())} window-frames: median |error| quadratic fit '
          f'{np.median(err_quad):.4f} px, Gauss-Newton {np.median(err_gn):.4f} px')
    assert np.median(err_gn) < 0.05
    assert np.median(err_gn) <= np.median(err_quad) + 0.005
