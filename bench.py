#!/usr/bin/env python3
"""
Benchmark of the framecor correlator on one MI355X.

Workload: the README shape of the reference (720 x 1280 uint8 frames, window 64, search +-48, skip 16:
2450 windows): 64 synthetic frames of a smoothed-noise texture with a 30-degree fault, camera shake, a
40 px slip ramp and two fixed text boxes (slipcam_amd.synthetic.fault_video), resident in HBM before
the timed region. A step is one correlate() of the 64 frames (the box-sum maps and the correlation
kernel with its fused peak epilogue), device-resident in and out.

Both forms are timed: int8 (exact int8-MFMA correlation; the headline value) and fft (the reference's
FP32 FFT algorithm). The line also carries the per-stage GPU times from hipEvents on the stream of the
calls, the roofline of the int8 correlation kernel, set_reference times, the PCIe-inclusive rate of the
host API, and a CPU baseline: the oracle's restatement of the reference CPU algorithm (FP32 FFT,
OpenMP over windows) on a bounded sample of the same frames. The 'batch' object times the batching
form: 16 uneven clips in one hipGraph launch per step (framecor.Batch, tuned on its first call). The
'periodicity', 'coarse_track' and 'diff' objects time slipcam-auto's periodicity of the reference frame,
its coarse tracking of a 720p video, and the differentiation form on its 1080p geometry.

  python bench.py [--steps K] [--warmup W] [--frames 64] [--no-cpu-baseline]
"""

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'ms/frame + windows/s @720p 2450 win 64² ±48: int8-MFMA vs FP32-FFT'
INT8_PEAK_TOPS = 5000.0        # MI355X dense int8 MFMA, TOP/s
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
FP32_PEAK_TFLOPS = 157.3       # MI355X FP32 vector (packed), TFLOP/s

# HBM traffic per launch of the two correlation kernels on this workload at 64 frames per launch, in KiB
# per dispatch as rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE read them (one pass each), and the factors
# that turn them into bytes: what the same counters read for a plain 256 MiB device copy
# (profiles/counters_int8Quad_fastCorrelate_720p_64frames.md holds the passes, the commands and the commit
# they were measured on; measure again whenever a kernel changes)
TRAFFIC_KIB = {'int8': (3.033e6, 1.201e6), 'fft': (1.676e6, 84680.0)}
TRAFFIC_SOURCE = 'profiles/counters_int8Quad_fastCorrelate_720p_64frames.md (commit 94ed00f)'
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0
TRAFFIC_FRAMES = 64


def traffic_bytes(method, frames):
    """HBM bytes per launch from the profiled counters, or None for a workload other than the profiled one."""
    if frames != TRAFFIC_FRAMES:
        return None
    fetch, write = TRAFFIC_KIB[method]
    return round(1024 * (FETCH_SCALE * fetch + WRITE_SCALE * write))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform's processor string"""
    try:
        with open('/proc/cpuinfo') as fh:
            for line in fh:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or 'unknown'


def time_method(torch, framecor, d, dev, method, steps, warmup):
    h, w = d['frames'].shape[1:]
    W, R, skip = d['window'], d['search'], d['skip']
    down, across = framecor.window_grid(h, w, W, R, skip)
    cor = framecor.Correlator(h, w, down, across, window=W, search=R, method=method)
    t0 = time.perf_counter()
    cor.set_reference(d['ref'])
    set_ref_ms = (time.perf_counter() - t0) * 1e3
    n = dev.shape[0]
    for _ in range(warmup):
        cor.correlate(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        cor.correlate(dev)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # the same steps again with per-stage hipEvents (the stream of the calls)
    cor.set_timing(True)
    stages = dict(box_sums=0.0, correlate=0.0, launches=0)
    for _ in range(steps):
        cor.correlate(dev)
        t = cor.timing()
        for k in stages:
            stages[k] += t[k]
    cor.set_timing(False)
    # the host API once (PCIe-inclusive: frames in, dict out)
    t0 = time.perf_counter()
    cor.correlate(d['frames'])
    host_ms = (time.perf_counter() - t0) * 1e3
    nwin = len(down) * len(across)
    ms_step = elapsed / steps * 1e3
    return dict(cor=cor, nwin=nwin, n=n, ms_per_step=ms_step, ms_per_frame=ms_step / n,
                frames_per_s=n * steps / elapsed, windows_per_s=nwin * n * steps / elapsed,
                set_reference_ms=set_ref_ms, host_api_ms_per_frame=host_ms / n,
                stage_ms_per_frame={k: stages[k] / (steps * n) for k in ('box_sums', 'correlate')},
                kernel_ms_per_launch=stages['correlate'] / max(stages['launches'], 1),
                frames_per_launch=n * steps / max(stages['launches'], 1))


def time_batch(torch, framecor, specs, vids, method, steps, warmup):
    """One graph launch over every clip per step (the batching form), tuned on its first call; then the
    first call of a second batch of the same clips that reads the tuned plans from the plan cache"""
    import tempfile
    cache = os.path.join(tempfile.mkdtemp(prefix='framecor_bench_'), 'plans.json')
    frames = [fr for _, fr in vids]

    def first_call():
        b = framecor.Batch(specs, method=method, tune=True, plan_cache=cache)
        for c, (ref, _) in enumerate(vids):
            b.set_reference(c, ref)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.correlate(frames)
        torch.cuda.synchronize()
        return b, (time.perf_counter() - t0) * 1e3

    batch, first_ms = first_call()
    cached, cached_ms = first_call()
    del cached
    for _ in range(warmup):
        batch.correlate(frames)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.correlate(frames)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    wf = sum(len(s['down']) * len(s['across']) * s['frames'] for s in specs)
    ops = sum(2.0 * s['window'] ** 2 * (2 * s['search'] + 1) ** 2 * len(s['down']) * len(s['across']) * s['frames']
              for s in specs)
    plans = {}
    for c, s in enumerate(specs):
        p = batch.plan(c)
        plans[f'{s["window"]}/{s["search"]}'] = f'{p["variant"]} x{p["frames_per_item"]} {p["blocks_per_cu"]}/CU'
    info = batch.info()
    return dict(clips=len(specs), window_frames=wf, ms_per_call=round(ms, 3), windows_per_s=round(wf / ms * 1e3, 1),
                first_call_ms=round(first_ms, 1), first_call_ms_with_cached_plans=round(cached_ms, 1),
                graph_nodes=info['graph_nodes'], stages=info['stages'],
                tile_shapes=info['shapes'], plans=plans,
                **({'tops': round(ops / ms * 1e3 / 1e12, 1)} if method == 'int8' else {}))


def batch_workload(torch, framecor, synthetic, steps, warmup):
    """The batching workload: 16 seeded clips of uneven sizes (resolutions 480x640 / 720x1280 /
    1080x1920, windows 32/64/128, search 16/32/64, 8-256 frames, skip = window / 4), frames made on the
    GPU; both forms over all clips"""
    specs = synthetic.clip_specs()
    vids = [synthetic.clip_video(s, 'cuda')[:2] for s in specs]
    torch.cuda.synchronize()
    out = {'workload': '16 clips: ' + ', '.join(f'{s["height"]}p W{s["window"]} R{s["search"]} x{s["frames"]}'
                                                for s in specs)}
    log('bench: batch int8')
    out['int8'] = time_batch(torch, framecor, specs, vids, 'int8', steps, warmup)
    log('bench: batch fft')
    out['fft'] = time_batch(torch, framecor, specs, vids, 'fft', steps, warmup)
    return out


def time_periodicity(d, repeats=3):
    """slipcam-auto's periodicity of the reference frame on this geometry (python/slipcam/auto_slip.py:308):
    the exact NCC surface of every window with its surroundings and its second peak, through the host
    API (one synchronous call: the frame in, one float64 per window out)"""
    from slipcam_amd import framecor
    h, w = d['ref'].shape
    W, R = d['window'], d['search']
    down, across = framecor.window_grid(h, w, W, R, d['skip'])
    cor = framecor.Correlator(h, w, down, across, window=W, search=R, method='int8')
    cor.set_reference(d['ref'])
    cor.periodicity(d['ref'])
    ms = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        per = cor.periodicity(d['ref'])
        ms.append((time.perf_counter() - t0) * 1e3)
    M = 2 * R + 1
    ops = 2.0 * W * W * M * M * len(down) * len(across)
    return {'ms': round(min(ms), 3), 'windows': int(per.size), 'tops': round(ops / (min(ms) * 1e-3) / 1e12, 1),
            'periodic_windows': int((per > 0.8).sum()), 'api': 'Correlator.periodicity (slipcam_amd.auto_slip.periodicity)'}


def time_coarse_track(frames=16):
    """slipcam-auto's coarse tracking (python/slipcam/stabilize.py:55-96) of a 720 x 1280 video whose
    camera moves by a known similarity transform per frame: window 96, step 96, search 32, per frame one
    GPU call for every window's NCC peak and the host's robust fit, through the host API"""
    from slipcam_amd import stabilize, synthetic
    v = synthetic.similarity_video(n=frames, height=720, width=1280, ref=frames // 2)
    stabilize.coarse_track(v['frames'][:3], 1, v['text'])
    t0 = time.perf_counter()
    params, used, ngrid = stabilize.coarse_track(v['frames'], v['ref'], v['text'])
    ms = (time.perf_counter() - t0) * 1e3
    return {'ms_per_frame': round(ms / (frames - 1), 3), 'frames': frames, 'windows': int(ngrid),
            'max_shift_error_px': round(float(np.abs(params[:, :2] - v['params'][:, :2]).max()), 4),
            'api': 'slipcam_amd.stabilize.coarse_track'}


def time_diff(torch, framecor, synthetic, frames=32, steps=3, repeats=7):
    """The differentiation form on the 1080 x 1920 geometry of its workload (window 64, search 32, skip 32:
    1624 windows; 32 of its frames, float32 on the GPU): the NCC of every window at continuous shifts,
    forward and backward (gradients of the summed NCC in the frames and the shifts), and Gauss-Newton
    refinement of the shifts, timed per frame"""
    from slipcam_amd.framecor import autograd
    d = synthetic.smooth_shifts_device(n=frames, device='cuda')
    h, w = d['ref'].shape
    down, across = framecor.window_grid(h, w, d['window'], d['search'], d['skip'])
    cor = framecor.Correlator(h, w, down, across, window=d['window'], search=d['search'], method='int8')
    cor.set_reference(d['ref'])
    dev = d['frames'].to(torch.float32).requires_grad_(True)
    shifts = torch.zeros((frames, len(down), len(across), 2), dtype=torch.float32, device='cuda',
                         requires_grad=True)
    # the integer offsets on the device once, as a caller that keeps them there
    gross = torch.tensor(d['gross'], dtype=torch.int32, device='cuda')

    def fwd_bwd():
        # the gradients of the summed NCC in the frames and the shifts (not accumulated into .grad)
        torch.autograd.grad(autograd.ncc_shift(cor, dev, shifts, gross=gross).sum(), [dev, shifts])

    def timed(fn):
        # one untimed call, then each call timed alone: mean and median of the same calls
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(np.mean(ts)), float(np.median(ts))

    fb = timed(fwd_bwd)
    with torch.no_grad():
        rf = timed(lambda: autograd.refine(cor, dev.detach(), shifts.detach(), gross=gross, steps=steps))
    return {'frames': frames, 'windows': len(down) * len(across), 'geometry': f'{h}x{w} W{d["window"]} R{d["search"]}',
            'forward_backward_ms_per_frame': round(fb[1] / frames, 4), 'refine_ms_per_frame': round(rf[1] / frames, 4),
            'forward_backward_ms_per_frame_mean': round(fb[0] / frames, 4),
            'refine_ms_per_frame_mean': round(rf[0] / frames, 4), 'calls': repeats,
            'refine_steps': steps, 'dtype': 'f32'}


def _agreement(got, want, valid):
    """offset, peak and SNR differences of two result dicts over the windows valid in both"""
    both = valid & np.isfinite(got['offsets'][..., 0]) & np.isfinite(want['offsets'][..., 0])
    off = np.abs(got['offsets'].astype(np.float64) - want['offsets'])[both]
    snr = np.abs(got['snr'].astype(np.float64) - want['snr'])[both] / np.maximum(np.abs(want['snr'][both]), 1e-30)
    return {'windows': int(both.sum()), 'offset_median_abs_px': float(np.median(off)) if off.size else None,
            'offset_max_abs_px': float(off.max()) if off.size else None,
            'peak_max_abs': float(np.abs(got['peak'].astype(np.float64) - want['peak'])[both].max()) if off.size else None,
            'snr_max_rel': float(snr.max()) if snr.size else None,
            'validity_mismatches': int((np.isfinite(got['offsets'][..., 0]) != np.isfinite(want['offsets'][..., 0])).sum())}


def cpu_baseline(d, seconds, gpu=None):
    """The oracle restatement of the reference CPU algorithm timed on a bounded sample of the workload;
    gpu: the GPU forms' results on the same frames ({'fft': dict, 'int8': dict}), which this leg also
    checks -- the FFT form against the restatement (the reference's own path) on the timed frames, the
    int8 form against the FP64 direct-NCC oracle on 96 sampled windows of 9 frames before, during and
    after the slip ramp"""
    import oracle
    from slipcam_amd import framecor
    h, w = d['frames'].shape[1:]
    W, R, skip = d['window'], d['search'], d['skip']
    down, across = framecor.window_grid(h, w, W, R, skip)
    o = oracle.OracleCorrelator(h, w, down, across, window=W, search=R)
    o.set_reference(d['ref'])
    done, t0 = 0, time.perf_counter()
    outs = []
    while done < len(d['frames']) and (done == 0 or time.perf_counter() - t0 < seconds):
        outs.append(o.correlate_fft(d['frames'][done:done + 1]))
        done += 1
    elapsed = time.perf_counter() - t0
    nwin = len(down) * len(across)
    line = dict(value=round(nwin * done / elapsed, 1), unit='windows/s', cores=oracle.threads(), kind='port',
                cpu_model=cpu_model(), ms_per_frame=round(elapsed / done * 1e3, 2),
                sample=f'{done} frames x {nwin} windows of the same workload, the oracle restatement of the '
                       f'reference CPU algorithm (own FP32 FFT, no FFTW), OpenMP over windows')
    if gpu:
        acc = {}
        ref_path = {k: np.concatenate([r[k] for r in outs]) for k in outs[0]}
        if 'fft' in gpu:
            g = {k: v[:done] for k, v in gpu['fft'].items()}
            acc['fft_vs_reference_algorithm'] = _agreement(g, ref_path, np.ones(g['peak'].shape, bool))
            # integer peaks: the rounded offsets agree except where the two FP32 surfaces near-tie
            both = np.isfinite(g['offsets'][..., 0]) & np.isfinite(ref_path['offsets'][..., 0])
            acc['fft_vs_reference_algorithm']['integer_peak_mismatches'] = int(
                (np.rint(g['offsets'][both]) != np.rint(ref_path['offsets'][both])).any(-1).sum())
        if 'int8' in gpu:
            # frames before, during and after the slip ramp (frames 20-40), 96 windows
            fsel = [f for f in (0, 16, 22, 26, 30, 34, 40, 48, 63) if f < len(d['frames'])]
            sample = np.random.default_rng(11).choice(nwin, 96, replace=False)
            want = o.correlate_direct(d['frames'][fsel], windows=sample)
            g = {k: v[fsel].reshape((len(fsel), nwin) + v.shape[3:])[:, sample] for k, v in gpu['int8'].items()}
            wnt = {k: v.reshape((len(fsel), nwin) + v.shape[3:])[:, sample] for k, v in want.items()}
            acc['int8_vs_fp64_oracle'] = _agreement(g, wnt, np.ones(g['peak'].shape, bool))
            acc['int8_vs_fp64_oracle']['sample'] = f'96 windows x frames {fsel}'
        line['accuracy'] = acc
    return line


def main():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=10)
    p.add_argument('--warmup', type=int, default=2)
    p.add_argument('--frames', type=int, default=64)
    p.add_argument('--cpu-seconds', type=float, default=15.0)
    p.add_argument('--no-cpu-baseline', action='store_true')
    p.add_argument('--no-batch', action='store_true', help='skip the batching workload (16 uneven clips)')
    p.add_argument('--batch-only', action='store_true', help='time only the batching workload')
    p.add_argument('--no-extras', action='store_true',
                   help='skip the periodicity, coarse tracking and differentiation timings')
    p.add_argument('--methods', default='int8,fft', help='the forms to time (the first is the headline)')
    p.add_argument('--copy-bytes', type=int, default=0,
                   help='instead: time a plain device copy of this many bytes (calibrates the HBM counters)')
    args = p.parse_args()
    if args.gpus != 1:
        print('bench.py: the correlator runs on one GPU (it is single-device; frames or '
              'windows would shard across GPUs with no collective on the hot loop). Run with --gpus 1.',
              file=sys.stderr)
        sys.exit(2)

    import torch

    from slipcam_amd import framecor, synthetic

    if args.copy_bytes:
        src = torch.randint(0, 255, (args.copy_bytes,), dtype=torch.uint8, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(args.warmup):
            dst.copy_(src)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        print(json.dumps({'copy_bytes': args.copy_bytes, 'ms': dt * 1e3, 'GB/s': 2 * args.copy_bytes / dt / 1e9}))
        return

    if args.batch_only:
        print(json.dumps({'batch': batch_workload(torch, framecor, synthetic, args.steps, args.warmup)}), flush=True)
        return
    methods = args.methods.split(',')
    log('bench: generating the workload')
    d = synthetic.fault_video(n=args.frames)
    dev = torch.from_numpy(d['frames']).cuda()
    torch.cuda.synchronize()
    res = {}
    for method in methods:
        log(f'bench: timing {method}')
        res[method] = time_method(torch, framecor, d, dev, method, args.steps, args.warmup)
        log(f'bench: {method} {res[method]["ms_per_frame"]:.4f} ms/frame')
    if methods != ['int8', 'fft']:
        print(json.dumps({m: {k: v for k, v in r.items() if k != 'cor'} for m, r in res.items()}), flush=True)
        return
    i8, ff = res['int8'], res['fft']
    W, R, nwin = d['window'], d['search'], i8['nwin']
    M = 2 * R + 1
    ops_per_frame = 2.0 * W * W * M * M * nwin
    ops_launch = ops_per_frame * i8['frames_per_launch']
    achieved = ops_launch / (i8['kernel_ms_per_launch'] * 1e-3) / 1e12
    L = W + 2 * R
    spectra_bytes = nwin * L * (L // 2 + 1) * 8
    # the FFT form: a block takes a run of frames of one window (at most 16, and at least 2048 blocks;
    # the launch rule of the library), so the spectra pass once per run; its FLOPs per window are the
    # forward and inverse real FFTs (2 x 2.5 L^2 log2 L^2), the spectrum product and the normalization
    fft_runs = max(-(-int(ff['frames_per_launch']) // 16), -(-2048 // nwin))
    fft_runs = min(fft_runs, int(ff['frames_per_launch']))
    fft_bytes_launch = spectra_bytes * fft_runs + ff['frames_per_launch'] * (d['frames'][0].nbytes + nwin * 7 * 4)
    fft_flops_frame = nwin * (5.0 * L * L * math.log2(L * L) + 6.0 * L * (L // 2 + 1) + 12.0 * M * M)
    fft_tflops = fft_flops_frame * ff['frames_per_launch'] / (ff['kernel_ms_per_launch'] * 1e-3) / 1e12
    # the int8 kernel's unique HBM bytes per launch: the reference windows, and per frame the frame,
    # its screen map (8 bytes) and sum-of-squares map (4 bytes) per window position, and the outputs
    bh, bw = d['frames'].shape[1] - W + 1, d['frames'].shape[2] - W + 1
    i8_bytes_launch = nwin * W * W + i8['frames_per_launch'] * (d['frames'][0].nbytes + 12 * bh * bw + nwin * 7 * 4)
    fft_gbs = fft_bytes_launch / (ff['kernel_ms_per_launch'] * 1e-3) / 1e9

    line = {
        'metric': METRIC,
        'value': round(i8['windows_per_s'], 1),
        'unit': 'windows/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(i8['ms_per_step'], 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'int8',
        'data': 'synthetic',
        'config': {'workload': f'fault_video {d["frames"].shape[1]}x{d["frames"].shape[2]} uint8, '
                               f'{args.frames} frames per step, window {W}, search {R}, skip {d["skip"]} '
                               f'({nwin} windows), frames resident in HBM',
                   'frames_per_step': args.frames, 'windows': nwin},
        'ms_per_frame': round(i8['ms_per_frame'], 5),
        'frames_per_s': round(i8['frames_per_s'], 2),
        'fft': {'ms_per_frame': round(ff['ms_per_frame'], 5), 'frames_per_s': round(ff['frames_per_s'], 2),
                'windows_per_s': round(ff['windows_per_s'], 1), 'dtype': 'f32',
                'stage_ms_per_frame': {k: round(v, 5) for k, v in ff['stage_ms_per_frame'].items()},
                'kernel_ms_per_launch': round(ff['kernel_ms_per_launch'], 4),
                'roofline': {'bound': 'fp32', 'achieved': round(fft_tflops, 2), 'peak': FP32_PEAK_TFLOPS,
                             'unit': 'TFLOP/s', 'frac': round(fft_tflops / FP32_PEAK_TFLOPS, 4),
                             'flops_per_launch': fft_flops_frame * ff['frames_per_launch'],
                             'hbm_gbs': round(fft_gbs, 1), 'hbm_frac': round(fft_gbs / HBM_PEAK_GBS, 4),
                             'traffic': traffic_bytes('fft', ff['frames_per_launch']), 'traffic_source': TRAFFIC_SOURCE,
                             'kernel': 'fastCorrelate<160>',
                             'algorithmic_bytes': fft_bytes_launch},
                'set_reference_ms': round(ff['set_reference_ms'], 1),
                'host_api_ms_per_frame': round(ff['host_api_ms_per_frame'], 4)},
        'int8_over_fft': round(ff['ms_per_frame'] / i8['ms_per_frame'], 3),
        'stage_ms_per_frame': {k: round(v, 5) for k, v in i8['stage_ms_per_frame'].items()},
        'set_reference_ms': round(i8['set_reference_ms'], 1),
        'host_api_ms_per_frame': round(i8['host_api_ms_per_frame'], 4),
        'roofline': {'bound': 'mfma', 'achieved': round(achieved, 1), 'peak': INT8_PEAK_TOPS, 'unit': 'TOP/s',
                     'frac': round(achieved / INT8_PEAK_TOPS, 4),
                     'traffic': traffic_bytes('int8', i8['frames_per_launch']), 'traffic_source': TRAFFIC_SOURCE,
                     'algorithmic_bytes': i8_bytes_launch,
                     'kernel': 'int8Quad<6,5,16>', 'kernel_ms_per_launch': round(i8['kernel_ms_per_launch'], 4),
                     'ops_per_launch': ops_launch},
    }
    if not args.no_extras:
        log('bench: periodicity, coarse tracking, differentiation')
        line['periodicity'] = time_periodicity(d)
        line['coarse_track'] = time_coarse_track()
        line['diff'] = time_diff(torch, framecor, synthetic)
    if not args.no_batch:
        line['batch'] = batch_workload(torch, framecor, synthetic, max(2, args.steps // 2), 1)
    if not args.no_cpu_baseline:
        log('bench: CPU baseline')
        gpu_out = {m: res[m]['cor'].correlate(d['frames']) for m in ('fft', 'int8')}
        line['cpu_baseline'] = cpu_baseline(d, args.cpu_seconds, gpu_out)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
