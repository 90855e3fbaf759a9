"""
The batching form (csrc/batch.hip, framecor.Batch): clips of uneven sizes in one hipGraph launch of
persistent work-queue kernels. Per clip, the batch's outputs are bit-identical to the clip run alone
through the Correlator with the same kernel variant (a size-independent property: the batch only
moves the same per-window, per-frame computation between blocks), for both forms, with and without
gross offsets, over several stages of frames, over more clips than one queue launch holds, after
tuning, and on replays of the captured graph.
"""

import os

import numpy as np
import pytest

from oracle import OracleCorrelator
from parity import assert_fft_parity, assert_int8_parity, flat_select
from slipcam_amd import framecor, synthetic

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

KEYS = ('offsets', 'peak', 'snr', 'cov')


def every_shape(frames=(1, 4), seed=11, resolutions=synthetic.RESOLUTIONS):
    """One clip per (window, search) pair of the batching workload, resolutions drawn at random"""
    specs = []
    i = 0
    for W in (32, 64, 128):
        for R in (16, 32, 64):
            s = synthetic.clip_specs(n=1, seed=seed + i, frames=frames, resolutions=resolutions,
                                     windows=(W,), searches=(R,))[0]
            specs.append(s)
            i += 1
    return specs


def videos(specs):
    out = []
    for s in specs:
        ref, fr, _ = synthetic.clip_video(s, 'cuda')
        out.append((ref, fr))
    return out


def solo(spec, ref, frames, method, variant='auto', gross=None):
    cor = framecor.Correlator(spec['height'], spec['width'], spec['down'], spec['across'], window=spec['window'],
                              search=spec['search'], method=method, stat_size=spec.get('stat_size', 21))
    if method == 'int8':
        cor.set_variant(variant)
    cor.set_reference(ref)
    r = cor.correlate(frames, gross=gross)
    return {k: (v.cpu().numpy() if isinstance(v, torch.Tensor) else v) for k, v in r.items()}


def same(a, b, where=''):
    for k in KEYS:
        x = a[k].cpu().numpy() if isinstance(a[k], torch.Tensor) else a[k]
        assert np.array_equal(x, b[k], equal_nan=True), (where, k)


def run_batch(specs, vids, method, tune=False, gross=None, stage_frames=None):
    batch = framecor.Batch(specs, method=method, tune=tune, plan_cache=None)
    for c, (ref, _) in enumerate(vids):
        batch.set_reference(c, ref)
    if stage_frames:
        batch.set_stage_frames(stage_frames)
    out = batch.correlate([fr for _, fr in vids], gross=gross)
    torch.cuda.synchronize()
    return batch, out


def test_int8_batch_matches_each_clip_alone(gpu):
    specs = every_shape()
    vids = videos(specs)
    batch, out = run_batch(specs, vids, 'int8')
    assert batch.info()['shapes'] == 9 and batch.info()['graph_nodes'] > 0
    for c, (s, (ref, fr)) in enumerate(zip(specs, vids)):
        expect = solo(s, ref, fr, 'int8', batch.plan(c)['variant'])
        same(out[c], expect, f'clip {c} W {s["window"]} R {s["search"]} {s["height"]}x{s["width"]}')
        assert np.isfinite(expect['offsets']).mean() > 0.5


def test_fft_batch_matches_each_clip_alone(gpu):
    specs = every_shape(seed=31)   # search windows 64 .. 256 (256: spectra in the global workspace)
    vids = videos(specs)
    batch, out = run_batch(specs, vids, 'fft')
    for c, (s, (ref, fr)) in enumerate(zip(specs, vids)):
        same(out[c], solo(s, ref, fr, 'fft'), f'clip {c}')


def test_tuned_batch_replays_the_same(gpu):
    """Tuning picks a plan per tile shape; the results are those of each clip alone with the chosen
    variant, and a replay of the captured graph gives the same bits"""
    specs = every_shape(frames=(2, 6), seed=51, resolutions=((480, 640), (720, 1280)))
    vids = videos(specs)
    batch, out = run_batch(specs, vids, 'int8', tune=True)
    first = [{k: v.clone() for k, v in o.items()} for o in out]
    nodes = batch.info()['graph_nodes']
    for c, (s, (ref, fr)) in enumerate(zip(specs, vids)):
        plan = batch.plan(c)
        assert plan['tuned'] and plan['frames_per_item'] >= 1
        same(out[c], solo(s, ref, fr, 'int8', plan['variant']), f'clip {c} plan {plan}')
    again = batch.correlate([fr for _, fr in vids])
    torch.cuda.synchronize()
    assert batch.info()['graph_nodes'] == nodes
    for c in range(len(specs)):
        same(again[c], {k: v.cpu().numpy() for k, v in first[c].items()}, f'replay clip {c}')


@pytest.mark.parametrize('variant', ['quad', 'simple'])
def test_int8_batch_plans_search_32_and_48(gpu, variant):
    """The four-wave kernel's shapes (windows 32 and 64, search 32 and 48: c = 4 and 6 core bands) under
    explicit plans, 1 and 16 frames per item, 4 blocks per CU: the same bits as each clip alone (the
    persistent form of c = 6 is compiled with each work item behind a call, which gives these bits)"""
    specs = []
    for i, (W, R) in enumerate(((32, 32), (32, 48), (64, 32), (64, 48))):
        specs.append(synthetic.clip_specs(n=1, seed=140 + i, frames=(5, 5), resolutions=((480, 640),), windows=(W,),
                                          searches=(R,))[0])
    vids = videos(specs)
    for fpb in (1, 16):
        batch = framecor.Batch(specs, method='int8', tune=False, plan_cache=None)
        for c, (ref, _) in enumerate(vids):
            batch.set_reference(c, ref)
            batch.set_plan(c, variant, fpb, 4)
        out = batch.correlate([fr for _, fr in vids])
        torch.cuda.synchronize()
        for c, (s, (ref, fr)) in enumerate(zip(specs, vids)):
            assert batch.plan(c)['variant'] == variant
            expect = solo(s, ref, fr, 'int8', variant)
            same(out[c], expect, f'clip {c} W {s["window"]} R {s["search"]} frames per item {fpb}')
            assert np.isfinite(expect['offsets']).mean() > 0.5


def test_int8_batch_one_wave_plans(gpu):
    """The one-wave kernel's shapes (windows up to 64; one, two and three full 16-row bands at search 4,
    12 and 23; two, three and four core bands with the edge row and column at search 16, 24 and 32)
    under explicit plans, 1 and 16 frames per item, 8 and 12 blocks per CU: the same bits as each clip
    alone with the same kernel and as the edge-wave kernel"""
    specs = []
    for i, (W, R) in enumerate(((16, 4), (32, 12), (32, 16), (64, 16), (48, 23), (48, 24), (32, 32))):
        specs.append(synthetic.clip_specs(n=1, seed=160 + i, frames=(5, 5), resolutions=((480, 640),), windows=(W,),
                                          searches=(R,))[0])
    vids = videos(specs)
    for fpb, per_cu in ((1, 12), (16, 8)):
        batch = framecor.Batch(specs, method='int8', tune=False, plan_cache=None)
        for c, (ref, _) in enumerate(vids):
            batch.set_reference(c, ref)
            batch.set_plan(c, 'solo', fpb, per_cu)
        out = batch.correlate([fr for _, fr in vids])
        torch.cuda.synchronize()
        for c, (s, (ref, fr)) in enumerate(zip(specs, vids)):
            assert batch.plan(c)['variant'] == 'solo'
            expect = solo(s, ref, fr, 'int8', 'solo')
            same(out[c], expect, f'clip {c} W {s["window"]} R {s["search"]} frames per item {fpb}')
            same(expect, solo(s, ref, fr, 'int8', 'simple'), f'clip {c} against the edge-wave kernel')
            assert np.isfinite(expect['offsets']).mean() > 0.5


@pytest.mark.parametrize('method', ['int8', 'fft'])
def test_batch_with_gross_offsets(gpu, method):
    specs = synthetic.clip_specs(n=4, seed=71, frames=(2, 5), windows=(32, 64), searches=(16, 32))
    vids = videos(specs)
    gross = []
    for c, s in enumerate(specs):
        n, ny, nx = s['frames'], len(s['down']), len(s['across'])
        g = synthetic.gross_offsets(n, ny, nx, seed=c) if c % 2 else np.random.default_rng(c).integers(-20, 21, (n, 2))
        gross.append(g.astype(np.int32))
    _, out = run_batch(specs, vids, method, gross=gross)
    for c, (s, (ref, fr)) in enumerate(zip(specs, vids)):
        same(out[c], solo(s, ref, fr, method, gross=gross[c]), f'clip {c}')


def test_many_clips_over_stages(gpu):
    """More clips of one shape than a queue launch holds (48), frames over three stages"""
    specs = []
    for i in range(50):
        h, w = 96 + 8 * (i % 3), 128 + 16 * (i % 4)
        down, across = framecor.window_grid(h, w, 16, 8, 12)
        specs.append(dict(height=h, width=w, window=16, search=8, down=down, across=across,
                          frames=5 + i % 3, sigma=1.0 + (i % 5) * 0.5, seed=900 + i))
    vids = videos(specs)
    batch, out = run_batch(specs, vids, 'int8', stage_frames=2)
    info = batch.info()
    # stages 0-2: all 50 clips (two queue launches each); stage 3: the 16 clips of 7 frames (one)
    assert info['stages'] == 4 and info['queues'] == 7
    for c, (s, (ref, fr)) in enumerate(zip(specs, vids)):
        same(out[c], solo(s, ref, fr, 'int8', batch.plan(c)['variant']), f'clip {c}')


def test_host_arrays(gpu):
    specs = synthetic.clip_specs(n=3, seed=81, frames=(1, 3), resolutions=((480, 640),))
    vids = videos(specs)
    batch = framecor.Batch(specs, method='int8', tune=False)
    for c, (ref, _) in enumerate(vids):
        batch.set_reference(c, ref)
    dev = batch.correlate([fr for _, fr in vids])
    dev = [{k: v.cpu().numpy() for k, v in o.items()} for o in dev]
    host = batch.correlate([fr.cpu().numpy() for _, fr in vids])
    for c in range(len(specs)):
        same(host[c], dev[c], f'clip {c}')


def test_stat_size_per_clip(gpu):
    """Clips of one window and search range with different SNR windows (21, an even 8, 3, and 40, larger
    than the 33 x 33 surface): each kernel launch sized for its clips' SNR box; the same bits as each clip
    alone"""
    base = synthetic.clip_specs(n=1, seed=91, frames=(3, 3), resolutions=((480, 640),), windows=(32,),
                                searches=(16,))[0]
    specs = [dict(base, stat_size=st, seed=base['seed'] + i) for i, st in enumerate((21, 8, 3, 40))]
    vids = videos(specs)
    batch, out = run_batch(specs, vids, 'int8')
    assert batch.info()['shapes'] == 4
    for c, (s, (ref, fr)) in enumerate(zip(specs, vids)):
        same(out[c], solo(s, ref, fr, 'int8', batch.plan(c)['variant']), f'clip {c} stat {s["stat_size"]}')


@pytest.mark.parametrize('method', ['int8', 'fft'])
def test_batching_workload_16_clips(gpu, method):
    """The batching workload itself: 16 seeded clips (480p-1080p, windows 32-128, search 16-64, 8-256 frames)
    in one graph launch (tuned: int8), each clip's outputs the bits of that clip alone through the
    Correlator, and per clip 6 windows of its first and last frames against the oracle (int8: the FP64
    direct NCC; fft: the restatement of the reference algorithm, near-ties classified on FP64 surfaces)"""
    specs = synthetic.clip_specs()
    vids = videos(specs)
    batch, out = run_batch(specs, vids, method, tune=method == 'int8')
    # the maps of all 16 clips fit on the device at once: one stage; one queue launch per tile shape and
    # the counters' fill kernel, and the box maps in one launch (int8) or one per tile shape (fft)
    info = batch.info()
    boxes = 1 if method == 'int8' else info['shapes']
    assert info['stages'] == 1 and info['graph_nodes'] == 1 + boxes + info['queues'], info
    rng = np.random.default_rng(17)
    for c, (s, (ref, fr)) in enumerate(zip(specs, vids)):
        where = f'clip {c}: {s["height"]}p W{s["window"]} R{s["search"]} x{s["frames"]}'
        expect = solo(s, ref, fr, method, batch.plan(c)['variant'])
        same(out[c], expect, where)
        assert np.isfinite(expect['offsets'][..., 0]).mean() > 0.5, where
        o = OracleCorrelator(s['height'], s['width'], s['down'], s['across'], window=s['window'], search=s['search'])
        o.set_reference(ref)
        fsel = [0, s['frames'] - 1]
        host = fr[fsel].cpu().numpy()
        nwin = len(s['down']) * len(s['across'])
        sample = np.sort(rng.choice(nwin, min(6, nwin), replace=False))
        got = flat_select({k: v[fsel] for k, v in expect.items()}, sample)
        if method == 'int8':
            assert_int8_parity(got, flat_select(o.correlate_direct(host, windows=sample), sample))
        else:
            want = o.correlate_fft(host, windows=sample)
            surf = o.correlate_direct(host, windows=sample, surfaces=True)['surface']
            keep = np.zeros(want['peak'].shape[1:], bool)
            keep.flat[sample] = True
            g2 = {k: v[fsel].copy() for k, v in expect.items()}
            for k in g2:   # the windows the oracle skipped: its fill values
                g2[k][:, ~keep] = want[k][:, ~keep]
            assert_fft_parity(g2, want, s['search'], surf=surf)


def test_plan_cache(gpu, tmp_path):
    """Tuned plans persist: a second batch of the same tile shapes reads them from the cache file, runs
    without tuning and gives the first batch's bits"""
    specs = every_shape(frames=(2, 4), seed=61, resolutions=((480, 640),))[:3]
    vids = videos(specs)
    path = str(tmp_path / 'plans.json')
    first = framecor.Batch(specs, method='int8', tune=True, plan_cache=path)
    assert first.plans_from_cache == 0 and not os.path.exists(path)
    for c, (ref, _) in enumerate(vids):
        first.set_reference(c, ref)
    out1 = [{k: v.cpu().numpy() for k, v in o.items()} for o in first.correlate([fr for _, fr in vids])]
    assert os.path.exists(path)
    plans = [first.plan(c) for c in range(len(specs))]
    second = framecor.Batch(specs, method='int8', tune=True, plan_cache=path)
    assert second.plans_from_cache == len(specs)
    assert [second.plan(c) for c in range(len(specs))] == plans
    for c, (ref, _) in enumerate(vids):
        second.set_reference(c, ref)
    out2 = second.correlate([fr for _, fr in vids])
    for c in range(len(specs)):
        same(out2[c], out1[c], f'clip {c}')
