"""
Host-side checks that need no GPU: the C ABI library loads and exports every entry point that
include/framecor_hip.h declares, window_grid, and the argument checks of the Correlator (which the
library makes before touching a GPU, with the reference's messages, cxx/common.cpp:15-28).
"""

import ctypes
import glob
import os
import re

import numpy as np
import pytest

from slipcam_amd import framecor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for header in glob.glob(os.path.join(ROOT, 'include', '*.h')):
        text = open(header).read()
        text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
        names |= set(re.findall(r'\b(framecor_\w+)\s*\(', text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(framecor._LIB_PATH)
    names = declared_symbols()
    assert 'framecor_create' in names and 'framecor_correlate_device' in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_probe():
    assert framecor.__version__ == '0.1.0'
    assert isinstance(framecor.has_gpu(), bool)


def test_window_grid_readme_shape():
    down, across = framecor.window_grid(720, 1280, 64, 48, 16)
    assert len(down) * len(across) == 2450           # README.md:88, 96 of the reference
    assert down[0] == 48 and across[0] == 48 and down[-1] == 592 and across[-1] == 1152


@pytest.mark.parametrize('h,w,W,R,skip,margin', [(240, 320, 32, 24, 24, 0), (256, 320, 32, 16, 32, 0),
                                                 (240, 320, 32, 6, 16, 32), (1080, 1920, 128, 64, 32, 0)])
def test_window_grid_search_windows_in_frame(h, w, W, R, skip, margin):
    down, across = framecor.window_grid(h, w, W, R, skip, margin=margin)
    assert down.dtype.kind == 'i'
    for y in down:
        assert y - R - margin >= 0 and y + W + R + margin <= h
    for x in across:
        assert x - R - margin >= 0 and x + W + R + margin <= w
    # one more window would not fit
    assert down[-1] + skip + W + R + margin > h - skip or len(down) == (h - 2 * (margin + R) - W) // skip


def test_window_grid_nothing_fits():
    with pytest.raises(ValueError, match='no window fits down'):
        framecor.window_grid(100, 1000, 64, 48, 16)
    with pytest.raises(ValueError, match='no window fits across'):
        framecor.window_grid(1000, 100, 64, 48, 16)


def test_arguments_without_gpu():
    with pytest.raises(ValueError, match='outside the frame'):
        framecor.Correlator(240, 320, [0], [300], window=32, search=8, backend='hip')
    with pytest.raises(ValueError, match='outside the frame'):
        framecor.Correlator(240, 320, [0], [300], window=32, search=8, backend='gpu')   # parameters first
    for backend in ('gpu', 'cpu', 'CUDA'):
        with pytest.raises(ValueError, match='backend'):
            framecor.Correlator(240, 320, [0], [0], window=32, search=8, backend=backend)
    # 'cuda' (the reference's name of its GPU backend) is accepted: parameters are still checked first
    with pytest.raises(ValueError, match='outside the frame'):
        framecor.Correlator(240, 320, [0], [300], window=32, search=8, backend='cuda')
    assert framecor.has_cuda() == framecor.has_gpu()
    with pytest.raises(ValueError, match='method'):
        framecor.Correlator(240, 320, [0], [0], window=32, search=8, method='fp16')
    with pytest.raises(ValueError, match='at least 3 pixels'):
        framecor.Correlator(240, 320, [0], [0], window=2, search=8)
    with pytest.raises(ValueError, match='half search range'):
        framecor.Correlator(240, 320, [0], [0], window=32, search=0)
    with pytest.raises(ValueError, match='SNR window'):
        framecor.Correlator(240, 320, [0], [0], window=32, search=8, stat_size=2)
    with pytest.raises(ValueError, match='batch'):
        framecor.Correlator(240, 320, [0], [0], window=32, search=8, batch=0)
    with pytest.raises(ValueError, match='no windows'):
        framecor.Correlator(240, 320, [], [0], window=32, search=8)
    with pytest.raises(ValueError, match='frame size'):
        framecor.Correlator(0, 320, [0], [0], window=32, search=8)
    # what the methods support
    with pytest.raises(ValueError, match='int8'):
        framecor.Correlator(1000, 1000, [0], [0], window=160, search=8, method='int8')
    with pytest.raises(ValueError, match='fft method needs a search window .* of at most 2048'):
        framecor.Correlator(3000, 3000, [0], [0], window=181, search=1000, method='fft')   # L = 2181
    # any L up to 2048 passes the checks (L = 47, a prime; L = 256, beyond LDS)
    for W, R in ((3, 22), (128, 64), (160, 48)):
        assert framecor._lib.framecor_check_params(1000, 1000, np.zeros(1, np.int32).ctypes.data, 1,
                                                   np.zeros(1, np.int32).ctypes.data, 1, W, R, 0, 1, 21) == 0


@pytest.mark.skipif(framecor.has_gpu(), reason='a GPU is present')
def test_no_gpu_is_a_runtime_error():
    with pytest.raises(RuntimeError, match='no GPU'):
        framecor.Correlator(240, 320, [0], [0], window=32, search=8)


def test_boxes_mask_brute_force():
    """The windows flagged by the text boxes (track_offsets.py:27-32): a window is masked iff some pixel
    of it lies inside some box"""
    from slipcam_amd import track
    down, across = framecor.window_grid(720, 1280, 64, 48, 16)
    boxes = [(870, 25, 1240, 68), (40, 650, 630, 700), (0, 0, 1, 1)]
    got = track.boxes_mask(down, across, 64, boxes)
    inside = np.zeros((720, 1280), bool)
    for x0, y0, x1, y1 in boxes:
        inside[y0:y1, x0:x1] = True
    cs = np.pad(inside.cumsum(0).cumsum(1), ((1, 0), (1, 0)))
    for i, y in enumerate(down):
        for j, x in enumerate(across):
            hits = cs[y + 64, x + 64] - cs[y, x + 64] - cs[y + 64, x] + cs[y, x]
            assert got[i, j] == (hits > 0), (y, x)
    assert 0 < got.sum() < got.size


def test_batch_arguments_without_gpu():
    """A batch checks every clip as a Correlator would, naming the clip, before any GPU call"""
    good = dict(height=240, width=320, down=[16, 48], across=[16, 80], window=32, search=8, frames=4)
    with pytest.raises(ValueError, match=r'window must be at least 3 pixels \(clip 1\)'):
        framecor.Batch([good, dict(good, window=2)], method='int8')
    with pytest.raises(ValueError, match=r'outside the frame \(across 400\) \(clip 2\)'):
        framecor.Batch([good, good, dict(good, across=[400])], method='int8')
    with pytest.raises(ValueError, match=r'fft method .* \(clip 0\)'):
        framecor.Batch([dict(good, height=3000, width=3000, window=128, search=1000)], method='fft')
    with pytest.raises(ValueError, match='frames must not be negative'):
        framecor.Batch([dict(good, frames=-1)], method='int8')
    with pytest.raises(ValueError, match='method'):
        framecor.Batch([good], method='fp16')
    with pytest.raises(ValueError, match='at least one clip'):
        framecor.Batch([], method='int8')


def test_product_never_uses_the_oracle():
    """The package neither imports nor loads anything from oracle/ (the checker)"""
    for path in glob.glob(os.path.join(ROOT, 'slipcam_amd', '**', '*.*'), recursive=True):
        if path.endswith(('.py', '.hip', '.h')):
            text = open(path).read()
            assert 'import oracle' not in text and 'from oracle' not in text and 'liboracle' not in text, path


C_CALLER = r'''
#include <stdio.h>
#include <string.h>
#include "framecor_hip.h"

int main(void)
{
    int down[3] = {16, 48, 80}, across[2] = {16, 48};
    int ok = framecor_check_params(256, 320, down, 3, across, 2, 32, 16, FRAMECOR_METHOD_INT8, 4096, 21);
    int bad = framecor_check_params(256, 320, down, 3, across, 2, 2, 16, FRAMECOR_METHOD_FFT, 4096, 21);
    printf("%s|%d|%d|%s\n", framecor_version(), ok, bad, framecor_last_error());
    return 0;
}
'''


def test_header_is_plain_c_and_a_c_program_calls_the_library(tmp_path):
    """The C ABI from C: the header compiles as strict C99, and a C program linked against the library
    checks parameters (no GPU needed) with the reference's statuses and messages"""
    import shutil
    import subprocess
    if shutil.which('gcc') is None:
        pytest.skip('no C compiler')
    src = tmp_path / 'caller.c'
    src.write_text(C_CALLER)
    exe = tmp_path / 'caller'
    libdir = os.path.dirname(framecor._LIB_PATH)
    subprocess.run(['gcc', '-std=c99', '-pedantic', '-Wall', '-Werror', '-I', os.path.join(ROOT, 'include'), str(src),
                    '-o', str(exe), '-L', libdir, '-l:' + os.path.basename(framecor._LIB_PATH), '-Wl,-rpath,' + libdir],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.strip()
    version, ok, bad, msg = out.split('|', 3)
    assert version == framecor.__version__ and int(ok) == 0 and int(bad) == 1
    assert msg == 'framecor: the window must be at least 3 pixels'


def test_framecor_module_without_gpu():
    """The pybind11 module of the reference's API (cxx/bindings.cpp:111-142): its names, keywords and
    defaults, and its argument errors before any GPU call, with the reference's messages"""
    from slipcam_amd.framecor import _framecor
    from slipcam_amd.framecor._framecor import Correlator, __version__, has_cuda  # noqa: F401
    assert __version__ == framecor.__version__ == '0.1.0'
    assert has_cuda() == framecor.has_gpu()
    doc = re.sub(r'typing\.SupportsInt \| typing\.SupportsIndex', 'int', Correlator.__init__.__doc__)
    assert ', *, window' in doc
    for kw in ("window: int = 64", "search: int = 48", "backend: str = 'auto'", "batch: int = 4096",
               "stat_size: int = 21", "device: int = 0", "threads: int = 0", "method: str = 'fft'"):
        assert kw in doc, kw
    with pytest.raises(ValueError, match=r'^framecor: a window is outside the frame \(across 300\)$'):
        Correlator(240, 320, [0], [300], window=32, search=8, backend='cpu')
    with pytest.raises(ValueError, match='^framecor: the backend must be auto, hip or cuda, not gpu$'):
        Correlator(240, 320, [0], [0], window=32, search=8, backend='gpu')
    with pytest.raises(ValueError, match='^framecor: the window must be at least 3 pixels$'):
        Correlator(240, 320, [0], [0], window=2, search=8)
    with pytest.raises(ValueError, match='SNR window'):
        Correlator(240, 320, [0], [0], window=32, search=8, stat_size=2)
    with pytest.raises(ValueError, match='method'):
        Correlator(240, 320, [0], [0], window=32, search=8, method='fp16')
    with pytest.raises(TypeError):
        Correlator(240, 320, [0], [0], 32, 8)    # window and search are keyword-only
    assert issubclass(framecor.Correlator, _framecor.Correlator)
