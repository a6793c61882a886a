"""The oracle's mip chain against the reference's own Image_MipReduce32.

The reference CLI (oracle/_ref/s2tc_compress, built from s2tc_compress.c)
loads oracle/_ref/libtxc_capture.so with -l; that library records every
level the CLI hands to tx_compress_dxtn (s2tc_compress.c:722-733).  The
oracle's chain from the captured level 0 must equal the captured levels, on
the reference's test images and on odd, one-texel-wide and one-texel-high
synthetic images where the truncating filter drops rows and columns.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import imgutil

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_CLI = os.path.join(ROOT, "oracle", "_ref", "s2tc_compress")
CAPTURE = os.path.join(ROOT, "oracle", "_ref", "libtxc_capture.so")
TGA_DIR = os.path.join(ROOT, "tests", "golden", "tga")

SIZES = [(1, 1), (1, 7), (7, 1), (5, 3), (13, 2), (2, 13), (33, 17), (64, 1), (3, 64), (31, 33)]


def _capture(tga, td):
    cap = os.path.join(td, "cap.bin")
    if os.path.exists(cap):
        os.remove(cap)
    subprocess.run([REF_CLI, "-i", tga, "-o", os.path.join(td, "o.dds"), "-t", "DXT5", "-l", CAPTURE],
                   check=True, capture_output=True, timeout=120, env=dict(os.environ, S2TC_CAPTURE_FILE=cap))
    return imgutil.read_capture(cap)


def _check(oracle, levels):
    assert all(c == 4 for _, _, c, _ in levels)
    w0, h0, _, l0 = levels[0]
    mine = oracle.mip_chain(l0, w0, h0)
    assert [(w, h) for w, h, _ in mine] == [(w, h) for w, h, _, _ in levels]
    for k, ((w, h, a), (_, _, _, b)) in enumerate(zip(mine, levels)):
        assert np.array_equal(a, b), (w0, h0, k, w, h)


needs_ref = pytest.mark.skipif(not (os.path.exists(REF_CLI) and os.path.exists(CAPTURE)),
                               reason="reference CLI / capture library not built here")


@needs_ref
@pytest.mark.parametrize("name", sorted(os.listdir(TGA_DIR)))
def test_mip_chain_matches_reference_on_test_images(oracle, name):
    with tempfile.TemporaryDirectory() as td:
        _check(oracle, _capture(os.path.join(TGA_DIR, name), td))


@needs_ref
def test_mip_chain_matches_reference_odd_sizes(oracle):
    rng = np.random.default_rng(11)
    with tempfile.TemporaryDirectory() as td:
        for w, h in SIZES:
            rgba = rng.integers(0, 256, (h, w, 4)).astype(np.uint8)
            rgba[..., 3][rng.random((h, w)) < 0.3] = 255
            tga = os.path.join(td, "in.tga")
            imgutil.write_tga32(tga, rgba, w, h)
            levels = _capture(tga, td)
            assert np.array_equal(levels[0][3], rgba.reshape(-1)), (w, h)
            _check(oracle, levels)


def test_mip_chain_closed_forms(oracle):
    """Properties the truncating box filter fixes without the reference:
    constant images stay constant, sizes halve with floor down to 1x1, and a
    2x2 image reduces to the floor of its per-channel mean."""
    for w, h in SIZES:
        c = np.tile(np.array([17, 200, 255, 3], np.uint8), w * h)
        lv = oracle.mip_chain(c, w, h)
        assert len(lv) == max(w, h).bit_length()
        for lw, lh, a in lv:
            assert np.array_equal(a, np.tile(c[:4], lw * lh))
    rng = np.random.default_rng(5)
    for _ in range(200):
        px = rng.integers(0, 256, (4, 4)).astype(np.uint8)
        lv = oracle.mip_chain(px.reshape(-1), 2, 2)
        assert np.array_equal(lv[1][2], px.astype(np.int64).sum(axis=0) // 4)
