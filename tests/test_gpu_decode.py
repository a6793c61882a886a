"""GPU decoding and the encoding-error metric.

* s2tc_amd_decode_image against the host fetch_2d_texel_* of the library and
  of the oracle, on random block bytes (every index, endpoint order and alpha
  mode occurs) and on real encodings;
* s2tc_amd_colordist_error against the oracle's colordist_error, every
  distance and format, quantized and RGBA8 sources, partial blocks.
"""
import itertools
import sys

import numpy as np
import pytest

import imgutil

pytestmark = pytest.mark.gpu

FMTS = [0x83F0, 0x83F1, 0x83F2, 0x83F3]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("fmt", FMTS)
def test_decode_random_blocks(gpu, oracle, fmt):
    rng = np.random.default_rng(fmt)
    for w, h in [(16, 8), (13, 7), (1, 1), (6, 5)]:
        bw, bh = (w + 3) // 4, (h + 3) // 4
        bs = 8 if fmt in (0x83F0, 0x83F1) else 16
        blocks = rng.integers(0, 256, bw * bh * bs).astype(np.uint8)
        _say(f"case decode fmt={fmt:#x} {w}x{h}")
        got = gpu.decode_image(blocks, fmt, w, h).reshape(h, w, 4)
        for j, i in itertools.product(range(h), range(w)):
            assert np.array_equal(got[j, i], oracle.fetch(fmt, w, blocks, i, j)), (fmt, i, j)


@pytest.mark.parametrize("cd", range(8))
def test_colordist_error_matches_oracle(gpu, oracle, cd):
    rng = np.random.default_rng(40 + cd)
    for (w, h), dxt in itertools.product([(16, 16), (13, 9), (2, 3)], range(3)):
        for kind in ("random", "smooth"):
            q = imgutil.random_quantized(rng, w, h, dxt, kind)
            for nr in (-1, 0):
                blocks = oracle.encode_quantized(q, w, h, dxt, cd, nr, 1, 0)
                _say(f"case error cd={cd} dxt={dxt} {w}x{h} {kind} nrandom={nr}")
                exp = oracle.colordist_error(q, w, h, dxt, cd, blocks)
                assert gpu.colordist_error(q, w, h, blocks, dxt, cd) == exp
            # arbitrary block bytes exercise checkerboard and transparent decodes
            junk = rng.integers(0, 256, blocks.size).astype(np.uint8)
            assert gpu.colordist_error(q, w, h, junk, dxt, cd) == oracle.colordist_error(q, w, h, dxt, cd, junk)


def test_colordist_error_rgba8_source(gpu, oracle):
    rng = np.random.default_rng(9)
    for dxt, cd in [(0, 5), (2, 3), (1, 7)]:
        w, h = 37, 23
        rgba = imgutil.smooth_rgba8(rng, w, h)
        q = oracle.rgb565_image(rgba, w, h, 4, imgutil.ALPHABITS[dxt], 0)
        blocks = oracle.encode_quantized(q, w, h, dxt, cd, -1, 1, 0)
        _say(f"case error rgba8 dxt={dxt} cd={cd}")
        assert gpu.colordist_error(rgba, w, h, blocks, dxt, cd, src_kind=gpu.SRC_RGBA8) == \
            oracle.colordist_error(q, w, h, dxt, cd, blocks)


@pytest.mark.parametrize("fmt", FMTS)
def test_decode_large_images(gpu, oracle, fmt):
    """Whole images of several waves per row against the oracle's decode:
    random blocks, DXT1 blocks with c1 >= c0 and DXT5 blocks in the 6-value
    mode mixed in, among them the opaque block (254, 255, every index 7)."""
    rng = np.random.default_rng(40 + fmt)
    for w, h in [(1030, 518), (512, 512)]:
        bw, bh = (w + 3) // 4, (h + 3) // 4
        bs = 8 if fmt in (0x83F0, 0x83F1) else 16
        blocks = rng.integers(0, 256, (bw * bh, bs)).astype(np.uint8)
        blocks[1::3, 8 if bs == 16 else 0] = 3
        blocks[2::5, 0] = 1
        if fmt == 0x83F3:
            blocks[3::4, :8] = [254, 255, 255, 255, 255, 255, 255, 255]
        blocks = blocks.reshape(-1)
        _say(f"case decode large fmt={fmt:#x} {w}x{h}")
        exp = oracle.decode_image(blocks, fmt, w, h)
        for rep in range(2):
            got = gpu.decode_image(blocks, fmt, w, h)
            bad = np.nonzero((got != exp).reshape(-1, 4).any(axis=1))[0]
            assert bad.size == 0, (fmt, w, h, rep, bad.size, int(bad[0]) % w, int(bad[0]) // w)


def test_decode_opaque_dxt5_encoding(gpu, oracle):
    """An opaque image encoded as DXT5 (alpha 255 everywhere: endpoints
    254, 255 and every index 7) decodes to alpha 255 at every texel, on
    repeated decodes of the same blocks, and to the oracle's colours."""
    rng = np.random.default_rng(77)
    w = h = 1024
    rgba = imgutil.smooth_rgba8(rng, w, h).reshape(h, w, 4).copy()
    rgba[..., 3] = 255
    _say(f"case decode opaque DXT5 {w}x{h}")
    blocks = gpu.encode_image(rgba.reshape(-1), w, h, 2, 5, -1, 1, src_kind=gpu.SRC_RGBA8)
    exp = oracle.decode_image(blocks, 0x83F3, w, h)
    assert (exp.reshape(-1, 4)[:, 3] == 255).all()
    for rep in range(3):
        got = gpu.decode_image(blocks, 0x83F3, w, h)
        bad = np.nonzero((got != exp).reshape(-1, 4).any(axis=1))[0]
        assert bad.size == 0, (rep, bad.size, int(bad[0]) % w, int(bad[0]) // w)
